// MI355X (gfx950) post-processing of one run's per-pair log10 likelihoods,
// so that a caller can take per-read decisions and per-allele likelihoods
// instead of copying the whole read x haplotype matrix to the host.
//
// Per read, in the order the reference driver applies them
// (pairhmm.hpp:83-93 normalize_likelihoods, :95-114
// filter_poorly_modeled_reads, both called from :116-141 and from
// pairhmm_avx.hpp:160-171):
//   best  = max_h L[h]
//   L'[h] = L[h] < best - 4.5 ? best - 4.5 : L[h]
//   keep  = best < min(2, ceil(len * 0.02)) * -4        (the filter as written)
// Per read and allele (genotyper.hpp:222-241 marginal_likelihoods, with the
// haplotype -> allele map of :200-209):
//   A[a]  = max over haplotypes h with map[h] == a of L'[h]; the lowest
//           finite double when no haplotype maps to a.
//
// One thread per read: a read's row is a contiguous run of its region's
// haplotype count (2-32 in short-read regions), so a wavefront reads a
// contiguous block of 64 rows and the per-read work is a handful of compares.
// The kernel is bound by HBM: 8 bytes read per pair, 8 bytes written per pair
// (normalised matrix) or per allele, one byte per read.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "pairhmm_common.h"

namespace bvp {

// pairhmm.hpp:27-36
constexpr double kBestAltDifference = -4.5;
constexpr double kErrorRatePerBase = 0.02;
constexpr double kLog10QualityPerBase = -4.0;
constexpr double kMaxErrorsPerRead = 2.0;

__device__ __forceinline__ int64_t region_of(const int64_t* region_read_off, int64_t n_regions,
                                             int64_t r) {
  // the g with off[g] <= r < off[g + 1] (regions without reads are skipped);
  // invariant: off[lo] <= r < off[hi]
  int64_t lo = 0, hi = n_regions;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (region_read_off[mid] <= r) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void post_kernel(PostArgs a) {
  const int64_t r = a.read0 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= a.read_end) return;
  const int64_t g = region_of(a.region_read_off, a.n_regions, r);
  const int64_t r_local = r - a.region_read_off[g];
  const int64_t h0 = a.region_hap_off[g];
  const int nh = static_cast<int>(a.region_hap_off[g + 1] - h0);
  if (nh <= 0) return;  // not reached: regions with reads have haplotypes
  const double* L = a.L + a.region_pair_off[g] + r_local * nh;
  double best = L[0];
  for (int h = 1; h < nh; ++h) best = L[h] > best ? L[h] : best;
  const double cap = best + kBestAltDifference;
  const bool norm = a.normalize != 0;
  if (a.keep != nullptr) {
    const double len = static_cast<double>(a.read_off[r + 1] - a.read_off[r]);
    const double thr = fmin(kMaxErrorsPerRead, ceil(len * kErrorRatePerBase)) * kLog10QualityPerBase;
    a.keep[r] = best < thr ? 1 : 0;
  }
  if (a.Lnorm != nullptr) {
    double* o = a.Lnorm + a.region_pair_off[g] + r_local * nh;
    for (int h = 0; h < nh; ++h) o[h] = (norm && L[h] < cap) ? cap : L[h];
  }
  if (a.A != nullptr) {
    const int na = a.n_alleles[g];
    double* o = a.A + a.allele_off[g] + r_local * na;
    const int32_t* map = a.hap_allele + h0;
    for (int al = 0; al < na; ++al) {
      double m = -DBL_MAX;
      for (int h = 0; h < nh; ++h) {
        const double v = (norm && L[h] < cap) ? cap : L[h];
        if (map[h] == al && v > m) m = v;
      }
      o[al] = m;
    }
  }
}

hipError_t launch_post(const PostArgs& a, hipStream_t st) {
  const int64_t n = a.read_end - a.read0;
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(post_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace bvp
