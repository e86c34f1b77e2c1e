// Banded PairHMM forward for gfx950: one pair per wavefront, one diagonal per
// lane, one read row per step.
//
// The recurrence is the reference's (avx-pairhmm-template.h:183-198, the same
// contract as the full-matrix kernel), restricted to the 64 diagonals
// d = j - r in [o_lo, o_lo + 63]; cells outside the band are zero.  With lane
// k holding diagonal o_lo + k, the three dependencies of row r are
//   M[r][j] <- cell (r-1, j-1): the same lane, previous step
//   I[r][j] <- cell (r-1, j):   lane k+1, previous step (DPP wave_shl:1)
//   D[r][j] <- cell (r, j-1):   lane k-1, this step
// The D dependency is a first-order linear recurrence across lanes,
// D_k = my*M_{k-1} + yy*D_{k-1}, solved as an inclusive scan with the decay
// folded in: four row_shr levels (1, 2, 4, 8) inside each 16-lane DPP row,
// then row_bcast:15 and row_bcast:31 with per-lane powers of yy.  This needs a
// constant gap-continuation penalty along the read (yy), which the dispatch
// checks; other reads take the full matrix.
//
// Acceptance.  Alongside the forward sums the kernel carries a second DP over
// the same band, T, whose cells hold the mass of in-band paths that have
// visited the band's edge zones (kBandEdge diagonals at each side, on rows
// where the band actually cuts the matrix).  T obeys the forward recurrence,
// except that on edge-zone cells it equals the forward value.  T(final)/S is
// the fraction of the banded likelihood carried by paths that ran along a
// cut; a band is accepted when it is at most `tol`, otherwise the pair is
// queued for the full-matrix pass.  Every path that leaves the band must first
// cross an edge zone, so little mass there means little is lost outside.
//
// Per-row parameters are computed 64 rows at a time (lane k: row i0+k) into an
// LDS table and read back with uniform-address loads; the haplotype's
// per-column class masks are staged in LDS once per pair.
#include <hip/hip_runtime.h>

#include <climits>

#include "pairhmm_band.h"

namespace bvp {

namespace {

constexpr float kK = 0x1p120f;  // Context.h:109 (FP32 initial constant)
constexpr double kLog10K = 120.0 * 0.30102999566398119521;

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, true));
}

constexpr int kWaveShl1 = 0x130;  // lane k <- lane k+1, lane 63 <- 0
constexpr int kWaveShr1 = 0x138;  // lane k <- lane k-1, lane 0 <- 0
constexpr int kRowShr = 0x110;    // + n: lane k <- lane k-n inside a 16-lane row
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;

struct alignas(16) RowRec {
  float pm;     // 1 - e (match prior)
  float pe;     // e / 3 (mismatch prior; pm for a wildcard read base)
  float mm;     // match -> match into this row
  float g;      // insert/delete -> match into this row (1 - ph2pr[gcp])
  float mx;     // match -> insert
  float xx;     // insert -> insert
  float my;     // match -> delete (along this row)
  uint32_t rmask;  // bit of this row's read class in the haplotype masks
};

struct BandLds {
  RowRec tab[kBandLanes];
  uint8_t hap[kBandHapMax + 8];  // hap[j] = class mask of column j for 1 <= j <= H; 0 at j = 0, H+1
};

// Inclusive decayed scan across the wavefront: returns y with
// y_k = sum_{m>=0} yy^m x_{k-m}.  f15 is yy^(k%16+1) on DPP rows 1 and 3 and
// 0 on rows 0 and 2; f31 is yy^(k-31) on rows 2 and 3 and 0 below (the two
// broadcasts then add exactly the lower rows' prefixes).
__device__ __forceinline__ float decay_scan(float x, float yy1, float yy2, float yy4, float yy8,
                                            float f15, float f31) {
  x = __builtin_fmaf(dpp<kRowShr + 1>(x), yy1, x);
  x = __builtin_fmaf(dpp<kRowShr + 2>(x), yy2, x);
  x = __builtin_fmaf(dpp<kRowShr + 4>(x), yy4, x);
  x = __builtin_fmaf(dpp<kRowShr + 8>(x), yy8, x);
  x = __builtin_fmaf(dpp<kRowBcast15>(x), f15, x);
  x = __builtin_fmaf(dpp<kRowBcast31>(x), f31, x);
  return x;
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Wave-wide exclusive prefix sum of v; total in `tot`.
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int& tot) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  tot = __shfl(x, 63);
  return x - v;
}

__device__ __forceinline__ bool band_eligible(const BArgs& a, const Task& tk, int h, bool gcp_const,
                                              int2& b) {
  b = a.band[tk.out0 + h];
  const int hap = tk.hap0 + h;
  const int64_t H = a.hap_off[hap + 1] - a.hap_off[hap];
  return gcp_const && b.x <= b.y && static_cast<int64_t>(b.y) - b.x + 1 <= kBandLanes &&
         H <= kBandHapMax && b.x >= -(1 << 26) && b.y <= (1 << 26);
}

}  // namespace

// Splits every task's pairs into banded items and full-matrix tasks.  One
// thread per task; appends are reserved once per wavefront.
__global__ __launch_bounds__(256) void band_dispatch_kernel(BArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  Task tk{};
  bool gcp_const = false;
  int nb = 0, nf = 0;
  if (t < a.n_tasks) {
    tk = a.tasks[t];
    const int64_t ro = a.read_off[tk.read];
    const int64_t R = a.read_off[tk.read + 1] - ro;
    const uint8_t c0 = a.gcp[ro] & 127;
    gcp_const = true;
    for (int64_t k = 1; k < R; ++k) {
      if ((a.gcp[ro + k] & 127) != c0) {
        gcp_const = false;
        break;
      }
    }
    for (int h = 0; h < tk.nhap; ++h) {
      int2 b;
      if (band_eligible(a, tk, h, gcp_const, b)) ++nb; else ++nf;
    }
  }
  int tb = 0, tf = 0;
  const int ob = wave_excl_scan(nb, lane, tb);
  const int of = wave_excl_scan(nf, lane, tf);
  int bb = 0, bf = 0;
  if (lane == 0) {
    if (tb) bb = atomicAdd(a.item_count, tb);
    if (tf) bf = atomicAdd(a.full_count, tf);
  }
  bb = __shfl(bb, 0) + ob;
  bf = __shfl(bf, 0) + of;
  if (t < a.n_tasks) {
    for (int h = 0; h < tk.nhap; ++h) {
      int2 b;
      const int64_t o = tk.out0 + h;
      if (band_eligible(a, tk, h, gcp_const, b)) {
        const int w = b.y - b.x + 1;
        a.items[bb++] = BandItem{tk.read, tk.hap0 + h, b.x - ((kBandLanes - w) >> 1), 0, o};
      } else {
        a.full[bf++] = Task{tk.read, tk.hap0 + h, 1, -1, o};
      }
    }
  }
}

__global__ __launch_bounds__(64) void band_forward_kernel(BArgs a) {
  __shared__ BandLds s;
  const int lane = threadIdx.x;
  const int n_items = *a.item_count;
  const bool left_lane = lane < kBandEdge;
  const bool right_lane = lane >= kBandLanes - kBandEdge;
  const int row16 = lane >> 4;
  int accepted = 0;
  unsigned long long cells = 0;
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(a.counter, 1);
    it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
    if (it >= n_items) break;
    const BandItem w = a.items[it];
    const int64_t ro = a.read_off[w.read];
    const int R = static_cast<int>(a.read_off[w.read + 1] - ro);
    const int64_t ho = a.hap_off[w.hap];
    const int H = static_cast<int>(a.hap_off[w.hap + 1] - ho);
    const int o_lo = w.o_lo;
    const int d = o_lo + lane;  // this lane's diagonal

    __syncthreads();  // the previous pair's LDS reads are done
    for (int c = lane; c < H; c += kBandLanes) s.hap[1 + c] = a.hmask[ho + c];
    if (lane == 0) {
      s.hap[0] = 0;
      s.hap[1 + H] = 0;
    }

    // Decay factors of the D scan (constant gap continuation along the read).
    const float yy = a.ph2pr[a.gcp[ro] & 127];
    const float yy2 = yy * yy, yy4 = yy2 * yy2, yy8 = yy4 * yy4;
    float f15 = (row16 & 1) ? 1.f : 0.f, f31 = (row16 >= 2) ? 1.f : 0.f;
    for (int m = 0; m <= (lane & 15); ++m) f15 *= yy;  // distance to lane 15 of the row below
    for (int m = 31; m < lane; ++m) f31 *= yy;        // distance to lane 31
    // Edge zones: the left one from the first row whose left neighbour column
    // is inside the matrix, the right one while the right neighbour column is.
    const int ez_from = left_lane ? 2 - o_lo : INT_MAX;
    const int ez_to = right_lane ? H - o_lo - kBandLanes : INT_MIN;

    // Row 0: M = I = 0, D = K/H for columns 0..H (Context.h:109, template :164-175).
    float Mp = 0.f, Ip = 0.f;
    float Dp = (d >= 0 && d <= H) ? kK / static_cast<float>(H) : 0.f;
    float TMp = 0.f, TIp = 0.f, TDp = 0.f;
    {
      const bool ez = (left_lane && o_lo - 1 >= 0) || (right_lane && o_lo + kBandLanes <= H);
      if (ez) TDp = Dp;
    }

    for (int i0 = 1; i0 <= R; i0 += kBandLanes) {
      __syncthreads();
      {
        const int r = i0 + lane;
        if (r <= R) {
          const int64_t p = ro + r - 1;
          const uint8_t rc = a.rclass[p];
          const int iq = a.ins[p] & 127, dq = a.del[p] & 127, cq = a.gcp[p] & 127;
          const int lo = iq < dq ? iq : dq, hi = iq < dq ? dq : iq;
          const float e = a.ph2pr[a.q[p] & 127];
          RowRec rr;
          rr.pm = 1.f - e;
          rr.pe = (rc & kWildcard) ? rr.pm : e / 3.f;
          rr.mm = a.mm[((hi * (hi + 1)) >> 1) + lo];
          rr.g = 1.f - a.ph2pr[cq];
          rr.mx = a.ph2pr[iq];
          rr.xx = a.ph2pr[cq];
          rr.my = a.ph2pr[dq];
          rr.rmask = 1u << (rc & 7);
          s.tab[lane] = rr;
        }
      }
      __syncthreads();
      const int n = min(kBandLanes, R - i0 + 1);
      // Software pipeline: the LDS reads for row u+1 are issued before row u
      // is computed.
      RowRec rr_next = s.tab[0];
      uint32_t hm_next = s.hap[min(max(i0 + d, 0), H + 1)];
      for (int u = 0; u < n; ++u) {
        const int i = i0 + u;
        const RowRec rr = rr_next;
        const uint32_t hm = hm_next;
        rr_next = s.tab[(u + 1) & (kBandLanes - 1)];
        hm_next = s.hap[min(max(i + 1 + d, 0), H + 1)];
        const float prior = (hm & rr.rmask) ? rr.pm : rr.pe;
        const bool ez = i >= ez_from || i <= ez_to;

        // forward
        float t = Ip * rr.g;
        t = __builtin_fmaf(Dp, rr.g, t);
        t = __builtin_fmaf(Mp, rr.mm, t);
        const float Mn = t * prior;
        const float In = __builtin_fmaf(dpp<kWaveShl1>(Mp), rr.mx, dpp<kWaveShl1>(Ip) * rr.xx);
        const float Dn = decay_scan(dpp<kWaveShr1>(Mn) * rr.my, yy, yy2, yy4, yy8, f15, f31);

        // touch mass
        float tt = TIp * rr.g;
        tt = __builtin_fmaf(TDp, rr.g, tt);
        tt = __builtin_fmaf(TMp, rr.mm, tt);
        float TMn = tt * prior;
        float TIn = __builtin_fmaf(dpp<kWaveShl1>(TMp), rr.mx, dpp<kWaveShl1>(TIp) * rr.xx);
        if (ez) {
          TMn = Mn;
          TIn = In;
        }
        float TDn = decay_scan(dpp<kWaveShr1>(TMn) * rr.my, yy, yy2, yy4, yy8, f15, f31);
        if (ez) TDn = Dn;

        Mp = Mn;
        Ip = In;
        Dp = Dn;
        TMp = TMn;
        TIp = TIn;
        TDp = TDn;
      }
    }
    // S = sum over the last row of M + I (template :328-343), columns 1..H.
    const int j = R + d;
    const bool in = j >= 1 && j <= H;
    const double S = wsum(in ? static_cast<double>(Mp) + static_cast<double>(Ip) : 0.0);
    const double Tm = wsum(in ? static_cast<double>(TMp) + static_cast<double>(TIp) : 0.0);
    cells += static_cast<unsigned long long>(R) * kBandLanes;
    if (lane == 0) {
      if (S >= static_cast<double>(kMinAccepted) && Tm <= static_cast<double>(a.tol) * S) {
        a.out[w.out] = log10(S) - kLog10K;
        a.form[w.out] = kFormBandFP32;
        ++accepted;
      } else {
        const int k = atomicAdd(a.full_count, 1);
        a.full[k] = Task{w.read, w.hap, 1, -2, w.out};
      }
    }
  }
  if (lane == 0) {
    if (accepted) atomicAdd(a.accepted, accepted);
    if (cells && a.cells_done) atomicAdd(a.cells_done, cells);
  }
}

hipError_t launch_band_dispatch(const BArgs& a, hipStream_t st) {
  const int blocks = (a.n_tasks + 255) / 256;
  if (blocks > 0) hipLaunchKernelGGL(band_dispatch_kernel, dim3(blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_band_forward(const BArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(band_forward_kernel, dim3(grid), dim3(kBandLanes), 0, st, a);
  return hipGetLastError();
}

int band_blocks_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, band_forward_kernel, kBandLanes, 0) != hipSuccess)
    return 8;
  return n > 0 ? n : 1;
}

}  // namespace bvp
