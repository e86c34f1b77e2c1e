// Band prediction on the device: per-pair seed features and the band MLP.
//
// Features (36 per pair): R/256, H/512, mean and min base quality / 40, and a
// 32-bin histogram of 12-mer seed-hit diagonal offsets (o = q - p for a read
// seed at p found at haplotype position q), 4 bp per bin, centred on the
// median hit offset ("anchor"), each bin divided by the number of seeds.
// The MLP (36 -> 64 -> 64 -> 2, ReLU, FP32) maps features to the band centre
// relative to the anchor and the log2 band width; the band is
// [anchor + dc - hw, anchor + dc + hw].
//
// Training kernels (forward, backward, Adam) are in this file too; the
// trainer in biovoltron_amd/band.py drives them through the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "band_common.h"
#include "pairhmm_band.h"
#include "pairhmm_common.h"

namespace bvp {

constexpr int kK = kSeedK;
constexpr unsigned long long kEmpty = ~0ull;
constexpr uint32_t kNoKey = ~0u;
constexpr int kMaxOcc = 4;    // occurrences of one seed used for the histogram
constexpr int kHapChunk = 4;  // haplotypes whose lookups are in flight together
constexpr int kXcds = 8;      // MI355X accelerator complex dies
constexpr int kFeatClaim = 2; // tasks a features wavefront claims at once

__device__ __forceinline__ int code_of_mask(uint8_t m) {
  // single-class masks 1,2,4,8 -> 0..3; anything else (N, 0) -> -1
  switch (m) {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    default: return -1;
  }
}

__device__ __forceinline__ uint32_t slot_of(uint32_t key, uint32_t mask) {
  return (key * 2654435761u >> 8) & mask;
}

// Wave-wide minimum and maximum by DPP (no LDS round trips): row_shr 1, 2,
// 4, 8 inside each 16-lane row, then row_bcast:15 into rows 1 and 3 and
// row_bcast:31 into rows 2 and 3.  Lanes a move does not reach keep the
// identity (bound_ctrl off, `old` = identity).
constexpr int kDppRowShr = 0x110, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

template <typename Op>
__device__ __forceinline__ int wave_scan_dpp(int x, int id, Op op) {
  x = op(x, __builtin_amdgcn_update_dpp(id, x, kDppRowShr + 1, 0xf, 0xf, false));
  x = op(x, __builtin_amdgcn_update_dpp(id, x, kDppRowShr + 2, 0xf, 0xf, false));
  x = op(x, __builtin_amdgcn_update_dpp(id, x, kDppRowShr + 4, 0xf, 0xf, false));
  x = op(x, __builtin_amdgcn_update_dpp(id, x, kDppRowShr + 8, 0xf, 0xf, false));
  x = op(x, __builtin_amdgcn_update_dpp(id, x, kDppRowBcast15, 0xa, 0xf, false));
  x = op(x, __builtin_amdgcn_update_dpp(id, x, kDppRowBcast31, 0xc, 0xf, false));
  return x;  // inclusive scan; lane 63 holds the total
}

__device__ __forceinline__ int wave_min(int x) {
  return __builtin_amdgcn_readlane(wave_scan_dpp(x, INT_MAX, [](int u, int v) { return min(u, v); }), 63);
}
__device__ __forceinline__ int wave_max(int x) {
  return __builtin_amdgcn_readlane(wave_scan_dpp(x, INT_MIN, [](int u, int v) { return max(u, v); }), 63);
}

// Per haplotype, an open-addressing table of its 12-mers: entry = key |
// position << 32, empty = ~0; haplotype h owns slots [tab_off[h],
// tab_off[h+1]), a power of two at least twice its length.  One block per
// haplotype, built once per plan.
__global__ __launch_bounds__(256) void hap_table_kernel(const uint8_t* __restrict__ hmask,
                                                        const int64_t* __restrict__ hap_off,
                                                        const int64_t* __restrict__ tab_off,
                                                        int64_t n_haps,
                                                        unsigned long long* __restrict__ tab) {
  for (int64_t h = blockIdx.x; h < n_haps; h += gridDim.x) {
    const int64_t ho = hap_off[h];
    const int H = static_cast<int>(hap_off[h + 1] - ho);
    unsigned long long* t = tab + tab_off[h];
    const uint32_t mask = static_cast<uint32_t>(tab_off[h + 1] - tab_off[h] - 1);
    for (uint32_t i = threadIdx.x; i <= mask; i += blockDim.x) t[i] = kEmpty;
    __syncthreads();
    // a haplotype shorter than a seed gets no table (one slot)
    for (int q = threadIdx.x; mask > 0 && q + kK <= H; q += blockDim.x) {
      uint32_t key = 0;
      bool ok = true;
      for (int k = 0; ok && k < kK; ++k) {
        const int c = code_of_mask(hmask[ho + q + k]);
        ok = c >= 0;
        key = (key << 2) | static_cast<uint32_t>(c & 3);
      }
      if (!ok) continue;
      const unsigned long long e = static_cast<unsigned long long>(key) |
                                   (static_cast<unsigned long long>(q) << 32);
      uint32_t sl = slot_of(key, mask);
      for (uint32_t probe = 0; probe <= mask; ++probe) {
        if (atomicCAS(&t[sl], kEmpty, e) == kEmpty) break;
        sl = (sl + 1) & mask;
      }
    }
    __syncthreads();
  }
}

// ---- seed features: helpers shared by the features kernel and the fused
// band prologue ---------------------------------------------------------------

// The read side of a task: quality summary and this lane's seed (a 12-mer,
// <= 64 spread over the read).
struct ReadSeeds {
  int R;
  int nvalid;       // seeds without a wildcard base
  int my_p;         // this lane's seed start, -1 for none
  uint32_t my_key;  // its 2-bit code
  float fq_mean, fq_min;
};

__device__ __forceinline__ ReadSeeds read_seeds(const uint8_t* __restrict__ rclass, const uint8_t* __restrict__ qual,
                                                int64_t ro, int R, int lane) {
  ReadSeeds s;
  s.R = R;
  double qsum = 0.0;
  int qmin = 255;
  for (int p = lane; p < R; p += 64) {
    const int q = qual[ro + p] & 127;
    qsum += q;
    qmin = min(qmin, q);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    qsum += __shfl_xor(qsum, o);
    qmin = min(qmin, __shfl_xor(qmin, o));
  }
  s.fq_mean = static_cast<float>(qsum / R) / 40.f;
  s.fq_min = static_cast<float>(qmin) / 40.f;
  const int nseeds = R >= kK ? min(64, R - kK + 1) : 0;
  const int stride = nseeds > 0 ? max(1, (R - kK + 1 + 63) / 64) : 1;
  s.my_p = -1;
  s.my_key = 0;
  if (lane < nseeds) {
    s.my_p = lane * stride;
    if (s.my_p + kK > R) s.my_p = -1;
    if (s.my_p >= 0) {
      uint8_t c[kK];
#pragma unroll
      for (int k = 0; k < kK; ++k) c[k] = rclass[ro + s.my_p + k];
      uint8_t wild = 0;
#pragma unroll
      for (int k = 0; k < kK; ++k) {
        wild |= c[k];
        s.my_key = (s.my_key << 2) | (c[k] & 3);
      }
      if (wild & kWildcard) s.my_p = -1;
    }
  }
  s.nvalid = __popcll(__ballot(s.my_p >= 0));
  return s;
}

// This lane's seed looked up in the tables of haplotypes hap0 .. hap0 + n - 1
// (n <= kHapChunk).  Per haplotype: the first occurrence (smallest position)
// and the offsets q - p of the first kMaxOcc occurrences met along the probe
// sequence, in occ[c][0 .. nocc[c] - 1], newest first (the histogram counts
// them as a set, so a shift register serves: an insert is kMaxOcc moves,
// where indexing by the count took a compare and select per register).
struct ChunkHits {
  int first[kHapChunk];
  int occ[kHapChunk][kMaxOcc];
  int nocc[kHapChunk];
};

__device__ __forceinline__ void chunk_lookups(const int64_t* __restrict__ hap_off, const int64_t* __restrict__ htab_off,
                                              const unsigned long long* __restrict__ htab, int hap0, int n,
                                              int my_p, uint32_t my_key, ChunkHits& ch) {
#pragma unroll
  for (int c = 0; c < kHapChunk; ++c) {
    ch.first[c] = INT_MAX;
    ch.nocc[c] = 0;
#pragma unroll
    for (int k = 0; k < kMaxOcc; ++k) ch.occ[c][k] = 0;
    const int hap = hap0 + c;
    if (c >= n || my_p < 0) continue;
    if (hap_off[hap + 1] - hap_off[hap] > kBandMaxHap) continue;  // beyond the banded kernel
    const int64_t t0 = htab_off[hap], t1 = htab_off[hap + 1];
    if (t1 - t0 < 2) continue;  // no table
    const unsigned long long* t = htab + t0;
    const uint32_t mask = static_cast<uint32_t>(t1 - t0 - 1);
    uint32_t sl = slot_of(my_key, mask);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
      const unsigned long long e = t[sl];
      if (e == kEmpty) break;
      if (static_cast<uint32_t>(e) == my_key) {
        const int q = static_cast<int>(e >> 32);
        ch.first[c] = min(ch.first[c], q);
        if (ch.nocc[c] < kMaxOcc) {
#pragma unroll
          for (int k = kMaxOcc - 1; k > 0; --k) ch.occ[c][k] = ch.occ[c][k - 1];
          ch.occ[c][0] = q - my_p;
          ++ch.nocc[c];
        }
      }
      sl = (sl + 1) & mask;
    }
  }
}

// The features of one pair from its lookups: the anchor (median of the
// seeds' first-occurrence offsets; (H - R) / 2 without hits), this lane's
// histogram bin count (lane b < kBins counts bin b: 4 bp over
// [anchor - 64, anchor + 64), every kept occurrence), and the read span
// covered by seeds hitting within kEndAnchorTol of the anchor ((-1, -1)
// without such seeds).
struct PairFeat {
  int anchor;
  int cnt;
  int pf, pl;
};

__device__ __forceinline__ PairFeat pair_features(int first, const int* occ, int nocc, int my_p, int R, int H,
                                                  int lane, uint32_t* hist) {
  PairFeat f;
  f.anchor = (H - R) / 2;
  const int found_off = first != INT_MAX ? first - my_p : INT_MIN;
  // Median of the hit offsets as the anchor: walk the distinct offsets in
  // ascending order (wave minimum, then the lanes holding it by ballot)
  // until more than half of the hits are behind -- one or two steps as a
  // rule, at most one per hit.
  const bool hit = found_off != INT_MIN;
  const int nhit = __popcll(__ballot(hit));
  if (nhit > 0) {
    const int half = (nhit - 1) / 2;
    int rest = hit ? found_off : INT_MAX;
    int behind = 0;
    for (;;) {
      const int v = wave_min(rest);
      behind += __popcll(__ballot(rest == v));
      if (behind > half) {
        f.anchor = v;
        break;
      }
      if (rest == v) rest = INT_MAX;
    }
  }
  // Bin counts by LDS atomics on the wavefront's own 32 counters (its LDS
  // operations run in order: clear, adds, read).
  __builtin_amdgcn_wave_barrier();
  if (lane < kBins) hist[lane] = 0u;
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int k = 0; k < kMaxOcc; ++k) {
    const int b = (occ[k] - f.anchor + 64) >> 2;
    if (k < nocc && b >= 0 && b < kBins)
      __hip_atomic_fetch_add(hist + b, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
  __builtin_amdgcn_wave_barrier();
  const int my_cnt = lane < kBins ? static_cast<int>(hist[lane]) : 0;
  f.cnt = my_cnt;
  // read ends anchored by seeds that hit next to the anchor
  const bool near = hit && abs(found_off - f.anchor) <= kEndAnchorTol;
  f.pf = wave_min(near ? my_p : INT_MAX);
  f.pl = wave_max(near ? my_p + kK : -1);
  if (f.pl < 0) f.pf = f.pl = -1;
  return f;
}

// Workgroups go round-robin to the 8 XCDs, each with its own L2.  Tasks are
// in region order; each XCD takes a contiguous eighth of them, and its
// wavefronts claim the next kFeatClaim tasks of that eighth from a counter
// of their own XCD, so that the tasks in flight on an XCD stay a few
// regions wide and those regions' haplotype tables stay in its L2.  (A fixed
// stride over the eighth lets fast wavefronts run ahead of slow ones: the
// tasks in flight then span dozens of regions and nearly every table probe
// missed the L2.)
struct TaskClaims {
  int begin, end, ctr;
};

__device__ __forceinline__ TaskClaims xcd_task_range(int n_tasks) {
  TaskClaims c{0, n_tasks, 0};
  if (gridDim.x % kXcds == 0 && gridDim.x >= kXcds) {
    c.ctr = blockIdx.x % kXcds;
    const int chunk = (n_tasks + kXcds - 1) / kXcds;
    c.begin = min(n_tasks, c.ctr * chunk);
    c.end = min(n_tasks, c.begin + chunk);
  }
  return c;
}

// The first task of the next claim of this wavefront (wave-uniform), or -1.
__device__ __forceinline__ int claim_tasks(int* next, const TaskClaims& c, int lane) {
  int c0 = 0;
  if (lane == 0) c0 = atomicAdd(next + c.ctr * kCtrStride, kFeatClaim);
  c0 = __builtin_amdgcn_readfirstlane(__shfl(c0, 0));
  return c0 >= c.end - c.begin ? -1 : c.begin + c0;
}

// One wavefront per task (a read and its region's haplotypes): features of
// every pair into X, anchor and ends (the band API's features, and the
// trainer's inputs).
__global__ __launch_bounds__(64) void band_features_kernel(FeatArgs a) {
  __shared__ uint32_t hist[kBins];
  const int lane = threadIdx.x;
  const TaskClaims cl = xcd_task_range(a.n_tasks);
  for (;;) {
    const int tc0 = claim_tasks(a.next, cl, lane);
    if (tc0 < 0) break;
    const int tc1 = min(cl.end, tc0 + kFeatClaim);
    for (int ti = tc0; ti < tc1; ++ti) {
      const Task tk = a.tasks[ti];
      const int64_t ro = a.read_off[tk.read];
      const int R = static_cast<int>(a.read_off[tk.read + 1] - ro);
      const ReadSeeds rs = read_seeds(a.rclass, a.q, ro, R, lane);
      const float inv = rs.nvalid > 0 ? 1.f / static_cast<float>(rs.nvalid) : 0.f;
      for (int h0 = 0; h0 < tk.nhap; h0 += kHapChunk) {
        ChunkHits ch;
        chunk_lookups(a.hap_off, a.htab_off, a.htab, tk.hap0 + h0, min(kHapChunk, tk.nhap - h0), rs.my_p,
                      rs.my_key, ch);
#pragma unroll
        for (int c = 0; c < kHapChunk; ++c) {
          const int h = h0 + c;
          if (h >= tk.nhap) break;
          const int hap = tk.hap0 + h;
          const int64_t o = tk.out0 + h;
          const int H = static_cast<int>(a.hap_off[hap + 1] - a.hap_off[hap]);
          const PairFeat f = pair_features(ch.first[c], ch.occ[c], ch.nocc[c], rs.my_p, R, H, lane, hist);
          if (a.rec != nullptr) {
            uint8_t* rp = a.rec + o * kRecBytes;
            if (lane < kBins) rp[lane] = static_cast<uint8_t>(f.cnt);
            if (lane == 0)
              *reinterpret_cast<int2*>(rp + kBins) =
                  int2{f.anchor, static_cast<int>((static_cast<uint32_t>(f.pf) & 0xffffu) |
                                                  (static_cast<uint32_t>(f.pl) << 16))};
            continue;
          }
          float* X = a.X + o * kFeat;
          if (lane == 0) {
            X[0] = static_cast<float>(R) / 256.f;
            X[1] = static_cast<float>(H) / 512.f;
            X[2] = rs.fq_mean;
            X[3] = rs.fq_min;
            a.anchor[o] = f.anchor;
            a.ends[o] = int2{f.pf, f.pl};
          }
          if (lane < kBins) X[4 + lane] = static_cast<float>(f.cnt) * inv;
        }
      }
      if (a.trec != nullptr && lane == 0) {
        a.trec[3 * ti] = inv;
        a.trec[3 * ti + 1] = rs.fq_mean;
        a.trec[3 * ti + 2] = rs.fq_min;
      }
    }
  }
}

// ---- MLP --------------------------------------------------------------------

struct MlpParams {
  const float* W1;  // [kHid][kFeat]
  const float* b1;  // [kHid]
  const float* W2;  // [kHid][kHid]
  const float* b2;
  const float* W3;  // [2][kHid]
  const float* b3;  // [2]
};

__device__ __forceinline__ MlpParams split_params(const float* p) {
  MlpParams m;
  m.W1 = p;
  m.b1 = m.W1 + kHid * kFeat;
  m.W2 = m.b1 + kHid;
  m.b2 = m.W2 + kHid * kHid;
  m.W3 = m.b2 + kHid;
  m.b3 = m.W3 + 2 * kHid;
  return m;
}

// Forward for one sample per thread; weights staged in LDS.  Optionally keeps
// the hidden activations for the backward pass.
__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* __restrict__ params,
                                                          const float* __restrict__ X, int64_t n,
                                                          float* __restrict__ Y,
                                                          float* __restrict__ H1,
                                                          float* __restrict__ H2) {
  __shared__ float w[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) w[i] = params[i];
  __syncthreads();
  const MlpParams m = split_params(w);
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    float x[kFeat];
#pragma unroll
    for (int k = 0; k < kFeat; ++k) x[k] = X[i * kFeat + k];
    float h1[kHid];
#pragma unroll
    for (int u = 0; u < kHid; ++u) {
      float acc = m.b1[u];
#pragma unroll
      for (int k = 0; k < kFeat; ++k) acc = __builtin_fmaf(m.W1[u * kFeat + k], x[k], acc);
      h1[u] = acc > 0.f ? acc : 0.f;
    }
    float h2[kHid];
#pragma unroll
    for (int u = 0; u < kHid; ++u) {
      float acc = m.b2[u];
#pragma unroll
      for (int k = 0; k < kHid; ++k) acc = __builtin_fmaf(m.W2[u * kHid + k], h1[k], acc);
      h2[u] = acc > 0.f ? acc : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      float acc = m.b3[u];
#pragma unroll
      for (int k = 0; k < kHid; ++k) acc = __builtin_fmaf(m.W3[u * kHid + k], h2[k], acc);
      Y[i * 2 + u] = acc;
    }
    if (H1 != nullptr) {
#pragma unroll
      for (int u = 0; u < kHid; ++u) H1[i * kHid + u] = h1[u];
#pragma unroll
      for (int u = 0; u < kHid; ++u) H2[i * kHid + u] = h2[u];
    }
  }
}

// Backward step 1: per-sample deltas.  Loss = mean over samples of
// 0.5 * sum_u wt_u * (Y_u - T_u)^2.  D3 = dL/dY, D2 = dL/d(pre-activation 2),
// D1 = dL/d(pre-activation 1).
__global__ __launch_bounds__(256) void mlp_delta_kernel(const float* __restrict__ params,
                                                        const float* __restrict__ Y,
                                                        const float* __restrict__ T, int64_t n,
                                                        const float* __restrict__ H1,
                                                        const float* __restrict__ H2,
                                                        float* __restrict__ D1,
                                                        float* __restrict__ D2,
                                                        float* __restrict__ D3, float w0, float w1,
                                                        float* __restrict__ loss) {
  __shared__ float w[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) w[i] = params[i];
  __syncthreads();
  const MlpParams m = split_params(w);
  const float invn = 1.f / static_cast<float>(n);
  float lsum = 0.f;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float e0 = Y[i * 2] - T[i * 2], e1 = Y[i * 2 + 1] - T[i * 2 + 1];
    lsum += 0.5f * (w0 * e0 * e0 + w1 * e1 * e1);
    const float d3[2] = {w0 * e0 * invn, w1 * e1 * invn};
    D3[i * 2] = d3[0];
    D3[i * 2 + 1] = d3[1];
    float d2[kHid];
#pragma unroll
    for (int k = 0; k < kHid; ++k) {
      const float g = m.W3[k] * d3[0] + m.W3[kHid + k] * d3[1];
      d2[k] = H2[i * kHid + k] > 0.f ? g : 0.f;
      D2[i * kHid + k] = d2[k];
    }
#pragma unroll 4
    for (int k = 0; k < kHid; ++k) {
      float g = 0.f;
#pragma unroll
      for (int u = 0; u < kHid; ++u) g = __builtin_fmaf(m.W2[u * kHid + k], d2[u], g);
      D1[i * kHid + k] = H1[i * kHid + k] > 0.f ? g : 0.f;
    }
  }
  if (loss != nullptr) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, lsum * invn);
  }
}

// Backward step 2: one thread per parameter, summing over the batch.
__global__ __launch_bounds__(256) void mlp_grad_kernel(const float* __restrict__ X, int64_t n,
                                                       const float* __restrict__ H1,
                                                       const float* __restrict__ H2,
                                                       const float* __restrict__ D1,
                                                       const float* __restrict__ D2,
                                                       const float* __restrict__ D3,
                                                       float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  // locate parameter p
  const float* src;  // input activations
  const float* dst;  // deltas
  int in_dim, out_dim, u, k;
  bool bias = false;
  int q = p;
  if (q < kHid * kFeat) {
    in_dim = kFeat; out_dim = kHid; u = q / kFeat; k = q % kFeat; src = X; dst = D1;
  } else if ((q -= kHid * kFeat) < kHid) {
    in_dim = kFeat; out_dim = kHid; u = q; k = 0; src = X; dst = D1; bias = true;
  } else if ((q -= kHid) < kHid * kHid) {
    in_dim = kHid; out_dim = kHid; u = q / kHid; k = q % kHid; src = H1; dst = D2;
  } else if ((q -= kHid * kHid) < kHid) {
    in_dim = kHid; out_dim = kHid; u = q; k = 0; src = H1; dst = D2; bias = true;
  } else if ((q -= kHid) < 2 * kHid) {
    in_dim = kHid; out_dim = 2; u = q / kHid; k = q % kHid; src = H2; dst = D3;
  } else {
    q -= 2 * kHid;
    in_dim = kHid; out_dim = 2; u = q; k = 0; src = H2; dst = D3; bias = true;
  }
  double g = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const float d = dst[i * out_dim + u];
    g += bias ? d : d * src[i * in_dim + k];
  }
  grad[p] = static_cast<float>(g);
}

__global__ void adam_kernel(float* __restrict__ params, const float* __restrict__ grad,
                            float* __restrict__ m1, float* __restrict__ m2, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  const float g = grad[p];
  m1[p] = b1 * m1[p] + (1.f - b1) * g;
  m2[p] = b2 * m2[p] + (1.f - b2) * g * g;
  const float mh = m1[p] / c1, vh = m2[p] / c2;
  params[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Bands from predictions.  The network predicts the posterior span's centre,
// anchor + 16 * Y0, and its log2 width, Y1; with `margin` diagonals added on
// each side that is the predicted span.  It gets:
//  - 32 diagonals centred on the anchor when the span fits between the narrow
//    band's edge zones, the centre is within 4 of the anchor, every seed hit
//    within 64 of the anchor lies inside those zones, and seeds that hit next
//    to the anchor reach within kEndSlack bases of both read ends (a read end
//    that matches nowhere is emitted as a run of insertions whose diagonals
//    sweep up to its length away from the alignment, out of sight of the
//    kernel's edge-zone test);
//  - otherwise 64 diagonals centred on the predicted centre when the span is
//    at most `max_width` wide, every seed hit lies inside the wide band's edge
//    zones, and the anchoring seeds cover at least half the read and reach
//    within kEndSlackWide bases of both read ends (a lone spurious 12-mer hit
//    can put the anchor far from the alignment);
//  - otherwise the full matrix, as for pairs without seed hits (no anchor) and
//    haplotypes longer than the banded kernel stages.
// Seed hits count every occurrence in the haplotype, so a repeat copy the
// read could align to shows up as a hit at its offset.
__device__ __forceinline__ bool hits_inside(const float* x, int lo, int hi) {
#pragma unroll
  for (int b = 0; b < kBins; ++b) {
    const int blo = -64 + 4 * b;
    if (x[4 + b] > 0.f && (blo < lo || blo + 3 > hi)) return false;
  }
  return true;
}

__global__ void band_from_prediction_kernel(const float* __restrict__ Y, const int* __restrict__ anchor,
                                            const int2* __restrict__ ends, const float* __restrict__ X,
                                            int64_t n, int margin, int max_width, int allow_narrow,
                                            int2* __restrict__ band) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const float* x = X + i * kFeat;
  const int R = static_cast<int>(x[0] * 256.f + 0.5f);
  const int H = static_cast<int>(x[1] * 512.f + 0.5f);
  float hits = 0.f;
#pragma unroll
  for (int b = 0; b < kBins; ++b) hits += x[4 + b];
  const float lw = fminf(fmaxf(Y[i * 2 + 1], 0.f), 12.f);
  const int width = static_cast<int>(ceilf(exp2f(lw))) + 2 * margin;
  if (H > kBandHapMax || hits <= 0.f || width > max_width) {
    band[i] = int2{1, 0};
    return;
  }
  const int dc = static_cast<int>(lrintf(16.f * fminf(fmaxf(Y[i * 2], -8.f), 8.f)));
  const int2 e = ends[i];
  constexpr int kIn = kBandNarrow / 2 - kBandEdge;  // narrow interior: [-kIn, kIn - 1]
  const bool narrow = allow_narrow && width <= kBandNarrow - 2 * kBandEdge && abs(dc) <= 4 &&
                      hits_inside(x, -kIn, kIn - 1) && e.x >= 0 && e.x <= kEndSlack &&
                      e.y >= R - kEndSlack;
  constexpr int kInW = kBandLanes / 2 - kBandEdge;  // wide interior: [-kInW, kInW - 1]
  const bool wide = hits_inside(x, dc - kInW, dc + kInW - 1) && e.x >= 0 && e.x <= kEndSlackWide &&
                    e.y >= R - kEndSlackWide && 2 * (e.y - e.x) >= R;
  if (narrow) {
    band[i] = int2{anchor[i] - kBandNarrow / 2, anchor[i] + kBandNarrow / 2 - 1};
  } else if (wide) {
    const int c = anchor[i] + dc;
    band[i] = int2{c - kBandLanes / 2, c + kBandLanes / 2 - 1};
  } else {
    band[i] = int2{1, 0};
  }
}

// ---- fused band prologue: features, MLP, band rules and dispatch ------------
//
// One pass over the tasks (region order, claimed per XCD as above) that
// writes the band table and the banded and full-matrix work lists directly:
// the 36 features, the prediction and the rule inputs of a pair live only in
// LDS.  Each wavefront buffers the pairs of whole tasks (up to 64 pairs; a
// task with more haplotypes is taken 64 at a time), then runs the MLP with one
// pair per lane -- the same FMA sequence per output as mlp_forward_kernel,
// skipping only histogram inputs that are zero for every buffered pair
// (adding w * 0 leaves a sum unchanged, and ReLU maps either zero to +0), so
// predictions and bands are bit for bit those of the separate kernels -- and
// applies the rules of band_from_prediction_kernel and the list split of
// band_dispatch_kernel (a task's wide pairs in groups of four for the quad
// shape, the rest to the one-pair list; narrow pairs padded to an even count
// per task) with one reservation per list per buffer.
constexpr int kProWaves = 8;  // wavefronts per workgroup, sharing the weights in LDS

struct ProlArgs {
  FeatArgs f;
  BArgs b;           // lists and counters written; b.band unused
  int2* band;        // [pairs]: the band table (short-read pairs)
  const float* params;
  int margin, max_width, allow_narrow;
};

// Per-wavefront buffer of up to 64 pairs (LDS).  A bin count fits a byte: a
// seed adds at most kMaxOcc occurrences to one 4-bp bin only if its 12-mer
// occurs 3 or 4 times within 4 bases (a homopolymer: 4 such keys, and the
// seeds sharing one share its kept occurrences, so with distinct seed
// positions each key adds at most 16), and at most 2 otherwise (period 2 or
// 3), so a bin holds at most 4 * 16 + 2 * 64 = 192.
struct ProBuf {
  uint32_t cnt[64][kBins / 4];  // histogram counts, one byte per bin
  int hap[64], H[64], anchor[64], pf[64], pl[64], read[64], R[64];
  float inv[64], fqm[64], fqn[64];
  long long out[64];
  int gc[64];  // the read's gap continuation is constant
  uint32_t hist[kBins];  // bin counters of the pair being featurised
};

__device__ __forceinline__ unsigned long long below(int n) {
  return n >= 64 ? ~0ull : ((1ull << n) - 1ull);
}

// MLP, rules and dispatch for the `fill` buffered pairs; seg_first marks the
// first pair of each task (or 64-haplotype piece of one).
__device__ __forceinline__ void prologue_flush(const ProlArgs& A, const float* __restrict__ w, ProBuf& B, int fill,
                                               unsigned long long seg_first, int lane) {
  const BArgs& a = A.b;
  const bool valid = lane < fill;
  const int s = valid ? lane : 0;
  uint32_t cw[kBins / 4];
#pragma unroll
  for (int k = 0; k < kBins / 4; ++k) cw[k] = B.cnt[s][k];
  const int R = B.R[s], H = B.H[s];
  const float inv = B.inv[s];
  // layer 1 (weights transposed in LDS: w1t[k][u]), input-major so that a
  // histogram bin that is zero for every pair is skipped
  const float* w1t = w;
  const float* bias1 = w1t + kFeat * kHid;
  const float* W2 = bias1 + kHid;
  const float* b2 = W2 + kHid * kHid;
  const float* W3 = b2 + kHid;
  const float* b3 = W3 + 2 * kHid;
  float h1[kHid];
#pragma unroll
  for (int u = 0; u < kHid; ++u) h1[u] = bias1[u];
  const float x0[4] = {static_cast<float>(R) / 256.f, static_cast<float>(H) / 512.f, B.fqm[s], B.fqn[s]};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
#pragma unroll
    for (int u = 0; u < kHid; ++u) h1[u] = __builtin_fmaf(w1t[k * kHid + u], x0[k], h1[u]);
  }
#pragma unroll
  for (int b = 0; b < kBins; ++b) {
    const uint32_t cb = (cw[b >> 2] >> (8 * (b & 3))) & 0xffu;
    if (__ballot(cb != 0) == 0) continue;
    const float x = static_cast<float>(cb) * inv;
#pragma unroll
    for (int u = 0; u < kHid; ++u) h1[u] = __builtin_fmaf(w1t[(4 + b) * kHid + u], x, h1[u]);
  }
#pragma unroll
  for (int u = 0; u < kHid; ++u) h1[u] = h1[u] > 0.f ? h1[u] : 0.f;
  // layers 2 and 3: output o sums W3[o][u] * h2[u] over u in order, so each
  // h2[u] is folded in as it is formed
  float y0 = b3[0], y1 = b3[1];
  for (int u = 0; u < kHid; ++u) {
    float acc = b2[u];
#pragma unroll
    for (int k = 0; k < kHid; ++k) acc = __builtin_fmaf(W2[u * kHid + k], h1[k], acc);
    const float h2 = acc > 0.f ? acc : 0.f;
    y0 = __builtin_fmaf(W3[u], h2, y0);
    y1 = __builtin_fmaf(W3[kHid + u], h2, y1);
  }
  // rules (band_from_prediction_kernel)
  bool hits = false;
  int2 band{1, 0};
  {
    const int anchor = B.anchor[s];
    const int pf = B.pf[s], pl = B.pl[s];
    auto inside = [&](int lo, int hi) {
      bool ok = true;
#pragma unroll
      for (int b = 0; b < kBins; ++b) {
        const uint32_t cb = (cw[b >> 2] >> (8 * (b & 3))) & 0xffu;
        const int blo = -64 + 4 * b;
        if (cb != 0 && (blo < lo || blo + 3 > hi)) ok = false;
      }
      return ok;
    };
#pragma unroll
    for (int k = 0; k < kBins / 4; ++k) hits |= cw[k] != 0;
    hits = hits && inv > 0.f;
    const float lw = fminf(fmaxf(y1, 0.f), 12.f);
    const int width = static_cast<int>(ceilf(exp2f(lw))) + 2 * A.margin;
    if (!(H > kBandHapMax || !hits || width > A.max_width)) {
      const int dc = static_cast<int>(lrintf(16.f * fminf(fmaxf(y0, -8.f), 8.f)));
      constexpr int kIn = kBandNarrow / 2 - kBandEdge;
      const bool narrow = A.allow_narrow && width <= kBandNarrow - 2 * kBandEdge && abs(dc) <= 4 &&
                          inside(-kIn, kIn - 1) && pf >= 0 && pf <= kEndSlack && pl >= R - kEndSlack;
      constexpr int kInW = kBandLanes / 2 - kBandEdge;
      const bool wide = inside(dc - kInW, dc + kInW - 1) && pf >= 0 && pf <= kEndSlackWide &&
                        pl >= R - kEndSlackWide && 2 * (pl - pf) >= R;
      if (narrow) {
        band = int2{anchor - kBandNarrow / 2, anchor + kBandNarrow / 2 - 1};
      } else if (wide) {
        const int c = anchor + dc;
        band = int2{c - kBandLanes / 2, c + kBandLanes / 2 - 1};
      }
    }
  }
  const long long out = B.out[s];
  const int read = B.read[s], hap = B.hap[s];
  if (valid) A.band[out] = band;
  // band class (band_dispatch_kernel)
  int cls = 0;
  if (valid && B.gc[s] && band.x <= band.y && H <= kBandHapMax && R <= kBandReadMax && band.x >= -(1 << 26) &&
      band.y <= (1 << 26)) {
    const int wd = band.y - band.x + 1;
    cls = wd <= kBandNarrow ? kBandNarrow : (wd <= kBandLanes ? kBandLanes : 0);
  }
  // this pair's task piece: lanes [start, end)
  const unsigned long long lt = below(lane);
  const unsigned long long upto = seg_first & below(lane + 1);
  const int start = 63 - __clzll(upto);
  const unsigned long long above = seg_first & ~below(lane + 1);
  const int end = above ? __ffsll(static_cast<long long>(above)) - 1 : fill;
  const unsigned long long seg = below(end) & ~below(start);
  const bool quad = a.wide_pairs == kQuadPairs;
  const unsigned long long w64 = __ballot(cls == kBandLanes);
  const int n64 = __popcll(w64 & seg), r64 = __popcll(w64 & seg & lt);
  // quad slots of the piece (band_dispatch_kernel's rule): whole units, and
  // one more, padded, when at least quad_fill wide pairs are left over
  const int rem64 = n64 & (kQuadPairs - 1);
  const int n64q = !quad ? n64 : (rem64 >= a.quad_fill ? n64 + kQuadPairs - rem64 : n64 - rem64);
  const bool to_q = cls == kBandLanes && r64 < n64q;
  const bool to_1 = cls == kBandLanes && !to_q;
  const unsigned long long w32 = __ballot(cls == kBandNarrow);
  const bool first = valid && ((seg_first >> lane) & 1ull);
  // the piece's padding slots (0-3), written by its first lane after its
  // quad items
  const int qpad = first ? max(0, n64q - n64) : 0;
  const unsigned long long mp1 = __ballot(qpad & 1), mp2 = __ballot(qpad & 2);
  const bool odd = first && (__popcll(w32 & seg) & 1);
  const bool to_f = valid && cls == 0 && H <= kStackHmax;
  const bool to_fl = valid && cls == 0 && H > kStackHmax;
  const unsigned long long mq = __ballot(to_q), m1 = __ballot(to_1), mo = __ballot(odd), mf = __ballot(to_f),
                           mfl = __ballot(to_fl);
  int bq = 0, b1 = 0, b32 = 0, bf = 0, bfl = 0;
  if (lane == 0) {
    if (mq) bq = atomicAdd(a.item_count, __popcll(mq) + __popcll(mp1) + 2 * __popcll(mp2));
    if (m1) b1 = atomicAdd(a.item1_count, __popcll(m1));
    if (w32) b32 = atomicAdd(a.item32_count, __popcll(w32) + __popcll(mo));
    if (mf) bf = atomicAdd(a.full_count, __popcll(mf));
    if (mfl) bfl = atomicAdd(a.full_long_count, __popcll(mfl));
  }
  bq = __shfl(bq, 0);
  b1 = __shfl(b1, 0);
  b32 = __shfl(b32, 0);
  bf = __shfl(bf, 0);
  bfl = __shfl(bfl, 0);
  const int pads_before = __popcll(mo & below(start));
  const int qpads_before = __popcll(mp1 & below(start)) + 2 * __popcll(mp2 & below(start));
  if (cls) {
    const int wd = band.y - band.x + 1;
    const BandItem it{read, hap, band.x - ((cls - wd) >> 1), 0, out};
    if (to_q) a.items[bq + __popcll(mq & lt) + qpads_before] = it;
    else if (to_1) a.items1[b1 + __popcll(m1 & lt)] = it;
    else a.items32[b32 + __popcll(w32 & lt) + pads_before] = it;
  } else if (to_f) {
    a.full[bf + __popcll(mf & lt)] = Task{read, hap, 1, -1, out};
  } else if (to_fl) {
    a.full_long[bfl + __popcll(mfl & lt)] = Task{read, hap, 1, -1, out};
  }
  if (odd) a.items32[b32 + __popcll(w32 & below(end)) + pads_before] = BandItem{read, -1, 0, 0, -1};
  for (int z = 0; z < qpad; ++z)
    a.items[bq + __popcll(mq & below(end)) + qpads_before + z] = BandItem{read, -1, 0, 0, -1};
}

// The two-kernel prologue's buffer filling: the features come as compact
// records from band_features_kernel.  A wavefront claims kRecClaim tasks at
// once and lane j loads task j's fields, so the loads of a claim overlap; the
// slots of the buffer (slot = lane) are then assigned to (task, haplotype)
// in task order with the same pieces and flush points as the one-kernel
// prologue, and every assigned slot loads its record in one pass before a
// flush or the next claim.  Claims are chunks of kRecClaim tasks, XCD x
// taking chunks x, x + 8, x + 16, ... from a counter of its own: the XCDs
// then move through the tasks (longest first) together, so the work lists
// come out in about that order, as the separate dispatch kernel writes them
// (the banded kernels finish on short pairs).  No table lookups here, so
// nothing is lost by spreading an XCD's tasks.
constexpr int kRecClaim = 64;

__device__ __forceinline__ void prologue_records(const ProlArgs& A, const float* __restrict__ w, ProBuf& B,
                                                 int lane) {
  const FeatArgs& a = A.f;
  const bool by_xcd = gridDim.x % kXcds == 0 && gridDim.x >= kXcds;
  const int xcd = by_xcd ? static_cast<int>(blockIdx.x % kXcds) : 0;
  const int n_chunks = (a.n_tasks + kRecClaim - 1) / kRecClaim;
  int fill = 0;
  unsigned long long seg_first = 0;
  for (;;) {
    int c0 = 0;
    if (lane == 0) c0 = atomicAdd(a.next + xcd * kCtrStride, 1);
    c0 = __builtin_amdgcn_readfirstlane(__shfl(c0, 0));
    const int chunk = by_xcd ? xcd + kXcds * c0 : c0;
    if (chunk >= n_chunks) break;
    const int tc0 = chunk * kRecClaim;
    const int nt = min(a.n_tasks, tc0 + kRecClaim) - tc0;
    Task tk{0, 0, 0, 0, 0};
    int R = 0, gc = 0;
    float inv = 0.f, fqm = 0.f, fqn = 0.f;
    if (lane < nt) {
      const int ti = tc0 + lane;
      tk = a.tasks[ti];
      const int64_t ro = a.read_off[tk.read];
      R = static_cast<int>(a.read_off[tk.read + 1] - ro);
      inv = a.trec[3 * ti];
      fqm = a.trec[3 * ti + 1];
      fqn = a.trec[3 * ti + 2];
      gc = A.b.gcp_const[tk.read] != 0;
    }
    int my_t = 0, my_h = 0, loaded = fill;
    // slots [loaded, fill) take their task's fields and their record
    auto load_slots = [&]() {
      const bool mine = lane >= loaded && lane < fill;
      const int read = __shfl(tk.read, my_t), hap0 = __shfl(tk.hap0, my_t);
      const long long out0 = __shfl(static_cast<long long>(tk.out0), my_t);
      const int Rs = __shfl(R, my_t), gcs = __shfl(gc, my_t);
      const float invs = __shfl(inv, my_t), fqms = __shfl(fqm, my_t), fqns = __shfl(fqn, my_t);
      if (mine) {
        const int hap = hap0 + my_h;
        const long long o = out0 + my_h;
        const uint4* rp = reinterpret_cast<const uint4*>(a.rec + o * kRecBytes);
        const uint4 q0 = rp[0], q1 = rp[1];
        const int2 ap = *reinterpret_cast<const int2*>(a.rec + o * kRecBytes + kBins);
        const int H = static_cast<int>(a.hap_off[hap + 1] - a.hap_off[hap]);
        B.cnt[lane][0] = q0.x; B.cnt[lane][1] = q0.y; B.cnt[lane][2] = q0.z; B.cnt[lane][3] = q0.w;
        B.cnt[lane][4] = q1.x; B.cnt[lane][5] = q1.y; B.cnt[lane][6] = q1.z; B.cnt[lane][7] = q1.w;
        B.hap[lane] = hap;
        B.H[lane] = H;
        B.anchor[lane] = ap.x;
        B.pf[lane] = static_cast<int16_t>(ap.y & 0xffff);
        B.pl[lane] = static_cast<int16_t>(static_cast<uint32_t>(ap.y) >> 16);
        B.read[lane] = read;
        B.R[lane] = Rs;
        B.inv[lane] = invs;
        B.fqm[lane] = fqms;
        B.fqn[lane] = fqns;
        B.out[lane] = o;
        B.gc[lane] = gcs;
      }
    };
    for (int t = 0; t < nt; ++t) {
      const int nhap = __builtin_amdgcn_readlane(tk.nhap, t);
      for (int hs = 0; hs < nhap; hs += 64) {
        const int nsub = min(64, nhap - hs);
        if (fill + nsub > 64) {
          load_slots();
          __builtin_amdgcn_wave_barrier();
          prologue_flush(A, w, B, fill, seg_first, lane);
          __builtin_amdgcn_wave_barrier();
          fill = 0;
          loaded = 0;
          seg_first = 0;
        }
        seg_first |= 1ull << fill;
        if (lane >= fill && lane < fill + nsub) {
          my_t = t;
          my_h = hs + lane - fill;
        }
        fill += nsub;
      }
    }
    load_slots();
  }
  if (fill > 0) {
    __builtin_amdgcn_wave_barrier();
    prologue_flush(A, w, B, fill, seg_first, lane);
  }
}

// kRecords: the features come as compact records from band_features_kernel
// (the two-kernel prologue) instead of being computed here.
template <bool kRecords>
__global__ __launch_bounds__(64 * kProWaves) __attribute__((amdgpu_waves_per_eu(4))) void band_prologue_kernel(
    ProlArgs A) {
  __shared__ float w[kParams];
  __shared__ ProBuf bufs[kProWaves];
  const FeatArgs& a = A.f;
  // weights: W1 transposed (input-major), the rest as stored
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) {
    float v;
    if (i < kHid * kFeat) {
      const int k = i / kHid, u = i % kHid;
      v = A.params[u * kFeat + k];
    } else {
      v = A.params[i];
    }
    w[i] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  ProBuf& B = bufs[threadIdx.x >> 6];
  const TaskClaims cl = xcd_task_range(a.n_tasks);
  if constexpr (kRecords) {
    prologue_records(A, w, B, lane);
    return;
  }
  int fill = 0;
  unsigned long long seg_first = 0;
  for (;;) {
    const int tc0 = claim_tasks(a.next, cl, lane);
    if (tc0 < 0) break;
    const int tc1 = min(cl.end, tc0 + kFeatClaim);
    for (int ti = tc0; ti < tc1; ++ti) {
      const Task tk = a.tasks[ti];
      const int64_t ro = a.read_off[tk.read];
      const int R = static_cast<int>(a.read_off[tk.read + 1] - ro);
      const ReadSeeds rs = read_seeds(a.rclass, a.q, ro, R, lane);
      const float inv = rs.nvalid > 0 ? 1.f / static_cast<float>(rs.nvalid) : 0.f;
      const int gc = A.b.gcp_const[tk.read] != 0;
      for (int hs = 0; hs < tk.nhap; hs += 64) {
        const int nsub = min(64, tk.nhap - hs);
        if (fill + nsub > 64) {
          __builtin_amdgcn_wave_barrier();
          prologue_flush(A, w, B, fill, seg_first, lane);
          __builtin_amdgcn_wave_barrier();
          fill = 0;
          seg_first = 0;
        }
        seg_first |= 1ull << fill;
        for (int h0 = 0; h0 < nsub; h0 += kHapChunk) {
          ChunkHits ch;
          chunk_lookups(a.hap_off, a.htab_off, a.htab, tk.hap0 + hs + h0, min(kHapChunk, nsub - h0), rs.my_p,
                        rs.my_key, ch);
#pragma unroll
          for (int c = 0; c < kHapChunk; ++c) {
            const int h = hs + h0 + c;
            if (h0 + c >= nsub) break;
            const int hap = tk.hap0 + h;
            const int H = static_cast<int>(a.hap_off[hap + 1] - a.hap_off[hap]);
            const PairFeat f = pair_features(ch.first[c], ch.occ[c], ch.nocc[c], rs.my_p, R, H, lane, B.hist);
            const int slot = fill + h0 + c;
            if (lane < kBins) reinterpret_cast<uint8_t*>(B.cnt[slot])[lane] = static_cast<uint8_t>(f.cnt);
            if (lane == 0) {
              B.hap[slot] = hap;
              B.H[slot] = H;
              B.anchor[slot] = f.anchor;
              B.pf[slot] = f.pf;
              B.pl[slot] = f.pl;
              B.read[slot] = tk.read;
              B.R[slot] = R;
              B.inv[slot] = inv;
              B.fqm[slot] = rs.fq_mean;
              B.fqn[slot] = rs.fq_min;
              B.out[slot] = tk.out0 + h;
              B.gc[slot] = gc;
            }
          }
        }
        fill += nsub;
      }
    }
  }
  if (fill > 0) {
    __builtin_amdgcn_wave_barrier();
    prologue_flush(A, w, B, fill, seg_first, lane);
  }
}

// Start bands of long reads (one task per pair, run in FP64 on the band that
// follows the alignment).  One wavefront per pair: lane l looks up the read's
// 12-mer at position 4l (the first 256 bases) in the haplotype's table; seeds
// found exactly once give diagonals q - p; the start diagonal is their median
// over the first 64 bases (at least kLongMinFirst seeds; else over all 256),
// and the band starts as the 64 diagonals from it - kLongStartLift (room
// above the alignment for a read start explained as a run of insertions).
// With fewer than kLongMinHits such seeds the pair keeps the full matrix.
constexpr int kLongSeedStride = 4;
constexpr int kLongMinHits = 4;
constexpr int kLongFirstLanes = 16;  // seeds of the first 64 bases
constexpr int kLongMinFirst = 3;

__global__ __launch_bounds__(64) void long_anchor_kernel(const uint8_t* __restrict__ rclass,
                                                         const int64_t* __restrict__ read_off,
                                                         const int64_t* __restrict__ htab_off,
                                                         const unsigned long long* __restrict__ htab,
                                                         const Task* __restrict__ tasks, int n_tasks,
                                                         int2* __restrict__ band) {
  const int lane = threadIdx.x;
  for (int ti = blockIdx.x; ti < n_tasks; ti += gridDim.x) {
    const Task tk = tasks[ti];
    const int64_t ro = read_off[tk.read];
    const int R = static_cast<int>(read_off[tk.read + 1] - ro);
    const int p = lane * kLongSeedStride;
    int diag = INT_MIN;
    const int64_t t0 = htab_off[tk.hap0], t1 = htab_off[tk.hap0 + 1];
    if (p + kK <= R && t1 - t0 >= 2) {
      uint32_t key = 0;
      uint8_t wild = 0;
      for (int k = 0; k < kK; ++k) {
        const uint8_t c = rclass[ro + p + k];
        wild |= c;
        key = (key << 2) | (c & 3);
      }
      if (!(wild & kWildcard)) {
        const unsigned long long* t = htab + t0;
        const uint32_t mask = static_cast<uint32_t>(t1 - t0 - 1);
        uint32_t sl = slot_of(key, mask);
        int found = 0, q = 0;
        for (uint32_t probe = 0; probe <= mask; ++probe) {
          const unsigned long long e = t[sl];
          if (e == kEmpty) break;
          if (static_cast<uint32_t>(e) == key) {
            q = static_cast<int>(e >> 32);
            if (++found > 1) break;
          }
          sl = (sl + 1) & mask;
        }
        if (found == 1) diag = q - p;
      }
    }
    const bool hit = diag != INT_MIN;
    const int nhit = __popcll(__ballot(hit));
    // the median over the lanes in `use`
    auto median = [&](bool use, int n_use) {
      const int half = (n_use - 1) / 2;
      int below = 0, equal = 0;
      for (int l = 0; l < 64; ++l) {
        const int v = __shfl(diag, l);
        const bool lu = __shfl(static_cast<int>(use), l) != 0;
        if (lu && use) {
          below += v < diag;
          equal += v == diag;
        }
      }
      const bool is_median = use && below <= half && half < below + equal;
      const unsigned long long mb = __ballot(is_median);
      return __shfl(diag, static_cast<int>(__ffsll(static_cast<long long>(mb)) - 1));
    };
    int2 b{1, 0};
    if (nhit >= kLongMinHits) {
      // the start diagonal: from the seeds of the first 64 bases when there
      // are enough of them (the alignment drifts along a long read)
      const bool first = hit && lane < kLongFirstLanes;
      const int nfirst = __popcll(__ballot(first));
      const int med = nfirst >= kLongMinFirst ? median(first, nfirst) : median(hit, nhit);
      b = int2{med - kLongStartLift, med - kLongStartLift + kBandLanes - 1};
    }
    if (lane == 0) band[tk.out0] = b;
  }
}

}  // namespace bvp

// ---- launchers ---------------------------------------------------------------
namespace bvp {

hipError_t launch_hap_tables(const uint8_t* hmask, const int64_t* hap_off, const int64_t* tab_off,
                             int64_t n_haps, unsigned long long* tab, hipStream_t st) {
  const int64_t blocks = std::min<int64_t>(std::max<int64_t>(n_haps, 1), 65536);
  if (n_haps > 0)
    hipLaunchKernelGGL(hap_table_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, hmask,
                       hap_off, tab_off, n_haps, tab);
  return hipGetLastError();
}

hipError_t launch_long_anchors(const uint8_t* rclass, const int64_t* read_off, const int64_t* htab_off,
                               const unsigned long long* htab, const Task* tasks, int n_tasks, int2* band,
                               int grid, hipStream_t st) {
  if (n_tasks > 0)
    hipLaunchKernelGGL(long_anchor_kernel, dim3(grid), dim3(64), 0, st, rclass, read_off, htab_off, htab,
                       tasks, n_tasks, band);
  return hipGetLastError();
}

hipError_t launch_band_features(const FeatArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(band_features_kernel, dim3(grid), dim3(64), 0, st, a);
  return hipGetLastError();
}

// The fused prologue's persistent grid: its occupancy on every CU, a
// multiple of the XCD count (per-XCD task claims).
int band_prologue_grid(int cus, int n_tasks) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, band_prologue_kernel<false>, 64 * kProWaves, 0) !=
          hipSuccess ||
      n < 1)
    n = 1;
  const int want = (n_tasks + kProWaves - 1) / kProWaves;
  int g = std::max(1, std::min(cus * n, want));
  if (g >= kXcds) g -= g % kXcds;
  return g;
}

hipError_t launch_band_prologue(const FeatArgs& f, const BArgs& b, int2* band, const float* params, int margin,
                                int max_width, int allow_narrow, int grid, hipStream_t st) {
  if (f.n_tasks <= 0) return hipSuccess;
  ProlArgs A{f, b, band, params, margin, max_width, allow_narrow};
  if (f.rec != nullptr)
    hipLaunchKernelGGL(band_prologue_kernel<true>, dim3(grid), dim3(64 * kProWaves), 0, st, A);
  else
    hipLaunchKernelGGL(band_prologue_kernel<false>, dim3(grid), dim3(64 * kProWaves), 0, st, A);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const float* params, const float* X, int64_t n, float* Y, float* H1,
                              float* H2, hipStream_t st) {
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(static_cast<unsigned>(std::max<int64_t>(1, blocks))),
                     dim3(256), 0, st, params, X, n, Y, H1, H2);
  return hipGetLastError();
}

hipError_t launch_mlp_backward(const float* params, const float* X, const float* Y, const float* T,
                               int64_t n, const float* H1, const float* H2, float* D1, float* D2,
                               float* D3, float w0, float w1, float* loss, float* grad,
                               hipStream_t st) {
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(mlp_delta_kernel, dim3(static_cast<unsigned>(std::max<int64_t>(1, blocks))),
                     dim3(256), 0, st, params, Y, T, n, H1, H2, D1, D2, D3, w0, w1, loss);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3((kParams + 255) / 256), dim3(256), 0, st, X, n, H1, H2,
                     D1, D2, D3, grad);
  return hipGetLastError();
}

hipError_t launch_adam(float* params, const float* grad, float* m1, float* m2, float lr, int t,
                       hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, static_cast<float>(t));
  const float c2 = 1.f - powf(b2, static_cast<float>(t));
  hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, st, params, grad, m1,
                     m2, lr, b1, b2, 1e-8f, c1, c2);
  return hipGetLastError();
}

hipError_t launch_band_from_prediction(const float* Y, const int* anchor, const int2* ends,
                                       const float* X, int64_t n, int margin, int max_width,
                                       int allow_narrow, int2* band, hipStream_t st) {
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0)
    hipLaunchKernelGGL(band_from_prediction_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st,
                       Y, anchor, ends, X, n, margin, max_width, allow_narrow, band);
  return hipGetLastError();
}

}  // namespace bvp
