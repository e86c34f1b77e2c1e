// MI355X (gfx950) PairHMM forward kernels.
//
// Arithmetic contract (reference: lib/intel_native/pairhmm/avx-pairhmm-template.h
// :83-128 per-row transitions, :136-177 boundary, :183-198 cell recurrence,
// :299-345 final sum; pairhmm.hpp:55-80 for the constant-transition mode):
//
//   M[r][j] = prior(r,j) * (M[r-1][j-1]*mm_r + I[r-1][j-1]*im_r + D[r-1][j-1]*dm_r)
//   I[r][j] = M[r-1][j]*mx_r + I[r-1][j]*xx_r
//   D[r][j] = M[r][j-1]*my_r + D[r][j-1]*yy_r
//   row 0: M = I = 0, D = K/H;  column 0: M = I = D = 0
//   S = sum_j (M[R][j] + I[R][j])
//
// Wavefront layout.  A wavefront sweeps anti-diagonals over a stripe of 63
// read rows: lane l (1..63) holds row r0-1+l and at step t works on column
// j = t - l.  Lane 0 is a feeder that holds the previous stripe's last row.
// Each lane publishes, per cell, the two terms its lower neighbour needs:
//   Z = M*mm' + I*im' + D*dm'   (' = next row's transitions)
//   J = M*mx' + I*xx'
// so the neighbour's M is one DPP-shifted multiply, dpp(Z two steps back) *
// prior, and its I is dpp(J one step back) -- I is never materialised.  Per
// cell this is 9 VALU instructions (1 prior, 1 M, 2 D, 3 Z, 2 J) with every
// cross-lane move folded into a v_*_dpp wave_shr:1 operand.
//
// Carry between stripes: the stripe's last row writes (Z, J) per column into a
// per-wavefront global scratch row, staged through LDS in 64-step blocks so
// that global traffic is coalesced and the per-step work stays in LDS.  The
// haplotype enters as a per-column match bitmask over up to 8 read-base
// classes, expanded into a sliding LDS window of f16 selectors; the prior is
// one fused multiply-add: prior = sel * (pm - pmis) + pmis.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "pairhmm_common.h"

namespace bvp {


// AVX alphabet (pairhmm_common.h:26-44): A0 C1 T2 G3, N a wildcard on either
// side, every other byte acts as A -- the class byte and haplotype mask the
// host otherwise builds once per plan.
__device__ __forceinline__ uint8_t avx_read_class(uint8_t b) {
  return b == 'C' ? 1 : b == 'T' ? 2 : b == 'G' ? 3 : b == 'N' ? kWildcard : 0;
}
__device__ __forceinline__ uint8_t avx_hap_mask(uint8_t b) {
  return b == 'C' ? 2 : b == 'T' ? 4 : b == 'G' ? 8 : b == 'N' ? 0xFF : 1;
}
__device__ __forceinline__ uint8_t read_class(const KArgs& a, int64_t p) {
  return a.rbases ? avx_read_class(a.rbases[p]) : a.rclass[p];
}
__device__ __forceinline__ uint8_t hap_mask(const KArgs& a, int64_t p) {
  return a.rbases ? avx_hap_mask(a.hbases[p]) : a.hmask[p];
}

template <typename T>
struct alignas(2 * sizeof(T)) Pair2 {
  T z, j;
};

template <typename T>
struct Lds {
  Pair2<T> cin[kBlock];       // carry-in, columns [64b, 64b+64)
  Pair2<T> cout[2 * kBlock];  // carry-out window, columns [64b-64, 64b+64)
  _Float16 win[2 * kBlock][kSlots];  // match selectors, columns [64b-64, 64b+64)
};

__device__ __forceinline__ float dpp_shr(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138 /*wave_shr:1*/, 0xf, 0xf, true));
}

__device__ __forceinline__ double dpp_shr(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), 0x138, 0xf, 0xf, true);
  return __longlong_as_double(static_cast<long long>(static_cast<unsigned>(lo)) |
                              (static_cast<long long>(hi) << 32));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

template <typename T>
struct Num;
template <>
struct Num<float> {
  static constexpr float K = 0x1p120f;  // Context.h:109 (FP32)
  static constexpr double log10K = 120.0 * 0.30102999566398119521;  // log10(2^120)
  __device__ static float ph2pr(const Tables& t, int q) { return t.ph2pr_f[q]; }
  __device__ static float mm(const Tables& t, int i) { return t.mm_f[i]; }
  __device__ static float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
};
template <>
struct Num<double> {
  static constexpr double K = 0x1p1020;  // Context.h:149 (FP64)
  static constexpr double log10K = 1020.0 * 0.30102999566398119521;  // log10(2^1020)
  __device__ static double ph2pr(const Tables& t, int q) { return t.ph2pr_d[q]; }
  __device__ static double mm(const Tables& t, int i) { return t.mm_d[i]; }
  __device__ static double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
};


constexpr int kRenormTarget = 1000;  // carry max is rescaled to 2^1000

template <typename T>
__device__ __forceinline__ void row_params(const KArgs& a, int64_t ro, int R, int r, int lane,
                                           T& pmis, T& dp, T& my, T& yy, int& cls, T& mm_n,
                                           T& im_n, T& dm_n, T& mx_n, T& xx_n) {
  using N = Num<T>;
  T pm = T(0);
  pmis = T(0); my = T(0); yy = T(0); cls = 0;
  mm_n = T(0); im_n = T(0); dm_n = T(0); mx_n = T(0); xx_n = T(0);
  if (lane >= 1 && r >= 1 && r <= R) {
    const int64_t p = ro + r - 1;
    const uint8_t rc = read_class(a, p);
    T e;
    if (a.mode == kModeAvx) {
      e = N::ph2pr(a.tab, a.q[p] & 127);
      my = N::ph2pr(a.tab, a.del[p] & 127);
      yy = N::ph2pr(a.tab, a.gcp[p] & 127);
    } else {
      e = static_cast<T>(a.tab.err_scalar[a.q[p]]);
      my = static_cast<T>(a.tab.trans[2]);
      yy = static_cast<T>(a.tab.trans[6]);
    }
    pm = T(1) - e;
    pmis = e / T(3);
    if (rc & kWildcard) pmis = pm;
    cls = rc & 7;
  }
  // prior = sel * dp + pmis gives pm for a match selector of 1.
  dp = pm - pmis;
  if (lane >= 1 && r == R) {
    // Last read row: J = M + I, the per-column summand of S.
    mx_n = T(1);
    xx_n = T(1);
  } else if (r + 1 >= 1 && r + 1 <= R) {
    if (a.mode == kModeAvx) {
      const int64_t p1 = ro + r;  // row r+1
      const int iq = a.ins[p1] & 127, dq = a.del[p1] & 127, cq = a.gcp[p1] & 127;
      const int lo = iq < dq ? iq : dq, hi = iq < dq ? dq : iq;
      mm_n = N::mm(a.tab, ((hi * (hi + 1)) >> 1) + lo);
      im_n = T(1) - N::ph2pr(a.tab, cq);
      dm_n = im_n;
      mx_n = N::ph2pr(a.tab, iq);
      xx_n = N::ph2pr(a.tab, cq);
    } else {
      mm_n = static_cast<T>(a.tab.trans[0]);
      mx_n = static_cast<T>(a.tab.trans[1]);
      im_n = static_cast<T>(a.tab.trans[3]);
      xx_n = static_cast<T>(a.tab.trans[4]);
      dm_n = static_cast<T>(a.tab.trans[5]);
    }
  }
}

// One anti-diagonal step for lane-local state.  ZS is the slot holding Z from
// two steps back; it receives this step's Z.
#define BVP_STEP(U, ZS)                                                       \
  {                                                                           \
    const T sel = static_cast<T>(selp[(U) * kSlots]);                          \
    const T prior = N::fma(sel, dp, pmis);                                     \
    const T M = dpp_shr(ZS) * prior;                                           \
    const T D = N::fma(D1, yy, M1 * my);                                       \
    const T jin = dpp_shr(J1);                                                 \
    T zn = N::fma(M, mm_n, N::fma(D, dm_n, jin * im_n));                       \
    T jn = N::fma(M, mx_n, jin * xx_n);                                        \
    if (lane == 0) {                                                           \
      const Pair2<T> c = s.cin[(U)];                                           \
      zn = c.z;                                                                \
      jn = c.j;                                                                \
    }                                                                          \
    if (lane == L) outp[(U)] = Pair2<T>{zn, jn};                               \
    ZS = zn;                                                                   \
    J1 = jn;                                                                   \
    M1 = M;                                                                    \
    D1 = D;                                                                    \
  }

#define BVP_STEP2(U) BVP_STEP((U), Za) BVP_STEP((U) + 1, Zb)
#define BVP_STEP16(U) \
  BVP_STEP2((U)) BVP_STEP2((U) + 2) BVP_STEP2((U) + 4) BVP_STEP2((U) + 6) \
  BVP_STEP2((U) + 8) BVP_STEP2((U) + 10) BVP_STEP2((U) + 12) BVP_STEP2((U) + 14)

// Forward probability of one (read, haplotype) pair.  Returns the scaled sum
// S (as a double) and the power-of-two exponent removed by renormalisation.
template <typename T>
__device__ double pair_forward(const KArgs& a, Lds<T>& s, Pair2<T>* carry, int lane,
                               int read, int hap, int& shift_total) {
  using N = Num<T>;
  const int64_t ro = a.read_off[read];
  const int R = static_cast<int>(a.read_off[read + 1] - ro);
  const int64_t ho = a.hap_off[hap];
  const int H = static_cast<int>(a.hap_off[hap + 1] - ho);
  const int nstripes = (R + kRowsPerStripe - 1) / kRowsPerStripe;
  shift_total = 0;
  T scale = T(1);
  double acc = 0.0;
  for (int st = 0; st < nstripes; ++st) {
    const int r0 = kRowsPerStripe * st + 1;
    const bool last = (st == nstripes - 1);
    const int L = last ? (R - r0 + 1) : kRowsPerStripe;
    const int r = r0 - 1 + lane;
    // Lane l at step t works on column t - l.
    const int c_hi = H;
    const int base = 0;
    T pmis, dp, my, yy, mm_n, im_n, dm_n, mx_n, xx_n;
    int cls;
    row_params<T>(a, ro, R, r, lane, pmis, dp, my, yy, cls, mm_n, im_n, dm_n, mx_n, xx_n);
    // Row 0 of the matrix as seen by row 1: Z = (K/H) * dm_1, J = 0.
    const T z_row0 = __shfl((T(N::K) / static_cast<T>(H)) * dm_n, 0);
    const int steps = (c_hi - base) + L + 1;
    const int nb = (steps + kBlock - 1) / kBlock;
    // Prefetch block 0.
    T pre_z, pre_j;
    uint8_t pre_m;
    {
      const int c = lane + base;
      if (st == 0) {
        pre_z = z_row0;
        pre_j = T(0);
      } else if (c <= H) {
        const Pair2<T> v = carry[c];
        pre_z = v.z * scale;
        pre_j = v.j * scale;
      } else {
        pre_z = T(0);
        pre_j = T(0);
      }
      pre_m = (c >= 1 && c <= H) ? hap_mask(a, ho + c - 1) : uint8_t(0);
    }
    T Za = T(0), Zb = T(0), J1 = T(0), M1 = T(0), D1 = T(0);
    T maxabs = T(0);
    const _Float16* selp = &s.win[kBlock - lane][cls];
    Pair2<T>* outp = &s.cout[kBlock - L];
    for (int b = 0; b < nb; ++b) {
      __syncthreads();
      s.cin[lane] = Pair2<T>{pre_z, pre_j};
      {
        _Float16 lo[kSlots];
#pragma unroll
        for (int k = 0; k < kSlots; ++k) lo[k] = (b == 0) ? _Float16(0) : s.win[kBlock + lane][k];
#pragma unroll
        for (int k = 0; k < kSlots; ++k) s.win[lane][k] = lo[k];
#pragma unroll
        for (int k = 0; k < kSlots; ++k)
          s.win[kBlock + lane][k] = ((pre_m >> k) & 1) ? _Float16(1) : _Float16(0);
      }
      {
        const int c = kBlock * (b + 1) + lane + base;
        if (st == 0) {
          pre_z = z_row0;
          pre_j = T(0);
        } else if (c <= H) {
          const Pair2<T> v = carry[c];
          pre_z = v.z * scale;
          pre_j = v.j * scale;
        } else {
          pre_z = T(0);
          pre_j = T(0);
        }
        pre_m = (c >= 1 && c <= H) ? hap_mask(a, ho + c - 1) : uint8_t(0);
      }
      __syncthreads();
      const int left = steps - kBlock * b;
      BVP_STEP16(0)
      if (left > 16) {
        BVP_STEP16(16)
        if (left > 32) {
          BVP_STEP16(32)
          if (left > 48) {
            BVP_STEP16(48)
          }
        }
      }
      __syncthreads();
      // Flush the 64 columns the last row produced in this block.
      const int c = kBlock * b - L + lane + base;
      const Pair2<T> v = s.cout[kBlock - L + lane];
      if (c >= base && c <= c_hi) {
        if (!last) {
          carry[c] = v;
          maxabs = fmax(maxabs, fmax(fabs(v.z), fabs(v.j)));
        } else if (c >= 1) {
          acc += static_cast<double>(v.j);
        }
      }
    }
    if (!last) {
      // Make this stripe's carry row visible to the next stripe's loads.
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      // Only the FP64 form renormalises; the FP32 form keeps the reference's
      // fixed 2^120 scaling so its underflow decisions match.
      if (sizeof(T) == 8 && a.renorm) {
        const double m = wave_max(static_cast<double>(maxabs));
        if (m > 0.0) {
          const int e = ilogb(m);
          const int k = kRenormTarget - e;
          scale = static_cast<T>(ldexp(1.0, k));
          shift_total += k;
        } else {
          scale = T(1);
        }
      }
    }
  }
  return wave_sum(acc);
}

template <typename T>
__global__ __launch_bounds__(64) void pairhmm_forward_kernel(KArgs a) {
  __shared__ Lds<T> s;
  const int lane = threadIdx.x;
  Pair2<T>* carry = static_cast<Pair2<T>*>(a.carry) + static_cast<int64_t>(blockIdx.x) * a.carry_stride;
  const int n_tasks = a.n_tasks_dev ? *a.n_tasks_dev : a.n_tasks;
  // Cell counts are summed per wavefront and added once at the end (a
  // same-address atomic per pair would queue at the L2).
  unsigned long long cells_w = 0;
  for (;;) {
    int ti = 0;
    if (lane == 0) ti = atomicAdd(a.counter, 1);
    ti = __builtin_amdgcn_readfirstlane(__shfl(ti, 0));
    if (ti >= n_tasks) break;
    const Task tk = a.identity_tasks ? Task{ti, ti, 1, 0, ti} : a.tasks[ti];
    for (int h = 0; h < tk.nhap; ++h) {
      int shift = 0;
      const int64_t o = tk.out0 + h;
      {
        const int64_t R = a.read_off[tk.read + 1] - a.read_off[tk.read];
        const int64_t H = a.hap_off[tk.hap0 + h + 1] - a.hap_off[tk.hap0 + h];
        if (R < 1 || H < 1 || H > a.hcap) {
          // no sweep past the carry row: the pair gets NaN
          if (lane == 0) {
            a.out[o] = __builtin_nan("");
            if (a.form) a.form[o] = sizeof(T) == 4 ? kFormFP32 : kFormFP64;
          }
          continue;
        }
      }
      const double S = pair_forward<T>(a, s, carry, lane, tk.read, tk.hap0 + h, shift);
      if (lane == 0) {
        cells_w += static_cast<unsigned long long>(a.read_off[tk.read + 1] - a.read_off[tk.read]) *
                   static_cast<unsigned long long>(a.hap_off[tk.hap0 + h + 1] - a.hap_off[tk.hap0 + h]);
        if (a.out_raw) {
          a.out[o] = ldexp(S, -shift);
          if (a.form) a.form[o] = sizeof(T) == 4 ? kFormFP32 : kFormFP64;
        } else if constexpr (sizeof(T) == 4) {
          if (a.verify_pass) {
            // the reference rule for a long read run in FP64 first: its FP32
            // result replaces the FP64 one unless it underflows
            if (S >= static_cast<double>(kMinAccepted)) {
              a.out[o] = log10(S) - Num<T>::log10K;
              a.form[o] = kFormFP32;
            }
          } else if (S < static_cast<double>(kMinAccepted) && a.rescue != nullptr) {
            const int k = atomicAdd(a.rescue_count, 1);
            a.rescue[k] = Task{tk.read, tk.hap0 + h, 1, -1, o};
            a.form[o] = kFormPending;
            const unsigned long long cells =
                static_cast<unsigned long long>(a.read_off[tk.read + 1] - a.read_off[tk.read]) *
                static_cast<unsigned long long>(a.hap_off[tk.hap0 + h + 1] - a.hap_off[tk.hap0 + h]);
            atomicAdd(a.rescue_cells, cells);
          } else {
            a.out[o] = log10(S) - Num<T>::log10K;
            a.form[o] = tk.band == -2 ? kFormBandFallback : kFormFP32;
          }
        } else {
          const double L = log10(S) - static_cast<double>(shift) * 0.30102999566398120 - Num<T>::log10K;
          a.out[o] = L;
          a.form[o] = kFormFP64;
          if (a.verify != nullptr && L >= a.verify_above) {
            const int k = atomicAdd(a.verify_count, 1);
            a.verify[k] = Task{tk.read, tk.hap0 + h, 1, -1, o};
          }
        }
      }
    }
  }
  if (lane == 0 && cells_w != 0 && a.cells_done != nullptr) {
    atomicAdd(a.cells_done, cells_w);
    if (sizeof(T) == 8 && a.fp64_cells != nullptr) atomicAdd(a.fp64_cells, cells_w);
  }
}

__global__ void zero_ints_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Per-base insertion / deletion / gap-continuation bytes from one triplet per
// read (reads whose three quality strings are each constant along the read,
// as SamRecord's GOP/GCP strings are, sam.hpp:451-475): the host uploads 3
// bytes per read instead of 3 per base.  One workgroup per read at a time.
__global__ __launch_bounds__(64) void expand_read_bytes_kernel(const int64_t* __restrict__ read_off,
                                                                int64_t n_reads,
                                                                const uint8_t* __restrict__ triplets,
                                                                uint8_t* __restrict__ ins,
                                                                uint8_t* __restrict__ del,
                                                                uint8_t* __restrict__ gcp) {
  for (int64_t r = blockIdx.x; r < n_reads; r += gridDim.x) {
    const int64_t a = read_off[r], b = read_off[r + 1];
    const uint8_t i = triplets[3 * r], d = triplets[3 * r + 1], c = triplets[3 * r + 2];
    for (int64_t p = a + threadIdx.x; p < b; p += blockDim.x) {
      ins[p] = i;
      del[p] = d;
      gcp[p] = c;
    }
  }
}

// Two-bit bases to the kernels' per-base bytes.  Reads: class = the code
// (A C T G -> 0..3), kReadEnd on a read's last base; haplotypes: match mask
// 1 << code.  Then the 'N' positions: the wildcard class bit for reads, the
// all-classes mask for haplotypes.  One workgroup per read at a time; the
// haplotype bytes grid-stride.
__global__ __launch_bounds__(256) void expand_bases_kernel(const int64_t* __restrict__ read_off, int64_t n_reads,
                                                           const uint8_t* __restrict__ rpack,
                                                           uint8_t* __restrict__ rclass, int64_t hbytes,
                                                           const uint8_t* __restrict__ hpack,
                                                           uint8_t* __restrict__ hmask) {
  for (int64_t r = blockIdx.x; r < n_reads; r += gridDim.x) {
    const int64_t a = read_off[r], b = read_off[r + 1];
    for (int64_t p = a + threadIdx.x; p < b; p += blockDim.x) {
      const uint8_t code = (rpack[p >> 2] >> (2 * (p & 3))) & 3u;
      rclass[p] = static_cast<uint8_t>(code | (p == b - 1 ? kReadEnd : 0));
    }
  }
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < hbytes;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    hmask[i] = static_cast<uint8_t>(1u << ((hpack[i >> 2] >> (2 * (i & 3))) & 3u));
}

__global__ __launch_bounds__(256) void mark_wildcards_kernel(const int64_t* __restrict__ rexc, int64_t n_rexc,
                                                             uint8_t* __restrict__ rclass,
                                                             const int64_t* __restrict__ hexc, int64_t n_hexc,
                                                             uint8_t* __restrict__ hmask) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n_rexc) rclass[rexc[i]] |= kWildcard;
  if (i < n_hexc) hmask[hexc[i]] = 0xFF;
}

}  // namespace bvp

// Device scratch of bv_compute_full_prob_noalloc, for callers that cannot
// allocate: the task queue counter and the per-wavefront carry rows
// (kNoallocScratchBytes, held by each device for the life of the process).
// Host code never passes this name to a kernel: it is a host-side stand-in.
// Its device address comes from noalloc_scratch_address, once per device.
static __device__ __align__(256) unsigned char scratch[bvp::kNoallocScratchBytes];

// ---- launchers (host) ----------------------------------------------------

namespace bvp {

// The device address of the scratch array on the current device; the
// library's init calls it once per device with that device current.
hipError_t noalloc_scratch_address(void** ptr) {
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

hipError_t launch_zero_ints(int* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_expand_read_bytes(const int64_t* read_off, int64_t n_reads, const uint8_t* triplets,
                                    uint8_t* ins, uint8_t* del, uint8_t* gcp, hipStream_t st) {
  if (n_reads <= 0) return hipSuccess;
  const int64_t grid = std::min<int64_t>(n_reads, 65536);
  hipLaunchKernelGGL(expand_read_bytes_kernel, dim3(static_cast<unsigned>(grid)), dim3(64), 0, st, read_off,
                     n_reads, triplets, ins, del, gcp);
  return hipGetLastError();
}

// Longest first on the device: a stable radix sort of the task indices by
// descending cost (ties keep the input order, as the host's stable sort
// did), then a gather of the tasks.  Costs are below 2^end_bit.
__global__ void iota_kernel(int32_t* __restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[i] = static_cast<int32_t>(i);
}

template <typename T>
__global__ void gather_kernel(const T* __restrict__ in, const int32_t* __restrict__ idx, int64_t n,
                              T* __restrict__ out) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[i] = in[idx[i]];
}

template <typename T>
hipError_t sort_longest_first(const T* d_in, const uint64_t* d_cost, int64_t n, int end_bit, T* d_out,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int num = static_cast<int>(n);
  uint64_t* keys_out = nullptr;
  int32_t *idx = nullptr, *idx_out = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp_bytes, d_cost, keys_out, idx, idx_out,
                                                              num, 0, end_bit, st);
  if (e == hipSuccess) e = hipMalloc(&keys_out, sizeof(uint64_t) * n);
  if (e == hipSuccess) e = hipMalloc(&idx, sizeof(int32_t) * n);
  if (e == hipSuccess) e = hipMalloc(&idx_out, sizeof(int32_t) * n);
  if (e == hipSuccess) e = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1));
  const unsigned grid = static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 65536));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(256), 0, st, idx, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairsDescending(tmp, tmp_bytes, d_cost, keys_out, idx, idx_out, num, 0, end_bit,
                                                     st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(gather_kernel<T>, dim3(grid), dim3(256), 0, st, d_in, idx_out, n, d_out);
    e = hipGetLastError();
  }
  const hipError_t es = hipStreamSynchronize(st);
  for (void* q : {static_cast<void*>(keys_out), static_cast<void*>(idx), static_cast<void*>(idx_out), tmp})
    if (q) (void)hipFree(q);
  return e != hipSuccess ? e : es;
}

template hipError_t sort_longest_first<Task>(const Task*, const uint64_t*, int64_t, int, Task*, hipStream_t);
template hipError_t sort_longest_first<StackTask>(const StackTask*, const uint64_t*, int64_t, int, StackTask*,
                                                  hipStream_t);

hipError_t launch_expand_bases(const int64_t* read_off, int64_t n_reads, const uint8_t* rpack, const int64_t* rexc,
                               int64_t n_rexc, uint8_t* rclass, int64_t hbytes, const uint8_t* hpack,
                               const int64_t* hexc, int64_t n_hexc, uint8_t* hmask, hipStream_t st) {
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(std::max(n_reads, (hbytes + 255) / 256), 65536));
  hipLaunchKernelGGL(expand_bases_kernel, dim3(static_cast<unsigned>(grid)), dim3(256), 0, st, read_off, n_reads,
                     rpack, rclass, hbytes, hpack, hmask);
  const int64_t n = std::max(n_rexc, n_hexc);
  if (n > 0)
    hipLaunchKernelGGL(mark_wildcards_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, rexc,
                       n_rexc, rclass, hexc, n_hexc, hmask);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_forward(const KArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(pairhmm_forward_kernel<T>, dim3(grid), dim3(64), 0, st, a);
  return hipGetLastError();
}

template hipError_t launch_forward<float>(const KArgs&, int, hipStream_t);
template hipError_t launch_forward<double>(const KArgs&, int, hipStream_t);

template <typename T>
int max_blocks_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pairhmm_forward_kernel<T>, 64, 0) != hipSuccess)
    return 8;
  return n > 0 ? n : 1;
}
template int max_blocks_per_cu<float>();

template int max_blocks_per_cu<double>();

}  // namespace bvp
