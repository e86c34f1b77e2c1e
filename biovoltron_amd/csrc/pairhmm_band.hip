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

#include <cfloat>
#include <climits>
#include <type_traits>

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

// The gap-to-match factor g = 1 - ph2pr[gcp] is constant along a banded read
// (the dispatch admits only reads with a constant, nonzero-quality gap
// continuation), so the kernel carries I and D times g: M then takes
// mm * M + (I + D) with no multiply by g, the insert and delete openings are
// mx * g and my * g, and the final sums divide I by g once.
struct alignas(16) RowRec {
  float dp;   // pm - pe, pm = 1 - e the match prior
  float pe;   // e / 3 (mismatch prior; pm for a wildcard read base)
  float mm;   // match -> match into this row
  float mx;   // g * (match -> insert)
  float xx;   // insert -> insert
  float my;   // g * (match -> delete, along this row)
  uint32_t selo;  // byte offset of this row's selectors: 2 * (cls * C + i - 1)
  float pad;
};

// Match selectors of the band's haplotype window, per lane group: slot c of
// the window is haplotype column j = wb + 1 + c, wb being the band's lowest
// offset, so row i, lane k (column i + o(i) + k) reads slot
// i - 1 + k + o(i) - wb; sel[g][cls][c] is 1 where that column matches read
// class cls (0 outside the haplotype).  The prior is then one mixed-precision
// FMA, sel * (pm - pe) + pe.  A lane's address is its part k + o(i) - wb
// (changed only by a slide) plus the row's part from the row record.  The
// window has
// C = a.sel_cols columns per class, sized by the plan's longest banded read
// (dynamic LDS: shorter reads leave room for more wavefronts per CU).
constexpr int kTabBytes = static_cast<int>(sizeof(RowRec)) * (kBandLanes + 1);
extern __shared__ __align__(16) unsigned char band_lds[];

// Inclusive decayed scan over each G-lane group: returns y with
// y_k = sum_{m>=0} yy^m x_{k-m}, k and k-m in the same group.  Four row_shr
// levels scan each 16-lane DPP row; then row_bcast:15 adds the lower row's
// prefix times f15 = yy^(k%16+1) (0 on rows 0 and 2), and for G = 64
// row_bcast:31 adds the lower half's prefix times f31 = yy^(k-31) (0 below).
template <int G>
__device__ __forceinline__ float decay_scan(float x, float yy1, float yy2, float yy4, float yy8,
                                            float f15, float f31) {
  x = __builtin_fmaf(dpp<kRowShr + 1>(x), yy1, x);
  x = __builtin_fmaf(dpp<kRowShr + 2>(x), yy2, x);
  x = __builtin_fmaf(dpp<kRowShr + 4>(x), yy4, x);
  x = __builtin_fmaf(dpp<kRowShr + 8>(x), yy8, x);
  x = __builtin_fmaf(dpp<kRowBcast15>(x), f15, x);
  if constexpr (G == 64) x = __builtin_fmaf(dpp<kRowBcast31>(x), f31, x);
  return x;
}

// Both decayed scans of a row (the forward D and the touch D), as
// decay_scan<G>, with each level's DPP operand folded into one v_fmac_f32_dpp
// (the compiler does not combine DPP into a multiply-accumulate and emits a
// separate v_mov_b32_dpp per level).  The two scans are interleaved so that
// every DPP read of the previous level's result is two wait states after its
// write (one instruction plus one s_nop), as gfx9 requires; the block starts
// and ends with s_nop 1 for the writes and reads around it.
template <int G>
__device__ __forceinline__ void decay_scan2(float& x, float& t, float y1, float y2, float y4, float y8,
                                            float f15, float f31) {
#define BVP_LVL(CTRL, Y)                                                           \
  "v_fmac_f32_dpp %0, %0, " Y " " CTRL " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t" \
  "v_fmac_f32_dpp %1, %1, " Y " " CTRL " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
  if constexpr (G == 64) {
    asm volatile("s_nop 1\n\t"
                 BVP_LVL("row_shr:1", "%2") "s_nop 0\n\t"
                 BVP_LVL("row_shr:2", "%3") "s_nop 0\n\t"
                 BVP_LVL("row_shr:4", "%4") "s_nop 0\n\t"
                 BVP_LVL("row_shr:8", "%5") "s_nop 0\n\t"
                 BVP_LVL("row_bcast:15", "%6") "s_nop 0\n\t"
                 BVP_LVL("row_bcast:31", "%7") "s_nop 1"
                 : "+v"(x), "+v"(t)
                 : "v"(y1), "v"(y2), "v"(y4), "v"(y8), "v"(f15), "v"(f31));
  } else {
    asm volatile("s_nop 1\n\t"
                 BVP_LVL("row_shr:1", "%2") "s_nop 0\n\t"
                 BVP_LVL("row_shr:2", "%3") "s_nop 0\n\t"
                 BVP_LVL("row_shr:4", "%4") "s_nop 0\n\t"
                 BVP_LVL("row_shr:8", "%5") "s_nop 0\n\t"
                 BVP_LVL("row_bcast:15", "%6") "s_nop 1"
                 : "+v"(x), "+v"(t)
                 : "v"(y1), "v"(y2), "v"(y4), "v"(y8), "v"(f15));
    (void)f31;
  }
#undef BVP_LVL
}

// I of the next row: dpp_shl1(m) * mx + dpp_shl1(i) * xx, the second DPP
// operand folded into a v_fmac_f32_dpp (s_nop 1 first: the DPP read of m must
// be two wait states after its write).
__device__ __forceinline__ float insert_next(float m, float i, float mx, float xx) {
  float r = dpp<kWaveShl1>(i) * xx;
  asm volatile("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"
               : "+v"(r)
               : "v"(m), "v"(mx));
  return r;
}

// Sum over the 64 lanes, wave-uniform: row_shr scans of each 16-lane row,
// row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2 and 3; lane 63
// then holds the total.  (DPP only: no lane-address registers.)
__device__ __forceinline__ float wtotalf(float v) {
  v += dpp<kRowShr + 1>(v);
  v += dpp<kRowShr + 2>(v);
  v += dpp<kRowShr + 4>(v);
  v += dpp<kRowShr + 8>(v);
  v += dpp<kRowBcast15, 0xa>(v);
  v += dpp<kRowBcast31, 0xc>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum over each G-lane slot, returned to every lane of the slot: row_shr
// scans of each 16-lane row, row_bcast:15 into rows 1 and 3 and (G = 64)
// row_bcast:31 into rows 2 and 3, after which the slot's top lane holds the
// total.
template <int G>
__device__ __forceinline__ float slot_total(float v, int grp) {
  v += dpp<kRowShr + 1>(v);
  v += dpp<kRowShr + 2>(v);
  v += dpp<kRowShr + 4>(v);
  v += dpp<kRowShr + 8>(v);
  v += dpp<kRowBcast15, 0xa>(v);
  if constexpr (G == kBandLanes) {
    v += dpp<kRowBcast31, 0xc>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
  } else {
    const float s0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
    const float s1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
    return grp ? s1 : s0;
  }
}

// log10 of a positive finite float: the exponent exactly, the mantissa's
// log2 in [-1, 0) by v_log_f32 (absolute error below 1e-7).
__device__ __forceinline__ double log10_pos(float x) {
  int e;
  const float m = frexpf(x, &e);
  return (static_cast<double>(e) + static_cast<double>(__log2f(m))) * 0.30102999566398119521;
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

// 0: full matrix, 32, 48 or 64: the banded list the pair goes to (48 only
// with the quad shape and the mid band on).
__device__ __forceinline__ int band_class(const BArgs& a, const Task& tk, int h, bool gcp_const,
                                          int2& b) {
  b = a.band[tk.out0 + h];
  const int hap = tk.hap0 + h;
  const int64_t H = a.hap_off[hap + 1] - a.hap_off[hap];
  const int64_t R = a.read_off[tk.read + 1] - a.read_off[tk.read];
  if (!gcp_const || b.x > b.y || H > kBandHapMax || R > kBandReadMax || b.x < -(1 << 26) ||
      b.y > (1 << 26))
    return 0;
  const int64_t w = static_cast<int64_t>(b.y) - b.x + 1;
  if (w <= kBandNarrow) return kBandNarrow;
  if (w <= kBandMid && a.mid_on && a.wide_pairs == kQuadPairs) return kBandMid;
  return w <= kBandLanes ? kBandLanes : 0;
}

}  // namespace

// Splits every task's pairs into the banded lists and full-matrix tasks.
// One thread per task; appends are reserved once per wavefront.  A task's
// narrow items are written in pairs (padded with an item of hap -1), and
// with the quad shape its wide items and its mid items in fours, the last
// fewer than four of each to the one-pair kernel's list (mid ones as 64-wide
// bands with the same centre).
__global__ __launch_bounds__(256) void band_dispatch_kernel(BArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  Task tk{};
  bool gcp_const = false;
  int n64 = 0, n48 = 0, n32 = 0, nf = 0, nfl = 0;
  if (t < a.n_tasks) {
    tk = a.tasks[t];
    gcp_const = a.gcp_const[tk.read] != 0;
    for (int h = 0; h < tk.nhap; ++h) {
      int2 b;
      const int c = band_class(a, tk, h, gcp_const, b);
      n64 += c == kBandLanes;
      n48 += c == kBandMid;
      n32 += c == kBandNarrow;
      if (c == 0) {
        const int hap = tk.hap0 + h;
        if (a.hap_off[hap + 1] - a.hap_off[hap] <= kStackHmax) ++nf; else ++nfl;
      }
    }
  }
  const int n32p = (n32 + 1) & ~1;
  // the quad shape takes a task's wide pairs four at a time; the rest (fewer
  // than four) fill one more unit, padded, when there are at least
  // quad_fill of them, else go to the one-pair kernel's list.  n64q: the
  // quad slots, padding included.
  const bool quad = a.wide_pairs == kQuadPairs;
  const int rem64 = n64 & (kQuadPairs - 1);
  const int n64q = !quad ? n64 : (rem64 >= a.quad_fill ? n64 + kQuadPairs - rem64 : n64 - rem64);
  const int n48q = n48 & ~(kQuadPairs - 1);  // mid pairs exist only with the quad shape
  const int n64r = max(0, n64 - n64q) + n48 - n48q;
  int t64 = 0, t48 = 0, t64r = 0, t32 = 0, tf = 0, tfl = 0;
  const int o64 = wave_excl_scan(n64q, lane, t64);
  const int o48 = wave_excl_scan(n48q, lane, t48);
  const int o64r = wave_excl_scan(n64r, lane, t64r);
  const int o32 = wave_excl_scan(n32p, lane, t32);
  const int of = wave_excl_scan(nf, lane, tf);
  const int ofl = wave_excl_scan(nfl, lane, tfl);
  int b64 = 0, b48 = 0, b64r = 0, b32 = 0, bf = 0, bfl = 0;
  if (lane == 0) {
    if (t64) b64 = atomicAdd(a.item_count, t64);
    if (t48) b48 = atomicAdd(a.item48_count, t48);
    if (t64r) b64r = atomicAdd(a.item1_count, t64r);
    if (t32) b32 = atomicAdd(a.item32_count, t32);
    if (tf) bf = atomicAdd(a.full_count, tf);
    if (tfl) bfl = atomicAdd(a.full_long_count, tfl);
  }
  b64 = __shfl(b64, 0) + o64;
  b48 = __shfl(b48, 0) + o48;
  b64r = __shfl(b64r, 0) + o64r;
  b32 = __shfl(b32, 0) + o32;
  bf = __shfl(bf, 0) + of;
  bfl = __shfl(bfl, 0) + ofl;
  if (t < a.n_tasks) {
    int k64 = 0, k48 = 0;  // wide and mid pairs of this task so far
    for (int h = 0; h < tk.nhap; ++h) {
      int2 b;
      const int64_t o = tk.out0 + h;
      const int c = band_class(a, tk, h, gcp_const, b);
      if (c) {
        const int w = b.y - b.x + 1;
        const BandItem it{tk.read, tk.hap0 + h, b.x - ((c - w) >> 1), 0, o};
        if (c == kBandLanes) {
          if (k64++ < n64q) a.items[b64++] = it; else a.items1[b64r++] = it;
        } else if (c == kBandMid) {
          if (k48++ < n48q) a.items48[b48++] = it;
          else a.items1[b64r++] = BandItem{tk.read, tk.hap0 + h, b.x - ((kBandLanes - w) >> 1), 0, o};
        } else {
          a.items32[b32++] = it;
        }
      } else if (a.hap_off[tk.hap0 + h + 1] - a.hap_off[tk.hap0 + h] <= kStackHmax) {
        a.full[bf++] = Task{tk.read, tk.hap0 + h, 1, -1, o};
      } else {
        a.full_long[bfl++] = Task{tk.read, tk.hap0 + h, 1, -1, o};
      }
    }
    for (int z = n64; z < n64q; ++z) a.items[b64++] = BandItem{tk.read, -1, 0, 0, -1};
    if (n32 & 1) a.items32[b32] = BandItem{tk.read, -1, 0, 0, -1};
  }
}

// A pair the FP32 band did not accept.  One whose band sum fell below 1e-28
// while still a normal float with the touch test passed (the band holds its
// alignment, its likelihood is beyond FP32's reach): to the FP64 band rescue
// list (a 64-diagonal band with the same centre) while it has room.  Any
// other (the touch test failed, or the sum is not a normal float): the
// full-matrix FP32 list, whose underflow sends it on to the FP64 pass.
__device__ __forceinline__ void band_reject(const BArgs& a, int read, int hap, int o_lo, int width, float S,
                                            float Tm, long long out) {
  if (a.band_rescue != nullptr && S < kMinAccepted && S >= FLT_MIN && Tm <= static_cast<float>(a.tol) * S) {
    const int q = atomicAdd(a.band_rescue_count, 1);
    if (q < a.band_rescue_cap) {
      a.band_rescue[q] = BandItem{read, hap, o_lo - ((kBandLanes - width) >> 1), 0, out};
      return;
    }
  }
  const int q = atomicAdd(a.full_count, 1);
  a.full[q] = Task{read, hap, 1, -2, out};
}

// Banded forward over one list.  G = 64: one pair per wavefront.  G = 32: the
// two pairs 2u, 2u+1 (one read, two haplotypes) in lanes 0-31 and 32-63.
template <int G>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void band_forward_kernel(BArgs a) {
  constexpr int P = kBandLanes / G;
  RowRec* tab = reinterpret_cast<RowRec*>(band_lds);
  const int C = a.sel_cols;
  const int lane = threadIdx.x;
  const int grp = lane / G;     // pair slot of this lane
  const int k = lane % G;       // lane within the slot: diagonal o_lo + k
  // the wide kernel takes the quad shape's leftovers when that shape is on
  const bool rest = G == kBandLanes && a.wide_pairs == kQuadPairs;
  const BandItem* items = G == kBandLanes ? (rest ? a.items1 : a.items) : a.items32;
  const int n_units = G == kBandLanes ? *(rest ? a.item1_count : a.item_count) : (*a.item32_count) / 2;
  const bool left_lane = k < kBandEdge;
  const bool right_lane = k >= G - kBandEdge;
  const bool bottom_lane = k == 0;
  const bool top_lane = k == G - 1;
  const int row16 = lane >> 4;
  int accepted = 0;
  unsigned long long cells = 0;
  // Units are claimed kBandChunk consecutive units (typically the pairs of
  // one read) at a time.  The wavefronts of XCD x take chunks x, x + 8,
  // x + 16, ... from a counter of their own, so each claim is one atomic on
  // one of eight addresses, and a wavefront that starts late (its slot held
  // by a kernel of the other stream) takes fewer chunks instead of delaying
  // the kernel's end.  (One shared counter with one claim per unit
  // serialised at the L2 at ~82 M units/s, slower than the units run.)
  const bool by_xcd = gridDim.x % kXcdCount == 0;
  const int xcd = by_xcd ? static_cast<int>(blockIdx.x % kXcdCount) : 0;
  int* const claim = a.claims + ((G == kBandLanes ? 0 : kXcdCount) + xcd) * kCtrStride;
  for (;;) {
    int c = 0;
    if (lane == 0) c = atomicAdd(claim, 1);
    c = __builtin_amdgcn_readfirstlane(__shfl(c, 0));
    const int chunk = by_xcd ? xcd + kXcdCount * c : c;
    if (chunk >= (n_units + kBandChunk - 1) / kBandChunk) break;
    const int u_end = min(n_units, (chunk + 1) * kBandChunk);
  for (int u = chunk * kBandChunk; u < u_end; ++u) {
    const BandItem w = items[P * u + grp];
    const bool live = w.hap >= 0;
    const int read = __builtin_amdgcn_readfirstlane(w.read);
    int hap = live ? w.hap : items[P * u].hap;  // padding mirrors slot 0
    int o_lo = live ? w.o_lo : items[P * u].o_lo;
    if constexpr (G == kBandLanes) {  // one pair: wave-uniform
      hap = __builtin_amdgcn_readfirstlane(hap);
      o_lo = __builtin_amdgcn_readfirstlane(o_lo);
    }
    const int64_t ro = a.read_off[read];
    const int R = static_cast<int>(a.read_off[read + 1] - ro);
    const int64_t ho = a.hap_off[hap];
    const int H = static_cast<int>(a.hap_off[hap + 1] - ho);
    // Start slide (wide band only): the band starts Pt diagonals higher and
    // steps down one diagonal per row to o_lo, so that a read start explained
    // as a run of insertions (which sweeps down one diagonal per row from
    // above) stays inside; capped where the matrix ends.  A step is held
    // while the top lane it would drop carries more than kBandHold of the
    // row's mass.  off(r) is the offset the schedule asks for on row r.
    constexpr int kSlideG = G == kBandLanes ? kBandSlide : 0;
    const int Pt = min(kSlideG, max(0, (H - 1) - (o_lo + G - 1)));
    auto off = [&](int r) { return o_lo + max(0, Pt + 1 - r); };
    const int wb = o_lo;  // lowest band offset: window slot c is column wb + 1 + c
    _Float16* selL = reinterpret_cast<_Float16*>(band_lds + kTabBytes) + grp * 4 * C;  // [4][C]

    __syncthreads();  // the previous unit's LDS reads are done
    for (int c = k; c < R + G + Pt; c += G) {
      const int j = wb + 1 + c;
      const uint32_t m = (j >= 1 && j <= H) ? a.hmask[ho + j - 1] : 0u;
#pragma unroll
      for (int q = 0; q < 4; ++q) selL[q * C + c] = ((m >> q) & 1u) ? _Float16(1) : _Float16(0);
    }

    // Decay factors of the D scan (constant gap continuation along the read)
    // and the gap-to-match factor g that I and D are carried times.
    const int cq0 = a.gcp[ro] & 127;
    const float yy = a.ph2pr[cq0];
    const float gx = 1.f - a.ph2pr[cq0];
    const double igx = 1.0 / static_cast<double>(gx);
    const float yy2 = yy * yy, yy4 = yy2 * yy2, yy8 = yy4 * yy4, yy16 = yy8 * yy8;
    auto ypow = [&](int e) {  // yy^e for 1 <= e <= 32
      float p = (e & 1) ? yy : 1.f;
      if (e & 2) p *= yy2;
      if (e & 4) p *= yy4;
      if (e & 8) p *= yy8;
      if (e & 16) p *= yy16;
      if (e & 32) p *= yy16 * yy16;
      return p;
    };
    const float f15 = (row16 & 1) ? ypow((lane & 15) + 1) : 0.f;  // distance to lane 15 of the row below
    const float f31 = (row16 >= 2) ? ypow(lane - 31) : 0.f;       // distance to lane 31
    // Edge zones (T = F there): the left one where the band's left
    // neighbour column is inside the matrix, the right one where its right
    // neighbour column is.  Without a slide the rows where each is active
    // are fixed per lane.
    const int ez_from = left_lane ? 2 - o_lo : INT_MAX;
    const int ez_to = right_lane ? H - o_lo - G : INT_MIN;

    // Row 0: M = I = 0, D = K/H for columns 0..H (Context.h:109, template :164-175).
    int o_cur = off(1);
    // this lane's selector slot k + o_cur - wb (the row adds its part)
    const _Float16* sl = selL + (k + o_cur - wb);
    float Mp = 0.f, Ip = 0.f;
    // A read end explained as a run of insertions leaves the band through
    // its bottom lane and sweeps down one diagonal per row; xg (the bottom
    // lane's value counts) carries the I mass that left, continued as pure
    // runs to the last row, and is added to the touch mass.
    float xg = 0.f;
    float Dp = (o_cur + k >= 0 && o_cur + k <= H) ? gx * (kK / static_cast<float>(H)) : 0.f;
    float TMp = 0.f, TIp = 0.f, TDp = 0.f;
    {
      const bool ez = (left_lane && o_cur - 1 >= 0) || (right_lane && o_cur + G <= H);
      if (ez) TDp = Dp;
    }

    for (int i0 = 1; i0 <= R; i0 += kBandLanes) {
      __syncthreads();
      {
        const int r = i0 + lane;
        if (r <= R) {
          const int64_t p = ro + r - 1;
          const uint8_t rc = a.rclass[p];
          const int iq = a.ins[p] & 127, dq = a.del[p] & 127;
          const int lo = iq < dq ? iq : dq, hi = iq < dq ? dq : iq;
          const float e = a.ph2pr[a.q[p] & 127];
          RowRec rr;
          const float pm = 1.f - e;
          rr.pe = (rc & kWildcard) ? pm : e / 3.f;
          rr.dp = pm - rr.pe;
          rr.mm = a.mm[((hi * (hi + 1)) >> 1) + lo];
          rr.mx = gx * a.ph2pr[iq];
          rr.xx = yy;
          rr.my = gx * a.ph2pr[dq];
          rr.selo = static_cast<uint32_t>(2 * ((rc & 3) * C + r - 1));
          rr.pad = 0.f;
          tab[lane] = rr;
        }
      }
      __syncthreads();
      const int n = __builtin_amdgcn_readfirstlane(min(kBandLanes, R - i0 + 1));
      // window slot i - 1 + k + o_cur - wb (in [0, R + G + Pt)): so is the
      // row's part in bytes
      auto sel_at = [&](uint32_t so) {
        return *reinterpret_cast<const _Float16*>(reinterpret_cast<const char*>(sl) + so);
      };
      // Row v of the block with its record, selector offset and selector
      // (read before the row starts).  kSlide: the row may take a step of
      // the start slide (a slide re-reads the selector, and the one read
      // ahead for the next row, nsel at offset nso, if given).  kExit: 0 no
      // exit-run term, 1 the row is in the exit zone, 2 check.
      auto row = [&](auto kSlide, auto kExit, int v, const RowRec& rr, uint32_t so, _Float16 selh,
                     _Float16* nsel, uint32_t nso) {
        const int i = i0 + v;
        if constexpr (decltype(kSlide)::value) {
          if (off(i) < o_cur) {
            // slide down one diagonal: lane k takes lane k - 1's cells and
            // the top lane is dropped, unless it holds more than kBandHold
            // of the row's mass
            const float wl = __builtin_fmaf(Ip + Dp, static_cast<float>(igx), Mp);
            const float wtop = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wl), kBandLanes - 1));
            if (wtop <= kBandHold * wtotalf(wl)) {
              Mp = dpp<kWaveShr1>(Mp);
              Ip = dpp<kWaveShr1>(Ip);
              Dp = dpp<kWaveShr1>(Dp);
              TMp = dpp<kWaveShr1>(TMp);
              TIp = dpp<kWaveShr1>(TIp);
              TDp = dpp<kWaveShr1>(TDp);
              --o_cur;
              --sl;
              selh = sel_at(so);
              if (nsel) *nsel = sel_at(nso);
            }
          }
        }
        const int oc = kSlideG > 0 ? o_cur : o_lo;  // without a slide off(i) = o_lo
        const float sel = static_cast<float>(selh);
        const float prior = __builtin_fmaf(sel, rr.dp, rr.pe);
        bool ez;
        if constexpr (kSlideG > 0) {
          const bool lon = i + oc >= 2, ron = i + oc + G <= H;  // wave-uniform
          ez = (left_lane && lon) || (right_lane && ron);
        } else {
          ez = i >= ez_from || i <= ez_to;
        }
        // the I mass leaving the bottom lane from row i - 1, times g (counted
        // on the last kEdgeRunRows rows; before them a run's factor yy^rows
        // to the last row is below 1e-46)
        constexpr int kX = decltype(kExit)::value;
        if (kX == 1 || (kX == 2 && i > R - kEdgeRunRows)) {
          const int c0 = i - 1 + oc;  // the bottom lane's column on row i - 1
          const float xe = __builtin_fmaf(Mp, rr.mx, Ip * rr.xx);
          xg = __builtin_fmaf(xg, yy, (c0 >= 1 && c0 <= H) ? xe : 0.f);
        }

        // forward (I and D are carried times g)
        const float Mn = __builtin_fmaf(Mp, rr.mm, Ip + Dp) * prior;
        float In = insert_next(Mp, Ip, rr.mx, rr.xx);
        float x = dpp<kWaveShr1>(Mn) * rr.my;
        if constexpr (G != kBandLanes) {
          // the wavefront shifts cross the slot boundary: cut them
          if (top_lane) In = 0.f;
          if (bottom_lane) x = 0.f;
        }

        // touch mass
        float TMn = __builtin_fmaf(TMp, rr.mm, TIp + TDp) * prior;
        float TIn = insert_next(TMp, TIp, rr.mx, rr.xx);
        if (ez) {
          TMn = Mn;
          TIn = In;
        }
        float tx = dpp<kWaveShr1>(TMn) * rr.my;
        if constexpr (G != kBandLanes) {
          if (top_lane) TIn = 0.f;
          if (bottom_lane) tx = 0.f;
        }
        // the D scans of both DPs (x becomes Dn, tx becomes TDn)
        decay_scan2<G>(x, tx, yy, yy2, yy4, yy8, f15, f31);
        const float Dn = x;
        float TDn = tx;
        if (ez) TDn = Dn;

        Mp = Mn;
        Ip = In;
        Dp = Dn;
        TMp = TMn;
        TIp = TIn;
        TDp = TDn;
      };
      // Rows [v, end) two per step, both rows' operands read at its start.
      auto rows = [&](auto kSlide, auto kExit, int v, int end) {
        for (; v + 1 < end; v += 2) {
          const RowRec ra = tab[v], rb = tab[v + 1];
          const uint32_t sa = ra.selo, sb = rb.selo;
          const _Float16 ha = sel_at(sa);
          _Float16 hb = sel_at(sb);
          row(kSlide, kExit, v, ra, sa, ha, &hb, sb);
          row(kSlide, kExit, v + 1, rb, sb, hb, nullptr, 0u);
        }
        if (v < end) {
          const RowRec ra = tab[v];
          row(kSlide, kExit, v, ra, ra.selo, sel_at(ra.selo), nullptr, 0u);
        }
      };
      using T_ = std::true_type;
      using F_ = std::false_type;
      using X0 = std::integral_constant<int, 0>;
      using X1 = std::integral_constant<int, 1>;
      using X2 = std::integral_constant<int, 2>;
      // Phases, so that the common rows carry no branch: the start slide
      // (rows up to 1 + the rows the band still stands above o_lo), the rows
      // before the exit zone, the rows in it.
      int v = 0;
      if constexpr (kSlideG > 0) {
        while (v < n && (i0 + v == 1 || o_cur > o_lo)) {
          const RowRec ra = tab[v];
          row(T_{}, X2{}, v, ra, ra.selo, sel_at(ra.selo), nullptr, 0u);
          ++v;
        }
      }
      const int vx = min(n, max(v, R - kEdgeRunRows + 1 - i0));  // first row of the exit zone
      rows(F_{}, X0{}, v, vx);
      rows(F_{}, X1{}, vx, n);
    }
    // S = sum over the last row of M + I (template :328-343), columns 1..H;
    // I is carried times g.
    const int j = R + (kSlideG > 0 ? o_cur : o_lo) + k;
    const bool in = j >= 1 && j <= H;
    const float ig = static_cast<float>(igx);
    const float S = slot_total<G>(in ? __builtin_fmaf(Ip, ig, Mp) : 0.f, grp);
    const float Tm = slot_total<G>((in ? __builtin_fmaf(TIp, ig, TMp) : 0.f) + (bottom_lane ? xg * ig : 0.f), grp);
    if (live) cells += static_cast<unsigned long long>(R) * G;
    if (k == 0 && live) {
      if (S >= kMinAccepted && Tm <= static_cast<float>(a.tol) * S) {
        a.out[w.out] = log10_pos(S) - kLog10K;
        a.form[w.out] = kFormBandFP32;
        ++accepted;
      } else {
        band_reject(a, read, hap, o_lo, G, S, Tm, w.out);
      }
    }
  }
  }
  // per-lane tallies live in the slot leaders (lanes 0 and 32)
  int acc_w = accepted;
  unsigned long long cells_w = cells;
  if constexpr (P == 2) {
    acc_w += __shfl_xor(acc_w, 32);
    cells_w = __shfl(cells, 0) + __shfl(cells, 32);
  }
  if (lane == 0) {
    if (acc_w) atomicAdd(a.accepted, acc_w);
    if (cells_w && a.cells_done) atomicAdd(a.cells_done, cells_w);
  }
}

// ---- the wide band, quad shape: four pairs of one read per wavefront --------
//
// Pair slot p (0..3) owns the 16-lane DPP row p of the wavefront; lane k of
// that row holds the four diagonals o + 4k + c (c = 0..3) of its pair's
// 64-diagonal band, one read row per step as above.  What changes is the D
// recurrence along a row, D_d = my*M_{d-1} + yy*D_{d-1}: each lane first runs
// it along its own four cells (a chain of three FMAs from the carry-in of the
// lane below), the lanes' top values are then combined by a decayed scan over
// the 16 lanes of the row (four row_shr levels with factors yy^4, yy^8, yy^16,
// yy^32; no row_bcast: the pairs of a wavefront never exchange data), and one
// FMA per cell adds the carry from the lanes below.  D is carried divided by
// its row's my (constant along a row), so the chain takes M itself and no
// multiply per cell: about 2.25 VALU per cell for D against 6 in the one-pair
// kernel, and I crosses a lane only for a lane's top cell.  A
// lane is exactly one edge zone (kQuadCells == kBandEdge): lane 0 of a row
// holds the band's 4 lowest diagonals, lane 15 its 4 highest.
//
// The template argument kC is the number of cells per lane: 4 for the wide
// band, 3 for the 48-diagonal mid band (its own list), whose edge zones are
// then one lane of three diagonals per side; the D scan's factors become
// yy^3 .. yy^24 and the carry from below enters cell 1 only.
//
// Touch mass with fewer overrides.  T equals F on active edge-zone cells.
// When a zone was active on the row before, the cells of its lane compute
// T = F by themselves (the same operations on the same inputs), except for
// the inputs that come from a lane outside the zone: the top cell's I of
// lane 0 (from lane 1) and the D values of lane 15 (carried up from below).
// A row after the start slide therefore overrides those five values only.
// The left zone is active on a suffix of the rows (from where the band's
// left edge column enters the matrix) and the right zone on a prefix, so
// the first row, the slide rows and the row where a left zone turns on take
// every override, exactly as the one-pair kernel does; T is then the same,
// bit for bit, as with every override on every row.
namespace {

constexpr int kRowShr1 = kRowShr + 1;
constexpr int kRowShl1 = 0x101;  // row_shl:1: lane k <- lane k+1 inside a 16-lane row, lane 15 <- 0
constexpr int kQuadChunk = 2;    // quad units per claim

// The two lane-crossing FMAs of a step that read the row before or the
// products just formed, for F and T: the top cell's I, r += dpp_shl1(m0)*mx
// (r arrives as dpp_shl1(i0)*xx; lane 15 gets 0, the diagonal above the band
// being outside it), and the first link of the D chain, y0 += dpp_shr1(y3)*yy
// (lane 0 has no lane below).  s_nop 1 first: a DPP operand may have been
// written by the instruction before.
__device__ __forceinline__ void quad_cross(float& r, float& tr, float& y0, float& ty0, float m0, float tm0,
                                           float y3, float ty3, float mx, float yy) {
  asm volatile(
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %4, %8 row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %1, %5, %8 row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %2, %6, %9 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_fmac_f32_dpp %3, %7, %9 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "+v"(r), "+v"(tr), "+v"(y0), "+v"(ty0)
      : "v"(m0), "v"(tm0), "v"(y3), "v"(ty3), "v"(mx), "v"(yy));
}

// The decayed scan of the lanes' top D values over each 16-lane row, for F
// (x) and T (tx): x_k += yy^4 x_{k-1}, += yy^8 x_{k-2}, += yy^16 x_{k-4},
// += yy^32 x_{k-8}; then the carry from the lanes below into cells 1 and 2,
// d1 += yy^2 x_{k-1}, d2 += yy^3 x_{k-1} (lane 0 has none).  F and T
// alternate inside each level, with one s_nop 0 between levels, so every DPP
// read of a level's result comes two wait states after its write.
#define BVQ_LVL(CTRL, Y)                                                             \
  "v_fmac_f32_dpp %0, %0, " Y " " CTRL " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t" \
  "v_fmac_f32_dpp %1, %1, " Y " " CTRL " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
__device__ __forceinline__ void quad_scan(float& x, float& tx, float& d1, float& td1, float& d2, float& td2,
                                          float y4, float y8, float y16, float y32, float y2, float y3) {
  asm volatile("s_nop 1\n\t"
               BVQ_LVL("row_shr:1", "%6") "s_nop 0\n\t"
               BVQ_LVL("row_shr:2", "%7") "s_nop 0\n\t"
               BVQ_LVL("row_shr:4", "%8") "s_nop 0\n\t"
               BVQ_LVL("row_shr:8", "%9") "s_nop 1\n\t"
               "v_fmac_f32_dpp %2, %0, %10 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
               "v_fmac_f32_dpp %3, %1, %10 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
               "v_fmac_f32_dpp %4, %0, %11 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
               "v_fmac_f32_dpp %5, %1, %11 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"
               : "+v"(x), "+v"(tx), "+v"(d1), "+v"(td1), "+v"(d2), "+v"(td2)
               : "v"(y4), "v"(y8), "v"(y16), "v"(y32), "v"(y2), "v"(y3));
}

// The same for the mid band (three cells per lane): factors yy^3, yy^6,
// yy^12, yy^24, and the carry from the lanes below into cell 1 only,
// d1 += yy^2 x_{k-1} (cell 2 is the lane's top, x itself).
__device__ __forceinline__ void quad_scan3(float& x, float& tx, float& d1, float& td1, float y3, float y6,
                                           float y12, float y24, float y2) {
  asm volatile("s_nop 1\n\t"
               BVQ_LVL("row_shr:1", "%4") "s_nop 0\n\t"
               BVQ_LVL("row_shr:2", "%5") "s_nop 0\n\t"
               BVQ_LVL("row_shr:4", "%6") "s_nop 0\n\t"
               BVQ_LVL("row_shr:8", "%7") "s_nop 1\n\t"
               "v_fmac_f32_dpp %2, %0, %8 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
               "v_fmac_f32_dpp %3, %1, %8 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"
               : "+v"(x), "+v"(tx), "+v"(d1), "+v"(td1)
               : "v"(y3), "v"(y6), "v"(y12), "v"(y24), "v"(y2));
}
#undef BVQ_LVL

// Sum over each 16-lane row; lane 15 of the row holds the total.
__device__ __forceinline__ float row_total(float v) {
  v += dpp<kRowShr + 1>(v);
  v += dpp<kRowShr + 2>(v);
  v += dpp<kRowShr + 4>(v);
  v += dpp<kRowShr + 8>(v);
  return v;
}

// Selector f16 units per pair slot: four class rows of sel_cols, plus two
// (4 bytes) so that the slots' windows start on different LDS banks.
__device__ __host__ constexpr int quad_sel_stride(int sel_cols) { return 4 * sel_cols + 2; }

}  // namespace

template <int kC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void band_quad_kernel(BArgs a) {
  static_assert(kC == kQuadCells || kC == kMidCells, "the wide or the mid band");
  constexpr int W = kQuadLanes * kC;  // the band's diagonals
  const BandItem* const items = kC == kQuadCells ? a.items : a.items48;
  RowRec* tab = reinterpret_cast<RowRec*>(band_lds);
  const int C = a.sel_cols;
  const int lane = threadIdx.x;
  const int p = lane >> 4;  // pair slot: the DPP row
  const int k = lane & 15;  // lane in the row: diagonals o + kC k .. o + kC k + kC - 1
  const bool left_lane = k == 0;
  const bool right_lane = k == kQuadLanes - 1;
  const int n_units = *(kC == kQuadCells ? a.item_count : a.item48_count) / kQuadPairs;
  const int n_chunks = (n_units + kQuadChunk - 1) / kQuadChunk;
  int accepted = 0;
  unsigned long long cells = 0;
  // chunks claimed from a counter per XCD, as in band_forward_kernel
  const bool by_xcd = gridDim.x % kXcdCount == 0;
  const int xcd = by_xcd ? static_cast<int>(blockIdx.x % kXcdCount) : 0;
  int* const claim = (kC == kQuadCells ? a.quad_claims : a.quad48_claims) + xcd * kCtrStride;
  _Float16* const selL = reinterpret_cast<_Float16*>(band_lds + kTabBytes) + p * quad_sel_stride(C);  // [4][C]
  for (;;) {
    int cl = 0;
    if (lane == 0) cl = atomicAdd(claim, 1);
    cl = __builtin_amdgcn_readfirstlane(__shfl(cl, 0));
    const int chunk = by_xcd ? xcd + kXcdCount * cl : cl;
    if (chunk >= n_chunks) break;
    const int u_end = min(n_units, (chunk + 1) * kQuadChunk);
    for (int u = chunk * kQuadChunk; u < u_end; ++u) {
      const BandItem w = items[kQuadPairs * u + p];
      const BandItem w0 = items[kQuadPairs * u];  // slot 0 of a unit is never padding
      const bool live = w.hap >= 0;
      const int read = __builtin_amdgcn_readfirstlane(w0.read);
      const int hap = live ? w.hap : w0.hap;  // padding mirrors slot 0
      const int o_lo = live ? w.o_lo : w0.o_lo;
      const int64_t ro = a.read_off[read];
      const int R = static_cast<int>(a.read_off[read + 1] - ro);
      const int64_t ho = a.hap_off[hap];
      const int H = static_cast<int>(a.hap_off[hap + 1] - ho);
      // start slide, as in band_forward_kernel<64>: off(r) is the offset the
      // schedule asks for on row r
      const int Pt = min(kBandSlide, max(0, (H - 1) - (o_lo + W - 1)));
      auto off = [&](int r) { return o_lo + max(0, Pt + 1 - r); };
      const int wb = o_lo;  // window slot s is haplotype column wb + 1 + s

      __syncthreads();  // the previous unit's LDS reads are done
      for (int s = k; s < R + W + Pt; s += kQuadLanes) {
        const int j = wb + 1 + s;
        const uint32_t m = (j >= 1 && j <= H) ? a.hmask[ho + j - 1] : 0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) selL[q * C + s] = ((m >> q) & 1u) ? _Float16(1) : _Float16(0);
      }

      const int cq0 = a.gcp[ro] & 127;
      const float yy = a.ph2pr[cq0];
      const float gx = 1.f - a.ph2pr[cq0];
      const double igx = 1.0 / static_cast<double>(gx);
      // decay factors of the D scan over lanes: yy^kC, yy^2kC, yy^4kC, yy^8kC
      const float yy2 = yy * yy, yy3 = yy2 * yy, yy4 = yy2 * yy2;
      const float yc = kC == kQuadCells ? yy4 : yy3;
      const float yc2 = yc * yc, yc4 = yc2 * yc2, yc8 = yc4 * yc4;

      // Row 0: M = I = 0, D = K/H for columns 0..H (times g); T = F on the
      // edge zones active there.
      int o_cur = off(1);
      const _Float16* sl = selL + (kC * k + o_cur - wb);  // this lane's cell 0 slot, less the row's part
      // D is carried divided by its row's my (D' = D / my, constant along a
      // row): the chain along a row is then D'_d = M_{d-1} + yy * D'_{d-1},
      // with no multiply per cell, and the next row takes my * D' in one FMA.
      // Row 0's D is kept as is (myp = 1).
      float M[kC], I[kC], D[kC], TM[kC], TI[kC], TD[kC];
      float myp = 1.f;
      {
        const float d0 = gx * (kK / static_cast<float>(H));
        const bool ez0 = (left_lane && o_cur - 1 >= 0) || (right_lane && o_cur + W <= H);
#pragma unroll
        for (int c = 0; c < kC; ++c) {
          const int d = o_cur + kC * k + c;
          M[c] = I[c] = TM[c] = TI[c] = 0.f;
          D[c] = (d >= 0 && d <= H) ? d0 : 0.f;
          TD[c] = ez0 ? D[c] : 0.f;
        }
      }
      float xg = 0.f;  // exit runs through the bottom lane, as in band_forward_kernel

      for (int i0 = 1; i0 <= R; i0 += kBandLanes) {
        __syncthreads();
        {
          const int r = i0 + lane;
          if (r <= R) {
            const int64_t q = ro + r - 1;
            const uint8_t rc = a.rclass[q];
            const int iq = a.ins[q] & 127, dq = a.del[q] & 127;
            const int lo = iq < dq ? iq : dq, hi = iq < dq ? dq : iq;
            const float e = a.ph2pr[a.q[q] & 127];
            RowRec rr;
            const float pm = 1.f - e;
            rr.pe = (rc & kWildcard) ? pm : e / 3.f;
            rr.dp = pm - rr.pe;
            rr.mm = a.mm[((hi * (hi + 1)) >> 1) + lo];
            rr.mx = gx * a.ph2pr[iq];
            rr.xx = yy;
            rr.my = gx * a.ph2pr[dq];
            rr.selo = static_cast<uint32_t>(2 * ((rc & 3) * C + r - 1));
            rr.pad = 0.f;
            tab[lane] = rr;
          }
        }
        __syncthreads();
        const int n = __builtin_amdgcn_readfirstlane(min(kBandLanes, R - i0 + 1));

        // Row v of the block.  kFull: every edge-zone override (else the
        // five of a steady row, given ez_from / ez_to); kExit: 0 no exit-run
        // term, 1 the row is in the exit zone, 2 check.
        int ez_from = INT_MAX, ez_to = INT_MIN;
        auto row = [&](auto kFull, auto kExit, int v) {
          constexpr bool full = decltype(kFull)::value;
          constexpr int kX = decltype(kExit)::value;
          const int i = i0 + v;
          const RowRec rr = tab[v];
          const char* sb = reinterpret_cast<const char*>(sl) + rr.selo;
          float pr[kC];
#pragma unroll
          for (int c = 0; c < kC; ++c)
            pr[c] = __builtin_fmaf(static_cast<float>(*reinterpret_cast<const _Float16*>(sb + 2 * c)), rr.dp, rr.pe);
          if (kX == 1 || (kX == 2 && i > R - kEdgeRunRows)) {
            const int c0 = i - 1 + o_cur;  // the bottom cell's column on row i - 1
            const float xe = __builtin_fmaf(M[0], rr.mx, I[0] * rr.xx);
            xg = __builtin_fmaf(xg, yy, (c0 >= 1 && c0 <= H) ? xe : 0.f);
          }
          bool ez = false;
          if constexpr (full) {
            const bool lon = i + o_cur >= 2, ron = i + o_cur + W <= H;
            ez = (left_lane && lon) || (right_lane && ron);
          }
          float Mn[kC], In[kC], TMn[kC], TIn[kC], Y[kC], TY[kC];
#pragma unroll
          for (int c = 0; c < kC; ++c) {
            Mn[c] = __builtin_fmaf(M[c], rr.mm, __builtin_fmaf(D[c], myp, I[c])) * pr[c];
            TMn[c] = __builtin_fmaf(TM[c], rr.mm, __builtin_fmaf(TD[c], myp, TI[c])) * pr[c];
          }
#pragma unroll
          for (int c = 0; c < kC - 1; ++c) {
            In[c] = __builtin_fmaf(M[c + 1], rr.mx, I[c + 1] * rr.xx);
            TIn[c] = __builtin_fmaf(TM[c + 1], rr.mx, TI[c + 1] * rr.xx);
          }
          In[kC - 1] = dpp<kRowShl1>(I[0]) * rr.xx;
          TIn[kC - 1] = dpp<kRowShl1>(TI[0]) * rr.xx;
          if constexpr (full) {
#pragma unroll
            for (int c = 0; c < kC; ++c)
              if (ez) TMn[c] = Mn[c];
          }
#pragma unroll
          for (int c = 0; c < kC; ++c) {
            Y[c] = Mn[c];
            TY[c] = TMn[c];
          }
          myp = rr.my;
          quad_cross(In[kC - 1], TIn[kC - 1], Y[0], TY[0], M[0], TM[0], Y[kC - 1], TY[kC - 1], rr.mx, yy);
          // D chain along the lane: cell 1 = Y0 + yy*carry (done above), then
          // the cells above; the lane's top value x goes into the scan
          float Dn[kC], TDn[kC];
          float x, tx;
          if constexpr (kC == 4) {
            float d1 = Y[0], td1 = TY[0];
            float d2 = __builtin_fmaf(d1, yy, Y[1]), td2 = __builtin_fmaf(td1, yy, TY[1]);
            x = __builtin_fmaf(d2, yy, Y[2]);
            tx = __builtin_fmaf(td2, yy, TY[2]);
            quad_scan(x, tx, d1, td1, d2, td2, yc, yc2, yc4, yc8, yy2, yy3);
            Dn[1] = d1;
            Dn[2] = d2;
            TDn[1] = td1;
            TDn[2] = td2;
          } else {
            float d1 = Y[0], td1 = TY[0];
            x = __builtin_fmaf(d1, yy, Y[1]);
            tx = __builtin_fmaf(td1, yy, TY[1]);
            quad_scan3(x, tx, d1, td1, yc, yc2, yc4, yc8, yy2);
            Dn[1] = d1;
            TDn[1] = td1;
          }
          Dn[kC - 1] = x;
          TDn[kC - 1] = tx;
          // cell 0 of the lane above: Y[kC - 1] + yy * (this lane's top D)
          const float z = __builtin_fmaf(x, yy, Y[kC - 1]), tz = __builtin_fmaf(tx, yy, TY[kC - 1]);
          Dn[0] = dpp<kRowShr1>(z);
          TDn[0] = dpp<kRowShr1>(tz);
          if constexpr (full) {
            if (ez) {
#pragma unroll
              for (int c = 0; c < kC; ++c) {
                TIn[c] = In[c];
                TDn[c] = Dn[c];
              }
            }
          } else {
            if (i >= ez_from) TIn[kC - 1] = In[kC - 1];  // left lanes whose zone is on
            if (i <= ez_to) {                            // right lanes whose zone is on
#pragma unroll
              for (int c = 0; c < kC; ++c) TDn[c] = Dn[c];
            }
          }
#pragma unroll
          for (int c = 0; c < kC; ++c) {
            M[c] = Mn[c];
            I[c] = In[c];
            TM[c] = TMn[c];
            TI[c] = TIn[c];
            D[c] = Dn[c];
            TD[c] = TDn[c];
          }
        };
        using T_ = std::true_type;
        using F_ = std::false_type;
        using X0 = std::integral_constant<int, 0>;
        using X1 = std::integral_constant<int, 1>;
        using X2 = std::integral_constant<int, 2>;

        // The first row and the start slide: every override, and a pair
        // whose schedule asks for it moves down one diagonal (lane k takes
        // the cells one diagonal below, the top cell is dropped) unless the
        // top cell holds more than kBandHold of the pair's row mass.
        int v = 0;
        while (v < n && (i0 + v == 1 || __builtin_amdgcn_ballot_w64(o_cur > o_lo) != 0)) {
          const int i = i0 + v;
          // straight-line code (no branch around the test or the move):
          // conditional state changes make the compiler copy the whole
          // state at the loop's end
          const bool want = off(i) < o_cur;
          const float ig = static_cast<float>(igx);
          float wl = 0.f;
#pragma unroll
          for (int c = 0; c < kC; ++c) wl += __builtin_fmaf(__builtin_fmaf(D[c], myp, I[c]), ig, M[c]);
          const float wtop = __builtin_fmaf(__builtin_fmaf(D[kC - 1], myp, I[kC - 1]), ig, M[kC - 1]);
          const float tot = __shfl(row_total(wl), lane | 15);
          const float top = __shfl(wtop, lane | 15);
          const bool go = want && top <= kBandHold * tot;
          auto shift = [&](float* s) {
            const float below = dpp<kRowShr1>(s[kC - 1]);
#pragma unroll
            for (int c = kC - 1; c > 0; --c) s[c] = go ? s[c - 1] : s[c];
            s[0] = go ? below : s[0];
          };
          shift(M);
          shift(I);
          shift(D);
          shift(TM);
          shift(TI);
          shift(TD);
          o_cur -= go ? 1 : 0;
          sl -= go ? 1 : 0;
          row(T_{}, X2{}, v);
          ++v;
        }
        // Steady rows: the overrides of a zone that was on the row before;
        // full rows where a left zone turns on (row 2 - o of its pair).
        ez_from = left_lane ? 2 - o_cur : INT_MAX;
        ez_to = right_lane ? H - o_cur - W : INT_MIN;
        const int act0 = __builtin_amdgcn_readlane(2 - o_cur, 0);
        const int act1 = __builtin_amdgcn_readlane(2 - o_cur, 16);
        const int act2 = __builtin_amdgcn_readlane(2 - o_cur, 32);
        const int act3 = __builtin_amdgcn_readlane(2 - o_cur, 48);
        auto steady = [&](auto kExit, int vb, int ve) {
          while (vb < ve) {
            const int i = i0 + vb;
            int na = INT_MAX;  // the next row at or after i where a left zone turns on
            if (act0 >= i) na = min(na, act0);
            if (act1 >= i) na = min(na, act1);
            if (act2 >= i) na = min(na, act2);
            if (act3 >= i) na = min(na, act3);
            const int stop = static_cast<int>(min(static_cast<int64_t>(ve), static_cast<int64_t>(na) - i0));
            // two rows per iteration: the state returns to the same registers
            for (; vb + 1 < stop; vb += 2) {
              row(F_{}, kExit, vb);
              row(F_{}, kExit, vb + 1);
            }
            if (vb < stop) {
              row(F_{}, kExit, vb);
              ++vb;
            }
            if (vb < ve) {
              row(T_{}, X2{}, vb);
              ++vb;
            }
          }
        };
        const int vx = min(n, max(v, R - kEdgeRunRows + 1 - i0));  // first row of the exit zone
        steady(X0{}, v, vx);
        steady(X1{}, vx, n);
      }
      // S = sum over the last row of M + I, columns 1..H (I carried times g);
      // lane 15 of the row gets its pair's totals.
      const float ig = static_cast<float>(igx);
      float sS = 0.f, sT = 0.f;
#pragma unroll
      for (int c = 0; c < kC; ++c) {
        const int j = R + o_cur + kC * k + c;
        if (j >= 1 && j <= H) {
          sS += __builtin_fmaf(I[c], ig, M[c]);
          sT += __builtin_fmaf(TI[c], ig, TM[c]);
        }
      }
      if (left_lane) sT += xg * ig;
      const float S = row_total(sS);
      const float Tm = row_total(sT);
      if (right_lane && live) {
        cells += static_cast<unsigned long long>(R) * W;
        if (S >= kMinAccepted && Tm <= static_cast<float>(a.tol) * S) {
          a.out[w.out] = log10_pos(S) - kLog10K;
          a.form[w.out] = kFormBandFP32;
          ++accepted;
        } else {
          band_reject(a, read, hap, o_lo, W, S, Tm, w.out);
        }
      }
    }
  }
  if (right_lane) {
    if (accepted) atomicAdd(a.accepted, accepted);
    if (cells && a.cells_done) atomicAdd(a.cells_done, cells);
  }
}

// ---- long reads: FP64, a band that follows the alignment -------------------
//
// Reads of 1024 bases or more run in FP64 (their FP32 sums underflow).  Their
// alignment drifts across diagonals (indels), so a fixed band would have to
// be as wide as the drift; instead the 64-diagonal band moves.  Before every
// 16th row the state (row i - 1) is summarised by the centroid c of the lane
// masses M + I + D, and the band moves by s = clamp(floor(c - 31.5 + 0.5),
// -4, 4) diagonals (the start band, from the seeds, leaves 16 extra
// diagonals above the alignment for a read start explained as a run of
// insertions, which sweeps down one diagonal per row): lane k takes lane
// k + s's cells, the lanes that fall out
// are dropped and the new lanes start at zero.  At most 4 = kBandEdge lanes
// go per move, all inside the edge zone whose touch mass T the kernel carries
// as in the FP32 kernel above.  A move is held when the lanes it would drop
// carry more than kBandHold of the row's forward mass (a cut mode can still
// have a large future; forward mass that small rarely does).  The touch mass
// also counts the insertion runs that leave through the bottom lane, as in
// the FP32 kernel.  A pair is accepted when T/S <= tol; otherwise, or when its
// likelihood is within reach of FP32 (the reference would keep an FP32
// result), it goes to the full-matrix FP64 pass.
//
// Precision: the DP starts from D = 1/H (the reference's K is a power of two
// and cancels), and the state is rescaled by a power of two every 64 rows
// (FP64) or 16 rows (FP32), which is exact; log10 L = log10 S + (removed
// exponent) * log10 2.  The FP32 form computes the same recurrence with the
// FP32 tables and the short-read kernels' DPP-folded I and D scans.
//
// Per 64-row block the row records (doubles) and a 192-column window of the
// haplotype's class masks are staged in LDS: with at most three moves of 4
// inside a block, row i and lane k read column i + o + k of the window
// [i0 + o0 - 12, i0 + o0 + 180).
namespace {

constexpr int kLongSlack = (kBandLanes / kLongEvery - 1) * kLongMove;  // moves inside a block
constexpr int kLongWin = 192;
static_assert(2 * kBandLanes + 2 * kLongSlack <= kLongWin, "long-band window too small");
static_assert(kLongMove <= kBandEdge, "a move may only drop edge-zone lanes");
constexpr double kLog10Two = 0.30102999566398119521;

template <typename T>
struct alignas(16) RowRecL {
  T dp;  // pm - pe
  T pe;  // e / 3 (pm for a wildcard read base)
  T mm;
  T g;   // 1 - ph2pr[gcp]: insert/delete -> match
  T mx;
  T xx;
  T my;
  uint32_t cls;
  uint32_t pad;
};

template <typename T>
struct LongLds {
  RowRecL<T> tab[kBandLanes];
  uint8_t hm[kLongWin];
};

template <int CTRL>
__device__ __forceinline__ double dpp64(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b)), CTRL, 0xf, 0xf,
                                             true);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b >> 32)), CTRL, 0xf,
                                             0xf, true);
  return __longlong_as_double(static_cast<long long>(
      (static_cast<unsigned long long>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo)));
}

// decay_scan<64> in FP64.
__device__ __forceinline__ double decay_scan64(double x, double y1, double y2, double y4, double y8,
                                               double f15, double f31) {
  x = fma(dpp64<kRowShr + 1>(x), y1, x);
  x = fma(dpp64<kRowShr + 2>(x), y2, x);
  x = fma(dpp64<kRowShr + 4>(x), y4, x);
  x = fma(dpp64<kRowShr + 8>(x), y8, x);
  x = fma(dpp64<kRowBcast15>(x), f15, x);
  x = fma(dpp64<kRowBcast31>(x), f31, x);
  return x;
}

template <typename T>
__device__ __forceinline__ T wmax(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

template <typename T>
__device__ __forceinline__ T wsumT(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Lane k takes lane k + s's value; lanes past the band edge take 0.
template <typename T>
__device__ __forceinline__ T take(T v, int k, int s) {
  const int src = k + s;
  const T x = __shfl(v, src & 63);
  return (src >= 0 && src < kBandLanes) ? x : T(0);
}

}  // namespace

// Splits the long-read tasks (one pair each) into the moving-band list and
// the full-matrix FP64 list.
__global__ __launch_bounds__(256) void band_dispatch_long_kernel(BArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_tasks64) return;
  const Task tk = a.tasks64[t];
  const int2 b = a.band[tk.out0];
  const int64_t w = static_cast<int64_t>(b.y) - b.x + 1;
  const bool ok = a.long_on && a.gcp_const[tk.read] != 0 && w >= 1 && w <= kBandLanes &&
                  b.x >= -(1 << 26) && b.y <= (1 << 26);
  if (ok) {
    const int q = atomicAdd(a.long_count, 1);
    a.long_items[q] = BandItem{tk.read, tk.hap0, b.x - ((kBandLanes - static_cast<int>(w)) >> 1), 0, tk.out0};
  } else {
    const int q = atomicAdd(a.long_full_count, 1);
    a.long_full[q] = tk;
  }
}

template <typename T>
__global__ __launch_bounds__(64) void band_long_kernel(BArgs a) {
  constexpr bool kF32 = sizeof(T) == 4;
  __shared__ LongLds<T> s;
  const T* const ph2pr = kF32 ? reinterpret_cast<const T*>(a.ph2pr) : reinterpret_cast<const T*>(a.ph2pr_d);
  const T* const mmt = kF32 ? reinterpret_cast<const T*>(a.mm) : reinterpret_cast<const T*>(a.mm_d);
  const int k = threadIdx.x;
  const bool left_lane = k < kBandEdge;
  const bool right_lane = k >= kBandLanes - kBandEdge;
  const int row16 = k >> 4;
  const int n_items = min(*a.long_count, a.long_cap);
  int accepted = 0;
  unsigned long long cells = 0;
  for (;;) {
    int u = 0;
    if (k == 0) u = atomicAdd(a.long_queue, 1);
    u = __builtin_amdgcn_readfirstlane(__shfl(u, 0));
    if (u >= n_items) break;
    const BandItem w = a.long_items[u];
    const int read = __builtin_amdgcn_readfirstlane(w.read);
    const int hap = __builtin_amdgcn_readfirstlane(w.hap);
    int o = __builtin_amdgcn_readfirstlane(w.o_lo);  // lane k holds diagonal o + k
    const int64_t ro = a.read_off[read];
    const int R = static_cast<int>(a.read_off[read + 1] - ro);
    const int64_t ho = a.hap_off[hap];
    const int H = static_cast<int>(a.hap_off[hap + 1] - ho);

    const T yy = ph2pr[a.gcp[ro] & 127];
    const T yy2 = yy * yy, yy4 = yy2 * yy2, yy8 = yy4 * yy4;
    T f15 = (row16 & 1) ? T(1) : T(0), f31 = (row16 >= 2) ? T(1) : T(0);
    for (int m = 0; m <= (k & 15); ++m) f15 *= yy;
    for (int m = 31; m < k; ++m) f31 *= yy;
    int ez_from = 2 - o, ez_to = H - o - kBandLanes;

    // Row 0: D = 1/H on columns 0..H.
    T Mp = 0, Ip = 0;
    T Dp = (o + k >= 0 && o + k <= H) ? T(1) / static_cast<T>(H) : T(0);
    T TMp = 0, TIp = 0, TDp = 0;
    if ((left_lane && o - 1 >= 0) || (right_lane && o + kBandLanes <= H)) TDp = Dp;
    int lg = 0;  // exponent removed by rescaling
    int jb = 0;  // haplotype column of window slot 0
    // a read end explained as insertions, leaving through the bottom lane
    // (as in band_forward_kernel)
    T xg = 0;

    for (int i0 = 1; i0 <= R; i0 += kBandLanes) {
      const int n = min(kBandLanes, R - i0 + 1);
      for (int v0 = 0; v0 < n; v0 += kLongEvery) {
        const int i = i0 + v0;  // next row; the state holds row i - 1
        if (i > 1) {
          // Lanes past the haplotype end hold values that never feed
          // in-matrix cells; clear them before the row statistics.
          if ((i - 1) + o + k > H) Mp = Ip = Dp = TMp = TIp = TDp = T(0);
          // the move statistics in FP64 in either precision
          const double wgt = static_cast<double>(Mp) + static_cast<double>(Ip) + static_cast<double>(Dp);
          const double W = wsum(wgt);
          const double C = wsum(static_cast<double>(k) * wgt);
          int sft = 0;
          if (W > 0.0) {
            sft = static_cast<int>(floor(C / W - 0.5 * (kBandLanes - 1) + 0.5));
            sft = min(kLongMove, max(-kLongMove, sft));
          }
          sft = __builtin_amdgcn_readfirstlane(sft);
          if (sft != 0) {
            // hold: no move that drops more than kBandHold of the row's mass
            const bool gone = k - sft < 0 || k - sft >= kBandLanes;
            if (wsum(gone ? wgt : 0.0) > static_cast<double>(kBandHold) * W) sft = 0;
          }
          if (sft != 0) {
            Mp = take(Mp, k, sft);
            Ip = take(Ip, k, sft);
            Dp = take(Dp, k, sft);
            TMp = take(TMp, k, sft);
            TIp = take(TIp, k, sft);
            TDp = take(TDp, k, sft);
            o += sft;
            ez_from = 2 - o;
            ez_to = H - o - kBandLanes;
          }
        }
        // Rescale by a power of two: every 64 rows in FP64, every 16 in
        // FP32 (a run of indels costs ~3e-5 a row, so 64 rows can take the
        // mass below FP32's range).
        if (kF32 ? i > 1 : (v0 == 0 && i0 > 1)) {
          const T m = wmax(fmax(Mp, fmax(Ip, Dp)));
          if (m > T(0)) {
            const int e = ilogb(m);
            const T sc = ldexp(T(1), -e);
            Mp *= sc; Ip *= sc; Dp *= sc;
            TMp *= sc; TIp *= sc; TDp *= sc;
            xg *= sc;
            lg += e;
          }
        }
        if (v0 == 0) {
          __syncthreads();  // the previous block's LDS reads are done
          if (k < n) {
            const int64_t p = ro + i0 - 1 + k;
            const uint8_t rc = a.rclass[p];
            const int iq = a.ins[p] & 127, dq = a.del[p] & 127, cq = a.gcp[p] & 127;
            const int lo = iq < dq ? iq : dq, hi = iq < dq ? dq : iq;
            const T e = ph2pr[a.q[p] & 127];
            RowRecL<T> rr;
            const T pm = T(1) - e;
            rr.pe = (rc & kWildcard) ? pm : e / T(3);
            rr.dp = pm - rr.pe;
            rr.mm = mmt[((hi * (hi + 1)) >> 1) + lo];
            rr.g = T(1) - ph2pr[cq];
            rr.mx = ph2pr[iq];
            rr.xx = ph2pr[cq];
            rr.my = ph2pr[dq];
            rr.cls = rc & 3;
            rr.pad = 0;
            s.tab[k] = rr;
          }
          jb = i0 + o - kLongSlack;
          for (int c = k; c < kLongWin; c += kBandLanes) {
            const int j = jb + c;
            s.hm[c] = (j >= 1 && j <= H) ? a.hmask[ho + j - 1] : uint8_t{0};
          }
          __syncthreads();
        }
        const int v1 = min(n, v0 + kLongEvery);
        for (int v = v0; v < v1; ++v) {
          const int ii = i0 + v;
          const RowRecL<T> rr = s.tab[v];
          const uint32_t hmv = s.hm[ii + o + k - jb];  // in [0, 151)
          const T sel = static_cast<T>((hmv >> rr.cls) & 1u);
          const T prior = fma(sel, rr.dp, rr.pe);
          const bool ez = (left_lane && ii >= ez_from) || (right_lane && ii <= ez_to);
          const int col = ii + o + k;
          if (ii > R - kEdgeRunRows)
            xg = fma(xg, yy, (k == 0 && col - 1 >= 1 && col - 1 <= H) ? fma(Mp, rr.mx, Ip * rr.xx) : T(0));

          T t = Ip * rr.g;
          t = fma(Dp, rr.g, t);
          t = fma(Mp, rr.mm, t);
          const T Mn = t * prior;
          T tt = TIp * rr.g;
          tt = fma(TDp, rr.g, tt);
          tt = fma(TMp, rr.mm, tt);
          T TMn = tt * prior;
          T In, TIn, Dn, TDn;
          if constexpr (kF32) {
            // the FP32 kernel's DPP-folded I and the two D scans in one block
            In = insert_next(Mp, Ip, rr.mx, rr.xx);
            TIn = insert_next(TMp, TIp, rr.mx, rr.xx);
            if (ez) {
              TMn = Mn;
              TIn = In;
            }
            Dn = dpp<kWaveShr1>(Mn) * rr.my;
            TDn = dpp<kWaveShr1>(TMn) * rr.my;
            decay_scan2<kBandLanes>(Dn, TDn, yy, yy2, yy4, yy8, f15, f31);
          } else {
            In = fma(dpp64<kWaveShl1>(Mp), rr.mx, dpp64<kWaveShl1>(Ip) * rr.xx);
            Dn = decay_scan64(dpp64<kWaveShr1>(Mn) * rr.my, yy, yy2, yy4, yy8, f15, f31);
            TIn = fma(dpp64<kWaveShl1>(TMp), rr.mx, dpp64<kWaveShl1>(TIp) * rr.xx);
            if (ez) {
              TMn = Mn;
              TIn = In;
            }
            TDn = decay_scan64(dpp64<kWaveShr1>(TMn) * rr.my, yy, yy2, yy4, yy8, f15, f31);
          }
          if (ez) TDn = Dn;

          Mp = Mn;
          Ip = In;
          Dp = Dn;
          TMp = TMn;
          TIp = TIn;
          TDp = TDn;
        }
      }
    }
    // S = sum over the last row of M + I, columns 1..H.
    const int j = R + o + k;
    const bool in = j >= 1 && j <= H;
    const T S = wsumT(in ? Mp + Ip : T(0));
    const T Tm = wsumT((in ? TMp + TIp : T(0)) + xg);
    cells += static_cast<unsigned long long>(R) * kBandLanes;
    if (k == 0) {
      const double L = log10(static_cast<double>(S)) + static_cast<double>(lg) * kLog10Two;
      if (S > T(0) && Tm <= static_cast<T>(a.tol) * S && L < a.verify_above) {
        a.out[w.out] = L;
        a.form[w.out] = kF32 ? kFormBandLong32 : kFormBandFP64;
        ++accepted;
      } else {
        const int q = atomicAdd(a.long_full_count, 1);
        a.long_full[q] = Task{read, hap, 1, -2, w.out};
      }
    }
  }
  if (k == 0) {
    if (accepted) atomicAdd(a.long_accepted, accepted);
    if (cells) {
      if (a.cells_done) atomicAdd(a.cells_done, cells);
      if (!kF32 && a.fp64_cells) atomicAdd(a.fp64_cells, cells);
    }
  }
}

// ---- long reads, quad shape (FP32): four pairs of one long read per
// wavefront ------------------------------------------------------------------
//
// The FP32 moving band over the quad kernel's layout: pair slot p owns the
// 16-lane DPP row p, lane k holds the diagonals o_p + 4k .. o_p + 4k + 3 of
// its pair's 64-diagonal band, and a row step is the quad kernel's (D
// carried divided by its row's my, I and D times g, the three-FMA chain and
// the decayed scan, T = F on the active edge zones of every row).  The four
// pairs share the read and so the row records; each has its own band that
// moves as in band_long_kernel: before every 16th row its centroid moves it
// by up to 4 diagonals unless the cells it would drop hold more than
// kBandHold of the pair's row mass (cells travel between lanes by
// ds_bpermute), and its state is rescaled by a power of two.  Per 64-row
// block a 192-column window of the haplotype's match selectors (f16, per
// read class) is staged per pair: with at most three moves of 4 in a block,
// row i lane k cell c reads window slot (i - i0) + (o - o0) + 12 + 4k + c.
// A read's pairs come in units of four (band_dispatch_long4_kernel); the
// rest run on band_long_kernel<float>.

// Splits the long reads' pairs: units of four pairs of one read with a band
// to the quad list, the rest with a band to the one-pair list, pairs
// without one to the full-matrix FP64 list.  One thread per long read.
__global__ __launch_bounds__(256) void band_dispatch_long4_kernel(BArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_tasks64r) return;
  const Task tk = a.tasks64r[t];
  const bool gc = a.long_on && a.gcp_const[tk.read] != 0;
  auto band_ok = [&](int h, int2& b) {
    b = a.band[tk.out0 + h];
    const int64_t w = static_cast<int64_t>(b.y) - b.x + 1;
    return gc && w >= 1 && w <= kBandLanes && b.x >= -(1 << 26) && b.y <= (1 << 26);
  };
  int n_ok = 0;
  for (int h = 0; h < tk.nhap; ++h) {
    int2 b;
    n_ok += band_ok(h, b);
  }
  const int nq = n_ok & ~(kQuadPairs - 1);
  int bq = nq > 0 ? atomicAdd(a.long4_count, nq) : 0;
  int kq = 0;
  for (int h = 0; h < tk.nhap; ++h) {
    int2 b;
    const int64_t o = tk.out0 + h;
    if (band_ok(h, b)) {
      const int w = b.y - b.x + 1;
      const BandItem it{tk.read, tk.hap0 + h, b.x - ((kBandLanes - w) >> 1), 0, o};
      if (kq < nq) {
        a.long_items4[bq + kq++] = it;
      } else {
        const int q = atomicAdd(a.long_count, 1);
        a.long_items[q] = it;
      }
    } else {
      const int q = atomicAdd(a.long_full_count, 1);
      a.long_full[q] = Task{tk.read, tk.hap0 + h, 1, 0, o};
    }
  }
}

// Maximum over each 16-lane row, returned to every lane of the row.
__device__ __forceinline__ float row_max(float v, int lane) {
  v = fmaxf(v, dpp<kRowShr + 1>(v));
  v = fmaxf(v, dpp<kRowShr + 2>(v));
  v = fmaxf(v, dpp<kRowShr + 4>(v));
  v = fmaxf(v, dpp<kRowShr + 8>(v));
  return __shfl(v, lane | 15);
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void band_long_quad_kernel(BArgs a) {
  __shared__ RowRec tab[kBandLanes];
  __shared__ _Float16 selw[kQuadPairs][4][kLongWin];
  const int lane = threadIdx.x;
  const int p = lane >> 4;  // pair slot: the DPP row
  const int k = lane & 15;  // lane in the row: diagonals o + 4k .. o + 4k + 3
  const bool left_lane = k == 0;
  const bool right_lane = k == kQuadLanes - 1;
  const int n_units = min(*a.long4_count, a.long_cap) / kQuadPairs;
  int accepted = 0;
  unsigned long long cells = 0;
  for (;;) {
    int u = 0;
    if (lane == 0) u = atomicAdd(a.long4_queue, 1);
    u = __builtin_amdgcn_readfirstlane(__shfl(u, 0));
    if (u >= n_units) break;
    const BandItem w = a.long_items4[kQuadPairs * u + p];
    const int read = __builtin_amdgcn_readfirstlane(a.long_items4[kQuadPairs * u].read);
    const int hap = w.hap;
    int o = w.o_lo;
    const int64_t ro = a.read_off[read];
    const int R = static_cast<int>(a.read_off[read + 1] - ro);
    const int64_t ho = a.hap_off[hap];
    const int H = static_cast<int>(a.hap_off[hap + 1] - ho);
    const int cq0 = a.gcp[ro] & 127;
    const float yy = a.ph2pr[cq0];
    const float gx = 1.f - a.ph2pr[cq0];
    const float ig = static_cast<float>(1.0 / static_cast<double>(gx));
    const float yy2 = yy * yy, yy3 = yy2 * yy, yy4 = yy2 * yy2, yy8 = yy4 * yy4, yy16 = yy8 * yy8,
                yy32 = yy16 * yy16;

    // Row 0: D = g / H on columns 0..H (the constant K of the reference
    // cancels); T = F on the edge zones active there.
    float M[4], I[4], D[4], TM[4], TI[4], TD[4];
    float myp = 1.f;
    {
      const float d0 = gx / static_cast<float>(H);
      const bool ez0 = (left_lane && o - 1 >= 0) || (right_lane && o + kBandLanes <= H);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int d = o + 4 * k + c;
        M[c] = I[c] = TM[c] = TI[c] = 0.f;
        D[c] = (d >= 0 && d <= H) ? d0 : 0.f;
        TD[c] = ez0 ? D[c] : 0.f;
      }
    }
    int lg = 0;       // exponent removed by rescaling (this pair's)
    float xg = 0.f;   // read-end insertion runs leaving through the bottom lane
    int o0 = o;       // the band's offset when the block's window was staged
    for (int i0 = 1; i0 <= R; i0 += kBandLanes) {
      const int n = min(kBandLanes, R - i0 + 1);
      for (int v0 = 0; v0 < n; v0 += kLongEvery) {
        const int i = i0 + v0;  // next row; the state holds row i - 1
        if (i > 1) {
          // cells past the haplotype end never feed in-matrix cells
#pragma unroll
          for (int c = 0; c < 4; ++c)
            if ((i - 1) + o + 4 * k + c > H) M[c] = I[c] = D[c] = TM[c] = TI[c] = TD[c] = 0.f;
          float wv[4], wl = 0.f, cl = 0.f;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            wv[c] = __builtin_fmaf(__builtin_fmaf(D[c], myp, I[c]), ig, M[c]);
            wl += wv[c];
            cl = __builtin_fmaf(static_cast<float>(4 * k + c), wv[c], cl);
          }
          const float W = __shfl(row_total(wl), lane | 15);
          const float Cn = __shfl(row_total(cl), lane | 15);
          int sft = 0;
          if (W > 0.f) {
            sft = static_cast<int>(floorf(Cn / W - 0.5f * (kBandLanes - 1) + 0.5f));
            sft = min(kLongMove, max(-kLongMove, sft));
          }
          // hold: no move that drops more than kBandHold of the pair's row mass
          float gone = 0.f;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int t = 4 * k + c;
            if (t - sft < 0 || t - sft >= kBandLanes) gone += wv[c];
          }
          const float G = __shfl(row_total(gone), lane | 15);
          if (sft != 0 && G > kBandHold * W) sft = 0;
          if (__ballot(sft != 0) != 0) {
            // cell t of the pair takes the old cell t + sft (0 outside): the
            // old cells of lanes kA and kA + 1 of the same row
            const int base = 4 * k + sft;              // in [-4, 67]
            const int kA = ((base + 64) >> 2) - 16;    // floor(base / 4)
            const int r = base - 4 * kA;               // 0..3
            const int la = (lane & ~15) | (kA & 15), lb = (lane & ~15) | ((kA + 1) & 15);
            const bool okA = kA >= 0 && kA < kQuadLanes, okB = kA + 1 >= 0 && kA + 1 < kQuadLanes;
            auto shift = [&](float* sv) {
              float A[4], B[4];
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                A[c] = __shfl(sv[c], la);
                B[c] = __shfl(sv[c], lb);
              }
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                const int q = r + c;  // 0..6
                const float va = q == 0 ? A[0] : q == 1 ? A[1] : q == 2 ? A[2] : A[3];
                const float vb = q == 4 ? B[0] : q == 5 ? B[1] : B[2];
                sv[c] = q < 4 ? (okA ? va : 0.f) : (okB ? vb : 0.f);
              }
            };
            shift(M);
            shift(I);
            shift(D);
            shift(TM);
            shift(TI);
            shift(TD);
            o += sft;
          }
          // rescale this pair by a power of two
          float m = 0.f;
#pragma unroll
          for (int c = 0; c < 4; ++c) m = fmaxf(m, fmaxf(M[c], fmaxf(I[c], D[c])));
          m = row_max(m, lane);
          if (m > 0.f) {
            const int e = ilogbf(m);
            const float sc = ldexpf(1.f, -e);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              M[c] *= sc; I[c] *= sc; D[c] *= sc;
              TM[c] *= sc; TI[c] *= sc; TD[c] *= sc;
            }
            xg *= sc;
            lg += e;
          }
        }
        if (v0 == 0) {
          __syncthreads();  // the previous block's LDS reads are done
          if (lane < n) {
            const int64_t q = ro + i0 - 1 + lane;
            const uint8_t rc = a.rclass[q];
            const int iq = a.ins[q] & 127, dq = a.del[q] & 127;
            const int lo = iq < dq ? iq : dq, hi = iq < dq ? dq : iq;
            const float e = a.ph2pr[a.q[q] & 127];
            RowRec rr;
            const float pm = 1.f - e;
            rr.pe = (rc & kWildcard) ? pm : e / 3.f;
            rr.dp = pm - rr.pe;
            rr.mm = a.mm[((hi * (hi + 1)) >> 1) + lo];
            rr.mx = gx * a.ph2pr[iq];
            rr.xx = yy;
            rr.my = gx * a.ph2pr[dq];
            rr.selo = static_cast<uint32_t>(2 * ((rc & 3) * kLongWin + lane));
            rr.pad = 0.f;
            tab[lane] = rr;
          }
          o0 = o;
          const int jb = i0 + o0 - kLongSlack;  // column of window slot 0
          for (int s = k; s < kLongWin; s += kQuadLanes) {
            const int j = jb + s;
            const uint32_t mk = (j >= 1 && j <= H) ? a.hmask[ho + j - 1] : 0u;
#pragma unroll
            for (int q = 0; q < 4; ++q) selw[p][q][s] = ((mk >> q) & 1u) ? _Float16(1) : _Float16(0);
          }
          __syncthreads();
        }
        const int v1 = min(n, v0 + kLongEvery);
        // this lane's cell 0 slot in its pair's window, less the row's part
        const _Float16* sl = &selw[p][0][0] + ((o - o0) + kLongSlack + 4 * k);
        for (int v = v0; v < v1; ++v) {
          const int ii = i0 + v;
          const RowRec rr = tab[v];
          const char* sb = reinterpret_cast<const char*>(sl) + rr.selo;
          float pr[4];
#pragma unroll
          for (int c = 0; c < 4; ++c)
            pr[c] = __builtin_fmaf(static_cast<float>(*reinterpret_cast<const _Float16*>(sb + 2 * c)), rr.dp, rr.pe);
          if (ii > R - kEdgeRunRows) {
            const int c0 = ii - 1 + o;  // the bottom cell's column on row ii - 1
            const float xe = __builtin_fmaf(M[0], rr.mx, I[0] * rr.xx);
            xg = __builtin_fmaf(xg, yy, (c0 >= 1 && c0 <= H) ? xe : 0.f);
          }
          const bool ez = (left_lane && ii + o >= 2) || (right_lane && ii + o + kBandLanes <= H);
          float Mn[4], In[4], TMn[4], TIn[4], Y[4], TY[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            Mn[c] = __builtin_fmaf(M[c], rr.mm, __builtin_fmaf(D[c], myp, I[c])) * pr[c];
            TMn[c] = __builtin_fmaf(TM[c], rr.mm, __builtin_fmaf(TD[c], myp, TI[c])) * pr[c];
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            In[c] = __builtin_fmaf(M[c + 1], rr.mx, I[c + 1] * rr.xx);
            TIn[c] = __builtin_fmaf(TM[c + 1], rr.mx, TI[c + 1] * rr.xx);
          }
          In[3] = dpp<kRowShl1>(I[0]) * rr.xx;
          TIn[3] = dpp<kRowShl1>(TI[0]) * rr.xx;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            if (ez) TMn[c] = Mn[c];
            Y[c] = Mn[c];
            TY[c] = TMn[c];
          }
          myp = rr.my;
          quad_cross(In[3], TIn[3], Y[0], TY[0], M[0], TM[0], Y[3], TY[3], rr.mx, yy);
          float d1 = Y[0], td1 = TY[0];
          float d2 = __builtin_fmaf(d1, yy, Y[1]), td2 = __builtin_fmaf(td1, yy, TY[1]);
          float x = __builtin_fmaf(d2, yy, Y[2]), tx = __builtin_fmaf(td2, yy, TY[2]);
          quad_scan(x, tx, d1, td1, d2, td2, yy4, yy8, yy16, yy32, yy2, yy3);
          const float z = __builtin_fmaf(x, yy, Y[3]), tz = __builtin_fmaf(tx, yy, TY[3]);
          const float Dn0 = dpp<kRowShr1>(z);
          float TDn0 = dpp<kRowShr1>(tz);
          float TDn1 = td1, TDn2 = td2, TDn3 = tx;
          if (ez) {
#pragma unroll
            for (int c = 0; c < 4; ++c) TIn[c] = In[c];
            TDn0 = Dn0;
            TDn1 = d1;
            TDn2 = d2;
            TDn3 = x;
          }
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            M[c] = Mn[c];
            I[c] = In[c];
            TM[c] = TMn[c];
            TI[c] = TIn[c];
          }
          D[0] = Dn0; D[1] = d1; D[2] = d2; D[3] = x;
          TD[0] = TDn0; TD[1] = TDn1; TD[2] = TDn2; TD[3] = TDn3;
        }
      }
    }
    // S = sum over the last row of M + I, columns 1..H (I carried times g);
    // lane 15 of the row gets its pair's totals.
    float sS = 0.f, sT = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = R + o + 4 * k + c;
      if (j >= 1 && j <= H) {
        sS += __builtin_fmaf(I[c], ig, M[c]);
        sT += __builtin_fmaf(TI[c], ig, TM[c]);
      }
    }
    if (left_lane) sT += xg * ig;
    const float S = row_total(sS);
    const float Tm = row_total(sT);
    if (right_lane) {
      cells += static_cast<unsigned long long>(R) * kBandLanes;
      const double L = log10(static_cast<double>(S)) + static_cast<double>(lg) * kLog10Two;
      if (S > 0.f && Tm <= static_cast<float>(a.tol) * S && L < a.verify_above) {
        a.out[w.out] = L;
        a.form[w.out] = kFormBandLong32;
        ++accepted;
      } else {
        const int q = atomicAdd(a.long_full_count, 1);
        a.long_full[q] = Task{read, hap, 1, -2, w.out};
      }
    }
  }
  if (right_lane) {
    if (accepted) atomicAdd(a.long_accepted, accepted);
    if (cells && a.cells_done) atomicAdd(a.cells_done, cells);
  }
}

// The FP64 band rescue (band_reject): the FP64 moving-band kernel over the
// rescue list, its rejects to the FP64 full-matrix rescue list.
hipError_t launch_band_rescue(const BArgs& b, Task* rescue, int* rescue_count, int* queue, int* accepted, int grid,
                              hipStream_t st) {
  if (b.band_rescue == nullptr) return hipSuccess;
  BArgs a = b;
  a.long_items = b.band_rescue;
  a.long_count = b.band_rescue_count;
  a.long_cap = b.band_rescue_cap;
  a.long_queue = queue;
  a.long_full = rescue;
  a.long_full_count = rescue_count;
  a.long_accepted = accepted;
  hipLaunchKernelGGL(band_long_kernel<double>, dim3(grid), dim3(kBandLanes), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_band_long(const BArgs& a, int grid, hipStream_t st) {
  if (a.n_tasks64 <= 0) return hipSuccess;
  if (a.long_fp32 && a.n_tasks64r > 0) {
    // FP32: a read's pairs in units of four on the quad kernel, the rest on
    // the one-pair kernel
    hipLaunchKernelGGL(band_dispatch_long4_kernel, dim3((a.n_tasks64r + 255) / 256), dim3(256), 0, st, a);
    hipLaunchKernelGGL(band_long_quad_kernel, dim3(grid), dim3(kBandLanes), 0, st, a);
    hipLaunchKernelGGL(band_long_kernel<float>, dim3(grid), dim3(kBandLanes), 0, st, a);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(band_dispatch_long_kernel, dim3((a.n_tasks64 + 255) / 256), dim3(256), 0, st, a);
  if (a.long_fp32)
    hipLaunchKernelGGL(band_long_kernel<float>, dim3(grid), dim3(kBandLanes), 0, st, a);
  else
    hipLaunchKernelGGL(band_long_kernel<double>, dim3(grid), dim3(kBandLanes), 0, st, a);
  return hipGetLastError();
}

int band_long_blocks_per_cu(int fp32) {
  int n = 0;
  const hipError_t e = fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, band_long_kernel<float>, kBandLanes, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, band_long_kernel<double>, kBandLanes, 0);
  if (e != hipSuccess) return 8;
  return n > 0 ? n : 1;
}

hipError_t launch_band_dispatch(const BArgs& a, hipStream_t st) {
  const int blocks = (a.n_tasks + 255) / 256;
  if (blocks > 0) hipLaunchKernelGGL(band_dispatch_kernel, dim3(blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

static size_t band_lds_bytes(int pairs, int sel_cols) {
  return static_cast<size_t>(kTabBytes) + static_cast<size_t>(pairs) * 4 * sel_cols * sizeof(_Float16);
}

static size_t quad_lds_bytes(int sel_cols) {
  return static_cast<size_t>(kTabBytes) + static_cast<size_t>(kQuadPairs) * quad_sel_stride(sel_cols) * sizeof(_Float16);
}

// The banded short-read kernels, in two launches: the quad kernels (quad
// shape only) over their groups of four, and the one-pair-per-wavefront
// kernels (every wide pair with the one-pair shape, else the quad shape's
// leftovers; the narrow band).  grid: the wide kernel of the plan's shape;
// grid1: the one-pair kernels that follow the quad shape.
static hipError_t check_band_args(const BArgs& a) {
  if (a.sel_cols < band_sel_cols(0) || a.sel_cols > band_sel_cols(kBandReadMax)) return hipErrorInvalidValue;
  if (a.wide_pairs != 1 && a.wide_pairs != kQuadPairs) return hipErrorInvalidValue;
  return hipSuccess;
}

hipError_t launch_band_quads(const BArgs& a, int grid, hipStream_t st) {
  if (hipError_t e = check_band_args(a); e != hipSuccess) return e;
  if (a.wide_pairs != kQuadPairs) return hipSuccess;
  hipLaunchKernelGGL(band_quad_kernel<kQuadCells>, dim3(grid), dim3(kBandLanes), quad_lds_bytes(a.sel_cols), st, a);
  if (a.mid_on)
    hipLaunchKernelGGL(band_quad_kernel<kMidCells>, dim3(grid), dim3(kBandLanes), quad_lds_bytes(a.sel_cols), st, a);
  return hipGetLastError();
}

hipError_t launch_band_one(const BArgs& a, int grid, int grid1, hipStream_t st) {
  if (hipError_t e = check_band_args(a); e != hipSuccess) return e;
  const bool quad = a.wide_pairs == kQuadPairs;
  hipLaunchKernelGGL(band_forward_kernel<kBandLanes>, dim3(quad ? grid1 : grid), dim3(kBandLanes),
                     band_lds_bytes(1, a.sel_cols), st, a);
  hipLaunchKernelGGL(band_forward_kernel<kBandNarrow>, dim3(grid1), dim3(kBandLanes),
                     band_lds_bytes(2, a.sel_cols), st, a);
  return hipGetLastError();
}

int band_blocks_per_cu(int sel_cols, int wide_pairs) {
  int n = 0;
  const hipError_t e =
      wide_pairs == kQuadPairs
          ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, band_quad_kernel<kQuadCells>, kBandLanes,
                                                         quad_lds_bytes(sel_cols))
          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, band_forward_kernel<kBandLanes>, kBandLanes,
                                                         band_lds_bytes(1, sel_cols));
  if (e != hipSuccess) return 8;
  return n > 0 ? n : 1;
}

}  // namespace bvp
