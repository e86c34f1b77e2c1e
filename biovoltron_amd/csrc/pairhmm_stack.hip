// MI355X (gfx950) full-matrix FP32 PairHMM over stacked reads.
//
// Same recurrence and driver rule as pairhmm_forward_kernel (reference:
// lib/intel_native/pairhmm/avx-pairhmm-template.h:83-128 per-row
// transitions, :136-177 boundary, :183-198 cell recurrence, :299-345 final
// sum; rescue rule pairhmm_avx.hpp:140-149), laid out so that a wavefront's
// lanes stay busy:
//
// * Stacked reads.  A task is a run of consecutive reads of one region; their
//   rows are stacked into one tall matrix and swept in stripes of 63 rows
//   against each haplotype, so a stripe only runs partly empty at the end of
//   the task, not at the end of every read.  A lane whose row is the first of
//   a read takes the row-0 boundary (M = I = 0, D = K/H) instead of the row
//   above; the lane holding a read's last row adds M + I per column to that
//   read's sum.  At most kStackEnds reads end inside a stripe's lanes 1-62
//   (their per-column sums go through a small LDS ring); lane 63 may end one
//   more (its sums are read back from the carry row).
// * Carry row in LDS.  Lane 63 (the stripe's last row) writes (Z, J) per
//   column to one LDS row; lane 0 of the next stripe replays it.  Column c is
//   read at step c and overwritten 63 steps later, so one row serves every
//   stripe of a haplotype in place.  Haplotypes up to kStackHmax bases.
// * Eight VALU instructions per anti-diagonal step.  With E = D * dm' (the
//   deletion term of the next row's Z carried pre-multiplied):
//     prior = clamp(sel * (pm - pmis) + pmis)        v_fma_mix_f32 clamp
//     t     = E * yy                                  v_mul
//     E    += dpp(J) * im'        (E now holds E + I*im')   v_fmac_dpp
//     w     = dpp(J) * xx'                            v_mul_dpp
//     M     = dpp(Z two steps back) * prior           v_mul_dpp
//     J     = M * mx' + w                             v_fmac
//     E'    = M * (my * dm') + t                      v_fmac  (next column)
//     Z     = M * mm' + E                             v_fma
//   (Z last: every input is read before it is written, so its register may
//   be one of theirs; every DPP operand was written at least two VALU
//   instructions earlier, the hazard distance DPP needs.)
//   The three DPP operations shift by wave_shr:1 without bound_ctrl, so lane
//   0 (no source lane) keeps its destination: loading the carry entry (Z, J)
//   into M and w before the step makes lane 0 replay the previous stripe's
//   last row with no select instructions (its mm' = 1, mx' = 0, a = 0).
// * Boundaries without branches.  A read's last row publishes Z = K/H * g of
//   the next read's first row (it keeps E at that constant: yy = 1, a = 0,
//   mm' = im' = 0) and J = M + I (mx' = xx' = 1); a first row ignores the J
//   it receives (im' = xx' = 0).  Lane 0 publishes the row-0 Z the same way
//   when the stripe starts a read.  A first row below a last row sees that Z
//   at every column, also left of column 1: the selector of columns outside
//   [1, H] is -65504, which the clamp turns into prior 0 (M = 0) for rows with
//   pm > pmis; a read whose first row is not such a row starts a stripe.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pairhmm_common.h"

namespace bvp {

namespace {

constexpr float kK32 = 0x1p120f;                                   // Context.h:109 (FP32)
constexpr double kLog10K32 = 120.0 * 0.30102999566398119521;       // log10(2^120)
constexpr unsigned short kSelOutside = 0xFBFF;  // f16 -65504 outside [1, H]: the clamped prior is 0
constexpr unsigned short kSelMatch = 0x3C00;    // f16 1.0
constexpr int kStackSlots = 4;                  // AVX read classes A C T G (N is a wildcard row)
// Two copies of the f16 selector window, the second shifted by one column,
// so that every lane reads the selectors of two consecutive steps as one
// aligned dword (even lanes from copy 0, odd lanes from copy 1).  Copy 1
// starts 17 dwords later modulo the 32 banks, so the even and odd lanes of a
// half wave read 32 different banks; class rows are 64 dwords apart (the
// same banks), so lanes of different classes do not conflict either.
constexpr int kWinRow = 2 * kBlock;             // f16 entries per class row
constexpr int kWinCopy1 = kStackSlots * kWinRow / 2 + 17;  // dword offset of copy 1
constexpr int kWinDwords = kWinCopy1 + kStackSlots * kWinRow / 2;

// LDS pointers with their address space spelled out, so that a base kept in
// a VGPR (see the sweep) still compiles to DS instructions.
typedef __attribute__((address_space(3))) float lds_f32;
typedef __attribute__((address_space(3))) const uint32_t lds_u32;
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) v2f lds_f2;

struct StackLds {
  uint32_t win[kWinDwords];                // f16 selectors, columns [64b - 64, 64b + 64), two copies
  float2 ring[kStackEnds][kBlock];         // per-column J (= M + I, in .y) of the rows that end a read
  float2 carry[1];                         // (Z, J) per column, kStackPad + carry_cols (dynamic)
};

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// One anti-diagonal step: ZS holds this lane's Z from two steps back (the
// DPP source of the M product) and receives this step's Z; J1 holds J from
// one step back and receives this step's J; E is the carried deletion term.
// The step's LDS operands (selector, carry Z and J for lane 0) were loaded
// one step earlier; this step loads the next step's.  Step U + 1 reads
// column 64b + U + 1 while step U's stores go to column 64b + U - 63 and the
// ring, so the early loads never see a value this block still writes.
#define STK_STEP(U, ZS, HALF)                                                             \
  {                                                                                       \
    float M = c_n.x, w = c_n.y, prior, t, zn;                                             \
    if ((U) + 1 < kBlock) c_n = crd[(U) + 1];                                             \
    if ((HALF) == 0 && (U) + 2 < kBlock) selw_n = selp[((U) + 2) / 2];                     \
    asm(                                                                                  \
        "v_fma_mix_f32 %[prior], %[sel], %[dp], %[pmis] op_sel:[" #HALF                   \
        ",0,0] op_sel_hi:[1,0,0] clamp\n\t"                                              \
        "v_mul_f32 %[t], %[E], %[yy]\n\t"                                                 \
        "v_fmac_f32_dpp %[E], %[J1], %[im] wave_shr:1 row_mask:0xf bank_mask:0xf\n\t"     \
        "v_mul_f32_dpp %[w], %[J1], %[xx] wave_shr:1 row_mask:0xf bank_mask:0xf\n\t"      \
        "v_mul_f32_dpp %[M], %[Z2], %[prior] wave_shr:1 row_mask:0xf bank_mask:0xf\n\t"   \
        "v_fmac_f32 %[w], %[M], %[mx]\n\t"                                                \
        "v_fmac_f32 %[t], %[M], %[aa]\n\t"                                                \
        "v_fma_f32 %[zn], %[M], %[mm], %[E]"                                              \
        : [prior] "=&v"(prior), [t] "=&v"(t), [zn] "=v"(zn), [M] "+v"(M), [E] "+v"(E),    \
          [w] "+v"(w)                                                                     \
        : [sel] "v"(selw), [dp] "v"(dp), [pmis] "v"(pmis), [yy] "v"(yy), [J1] "v"(J1),     \
          [im] "v"(im), [xx] "v"(xx), [Z2] "v"(ZS), [mm] "v"(mm), [mx] "v"(mx),            \
          [aa] "v"(aa));                                                                  \
    if (store) {                                                                          \
      stp[2 * (U)] = zn;                                                                  \
      stp[2 * (U) + 1] = w;                                                               \
    }                                                                                     \
    E = t;                                                                                \
    ZS = zn;                                                                              \
    J1 = w;                                                                               \
    if ((HALF) == 1) selw = selw_n;                                                       \
  }

// an even step takes the selector word's low half, the odd step after it the high half
#define STK_STEP2(U) STK_STEP((U), Za, 0) STK_STEP((U) + 1, Zb, 1)
#define STK_STEP16(U) \
  STK_STEP2((U)) STK_STEP2((U) + 2) STK_STEP2((U) + 4) STK_STEP2((U) + 6) \
  STK_STEP2((U) + 8) STK_STEP2((U) + 10) STK_STEP2((U) + 12) STK_STEP2((U) + 14)

}  // namespace

__global__ __launch_bounds__(64) void pairhmm_stack_kernel(KArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  StackLds& s = *reinterpret_cast<StackLds*>(lds_raw);
  const int lane = threadIdx.x;
  const int ncarry = kStackPad + a.carry_cols;
  float2* carry = s.carry;  // (Z, J) of the stripe's last row per column (+ kStackPad)
  // every carry entry is finite before the first read (lane 0 multiplies
  // stale entries by 0 where a stripe starts a read)
  for (int i = lane; i < ncarry; i += kLanes) carry[i] = float2{0.f, 0.f};
  __syncthreads();
  const int n_tasks = a.n_tasks_dev ? *a.n_tasks_dev : a.n_tasks;
  unsigned long long cells_w = 0;
  for (;;) {
    int ti = 0;
    if (lane == 0) ti = atomicAdd(a.counter, 1);
    ti = __builtin_amdgcn_readfirstlane(__shfl(ti, 0));
    if (ti >= n_tasks) break;
    // a stacked task, or one read against a run of haplotypes from a pair list
    // (the learned-band run's full-matrix list; band -2: a band it rejected)
    StackTask tk;
    uint8_t form_ok = kFormFP32;
    if (a.stasks) {
      tk = a.stasks[ti];
    } else {
      const Task t1 = a.tasks[ti];
      tk = StackTask{t1.read, 1, t1.hap0, t1.nhap, t1.out0, 0, 0};
      if (t1.band == -2) form_ok = kFormBandFallback;
    }
    const int64_t p0 = a.read_off[tk.read0], p1 = a.read_off[tk.read0 + tk.nreads];
    for (int h = 0; h < tk.nhap; ++h) {
      const int hap = tk.hap0 + h;
      const int64_t ho = a.hap_off[hap];
      const int H = static_cast<int>(a.hap_off[hap + 1] - ho);
      if (H + 128 > a.carry_cols) {
        // longer than the LDS carry row (lists are split so that this does not
        // happen): NaN rather than a sweep past the row
        if (lane == 0)
          for (int r = 0; r < tk.nreads; ++r) {
            const int64_t o = tk.out0 + static_cast<int64_t>(r) * tk.stride + h;
            a.out[o] = __builtin_nan("");
            a.form[o] = form_ok;
          }
        continue;
      }
      if (lane == 0) cells_w += static_cast<unsigned long long>(p1 - p0) * static_cast<unsigned long long>(H);
      const float kh = kK32 / static_cast<float>(H);  // D[0][j] = K/H (avx-pairhmm-template.h:86-92)
      int64_t ps = p0;         // first base of this stripe (lane 1)
      int rd = 0;              // reads of the task that ended before this stripe
      bool starts_read = true;  // lane 1 holds the first row of a read
      while (ps < p1) {
        // ---- stripe set-up: rows, read ends, per-lane transitions -------------
        const int64_t p = ps + lane - 1;
        bool act = lane >= 1 && p < p1;
        const uint8_t rc = act ? a.rclass[p] : uint8_t(0);
        bool last = act && (rc & kReadEnd);
        const float e = act ? a.tab.ph2pr_f[a.q[p] & 127] : 0.f;
        const float pm = 1.f - e;
        const float pmis_r = (rc & kWildcard) ? pm : e / 3.f;
        int L = static_cast<int>(std::min<int64_t>(63, p1 - ps));
        {
          // at most kStackEnds read ends in lanes 1-62: cut the stripe after the last one allowed
          const unsigned long long em = __ballot(last && lane <= 62);
          if (__popcll(em) > kStackEnds) {
            unsigned long long m = em;
            for (int k = 1; k < kStackEnds; ++k) m &= m - 1;  // drop the first kStackEnds - 1 ends
            L = __builtin_ctzll(m);
          }
          // A read's first row right below another read's last row gets that
          // row's Z = K/H * g at every column, so its prior must be 0 left of
          // column 1; the clamped selector gives that only when pm > pmis.  A
          // read starting with an N base or a quality of 0 or 1 starts the
          // next stripe instead, below the feeder's row 0.
          const unsigned long long bad = __ballot(act && !(pm - pmis_r > 0.f)) & (__ballot(last) << 1);
          if (bad) L = std::min(L, __builtin_ctzll(bad) - 1);
        }
        act = act && lane <= L;
        last = last && lane <= L;
        const unsigned long long ends = __ballot(last);
        const unsigned long long ring_ends = ends & ~(1ull << 63);
        const int n_ring = __popcll(ring_ends);
        const bool end63 = (ends >> 63) & 1ull;
        // the row above is the previous read's last row (or row 0)
        const bool first = act && (lane == 1 ? starts_read : ((ends >> (lane - 1)) & 1ull) != 0);
        float dp = 0.f, pmis = 0.f, yy = 0.f, aa = 0.f, mm = 0.f, im = 0.f, mx = 0.f, xx = 0.f, E = 0.f;
        uint32_t cls = 0;
        if (act) {
          pmis = pmis_r;
          dp = pm - pmis;
          cls = rc & 3;
          const float my = a.tab.ph2pr_f[a.del[p] & 127];
          yy = a.tab.ph2pr_f[a.gcp[p] & 127];
          if (!last) {
            // the next row's transitions (row r + 1 of the same read)
            const int iq = a.ins[p + 1] & 127, dq = a.del[p + 1] & 127, cq = a.gcp[p + 1] & 127;
            const int lo = iq < dq ? iq : dq, hi = iq < dq ? dq : iq;
            mm = a.tab.mm_f[((hi * (hi + 1)) >> 1) + lo];
            const float g = 1.f - a.tab.ph2pr_f[cq];
            im = first ? 0.f : g;
            mx = a.tab.ph2pr_f[iq];
            xx = first ? 0.f : a.tab.ph2pr_f[cq];
            aa = my * g;
          } else {
            // J = M + I per column; Z = row 0 of the next read (K/H * g of its first row)
            mx = 1.f;
            xx = first ? 0.f : 1.f;
            yy = 1.f;
            E = (p + 1 < p1) ? kh * (1.f - a.tab.ph2pr_f[a.gcp[p + 1] & 127]) : 0.f;
          }
        } else if (lane == 0) {
          // feeder: replays the carry row, or publishes row 0 when lane 1 starts a read
          if (starts_read) {
            yy = 1.f;
            E = kh * (1.f - a.tab.ph2pr_f[a.gcp[ps] & 127]);
          } else {
            mm = 1.f;
          }
        }
        // ring slot of an end lane (ends before it), and where lane 63 / end lanes store
        const int slot = __popcll(ring_ends & ((1ull << lane) - 1ull));
        const bool store = (lane == 63) || (last && lane <= 62);
        // ---- sweep: columns 0..H for lanes 1..L, 64 steps per block ----------------
        const int steps = H + L + 1;
        const int nb = (steps + kBlock - 1) / kBlock;
        uint8_t pre_m = (lane >= 1 && lane <= H) ? a.hmask[ho + lane - 1] : uint8_t(0);
        float Za = 0.f, Zb = 0.f, J1 = 0.f;
        double acc_k[kStackEnds] = {};  // this lane's columns of ring slot k's sum
        double acc63 = 0.0;
        // this lane's selector dword for steps 0 and 1 (window index 64 - lane and the next)
        const uint32_t* selp0 = (lane & 1) ? &s.win[kWinCopy1 + (cls * kWinRow + kBlock - 1 - lane) / 2]
                                           : &s.win[(cls * kWinRow + kBlock - lane) / 2];
        unsigned short* win16 = reinterpret_cast<unsigned short*>(s.win);
        unsigned short lo_[kStackSlots];  // selectors of window index lane (the previous block's columns)
#pragma unroll
        for (int k = 0; k < kStackSlots; ++k) lo_[k] = kSelOutside;
        for (int b = 0; b < nb; ++b) {
          __syncthreads();
          {
            // selector window: the previous block's columns at index lane, this
            // block's at 64 + lane (copy 1 holds index i + 1 at entry i)
            const int c = kBlock * b + lane;
            const bool inside = c >= 1 && c <= H;
#pragma unroll
            for (int k = 0; k < kStackSlots; ++k) {
              const unsigned short v = !inside ? kSelOutside : (((pre_m >> k) & 1) ? kSelMatch : uint16_t(0));
              unsigned short* c0 = win16 + k * kWinRow;
              unsigned short* c1 = win16 + 2 * kWinCopy1 + k * kWinRow;
              c0[lane] = lo_[k];
              c0[kBlock + lane] = v;
              if (lane >= 1) c1[lane - 1] = lo_[k];
              c1[kBlock - 1 + lane] = v;
              lo_[k] = v;
            }
            const int cn = c + kBlock;
            pre_m = (cn >= 1 && cn <= H) ? a.hmask[ho + cn - 1] : uint8_t(0);
          }
          __syncthreads();
          // the carry loads are the same address on every lane, but a DS
          // instruction takes its address from a VGPR: keep the bases there
          lds_u32* selp = (lds_u32*)(selp0);
          lds_f2* crd = (lds_f2*)(&carry[kStackPad + kBlock * b]);
          // lane 63 stores (Z, J) to the carry row, an ending row to the ring
          lds_f32* stp = (lane == 63) ? (lds_f32*)(&carry[kStackPad + kBlock * b - 63].x)
                                      : (lds_f32*)(&s.ring[slot < kStackEnds ? slot : 0][0].x);
          asm("" : "+v"(crd), "+v"(stp), "+v"(selp));
          uint32_t selw = selp[0], selw_n = 0;
          v2f c_n = crd[0];
          const int left = steps - kBlock * b;
          STK_STEP16(0)
          if (left > 16) {
            STK_STEP16(16)
            if (left > 32) {
              STK_STEP16(32)
              if (left > 48) {
                STK_STEP16(48)
              }
            }
          }
          __syncthreads();
          // per-column sums of the rows that end a read: slot k's row is on the lane of the
          // k-th ring end; lane u holds the column that row reached at step u of this block
          {
            unsigned long long m = ring_ends;
#pragma unroll
            for (int k = 0; k < kStackEnds; ++k) {
              if (m) {
                const int le = __builtin_ctzll(m);
                m &= m - 1;
                const int c = kBlock * b + lane - le;
                if (c >= 1 && c <= H) acc_k[k] += static_cast<double>(s.ring[k][lane].y);
              }
            }
            if (end63) {
              const int c = kBlock * b + lane - 63;
              if (c >= 1 && c <= H) acc63 += static_cast<double>(carry[kStackPad + c].y);
            }
          }
        }
        // ---- results of the reads that ended in this stripe ------------------------
        {
          unsigned long long m = ends;
          for (int k = 0; k < n_ring + (end63 ? 1 : 0); ++k) {
            const int le = __builtin_ctzll(m);
            m &= m - 1;
            double v = 0.0;
#pragma unroll
            for (int j = 0; j < kStackEnds; ++j)
              if (j == k) v = acc_k[j];
            if (le == 63) v = acc63;
            const double S = wave_sum_d(v);
            if (lane == 0) {
              const int read = tk.read0 + rd + k;
              const int64_t o = tk.out0 + static_cast<int64_t>(rd + k) * tk.stride + h;
              if (S < static_cast<double>(kMinAccepted) && a.rescue != nullptr) {
                const int slot_r = atomicAdd(a.rescue_count, 1);
                a.rescue[slot_r] = Task{read, hap, 1, -1, o};
                a.form[o] = kFormPending;
                const unsigned long long cells =
                    static_cast<unsigned long long>(a.read_off[read + 1] - a.read_off[read]) *
                    static_cast<unsigned long long>(H);
                atomicAdd(a.rescue_cells, cells);
              } else {
                a.out[o] = log10(S) - kLog10K32;
                a.form[o] = form_ok;
              }
            }
          }
        }
        rd += n_ring + (end63 ? 1 : 0);
        starts_read = ((ends >> L) & 1ull) != 0;
        ps += L;
      }
    }
  }
  if (lane == 0 && cells_w != 0 && a.cells_done != nullptr) atomicAdd(a.cells_done, cells_w);
}

// ---- launchers (host) ----------------------------------------------------

size_t stack_lds_bytes(int carry_cols) {
  return offsetof(StackLds, carry) + sizeof(float2) * static_cast<size_t>(kStackPad + carry_cols);
}

hipError_t launch_stack(const KArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(pairhmm_stack_kernel, dim3(grid), dim3(64), stack_lds_bytes(a.carry_cols), st, a);
  return hipGetLastError();
}

int stack_blocks_per_cu(int carry_cols) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pairhmm_stack_kernel, 64, stack_lds_bytes(carry_cols)) !=
      hipSuccess)
    return 4;
  return n > 0 ? n : 1;
}

}  // namespace bvp
