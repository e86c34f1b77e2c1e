// MI355X (gfx950) Smith-Waterman alignment of haplotypes to a reference
// window, the integer DP upstream of PairHMM in the haplotype caller
// (SmithWaterman::align, include/biovoltron/algo/align/inexact_match/
// smithwaterman.hpp:54-258; called from haplotype_graph.hpp:219).
//
// Per pair: reference bases are rows i = 1..R, alternate bases columns
// j = 1..H; row and column 0 score 0 and scores are not clamped at zero.
//   diag  = S[i-1][j-1] + (ref[i-1] == alt[j-1] ? w_match : w_mismatch)
//   down  = max(S[i-1][j] + w_open, down_prev[j] + w_extend) per column,
//           its length reset to 1 when the open wins (strictly), else + 1
//   right = the same per row from S[i][j-1]
//   S = diag if diag >= down and diag >= right, else right if right >= down,
//       else down; trace 0 / -right length / +down length
// The end cell, the traceback and the CIGAR follow calculate_cigar exactly
// (:135-222), and equal-length pairs with at most 2 mismatches take the
// reference's shortcut (:54-60, :238-239).
//
// Layout: one wavefront per pair (persistent grid, pairs from an atomic
// queue over the launch's size class: the host launches pairs of up to 512,
// 1024, 2048 and 4096 bases separately, so that the per-wavefront scratch,
// sized for the class's longest pair, leaves room for a full grid).  The wavefront sweeps 64-row stripes by anti-diagonals: lane l holds
// row r0 + l and at step t computes column t - l + 1.  The row above comes from
// the lane above by a DPP shift of its previous step's (S, down, down length);
// lane 0 reads the previous stripe's last row from a carry row in the
// wavefront's scratch.  Trace values go to scratch in anti-diagonal-major order
// (cell (i, j) at (i + j) * (R + 1) + i), so a step's stores are contiguous.
// Lane 0 then finds the end cell and the wavefront walks the trace, 64 cells
// of a diagonal per load.  Integer work: results are bit-identical to the
// reference's.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "sw_common.h"

namespace bvp {

__device__ __forceinline__ int sw_shr(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x138 /*wave_shr:1*/, 0xf, 0xf, true);
}

__global__ __launch_bounds__(64) void sw_align_kernel(SwArgs a) {
  const int lane = threadIdx.x;
  char* scr = a.scratch + static_cast<int64_t>(blockIdx.x) * a.scratch_stride;
  for (;;) {
    int q = 0;
    if (lane == 0) q = atomicAdd(a.queue, 1);
    q = __builtin_amdgcn_readfirstlane(__shfl(q, 0));
    if (q >= a.n_pairs) break;
    const int pk = a.order[q];
    const int64_t ro = a.ref_off[pk], ao = a.alt_off[pk];
    const int R = static_cast<int>(a.ref_off[pk + 1] - ro);
    const int H = static_cast<int>(a.alt_off[pk + 1] - ao);
    const uint8_t* ref = a.ref + ro;
    const uint8_t* alt = a.alt + ao;
    const int64_t slot = a.out_off[pk];
    // well_match shortcut (equal lengths, at most 2 mismatches)
    if (R == H) {
      int mism = 0;
      for (int p0 = 0; p0 < R && mism <= 2; p0 += 64) {
        const int p = p0 + lane;
        mism += __popcll(__ballot(p < R && ref[p] != alt[p]));
      }
      if (mism <= 2) {
        if (lane == 0) {
          a.offset[pk] = 0;
          a.n_elems[pk] = 1;
          a.ops[slot] = 'M';
          a.lens[slot] = R;
        }
        continue;
      }
    }
    // scratch: trace (int16), carry rows S / down / down length, last row,
    // last column
    int16_t* trace = reinterpret_cast<int16_t*>(scr);
    const int64_t tr_bytes = ((static_cast<int64_t>(R + H + 1) * (R + 1) * 2 + 15) / 16) * 16;
    int* cS = reinterpret_cast<int*>(scr + tr_bytes);
    int* cB = cS + (H + 1);
    int* cG = cB + (H + 1);
    int* lastrow = cG + (H + 1);
    int* lastcol = lastrow + (H + 1);
    const int nstripes = (R + 63) / 64;
    for (int st = 0; st < nstripes; ++st) {
      const int r0 = 64 * st + 1;
      const int i = r0 + lane;
      const bool row_ok = i <= R;
      const uint8_t rb = row_ok ? ref[i - 1] : uint8_t(0);
      int s_left = 0, up_prev = 0;  // S[i][j-1], S[i-1][j-1]
      int bgr = INT_MIN / 2, gsr = 0;
      int s_out = 0, b_out = INT_MIN / 2, g_out = 0;  // this lane's last cell
      // the carry row for the next 64 columns lane 0 reaches, one per lane
      // (loaded every 64 steps, handed to lane 0 by v_readlane)
      int pS = 0, pB = INT_MIN / 2, pG = 0;
      const int steps = H + 63;
      for (int t = 0; t < steps; ++t) {
        const int j = t - lane + 1;
        const int k = t & 63;
        if (k == 0 && st > 0) {
          // lane 63 of this stripe has written columns up to t - 62 so far,
          // all below the columns t + 1 .. t + 64 read here
          const int jc = t + 1 + lane;
          if (jc <= H) {
            pS = cS[jc];
            pB = cB[jc];
            pG = cG[jc];
          }
        }
        // the row above at column j: the lane above's previous step, or the
        // carry row for lane 0
        int sU = sw_shr(s_out), bU = sw_shr(b_out), gU = sw_shr(g_out);
        const int cs = __builtin_amdgcn_readlane(pS, k), cb = __builtin_amdgcn_readlane(pB, k),
                  cg = __builtin_amdgcn_readlane(pG, k);
        if (lane == 0) {
          if (st == 0) {
            sU = 0;
            bU = INT_MIN / 2;
            gU = 0;
          } else {
            sU = cs;
            bU = cb;
            gU = cg;
          }
        }
        if (row_ok && j >= 1 && j <= H) {
          const int diag = up_prev + (rb == alt[j - 1] ? a.wm : a.wx);
          int bgd = bU + a.we, gsd;
          const int open_d = sU + a.wo;
          if (open_d > bgd) {
            bgd = open_d;
            gsd = 1;
          } else {
            gsd = gU + 1;
          }
          bgr += a.we;
          const int open_r = s_left + a.wo;
          if (open_r > bgr) {
            bgr = open_r;
            gsr = 1;
          } else {
            ++gsr;
          }
          int s, tv;
          if (diag >= bgd && diag >= bgr) {
            s = diag;
            tv = 0;
          } else if (bgr >= bgd) {
            s = bgr;
            tv = -gsr;
          } else {
            s = bgd;
            tv = gsd;
          }
          trace[static_cast<int64_t>(i + j) * (R + 1) + i] = static_cast<int16_t>(tv);
          if (i == R) lastrow[j] = s;
          if (j == H) lastcol[i] = s;
          if (lane == 63) {
            cS[j] = s;
            cB[j] = bgd;
            cG[j] = gsd;
          }
          up_prev = sU;
          s_left = s;
          s_out = s;
          b_out = bgd;
          g_out = gsd;
        }
      }
      // the stripe's carry row (lane 63) before the next stripe's lane 0
      // reads it; lane 63 of this stripe finishes its last column at the
      // last step, so the reads above never run ahead of the writes
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    // End cell (calculate_cigar :141-172), in the reference's order: the
    // last column and then the last row are read 64 entries per load (one
    // per lane) and scanned in step by every lane through v_readlane.
    int pi = 0, pj = H, seg = 0;
    {
      int best = INT_MIN;
      for (int r0 = 1; r0 <= R; r0 += 64) {
        const int v = r0 + lane <= R ? lastcol[r0 + lane] : 0;
        const int m = min(64, R - r0 + 1);
        for (int k = 0; k < m; ++k) {
          const int x = __builtin_amdgcn_readlane(v, k);
          if (x >= best) {
            best = x;
            pi = r0 + k;
          }
        }
      }
      for (int c0 = 1; c0 <= H; c0 += 64) {
        const int v = c0 + lane <= H ? lastrow[c0 + lane] : 0;
        const int m = min(64, H - c0 + 1);
        for (int k = 0; k < m; ++k) {
          const int x = __builtin_amdgcn_readlane(v, k);
          const int c = c0 + k;
          const int dn = R > c ? R - c : c - R, dc = pi > pj ? pi - pj : pj - pi;
          if (x > best || (x == best && dn < dc)) {
            best = x;
            pi = R;
            pj = c;
            seg = H - c;
          }
        }
      }
    }
    // Walk (:174-210) by the whole wavefront in step: lane k holds the trace
    // of cell (bi - k, bj - k), 64 cells down the current diagonal loaded in
    // one round trip, and the walk reads them by v_readlane until a gap
    // takes it off that diagonal.  Lane 0 writes the elements, in reverse
    // order, then reverses them.
    char* ops = a.ops + slot;
    int32_t* lens = a.lens + slot;
    int n = 0;
    if (seg > 0) {
      if (lane == 0) {
        ops[n] = 'S';
        lens[n] = seg;
      }
      ++n;
      seg = 0;
    }
    char state = 'M';
    int bi = -1, bj = -1, win = 0;
    do {
      if (pi - pj != bi - bj || bi - pi < 0 || bi - pi >= 64) {
        const int ci = pi - lane, cj = pj - lane;
        win = (ci >= 1 && cj >= 1) ? trace[static_cast<int64_t>(ci + cj) * (R + 1) + ci] : 0;
        bi = pi;
        bj = pj;
      }
      const int tv = __builtin_amdgcn_readlane(win, bi - pi);
      char ns;
      int step;
      if (tv > 0) {
        ns = 'D';
        step = tv;
        pi -= step;
      } else if (tv < 0) {
        ns = 'I';
        step = -tv;
        pj -= step;
      } else {
        ns = 'M';
        step = 1;
        --pi;
        --pj;
      }
      if (ns == state) {
        seg += step;
      } else {
        if (lane == 0) {
          ops[n] = state;
          lens[n] = seg;
        }
        ++n;
        seg = step;
        state = ns;
      }
    } while (pi > 0 && pj > 0);
    if (lane == 0) {
      ops[n] = state;
      lens[n] = seg;
    }
    ++n;
    if (pj > 0) {
      if (lane == 0) {
        ops[n] = 'S';
        lens[n] = pj;
      }
      ++n;
    }
    if (lane == 0) {
      a.offset[pk] = pi;
      for (int k = 0; k < n / 2; ++k) {
        const char o = ops[k];
        ops[k] = ops[n - 1 - k];
        ops[n - 1 - k] = o;
        const int32_t l = lens[k];
        lens[k] = lens[n - 1 - k];
        lens[n - 1 - k] = l;
      }
      a.n_elems[pk] = n;
    }
    // the scratch is reused by the next pair
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
}

// Scratch bytes one wavefront needs for pairs of at most R reference and H
// alternate bases.
int64_t sw_scratch_bytes(int R, int H) {
  const int64_t tr = ((static_cast<int64_t>(R + H + 1) * (R + 1) * 2 + 15) / 16) * 16;
  return tr + 4 * (4 * static_cast<int64_t>(H + 1) + (R + 1)) + 64;
}

hipError_t launch_sw(const SwArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(sw_align_kernel, dim3(grid), dim3(64), 0, st, a);
  return hipGetLastError();
}

}  // namespace bvp
