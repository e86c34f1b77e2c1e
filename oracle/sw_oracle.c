/* Smith-Waterman oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of SmithWaterman::align
 * (include/biovoltron/algo/align/inexact_match/smithwaterman.hpp):
 *   :54-60   well_match: equal lengths and at most 2 mismatches give
 *            offset 0 and "<len>M" without a matrix;
 *   :68-125  calculate_matrix: rows are reference bases i = 1..R, columns
 *            alternate bases j = 1..H, row and column 0 score 0 (no clamp at
 *            zero); a vertical gap state per column and a horizontal one per
 *            row, opened at score + w_open when that beats the extended gap
 *            (strictly), else extended by w_extend; the move priority is
 *            diagonal > right (insertion) > down (deletion) with >= ties;
 *            trace 0 / -size / +size;
 *   :135-222 calculate_cigar: the end cell is the last maximum of the last
 *            column (>=), replaced by a cell of the last row with a larger
 *            score, or an equal one nearer the diagonal; a soft clip for the
 *            alternate bases past it; the walk emits an element whenever the
 *            state changes (including a first zero-length 'M' when the first
 *            move is a gap, as the reference does), a soft clip for the
 *            alternate bases before the start, then reverses.
 * Elements are written in CIGAR order to ops/lens (capacity R + H + 3). */
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>

#include "pairhmm_oracle.h"

int bvo_sw_align(int R, const uint8_t* ref, int H, const uint8_t* alt, const int32_t* params,
                 int32_t* offset, char* ops, int32_t* lens) {
  const int wm = params[0], wx = params[1], wo = params[2], we = params[3];
  if (R == H) {
    int mism = 0;
    for (int i = 0; mism <= 2 && i < R; ++i) mism += alt[i] != ref[i];
    if (mism <= 2) {
      *offset = 0;
      ops[0] = 'M';
      lens[0] = R;
      return 1;
    }
  }
  const size_t W = (size_t)H + 1;
  int* S = (int*)calloc((size_t)(R + 1) * W, sizeof(int));
  int* T = (int*)calloc((size_t)(R + 1) * W, sizeof(int));
  int* bgd = (int*)malloc((W + 1) * sizeof(int));
  int* gsd = (int*)calloc(W + 1, sizeof(int));
  for (size_t j = 0; j <= W; ++j) bgd[j] = INT_MIN / 2;
  for (int i = 1; i <= R; ++i) {
    int bgr = INT_MIN / 2, gsr = 0;
    for (int j = 1; j <= H; ++j) {
      const int diag = S[(size_t)(i - 1) * W + j - 1] + (ref[i - 1] == alt[j - 1] ? wm : wx);
      const int open_d = S[(size_t)(i - 1) * W + j] + wo;
      bgd[j] += we;
      if (open_d > bgd[j]) {
        bgd[j] = open_d;
        gsd[j] = 1;
      } else {
        gsd[j]++;
      }
      const int open_r = S[(size_t)i * W + j - 1] + wo;
      bgr += we;
      if (open_r > bgr) {
        bgr = open_r;
        gsr = 1;
      } else {
        gsr++;
      }
      int s, t;
      if (diag >= bgd[j] && diag >= bgr) {
        s = diag;
        t = 0;
      } else if (bgr >= bgd[j]) {
        s = bgr;
        t = -gsr;
      } else {
        s = bgd[j];
        t = gsd[j];
      }
      S[(size_t)i * W + j] = s;
      T[(size_t)i * W + j] = t;
    }
  }
  int best = INT_MIN, pi = 0, pj = H, seg = 0;
  for (int i = 1; i <= R; ++i)
    if (S[(size_t)i * W + H] >= best) {
      best = S[(size_t)i * W + H];
      pi = i;
    }
  for (int j = 1; j <= H; ++j) {
    const int s = S[(size_t)R * W + j];
    const int dn = R > j ? R - j : j - R, dc = pi > pj ? pi - pj : pj - pi;
    if (s > best || (s == best && dn < dc)) {
      best = s;
      pi = R;
      pj = j;
      seg = H - j;
    }
  }
  int n = 0;  /* elements in reverse order */
  char* rops = (char*)malloc((size_t)(R + H + 3));
  int32_t* rlens = (int32_t*)malloc(sizeof(int32_t) * (size_t)(R + H + 3));
  if (seg > 0) {
    rops[n] = 'S';
    rlens[n++] = seg;
    seg = 0;
  }
  char state = 'M';
  do {
    const int t = T[(size_t)pi * W + pj];
    char ns;
    int step;
    if (t > 0) {
      ns = 'D';
      step = t;
      pi -= step;
    } else if (t < 0) {
      ns = 'I';
      step = -t;
      pj -= step;
    } else {
      ns = 'M';
      step = 1;
      pi--;
      pj--;
    }
    if (ns == state) {
      seg += step;
    } else {
      rops[n] = state;
      rlens[n++] = seg;
      seg = step;
      state = ns;
    }
  } while (pi > 0 && pj > 0);
  rops[n] = state;
  rlens[n++] = seg;
  *offset = pi;
  if (pj > 0) {
    rops[n] = 'S';
    rlens[n++] = pj;
  }
  for (int k = 0; k < n; ++k) {
    ops[k] = rops[n - 1 - k];
    lens[k] = rlens[n - 1 - k];
  }
  free(rops);
  free(rlens);
  free(S);
  free(T);
  free(bgd);
  free(gsd);
  return n;
}
