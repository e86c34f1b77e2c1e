/*
 * PairHMM oracle -- TEST INFRASTRUCTURE ONLY (see pairhmm_oracle.h).
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include "pairhmm_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>

/* Context.h:7-11 */
#define JAC_STEP 0.0001
#define JAC_SIZE 80001 /* (int)(MAX_JACOBIAN_TOLERANCE / STEP) + 1 */
#define MAXQ 254
#define MM_SIZE (((MAXQ + 1) * (MAXQ + 2)) >> 1)

static float jac_f[JAC_SIZE];
static double jac_d[JAC_SIZE];
static float mm_f[MM_SIZE];
static double mm_d[MM_SIZE];
static float ph2pr_f[128];
static double ph2pr_d[128];
static uint8_t base_code[256];
static int g_ready = 0;

/* Context.h:63-90: log10(10^a + 10^b) with a quantised Jacobian correction.
 * The float context evaluates it in float (NUMBER = float). */
static float log10sum_f(float a, float b) {
  float lo = a, hi = b;
  if (lo > hi) { lo = b; hi = a; }
  float diff = hi - lo;
  if (diff >= 8.0f) return hi;
  float x = diff * (float)(1.0 / JAC_STEP);
  int k = (x > 0.0f) ? (int)(x + 0.5f) : (int)(x - 0.5f);
  return hi + jac_f[k];
}

static double log10sum_d(double a, double b) {
  double lo = a, hi = b;
  if (lo > hi) { lo = b; hi = a; }
  double diff = hi - lo;
  if (diff >= 8.0) return hi;
  double x = diff * (1.0 / JAC_STEP);
  int k = (x > 0.0) ? (int)(x + 0.5) : (int)(x - 0.5);
  return hi + jac_d[k];
}

static inline int mm_index(int iq, int dq) {
  int lo = iq < dq ? iq : dq, hi = iq < dq ? dq : iq;
  return ((hi * (hi + 1)) >> 1) + lo; /* Context.h:123-134 */
}

void bvo_init(void) {
  if (g_ready) return;
  /* Context.h:42-47 */
  for (int k = 0; k < JAC_SIZE; ++k) {
    double v = log10(1.0 + pow(10.0, -((double)k) * JAC_STEP));
    jac_d[k] = v;
    jac_f[k] = (float)v;
  }
  /* Context.h:50-61, once per precision. */
  const double inv_ln10 = 1.0 / log(10.0);
  for (int i = 0, off = 0; i <= MAXQ; off += ++i) {
    for (int j = 0; j <= i; ++j) {
      double sf = (double)log10sum_f((float)(-0.1 * i), (float)(-0.1 * j));
      double lf = log1p(-fmin(1.0, pow(10.0, sf))) * inv_ln10;
      mm_f[off + j] = (float)pow(10.0, lf);
      double sd = log10sum_d(-0.1 * i, -0.1 * j);
      double ld = log1p(-fmin(1.0, pow(10.0, sd))) * inv_ln10;
      mm_d[off + j] = pow(10.0, ld);
    }
  }
  /* Context.h:105-107,145-147 */
  for (int x = 0; x < 128; ++x) {
    ph2pr_d[x] = pow(10.0, -((double)x) / 10.0);
    ph2pr_f[x] = powf(10.f, -((float)x) / 10.f);
  }
  /* pairhmm_common.h:26-44: A0 C1 T2 G3 N4, every other byte 0. */
  memset(base_code, 0, sizeof(base_code));
  base_code['A'] = 0;
  base_code['C'] = 1;
  base_code['T'] = 2;
  base_code['G'] = 3;
  base_code['N'] = 4;
  g_ready = 1;
}

float bvo_ph2pr_f32(int q) { bvo_init(); return ph2pr_f[q & 127]; }
double bvo_ph2pr_f64(int q) { bvo_init(); return ph2pr_d[q & 127]; }
float bvo_mm_f32(int iq, int dq) { bvo_init(); return mm_f[mm_index(iq & 127, dq & 127)]; }
double bvo_mm_f64(int iq, int dq) { bvo_init(); return mm_d[mm_index(iq & 127, dq & 127)]; }

/* ---- stripe restatement, FP32 and FP64 -------------------------------- */
#define T float
#define FN prob_stripe_f32
#define INIT_K ldexpf(1.f, 120)
#define MM_TAB mm_f
#define PH2PR ph2pr_f
#include "dp_stripe.inc"
#undef T
#undef FN
#undef INIT_K
#undef MM_TAB
#undef PH2PR

#define T double
#define FN prob_stripe_f64
#define INIT_K ldexp(1.0, 1020)
#define MM_TAB mm_d
#define PH2PR ph2pr_d
#include "dp_stripe.inc"
#undef T
#undef FN
#undef INIT_K
#undef MM_TAB
#undef PH2PR

static unsigned set_ftz(int on) {
  unsigned old = _mm_getcsr();
  if (on) _mm_setcsr(old | _MM_FLUSH_ZERO_ON);
  return old;
}

float bvo_prob_f32(int R, int H, const uint8_t* rs, const uint8_t* q,
                   const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                   const uint8_t* hap, int ftz) {
  bvo_init();
  unsigned old = set_ftz(ftz);
  float s = prob_stripe_f32(R, H, rs, q, ins, del, gcp, hap);
  _mm_setcsr(old);
  return s;
}

double bvo_prob_f64(int R, int H, const uint8_t* rs, const uint8_t* q,
                    const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                    const uint8_t* hap, int ftz) {
  bvo_init();
  unsigned old = set_ftz(ftz);
  double s = prob_stripe_f64(R, H, rs, q, ins, del, gcp, hap);
  _mm_setcsr(old);
  return s;
}

/* pairhmm_avx.hpp:140-148 and pairhmm_common.h:16 (MIN_ACCEPTED). */
double bvo_log10_reference(int R, int H, const uint8_t* rs, const uint8_t* q,
                           const uint8_t* ins, const uint8_t* del,
                           const uint8_t* gcp, const uint8_t* hap, int ftz,
                           int* rescued) {
  bvo_init();
  float sf = bvo_prob_f32(R, H, rs, q, ins, del, gcp, hap, ftz);
  if (sf < 1e-28f) {
    double sd = bvo_prob_f64(R, H, rs, q, ins, del, gcp, hap, ftz);
    if (rescued) *rescued = 1;
    return log10(sd) - log10(ldexp(1.0, 1020));
  }
  if (rescued) *rescued = 0;
  return (double)(log10f(sf) - log10f(ldexpf(1.f, 120)));
}

/* ---- FP64 textbook DP with per-row renormalisation -------------------- */

typedef struct {
  double mm, im, dm, mi, ii, md, dd, pm, pmis;
  uint8_t code; /* mode 0: base code; mode 1: raw byte */
} row_par;

static void row_params(int mode, int r, const uint8_t* rs, const uint8_t* q,
                       const uint8_t* ins, const uint8_t* del,
                       const uint8_t* gcp, row_par* p) {
  if (mode == 0) {
    int iq = ins[r] & 127, dq = del[r] & 127, cq = gcp[r] & 127, qq = q[r] & 127;
    p->mm = mm_d[mm_index(iq, dq)];
    p->im = 1.0 - ph2pr_d[cq];
    p->dm = p->im;
    p->mi = ph2pr_d[iq];
    p->ii = ph2pr_d[cq];
    p->md = ph2pr_d[dq];
    p->dd = ph2pr_d[cq];
    double e = ph2pr_d[qq];
    p->pm = 1.0 - e;
    p->pmis = e / 3.0;
    p->code = base_code[rs[r]];
  } else {
    /* pairhmm.hpp:32-33 constants, quality_utils.hpp:25-28,71-76 priors. */
    p->mm = 0.9998; p->mi = 0.0001; p->md = 0.0001;
    p->im = 0.9; p->ii = 0.1; p->dm = 0.9; p->dd = 0.1;
    int qq = (int)(int8_t)q[r];
    double e = pow(10.0, qq / -10.0);
    p->pm = 1.0 - e;
    p->pmis = e / 3.0;
    p->code = rs[r];
  }
}

static inline int is_match(int mode, uint8_t rc, uint8_t hb) {
  if (mode == 0) {
    uint8_t hc = base_code[hb];
    return rc == hc || rc == 4 || hc == 4;
  }
  return rc == hb || rc == 'N' || hb == 'N';
}

double bvo_log10_exact(int R, int H, const uint8_t* rs, const uint8_t* q,
                       const uint8_t* ins, const uint8_t* del,
                       const uint8_t* gcp, const uint8_t* hap, int mode) {
  bvo_init();
  double* buf = (double*)calloc((size_t)6 * (size_t)(H + 1), sizeof(double));
  double *Mp = buf, *Ip = buf + (H + 1), *Dp = buf + 2 * (H + 1);
  double *Mc = buf + 3 * (H + 1), *Ic = buf + 4 * (H + 1), *Dc = buf + 5 * (H + 1);
  for (int j = 0; j <= H; ++j) Dp[j] = 1.0 / (double)H;
  long log2scale = 0; /* values are stored times 2^-log2scale */
  for (int r = 1; r <= R; ++r) {
    row_par p;
    row_params(mode, r - 1, rs, q, ins, del, gcp, &p);
    Mc[0] = Ic[0] = Dc[0] = 0.0;
    double mx = 0.0;
    for (int j = 1; j <= H; ++j) {
      double prior = is_match(mode, p.code, hap[j - 1]) ? p.pm : p.pmis;
      Mc[j] = (Mp[j - 1] * p.mm + Ip[j - 1] * p.im + Dp[j - 1] * p.dm) * prior;
      Ic[j] = Mp[j] * p.mi + Ip[j] * p.ii;
      Dc[j] = Mc[j - 1] * p.md + Dc[j - 1] * p.dd;
      double v = Mc[j] > Ic[j] ? Mc[j] : Ic[j];
      v = v > Dc[j] ? v : Dc[j];
      mx = mx > v ? mx : v;
    }
    if (mx > 0.0) {
      int e;
      frexp(mx, &e);
      for (int j = 0; j <= H; ++j) {
        Mc[j] = ldexp(Mc[j], -e);
        Ic[j] = ldexp(Ic[j], -e);
        Dc[j] = ldexp(Dc[j], -e);
      }
      log2scale += e;
    }
    double* t;
    t = Mp; Mp = Mc; Mc = t;
    t = Ip; Ip = Ic; Ic = t;
    t = Dp; Dp = Dc; Dc = t;
  }
  double s = 0.0;
  for (int j = 1; j <= H; ++j) s += Mp[j] + Ip[j];
  free(buf);
  return log10(s) + (double)log2scale * log10(2.0);
}

/* pairhmm.hpp:55-80: no renormalisation, K = 2^1020. */
double bvo_scalar_log10(int R, int H, const uint8_t* rs, const uint8_t* q,
                        const uint8_t* hap, const double* tr) {
  bvo_init();
  double* buf = (double*)calloc((size_t)6 * (size_t)(H + 1), sizeof(double));
  double *Mp = buf, *Ip = buf + (H + 1), *Dp = buf + 2 * (H + 1);
  double *Mc = buf + 3 * (H + 1), *Ic = buf + 4 * (H + 1), *Dc = buf + 5 * (H + 1);
  const double K = pow(2.0, 1020);
  for (int j = 0; j <= H; ++j) Dp[j] = K / (double)H;
  for (int r = 1; r <= R; ++r) {
    int qq = (int)(int8_t)q[r - 1];
    double e = pow(10.0, qq / -10.0);
    uint8_t x = rs[r - 1];
    Mc[0] = Ic[0] = Dc[0] = 0.0;
    for (int j = 1; j <= H; ++j) {
      uint8_t y = hap[j - 1];
      double prior = (x == y || x == 'N' || y == 'N') ? 1.0 - e : e / 3.0;
      Mc[j] = prior * (Mp[j - 1] * tr[0] + Ip[j - 1] * tr[3] + Dp[j - 1] * tr[5]);
      Ic[j] = Mp[j] * tr[1] + Ip[j] * tr[4];
      Dc[j] = Mc[j - 1] * tr[2] + Dc[j - 1] * tr[6];
    }
    double* t;
    t = Mp; Mp = Mc; Mc = t;
    t = Ip; Ip = Ic; Ic = t;
    t = Dp; Dp = Dc; Dc = t;
  }
  double s = 0.0;
  for (int j = 1; j <= H; ++j) s += Mp[j] + Ip[j];
  free(buf);
  return log10(s) - log10(K);
}

/* ---- forward-backward posteriors (FP64, per-row scaling) -------------- */

double bvo_band_posterior(int R, int H, const uint8_t* rs, const uint8_t* q,
                          const uint8_t* ins, const uint8_t* del,
                          const uint8_t* gcp, const uint8_t* hap,
                          double* mass, double* row_sum) {
  bvo_init();
  const size_t W = (size_t)H + 2;
  const size_t N = ((size_t)R + 2) * W;
  double* FM = (double*)calloc(N, sizeof(double));
  double* FI = (double*)calloc(N, sizeof(double));
  double* FD = (double*)calloc(N, sizeof(double));
  double* BM = (double*)calloc(N, sizeof(double));
  double* BI = (double*)calloc(N, sizeof(double));
  double* BD = (double*)calloc(N, sizeof(double));
  row_par* P = (row_par*)malloc(sizeof(row_par) * (size_t)(R + 2));
  double* fs = (double*)calloc((size_t)R + 2, sizeof(double)); /* log2 scale of forward row r */
  double* bs = (double*)calloc((size_t)R + 2, sizeof(double)); /* log2 scale of backward row r */
#define AT(a, r, j) a[(size_t)(r) * W + (size_t)(j)]
  for (int r = 1; r <= R; ++r) row_params(0, r - 1, rs, q, ins, del, gcp, &P[r]);
  /* forward */
  for (int j = 0; j <= H; ++j) AT(FD, 0, j) = 1.0 / (double)H;
  for (int r = 1; r <= R; ++r) {
    double mx = 0.0;
    for (int j = 1; j <= H; ++j) {
      double prior = is_match(0, P[r].code, hap[j - 1]) ? P[r].pm : P[r].pmis;
      AT(FM, r, j) = (AT(FM, r - 1, j - 1) * P[r].mm + AT(FI, r - 1, j - 1) * P[r].im +
                      AT(FD, r - 1, j - 1) * P[r].dm) * prior;
      AT(FI, r, j) = AT(FM, r - 1, j) * P[r].mi + AT(FI, r - 1, j) * P[r].ii;
      AT(FD, r, j) = AT(FM, r, j - 1) * P[r].md + AT(FD, r, j - 1) * P[r].dd;
      double v = fmax(AT(FM, r, j), fmax(AT(FI, r, j), AT(FD, r, j)));
      mx = fmax(mx, v);
    }
    int e = 0;
    if (mx > 0.0) frexp(mx, &e);
    for (int j = 0; j <= H; ++j) {
      AT(FM, r, j) = ldexp(AT(FM, r, j), -e);
      AT(FI, r, j) = ldexp(AT(FI, r, j), -e);
      AT(FD, r, j) = ldexp(AT(FD, r, j), -e);
    }
    fs[r] = fs[r - 1] + e;
  }
  double total = 0.0;
  for (int j = 1; j <= H; ++j) total += AT(FM, R, j) + AT(FI, R, j);
  double log2L = log2(total) + fs[R];
  /* backward: B(r,j) = P(rows r+1..R emitted | state at (r,j)); the end
   * condition is B_M(R,j) = B_I(R,j) = 1, B_D(R,j) = 0. */
  for (int j = 1; j <= H; ++j) { AT(BM, R, j) = 1.0; AT(BI, R, j) = 1.0; }
  bs[R] = 0.0;
  for (int r = R - 1; r >= 1; --r) {
    double mx = 0.0;
    for (int j = H; j >= 1; --j) {
      double pr_next = (j + 1 <= H)
          ? (is_match(0, P[r + 1].code, hap[j]) ? P[r + 1].pm : P[r + 1].pmis) : 0.0;
      double mnext = (j + 1 <= H) ? AT(BM, r + 1, j + 1) * pr_next : 0.0;
      double dnext = (j + 1 <= H) ? AT(BD, r, j + 1) : 0.0;
      AT(BM, r, j) = P[r + 1].mm * mnext + P[r + 1].mi * AT(BI, r + 1, j) + P[r].md * dnext;
      AT(BI, r, j) = P[r + 1].im * mnext + P[r + 1].ii * AT(BI, r + 1, j);
      AT(BD, r, j) = P[r + 1].dm * mnext + P[r].dd * dnext;
      mx = fmax(mx, fmax(AT(BM, r, j), fmax(AT(BI, r, j), AT(BD, r, j))));
    }
    int e = 0;
    if (mx > 0.0) frexp(mx, &e);
    for (int j = 1; j <= H; ++j) {
      AT(BM, r, j) = ldexp(AT(BM, r, j), -e);
      AT(BI, r, j) = ldexp(AT(BI, r, j), -e);
      AT(BD, r, j) = ldexp(AT(BD, r, j), -e);
    }
    bs[r] = bs[r + 1] + e;
  }
  for (int o = 0; o < R + H - 1; ++o) mass[o] = 0.0;
  for (int r = 1; r <= R; ++r) {
    double rsum = 0.0;
    double sc = exp2(fs[r] + bs[r] - log2L);
    for (int j = 1; j <= H; ++j) {
      double pmi = (AT(FM, r, j) * AT(BM, r, j) + AT(FI, r, j) * AT(BI, r, j)) * sc;
      double pd = AT(FD, r, j) * AT(BD, r, j) * sc;
      rsum += pmi;
      mass[(j - r) + R - 1] += pmi + pd;
    }
    if (row_sum) row_sum[r - 1] = rsum;
  }
  /* mass[o] is the expected number of visits of the alignment path to
   * diagonal o - (R - 1) (a run of deletions visits several cells of one row,
   * so the total exceeds R). */
#undef AT
  free(FM); free(FI); free(FD); free(BM); free(BI); free(BD);
  free(P); free(fs); free(bs);
  return log2L * log10(2.0);
}

/* ---- region driver (pairhmm_avx.hpp:121-157) --------------------------- */

long bvo_region_log10(int n_reads, const int64_t* read_off, const uint8_t* rs,
                      const uint8_t* q, const uint8_t* ins, const uint8_t* del,
                      const uint8_t* gcp, int n_haps, const int64_t* hap_off,
                      const uint8_t* hap, int n_threads, double* out) {
  bvo_init();
  long rescued = 0;
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads) reduction(+ : rescued)
  for (int i = 0; i < n_reads; ++i) {
    /* FTZ on every worker, as the reference's calling thread has it. */
    _mm_setcsr(_mm_getcsr() | _MM_FLUSH_ZERO_ON);
    const int64_t ro = read_off[i];
    const int R = (int)(read_off[i + 1] - ro);
    for (int h = 0; h < n_haps; ++h) {
      const int64_t ho = hap_off[h];
      const int H = (int)(hap_off[h + 1] - ho);
      int resc = 0;
      out[(size_t)i * (size_t)n_haps + (size_t)h] = bvo_log10_reference(
          R, H, rs + ro, q + ro, ins + ro, del + ro, gcp + ro, hap + ho, 1, &resc);
      rescued += resc;
    }
  }
  return rescued;
}

/* ---- banded FP64 forward with the edge-zone touch mass ------------------
 * The band of the banded kernel (biovoltron_amd/csrc/pairhmm_band.hip):
 * cells with d = j - r in [o_lo, o_lo + width - 1]; outside cells are zero,
 * row 0 included (D(0, j) = 1/H only for j in the band).  The touch DP T has
 * the forward recurrence but equals the forward value on edge-zone cells:
 * the `edge` lowest diagonals on rows where the column left of the band is a
 * real cell (r + o_lo - 1 >= 1; row 0: o_lo - 1 >= 0), and the `edge` highest
 * ones while the column right of it is (r + o_lo + width <= H).  Rows are
 * renormalised by a power of two.  Returns log10 of the banded likelihood
 * (-inf when it is zero) and, in touch_ratio, T's final sum over S. */
double bvo_band_forward(int R, int H, const uint8_t* rs, const uint8_t* q,
                        const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                        const uint8_t* hap, int o_lo, int width, int edge,
                        double* touch_ratio) {
  bvo_init();
  const size_t W = (size_t)H + 2;
  double* buf = (double*)calloc(12 * W, sizeof(double));
  double *fM0 = buf, *fI0 = buf + W, *fD0 = buf + 2 * W, *tM0 = buf + 3 * W, *tI0 = buf + 4 * W,
         *tD0 = buf + 5 * W, *fM1 = buf + 6 * W, *fI1 = buf + 7 * W, *fD1 = buf + 8 * W,
         *tM1 = buf + 9 * W, *tI1 = buf + 10 * W, *tD1 = buf + 11 * W;
  const int o_hi = o_lo + width - 1;
  for (int j = 0; j <= H; ++j) {
    if (j < o_lo || j > o_hi) continue;
    fD0[j] = 1.0 / (double)H;
    int zone = (j < o_lo + edge && o_lo - 1 >= 0) || (j > o_hi - edge && o_lo + width <= H);
    if (zone) tD0[j] = fD0[j];
  }
  double log2scale = 0.0;
  for (int r = 1; r <= R; ++r) {
    row_par p;
    row_params(0, r - 1, rs, q, ins, del, gcp, &p);
    const int lact = r + o_lo - 1 >= 1, ract = r + o_lo + width <= H;
    fM1[0] = fI1[0] = fD1[0] = tM1[0] = tI1[0] = tD1[0] = 0.0;
    double mx = 0.0;
    for (int j = 1; j <= H; ++j) {
      const int d = j - r;
      if (d < o_lo || d > o_hi) {
        fM1[j] = fI1[j] = fD1[j] = tM1[j] = tI1[j] = tD1[j] = 0.0;
        continue;
      }
      const int zone = (d < o_lo + edge && lact) || (d > o_hi - edge && ract);
      const double prior = is_match(0, p.code, hap[j - 1]) ? p.pm : p.pmis;
      fM1[j] = (fM0[j - 1] * p.mm + fI0[j - 1] * p.im + fD0[j - 1] * p.dm) * prior;
      fI1[j] = fM0[j] * p.mi + fI0[j] * p.ii;
      fD1[j] = fM1[j - 1] * p.md + fD1[j - 1] * p.dd;
      if (zone) {
        tM1[j] = fM1[j];
        tI1[j] = fI1[j];
        tD1[j] = fD1[j];
      } else {
        tM1[j] = (tM0[j - 1] * p.mm + tI0[j - 1] * p.im + tD0[j - 1] * p.dm) * prior;
        tI1[j] = tM0[j] * p.mi + tI0[j] * p.ii;
        tD1[j] = tM1[j - 1] * p.md + tD1[j - 1] * p.dd;
      }
      mx = fmax(mx, fmax(fM1[j], fmax(fI1[j], fD1[j])));
    }
    int e = 0;
    if (mx > 0.0) frexp(mx, &e);
    for (int j = 0; j <= H; ++j) {
      fM0[j] = ldexp(fM1[j], -e); fI0[j] = ldexp(fI1[j], -e); fD0[j] = ldexp(fD1[j], -e);
      tM0[j] = ldexp(tM1[j], -e); tI0[j] = ldexp(tI1[j], -e); tD0[j] = ldexp(tD1[j], -e);
    }
    log2scale += e;
  }
  double S = 0.0, T = 0.0;
  for (int j = 1; j <= H; ++j) {
    S += fM0[j] + fI0[j];
    T += tM0[j] + tI0[j];
  }
  free(buf);
  if (touch_ratio) *touch_ratio = S > 0.0 ? T / S : 1.0;
  if (!(S > 0.0)) return -INFINITY;
  return (log2(S) + log2scale) * log10(2.0);
}

/* Banded FP64 forward whose band moves: `width` diagonals, one array slot per
 * diagonal of the current band.  Before row r the band may move by s
 * diagonals (slot k then holds the cells of slot k + s, the slots that fall
 * out are dropped, the new ones start at zero):
 *  - sched == NULL: when r > 1 and (r - 1) % every == 0, towards the
 *    centroid c of the slot masses M + I + D of row r - 1:
 *    s = clamp(floor(c - t + 0.5), -max_move, max_move), t = (width - 1) / 2,
 *    plus end_bias on rows r > R - end_rows; the band starts at o_init;
 *  - sched != NULL: towards the scheduled offset, s = clamp(sched[r] - o,
 *    -max_move, max_move); the band starts at sched[0].
 * Hold rule (hold_eps > 0): a move whose dropped slots carry more than
 * hold_eps of row r - 1's mass is not made.  The edge zones and the touch DP
 * are those of bvo_band_forward at the band's current position.  offs
 * (optional, R + 1 entries) receives the offset used on each row. */
static double leak_rho = 2.5e-7;
void bvo_set_leak_rho(double rho) { leak_rho = rho; }

double bvo_band_follow_forward(int R, int H, const uint8_t* rs, const uint8_t* q,
                               const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                               const uint8_t* hap, int o_init, int width, int edge, int every,
                               int max_move, int end_rows, int end_bias, double hold_eps,
                               int irun, const int* sched, double* touch_ratio, double* lost,
                               int* offs) {
  bvo_init();
  /* edge < 0: zones of -edge diagonals that are active on every row (the
   * anti-diagonal banded kernel's rule), not only where the band cuts */
  const int always = edge < 0;
  if (always) edge = -edge;
  /* irun bit 2: the anti-diagonal kernel's rule -- T follows the recurrence
   * on zone cells too, except that the innermost zone cells take the
   * forward value of their input from the band's middle (I on diagonal
   * edge - 1, D on width - edge); with T = F on the zones at row 0 that is
   * T = F on every zone cell, and the entry runs added at the top slot then
   * propagate through the zone instead of being overwritten (with edge < 0) */
  const int kernel_rule = (irun & 4) != 0;
  /* irun bit 3: *touch_ratio receives the leak bound E / S instead of T / S.
   * On an active zone cell T = F, so the touched mass can leave the zones'
   * diagonals only by: D from the left zone into diagonal o + edge (rows
   * where the left zone is active), I from the right zone into diagonal
   * o + width - edge - 1 (the row after one where the right zone is active),
   * or a zone cell that the next row does not keep in an active zone of its
   * side (the zone turns off, or a move carries the cell elsewhere).  E adds those F flows, the
   * zones' M + I on the last row (where active) and the exit runs (bit 1).
   * Every continuation is a probability, at most 1 -- times phi, the
   * product over rows of max(1, the out-weight sums of M and D), which is 1
   * when the gap penalties are constant along the read -- so E * phi >= T,
   * and a band that passes with E passes with T. */
  const int leak = (irun & 8) != 0;
  /* irun bits 4-7 drop one term of E each (left leak, right leak, turn-off
   * and move leaks, last-row zone mass): planted mistakes for the tests */
  const int drop = (irun >> 4) & 15;
  /* irun >> 8 = r0: with bit 3, T is carried exactly for rows 1..r0 and its
   * mass on row r0 (every cell, continuation at most 1) starts E; the leak
   * terms count from row r0 + 1 on (r0 = 0: E alone) */
  const int r0 = irun >> 8;
  double phi = 1.0;
  if (leak) {
    for (int r = 1; r < R; ++r) {
      row_par a, b;
      row_params(0, r - 1, rs, q, ins, del, gcp, &a);
      row_params(0, r, rs, q, ins, del, gcp, &b);
      double f;
      if (a.dm > 0.0) {
        const double fd = b.dm / a.dm;
        const double fm = b.mm + b.mi + a.md * fd;
        f = fm > fd ? fm : fd;
      } else {
        f = 1.0 + (double)H; /* D runs along the whole row */
      }
      if (f > 1.0) phi *= f;
    }
  }
  double E = 0.0;
#define LACT_AT(rr, oo) (always || ((rr) == 0 ? (oo) - 1 >= 0 : (rr) + (oo) - 1 >= 1))
#define RACT_AT(rr, oo) (always || (rr) + (oo) + width <= H)
  const size_t W = (size_t)width;
  double* buf = (double*)calloc(12 * W, sizeof(double));
  double *fM = buf, *fI = buf + W, *fD = buf + 2 * W, *tM = buf + 3 * W, *tI = buf + 4 * W,
         *tD = buf + 5 * W, *nM = buf + 6 * W, *nI = buf + 7 * W, *nD = buf + 8 * W,
         *uM = buf + 9 * W, *uI = buf + 10 * W, *uD = buf + 11 * W;
  int o = sched ? sched[0] : o_init;
  double lost_sum = 0.0, log2scale = 0.0;
  /* row 0: D = 1/H on columns 0..H (the scale K is carried as log2scale = 0) */
  for (int k = 0; k < width; ++k) {
    const int j = o + k;
    if (j < 0 || j > H) continue;
    fD[k] = 1.0 / (double)H;
    const int zone = (k < edge && (always || o - 1 >= 0)) || (k >= width - edge && (always || o + width <= H));
    if (zone) tD[k] = fD[k];
  }
  if (offs) offs[0] = o;
  /* irun: insertion runs across the band's long edges.  A read start
   * explained as insertions is a path M(1, c) then I down column c; where c
   * lies above the band it enters at the top slot (row r, column c): that
   * pure run is added to the touch DP there.  A read end explained as
   * insertions leaves through the bottom slot: the I mass that leaves is
   * carried on as pure runs (x, times ii per row) and added to T at the end. */
  row_par p1, p2;
  row_params(0, 0, rs, q, ins, del, gcp, &p1);
  row_params(0, R >= 2 ? 1 : 0, rs, q, ins, del, gcp, &p2);
  double run_prod = 1.0, x_tail = 0.0;
  for (int r = 1; r <= R; ++r) {
    int s = 0;
    double Wt = 0.0, C = 0.0;
    for (int k = 0; k < width; ++k) {
      const double w = fM[k] + fI[k] + fD[k];
      Wt += w;
      C += (double)k * w;
    }
    if (sched) {
      s = sched[r] - o;
    } else if (r > 1 && (r - 1) % every == 0 && Wt > 0.0) {
      const double target = 0.5 * (double)(width - 1) + (r > R - end_rows ? (double)end_bias : 0.0);
      s = (int)floor(C / Wt - target + 0.5);
    }
    if (s > max_move) s = max_move;
    if (s < -max_move) s = -max_move;
    if (s != 0) {
      double dropped = 0.0;
      for (int k = 0; k < width; ++k)
        if (!(k - s >= 0 && k - s < width)) dropped += fM[k] + fI[k] + fD[k];  /* slot k lands on k - s */
      if (hold_eps > 0.0 && dropped > hold_eps * Wt) s = 0;
      else if (Wt > 0.0) lost_sum += dropped / Wt;
    }
    const int lprev = LACT_AT(r - 1, o), rprev = RACT_AT(r - 1, o);
    if (leak && r - 1 >= r0) {
      /* A zone cell of row r - 1 stays tracked only where row r puts it in
       * the same side's zone, active on row r; otherwise (the zone turns
       * off, or a move carries the cell into the interior or the other
       * side's zone) its mass leaves the tracked diagonals.  Cells a move
       * drops from the band are not counted (nor are they in T). */
      const int lnow = LACT_AT(r, o + s), rnow = RACT_AT(r, o + s);
      for (int k = 0; k < width; ++k) {
        const int lz = k < edge && lprev, rz = k >= width - edge && rprev;
        const int kn = k - s;
        if ((!lz && !rz) || kn < 0 || kn >= width) continue;
        const int kept = (lz && kn < edge && lnow) || (rz && kn >= width - edge && rnow);
        if (!kept && !(drop & 4)) E += fM[k] + fI[k] + fD[k];
      }
    }
    if (s != 0) {
      double* arrs[6] = {fM, fI, fD, tM, tI, tD};
      for (int a = 0; a < 6; ++a) {
        double* x = arrs[a];
        if (s > 0) {
          for (int k = 0; k < width; ++k) x[k] = k + s < width ? x[k + s] : 0.0;
        } else {
          for (int k = width - 1; k >= 0; --k) x[k] = k + s >= 0 ? x[k + s] : 0.0;
        }
      }
      o += s;
    }
    if (offs) offs[r] = o;
    row_par p;
    row_params(0, r - 1, rs, q, ins, del, gcp, &p);
    const int lact = always || r + o - 1 >= 1, ract = always || r + o + width <= H;
    double top_in = 0.0;
    if (irun) {
      /* counted on rows 2..48 (entries) and from the last 48 rows (exits):
       * beyond them a run's factor ii^rows is below 1e-46 */
      if (r >= 3) run_prod *= p.ii;
      const int c = r + o + width - 1;
      if ((irun & 5) && r >= 2 && r <= 48 && c >= 1 && c <= H) {
        const double pr1 = is_match(0, p1.code, hap[c - 1]) ? p1.pm : p1.pmis;
        top_in = ldexp(p1.dm * pr1 * p2.mi * run_prod / (double)H, -(int)log2scale);
      }
      if ((irun & 2) && r > R - 48) {
        x_tail *= p.ii;
        const int c0 = r - 1 + o;
        if (c0 >= 1 && c0 <= H) x_tail += fM[0] * p.mi + fI[0] * p.ii;
      }
    }
    double mx = 0.0;
    for (int k = 0; k < width; ++k) {
      const int j = r + o + k;
      if (j < 1 || j > H) {
        nM[k] = nI[k] = nD[k] = uM[k] = uI[k] = uD[k] = 0.0;
        continue;
      }
      const double prior = is_match(0, p.code, hap[j - 1]) ? p.pm : p.pmis;
      const double upM = k + 1 < width ? fM[k + 1] : 0.0, upI = k + 1 < width ? fI[k + 1] : 0.0;
      const double lfM = k >= 1 ? nM[k - 1] : 0.0, lfD = k >= 1 ? nD[k - 1] : 0.0;
      nM[k] = (fM[k] * p.mm + fI[k] * p.im + fD[k] * p.dm) * prior;
      nI[k] = upM * p.mi + upI * p.ii;
      nD[k] = lfM * p.md + lfD * p.dd;
      const int zone = (k < edge && lact) || (k >= width - edge && ract);
      if (zone && !kernel_rule) {
        uM[k] = nM[k];
        uI[k] = nI[k];
        uD[k] = nD[k];
      } else {
        const double tuM = k + 1 < width ? tM[k + 1] : 0.0, tuI = k + 1 < width ? tI[k + 1] : 0.0;
        const double tlM = k >= 1 ? uM[k - 1] : 0.0, tlD = k >= 1 ? uD[k - 1] : 0.0;
        uM[k] = (tM[k] * p.mm + tI[k] * p.im + tD[k] * p.dm) * prior;
        uI[k] = tuM * p.mi + tuI * p.ii;
        uD[k] = tlM * p.md + tlD * p.dd;
      }
      if (kernel_rule) {
        /* the zones' innermost inputs from the band's middle: T = F there */
        if (k == edge - 1) uI[k] = nI[k];
        if (k == width - edge) uD[k] = nD[k];
      }
      mx = fmax(mx, fmax(nM[k], fmax(nI[k], nD[k])));
    }
    uI[width - 1] += top_in;
    if (leak && r > r0) {
      if (lact && edge < width && !(drop & 1)) E += nD[edge];
      if (rprev && width - edge - 1 >= 0 && !(drop & 2)) E += nI[width - edge - 1];
    }
    int e = 0;
    if (mx > 0.0) frexp(mx, &e);
    x_tail = ldexp(x_tail, -e);
    E = ldexp(E, -e);
    for (int k = 0; k < width; ++k) {
      fM[k] = ldexp(nM[k], -e); fI[k] = ldexp(nI[k], -e); fD[k] = ldexp(nD[k], -e);
      tM[k] = ldexp(uM[k], -e); tI[k] = ldexp(uI[k], -e); tD[k] = ldexp(uD[k], -e);
    }
    log2scale += e;
    if (leak && r == r0 && r < R) {
      /* Every path crosses from row r0 to row r0 + 1 once, from one cell and
       * state, and T and F continue alike from there: the cells where
       * T <= rho * F carry at most rho * S of T, the others at most their T. */
      for (int k = 0; k < width; ++k) {
        if (tM[k] > leak_rho * fM[k]) E += tM[k];
        if (tI[k] > leak_rho * fI[k]) E += tI[k];
        if (tD[k] > leak_rho * fD[k]) E += tD[k];
      }
    }
  }
  double S = 0.0, T = 0.0;
  for (int k = 0; k < width; ++k) {
    const int j = R + o + k;
    if (j < 1 || j > H) continue;
    S += fM[k] + fI[k];
    T += tM[k] + tI[k];
  }
  T += x_tail;
  if (leak && r0 >= R) {
    E = T - x_tail; /* T carried to the last row: E is T */
  } else if (leak) {
    for (int k = 0; k < width; ++k) {
      const int j = R + o + k;
      if (j < 1 || j > H) continue;
      if (((k < edge && LACT_AT(R, o)) || (k >= width - edge && RACT_AT(R, o))) && !(drop & 8))
        E += fM[k] + fI[k];
    }
  }
  if (leak) E = (E + x_tail) * phi;
#undef LACT_AT
#undef RACT_AT
  free(buf);
  if (touch_ratio)
    *touch_ratio = S > 0.0 ? (leak ? E / S + (r0 > 0 && r0 < R ? leak_rho : 0.0) : T / S) : 1.0;
  if (lost) *lost = lost_sum;
  if (!(S > 0.0)) return -INFINITY;
  return (log2(S) + log2scale) * log10(2.0);
}

/* pairhmm.hpp:27-36 constants; the loops follow pairhmm.hpp:83-114 and
 * genotyper.hpp:222-241 (std::ranges::max over the row, the cap, the filter
 * as written, then the per-allele maximum starting from lowest()). */
void bvo_postprocess_region(int n_reads, int n_haps, const double* L, const int64_t* read_len,
                            int normalize, const int32_t* hap_allele, int n_alleles,
                            uint8_t* keep, double* Lnorm, double* A) {
  for (int r = 0; r < n_reads; ++r) {
    const double* row = L + (size_t)r * (size_t)n_haps;
    double best = row[0];
    for (int h = 1; h < n_haps; ++h)
      if (best < row[h]) best = row[h];
    const double cap = best + -4.5;
    if (keep) {
      const double e = ceil((double)read_len[r] * 0.02);
      const double thr = (2.0 < e ? 2.0 : e) * -4.0;
      keep[r] = best < thr ? 1 : 0;
    }
    for (int h = 0; h < n_haps; ++h) {
      const double v = (normalize && row[h] < cap) ? cap : row[h];
      if (Lnorm) Lnorm[(size_t)r * (size_t)n_haps + (size_t)h] = v;
    }
    if (A) {
      double* out = A + (size_t)r * (size_t)n_alleles;
      for (int a = 0; a < n_alleles; ++a) out[a] = -DBL_MAX;
      for (int h = 0; h < n_haps; ++h) {
        const double v = (normalize && row[h] < cap) ? cap : row[h];
        if (v > out[hap_allele[h]]) out[hap_allele[h]] = v;
      }
    }
  }
}
