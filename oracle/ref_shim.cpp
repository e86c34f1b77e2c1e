// PairHMM oracle -- TEST INFRASTRUCTURE ONLY.
//
// C entry points around the reference's own per-pair kernels, compiled from
// their sources where they lie (lib/intel_native/pairhmm/avx-pairhmm.h) by
// oracle/Makefile into oracle/_ref/libref_pairhmm.so.  No reference source is
// copied; this file only builds a testcase and calls
// compute_full_prob_avx{s,d} (avx-pairhmm-template.h:210, instantiated by
// avx-pairhmm.h:7-15).  The region loop restates the driver rule of
// pairhmm_avx.hpp:121-157 (that header needs spdlog, which is absent).
#include <sys/time.h>

#include <cmath>
#include <cstdint>
#include <omp.h>
#include <xmmintrin.h>

#include "intel_native/pairhmm/avx-pairhmm.h"

namespace {
struct Init {
  Init() { ConvertChar::init(); }  // pairhmm_avx.hpp:117
};
Init g_init;

testcase make_tc(int R, int H, const uint8_t* rs, const uint8_t* q,
                 const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                 const uint8_t* hap) {
  testcase tc;
  tc.rslen = R;
  tc.haplen = H;
  tc.q = reinterpret_cast<const char*>(q);
  tc.i = reinterpret_cast<const char*>(ins);
  tc.d = reinterpret_cast<const char*>(del);
  tc.c = reinterpret_cast<const char*>(gcp);
  tc.hap = reinterpret_cast<const char*>(hap);
  tc.rs = reinterpret_cast<const char*>(rs);
  return tc;
}

unsigned set_ftz(int on) {
  unsigned old = _mm_getcsr();
  if (on) _mm_setcsr(old | _MM_FLUSH_ZERO_ON);
  return old;
}
}  // namespace

extern "C" {

float ref_prob_f32(int R, int H, const uint8_t* rs, const uint8_t* q,
                   const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                   const uint8_t* hap, int ftz) {
  testcase tc = make_tc(R, H, rs, q, ins, del, gcp, hap);
  unsigned old = set_ftz(ftz);
  float s = compute_full_prob_avxs<float>(&tc);
  _mm_setcsr(old);
  return s;
}

double ref_prob_f64(int R, int H, const uint8_t* rs, const uint8_t* q,
                    const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                    const uint8_t* hap, int ftz) {
  testcase tc = make_tc(R, H, rs, q, ins, del, gcp, hap);
  unsigned old = set_ftz(ftz);
  double s = compute_full_prob_avxd<double>(&tc);
  _mm_setcsr(old);
  return s;
}

float ref_mm_f32(int iq, int dq) {
  Context<float> ctx;
  return ctx.set_mm_prob(iq, dq);
}

double ref_mm_f64(int iq, int dq) {
  Context<double> ctx;
  return ctx.set_mm_prob(iq, dq);
}

float ref_ph2pr_f32(int q) {
  Context<float> ctx;
  return ctx.ph2pr[q & 127];
}

double ref_ph2pr_f64(int q) {
  Context<double> ctx;
  return ctx.ph2pr[q & 127];
}

// Driver rule of pairhmm_avx.hpp:133-152 over one region: OpenMP
// schedule(dynamic,1) over reads, haplotypes serial, FP32 then FP64 when
// S_f < MIN_ACCEPTED, log10 minus log10(K).  out is [n_reads][n_haps].
long ref_region_log10(int n_reads, const int64_t* read_off, const uint8_t* rs,
                      const uint8_t* q, const uint8_t* ins, const uint8_t* del,
                      const uint8_t* gcp, int n_haps, const int64_t* hap_off,
                      const uint8_t* hap, int n_threads, double* out) {
  Context<float> ctxf;
  Context<double> ctxd;
  long rescued = 0;
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads) reduction(+ : rescued)
  for (int i = 0; i < n_reads; ++i) {
    _mm_setcsr(_mm_getcsr() | _MM_FLUSH_ZERO_ON);
    const int64_t ro = read_off[i];
    const int R = static_cast<int>(read_off[i + 1] - ro);
    for (int h = 0; h < n_haps; ++h) {
      const int64_t ho = hap_off[h];
      const int H = static_cast<int>(hap_off[h + 1] - ho);
      testcase tc = make_tc(R, H, rs + ro, q + ro, ins + ro, del + ro, gcp + ro, hap + ho);
      const float sf = compute_full_prob_avxs<float>(&tc);
      double L;
      if (sf < MIN_ACCEPTED) {
        const double sd = compute_full_prob_avxd<double>(&tc);
        L = std::log10(sd) - ctxd.LOG10_INITIAL_CONSTANT;
        ++rescued;
      } else {
        L = ::log10f(sf) - ctxf.LOG10_INITIAL_CONSTANT;
      }
      out[static_cast<size_t>(i) * n_haps + h] = L;
    }
  }
  return rescued;
}

}  // extern "C"
