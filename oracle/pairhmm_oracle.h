/*
 * PairHMM oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU PairHMM, used as the parity
 * checker for the HIP path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product path
 * (biovoltron_amd/) links or calls it.
 *
 * Restated reference (joyeli/Biovoltron):
 *   tables         lib/intel_native/pairhmm/Context.h:42-61,63-90,100-175
 *   base codes     lib/intel_native/pairhmm/pairhmm_common.h:16,26-44
 *   DP + striping  lib/intel_native/pairhmm/avx-pairhmm-template.h:83-346
 *   carry shifts   lib/intel_native/pairhmm/avx-vector-shift.h:3-56
 *   driver rule    include/biovoltron/algo/align/inexact_match/pairhmm_avx.hpp:121-157
 *   scalar mode    include/biovoltron/algo/align/inexact_match/pairhmm.hpp:39-80
 *   post-process   include/biovoltron/algo/align/inexact_match/pairhmm.hpp:83-114
 *
 * How each function is pinned (what checks it besides its own formula):
 *   bvo_prob_f32 / bvo_prob_f64   bit-exact vs the reference kernels compiled
 *                                 from their own sources (oracle/_ref), KATs,
 *                                 and the closed forms for R=1 and R=H=1.
 *   bvo_log10_exact               brute-force path enumeration (R,H<=4),
 *                                 invariance to the scale K, 1e-9 agreement
 *                                 with bvo_prob_f64 where that is finite.
 *   bvo_scalar_log10              brute-force enumeration with the constant
 *                                 transition matrix.
 *   bvo_band_posterior            per-row posterior mass sums to 1 (pins the
 *                                 backward pass), band covering everything.
 *   bvo_band_forward              band covering the matrix equals the full
 *                                 DP with zero touch; brute-force enumeration
 *                                 of in-band paths and of those touching an
 *                                 edge zone (R, H <= 4).
 *   bvo_band_follow_forward       no moves: equals bvo_band_forward; with
 *                                 moves: brute-force enumeration of the paths
 *                                 that stay in each row's band (from the
 *                                 returned offsets), touch mass included;
 *                                 dropped share against a full-matrix DP with
 *                                 the same row masks; a planted deletion and an
 *                                 off-centre start are followed.
 */
#ifndef BV_PAIRHMM_ORACLE_H
#define BV_PAIRHMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Build the tables (idempotent). */
void bvo_init(void);

/* Table access for cross-checking the product's own tables. */
float bvo_ph2pr_f32(int q);
double bvo_ph2pr_f64(int q);
float bvo_mm_f32(int iq, int dq);
double bvo_mm_f64(int iq, int dq);

/* Reference stripe semantics.  Returns the scaled sum S (not a log), exactly
 * as compute_full_prob_avx{s,d} does: FP32 with K=2^120 and 8-row stripes,
 * FP64 with K=2^1020 and 4-row stripes.  ftz!=0 sets the x86 flush-to-zero
 * mode during the call, as the reference driver does on its calling thread. */
float bvo_prob_f32(int R, int H, const uint8_t* rs, const uint8_t* q,
                   const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                   const uint8_t* hap, int ftz);
double bvo_prob_f64(int R, int H, const uint8_t* rs, const uint8_t* q,
                    const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                    const uint8_t* hap, int ftz);

/* Driver rule: FP32 first, FP64 re-run when S_f < 1e-28f, then log10 minus
 * log10(K).  *rescued (optional) receives 1 when the FP64 run was used. */
double bvo_log10_reference(int R, int H, const uint8_t* rs, const uint8_t* q,
                           const uint8_t* ins, const uint8_t* del,
                           const uint8_t* gcp, const uint8_t* hap, int ftz,
                           int* rescued);

/* FP64 textbook forward DP with per-row renormalisation (never underflows).
 * mode 0: per-base transitions as in the AVX path (matchToMatch table,
 *         ph2pr); mode 1: the scalar PairHMM constant transition matrix with
 *         priors from 10^(-q/10) and wildcard 'N' only.  Returns log10 L. */
double bvo_log10_exact(int R, int H, const uint8_t* rs, const uint8_t* q,
                       const uint8_t* ins, const uint8_t* del,
                       const uint8_t* gcp, const uint8_t* hap, int mode);

/* Scalar PairHMM (pairhmm.hpp:55-80): FP64, constant transitions
 * trans[7] = {MM, MI, MD, IM, II, DM, DD}, K = 2^1020, no rescue. */
double bvo_scalar_log10(int R, int H, const uint8_t* rs, const uint8_t* q,
                        const uint8_t* hap, const double* trans);

/* Forward-backward posterior per diagonal offset o = j - r, for
 * o in [-(R-1), H-1] (length R+H-1), written to mass[o + R - 1]: the expected
 * number of cells (M, I or D) the alignment path visits on that diagonal, so
 * the values sum to R plus the expected number of deleted columns.  row_sum
 * (optional, length R) receives the per-row total posterior over M+I (must
 * be 1 for every row).  Returns log10 L. */
double bvo_band_posterior(int R, int H, const uint8_t* rs, const uint8_t* q,
                          const uint8_t* ins, const uint8_t* del,
                          const uint8_t* gcp, const uint8_t* hap,
                          double* mass, double* row_sum);

/* Region driver as the reference runs it (pairhmm_avx.hpp:121-157): OpenMP
 * schedule(dynamic,1) over reads, haplotypes serial inside, FP32 then FP64.
 * Reads and haplotypes are packed byte arrays with offsets (length n+1).
 * out is read-major [n_reads][n_haps].  Returns the number of rescued pairs. */
long bvo_region_log10(int n_reads, const int64_t* read_off, const uint8_t* rs,
                      const uint8_t* q, const uint8_t* ins, const uint8_t* del,
                      const uint8_t* gcp, int n_haps, const int64_t* hap_off,
                      const uint8_t* hap, int n_threads, double* out);

/* Banded FP64 forward restricted to diagonals [o_lo, o_lo + width - 1] with
 * the edge-zone touch mass of the banded kernel (pairhmm_band.hip); returns
 * log10 of the banded likelihood and the touch ratio T/S. */
double bvo_band_forward(int R, int H, const uint8_t* rs, const uint8_t* q,
                        const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                        const uint8_t* hap, int o_lo, int width, int edge,
                        double* touch_ratio);

/* Banded FP64 forward whose band moves (the long-read kernel's band, and the
 * wide band with its end slides, both in pairhmm_band.hip): `width` diagonals
 * starting at o_init (or sched[0]); before a row the band moves towards the
 * centroid of the previous row's mass every `every` rows (by at most
 * max_move, aiming end_bias slots above the centre on the last end_rows
 * rows), or, when sched is not NULL, towards the offsets sched[0..R] (by at
 * most max_move per row).  A move is not made when the slots it would drop
 * hold more than hold_eps of the row's mass (hold_eps > 0).  irun adds to the
 * touch mass the pure insertion runs that enter through the top edge from a
 * read start above the band (bit 0, rows 2..48) and those that leave through
 * the bottom edge to end the read below it (bit 1, the last 48 rows; the
 * kernels count these).  edge < 0 gives zones of -edge diagonals active on
 * every row (not only where the band cuts the matrix).  irun bit 2 (with
 * edge < 0): the anti-diagonal banded kernel's rule -- T = F on the zones is
 * kept by taking the forward value only for the innermost zone cells' inputs
 * from the band's middle, and the entry runs (bit 0 implied) then propagate
 * through the zone into T.  irun bit 3: the touch ratio is replaced by
 * the leak bound E / S, an upper bound on T / S computed from forward values
 * only (the banded kernels' accept test; see the definition); bits 4-7 then
 * drop one of its terms each (test use: planted mistakes).  Returns log10 of
 * the banded likelihood, the touch ratio T/S (the accept test), the mass share
 * the moves dropped (*lost) and, if offs is not NULL, the band offset used on
 * each row 0..R. */
/* The split level rho of the leak bound with r0 > 0 (default 2.5e-7). */
void bvo_set_leak_rho(double rho);
double bvo_band_follow_forward(int R, int H, const uint8_t* rs, const uint8_t* q,
                               const uint8_t* ins, const uint8_t* del, const uint8_t* gcp,
                               const uint8_t* hap, int o_init, int width, int edge, int every,
                               int max_move, int end_rows, int end_bias, double hold_eps,
                               int irun, const int* sched, double* touch_ratio, double* lost,
                               int* offs);

/* Post-processing of one region's read x haplotype matrix L (n_reads rows of
 * n_haps, read-major), in the order the reference applies it:
 *   normalize_likelihoods (pairhmm.hpp:83-93), when normalize != 0: every
 *     entry below the row's best - 4.5 is raised to it (written to Lnorm);
 *   filter_poorly_modeled_reads (pairhmm.hpp:95-114): keep[r] = 1 when the
 *     row's best is below min(2, ceil(read_len[r] * 0.02)) * -4;
 *   marginal_likelihoods (genotyper.hpp:222-241): A[r * n_alleles + a] = the
 *     largest (normalised) entry of row r over the haplotypes h with
 *     hap_allele[h] == a, std::numeric_limits<double>::lowest() if none.
 * Any output may be NULL. */
void bvo_postprocess_region(int n_reads, int n_haps, const double* L, const int64_t* read_len,
                            int normalize, const int32_t* hap_allele, int n_alleles,
                            uint8_t* keep, double* Lnorm, double* A);

/* Smith-Waterman of SmithWaterman::align (smithwaterman.hpp:54-258; see
 * sw_oracle.c): reference bases are rows, alternate bases columns;
 * params = {w_match, w_mismatch, w_open, w_extend}.  Writes the CIGAR
 * elements in order to ops/lens (capacity R + H + 3), the alignment offset
 * to *offset, and returns the element count. */
int bvo_sw_align(int R, const uint8_t* ref, int H, const uint8_t* alt, const int32_t* params,
                 int32_t* offset, char* ops, int32_t* lens);

#ifdef __cplusplus
}
#endif
#endif
