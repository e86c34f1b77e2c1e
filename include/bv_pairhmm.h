/*
 * bv_pairhmm.h -- C ABI of the MI355X PairHMM read-likelihood stage.
 *
 * Drop-in boundary for joyeli/Biovoltron's PairHMM path.  Plain C types only
 * (no HIP or torch types); every function returns 0 on success or a negative
 * status (BV_E*), with a per-thread message from bv_pairhmm_last_error().
 * Buffers passed in are host memory unless a name says otherwise; the library
 * copies its inputs and keeps no caller pointer after a call returns.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   bv_compute_full_prob        compute_full_prob_avxs<float> / _avxd<double>
 *                               (lib/intel_native/pairhmm/avx-pairhmm-template.h:210,
 *                               instantiated in avx-pairhmm.h:7-15, called through
 *                               the pointers of pairhmm_avx.hpp:33-34,113-114)
 *   bv_pairhmm_batch /          AvxPairHMM::computeLikelihoodsNative
 *   bv_plan_*                   (include/biovoltron/algo/align/inexact_match/
 *                               pairhmm_avx.hpp:121-157): FP32 first, FP64 re-run
 *                               when S_f < 1e-28f, log10 minus log10(K); for many
 *                               regions in one launch sequence
 *   BV_MODE_SCALAR              PairHMM::compute_likelihoods with a TrasMatrx
 *                               (pairhmm.hpp:116-141, recurrence :55-80)
 *   bv_pairhmm_init             AvxPairHMM::initNative (pairhmm_avx.hpp:83-119)
 *   (C++ linkage, in the same   shacc_pairhmm::calculate
 *    library)                   (lib/intel_native/pairhmm/shacc_pairhmm.h:10-36)
 *
 * Normalisation and filtering (pairhmm.hpp:83-114) are host-side and live in
 * the C++ adapter include/biovoltron_amd/gpu_pairhmm.hpp.
 */
#ifndef BV_PAIRHMM_H
#define BV_PAIRHMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BV_OK 0
#define BV_EINVAL (-22)
#define BV_ENOMEM (-12)
#define BV_ENODEV (-19)
#define BV_EHIP (-5)
#define BV_ESTATE (-1)

/* Transition/emission model. */
#define BV_MODE_AVX 0    /* per-base GOP/GCP transitions, FP32 + FP64 rescue */
#define BV_MODE_SCALAR 1 /* constant transition matrix, FP64 (scalar PairHMM) */

/* Per-pair result forms (bv_plan_results form[] values). */
#define BV_FORM_FP32 0
#define BV_FORM_FP64 1
#define BV_FORM_BAND_FP32 2
#define BV_FORM_BAND_FALLBACK 3
#define BV_FORM_BAND_FP64 4 /* moving-band FP64 forward: a long read (>= 1024 bases) with
                               bv_plan_set_long_precision(plan, 64), or a short-read pair whose
                               FP32 band sum fell below 1e-28 (the reference's FP64 re-run) */
#define BV_FORM_BAND_LONG32 5 /* long read: moving-band FP32 forward, rescaled by a power of
                                 two every 64 rows (bv_plan_set_long_precision, default) */

/* Initialise the device context and upload the tables (idempotent).  device
 * is a HIP ordinal; -1 keeps the current device. */
int bv_pairhmm_init(int device);
void bv_pairhmm_shutdown(void);
const char* bv_pairhmm_last_error(void);

typedef struct bv_device_info {
  int32_t device;
  int32_t compute_units;
  int32_t blocks_per_cu_f32;
  int32_t blocks_per_cu_f64;
  char arch[64];
} bv_device_info;
int bv_pairhmm_device_info(bv_device_info* info);

/* Host copies of the lookup tables the kernels use (ph2pr[128] per
 * precision; match-to-match for insertion/deletion quals < 128, 8256 entries
 * indexed max*(max+1)/2 + min), for checking against the reference's
 * Context<T> tables (lib/intel_native/pairhmm/Context.h:42-61,105-109,145-149).
 * Any pointer may be NULL. */
int bv_pairhmm_tables(float* ph2pr_f, double* ph2pr_d, float* mm_f, double* mm_d);

/* The reference's per-pair descriptor (pairhmm_common.h:20-24). */
typedef struct bv_testcase {
  int rslen, haplen;
  const char *q, *i, *d, *c;
  const char *hap, *rs;
} bv_testcase;

/* Per-pair kernel ABI: scaled sum S for each testcase, FP32 (precision 32,
 * K = 2^120) or FP64 (precision 64, K = 2^1020), exactly the quantity the
 * reference kernels return (not a log). */
int bv_compute_full_prob(const bv_testcase* tcs, int n, int precision, double* out_scaled);

/* Allocation-free form of bv_compute_full_prob.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Pair k is read bytes [read_off[k], read_off[k+1]) of the five read arrays
 * against haplotype bytes [hap_off[k], hap_off[k+1]); every pointer,
 * offsets included, is DEVICE memory, and every haplotype must be at most
 * max_hap_len bases (a pair with a longer or empty haplotype, or an empty
 * read, gets NaN).  Writes the scaled sum S per pair to out_scaled (device,
 * n_pairs doubles), as bv_compute_full_prob does, enqueued on `stream` (a
 * hipStream_t; NULL is the null stream) and asynchronous: synchronise the
 * stream before reading.  Needs bv_pairhmm_init first (BV_ESTATE otherwise).
 * Calls no device allocator: its scratch (queue counter and carry rows) is a
 * static device array of BV_NOALLOC_SCRATCH_BYTES per device, held until the
 * process ends; max_hap_len is rejected (BV_EINVAL) when one carry row does
 * not fit it.  Every call on a device shares that array, so calls that
 * overlap on one device race: serialise them (one stream, or events). */
#define BV_NOALLOC_SCRATCH_BYTES (32u << 20)
typedef struct bv_device_pairs {
  int64_t n_pairs;
  const int64_t* read_off; /* n_pairs + 1 */
  const uint8_t* read_bases;
  const uint8_t* read_qual;
  const uint8_t* read_ins;
  const uint8_t* read_del;
  const uint8_t* read_gcp;
  const int64_t* hap_off; /* n_pairs + 1 */
  const uint8_t* hap_bases;
} bv_device_pairs;
int bv_compute_full_prob_noalloc(const bv_device_pairs* pairs, int precision, int max_hap_len,
                                 double* out_scaled, void* stream);

/* Regions of reads x haplotypes in packed structure-of-arrays form.  Region g
 * owns reads [region_read_off[g], region_read_off[g+1]) and haplotypes
 * [region_hap_off[g], region_hap_off[g+1]); read r's bytes are
 * [read_off[r], read_off[r+1]) of the five read arrays (quality bytes are raw
 * Phred values as the reference's testcase carries them); haplotype h's bytes
 * are [hap_off[h], hap_off[h+1]) of hap_bases.  Results are numbered region
 * by region, read-major: pair = sum of earlier regions' reads*haps
 * + r_local * n_haps(g) + h_local. */
typedef struct bv_region_batch {
  int64_t n_regions;
  const int64_t* region_read_off;
  const int64_t* region_hap_off;
  int64_t n_reads;
  const int64_t* read_off;
  const uint8_t* read_bases;
  const uint8_t* read_qual;
  const uint8_t* read_ins;
  const uint8_t* read_del;
  const uint8_t* read_gcp;
  int64_t n_haps;
  const int64_t* hap_off;
  const uint8_t* hap_bases;
} bv_region_batch;

typedef struct bv_options {
  int32_t mode;        /* BV_MODE_* */
  int32_t fp64_renorm; /* FP64: rescale per stripe so it never underflows (1),
                          or reproduce the reference's 2^1020-only scaling (0) */
  int32_t force_fp64;  /* skip the FP32 pass */
  int32_t grid_f32;    /* persistent grid override, 0 = occupancy-derived */
  int32_t grid_f64;
  float band_tol;      /* banded pairs whose edge-zone touch mass exceeds this
                          fraction of the banded likelihood are re-run on the
                          full matrix (default 5e-7: at 1e-6, one short-read
                          pair in 4.04e8 with a touch ratio of 5.8e-7 lost
                          2.4e-3 of its likelihood, |dlog10 L| = 1.05e-3) */
  int32_t band_narrow; /* learned bands may use the 32-diagonal band (two
                          pairs per wavefront; default 1).  A read that
                          aligns partly to an inexact repeat copy 17-40
                          diagonals away can pass every seed test, so the
                          narrow band needs the network's width gate: use
                          margin >= 10 in bv_plan_set_band_model */
  int32_t use_graph;   /* bv_plan_run replays the run as one captured hipGraph
                          (default 1); the graph is rebuilt after the bands,
                          the band model or its parameters change */
  double trans[7];     /* BV_MODE_SCALAR: {MM, MI, MD, IM, II, DM, DD}
                          (pairhmm.hpp:15-33 order); all zero = ORIGINAL_DEFAULT */
} bv_options;

void bv_options_default(bv_options* opt);

typedef struct bv_plan bv_plan;

typedef struct bv_plan_stats {
  int64_t n_pairs;
  int64_t cells;        /* sum over pairs of R*H */
  int64_t fp64_pairs;   /* pairs computed by the FP64 pass in the last run */
  int64_t fp64_cells;   /* cells computed in FP64 (full-matrix and banded) */
  int64_t tasks;
  int32_t grid_f32, grid_f64;
  float last_ms_f32, last_ms_f64; /* kernel times of the last timed run */
  int64_t band_pairs;     /* banded pairs accepted in the last run */
  int64_t fallback_pairs; /* banded pairs rejected and re-run on the full matrix */
  int64_t cells_swept;    /* cells actually computed in the last run, all passes */
  float last_ms_fallback;
  float last_ms_band;     /* banded kernels alone (both widths), last timed run */
  int64_t long_band_pairs; /* long reads accepted on the moving band (also in band_pairs) */
  float last_ms_long;      /* moving-band kernel (long reads), last timed run */
  float last_ms_post;      /* post-processing kernel, last bv_plan_postprocess */
  float setup_ms_host;     /* bv_plan_create: host work (classes, work lists) */
  float setup_ms_upload;   /* bv_plan_create: uploads and their expansion on the device */
  int64_t setup_bytes;     /* bytes bv_plan_create uploaded */
} bv_plan_stats;

/* Uploads a batch to device memory and builds the length-ordered work list. */
int bv_plan_create(const bv_region_batch* batch, const bv_options* opt, bv_plan** plan);
/* Enqueues one full evaluation (FP32 pass, FP64 re-run of underflowed pairs)
 * on the plan's stream; asynchronous. */
int bv_plan_run(bv_plan* plan);
/* Runs the plan iters times back to back and returns the mean wall time per
 * run measured with HIP events on the plan's stream (inputs resident). */
int bv_plan_time(bv_plan* plan, int iters, float* ms_per_run);
/* Waits for the plan's stream and copies results to host buffers sized
 * n_pairs (either pointer may be NULL). */
int bv_plan_results(bv_plan* plan, double* log10_out, uint8_t* form_out);
int bv_plan_stats_get(bv_plan* plan, bv_plan_stats* stats);
/* Per-pair diagonal bands for the FP32 pass: bands[2k], bands[2k+1] =
 * (o_lo, o_hi) of j - r for pair k (host memory, n_pairs x 2 int32); a pair
 * with o_lo > o_hi is evaluated on the full matrix.  NULL removes all bands.
 * For a read of 1024 bases or more (run in FP64) the band is where a
 * 64-diagonal band starts: it then follows the alignment, moving at most 4
 * diagonals every 16 rows (BV_FORM_BAND_FP64); wider bands take the full
 * matrix.  No reference counterpart: the reference computes every cell. */
int bv_plan_set_bands(bv_plan* plan, const int32_t* bands);
void bv_plan_destroy(bv_plan* plan);

/* Post-processing of the last run's results on the device, so that only
 * per-read decisions and per-allele likelihoods need to leave it.  Replaces,
 * per region:
 *   PairHMM::normalize_likelihoods (pairhmm.hpp:83-93): with
 *     BV_POST_NORMALIZE every entry below (row best - 4.5) is raised to it;
 *   PairHMM::filter_poorly_modeled_reads (pairhmm.hpp:95-114): keep_out[r] =
 *     1 when the row best is below min(2, ceil(0.02 * read length)) * -4,
 *     the reference's test as written (it keeps the poorly modelled reads);
 *     reads outside every region get 0.  The caller drops rows with 0;
 *   Genotyper::marginal_likelihoods (genotyper.hpp:222-241, map of :200-209):
 *     allele_out[allele_off(g) + r_local * n_alleles[g] + a] = max over the
 *     region's haplotypes h with hap_allele[h] == a of the (normalised)
 *     entry, or -DBL_MAX (std::numeric_limits<double>::lowest()) when none
 *     maps to a; allele_off(g) = sum over earlier regions of reads *
 *     n_alleles.  Rows are computed for every read; the reference computes
 *     them for the reads the genotyper keeps, a subset of these rows.
 * hap_allele (n_haps, indexed by global haplotype) and n_alleles (n_regions)
 * are given together or both NULL; hap_allele[h] must lie in
 * [0, n_alleles[g]) for h in region g (the reference maps an unmapped
 * haplotype to allele 0).  keep_out (n_reads), normalized_out (n_pairs) and
 * allele_out may each be NULL.  Synchronous; call after bv_plan_run.  No
 * reference counterpart as one call: the reference runs these loops on the
 * host over the whole matrix. */
/* Smith-Waterman alignment of alternate sequences (haplotypes) to reference
 * sequences, the integer DP upstream of PairHMM in the haplotype caller.
 * Replaces SmithWaterman::align(ref, alt, params)
 * (include/biovoltron/algo/align/inexact_match/smithwaterman.hpp:229-255),
 * called from haplotype_graph.hpp:219, for many pairs at once; same results
 * (offset and CIGAR elements, including the reference's zero-length 'M'
 * element when the walk starts with a gap).  Pair k aligns alt bytes
 * [alt_off[k], alt_off[k+1]) to ref bytes [ref_off[k], ref_off[k+1]);
 * params = {w_match, w_mismatch, w_open, w_extend} (|w| <= 65536; the
 * reference's default is NEW_SW_PARAMETERS = {200, -150, -260, -11}).
 * Outputs: offset_out[k], n_elems_out[k], and the CIGAR elements of pair k
 * in ops_out / lens_out from slot sum_{m<k} (R_m + H_m + 3) on.  Sequences
 * of 1..4096 bases (the reference asserts non-empty ones).  kernel_ms
 * (optional) receives the kernel time.  Synchronous.  The per-wavefront
 * scratch (sized by the call's longest pairs, at most 16 GiB) stays
 * allocated for the next call, grown when a call needs more, and is freed
 * by bv_sw_release or bv_pairhmm_shutdown; concurrent calls take turns on
 * it. */
typedef struct bv_sw_batch {
  int64_t n_pairs;
  const int64_t* ref_off;
  const uint8_t* ref;
  const int64_t* alt_off;
  const uint8_t* alt;
} bv_sw_batch;
int bv_sw_align(const bv_sw_batch* batch, const int32_t* params, int32_t* offset_out, int32_t* n_elems_out,
                char* ops_out, int32_t* lens_out, float* kernel_ms);
/* Frees bv_sw_align's kept scratch (the next call allocates it again).  No
 * reference counterpart: for callers that align once and then need the
 * device memory back, e.g. for a large resident PairHMM plan. */
void bv_sw_release(void);

#define BV_POST_NORMALIZE 1
int bv_plan_postprocess(bv_plan* plan, int32_t flags, const int32_t* hap_allele,
                        const int32_t* n_alleles, uint8_t* keep_out, double* normalized_out,
                        double* allele_out);

/* ---- learned bands ----------------------------------------------------------
 * None of these has a reference counterpart: the reference computes every
 * cell.  A band model is a 36 -> 64 -> 64 -> 2 ReLU MLP (FP32) over per-pair
 * features (read/haplotype lengths, mean/min base quality, a 32-bin 12-mer
 * seed-hit offset histogram around the median hit offset, the "anchor");
 * its outputs are (band centre - anchor) / 16 and log2(band width). */
#define BV_BAND_FEATURES 36
#define BV_BAND_PARAMS 6658
typedef struct bv_band_model bv_band_model;
int bv_band_model_create(bv_band_model** model, uint64_t seed);
void bv_band_model_destroy(bv_band_model* model);
int bv_band_model_get_params(const bv_band_model* model, float* params /* BV_BAND_PARAMS */);
int bv_band_model_set_params(bv_band_model* model, const float* params);
/* X [n][36] -> Y [n][2] (host buffers), device forward kernel. */
int bv_band_model_predict(bv_band_model* model, const float* X, int64_t n, float* Y);
/* Adam state (first and second moments, BV_BAND_PARAMS each, and the step
 * count), so that training can resume where it stopped.  Any pointer of
 * get_state may be NULL. */
int bv_band_model_get_state(const bv_band_model* model, float* m1, float* m2, int64_t* step);
int bv_band_model_set_state(bv_band_model* model, const float* m1, const float* m2, int64_t step);
/* Band model files: "BVBANDM\0", uint32 version 1, uint32 BV_BAND_PARAMS,
 * the parameters as little-endian float32, then the SHA-256 of those
 * parameter bytes (the digest band.py records).  read checks the file and
 * fills params (host memory, no device use); load makes a device model from
 * it; write saves a model.  A NULL path means the trained model shipped with
 * the library, models/band_mlp.bin in the library's own directory
 * (bv_band_model_default_path).  Wrong magic, version, size or digest is
 * BV_EINVAL. */
int bv_band_model_read(const char* path, float* params /* BV_BAND_PARAMS */);
int bv_band_model_load(const char* path, bv_band_model** model);
int bv_band_model_write(const bv_band_model* model, const char* path);
int bv_band_model_default_path(char* buf, int64_t size);
/* Loss 0.5*mean(w0*(Y0-T0)^2 + w1*(Y1-T1)^2) and its gradient (device kernels). */
int bv_band_model_loss_grad(bv_band_model* model, const float* X, const float* T, int64_t n,
                            float w0, float w1, float* loss, float* grad /* BV_BAND_PARAMS */);
/* Adam on the device over `epochs` passes of minibatches; order (optional,
 * epochs x n sample indices) fixes the sample order of every epoch. */
int bv_band_model_train(bv_band_model* model, const float* X, const float* T, int64_t n,
                        int epochs, int batch, float lr, float w0, float w1,
                        const int64_t* order, float* epoch_loss);
/* Per-pair features (device kernel) copied to host: X [n_pairs][36], the
 * seed anchor per pair, and (optional) ends_out [n_pairs][2]: the first read
 * position and last read end covered by seeds hitting within 8 of the anchor
 * (-1, -1 without such hits).  Any output may be NULL. */
int bv_plan_features(bv_plan* plan, float* X_out, int32_t* anchor_out, int32_t* ends_out);
/* Every later run predicts bands on the device (features, MLP, band table)
 * before the FP32 pass.  With the predicted span widened by `margin` per
 * side: 32 diagonals around the anchor when it fits the narrow band and the
 * seeds anchor both read ends inside it, else 64 around the predicted centre
 * when it is at most max_width wide, else the full matrix (also for pairs
 * without seed hits or with haplotypes longer than 1024).  NULL model turns
 * learned bands off. */
int bv_plan_set_band_model(bv_plan* plan, const bv_band_model* model, int margin, int max_width);
/* Autotuner: picks the persistent grid of the banded kernels, the FP32 pass,
 * the FP64 pass and the long-read moving band (a coordinate search over
 * fractions of each kernel's occupancy-derived grid, timing whole runs of
 * this plan, `iters` each) and keeps the fastest for later runs.  grids_out
 * (optional) receives {banded, fp32, fp64, long}.  Tune after setting bands
 * or a band model. */
int bv_plan_autotune(bv_plan* plan, int iters, int32_t* grids_out);
/* Sets the persistent grids {banded, fp32, fp64, long} (0 keeps a grid; each
 * is clamped to the plan's occupancy-derived maximum), e.g. to reuse grids
 * tuned on a smaller plan of the same workload. */
int bv_plan_set_grids(bv_plan* plan, const int32_t* grids);
/* Shape of the 64-diagonal banded kernel: pairs per wavefront.  4 (the
 * default, also for 0): four pairs of one read share a wavefront, 16 lanes
 * and four diagonals per lane each; 1: one pair, one diagonal per lane.
 * The two compute the same recurrence and edge-zone test; results agree to
 * FP32 rounding.  Resets the banded grid to the shape's occupancy-derived
 * maximum; bv_plan_autotune also searches it.  No reference counterpart. */
int bv_plan_set_band_shape(bv_plan* plan, int32_t wide_pairs);
int bv_plan_band_shape(bv_plan* plan, int32_t* wide_pairs);
/* The quad shape's mid band (default off): with it on, bands of 33-48
 * diagonals from bv_plan_set_bands run as 48-diagonal bands, four pairs of
 * one read per wavefront with three diagonals per lane and edge zones of
 * three diagonals (about 25 % less banded kernel time than 64); with it off,
 * or with the one-pair shape, such bands run as 64-diagonal bands with the
 * same centre.  Learned bands never take 48 diagonals: on the combined
 * workload 48-diagonal learned bands missed 20 pairs of 4.04e8 by more than
 * 1e-3 (worst 2.98 log10 units), alignments 20-30 diagonals from the
 * predicted centre that a 48-diagonal band's edge zones do not see.  No
 * reference counterpart. */
int bv_plan_set_band_mid(bv_plan* plan, int32_t on);
/* Arithmetic of the long reads' moving band: 32 (default) FP32 with the
 * state rescaled by a power of two every 16 rows (exact, so it never
 * underflows; log10 L adds the removed exponent), form BV_FORM_BAND_LONG32;
 * or 64, FP64 as before, form BV_FORM_BAND_FP64.  Either way a pair whose
 * touch mass exceeds the tolerance, or whose likelihood FP32 would reach
 * unscaled (the reference's FP32 rule), takes the full-matrix FP64 pass.
 * No reference counterpart: the reference runs long reads in FP32 and,
 * after underflow, in FP64 on the full matrix. */
int bv_plan_set_long_precision(bv_plan* plan, int32_t bits);
int bv_plan_long_precision(bv_plan* plan, int32_t* bits);
int bv_plan_band_mid(bv_plan* plan, int32_t* on);
/* The quad shape's leftovers: a read's wide pairs run four to a wavefront,
 * and the last one to three of them either fill one more wavefront, its
 * empty slots idle, when there are at least min_pairs of them, or run one
 * pair per wavefront.  min_pairs 1-4; 4: never fill; default 3 (three
 * pairs in a quad wavefront cost about 0.7 of three one-pair wavefronts).
 * Results are the same banded likelihoods either way, to FP32 rounding.
 * No reference counterpart. */
int bv_plan_set_quad_fill(bv_plan* plan, int32_t min_pairs);
int bv_plan_quad_fill(bv_plan* plan, int32_t* min_pairs);
/* How a run with a band model builds its bands and work lists.  2 (the
 * default): the features kernel writes a compact record per pair (bin
 * counts as bytes, anchor, seed-end span), then one kernel runs the MLP,
 * the band rules, the band table and the banded and full-matrix lists from
 * them; 1: one kernel does all of it, the features kept in LDS; 0: the
 * separate features, MLP, band-rule and dispatch kernels.  All give the same
 * band table and the same pairs in each list (list order may differ).  No
 * reference counterpart. */
int bv_plan_set_band_prologue(bv_plan* plan, int32_t mode);
/* The band table of the last run (n_pairs x 2 int32). */
int bv_plan_bands(bv_plan* plan, int32_t* bands_out);
/* Training labels (FP64 forward-backward on the host), lo_hi [n_pairs][2]:
 * the narrowest diagonal span outside which the alignment path makes at most
 * 1 - mass expected visits, a bound on the probability that it leaves. */
int bv_band_labels(const bv_region_batch* batch, double mass, int n_threads, int32_t* lo_hi);

/* One-shot: create, run, fetch, destroy. */
int bv_pairhmm_batch(const bv_region_batch* batch, const bv_options* opt,
                     double* log10_out, uint8_t* form_out);

#ifdef __cplusplus
}
#endif
#endif /* BV_PAIRHMM_H */
