// Host side of the MI355X PairHMM: tables, device context, plans, C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "band_common.h"
#include "bv_pairhmm.h"
#include "pairhmm_band.h"
#include "pairhmm_common.h"
#include "sw_common.h"

namespace bvp {
hipError_t launch_band_dispatch(const BArgs& a, hipStream_t st);
hipError_t launch_band_quads(const BArgs& a, int grid, hipStream_t st);
hipError_t launch_band_one(const BArgs& a, int grid, int grid1, hipStream_t st);
int band_blocks_per_cu(int sel_cols, int wide_pairs);
hipError_t launch_band_long(const BArgs& a, int grid, hipStream_t st);
hipError_t launch_band_rescue(const BArgs& b, Task* rescue, int* rescue_count, int* queue, int* accepted, int grid,
                              hipStream_t st);
int band_long_blocks_per_cu(int fp32);
hipError_t launch_long_anchors(const uint8_t* rclass, const int64_t* read_off, const int64_t* htab_off,
                               const unsigned long long* htab, const Task* tasks, int n_tasks, int2* band,
                               int grid, hipStream_t st);
hipError_t launch_band_features(const FeatArgs& a, int grid, hipStream_t st);
template <typename T>
hipError_t sort_longest_first(const T* d_in, const uint64_t* d_cost, int64_t n, int end_bit, T* d_out,
                              hipStream_t st);
hipError_t launch_expand_bases(const int64_t* read_off, int64_t n_reads, const uint8_t* rpack, const int64_t* rexc,
                               int64_t n_rexc, uint8_t* rclass, int64_t hbytes, const uint8_t* hpack,
                               const int64_t* hexc, int64_t n_hexc, uint8_t* hmask, hipStream_t st);
int band_prologue_grid(int cus, int n_tasks);
hipError_t launch_band_prologue(const FeatArgs& f, const BArgs& b, int2* band, const float* params, int margin,
                                int max_width, int allow_narrow, int grid, hipStream_t st);
hipError_t launch_mlp_forward(const float* params, const float* X, int64_t n, float* Y, float* H1,
                              float* H2, hipStream_t st);
hipError_t launch_mlp_backward(const float* params, const float* X, const float* Y, const float* T,
                               int64_t n, const float* H1, const float* H2, float* D1, float* D2,
                               float* D3, float w0, float w1, float* loss, float* grad,
                               hipStream_t st);
hipError_t launch_adam(float* params, const float* grad, float* m1, float* m2, float lr, int t,
                       hipStream_t st);
hipError_t launch_band_from_prediction(const float* Y, const int* anchor, const int2* ends,
                                       const float* X, int64_t n, int margin, int max_width,
                                       int allow_narrow, int2* band, hipStream_t st);
hipError_t launch_hap_tables(const uint8_t* hmask, const int64_t* hap_off, const int64_t* tab_off,
                             int64_t n_haps, unsigned long long* tab, hipStream_t st);
template <typename T>
hipError_t launch_forward(const KArgs& a, int grid, hipStream_t st);
template <typename T>
int max_blocks_per_cu();
hipError_t launch_zero_ints(int* p, int n, hipStream_t st);
hipError_t noalloc_scratch_address(void** ptr);
hipError_t launch_stack(const KArgs& a, int grid, hipStream_t st);
int read_model_file(const char* path, float* params, int count, std::string& err);
int write_model_file(const char* path, const float* params, int count, std::string& err);
std::string default_model_path();
int stack_blocks_per_cu(int carry_cols);
hipError_t launch_post(const PostArgs& a, hipStream_t st);
hipError_t launch_expand_read_bytes(const int64_t* read_off, int64_t n_reads, const uint8_t* triplets,
                                    uint8_t* ins, uint8_t* del, uint8_t* gcp, hipStream_t st);
}  // namespace bvp

static_assert(bvp::kBandHapMax <= bvp::kStackHmax,
              "rejected banded pairs go to the stacked kernel's list");

namespace {

using namespace bvp;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess)                                                        \
      return fail(BV_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
  } while (0)

// ---- tables -------------------------------------------------------------
// Same quantities as the reference's Context<T> (lib/intel_native/pairhmm/
// Context.h:42-61 Jacobian and match-to-match tables, :63-90 quantised
// log10-sum, :105-109 / :145-149 ph2pr and K), computed once on the host and
// uploaded.  Only qualities below 128 are reachable (the kernel masks & 127,
// avx-pairhmm-template.h:110-125), so the match-to-match table stops there.
constexpr int kCounters = 88;  // device counters of a plan, kCtrStride ints apart (see bv_plan)
constexpr int kFeatCounters = 24;  // [24..31]: the features kernel's per-XCD task claims
constexpr int kPredCounters = 72;  // [72..79]: the split prologue's second kernel's task claims
constexpr int kBandClaims = 32;    // [32..47]: the banded kernels' per-XCD chunk claims
constexpr int kQuadClaims = 56;    // [56..63]: the quad banded kernel's per-XCD chunk claims
constexpr int kItem1Count = 64;    // [64]: the quad shape's leftover wide pairs
constexpr int kItem48Count = 13;   // [13]: the mid band's item count
constexpr int kQuad48Claims = 80;  // [80..87]: the mid-band quad kernel's per-XCD chunk claims
constexpr int kBandRescueCount = 50;  // [50]: short pairs whose FP32 band underflowed, on the FP64 band
constexpr int kBandRescueQueue = 51;  // [51]: its queue
constexpr int kBandRescueAccepted = 52;  // [52]: the pairs it accepted
constexpr int kLong4Count = 53;  // [53]: the FP32 moving band's quad-unit items
constexpr int kLong4Queue = 54;  // [54]: their queue
constexpr int kJac = 80001;
constexpr int kMmEntries = (128 * 129) / 2;

struct HostTables {
  float ph2pr_f[128];
  double ph2pr_d[128];
  float mm_f[kMmEntries];
  double mm_d[kMmEntries];
  double err_scalar[256];
};

template <typename T>
T log10_sum_quantised(const std::vector<T>& jac, T a, T b) {
  T lo = a, hi = b;
  if (lo > hi) std::swap(lo, hi);
  const T diff = hi - lo;
  if (diff >= T(8)) return hi;
  const T x = diff * static_cast<T>(1.0 / 0.0001);
  const int k = (x > T(0)) ? static_cast<int>(x + T(0.5)) : static_cast<int>(x - T(0.5));
  return hi + jac[k];
}

void build_tables(HostTables& t) {
  std::vector<float> jf(kJac);
  std::vector<double> jd(kJac);
  for (int k = 0; k < kJac; ++k) {
    const double v = std::log10(1.0 + std::pow(10.0, -static_cast<double>(k) * 0.0001));
    jd[k] = v;
    jf[k] = static_cast<float>(v);
  }
  const double inv_ln10 = 1.0 / std::log(10.0);
  for (int i = 0, off = 0; i < 128; off += ++i) {
    for (int j = 0; j <= i; ++j) {
      const double sf = static_cast<double>(log10_sum_quantised<float>(
          jf, static_cast<float>(-0.1 * i), static_cast<float>(-0.1 * j)));
      t.mm_f[off + j] = static_cast<float>(
          std::pow(10.0, std::log1p(-std::min(1.0, std::pow(10.0, sf))) * inv_ln10));
      const double sd = log10_sum_quantised<double>(jd, -0.1 * i, -0.1 * j);
      t.mm_d[off + j] = std::pow(10.0, std::log1p(-std::min(1.0, std::pow(10.0, sd))) * inv_ln10);
    }
  }
  for (int x = 0; x < 128; ++x) {
    t.ph2pr_d[x] = std::pow(10.0, -static_cast<double>(x) / 10.0);
    t.ph2pr_f[x] = powf(10.f, -static_cast<float>(x) / 10.f);
  }
  // Scalar mode: QualityUtils::qual_to_error_prob (quality_utils.hpp:25-28,
  // 71-76) indexes a 128-entry table with the raw char; bytes >= 128 are
  // rejected at plan creation.
  for (int x = 0; x < 256; ++x) t.err_scalar[x] = x < 128 ? std::pow(10.0, x / -10.0) : 1.0;
}

// ---- device context -----------------------------------------------------
struct DeviceCtx {
  bool ready = false;
  int device = 0;
  int cus = 0;
  int bpc_f32 = 1, bpc_f64 = 1, bpc_long = 1;
  std::string arch;
  HostTables host{};
  float* d_ph2pr_f = nullptr;
  double* d_ph2pr_d = nullptr;
  float* d_mm_f = nullptr;
  double* d_mm_d = nullptr;
  double* d_err_scalar = nullptr;
  // device address of the allocation-free entry point's static scratch on
  // this device (looked up once at init, with the device current)
  void* d_noalloc_scratch = nullptr;
  // Smith-Waterman workspace kept across bv_sw_align calls (grown, never
  // shrunk, until bv_pairhmm_shutdown): per-wavefront scratch and the queues
  char* d_sw_scratch = nullptr;
  int64_t sw_scratch_bytes = 0;
  int* d_sw_queue = nullptr;
};

std::mutex g_mu;
std::mutex g_sw_mu;  // one bv_sw_align at a time owns the workspace
DeviceCtx g_ctx;

int ensure_init_locked(int device) {
  if (g_ctx.ready) return BV_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(BV_ENODEV, "no HIP device");
  if (device < 0) HIP_TRY(hipGetDevice(&device));
  if (device >= n) return fail(BV_ENODEV, "device ordinal out of range");
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  g_ctx.device = device;
  g_ctx.cus = prop.multiProcessorCount;
  g_ctx.arch = prop.gcnArchName;
  if (g_ctx.arch.rfind("gfx950", 0) != 0)
    return fail(BV_ENODEV, "this library is built for gfx950, device is " + g_ctx.arch);
  build_tables(g_ctx.host);
  auto& h = g_ctx.host;
  HIP_TRY(hipMalloc(&g_ctx.d_ph2pr_f, sizeof(h.ph2pr_f)));
  HIP_TRY(hipMalloc(&g_ctx.d_ph2pr_d, sizeof(h.ph2pr_d)));
  HIP_TRY(hipMalloc(&g_ctx.d_mm_f, sizeof(h.mm_f)));
  HIP_TRY(hipMalloc(&g_ctx.d_mm_d, sizeof(h.mm_d)));
  HIP_TRY(hipMalloc(&g_ctx.d_err_scalar, sizeof(h.err_scalar)));
  HIP_TRY(hipMemcpy(g_ctx.d_ph2pr_f, h.ph2pr_f, sizeof(h.ph2pr_f), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(g_ctx.d_ph2pr_d, h.ph2pr_d, sizeof(h.ph2pr_d), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(g_ctx.d_mm_f, h.mm_f, sizeof(h.mm_f), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(g_ctx.d_mm_d, h.mm_d, sizeof(h.mm_d), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(g_ctx.d_err_scalar, h.err_scalar, sizeof(h.err_scalar), hipMemcpyHostToDevice));
  HIP_TRY(noalloc_scratch_address(&g_ctx.d_noalloc_scratch));
  if (!g_ctx.d_noalloc_scratch) return fail(BV_EHIP, "no device address for the allocation-free scratch");
  g_ctx.bpc_f32 = max_blocks_per_cu<float>();
  g_ctx.bpc_f64 = max_blocks_per_cu<double>();
  g_ctx.bpc_long = std::max(band_long_blocks_per_cu(1), band_long_blocks_per_cu(0));
  g_ctx.ready = true;
  return BV_OK;
}

// ---- read/haplotype encoding ---------------------------------------------
// AVX mode (pairhmm_common.h:26-44, avx-pairhmm-template.h:3-35): codes
// A0 C1 T2 G3, 'N' is a wildcard on either side, every other byte acts as
// 'A'.  Scalar mode (pairhmm.hpp:48): raw byte equality, uppercase 'N' is the
// only wildcard; every distinct non-'N' read byte gets a match class.
struct Alphabet {
  uint8_t read_class[256];
  uint8_t hap_mask[256];
};

void alphabet_avx(Alphabet& a) {
  for (int b = 0; b < 256; ++b) {
    a.read_class[b] = 0;
    a.hap_mask[b] = 1;  // acts as 'A'
  }
  const char codes[4] = {'A', 'C', 'T', 'G'};
  for (int k = 0; k < 4; ++k) {
    a.read_class[static_cast<uint8_t>(codes[k])] = static_cast<uint8_t>(k);
    a.hap_mask[static_cast<uint8_t>(codes[k])] = static_cast<uint8_t>(1u << k);
  }
  a.read_class[static_cast<uint8_t>('N')] = kWildcard;
  a.hap_mask[static_cast<uint8_t>('N')] = 0xFF;
}

int alphabet_scalar(Alphabet& a, const uint8_t* bases, int64_t n) {
  bool seen[256] = {false};
  for (int64_t i = 0; i < n; ++i) seen[bases[i]] = true;
  int next = 0;
  for (int b = 0; b < 256; ++b) {
    a.read_class[b] = 0;
    a.hap_mask[b] = 0;
  }
  for (int b = 0; b < 256; ++b) {
    if (!seen[b] || b == 'N') continue;
    if (next >= kSlots)
      return fail(BV_EINVAL, "scalar mode supports at most 8 distinct read base bytes");
    a.read_class[b] = static_cast<uint8_t>(next);
    a.hap_mask[b] = static_cast<uint8_t>(1u << next);
    ++next;
  }
  a.read_class[static_cast<uint8_t>('N')] = kWildcard;
  a.hap_mask[static_cast<uint8_t>('N')] = 0xFF;
  return BV_OK;
}

template <typename F>
void parallel_for(int64_t n, F&& f, int64_t serial_below = 1 << 16) {
  const int nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < serial_below || nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int64_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([&, a, b] { f(a, b); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

// ---- plan ---------------------------------------------------------------
struct bv_plan {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[9] = {};
  bv_options opt{};
  int64_t n_pairs = 0, cells = 0, n_reads = 0, n_haps = 0, read_bytes = 0, hap_bytes = 0;
  int hmax = 0;
  int n_tasks = 0;
  int64_t n_subtasks = 0;  // tasks counted per 64 haplotypes (the fused prologue's pieces)
  int grid32 = 0, grid64 = 0, grid_band = 0, grid_long = 1;
  int grid_band1 = 1;  // one-pair banded kernels after the quad one (leftovers, narrow band)
  int grid32_max = 0, grid64_max = 0, grid_band_max = 0, grid_long_max = 1;  // scratch is sized for these
  int sel_cols = 0;  // short-read banded kernel: selector columns per class
  int wide_pairs = kQuadPairs;  // wide-band shape: pairs per wavefront (1 or kQuadPairs)
  bool mid_on = false;          // bands of 33-48 diagonals from bv_plan_set_bands on the 48-diagonal quad kernel
  int quad_fill = 3;            // quad shape: a read's last wide pairs fill a padded unit when at least this many
  bool long_fp32 = true;        // the long reads' moving band in FP32 (rescaled every 64 rows), else FP64
  int64_t stride32 = 0, stride64 = 0;
  uint8_t* d_rclass = nullptr;
  uint8_t* d_q = nullptr;
  uint8_t* d_ins = nullptr;
  uint8_t* d_del = nullptr;
  uint8_t* d_gcp = nullptr;
  uint8_t* d_gcp_const = nullptr;  // per read: 1 if its gap continuation is constant
  int64_t* d_read_off = nullptr;
  uint8_t* d_hmask = nullptr;
  int64_t* d_hap_off = nullptr;
  Task* d_tasks = nullptr;         // longest first (forward kernels)
  Task* d_tasks_region = nullptr;  // region order (seed features: haplotype tables stay in L2)
  // FP32 full-matrix pass: stacked tasks (runs of reads of a region against
  // its haplotypes, pairhmm_stack_kernel), longest first, and the tasks of
  // the reads they leave out (a haplotype longer than kStackHmax), for the
  // striped kernel
  StackTask* d_stasks = nullptr;
  int n_stasks = 0;
  Task* d_tasks_rest = nullptr;
  int n_tasks_rest = 0;
  int grid_stack = 1, grid_stack_max = 1;
  int grid_fbstack = 1;  // stacked kernel over the learned-band run's full-matrix list
  int stack_cols = 0;  // LDS carry row length of the stacked kernel
  Task* d_tasks64 = nullptr;       // long reads, FP64 first (FP32 mode), one pair each
  Task* d_tasks64r = nullptr;      // the same long reads, one task per read (all its haplotypes)
  int n_tasks64r = 0;
  BandItem* d_items_long4 = nullptr;  // the FP32 moving band's quad units (four pairs of a read)
  Task* d_verify = nullptr;        // their pairs that may not underflow FP32
  int n_tasks64 = 0;
  Task* d_rescue = nullptr;
  Task* d_fallback = nullptr;       // full-matrix list of the learned-band run (stacked kernel)
  Task* d_fallback_long = nullptr;  // its pairs with haplotypes over kStackHmax (striped kernel)
  int64_t n_longhap_pairs = 0;      // short-read pairs with such haplotypes (that list's capacity)
  int2* d_band = nullptr;
  BandItem* d_items = nullptr;    // banded work lists, built on the device each run
  BandItem* d_items1 = nullptr;   // the quad shape's leftover wide and mid pairs
  BandItem* d_items48 = nullptr;  // the mid band's quad groups
  BandItem* d_band_rescue = nullptr;  // short pairs whose FP32 band sum underflowed
  int band_rescue_cap = 0;
  int grid_rescue = 1;
  BandItem* d_items32 = nullptr;
  BandItem* d_items_long = nullptr;  // long reads on the moving band
  Task* d_long_full = nullptr;       // long reads left to the full-matrix FP64 pass
  // learned bands: features and predictions computed on the device each run
  float* d_X = nullptr;
  int* d_anchor = nullptr;
  int2* d_ends = nullptr;
  unsigned long long* d_htab = nullptr;  // per haplotype 12-mer hash tables
  int64_t* d_htab_off = nullptr;
  std::vector<int64_t> hap_len;          // host copy of haplotype lengths
  float* d_Y = nullptr;
  float* d_model = nullptr;  // MLP parameters (copy)
  bool has_model = false;
  // learned bands: 0 the separate features, MLP, band-rule and dispatch
  // kernels; 1 one kernel from features to work lists; 2 the features
  // kernel writing compact records, then one kernel from them to work lists
  int prologue_mode = 2;
  uint8_t* d_rec = nullptr;  // split prologue: per pair records (kRecBytes)
  float* d_trec = nullptr;   // and per task records
  int grid_prologue = 1;
  int band_margin = 8, band_max_width = 256;
  bool has_band = false;
  // Counter i lives at d_counters + i * kCtrStride, on a line of its own.
  // [0] FP32 queue, [1] FP64 queue, [2] rescue count, [3] full-matrix FP32
  // list count, [4] its queue, [5] accepted banded pairs, [6] rescued cells
  // (u64), [8] swept cells (u64), [10] wide banded item count, [11] stacked
  // FP32 queue, [12] narrow banded item count, [13] mid banded item count,
  // [14] verify count, [15] its queue, [16] FP64-first queue, [17] moving-band item count, [18] its
  // queue, [19] long full-matrix count, [20] its queue, [21] accepted
  // moving-band pairs, [22] FP64 cells (u64), [24..31] features task claims
  // (one per XCD), [32..47] banded chunk claims (per XCD, wide then narrow
  // kernel), [48] full-matrix list with haplotypes longer than kStackHmax,
  // [49] its queue, [50] FP64 band rescue count, [51] its queue, [52] its
  // accepted pairs, [53] the FP32 moving band's quad-unit items, [54] their
  // queue, [56..63]
  // quad banded chunk claims (per XCD), [64] the
  // quad shape's leftover wide and mid pairs, [72..79] the split prologue's
  // second kernel's task claims (per XCD), [80..87] the mid-band quad
  // kernel's chunk claims (per XCD); the rest unused
  int* d_counters = nullptr;
  unsigned long long* d_rescue_cells = nullptr;
  double* d_out = nullptr;
  uint8_t* d_form = nullptr;
  void* d_carry32 = nullptr;
  void* d_carry64 = nullptr;
  float last_ms32 = 0.f, last_ms64 = 0.f, last_msfb = 0.f, last_msband = 0.f, last_mslong = 0.f;
  // post-processing (bv_plan_postprocess): region tables (host copies kept
  // from bv_plan_create, device copies made on first use) and its buffers
  int64_t n_regions = 0;
  std::vector<int64_t> region_read_off, region_hap_off, region_pair_off;
  int64_t* d_region_tab = nullptr;  // [read_off | hap_off | pair_off], 3 x (n_regions + 1)
  int32_t* d_hap_allele = nullptr;  // n_haps
  int32_t* d_n_alleles = nullptr;   // n_regions
  int64_t* d_allele_off = nullptr;  // n_regions + 1
  double* d_lnorm = nullptr;        // n_pairs
  uint8_t* d_keep = nullptr;        // n_reads
  double* d_allele = nullptr;
  int64_t allele_cap = 0;
  hipEvent_t ev_post[2] = {nullptr, nullptr};
  float last_mspost = 0.f;
  float setup_ms_host = 0.f, setup_ms_upload = 0.f;  // bv_plan_create: host work, uploads (+ expansion)
  int64_t setup_bytes = 0;                           // bytes it uploaded
  // long reads on the band run on a second stream beside the short-read
  // passes that follow the banded kernels, with their own carry scratch
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void* d_carry32_long = nullptr;
  void* d_carry64_long = nullptr;
  // one run captured as a graph (bv_plan_run), rebuilt when the band setup changes
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool timed = false;
};

namespace {

// Counter i of a plan's counter block.
int* ctr(const bv_plan* p, int i) { return p->d_counters + i * kCtrStride; }

void free_plan(bv_plan* p) {
  if (!p) return;
  // d_rescue_cells points into d_counters and is not freed on its own.
  void* bufs[] = {p->d_rclass, p->d_q, p->d_ins, p->d_del, p->d_gcp, p->d_gcp_const, p->d_read_off,
                  p->d_hmask, p->d_hap_off, p->d_tasks, p->d_tasks_region, p->d_tasks64, p->d_tasks64r,
                  p->d_items_long4, p->d_verify,
                  p->d_stasks, p->d_tasks_rest, p->d_fallback_long,
                  p->d_rescue, p->d_counters,
                  p->d_out, p->d_form, p->d_carry32, p->d_carry64, p->d_fallback, p->d_band, p->d_items, p->d_items48, p->d_band_rescue,
                  p->d_items1, p->d_items32, p->d_items_long, p->d_long_full,
                  p->d_X, p->d_anchor, p->d_Y, p->d_model, p->d_ends, p->d_htab, p->d_htab_off, p->d_rec, p->d_trec,
                  p->d_region_tab, p->d_hap_allele, p->d_n_alleles, p->d_allele_off, p->d_lnorm,
                  p->d_keep, p->d_allele, p->d_carry32_long, p->d_carry64_long};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto& e : p->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : p->ev_post)
    if (e) (void)hipEventDestroy(e);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  if (p->side) (void)hipStreamDestroy(p->side);
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

// Allocates max(n, capacity, 1) elements and copies n of them from src.
template <typename T>
int dev_upload(T** dst, const T* src, int64_t n, hipStream_t st, int64_t capacity = 0, int64_t* counted = nullptr) {
  const size_t bytes = sizeof(T) * static_cast<size_t>(std::max<int64_t>(std::max<int64_t>(n, capacity), 1));
  HIP_TRY(hipMalloc(dst, bytes));
  if (n > 0) HIP_TRY(hipMemcpyAsync(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice, st));
  if (counted && n > 0) *counted += static_cast<int64_t>(sizeof(T)) * n;
  return BV_OK;
}

// A task list written longest first into *dst: uploaded with its costs (or
// taken from src_dev, the same list already on the device in input order)
// and ordered there by a stable sort on descending cost.
template <typename T>
int upload_longest_first(T** dst, const T* src_dev, const std::vector<T>& t, const std::vector<uint64_t>& cost,
                         hipStream_t st, int64_t* up) {
  const int64_t n = static_cast<int64_t>(t.size());
  HIP_TRY(hipMalloc(dst, sizeof(T) * static_cast<size_t>(std::max<int64_t>(n, 1))));
  if (n == 0) return BV_OK;
  T* tmp = nullptr;
  uint64_t* dcost = nullptr;
  int rc = BV_OK;
  if (!src_dev) {
    rc = dev_upload(&tmp, t.data(), n, st, 0, up);
    src_dev = tmp;
  }
  if (!rc) rc = dev_upload(&dcost, cost.data(), n, st, 0, up);
  uint64_t mx = 0;
  for (uint64_t c : cost) mx = std::max(mx, c);
  int bits = 1;
  while (bits < 64 && (mx >> bits) != 0) ++bits;
  if (!rc && sort_longest_first(src_dev, dcost, n, bits, *dst, st) != hipSuccess)
    rc = fail(BV_EHIP, "task sort failed");
  if (tmp) (void)hipFree(tmp);
  if (dcost) (void)hipFree(dcost);
  return rc;
}

// Bases at two bits each (AVX mode: A C T G are classes 0..3, any other byte
// acts as A), four per byte, base i in bits 2*(i%4); the positions whose
// byte is 'N' (the wildcard) listed apart.  The device expands them to the
// per-base class bytes and haplotype match masks the kernels read.
void pack_bases(const uint8_t* bases, int64_t n, const uint8_t* code_of, uint8_t nbyte,
                std::vector<uint8_t>& packed, std::vector<int64_t>& exc) {
  packed.assign(static_cast<size_t>((n + 3) / 4), 0);
  exc.clear();
  std::mutex mu;
  parallel_for(static_cast<int64_t>(packed.size()), [&](int64_t a0, int64_t a1) {
    std::vector<int64_t> mine;
    for (int64_t k = a0; k < a1; ++k) {
      uint8_t v = 0;
      for (int64_t i = 4 * k, e = std::min<int64_t>(n, 4 * k + 4); i < e; ++i) {
        v |= static_cast<uint8_t>((code_of[bases[i]] & 3u) << (2 * (i & 3)));
        if (bases[i] == nbyte) mine.push_back(i);
      }
      packed[k] = v;
    }
    std::lock_guard<std::mutex> lk(mu);
    exc.insert(exc.end(), mine.begin(), mine.end());
  });
}

int validate(const bv_region_batch* b) {
  if (!b) return fail(BV_EINVAL, "batch is NULL");
  if (b->n_regions < 0 || b->n_reads < 0 || b->n_haps < 0) return fail(BV_EINVAL, "negative count");
  if (b->n_regions > 0 && (!b->region_read_off || !b->region_hap_off))
    return fail(BV_EINVAL, "region offsets are NULL");
  if (b->n_reads > 0 && (!b->read_off || !b->read_bases || !b->read_qual || !b->read_ins ||
                         !b->read_del || !b->read_gcp))
    return fail(BV_EINVAL, "read arrays are NULL");
  if (b->n_haps > 0 && (!b->hap_off || !b->hap_bases)) return fail(BV_EINVAL, "hap arrays are NULL");
  if (b->n_reads >= (1ll << 31) || b->n_haps >= (1ll << 31))
    return fail(BV_EINVAL, "more than 2^31 reads or haplotypes");
  for (int64_t g = 0; g < b->n_regions; ++g) {
    if (b->region_read_off[g] > b->region_read_off[g + 1] ||
        b->region_hap_off[g] > b->region_hap_off[g + 1])
      return fail(BV_EINVAL, "region offsets decrease at region " + std::to_string(g));
    const bool has_reads = b->region_read_off[g + 1] > b->region_read_off[g];
    const bool has_haps = b->region_hap_off[g + 1] > b->region_hap_off[g];
    if (has_reads && !has_haps)
      return fail(BV_EINVAL, "region " + std::to_string(g) + " has reads but no haplotypes");
  }
  if (b->n_regions > 0 &&
      (b->region_read_off[0] < 0 || b->region_read_off[b->n_regions] > b->n_reads ||
       b->region_hap_off[0] < 0 || b->region_hap_off[b->n_regions] > b->n_haps))
    return fail(BV_EINVAL, "region offsets out of range");
  for (int64_t r = 0; r < b->n_reads; ++r) {
    const int64_t len = b->read_off[r + 1] - b->read_off[r];
    if (len <= 0) return fail(BV_EINVAL, "read " + std::to_string(r) + " is empty");
    if (len > (1 << 24)) return fail(BV_EINVAL, "read " + std::to_string(r) + " is longer than 2^24");
  }
  for (int64_t h = 0; h < b->n_haps; ++h) {
    const int64_t len = b->hap_off[h + 1] - b->hap_off[h];
    if (len <= 0) return fail(BV_EINVAL, "haplotype " + std::to_string(h) + " is empty");
    if (len > (1 << 24)) return fail(BV_EINVAL, "haplotype " + std::to_string(h) + " is longer than 2^24");
  }
  return BV_OK;
}

// Long reads run in FP64 first; a pair whose log10 L is at least this may
// not underflow FP32 (S_f >= 1e-28 means log10 L >= -28 - log10(2^120)), so
// the reference would keep its FP32 result.  The margin covers FP32 rounding.
constexpr double kVerifyAbove = -28.0 - 120.0 * 0.30102999566398119521 - 0.05;

KArgs base_args(const bv_plan* p) {
  KArgs a{};
  a.rclass = p->d_rclass;
  a.q = p->d_q;
  a.ins = p->d_ins;
  a.del = p->d_del;
  a.gcp = p->d_gcp;
  a.read_off = p->d_read_off;
  a.hmask = p->d_hmask;
  a.hap_off = p->d_hap_off;
  a.out = p->d_out;
  a.form = p->d_form;
  a.tab.ph2pr_f = g_ctx.d_ph2pr_f;
  a.tab.ph2pr_d = g_ctx.d_ph2pr_d;
  a.tab.mm_f = g_ctx.d_mm_f;
  a.tab.mm_d = g_ctx.d_mm_d;
  a.tab.err_scalar = g_ctx.d_err_scalar;
  for (int k = 0; k < 7; ++k) a.tab.trans[k] = p->opt.trans[k];
  a.mode = p->opt.mode;
  a.renorm = p->opt.fp64_renorm;
  a.rescue_cells = p->d_rescue_cells;
  a.cells_done = reinterpret_cast<unsigned long long*>(ctr(p, 8));
  a.fp64_cells = reinterpret_cast<unsigned long long*>(ctr(p, 22));
  a.hcap = p->hmax;
  return a;
}

FeatArgs feat_args(const bv_plan* p) {
  FeatArgs f{};
  f.rclass = p->d_rclass;
  f.q = p->d_q;
  f.read_off = p->d_read_off;
  f.hmask = p->d_hmask;
  f.hap_off = p->d_hap_off;
  f.tasks = p->d_tasks_region;
  f.n_tasks = p->n_tasks;
  f.X = p->d_X;
  f.anchor = p->d_anchor;
  f.ends = p->d_ends;
  f.htab = p->d_htab;
  f.htab_off = p->d_htab_off;
  f.next = ctr(p, kFeatCounters);
  f.rec = nullptr;
  f.trec = nullptr;
  return f;
}

hipError_t launch_features(bv_plan* p) {
  const FeatArgs f = feat_args(p);  // writes X (allocated by ensure_x)
  // the claim counters are cleared by a kernel (graph-safe, no memset node)
  if (hipError_t e = launch_zero_ints(f.next, kXcdCount * kCtrStride, p->stream); e != hipSuccess) return e;
  const int grid = std::max(1, std::min(p->n_tasks, g_ctx.cus * 32));
  return launch_band_features(f, grid, p->stream);
}

// The banded kernels' persistent grid: the wide kernel's occupancy (for the
// plan's shape and selector window) on every CU, at most one wavefront per pair.
int band_grid_cap(const bv_plan* p, int wide_pairs) {
  return static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(static_cast<int64_t>(g_ctx.cus) * band_blocks_per_cu(p->sel_cols, wide_pairs),
                           std::max<int64_t>(1, p->n_pairs))));
}

int ensure_band_lists(bv_plan* p) {
  const int64_t n = std::max<int64_t>(1, p->n_pairs);
  // the quad shape leaves at most kQuadPairs - 1 wide and as many mid pairs
  // per task (per 64 haplotypes of a task in the fused prologue) to the
  // one-pair kernel; narrow items are padded to an even count per task (or
  // piece)
  // the quad list: the pairs, and at most kQuadPairs - 1 padding slots per
  // task piece (a task or a 64-haplotype part of it; the prologue never
  // splits a piece over two flushes)
  if (!p->d_items &&
      hipMalloc(&p->d_items, sizeof(BandItem) * (n + (kQuadPairs - 1) * p->n_subtasks)) != hipSuccess)
    return fail(BV_ENOMEM, "band work list allocation failed");
  if (p->mid_on && !p->d_items48 && hipMalloc(&p->d_items48, sizeof(BandItem) * n) != hipSuccess)
    return fail(BV_ENOMEM, "band work list allocation failed");
  // FP64 band rescue: room for a sixteenth of the pairs (more go the
  // full-matrix way)
  if (!p->d_band_rescue) {
    p->band_rescue_cap = static_cast<int>(std::min<int64_t>(INT_MAX / 2, std::max<int64_t>(4096, n / 16)));
    if (hipMalloc(&p->d_band_rescue, sizeof(BandItem) * p->band_rescue_cap) != hipSuccess)
      return fail(BV_ENOMEM, "band work list allocation failed");
  }
  if (!p->d_items1 &&
      hipMalloc(&p->d_items1, sizeof(BandItem) * std::max<int64_t>(1, 2 * (kQuadPairs - 1) * p->n_subtasks)) !=
          hipSuccess)
    return fail(BV_ENOMEM, "band work list allocation failed");
  if (!p->d_items32 && hipMalloc(&p->d_items32, sizeof(BandItem) * (n + p->n_subtasks)) != hipSuccess)
    return fail(BV_ENOMEM, "band work list allocation failed");
  if (!p->d_fallback_long &&
      hipMalloc(&p->d_fallback_long, sizeof(Task) * std::max<int64_t>(1, p->n_longhap_pairs)) != hipSuccess)
    return fail(BV_ENOMEM, "band work list allocation failed");
  const int64_t n64 = std::max(1, p->n_tasks64);
  if (!p->d_items_long && hipMalloc(&p->d_items_long, sizeof(BandItem) * n64) != hipSuccess)
    return fail(BV_ENOMEM, "band work list allocation failed");
  if (!p->d_items_long4 && hipMalloc(&p->d_items_long4, sizeof(BandItem) * n64) != hipSuccess)
    return fail(BV_ENOMEM, "band work list allocation failed");
  if (!p->d_long_full && hipMalloc(&p->d_long_full, sizeof(Task) * n64) != hipSuccess)
    return fail(BV_ENOMEM, "band work list allocation failed");
  if (p->n_tasks64 > 0) {
    if (!p->d_carry32_long &&
        hipMalloc(&p->d_carry32_long, sizeof(float) * 2 * p->stride32 * p->grid32_max) != hipSuccess)
      return fail(BV_ENOMEM, "carry scratch allocation failed");
    if (!p->d_carry64_long &&
        hipMalloc(&p->d_carry64_long, sizeof(double) * 2 * p->stride64 * p->grid64_max) != hipSuccess)
      return fail(BV_ENOMEM, "carry scratch allocation failed");
  }
  return BV_OK;
}

// The per-pair feature matrix (kFeat floats a pair), anchors, seed-end
// spans and predictions: for bv_plan_features and the separate prologue only.
int ensure_x(bv_plan* p) {
  const int64_t n = std::max<int64_t>(1, p->n_pairs);
  if (!p->d_X && hipMalloc(&p->d_X, sizeof(float) * kFeat * n) != hipSuccess)
    return fail(BV_ENOMEM, "feature buffer allocation failed");
  if (!p->d_anchor && hipMalloc(&p->d_anchor, sizeof(int) * n) != hipSuccess)
    return fail(BV_ENOMEM, "anchor buffer allocation failed");
  if (!p->d_ends && hipMalloc(&p->d_ends, sizeof(int2) * n) != hipSuccess)
    return fail(BV_ENOMEM, "seed end buffer allocation failed");
  if (!p->d_Y && hipMalloc(&p->d_Y, sizeof(float) * 2 * n) != hipSuccess)
    return fail(BV_ENOMEM, "prediction buffer allocation failed");
  return BV_OK;
}

int ensure_feature_buffers(bv_plan* p) {
  const int64_t n = std::max<int64_t>(1, p->n_pairs);
  int rc0 = BV_OK;
  if (p->prologue_mode == 0 && (rc0 = ensure_x(p))) return rc0;
  if (p->prologue_mode == 2 && !p->d_rec) {
    if (hipMalloc(&p->d_rec, static_cast<size_t>(kRecBytes) * n) != hipSuccess ||
        hipMalloc(&p->d_trec, sizeof(float) * 3 * std::max(1, p->n_tasks)) != hipSuccess)
      return fail(BV_ENOMEM, "feature record allocation failed");
  }
  if (!p->d_htab) {
    // haplotype 12-mer tables: fixed for the plan, built once
    std::vector<int64_t> off(static_cast<size_t>(p->n_haps) + 1, 0);
    for (int64_t h = 0; h < p->n_haps; ++h) {
      int64_t sz = 64;
      while (sz < 2 * p->hap_len[h]) sz <<= 1;
      off[h + 1] = off[h] + (p->hap_len[h] >= kSeedK ? sz : 1);
    }
    if (hipMalloc(&p->d_htab, sizeof(unsigned long long) * std::max<int64_t>(1, off.back())) != hipSuccess ||
        hipMalloc(&p->d_htab_off, sizeof(int64_t) * off.size()) != hipSuccess)
      return fail(BV_ENOMEM, "haplotype table allocation failed");
    HIP_TRY(hipMemcpyAsync(p->d_htab_off, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice,
                           p->stream));
    HIP_TRY(launch_hap_tables(p->d_hmask, p->d_hap_off, p->d_htab_off, p->n_haps, p->d_htab, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
  }
  if (!p->d_band && hipMalloc(&p->d_band, sizeof(int2) * n) != hipSuccess)
    return fail(BV_ENOMEM, "band buffer allocation failed");
  return ensure_band_lists(p);
}

BArgs band_args(const bv_plan* p) {
  BArgs b{};
  b.rclass = p->d_rclass;
  b.q = p->d_q;
  b.ins = p->d_ins;
  b.del = p->d_del;
  b.gcp = p->d_gcp;
  b.gcp_const = p->d_gcp_const;
  b.read_off = p->d_read_off;
  b.hmask = p->d_hmask;
  b.hap_off = p->d_hap_off;
  b.tasks = p->d_tasks;
  b.n_tasks = p->n_tasks;
  b.band = p->d_band;
  b.items = p->d_items;
  b.item_count = ctr(p, 10);
  b.items1 = p->d_items1;
  b.item1_count = ctr(p, kItem1Count);
  b.items32 = p->d_items32;
  b.item32_count = ctr(p, 12);
  b.items48 = p->d_items48;
  b.item48_count = ctr(p, kItem48Count);
  b.full = p->d_fallback;
  b.full_count = ctr(p, 3);
  b.full_long = p->d_fallback_long;
  b.full_long_count = ctr(p, 48);
  b.ph2pr = g_ctx.d_ph2pr_f;
  b.mm = g_ctx.d_mm_f;
  b.out = p->d_out;
  b.form = p->d_form;
  b.tol = p->opt.band_tol;
  b.accepted = ctr(p, 5);
  b.claims = ctr(p, kBandClaims);
  b.quad_claims = ctr(p, kQuadClaims);
  b.quad48_claims = ctr(p, kQuad48Claims);
  b.cells_done = reinterpret_cast<unsigned long long*>(ctr(p, 8));
  b.tasks64 = p->d_tasks64;
  b.n_tasks64 = p->n_tasks64;
  b.long_on = p->opt.fp64_renorm != 0;
  b.long_fp32 = p->long_fp32 ? 1 : 0;
  b.long_items = p->d_items_long;
  b.long_count = ctr(p, 17);
  b.long_queue = ctr(p, 18);
  b.long_full = p->d_long_full;
  b.long_full_count = ctr(p, 19);
  b.long_accepted = ctr(p, 21);
  b.ph2pr_d = g_ctx.d_ph2pr_d;
  b.mm_d = g_ctx.d_mm_d;
  b.verify_above = kVerifyAbove;
  b.fp64_cells = reinterpret_cast<unsigned long long*>(ctr(p, 22));
  b.sel_cols = p->sel_cols;
  b.wide_pairs = p->wide_pairs;
  b.mid_on = p->mid_on && p->d_items48 != nullptr;
  b.quad_fill = p->quad_fill;
  b.long_cap = std::max(1, p->n_tasks64);
  b.tasks64r = p->d_tasks64r;
  b.n_tasks64r = p->n_tasks64r;
  b.long_items4 = p->d_items_long4;
  b.long4_count = ctr(p, kLong4Count);
  b.long4_queue = ctr(p, kLong4Queue);
  b.band_rescue = p->d_band_rescue;
  b.band_rescue_count = ctr(p, kBandRescueCount);
  b.band_rescue_cap = p->band_rescue_cap;
  return b;
}


// Long reads in FP64 first; pairs within reach of FP32's range are queued
// for FP32 and keep the FP32 result when it does not underflow (the
// reference's rule, pairhmm_avx.hpp:140-149).  S_f >= 1e-28 means
// log10 L >= -28 - log10(2^120); the FP64 value screens with a margin.
// With bands, the moving-band kernel takes the long pairs first and leaves
// the full-matrix pass those without a band or rejected (and any within
// FP32's reach, so that the rule above applies to them).
int enqueue_long(bv_plan* p, bool timed, hipStream_t st, void* carry32, void* carry64) {
  KArgs a = base_args(p);
  if (p->has_band) {
    const BArgs b = band_args(p);
    if (timed) HIP_TRY(hipEventRecord(p->ev[7], st));
    HIP_TRY(launch_band_long(b, p->grid_long, st));
    if (timed) HIP_TRY(hipEventRecord(p->ev[8], st));
    a.tasks = p->d_long_full;
    a.n_tasks = 0;
    a.n_tasks_dev = ctr(p, 19);
    a.counter = ctr(p, 20);
  } else {
    a.tasks = p->d_tasks64;
    a.n_tasks = p->n_tasks64;
    a.counter = ctr(p, 16);
  }
  a.carry = carry64;
  a.carry_stride = p->stride64;
  a.verify = p->d_verify;
  a.verify_count = ctr(p, 14);
  a.verify_above = kVerifyAbove;
  HIP_TRY(launch_forward<double>(a, p->grid64, st));
  KArgs v = base_args(p);
  v.tasks = p->d_verify;
  v.n_tasks = 0;
  v.n_tasks_dev = ctr(p, 14);
  v.counter = ctr(p, 15);
  v.carry = carry32;
  v.carry_stride = p->stride32;
  v.verify_pass = 1;
  HIP_TRY(launch_forward<float>(v, p->grid32, st));
  return BV_OK;
}

int enqueue_run(bv_plan* p, bool timed) {
  HIP_TRY(launch_zero_ints(p->d_counters, kCounters * kCtrStride, p->stream));
  if (p->n_tasks == 0 && p->n_tasks64 == 0) return BV_OK;
  const bool fp32 = (p->opt.mode == BV_MODE_AVX) && !p->opt.force_fp64;
  if (timed) HIP_TRY(hipEventRecord(p->ev[0], p->stream));
  const bool fused = fp32 && p->has_model && p->has_band && p->prologue_mode != 0;
  if (fused) {
    // Learned bands: seed features, MLP, band rules, band table and the
    // banded / full-matrix work lists, in one kernel or (split) as the
    // features kernel writing compact records and one kernel from them on.
    FeatArgs f = feat_args(p);
    if (p->prologue_mode == 2) {
      f.rec = p->d_rec;
      f.trec = p->d_trec;
      const int grid = std::max(1, std::min(p->n_tasks, g_ctx.cus * 32));
      HIP_TRY(launch_band_features(f, grid, p->stream));
      f.next = ctr(p, kPredCounters);
    }
    HIP_TRY(launch_band_prologue(f, band_args(p), p->d_band, p->d_model, p->band_margin, p->band_max_width,
                                 p->opt.band_narrow, p->grid_prologue, p->stream));
  } else if (fp32 && p->has_model) {
    // Learned bands: seed features, MLP, band table -- all on the device.
    HIP_TRY(launch_features(p));
    HIP_TRY(launch_mlp_forward(p->d_model, p->d_X, p->n_pairs, p->d_Y, nullptr, nullptr, p->stream));
    HIP_TRY(launch_band_from_prediction(p->d_Y, p->d_anchor, p->d_ends, p->d_X, p->n_pairs, p->band_margin,
                                        p->band_max_width, p->opt.band_narrow, p->d_band, p->stream));
  }
  if (fp32 && p->has_model) {
    // Long reads (outside the network's training range): start bands from
    // the seeds of their first bases; the moving band follows from there.
    // (After the prologue: its separate-kernel form writes every pair's band.)
    const int grid_anchor = std::max(1, std::min(p->n_tasks64, g_ctx.cus * 32));
    HIP_TRY(launch_long_anchors(p->d_rclass, p->d_read_off, p->d_htab_off, p->d_htab, p->d_tasks64,
                                p->n_tasks64, p->d_band, grid_anchor, p->stream));
  }
  // Long reads with bands: moving band, full-matrix FP64 for the pairs it
  // leaves, FP32 verify.  They run on the second stream, beside every
  // short-read pass (the banded kernels claim their work per XCD, so waves
  // that start late take less of it), with their own carry scratch; the run
  // joins the two streams before its end.
  const bool long_side = fp32 && p->has_band && p->n_tasks64 > 0 && p->d_carry32_long && p->d_carry64_long;
  if (long_side) {
    HIP_TRY(hipEventRecord(p->ev_fork, p->stream));
    HIP_TRY(hipStreamWaitEvent(p->side, p->ev_fork, 0));
    if (int rc = enqueue_long(p, timed, p->side, p->d_carry32_long, p->d_carry64_long)) return rc;
    HIP_TRY(hipEventRecord(p->ev_join, p->side));
  }
  if (fp32 && p->has_band) {
    // Pass 1: pairs with an eligible band on the banded kernel; the rest,
    // and every band it rejects, go to the full-matrix list.
    const BArgs b = band_args(p);
    if (!fused) HIP_TRY(launch_band_dispatch(b, p->stream));  // (the fused prologue wrote the lists)
    if (timed) HIP_TRY(hipEventRecord(p->ev[5], p->stream));
    HIP_TRY(launch_band_quads(b, p->grid_band, p->stream));
    HIP_TRY(launch_band_one(b, p->grid_band, p->grid_band1, p->stream));
    if (timed) HIP_TRY(hipEventRecord(p->ev[6], p->stream));
  } else if (fp32) {
    // Pass 1: every pair in FP32 on the full matrix: stacked reads, and the
    // reads of regions with haplotypes too long for them on the striped kernel.
    KArgs a = base_args(p);
    a.rescue = p->d_rescue;
    a.rescue_count = ctr(p, 2);
    if (p->n_stasks > 0) {
      a.stasks = p->d_stasks;
      a.n_tasks = p->n_stasks;
      a.counter = ctr(p, 11);
      a.carry_cols = p->stack_cols;
      HIP_TRY(launch_stack(a, p->grid_stack, p->stream));
    }
    if (p->n_tasks_rest > 0) {
      a.tasks = p->d_tasks_rest;
      a.n_tasks = p->n_tasks_rest;
      a.counter = ctr(p, 0);
      a.carry = p->d_carry32;
      a.carry_stride = p->stride32;
      HIP_TRY(launch_forward<float>(a, p->grid32, p->stream));
    }
  }
  if (timed) HIP_TRY(hipEventRecord(p->ev[1], p->stream));
  if (fp32 && p->has_band) {
    // Pass 2: the full-matrix list (FP32, same rescue rule): pairs on the
    // stacked kernel (one read each), those with haplotypes too long for its
    // LDS carry row on the striped kernel.
    KArgs a = base_args(p);
    a.tasks = p->d_fallback;
    a.n_tasks = 0;
    a.n_tasks_dev = ctr(p, 3);
    a.counter = ctr(p, 4);
    a.rescue = p->d_rescue;
    a.rescue_count = ctr(p, 2);
    a.carry_cols = kStackHmax + 128;
    HIP_TRY(launch_stack(a, p->grid_fbstack, p->stream));
    if (p->n_longhap_pairs > 0) {
      a.tasks = p->d_fallback_long;
      a.n_tasks_dev = ctr(p, 48);
      a.counter = ctr(p, 49);
      a.carry = p->d_carry32;
      a.carry_stride = p->stride32;
      HIP_TRY(launch_forward<float>(a, p->grid32, p->stream));
    }
  }
  if (timed) HIP_TRY(hipEventRecord(p->ev[4], p->stream));
  if (fp32 && p->has_band) {
    // Short pairs whose FP32 band sum fell below 1e-28: the FP64 moving band
    // (64 diagonals, same centre); the pairs it does not accept join the
    // FP64 full-matrix list.
    HIP_TRY(launch_band_rescue(band_args(p), p->d_rescue, ctr(p, 2), ctr(p, kBandRescueQueue),
                               ctr(p, kBandRescueAccepted), p->grid_rescue, p->stream));
  }
  {
    // Pass 3: FP64 over the pairs whose FP32 sum fell below 1e-28 (or all
    // pairs in the FP64-only modes).
    KArgs a = base_args(p);
    if (fp32) {
      a.tasks = p->d_rescue;
      a.n_tasks = 0;
      a.n_tasks_dev = ctr(p, 2);
    } else {
      a.tasks = p->d_tasks;
      a.n_tasks = p->n_tasks;
    }
    a.counter = ctr(p, 1);
    a.carry = p->d_carry64;
    a.carry_stride = p->stride64;
    HIP_TRY(launch_forward<double>(a, p->grid64, p->stream));
  }
  if (p->n_tasks64 > 0 && !long_side) {
    if (int rc = enqueue_long(p, timed, p->stream, p->d_carry32, p->d_carry64)) return rc;
  }
  if (long_side) HIP_TRY(hipStreamWaitEvent(p->stream, p->ev_join, 0));
  if (timed) HIP_TRY(hipEventRecord(p->ev[2], p->stream));
  return BV_OK;
}

}  // namespace

extern "C" {

const char* bv_pairhmm_last_error(void) { return g_err.c_str(); }

int bv_pairhmm_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  return ensure_init_locked(device);
}

void bv_pairhmm_shutdown(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx.ready) return;
  (void)hipFree(g_ctx.d_ph2pr_f);
  (void)hipFree(g_ctx.d_ph2pr_d);
  (void)hipFree(g_ctx.d_mm_f);
  (void)hipFree(g_ctx.d_mm_d);
  (void)hipFree(g_ctx.d_err_scalar);
  {
    std::lock_guard<std::mutex> sk(g_sw_mu);
    if (g_ctx.d_sw_scratch) (void)hipFree(g_ctx.d_sw_scratch);
    if (g_ctx.d_sw_queue) (void)hipFree(g_ctx.d_sw_queue);
  }
  g_ctx = DeviceCtx{};
}

int bv_pairhmm_device_info(bv_device_info* info) {
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = ensure_init_locked(-1);
  if (rc) return rc;
  if (!info) return fail(BV_EINVAL, "info is NULL");
  info->device = g_ctx.device;
  info->compute_units = g_ctx.cus;
  info->blocks_per_cu_f32 = g_ctx.bpc_f32;
  info->blocks_per_cu_f64 = g_ctx.bpc_f64;
  std::snprintf(info->arch, sizeof(info->arch), "%s", g_ctx.arch.c_str());
  return BV_OK;
}

// Host copies of the tables, for cross-checking against the reference's.
int bv_pairhmm_tables(float* ph2pr_f, double* ph2pr_d, float* mm_f, double* mm_d) {
  static HostTables t;
  static std::once_flag once;
  std::call_once(once, [] { build_tables(t); });
  if (ph2pr_f) std::memcpy(ph2pr_f, t.ph2pr_f, sizeof(t.ph2pr_f));
  if (ph2pr_d) std::memcpy(ph2pr_d, t.ph2pr_d, sizeof(t.ph2pr_d));
  if (mm_f) std::memcpy(mm_f, t.mm_f, sizeof(t.mm_f));
  if (mm_d) std::memcpy(mm_d, t.mm_d, sizeof(t.mm_d));
  return BV_OK;
}

void bv_options_default(bv_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->mode = BV_MODE_AVX;
  opt->fp64_renorm = 1;
  opt->band_tol = 5e-7f;
  opt->band_narrow = 1;
  opt->use_graph = 1;
}

int bv_plan_create(const bv_region_batch* b, const bv_options* optp, bv_plan** out) {
  if (!out) return fail(BV_EINVAL, "plan output pointer is NULL");
  *out = nullptr;
  bv_options opt;
  if (optp) opt = *optp; else bv_options_default(&opt);
  if (opt.mode != BV_MODE_AVX && opt.mode != BV_MODE_SCALAR) return fail(BV_EINVAL, "unknown mode");
  if (opt.mode == BV_MODE_SCALAR) {
    bool all_zero = true;
    for (double v : opt.trans) all_zero = all_zero && v == 0.0;
    if (all_zero) {  // PairHMM::ORIGINAL_DEFAULT (pairhmm.hpp:32-33)
      const double d[7] = {0.9998, 0.0001, 0.0001, 0.9, 0.1, 0.9, 0.1};
      std::memcpy(opt.trans, d, sizeof(d));
    }
  }
  int rc = validate(b);
  if (rc) return rc;
  Alphabet al;
  if (opt.mode == BV_MODE_AVX) {
    alphabet_avx(al);
  } else {
    const int64_t nb = b->n_reads ? b->read_off[b->n_reads] - b->read_off[0] : 0;
    rc = alphabet_scalar(al, b->read_bases + (b->n_reads ? b->read_off[0] : 0), nb);
    if (rc) return rc;
    for (int64_t i = 0; i < nb; ++i)
      if (b->read_qual[b->read_off[0] + i] >= 128)
        return fail(BV_EINVAL, "scalar mode: quality byte >= 128 (negative char index in the reference)");
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    rc = ensure_init_locked(-1);
    if (rc) return rc;
  }
  HIP_TRY(hipSetDevice(g_ctx.device));

  auto* p = new bv_plan();
  p->opt = opt;
  p->n_reads = b->n_reads;
  p->n_haps = b->n_haps;
  auto cleanup = [&](int code) {
    free_plan(p);
    return code;
  };
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(BV_EHIP, "hipStreamCreate failed"));
  for (auto& e : p->ev)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(BV_EHIP, "hipEventCreate failed"));
  if (hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess)
    return cleanup(fail(BV_EHIP, "hipEventCreate failed"));

  // Rebase read and hap byte ranges to start at zero.
  const int64_t rb0 = b->n_reads ? b->read_off[0] : 0;
  const int64_t rbytes = b->n_reads ? b->read_off[b->n_reads] - rb0 : 0;
  const int64_t hb0 = b->n_haps ? b->hap_off[0] : 0;
  const int64_t hbytes = b->n_haps ? b->hap_off[b->n_haps] - hb0 : 0;
  p->read_bytes = rbytes;
  p->hap_bytes = hbytes;
  std::vector<int64_t> roff(b->n_reads + 1), hoff(b->n_haps + 1);
  for (int64_t r = 0; r <= b->n_reads; ++r) roff[r] = b->n_reads ? b->read_off[r] - rb0 : 0;
  for (int64_t h = 0; h <= b->n_haps; ++h) hoff[h] = b->n_haps ? b->hap_off[h] - hb0 : 0;
  const auto t_host0 = std::chrono::steady_clock::now();
  // AVX mode: bases travel at two bits each (plus the 'N' positions) and are
  // expanded on the device; the scalar mode's classes (up to 8 raw bytes)
  // travel as class bytes.
  const bool packed_bases = opt.mode == BV_MODE_AVX;
  std::vector<uint8_t> rclass, hmask, rpack, hpack;
  std::vector<int64_t> rexc, hexc;
  if (packed_bases) {
    uint8_t code_of[256];
    for (int k = 0; k < 256; ++k) code_of[k] = al.read_class[k] & 3u;
    pack_bases(b->read_bases + rb0, rbytes, code_of, 'N', rpack, rexc);
    pack_bases(b->hap_bases + hb0, hbytes, code_of, 'N', hpack, hexc);
  } else {
    rclass.resize(static_cast<size_t>(rbytes));
    hmask.resize(static_cast<size_t>(hbytes));
    parallel_for(rbytes, [&](int64_t a0, int64_t a1) {
      for (int64_t i = a0; i < a1; ++i) rclass[i] = al.read_class[b->read_bases[rb0 + i]];
    });
    // the last base of every read carries kReadEnd (the stacked kernel's read boundaries)
    for (int64_t r = 0; r < b->n_reads; ++r)
      if (roff[r + 1] > roff[r]) rclass[roff[r + 1] - 1] |= kReadEnd;
    parallel_for(hbytes, [&](int64_t a0, int64_t a1) {
      for (int64_t i = a0; i < a1; ++i) hmask[i] = al.hap_mask[b->hap_bases[hb0 + i]];
    });
  }
  // One pass over the gap bytes.  Reads whose gap-continuation quality is the
  // same at every base and not 0 (the banded kernels' deletion scan needs a
  // constant decay along the read, and the short-read kernel carries I and D
  // times the gap-to-match factor 1 - ph2pr[q], which is 0 for q = 0).  And
  // whether the insertion, deletion and gap-continuation bytes are each
  // constant along every read (every SamRecord's GOP/GCP strings, sam.hpp:
  // 451-475): then they travel as one triplet per read and a kernel writes
  // the per-base bytes on the device.
  std::vector<uint8_t> gcp_const(static_cast<size_t>(std::max<int64_t>(1, b->n_reads)), 0);
  std::vector<uint8_t> triplets(static_cast<size_t>(3 * std::max<int64_t>(1, b->n_reads)), 0);
  std::atomic<bool> all_uniform{true};
  parallel_for(b->n_reads, [&](int64_t a0, int64_t a1) {
    bool ok = true;
    for (int64_t r = a0; r < a1; ++r) {
      const int64_t o = b->read_off[r], R = b->read_off[r + 1] - o;
      const uint8_t *i = b->read_ins + o, *d = b->read_del + o, *c = b->read_gcp + o;
      bool same_c = true, same_id = true;
      for (int64_t k = 1; k < R; ++k) {
        same_c = same_c && c[k] == c[0];
        same_id = same_id && i[k] == i[0] && d[k] == d[0];
      }
      // the banded kernels compare the 7-bit quality (the byte & 127)
      bool same_q = R == 0 || (c[0] & 127) != 0;
      for (int64_t k = 1; k < R && same_q && !same_c; ++k) same_q = (c[k] & 127) == (c[0] & 127);
      gcp_const[r] = same_q ? 1 : 0;
      ok = ok && same_c && same_id;
      triplets[3 * r] = i[0];
      triplets[3 * r + 1] = d[0];
      triplets[3 * r + 2] = c[0];
    }
    if (!ok) all_uniform.store(false, std::memory_order_relaxed);
  });
  const bool gaps_per_read = b->n_reads > 0 && all_uniform.load();
  // The short-read banded kernel's selector window: the longest read it may get.
  int rmax_band = 0;
  for (int64_t r = 0; r < b->n_reads; ++r) {
    const int64_t R = b->read_off[r + 1] - b->read_off[r];
    if (R <= kBandReadMax) rmax_band = std::max(rmax_band, static_cast<int>(R));
  }
  p->sel_cols = band_sel_cols(rmax_band);
  int hmax = 0;
  p->hap_len.resize(static_cast<size_t>(b->n_haps));
  for (int64_t h = 0; h < b->n_haps; ++h) {
    p->hap_len[h] = hoff[h + 1] - hoff[h];
    hmax = std::max<int>(hmax, static_cast<int>(p->hap_len[h]));
  }
  p->hmax = hmax;

  // Tasks: a read against its region's haplotypes (one task per haplotype
  // for reads of kSplitRows or more, so long reads spread over the GPU),
  // longest first.  In the FP32 mode reads of kFp64FirstRows or more go to
  // the FP64 list instead: they underflow FP32's range at these lengths, and
  // the few that would not are re-run in FP32 afterwards (verify pass).
  const bool fp32_mode = opt.mode == BV_MODE_AVX && !opt.force_fp64;
  // Built per chunk of consecutive regions on its own thread, then joined
  // in region order (the same lists a single pass would build).
  struct Built {
    std::vector<Task> tasks, tasks64, tasks64r, rest;
    std::vector<uint64_t> cost, cost64, cost_rest, cost_stack;
    std::vector<StackTask> stasks;
    int64_t cells = 0, longhap = 0;
    int stack_hmax = 0;
  };
  p->n_regions = b->n_regions;
  p->region_read_off.assign(b->region_read_off, b->region_read_off + (b->n_regions ? b->n_regions + 1 : 0));
  p->region_hap_off.assign(b->region_hap_off, b->region_hap_off + (b->n_regions ? b->n_regions + 1 : 0));
  p->region_pair_off.assign(static_cast<size_t>(b->n_regions ? b->n_regions + 1 : 0), 0);
  for (int64_t g = 0; g < b->n_regions; ++g)
    p->region_pair_off[g + 1] = p->region_pair_off[g] + (b->region_read_off[g + 1] - b->region_read_off[g]) *
                                                            (b->region_hap_off[g + 1] - b->region_hap_off[g]);
  // FP32 mode: runs of consecutive reads (shorter than kFp64FirstRows) of a
  // region whose haplotypes all fit the stacked kernel's LDS carry row become
  // stacked tasks of at most kStackRows rows.  A run is cut early once it
  // fills its last 63-row stripe to within 6 rows (lanes left empty at the
  // task's end are the kernel's only partial stripe).
  auto build_region = [&](int64_t g, Built& o) {
    const int64_t pair_base = p->region_pair_off[g];
    const int64_t r0 = b->region_read_off[g], r1 = b->region_read_off[g + 1];
    const int64_t h0 = b->region_hap_off[g], h1 = b->region_hap_off[g + 1];
    const int64_t nh = h1 - h0;
    int64_t hsum = 0, hmax_g = 0;
    for (int64_t h = h0; h < h1; ++h) {
      hsum += hoff[h + 1] - hoff[h];
      hmax_g = std::max<int64_t>(hmax_g, hoff[h + 1] - hoff[h]);
    }
    const bool stackable = fp32_mode && nh > 0 && hmax_g <= kStackHmax;
    if (stackable) {
      o.stack_hmax = std::max<int>(o.stack_hmax, static_cast<int>(hmax_g));
      int64_t run0 = -1, rows = 0;
      auto flush = [&](int64_t r_end) {
        if (run0 < 0) return;
        o.stasks.push_back(StackTask{static_cast<int32_t>(run0), static_cast<int32_t>(r_end - run0),
                                     static_cast<int32_t>(h0), static_cast<int32_t>(nh),
                                     pair_base + (run0 - r0) * nh, static_cast<int32_t>(nh),
                                     static_cast<int32_t>(rows)});
        o.cost_stack.push_back(static_cast<uint64_t>(rows * hsum));
        run0 = -1;
        rows = 0;
      };
      for (int64_t r = r0; r < r1; ++r) {
        const int64_t R = roff[r + 1] - roff[r];
        if (R >= kFp64FirstRows) {
          flush(r);
          continue;
        }
        if (run0 >= 0 && rows + R > kStackRows) flush(r);
        if (run0 < 0) run0 = r;
        rows += R;
        if (rows >= 8 * kRowsPerStripe && (kRowsPerStripe - rows % kRowsPerStripe) % kRowsPerStripe <= 6)
          flush(r + 1);
      }
      flush(r1);
    }
    int64_t longhaps = 0;
    for (int64_t h = h0; h < h1; ++h) longhaps += hoff[h + 1] - hoff[h] > kStackHmax;
    for (int64_t r = r0; r < r1; ++r) {
      if (nh == 0) continue;
      const int64_t R = roff[r + 1] - roff[r];
      const int64_t out0 = pair_base + (r - r0) * nh;
      o.cells += R * hsum;
      if (fp32_mode && R >= kFp64FirstRows) {
        o.tasks64r.push_back(Task{static_cast<int32_t>(r), static_cast<int32_t>(h0), static_cast<int32_t>(nh), 0, out0});
        for (int64_t h = h0; h < h1; ++h) {
          o.tasks64.push_back(Task{static_cast<int32_t>(r), static_cast<int32_t>(h), 1, 0, out0 + (h - h0)});
          o.cost64.push_back(static_cast<uint64_t>(R * (hoff[h + 1] - hoff[h])));
        }
      } else if (R >= kSplitRows) {
        o.longhap += longhaps;
        for (int64_t h = h0; h < h1; ++h) {
          o.tasks.push_back(Task{static_cast<int32_t>(r), static_cast<int32_t>(h), 1, 0, out0 + (h - h0)});
          o.cost.push_back(static_cast<uint64_t>(R * (hoff[h + 1] - hoff[h])));
          if (fp32_mode && !stackable) {
            o.rest.push_back(o.tasks.back());
            o.cost_rest.push_back(o.cost.back());
          }
        }
      } else {
        o.longhap += longhaps;
        o.tasks.push_back(Task{static_cast<int32_t>(r), static_cast<int32_t>(h0), static_cast<int32_t>(nh), 0, out0});
        o.cost.push_back(static_cast<uint64_t>(R * hsum));
        if (fp32_mode && !stackable) {
          o.rest.push_back(o.tasks.back());
          o.cost_rest.push_back(o.cost.back());
        }
      }
    }
  };
  // chunks of regions with about equal read counts
  const int64_t n_chunks = b->n_regions < 256 ? 1 : 64;
  std::vector<int64_t> cut(static_cast<size_t>(n_chunks) + 1, b->n_regions);
  cut[0] = 0;
  for (int64_t c = 1, g = 0; c < n_chunks; ++c) {
    const int64_t want = b->n_reads * c / n_chunks;
    while (g < b->n_regions && b->region_read_off[g] < want) ++g;
    cut[c] = std::max(cut[c - 1], g);
  }
  std::vector<Built> parts(static_cast<size_t>(n_chunks));
  parallel_for(n_chunks, [&](int64_t c0, int64_t c1) {
    for (int64_t c = c0; c < c1; ++c)
      for (int64_t g = cut[c]; g < cut[c + 1]; ++g) build_region(g, parts[c]);
  }, 1);
  // join in chunk order, each list's parts copied in parallel
  auto join = [&](auto member, auto& dst) {
    std::vector<int64_t> at(static_cast<size_t>(n_chunks) + 1, 0);
    for (int64_t c = 0; c < n_chunks; ++c) at[c + 1] = at[c] + static_cast<int64_t>((parts[c].*member).size());
    dst.resize(static_cast<size_t>(at[n_chunks]));
    parallel_for(n_chunks, [&](int64_t c0, int64_t c1) {
      for (int64_t c = c0; c < c1; ++c)
        std::copy((parts[c].*member).begin(), (parts[c].*member).end(), dst.begin() + at[c]);
    }, 1);
  };
  std::vector<Task> tasks, tasks64, tasks64r, tasks_rest;
  std::vector<uint64_t> cost, cost64, cost_rest, cost_stack;
  std::vector<StackTask> stasks;
  join(&Built::tasks, tasks);
  join(&Built::cost, cost);
  join(&Built::tasks64, tasks64);
  join(&Built::tasks64r, tasks64r);
  join(&Built::cost64, cost64);
  join(&Built::rest, tasks_rest);
  join(&Built::cost_rest, cost_rest);
  join(&Built::stasks, stasks);
  join(&Built::cost_stack, cost_stack);
  int stack_hmax = 0;
  int64_t pair_base = b->n_regions ? p->region_pair_off[b->n_regions] : 0, cells = 0;
  for (const Built& o : parts) {
    cells += o.cells;
    p->n_longhap_pairs += o.longhap;
    stack_hmax = std::max(stack_hmax, o.stack_hmax);
  }
  parts.clear();
  p->n_pairs = pair_base;
  p->cells = cells;
  if (tasks.size() >= (1u << 31) || tasks64.size() >= (1u << 31))
    return cleanup(fail(BV_EINVAL, "too many tasks"));
  p->n_stasks = static_cast<int>(stasks.size());
  p->n_tasks_rest = static_cast<int>(tasks_rest.size());
  p->stack_cols = stack_hmax + 128;  // lane 0 reads up to column H + 126
  p->n_tasks = static_cast<int>(tasks.size());
  p->n_subtasks = 0;
  for (const Task& t : tasks) p->n_subtasks += (t.nhap + 63) / 64;
  p->grid_prologue = band_prologue_grid(g_ctx.cus, p->n_tasks);
  p->n_tasks64 = static_cast<int>(tasks64.size());
  p->n_tasks64r = static_cast<int>(tasks64r.size());

  const int bpc32 = g_ctx.bpc_f32, bpc64 = g_ctx.bpc_f64;
  p->grid32 = opt.grid_f32 > 0 ? opt.grid_f32 : g_ctx.cus * bpc32;
  p->grid64 = opt.grid_f64 > 0 ? opt.grid_f64 : g_ctx.cus * bpc64;
  // the FP32 passes also run the full-matrix and verify lists (up to one task per pair)
  p->grid32 = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(p->grid32, std::max<int64_t>(1, p->n_pairs))));
  p->grid_band = band_grid_cap(p, p->wide_pairs);
  p->grid_band1 = band_grid_cap(p, 1);
  const int64_t max_fp64_tasks = (opt.mode == BV_MODE_AVX && !opt.force_fp64) ? p->n_pairs : p->n_tasks;
  p->grid64 = std::max(1, static_cast<int>(std::min<int64_t>(p->grid64, std::max<int64_t>(1, max_fp64_tasks))));
  p->grid32_max = p->grid32;
  p->grid64_max = p->grid64;
  p->grid_band_max = p->grid_band;
  p->grid_long = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(static_cast<int64_t>(g_ctx.cus) * g_ctx.bpc_long, p->n_tasks64)));
  p->grid_long_max = p->grid_long;
  p->grid_rescue = std::max(1, g_ctx.cus * g_ctx.bpc_long);
  p->grid_stack = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(static_cast<int64_t>(g_ctx.cus) * stack_blocks_per_cu(p->stack_cols),
                           std::max(1, p->n_stasks))));
  p->grid_stack_max = p->grid_stack;
  p->grid_fbstack = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(static_cast<int64_t>(g_ctx.cus) * stack_blocks_per_cu(kStackHmax + 128),
                           std::max<int64_t>(1, p->n_pairs))));
  p->stride32 = hmax + kCarryPad;
  p->stride64 = hmax + kCarryPad;

  hipStream_t st = p->stream;
  const auto t_up0 = std::chrono::steady_clock::now();
  int64_t* up = &p->setup_bytes;
  if ((rc = dev_upload(&p->d_rclass, packed_bases ? nullptr : rclass.data(), packed_bases ? 0 : rbytes, st, rbytes,
                       up)) ||
      (rc = dev_upload(&p->d_q, b->read_qual + rb0, rbytes, st, 0, up)) ||
      (rc = dev_upload(&p->d_ins, gaps_per_read ? nullptr : b->read_ins + rb0, gaps_per_read ? 0 : rbytes, st,
                       rbytes, up)) ||
      (rc = dev_upload(&p->d_del, gaps_per_read ? nullptr : b->read_del + rb0, gaps_per_read ? 0 : rbytes, st,
                       rbytes, up)) ||
      (rc = dev_upload(&p->d_gcp, gaps_per_read ? nullptr : b->read_gcp + rb0, gaps_per_read ? 0 : rbytes, st,
                       rbytes, up)) ||
      (rc = dev_upload(&p->d_gcp_const, gcp_const.data(), b->n_reads, st, 0, up)) ||
      (rc = dev_upload(&p->d_read_off, roff.data(), b->n_reads + 1, st, 0, up)) ||
      (rc = dev_upload(&p->d_hmask, packed_bases ? nullptr : hmask.data(), packed_bases ? 0 : hbytes, st, hbytes,
                       up)) ||
      (rc = dev_upload(&p->d_hap_off, hoff.data(), b->n_haps + 1, st, 0, up)) ||
      (rc = dev_upload(&p->d_tasks_region, tasks.data(), static_cast<int64_t>(tasks.size()), st, 0, up)) ||
      (rc = upload_longest_first(&p->d_tasks, p->d_tasks_region, tasks, cost, st, up)) ||
      (rc = upload_longest_first(&p->d_stasks, static_cast<const StackTask*>(nullptr), stasks, cost_stack, st, up)) ||
      (rc = upload_longest_first(&p->d_tasks_rest, static_cast<const Task*>(nullptr), tasks_rest, cost_rest, st,
                                 up)) ||
      (rc = upload_longest_first(&p->d_tasks64, static_cast<const Task*>(nullptr), tasks64, cost64, st, up)) ||
      (rc = dev_upload(&p->d_tasks64r, tasks64r.data(), static_cast<int64_t>(tasks64r.size()), st, 0, up)) ||
      (rc = dev_upload(&p->d_verify, static_cast<const Task*>(nullptr), 0, st,
                       std::max<int64_t>(1, static_cast<int64_t>(tasks64.size())))))
    return cleanup(rc);
  if (hipMalloc(&p->d_rescue, sizeof(Task) * std::max<int64_t>(1, p->n_pairs)) != hipSuccess ||
      hipMalloc(&p->d_counters, sizeof(int) * kCounters * kCtrStride) != hipSuccess ||
      hipMalloc(&p->d_fallback, sizeof(Task) * std::max<int64_t>(1, p->n_pairs)) != hipSuccess ||
      hipMalloc(&p->d_out, sizeof(double) * std::max<int64_t>(1, p->n_pairs)) != hipSuccess ||
      hipMalloc(&p->d_form, std::max<int64_t>(1, p->n_pairs)) != hipSuccess ||
      hipMalloc(&p->d_carry32, sizeof(float) * 2 * p->stride32 * p->grid32) != hipSuccess ||
      hipMalloc(&p->d_carry64, sizeof(double) * 2 * p->stride64 * p->grid64) != hipSuccess)
    return cleanup(fail(BV_ENOMEM, "device allocation failed"));
  if (packed_bases) {
    // two-bit bases and the 'N' positions, expanded to class bytes (with the
    // read-end flag) and haplotype match masks on the device
    uint8_t *d_rpack = nullptr, *d_hpack = nullptr;
    int64_t *d_rexc = nullptr, *d_hexc = nullptr;
    auto drop = [&] {
      for (void* q : {static_cast<void*>(d_rpack), static_cast<void*>(d_hpack), static_cast<void*>(d_rexc),
                      static_cast<void*>(d_hexc)})
        if (q) (void)hipFree(q);
    };
    if ((rc = dev_upload(&d_rpack, rpack.data(), static_cast<int64_t>(rpack.size()), st, 0, up)) ||
        (rc = dev_upload(&d_hpack, hpack.data(), static_cast<int64_t>(hpack.size()), st, 0, up)) ||
        (rc = dev_upload(&d_rexc, rexc.data(), static_cast<int64_t>(rexc.size()), st, 0, up)) ||
        (rc = dev_upload(&d_hexc, hexc.data(), static_cast<int64_t>(hexc.size()), st, 0, up))) {
      drop();
      return cleanup(rc);
    }
    const hipError_t e = launch_expand_bases(p->d_read_off, b->n_reads, d_rpack, d_rexc,
                                             static_cast<int64_t>(rexc.size()), p->d_rclass, hbytes, d_hpack, d_hexc,
                                             static_cast<int64_t>(hexc.size()), p->d_hmask, st);
    const hipError_t es = hipStreamSynchronize(st);
    drop();
    if (e != hipSuccess || es != hipSuccess) return cleanup(fail(BV_EHIP, "base expansion failed"));
  }
  if (gaps_per_read) {
    uint8_t* d_trip = nullptr;
    if ((rc = dev_upload(&d_trip, triplets.data(), 3 * b->n_reads, st, 0, up))) {
      if (d_trip) (void)hipFree(d_trip);
      return cleanup(rc);
    }
    const hipError_t e = launch_expand_read_bytes(p->d_read_off, b->n_reads, d_trip, p->d_ins, p->d_del,
                                                  p->d_gcp, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipFree(d_trip);
    if (e != hipSuccess || es != hipSuccess) return cleanup(fail(BV_EHIP, "read byte expansion failed"));
  }
  // The 64-bit rescue cell counter shares the zeroed counter block.
  p->d_rescue_cells = reinterpret_cast<unsigned long long*>(ctr(p, 6));
  if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(BV_EHIP, "upload failed"));
  const auto t_end = std::chrono::steady_clock::now();
  p->setup_ms_host = std::chrono::duration<float, std::milli>(t_up0 - t_host0).count();
  p->setup_ms_upload = std::chrono::duration<float, std::milli>(t_end - t_up0).count();
  *out = p;
  return BV_OK;
}

static void drop_graph(bv_plan* p) {
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
}

int bv_plan_autotune(bv_plan* p, int iters, int32_t* grids_out) {
  if (!p || iters <= 0) return fail(BV_EINVAL, "bad plan or iteration count");
  HIP_TRY(hipSetDevice(g_ctx.device));
  int* knobs[4] = {&p->grid_band, &p->grid32, &p->grid64, &p->grid_long};
  float best_ms = 0.f;
  int rc = bv_plan_time(p, iters, &best_ms);
  if (rc) return rc;
  // The wide band's shape (pairs per wavefront) first, each at its own
  // occupancy-derived grid; then the grids of the chosen shape.
  if (p->has_band) {
    const int cur = p->wide_pairs;
    for (int wp : {1, kQuadPairs}) {
      if (wp == p->wide_pairs) continue;
      const int keep = p->wide_pairs, keep_grid = p->grid_band;
      if ((rc = bv_plan_set_band_shape(p, wp))) return rc;
      float ms = 0.f;
      if ((rc = bv_plan_time(p, iters, &ms))) return rc;
      if (ms < best_ms) {
        best_ms = ms;
      } else {
        if ((rc = bv_plan_set_band_shape(p, keep))) return rc;
        p->grid_band = keep_grid;
      }
    }
    (void)cur;
  }
  const int caps[4] = {p->grid_band_max, p->grid32_max, p->grid64_max, p->grid_long_max};
  // Coordinate search: each persistent grid over fractions of its
  // occupancy-derived maximum, timing whole runs on the plan's own data.
  for (int k = 0; k < 4; ++k) {
    const int cur = *knobs[k];
    int best = cur;
    for (int num : {2, 4, 6, 8}) {
      const int g = std::max(g_ctx.cus, caps[k] * num / 8);
      if (g == cur || g > caps[k]) continue;
      *knobs[k] = g;
      float ms = 0.f;
      if ((rc = bv_plan_time(p, iters, &ms))) {
        *knobs[k] = cur;
        return rc;
      }
      if (ms < best_ms) {
        best_ms = ms;
        best = g;
      }
    }
    *knobs[k] = best;
  }
  drop_graph(p);
  if (grids_out) {
    grids_out[0] = p->grid_band;
    grids_out[1] = p->grid32;
    grids_out[2] = p->grid64;
    grids_out[3] = p->grid_long;
  }
  return BV_OK;
}

int bv_plan_set_grids(bv_plan* p, const int32_t* grids) {
  if (!p || !grids) return fail(BV_EINVAL, "NULL argument");
  int* knobs[4] = {&p->grid_band, &p->grid32, &p->grid64, &p->grid_long};
  const int caps[4] = {p->grid_band_max, p->grid32_max, p->grid64_max, p->grid_long_max};
  for (int k = 0; k < 4; ++k) {
    if (grids[k] < 0) return fail(BV_EINVAL, "negative grid");
    if (grids[k] > 0) *knobs[k] = std::min(grids[k], std::max(1, caps[k]));
  }
  drop_graph(p);
  return BV_OK;
}

int bv_plan_set_band_shape(bv_plan* p, int32_t wide_pairs) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  if (wide_pairs == 0) wide_pairs = kQuadPairs;
  if (wide_pairs != 1 && wide_pairs != kQuadPairs)
    return fail(BV_EINVAL, "the wide band takes 1 or 4 pairs per wavefront");
  HIP_TRY(hipSetDevice(g_ctx.device));
  if (wide_pairs != p->wide_pairs) {
    p->wide_pairs = wide_pairs;
    p->grid_band = p->grid_band_max = band_grid_cap(p, wide_pairs);
    drop_graph(p);
  }
  return BV_OK;
}

int bv_plan_set_band_mid(bv_plan* p, int32_t on) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  HIP_TRY(hipSetDevice(g_ctx.device));
  if ((on != 0) != p->mid_on) {
    p->mid_on = on != 0;
    drop_graph(p);
  }
  // the mid band's list, once the plan has bands (else when they are set)
  if (p->mid_on && p->d_items)
    if (int rc = ensure_band_lists(p)) return rc;
  return BV_OK;
}

int bv_plan_set_band_prologue(bv_plan* p, int32_t mode) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  if (mode < 0 || mode > 2) return fail(BV_EINVAL, "prologue mode is 0, 1 or 2");
  HIP_TRY(hipSetDevice(g_ctx.device));
  if (mode != p->prologue_mode) {
    p->prologue_mode = mode;
    drop_graph(p);
  }
  // the buffers the mode needs, once a band model is set (not under capture)
  if (p->has_model)
    if (int rc = ensure_feature_buffers(p)) return rc;
  return BV_OK;
}

int bv_plan_set_long_precision(bv_plan* p, int32_t bits) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  if (bits != 32 && bits != 64) return fail(BV_EINVAL, "the moving band runs in 32 or 64 bits");
  if ((bits == 32) != p->long_fp32) {
    p->long_fp32 = bits == 32;
    drop_graph(p);
  }
  return BV_OK;
}

int bv_plan_long_precision(bv_plan* p, int32_t* bits) {
  if (!p || !bits) return fail(BV_EINVAL, "NULL argument");
  *bits = p->long_fp32 ? 32 : 64;
  return BV_OK;
}

int bv_plan_set_quad_fill(bv_plan* p, int32_t min_pairs) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  if (min_pairs < 1 || min_pairs > kQuadPairs) return fail(BV_EINVAL, "quad fill is 1 to 4 pairs");
  HIP_TRY(hipSetDevice(g_ctx.device));
  if (min_pairs != p->quad_fill) {
    p->quad_fill = min_pairs;
    drop_graph(p);
  }
  return BV_OK;
}

int bv_plan_quad_fill(bv_plan* p, int32_t* min_pairs) {
  if (!p || !min_pairs) return fail(BV_EINVAL, "NULL argument");
  *min_pairs = p->quad_fill;
  return BV_OK;
}

int bv_plan_band_mid(bv_plan* p, int32_t* on) {
  if (!p || !on) return fail(BV_EINVAL, "NULL argument");
  *on = p->mid_on ? 1 : 0;
  return BV_OK;
}

int bv_plan_band_shape(bv_plan* p, int32_t* wide_pairs) {
  if (!p || !wide_pairs) return fail(BV_EINVAL, "NULL argument");
  *wide_pairs = p->wide_pairs;
  return BV_OK;
}

int bv_plan_run(bv_plan* p) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  HIP_TRY(hipSetDevice(g_ctx.device));
  p->timed = false;
  if (!p->opt.use_graph) return enqueue_run(p, false);
  if (!p->graph_exec) {
    // Capture one run: only kernel launches (counters are cleared by a
    // kernel, task counts are read on the device), so it replays as is.
    HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_run(p, false);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(p->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIP_TRY(ec);
    p->graph = g;
    HIP_TRY(hipGraphInstantiate(&p->graph_exec, p->graph, nullptr, nullptr, 0));
  }
  HIP_TRY(hipGraphLaunch(p->graph_exec, p->stream));
  return BV_OK;
}

int bv_plan_time(bv_plan* p, int iters, float* ms_per_run) {
  if (!p || iters <= 0) return fail(BV_EINVAL, "bad plan or iteration count");
  HIP_TRY(hipSetDevice(g_ctx.device));
  float sum32 = 0.f, sum64 = 0.f, sumfb = 0.f, sumband = 0.f, sumlong = 0.f;
  const bool banded = (p->opt.mode == BV_MODE_AVX) && !p->opt.force_fp64 && p->has_band;
  const bool long_banded = banded && p->n_tasks64 > 0;
  HIP_TRY(hipEventRecord(p->ev[3], p->stream));
  hipEvent_t start = p->ev[3];
  // Per-run split between the FP32 and FP64 passes from events between them.
  for (int i = 0; i < iters; ++i) {
    int rc = enqueue_run(p, true);
    if (rc) return rc;
    HIP_TRY(hipEventSynchronize(p->ev[2]));
    float a = 0.f, fb = 0.f, c = 0.f;
    HIP_TRY(hipEventElapsedTime(&a, p->ev[0], p->ev[1]));
    HIP_TRY(hipEventElapsedTime(&fb, p->ev[1], p->ev[4]));
    HIP_TRY(hipEventElapsedTime(&c, p->ev[4], p->ev[2]));
    sum32 += a;
    sumfb += fb;
    sum64 += c;
    if (banded) {
      float bk = 0.f;
      HIP_TRY(hipEventElapsedTime(&bk, p->ev[5], p->ev[6]));
      sumband += bk;
    }
    if (long_banded) {
      float lk = 0.f;
      HIP_TRY(hipEventElapsedTime(&lk, p->ev[7], p->ev[8]));
      sumlong += lk;
    }
  }
  float total = 0.f;
  HIP_TRY(hipEventElapsedTime(&total, start, p->ev[2]));
  p->last_ms32 = sum32 / iters;
  p->last_ms64 = sum64 / iters;
  p->last_msfb = sumfb / iters;
  p->last_msband = sumband / iters;
  p->last_mslong = sumlong / iters;
  if (ms_per_run) *ms_per_run = total / iters;
  return BV_OK;
}

int bv_plan_results(bv_plan* p, double* log10_out, uint8_t* form_out) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  HIP_TRY(hipSetDevice(g_ctx.device));
  if (log10_out && p->n_pairs)
    HIP_TRY(hipMemcpyAsync(log10_out, p->d_out, sizeof(double) * p->n_pairs, hipMemcpyDeviceToHost, p->stream));
  if (form_out && p->n_pairs)
    HIP_TRY(hipMemcpyAsync(form_out, p->d_form, p->n_pairs, hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return BV_OK;
}

int bv_plan_stats_get(bv_plan* p, bv_plan_stats* s) {
  if (!p || !s) return fail(BV_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(g_ctx.device));
  HIP_TRY(hipStreamSynchronize(p->stream));
  std::vector<int> line(kCounters * kCtrStride);
  HIP_TRY(hipMemcpy(line.data(), p->d_counters, sizeof(int) * line.size(), hipMemcpyDeviceToHost));
  int cnt[kCounters];
  for (int i = 0; i < kCounters; ++i) cnt[i] = line[i * kCtrStride];
  unsigned long long c64 = 0, swept = 0;
  std::memcpy(&c64, line.data() + 22 * kCtrStride, sizeof(c64));
  std::memcpy(&swept, line.data() + 8 * kCtrStride, sizeof(swept));
  const bool fp32 = (p->opt.mode == BV_MODE_AVX) && !p->opt.force_fp64;
  s->n_pairs = p->n_pairs;
  s->cells = p->cells;
  const bool long_banded = fp32 && p->has_band && p->n_tasks64 > 0;
  // FP64 pairs: rescued ones and every long read (full matrix, or the moving
  // band when it runs in FP64)
  s->fp64_pairs =
      fp32 ? cnt[2] + cnt[kBandRescueAccepted] + p->n_tasks64 - (long_banded && p->long_fp32 ? cnt[21] : 0) : p->n_pairs;
  s->fp64_cells = static_cast<int64_t>(c64);
  s->tasks = p->n_tasks;
  s->grid_f32 = p->grid32;
  s->grid_f64 = p->grid64;
  s->last_ms_f32 = p->last_ms32;
  s->last_ms_f64 = p->last_ms64;
  s->band_pairs = p->has_band ? cnt[5] + cnt[21] + cnt[kBandRescueAccepted] : 0;
  s->fallback_pairs = p->has_band ? cnt[3] + cnt[48] : 0;
  s->cells_swept = static_cast<int64_t>(swept);
  s->last_ms_fallback = p->last_msfb;
  s->last_ms_band = p->last_msband;
  s->last_ms_long = p->last_mslong;
  s->long_band_pairs = long_banded ? cnt[21] : 0;
  s->last_ms_post = p->last_mspost;
  s->setup_ms_host = p->setup_ms_host;
  s->setup_ms_upload = p->setup_ms_upload;
  s->setup_bytes = p->setup_bytes;
  return BV_OK;
}

int bv_plan_postprocess(bv_plan* p, int32_t flags, const int32_t* hap_allele, const int32_t* n_alleles,
                        uint8_t* keep_out, double* normalized_out, double* allele_out) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  if ((flags & ~BV_POST_NORMALIZE) != 0) return fail(BV_EINVAL, "unknown post-processing flags");
  if ((hap_allele == nullptr) != (n_alleles == nullptr))
    return fail(BV_EINVAL, "hap_allele and n_alleles are given together");
  if (allele_out && !hap_allele) return fail(BV_EINVAL, "allele_out needs hap_allele and n_alleles");
  HIP_TRY(hipSetDevice(g_ctx.device));
  const int64_t G = p->n_regions;
  if (G == 0 || p->n_pairs == 0) {
    // no pair to decide on: every read gets 0, as reads outside every region do
    if (keep_out && p->n_reads > 0) std::memset(keep_out, 0, static_cast<size_t>(p->n_reads));
    return BV_OK;
  }
  // the haplotype -> allele map must name an allele of the haplotype's region
  // (genotyper.hpp:200-209 maps every haplotype; unmapped ones go to allele 0)
  std::vector<int64_t> aoff;
  if (hap_allele) {
    aoff.assign(static_cast<size_t>(G + 1), 0);
    for (int64_t g = 0; g < G; ++g) {
      const int32_t na = n_alleles[g];
      if (na < 0) return fail(BV_EINVAL, "region " + std::to_string(g) + " has a negative allele count");
      for (int64_t h = p->region_hap_off[g]; h < p->region_hap_off[g + 1]; ++h)
        if (hap_allele[h] < 0 || hap_allele[h] >= na)
          return fail(BV_EINVAL, "haplotype " + std::to_string(h) + " maps to allele " +
                                     std::to_string(hap_allele[h]) + ", region " + std::to_string(g) +
                                     " has " + std::to_string(na));
      aoff[g + 1] = aoff[g] + (p->region_read_off[g + 1] - p->region_read_off[g]) * na;
    }
  }
  hipStream_t st = p->stream;
  if (!p->d_region_tab) {
    const size_t n1 = static_cast<size_t>(G + 1);
    std::vector<int64_t> tab(3 * n1);
    std::copy(p->region_read_off.begin(), p->region_read_off.end(), tab.begin());
    std::copy(p->region_hap_off.begin(), p->region_hap_off.end(), tab.begin() + n1);
    std::copy(p->region_pair_off.begin(), p->region_pair_off.end(), tab.begin() + 2 * n1);
    if (int rc = dev_upload(&p->d_region_tab, tab.data(), static_cast<int64_t>(tab.size()), st)) return rc;
    HIP_TRY(hipStreamSynchronize(st));
  }
  if (!p->ev_post[0]) {
    HIP_TRY(hipEventCreate(&p->ev_post[0]));
    HIP_TRY(hipEventCreate(&p->ev_post[1]));
  }
  if (keep_out && !p->d_keep) HIP_TRY(hipMalloc(&p->d_keep, std::max<int64_t>(1, p->n_reads)));
  if (normalized_out && !p->d_lnorm) HIP_TRY(hipMalloc(&p->d_lnorm, sizeof(double) * p->n_pairs));
  if (hap_allele) {
    if (!p->d_hap_allele) {
      HIP_TRY(hipMalloc(&p->d_hap_allele, sizeof(int32_t) * std::max<int64_t>(1, p->n_haps)));
      HIP_TRY(hipMalloc(&p->d_n_alleles, sizeof(int32_t) * G));
      HIP_TRY(hipMalloc(&p->d_allele_off, sizeof(int64_t) * (G + 1)));
    }
    if (allele_out && aoff[G] > p->allele_cap) {
      if (p->d_allele) HIP_TRY(hipFree(p->d_allele));
      p->d_allele = nullptr;
      p->allele_cap = 0;
      HIP_TRY(hipMalloc(&p->d_allele, sizeof(double) * aoff[G]));
      p->allele_cap = aoff[G];
    }
    if (p->n_haps)
      HIP_TRY(hipMemcpyAsync(p->d_hap_allele, hap_allele, sizeof(int32_t) * p->n_haps, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p->d_n_alleles, n_alleles, sizeof(int32_t) * G, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p->d_allele_off, aoff.data(), sizeof(int64_t) * (G + 1), hipMemcpyHostToDevice, st));
  }
  PostArgs a{};
  a.L = p->d_out;
  a.read0 = p->region_read_off[0];
  a.read_end = p->region_read_off[G];
  a.n_regions = G;
  a.region_read_off = p->d_region_tab;
  a.region_hap_off = p->d_region_tab + (G + 1);
  a.region_pair_off = p->d_region_tab + 2 * (G + 1);
  a.read_off = p->d_read_off;
  a.normalize = (flags & BV_POST_NORMALIZE) ? 1 : 0;
  a.keep = keep_out ? p->d_keep : nullptr;
  a.Lnorm = normalized_out ? p->d_lnorm : nullptr;
  a.hap_allele = p->d_hap_allele;
  a.n_alleles = p->d_n_alleles;
  a.allele_off = p->d_allele_off;
  a.A = (allele_out && aoff[G] > 0) ? p->d_allele : nullptr;
  if (keep_out) {
    // reads outside every region get no decision: 0
    HIP_TRY(hipMemsetAsync(p->d_keep, 0, std::max<int64_t>(1, p->n_reads), st));
  }
  HIP_TRY(hipEventRecord(p->ev_post[0], st));
  HIP_TRY(launch_post(a, st));
  HIP_TRY(hipEventRecord(p->ev_post[1], st));
  if (keep_out && p->n_reads)
    HIP_TRY(hipMemcpyAsync(keep_out, p->d_keep, p->n_reads, hipMemcpyDeviceToHost, st));
  if (normalized_out)
    HIP_TRY(hipMemcpyAsync(normalized_out, p->d_lnorm, sizeof(double) * p->n_pairs, hipMemcpyDeviceToHost, st));
  if (allele_out && aoff[G] > 0)
    HIP_TRY(hipMemcpyAsync(allele_out, p->d_allele, sizeof(double) * aoff[G], hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipEventElapsedTime(&p->last_mspost, p->ev_post[0], p->ev_post[1]));
  return BV_OK;
}

int bv_plan_set_bands(bv_plan* p, const int32_t* bands) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  HIP_TRY(hipSetDevice(g_ctx.device));
  drop_graph(p);
  if (!bands) {
    p->has_band = false;
    return BV_OK;
  }
  const int64_t n = std::max<int64_t>(1, p->n_pairs);
  if (!p->d_band) HIP_TRY(hipMalloc(&p->d_band, sizeof(int2) * n));
  if (int rc = ensure_band_lists(p)) return rc;
  HIP_TRY(hipMemcpyAsync(p->d_band, bands, sizeof(int2) * p->n_pairs, hipMemcpyHostToDevice, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->has_model = false;
  p->has_band = true;
  return BV_OK;
}

void bv_plan_destroy(bv_plan* p) {
  if (!p) return;
  (void)hipSetDevice(g_ctx.device);
  (void)hipStreamSynchronize(p->stream);
  free_plan(p);
}

int bv_pairhmm_batch(const bv_region_batch* b, const bv_options* opt, double* log10_out,
                     uint8_t* form_out) {
  bv_plan* p = nullptr;
  int rc = bv_plan_create(b, opt, &p);
  if (rc) return rc;
  rc = bv_plan_run(p);
  if (!rc) rc = bv_plan_results(p, log10_out, form_out);
  bv_plan_destroy(p);
  return rc;
}

void bv_sw_release(void) {
  std::lock_guard<std::mutex> sk(g_sw_mu);
  if (g_ctx.d_sw_scratch) (void)hipFree(g_ctx.d_sw_scratch);
  if (g_ctx.d_sw_queue) (void)hipFree(g_ctx.d_sw_queue);
  g_ctx.d_sw_scratch = nullptr;
  g_ctx.d_sw_queue = nullptr;
  g_ctx.sw_scratch_bytes = 0;
}

int bv_sw_align(const bv_sw_batch* b, const int32_t* params, int32_t* offset_out, int32_t* n_elems_out,
                char* ops_out, int32_t* lens_out, float* kernel_ms) {
  if (!b || !params || !offset_out || !n_elems_out || !ops_out || !lens_out)
    return fail(BV_EINVAL, "NULL argument");
  if (b->n_pairs < 0) return fail(BV_EINVAL, "negative pair count");
  if (b->n_pairs == 0) return BV_OK;
  if (!b->ref_off || !b->ref || !b->alt_off || !b->alt) return fail(BV_EINVAL, "sequence arrays are NULL");
  for (int k = 0; k < 4; ++k)
    if (params[k] > kSwMaxWeight || params[k] < -kSwMaxWeight)
      return fail(BV_EINVAL, "Smith-Waterman weight beyond +-65536");
  const int64_t n = b->n_pairs;
  if (n >= (int64_t(1) << 31)) return fail(BV_EINVAL, "more than 2^31 pairs");
  std::vector<int64_t> out_off(static_cast<size_t>(n) + 1, 0);
  for (int64_t k = 0; k < n; ++k) {
    const int64_t R = b->ref_off[k + 1] - b->ref_off[k], H = b->alt_off[k + 1] - b->alt_off[k];
    // SmithWaterman::align asserts non-empty sequences (smithwaterman.hpp:236)
    if (R < 1 || H < 1) return fail(BV_EINVAL, "pair " + std::to_string(k) + " has an empty sequence");
    if (R > kSwMaxLen || H > kSwMaxLen)
      return fail(BV_EINVAL, "pair " + std::to_string(k) + " is longer than 4096 bases");
    out_off[k + 1] = out_off[k] + R + H + 3;
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = ensure_init_locked(-1)) return rc;
  }
  HIP_TRY(hipSetDevice(g_ctx.device));
  const int64_t rb0 = b->ref_off[0], ab0 = b->alt_off[0];
  const int64_t rbytes = b->ref_off[n] - rb0, abytes = b->alt_off[n] - ab0;
  std::vector<int64_t> roff(static_cast<size_t>(n) + 1), aoff(static_cast<size_t>(n) + 1);
  for (int64_t k = 0; k <= n; ++k) {
    roff[k] = b->ref_off[k] - rb0;
    aoff[k] = b->alt_off[k] - ab0;
  }
  // Size classes (longest side up to 512, 1024, 2048, 4096 bases), one
  // launch each: a launch's persistent grid is at most 32 wavefronts per CU
  // (lane 0's carry loads and the trace stores need other wavefronts to hide
  // behind) and at most 16 GiB of per-wavefront scratch (trace matrix, carry
  // row, last row and column) sized for the class's longest pair.
  constexpr int kClasses = 4;
  std::vector<int32_t> order;
  order.reserve(static_cast<size_t>(n));
  int64_t cls_begin[kClasses + 1] = {0};
  int cls_r[kClasses] = {0}, cls_h[kClasses] = {0};
  for (int c = 0; c < kClasses; ++c) {
    const int lim = 512 << c;
    for (int64_t k = 0; k < n; ++k) {
      const int R = static_cast<int>(b->ref_off[k + 1] - b->ref_off[k]);
      const int H = static_cast<int>(b->alt_off[k + 1] - b->alt_off[k]);
      const int m = std::max(R, H);
      if (m <= lim && (c == 0 || m > (256 << c))) {
        order.push_back(static_cast<int32_t>(k));
        cls_r[c] = std::max(cls_r[c], R);
        cls_h[c] = std::max(cls_h[c], H);
      }
    }
    cls_begin[c + 1] = static_cast<int64_t>(order.size());
  }
  const int64_t budget = int64_t(16) << 30;
  int64_t scratch_bytes = 0;
  int cls_grid[kClasses] = {0};
  int64_t cls_stride[kClasses] = {0};
  for (int c = 0; c < kClasses; ++c) {
    const int64_t nc = cls_begin[c + 1] - cls_begin[c];
    if (nc == 0) continue;
    cls_stride[c] = sw_scratch_bytes(cls_r[c], cls_h[c]);
    cls_grid[c] = static_cast<int>(std::max<int64_t>(
        1, std::min<int64_t>({nc, static_cast<int64_t>(g_ctx.cus) * 32, budget / cls_stride[c]})));
    scratch_bytes = std::max(scratch_bytes, cls_stride[c] * cls_grid[c]);
  }
  // The per-wavefront scratch and the queues live in the library's
  // workspace, grown to this call's need; calls take turns on it.
  std::unique_lock<std::mutex> sw_lock(g_sw_mu);
  if (g_ctx.sw_scratch_bytes < scratch_bytes) {
    if (g_ctx.d_sw_scratch) (void)hipFree(g_ctx.d_sw_scratch);
    g_ctx.d_sw_scratch = nullptr;
    g_ctx.sw_scratch_bytes = 0;
    if (hipMalloc(&g_ctx.d_sw_scratch, static_cast<size_t>(scratch_bytes)) != hipSuccess) {
      g_ctx.d_sw_scratch = nullptr;
      return fail(BV_ENOMEM, "Smith-Waterman scratch allocation failed");
    }
    g_ctx.sw_scratch_bytes = scratch_bytes;
  }
  if (!g_ctx.d_sw_queue && hipMalloc(&g_ctx.d_sw_queue, sizeof(int) * kClasses) != hipSuccess) {
    g_ctx.d_sw_queue = nullptr;
    return fail(BV_ENOMEM, "Smith-Waterman queue allocation failed");
  }
  char* const d_scr = g_ctx.d_sw_scratch;
  int* const d_queue = g_ctx.d_sw_queue;
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  int64_t *d_roff = nullptr, *d_aoff = nullptr, *d_out = nullptr;
  uint8_t *d_ref = nullptr, *d_alt = nullptr;
  int32_t *d_offset = nullptr, *d_n = nullptr, *d_lens = nullptr, *d_order = nullptr;
  char* d_ops = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&](int code) {
    void* bufs[] = {d_roff, d_aoff, d_out, d_ref, d_alt, d_offset, d_n, d_lens, d_ops, d_order};
    for (void* x : bufs)
      if (x) (void)hipFree(x);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    return code;
  };
  int rc = 0;
  const int64_t slots = out_off[n];
  if ((rc = dev_upload(&d_roff, roff.data(), n + 1, st)) || (rc = dev_upload(&d_aoff, aoff.data(), n + 1, st)) ||
      (rc = dev_upload(&d_out, out_off.data(), n + 1, st)) ||
      (rc = dev_upload(&d_ref, b->ref + rb0, rbytes, st)) || (rc = dev_upload(&d_alt, b->alt + ab0, abytes, st)) ||
      (rc = dev_upload(&d_order, order.data(), n, st)))
    return cleanup(rc);
  if (hipMalloc(&d_offset, sizeof(int32_t) * n) != hipSuccess || hipMalloc(&d_n, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc(&d_lens, sizeof(int32_t) * slots) != hipSuccess || hipMalloc(&d_ops, slots) != hipSuccess)
    return cleanup(fail(BV_ENOMEM, "Smith-Waterman device allocation failed"));
  if (hipMemsetAsync(d_queue, 0, sizeof(int) * kClasses, st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    return cleanup(fail(BV_EHIP, "Smith-Waterman set-up failed"));
  SwArgs a{};
  a.n_pairs = n;
  a.ref_off = d_roff;
  a.ref = d_ref;
  a.alt_off = d_aoff;
  a.alt = d_alt;
  a.wm = params[0];
  a.wx = params[1];
  a.wo = params[2];
  a.we = params[3];
  a.scratch = d_scr;
  a.out_off = d_out;
  a.offset = d_offset;
  a.n_elems = d_n;
  a.ops = d_ops;
  a.lens = d_lens;
  if (hipEventRecord(e0, st) != hipSuccess) return cleanup(fail(BV_EHIP, "Smith-Waterman launch failed"));
  for (int c = 0; c < kClasses; ++c) {
    // the classes run one after another on the stream and share the scratch
    const int64_t nc = cls_begin[c + 1] - cls_begin[c];
    if (nc == 0) continue;
    a.n_pairs = nc;
    a.order = d_order + cls_begin[c];
    a.queue = d_queue + c;
    a.scratch_stride = cls_stride[c];
    if (launch_sw(a, cls_grid[c], st) != hipSuccess) return cleanup(fail(BV_EHIP, "Smith-Waterman launch failed"));
  }
  if (hipEventRecord(e1, st) != hipSuccess) return cleanup(fail(BV_EHIP, "Smith-Waterman launch failed"));
  if (hipMemcpyAsync(offset_out, d_offset, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(n_elems_out, d_n, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(lens_out, d_lens, sizeof(int32_t) * slots, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(ops_out, d_ops, slots, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return cleanup(fail(BV_EHIP, "Smith-Waterman run failed"));
  if (kernel_ms && hipEventElapsedTime(kernel_ms, e0, e1) != hipSuccess)
    return cleanup(fail(BV_EHIP, "Smith-Waterman timing failed"));
  return cleanup(BV_OK);
}

int bv_compute_full_prob(const bv_testcase* tcs, int n, int precision, double* out_scaled) {
  if (n < 0 || (n > 0 && (!tcs || !out_scaled))) return fail(BV_EINVAL, "bad arguments");
  if (precision != 32 && precision != 64) return fail(BV_EINVAL, "precision must be 32 or 64");
  if (n == 0) return BV_OK;
  // Each testcase becomes a one-read, one-haplotype region.
  std::vector<int64_t> rro(n + 1), rho(n + 1), ro(n + 1), ho(n + 1);
  std::vector<uint8_t> rs, q, ins, del, gcp, hap;
  for (int k = 0; k < n; ++k) {
    const bv_testcase& t = tcs[k];
    if (t.rslen <= 0 || t.haplen <= 0) return fail(BV_EINVAL, "testcase " + std::to_string(k) + " has an empty read or haplotype");
    rro[k] = k;
    rho[k] = k;
    ro[k] = static_cast<int64_t>(rs.size());
    ho[k] = static_cast<int64_t>(hap.size());
    rs.insert(rs.end(), t.rs, t.rs + t.rslen);
    q.insert(q.end(), t.q, t.q + t.rslen);
    ins.insert(ins.end(), t.i, t.i + t.rslen);
    del.insert(del.end(), t.d, t.d + t.rslen);
    gcp.insert(gcp.end(), t.c, t.c + t.rslen);
    hap.insert(hap.end(), t.hap, t.hap + t.haplen);
  }
  rro[n] = n;
  rho[n] = n;
  ro[n] = static_cast<int64_t>(rs.size());
  ho[n] = static_cast<int64_t>(hap.size());
  bv_region_batch b{n, rro.data(), rho.data(), n, ro.data(), rs.data(), q.data(), ins.data(),
                    del.data(), gcp.data(), n, ho.data(), hap.data()};
  bv_options opt;
  bv_options_default(&opt);
  opt.fp64_renorm = 0;
  opt.force_fp64 = precision == 64;
  bv_plan* p = nullptr;
  int rc = bv_plan_create(&b, &opt, &p);
  if (rc) return rc;
  auto run = [&]() -> int {
    HIP_TRY(launch_zero_ints(p->d_counters, 12 * kCtrStride, p->stream));
    KArgs a = base_args(p);
    a.tasks = p->d_tasks;
    a.n_tasks = p->n_tasks;
    a.counter = ctr(p, 0);
    a.out_raw = 1;
    if (precision == 32) {
      a.carry = p->d_carry32;
      a.carry_stride = p->stride32;
      HIP_TRY(launch_forward<float>(a, p->grid32, p->stream));
    } else {
      a.carry = p->d_carry64;
      a.carry_stride = p->stride64;
      HIP_TRY(launch_forward<double>(a, p->grid64, p->stream));
    }
    return BV_OK;
  };
  rc = run();
  if (!rc) rc = bv_plan_results(p, out_scaled, nullptr);
  bv_plan_destroy(p);
  return rc;
}

static_assert(BV_NOALLOC_SCRATCH_BYTES == kNoallocScratchBytes, "header and kernel scratch sizes differ");

int bv_compute_full_prob_noalloc(const bv_device_pairs* pr, int precision, int max_hap_len,
                                 double* out_scaled, void* stream) {
  if (!pr || !out_scaled) return fail(BV_EINVAL, "NULL argument");
  if (precision != 32 && precision != 64) return fail(BV_EINVAL, "precision must be 32 or 64");
  if (pr->n_pairs < 0 || pr->n_pairs >= (int64_t(1) << 31)) return fail(BV_EINVAL, "pair count out of range");
  if (pr->n_pairs == 0) return BV_OK;
  if (!pr->read_off || !pr->read_bases || !pr->read_qual || !pr->read_ins || !pr->read_del ||
      !pr->read_gcp || !pr->hap_off || !pr->hap_bases)
    return fail(BV_EINVAL, "device arrays are NULL");
  if (max_hap_len < 1) return fail(BV_EINVAL, "max_hap_len must be at least 1");
  struct {
    int device, cus, bpc_f32, bpc_f64;
    void* scratch;
    Tables tab;
  } ctx{};
  {
    // the set-up (tables, the scratch address) allocates, so it is not done here
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx.ready) return fail(BV_ESTATE, "call bv_pairhmm_init before bv_compute_full_prob_noalloc");
    ctx.device = g_ctx.device;
    ctx.cus = g_ctx.cus;
    ctx.bpc_f32 = g_ctx.bpc_f32;
    ctx.bpc_f64 = g_ctx.bpc_f64;
    ctx.scratch = g_ctx.d_noalloc_scratch;
    ctx.tab.ph2pr_f = g_ctx.d_ph2pr_f;
    ctx.tab.ph2pr_d = g_ctx.d_ph2pr_d;
    ctx.tab.mm_f = g_ctx.d_mm_f;
    ctx.tab.mm_d = g_ctx.d_mm_d;
    ctx.tab.err_scalar = g_ctx.d_err_scalar;
  }
  const int64_t entry = precision == 32 ? 2 * sizeof(float) : 2 * sizeof(double);
  const int64_t stride = static_cast<int64_t>(max_hap_len) + kCarryPad;
  const int64_t waves = (kNoallocScratchBytes - kNoallocHeader) / (stride * entry);
  if (waves < 1)
    return fail(BV_EINVAL, "max_hap_len " + std::to_string(max_hap_len) + " needs a carry row beyond the " +
                               std::to_string(kNoallocScratchBytes >> 20) + " MiB scratch");
  const int bpc = precision == 32 ? ctx.bpc_f32 : ctx.bpc_f64;
  const int grid = static_cast<int>(
      std::max<int64_t>(1, std::min<int64_t>({pr->n_pairs, static_cast<int64_t>(ctx.cus) * bpc, waves})));
  HIP_TRY(hipSetDevice(ctx.device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(ctx.scratch);
  int* counter = reinterpret_cast<int*>(base);
  // the queue counter keeps the last call's value: cleared by a kernel (no memset node under capture)
  HIP_TRY(launch_zero_ints(counter, 1, st));
  KArgs a{};
  a.q = pr->read_qual;
  a.ins = pr->read_ins;
  a.del = pr->read_del;
  a.gcp = pr->read_gcp;
  a.read_off = pr->read_off;
  a.hap_off = pr->hap_off;
  a.rbases = pr->read_bases;
  a.hbases = pr->hap_bases;
  a.identity_tasks = 1;
  a.n_tasks = static_cast<int>(pr->n_pairs);
  a.counter = counter;
  a.out = out_scaled;
  a.carry = base + kNoallocHeader;
  a.carry_stride = stride;
  a.hcap = max_hap_len;
  a.tab = ctx.tab;
  a.mode = BV_MODE_AVX;
  a.renorm = 0;  // the reference's fixed 2^120 / 2^1020 scaling, as bv_compute_full_prob
  a.out_raw = 1;
  if (precision == 32) {
    HIP_TRY(launch_forward<float>(a, grid, st));
  } else {
    HIP_TRY(launch_forward<double>(a, grid, st));
  }
  return BV_OK;
}

// ---- band model (MLP) ---------------------------------------------------------

struct bv_band_model {
  float* d_params = nullptr;
  float* d_m1 = nullptr;
  float* d_m2 = nullptr;
  float* d_grad = nullptr;
  float* d_loss = nullptr;
  int step = 0;
  hipStream_t stream = nullptr;
};

static void free_model(bv_band_model* m) {
  if (!m) return;
  for (void* b : {static_cast<void*>(m->d_params), static_cast<void*>(m->d_m1),
                  static_cast<void*>(m->d_m2), static_cast<void*>(m->d_grad),
                  static_cast<void*>(m->d_loss)})
    if (b) (void)hipFree(b);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int bv_band_model_create(bv_band_model** out, uint64_t seed) {
  if (!out) return fail(BV_EINVAL, "model output pointer is NULL");
  *out = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_init_locked(-1);
    if (rc) return rc;
  }
  HIP_TRY(hipSetDevice(g_ctx.device));
  auto* m = new bv_band_model();
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&m->d_params, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc(&m->d_m1, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc(&m->d_m2, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc(&m->d_grad, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc(&m->d_loss, sizeof(float)) != hipSuccess) {
    free_model(m);
    return fail(BV_ENOMEM, "model allocation failed");
  }
  // He-uniform weights, zero biases, from a splitmix64 stream.
  std::vector<float> p(kParams, 0.f);
  uint64_t st = seed;
  auto next = [&st]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return static_cast<double>(z >> 11) * 0x1.0p-53;
  };
  auto fill = [&](int off, int fan_in, int count) {
    const double lim = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < count; ++i) p[off + i] = static_cast<float>((2.0 * next() - 1.0) * lim);
  };
  fill(0, kFeat, kHid * kFeat);
  fill(kHid * kFeat + kHid, kHid, kHid * kHid);
  fill(kHid * kFeat + kHid + kHid * kHid + kHid, kHid, 2 * kHid);
  std::vector<float> zero(kParams, 0.f);
  if (hipMemcpy(m->d_params, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_m1, zero.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_m2, zero.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess) {
    free_model(m);
    return fail(BV_EHIP, "model upload failed");
  }
  *out = m;
  return BV_OK;
}

void bv_band_model_destroy(bv_band_model* m) {
  if (!m) return;
  (void)hipSetDevice(g_ctx.device);
  (void)hipStreamSynchronize(m->stream);
  free_model(m);
}

int bv_band_model_get_params(const bv_band_model* m, float* out) {
  if (!m || !out) return fail(BV_EINVAL, "NULL argument");
  HIP_TRY(hipStreamSynchronize(m->stream));
  HIP_TRY(hipMemcpy(out, m->d_params, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  return BV_OK;
}

int bv_band_model_set_params(bv_band_model* m, const float* in) {
  if (!m || !in) return fail(BV_EINVAL, "NULL argument");
  HIP_TRY(hipStreamSynchronize(m->stream));
  HIP_TRY(hipMemcpy(m->d_params, in, sizeof(float) * kParams, hipMemcpyHostToDevice));
  m->step = 0;
  std::vector<float> zero(kParams, 0.f);
  HIP_TRY(hipMemcpy(m->d_m1, zero.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m->d_m2, zero.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  return BV_OK;
}

int bv_band_model_get_state(const bv_band_model* m, float* m1, float* m2, int64_t* step) {
  if (!m) return fail(BV_EINVAL, "NULL argument");
  HIP_TRY(hipStreamSynchronize(m->stream));
  if (m1) HIP_TRY(hipMemcpy(m1, m->d_m1, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  if (m2) HIP_TRY(hipMemcpy(m2, m->d_m2, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  if (step) *step = m->step;
  return BV_OK;
}

int bv_band_model_set_state(bv_band_model* m, const float* m1, const float* m2, int64_t step) {
  if (!m || !m1 || !m2 || step < 0 || step > INT32_MAX) return fail(BV_EINVAL, "bad argument");
  HIP_TRY(hipStreamSynchronize(m->stream));
  HIP_TRY(hipMemcpy(m->d_m1, m1, sizeof(float) * kParams, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m->d_m2, m2, sizeof(float) * kParams, hipMemcpyHostToDevice));
  m->step = static_cast<int>(step);
  return BV_OK;
}

int bv_band_model_read(const char* path, float* params) {
  if (!params) return fail(BV_EINVAL, "params is NULL");
  std::string err;
  const std::string def = path ? std::string() : default_model_path();
  if (int rc = read_model_file(path ? path : def.c_str(), params, kParams, err)) return fail(rc, err);
  return BV_OK;
}

int bv_band_model_load(const char* path, bv_band_model** out) {
  if (!out) return fail(BV_EINVAL, "model output pointer is NULL");
  *out = nullptr;
  std::vector<float> p(kParams);
  if (int rc = bv_band_model_read(path, p.data())) return rc;
  bv_band_model* m = nullptr;
  if (int rc = bv_band_model_create(&m, 0)) return rc;
  if (int rc = bv_band_model_set_params(m, p.data())) {
    bv_band_model_destroy(m);
    return rc;
  }
  *out = m;
  return BV_OK;
}

int bv_band_model_write(const bv_band_model* m, const char* path) {
  std::vector<float> p(kParams);
  if (int rc = bv_band_model_get_params(m, p.data())) return rc;
  std::string err;
  if (int rc = write_model_file(path, p.data(), kParams, err)) return fail(rc, err);
  return BV_OK;
}

int bv_band_model_default_path(char* buf, int64_t size) {
  const std::string p = default_model_path();
  if (p.empty()) return fail(BV_ESTATE, "cannot locate the library to find its model directory");
  if (!buf || size < static_cast<int64_t>(p.size()) + 1) return fail(BV_EINVAL, "buffer too small");
  std::memcpy(buf, p.c_str(), p.size() + 1);
  return BV_OK;
}

namespace {
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  float* f() { return static_cast<float*>(p); }
};
int dev_alloc(DevBuf& b, size_t bytes) {
  if (hipMalloc(&b.p, std::max<size_t>(bytes, 4)) != hipSuccess) return fail(BV_ENOMEM, "device allocation failed");
  return BV_OK;
}
}  // namespace

int bv_band_model_predict(bv_band_model* m, const float* X, int64_t n, float* Y) {
  if (!m || n < 0 || (n > 0 && (!X || !Y))) return fail(BV_EINVAL, "bad arguments");
  if (n == 0) return BV_OK;
  HIP_TRY(hipSetDevice(g_ctx.device));
  DevBuf dx, dy;
  int rc;
  if ((rc = dev_alloc(dx, sizeof(float) * kFeat * n)) || (rc = dev_alloc(dy, sizeof(float) * 2 * n))) return rc;
  HIP_TRY(hipMemcpyAsync(dx.p, X, sizeof(float) * kFeat * n, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(launch_mlp_forward(m->d_params, dx.f(), n, dy.f(), nullptr, nullptr, m->stream));
  HIP_TRY(hipMemcpyAsync(Y, dy.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return BV_OK;
}

// Loss 0.5*mean(w0*e0^2 + w1*e1^2) and its gradient for one batch.
int bv_band_model_loss_grad(bv_band_model* m, const float* X, const float* T, int64_t n, float w0,
                            float w1, float* loss, float* grad) {
  if (!m || n <= 0 || !X || !T) return fail(BV_EINVAL, "bad arguments");
  HIP_TRY(hipSetDevice(g_ctx.device));
  DevBuf dx, dt, dy, h1, h2, d1, d2, d3;
  int rc;
  if ((rc = dev_alloc(dx, sizeof(float) * kFeat * n)) || (rc = dev_alloc(dt, sizeof(float) * 2 * n)) ||
      (rc = dev_alloc(dy, sizeof(float) * 2 * n)) || (rc = dev_alloc(h1, sizeof(float) * kHid * n)) ||
      (rc = dev_alloc(h2, sizeof(float) * kHid * n)) || (rc = dev_alloc(d1, sizeof(float) * kHid * n)) ||
      (rc = dev_alloc(d2, sizeof(float) * kHid * n)) || (rc = dev_alloc(d3, sizeof(float) * 2 * n)))
    return rc;
  HIP_TRY(hipMemcpyAsync(dx.p, X, sizeof(float) * kFeat * n, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemcpyAsync(dt.p, T, sizeof(float) * 2 * n, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(launch_zero_ints(reinterpret_cast<int*>(m->d_loss), 1, m->stream));
  HIP_TRY(launch_mlp_forward(m->d_params, dx.f(), n, dy.f(), h1.f(), h2.f(), m->stream));
  HIP_TRY(launch_mlp_backward(m->d_params, dx.f(), dy.f(), dt.f(), n, h1.f(),
                              h2.f(), d1.f(), d2.f(), d3.f(), w0, w1,
                              m->d_loss, m->d_grad, m->stream));
  if (loss) HIP_TRY(hipMemcpyAsync(loss, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (grad) HIP_TRY(hipMemcpyAsync(grad, m->d_grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return BV_OK;
}

// Adam training on the device.  X [n][36], T [n][2] (host); the data is
// uploaded once and every minibatch step (forward, deltas, gradient, Adam)
// runs as device kernels.  order (host, epochs x n indices, optional) gives
// the sample order; NULL keeps the given order every epoch.
int bv_band_model_train(bv_band_model* m, const float* X, const float* T, int64_t n, int epochs,
                        int batch, float lr, float w0, float w1, const int64_t* order,
                        float* epoch_loss) {
  if (!m || n <= 0 || !X || !T || epochs <= 0 || batch <= 0) return fail(BV_EINVAL, "bad arguments");
  HIP_TRY(hipSetDevice(g_ctx.device));
  const int64_t bsz = std::min<int64_t>(batch, n);
  DevBuf dx, dt, bx, bt, dy, h1, h2, d1, d2, d3;
  int rc;
  if ((rc = dev_alloc(dx, sizeof(float) * kFeat * n)) || (rc = dev_alloc(dt, sizeof(float) * 2 * n)) ||
      (rc = dev_alloc(bx, sizeof(float) * kFeat * bsz)) || (rc = dev_alloc(bt, sizeof(float) * 2 * bsz)) ||
      (rc = dev_alloc(dy, sizeof(float) * 2 * bsz)) || (rc = dev_alloc(h1, sizeof(float) * kHid * bsz)) ||
      (rc = dev_alloc(h2, sizeof(float) * kHid * bsz)) || (rc = dev_alloc(d1, sizeof(float) * kHid * bsz)) ||
      (rc = dev_alloc(d2, sizeof(float) * kHid * bsz)) || (rc = dev_alloc(d3, sizeof(float) * 2 * bsz)))
    return rc;
  // Shuffled copies per epoch are made on the host side of the order array;
  // the device data is permuted once per epoch by host-built gathers.
  std::vector<float> xs(static_cast<size_t>(kFeat * n)), ts(static_cast<size_t>(2 * n));
  for (int ep = 0; ep < epochs; ++ep) {
    const int64_t* ord = order ? order + static_cast<int64_t>(ep) * n : nullptr;
    for (int64_t i = 0; i < n; ++i) {
      const int64_t k = ord ? ord[i] : i;
      std::memcpy(&xs[i * kFeat], X + k * kFeat, sizeof(float) * kFeat);
      ts[2 * i] = T[2 * k];
      ts[2 * i + 1] = T[2 * k + 1];
    }
    HIP_TRY(hipMemcpyAsync(dx.p, xs.data(), sizeof(float) * kFeat * n, hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemcpyAsync(dt.p, ts.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, m->stream));
    HIP_TRY(launch_zero_ints(reinterpret_cast<int*>(m->d_loss), 1, m->stream));
    int nb = 0;
    for (int64_t off = 0; off + bsz <= n; off += bsz, ++nb) {
      const float* xb = dx.f() + off * kFeat;
      const float* tb = dt.f() + off * 2;
      HIP_TRY(launch_mlp_forward(m->d_params, xb, bsz, dy.f(), h1.f(), h2.f(), m->stream));
      HIP_TRY(launch_mlp_backward(m->d_params, xb, dy.f(), tb, bsz, h1.f(), h2.f(),
                                  d1.f(), d2.f(), d3.f(), w0, w1, m->d_loss,
                                  m->d_grad, m->stream));
      ++m->step;
      HIP_TRY(launch_adam(m->d_params, m->d_grad, m->d_m1, m->d_m2, lr, m->step, m->stream));
    }
    if (epoch_loss) {
      float l = 0.f;
      HIP_TRY(hipMemcpyAsync(&l, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
      HIP_TRY(hipStreamSynchronize(m->stream));
      epoch_loss[ep] = nb > 0 ? l / nb : 0.f;
    }
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  return BV_OK;
}

int bv_plan_features(bv_plan* p, float* X_out, int32_t* anchor_out, int32_t* ends_out) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  HIP_TRY(hipSetDevice(g_ctx.device));
  int rc = ensure_feature_buffers(p);
  if (!rc) rc = ensure_x(p);
  if (rc) return rc;
  if (p->n_tasks > 0) HIP_TRY(launch_features(p));
  if (X_out && p->n_pairs)
    HIP_TRY(hipMemcpyAsync(X_out, p->d_X, sizeof(float) * kFeat * p->n_pairs, hipMemcpyDeviceToHost, p->stream));
  if (anchor_out && p->n_pairs)
    HIP_TRY(hipMemcpyAsync(anchor_out, p->d_anchor, sizeof(int) * p->n_pairs, hipMemcpyDeviceToHost, p->stream));
  if (ends_out && p->n_pairs)
    HIP_TRY(hipMemcpyAsync(ends_out, p->d_ends, sizeof(int2) * p->n_pairs, hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return BV_OK;
}

int bv_plan_set_band_model(bv_plan* p, const bv_band_model* m, int margin, int max_width) {
  if (!p) return fail(BV_EINVAL, "plan is NULL");
  HIP_TRY(hipSetDevice(g_ctx.device));
  drop_graph(p);
  if (!m) {
    p->has_model = false;
    p->has_band = false;
    return BV_OK;
  }
  if (margin < 0 || max_width < 1) return fail(BV_EINVAL, "bad band margin or width");
  int rc = ensure_feature_buffers(p);
  if (rc) return rc;
  if (!p->d_model) HIP_TRY(hipMalloc(&p->d_model, sizeof(float) * kParams));
  HIP_TRY(hipStreamSynchronize(m->stream));
  HIP_TRY(hipMemcpy(p->d_model, m->d_params, sizeof(float) * kParams, hipMemcpyDeviceToDevice));
  p->has_model = true;
  p->has_band = true;
  p->band_margin = margin;
  p->band_max_width = max_width;
  return BV_OK;
}

int bv_plan_bands(bv_plan* p, int32_t* bands_out) {
  if (!p || !bands_out) return fail(BV_EINVAL, "NULL argument");
  if (!p->has_band || !p->d_band) return fail(BV_ESTATE, "plan has no bands");
  HIP_TRY(hipSetDevice(g_ctx.device));
  HIP_TRY(hipMemcpyAsync(bands_out, p->d_band, sizeof(int2) * p->n_pairs, hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return BV_OK;
}

// Band labels: the narrowest diagonal span [o_lo, o_hi] of j - r outside
// which the alignment path makes at most 1 - mass expected visits (posterior
// of M, I and D cells summed over rows), from an FP64 forward-backward pass
// with per-row renormalisation on the host.
int bv_band_labels(const bv_region_batch* b, double mass, int n_threads, int32_t* lo_hi) {
  int rc = validate(b);
  if (rc) return rc;
  if (!lo_hi || !(mass > 0.0 && mass <= 1.0)) return fail(BV_EINVAL, "bad arguments");
  static HostTables t;
  static std::once_flag once;
  std::call_once(once, [] { build_tables(t); });
  Alphabet al;
  alphabet_avx(al);
  struct Job { int64_t read, hap, out; };
  std::vector<Job> jobs;
  int64_t base = 0;
  for (int64_t g = 0; g < b->n_regions; ++g) {
    const int64_t r0 = b->region_read_off[g], r1 = b->region_read_off[g + 1];
    const int64_t h0 = b->region_hap_off[g], h1 = b->region_hap_off[g + 1];
    for (int64_t r = r0; r < r1; ++r)
      for (int64_t h = h0; h < h1; ++h) jobs.push_back(Job{r, h, base + (r - r0) * (h1 - h0) + (h - h0)});
    base += (r1 - r0) * (h1 - h0);
  }
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    std::vector<double> F, B, diag;
    std::vector<double> fs, bs;
    for (;;) {
      const int64_t k = next.fetch_add(1);
      if (k >= static_cast<int64_t>(jobs.size())) break;
      const Job& jb = jobs[k];
      const int64_t ro = b->read_off[jb.read], ho = b->hap_off[jb.hap];
      const int R = static_cast<int>(b->read_off[jb.read + 1] - ro);
      const int H = static_cast<int>(b->hap_off[jb.hap + 1] - ho);
      const size_t W = static_cast<size_t>(H) + 2;
      const size_t N = (static_cast<size_t>(R) + 2) * W;
      F.assign(3 * N, 0.0);
      B.assign(3 * N, 0.0);
      fs.assign(R + 2, 0.0);
      bs.assign(R + 2, 0.0);
      auto fm = [&](int r, int j) -> double& { return F[(static_cast<size_t>(r) * W + j) * 3]; };
      auto fi = [&](int r, int j) -> double& { return F[(static_cast<size_t>(r) * W + j) * 3 + 1]; };
      auto fd = [&](int r, int j) -> double& { return F[(static_cast<size_t>(r) * W + j) * 3 + 2]; };
      auto bm = [&](int r, int j) -> double& { return B[(static_cast<size_t>(r) * W + j) * 3]; };
      auto bi = [&](int r, int j) -> double& { return B[(static_cast<size_t>(r) * W + j) * 3 + 1]; };
      auto bd = [&](int r, int j) -> double& { return B[(static_cast<size_t>(r) * W + j) * 3 + 2]; };
      struct P { double mm, g, mi, ii, md, dd, pm, pmis; uint8_t cls; };
      std::vector<P> prm(R + 2);
      for (int r = 1; r <= R; ++r) {
        const int64_t p = ro + r - 1;
        const int iq = b->read_ins[p] & 127, dq = b->read_del[p] & 127, cq = b->read_gcp[p] & 127;
        const int lo = std::min(iq, dq), hi = std::max(iq, dq);
        const double e = t.ph2pr_d[b->read_qual[p] & 127];
        const uint8_t rc8 = al.read_class[b->read_bases[p]];
        prm[r] = P{t.mm_d[((hi * (hi + 1)) >> 1) + lo], 1.0 - t.ph2pr_d[cq], t.ph2pr_d[iq], t.ph2pr_d[cq],
                   t.ph2pr_d[dq], t.ph2pr_d[cq], 1.0 - e, (rc8 & kWildcard) ? 1.0 - e : e / 3.0, rc8};
      }
      auto prior = [&](int r, int j) {
        const uint8_t hm = al.hap_mask[b->hap_bases[ho + j - 1]];
        return ((hm >> (prm[r].cls & 7)) & 1) ? prm[r].pm : prm[r].pmis;
      };
      for (int j = 0; j <= H; ++j) fd(0, j) = 1.0 / H;
      for (int r = 1; r <= R; ++r) {
        double mx = 0.0;
        for (int j = 1; j <= H; ++j) {
          fm(r, j) = (fm(r - 1, j - 1) * prm[r].mm + (fi(r - 1, j - 1) + fd(r - 1, j - 1)) * prm[r].g) * prior(r, j);
          fi(r, j) = fm(r - 1, j) * prm[r].mi + fi(r - 1, j) * prm[r].ii;
          fd(r, j) = fm(r, j - 1) * prm[r].md + fd(r, j - 1) * prm[r].dd;
          mx = std::max(mx, std::max(fm(r, j), std::max(fi(r, j), fd(r, j))));
        }
        int e = 0;
        if (mx > 0.0) std::frexp(mx, &e);
        for (int j = 0; j <= H; ++j) {
          fm(r, j) = std::ldexp(fm(r, j), -e);
          fi(r, j) = std::ldexp(fi(r, j), -e);
          fd(r, j) = std::ldexp(fd(r, j), -e);
        }
        fs[r] = fs[r - 1] + e;
      }
      double tot = 0.0;
      for (int j = 1; j <= H; ++j) tot += fm(R, j) + fi(R, j);
      const double log2L = std::log2(tot) + fs[R];
      for (int j = 1; j <= H; ++j) {
        bm(R, j) = 1.0;
        bi(R, j) = 1.0;
      }
      for (int r = R - 1; r >= 1; --r) {
        double mx = 0.0;
        for (int j = H; j >= 1; --j) {
          const double mn = j < H ? bm(r + 1, j + 1) * prior(r + 1, j + 1) : 0.0;
          const double dn = j < H ? bd(r, j + 1) : 0.0;
          bm(r, j) = prm[r + 1].mm * mn + prm[r + 1].mi * bi(r + 1, j) + prm[r].md * dn;
          bi(r, j) = prm[r + 1].g * mn + prm[r + 1].ii * bi(r + 1, j);
          bd(r, j) = prm[r + 1].g * mn + prm[r].dd * dn;
          mx = std::max(mx, std::max(bm(r, j), std::max(bi(r, j), bd(r, j))));
        }
        int e = 0;
        if (mx > 0.0) std::frexp(mx, &e);
        for (int j = 1; j <= H; ++j) {
          bm(r, j) = std::ldexp(bm(r, j), -e);
          bi(r, j) = std::ldexp(bi(r, j), -e);
          bd(r, j) = std::ldexp(bd(r, j), -e);
        }
        bs[r] = bs[r + 1] + e;
      }
      diag.assign(static_cast<size_t>(R + H), 0.0);
      double all = 0.0;
      for (int r = 1; r <= R; ++r) {
        const double sc = std::exp2(fs[r] + bs[r] - log2L);
        for (int j = 1; j <= H; ++j) {
          const double v = (fm(r, j) * bm(r, j) + fi(r, j) * bi(r, j) + fd(r, j) * bd(r, j)) * sc;
          diag[static_cast<size_t>(j - r + R - 1)] += v;
          all += v;
        }
      }
      // Narrowest window outside which the expected number of path visits is
      // at most 1 - mass.  A path that leaves a diagonal window visits a
      // diagonal next to it (every move changes j - r by at most one), so
      // this bounds the probability that the alignment leaves the window.
      const int nd = R + H - 1;
      const double need = all - (1.0 - mass);
      int best_lo = 0, best_hi = nd - 1;
      double win = 0.0;
      int lo = 0;
      for (int hi = 0; hi < nd; ++hi) {
        win += diag[hi];
        while (lo < hi && win - diag[lo] >= need) win -= diag[lo++];
        if (win >= need && hi - lo < best_hi - best_lo) {
          best_lo = lo;
          best_hi = hi;
        }
      }
      lo_hi[2 * jb.out] = best_lo - (R - 1);
      lo_hi[2 * jb.out + 1] = best_hi - (R - 1);
    }
  };
  if (n_threads <= 0) n_threads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int i = 0; i < n_threads; ++i) th.emplace_back(worker);
  for (auto& x : th) x.join();
  return BV_OK;
}

}  // extern "C"

// ---- batch accelerator hook (lib/intel_native/pairhmm/shacc_pairhmm.h:10-36) --
// The reference declares shacc_pairhmm::calculate as a weak C++ symbol and
// never defines it; this is the strong definition.  results receives
// log10 likelihoods (FP32 with FP64 rescue), read-major [num_reads][num_haps],
// as float.  Returns false on any error (message in bv_pairhmm_last_error).
namespace shacc_pairhmm {
struct Read {
  int length;
  const char* bases;
  const char* q;
  const char* i;
  const char* d;
  const char* c;
};
struct Haplotype {
  int length;
  const char* bases;
};
struct Batch {
  int num_reads;
  int num_haps;
  long num_cells;
  Read* reads;
  Haplotype* haps;
  float* results;
};

bool calculate(Batch& batch) {
  const int nr = batch.num_reads, nh = batch.num_haps;
  if (nr < 0 || nh < 0 || (nr > 0 && nh > 0 && (!batch.reads || !batch.haps || !batch.results))) {
    g_err = "shacc_pairhmm::calculate: bad batch";
    return false;
  }
  if (nr == 0 || nh == 0) return true;
  std::vector<int64_t> rro{0, nr}, rho{0, nh}, ro(nr + 1), ho(nh + 1);
  std::vector<uint8_t> rs, q, ins, del, gcp, hap;
  for (int r = 0; r < nr; ++r) {
    const Read& x = batch.reads[r];
    ro[r] = static_cast<int64_t>(rs.size());
    if (x.length <= 0) {
      g_err = "shacc_pairhmm::calculate: empty read";
      return false;
    }
    rs.insert(rs.end(), x.bases, x.bases + x.length);
    q.insert(q.end(), x.q, x.q + x.length);
    ins.insert(ins.end(), x.i, x.i + x.length);
    del.insert(del.end(), x.d, x.d + x.length);
    gcp.insert(gcp.end(), x.c, x.c + x.length);
  }
  ro[nr] = static_cast<int64_t>(rs.size());
  for (int h = 0; h < nh; ++h) {
    ho[h] = static_cast<int64_t>(hap.size());
    const Haplotype& y = batch.haps[h];
    if (y.length <= 0) {
      g_err = "shacc_pairhmm::calculate: empty haplotype";
      return false;
    }
    hap.insert(hap.end(), y.bases, y.bases + y.length);
  }
  ho[nh] = static_cast<int64_t>(hap.size());
  bv_region_batch b{1, rro.data(), rho.data(), nr, ro.data(), rs.data(), q.data(), ins.data(),
                    del.data(), gcp.data(), nh, ho.data(), hap.data()};
  std::vector<double> out(static_cast<size_t>(nr) * nh);
  if (bv_pairhmm_batch(&b, nullptr, out.data(), nullptr) != BV_OK) return false;
  for (size_t k = 0; k < out.size(); ++k) batch.results[k] = static_cast<float>(out[k]);
  return true;
}
}  // namespace shacc_pairhmm
