// Shared host/device definitions for the MI355X PairHMM.
#pragma once

#include <cstdint>

namespace bvp {

// One unit of kernel work: a read against a contiguous run of haplotypes
// (all haplotypes of its region in the FP32 pass, one haplotype in the FP64
// pass).  Results go to out[out0 + h] for h in [0, nhap).
struct alignas(8) Task {
  int32_t read;
  int32_t hap0;
  int32_t nhap;
  int32_t band;  // -2: a pair the banded kernel rejected; otherwise unused
  int64_t out0;
};

// Per-pair result form, written next to every log10 likelihood.
enum Form : uint8_t {
  kFormFP32 = 0,        // FP32 forward, S_f >= 1e-28f
  kFormFP64 = 1,        // FP32 underflowed, FP64 re-run
  kFormBandFP32 = 2,    // banded FP32 forward accepted
  kFormBandFallback = 3,// band rejected (boundary mass), full FP32 used
  kFormBandFP64 = 4,    // moving-band FP64 forward accepted (long read; short pair whose FP32 band underflowed)
  kFormBandLong32 = 5,  // long read: moving-band FP32 forward (rescaled per 64 rows) accepted
  kFormPending = 255,   // internal: queued for the FP64 pass
};

// Read base classes.  Bits 0-2: match class (slot in the haplotype match
// window); bit 6: the last base of its read; bit 7: wildcard (the read base
// matches everything).
constexpr uint8_t kWildcard = 0x80;
constexpr uint8_t kReadEnd = 0x40;

// One unit of the stacked full-matrix kernel: reads [read0, read0 + nreads)
// of one region, consecutive in the read arrays, stacked row after row,
// against haplotypes [hap0, hap0 + nhap).  The result of read r and
// haplotype h goes to out[out0 + (r - read0) * stride + (h - hap0)].
struct alignas(16) StackTask {
  int32_t read0;
  int32_t nreads;
  int32_t hap0;
  int32_t nhap;
  int64_t out0;
  int32_t stride;
  int32_t rows;  // sum of the reads' lengths
};
constexpr int kStackEnds = 2;     // read ends a stripe holds besides its last lane
constexpr int kStackHmax = 1024;  // longest haplotype of the stacked kernel (LDS carry row)
constexpr int kStackPad = 64;     // carry entries before column 0
constexpr int kStackRows = 1008;  // a stacked task holds at most this many rows (16 stripes)

constexpr int kLanes = 64;        // wavefront width
constexpr int kRowsPerStripe = 63;// lane 0 carries the previous stripe's last row
constexpr int kBlock = 64;        // anti-diagonal steps per staging block
constexpr int kSlots = 8;         // match classes per haplotype column
constexpr int kCarryPad = 192;    // extra carry entries past the haplotype end
constexpr int kSplitRows = 256;   // reads at least this long get one task per haplotype
constexpr int kFp64FirstRows = 1024;  // reads at least this long go to FP64 first (FP32 mode)

// Rescue threshold of the reference driver (pairhmm_common.h:16).
constexpr float kMinAccepted = 1e-28f;

// A plan's device counters (list lengths, work-claim counters, tallies) are
// kCtrStride ints apart, one 128-byte line each: atomics on one line from the
// whole GPU serialise at about 88 M per second (measured on MI355X for one
// line), while atomics on separate lines proceed in parallel.  Per-XCD claim
// counters are arrays with this stride as well.
constexpr int kCtrStride = 32;

// Transition mode.
enum Mode : int {
  kModeAvx = 0,     // per-base transitions from the quality bytes (AvxPairHMM)
  kModeScalar = 1,  // constant transition matrix (scalar PairHMM)
};

struct Tables {
  const float* ph2pr_f;   // [128]
  const double* ph2pr_d;  // [128]
  const float* mm_f;      // [8256]: match-to-match for ins,del quals < 128
  const double* mm_d;     // [8256]
  const double* err_scalar;  // [256]: scalar-mode error prob by raw byte
  double trans[7];        // scalar-mode {MM, MI, MD, IM, II, DM, DD}
};

// Kernel arguments (one struct, passed by value).
struct KArgs {
  // reads: per-position class byte and raw quality bytes
  const uint8_t* rclass;
  const uint8_t* q;
  const uint8_t* ins;
  const uint8_t* del;
  const uint8_t* gcp;
  const int64_t* read_off;
  // haplotypes: per-position match bitmask over read classes
  const uint8_t* hmask;
  const int64_t* hap_off;
  // work
  const Task* tasks;
  int n_tasks;
  const int* n_tasks_dev;  // when set, the task count is read from device memory
  int* counter;
  // results
  double* out;
  uint8_t* form;
  Task* rescue;         // FP32 pass: pairs to re-run in FP64
  int* rescue_count;
  // per-wavefront carry scratch
  void* carry;
  int64_t carry_stride;  // Pair2<T> entries per wavefront
  Tables tab;
  int mode;
  int renorm;   // FP64: rescale the carry row after every stripe
  int out_raw;  // write the scaled sum S instead of log10 L (testcase ABI)
  unsigned long long* rescue_cells;  // FP32 pass: sum of R*H over rescued pairs
  unsigned long long* cells_done;    // cells computed (R*H per pair)
  // FP64 pass over long reads run before any FP32 attempt: pairs whose log10 L
  // is at least verify_above may not underflow in FP32 and are queued here
  Task* verify;
  int* verify_count;
  double verify_above;
  int verify_pass;  // FP32 pass over those pairs: keep the FP64 result when S < 1e-28
  unsigned long long* fp64_cells;  // FP64 passes: sum of R*H over the pairs they compute
  // Longest haplotype the carry rows hold; a pair with a longer (or empty)
  // haplotype or an empty read gets NaN instead of a sweep past its row.
  int hcap;
  // Allocation-free per-pair form: when rbases is set, read and haplotype
  // bytes (AVX alphabet) are turned into classes and match masks as they are
  // loaded (rclass and hmask are unused), and task i is pair i.
  const uint8_t* rbases;
  const uint8_t* hbases;
  int identity_tasks;
  // Stacked full-matrix kernel (FP32): its tasks and the length of its LDS
  // carry row (columns, at least the longest haplotype + 65)
  const StackTask* stasks;
  int carry_cols;
};

// Scratch of the allocation-free entry point: a 256-byte header (the task
// queue counter) followed by the per-wavefront carry rows.
constexpr int64_t kNoallocScratchBytes = int64_t(32) << 20;
constexpr int64_t kNoallocHeader = 256;

// Post-processing of one run's results (post_kernel): per read, the
// normalised row, the filter decision and the per-allele maxima.
struct PostArgs {
  const double* L;                // log10 L per pair, region by region, read-major
  int64_t read0, read_end;        // the reads the regions cover: [read0, read_end)
  int64_t n_regions;
  const int64_t* region_read_off; // n_regions + 1 (read indices)
  const int64_t* region_hap_off;  // n_regions + 1
  const int64_t* region_pair_off; // n_regions + 1
  const int64_t* read_off;        // n_reads + 1 (read lengths)
  int normalize;                  // cap rows at best - 4.5 before writing Lnorm / A
  uint8_t* keep;                  // n_reads, or null
  double* Lnorm;                  // n_pairs, or null
  const int32_t* hap_allele;      // per haplotype: allele index in its region
  const int32_t* n_alleles;       // per region
  const int64_t* allele_off;      // per region: first entry of its block in A
  double* A;                      // sum over regions of reads * alleles, or null
};

}  // namespace bvp
