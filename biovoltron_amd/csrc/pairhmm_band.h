// Banded PairHMM: shared host/device definitions.
#pragma once

#include <cstdint>

#include "pairhmm_common.h"

namespace bvp {

constexpr int kBandLanes = 64;    // wavefront width; the wide band's diagonals (one per lane)
constexpr int kBandNarrow = 32;   // the narrow band: two pairs of one read per wavefront
constexpr int kBandEdge = 4;      // diagonals per side whose touch mass is measured
constexpr int kBandHapMax = 1024; // longest haplotype given a band (seed tables, features)
constexpr int kBandReadMax = 448; // longest read given a band (LDS haplotype window)
constexpr int kBandSlide = 8;     // wide band: start slide (diagonals)
constexpr float kBandHold = 1e-9f;// a move that would drop more of a row's mass is held
constexpr int kEdgeRunRows = 48;  // insertion runs leaving the band bottom: last rows counted
constexpr int kLongEvery = 16;    // long-read band: rows between moves
constexpr int kLongMove = kBandEdge;  // long-read band: diagonals per move, at most
constexpr int kXcdCount = 8;   // MI355X accelerator complex dies (workgroups go round-robin over them)
constexpr int kBandChunk = 8;  // consecutive units per wavefront assignment (short-read kernel)
constexpr int kLongStartLift = 16;// long-read start band: the anchor sits this far above lane 0
                                  // less than the centre (room above the alignment)
constexpr int kQuadPairs = 4;     // wide band, quad shape: pairs of one read per wavefront
constexpr int kQuadLanes = 16;    // lanes per pair (one DPP row)
constexpr int kQuadCells = 4;     // diagonals per lane
static_assert(kQuadLanes * kQuadCells == kBandLanes, "a quad pair covers the wide band");
static_assert(kQuadCells == kBandEdge, "an edge zone is one lane of the quad shape");
// The mid band: 48 diagonals, the quad shape with three diagonals per lane;
// its edge zones are one lane (three diagonals) per side.  Bands of 33-48
// diagonals a caller sets take it when it is on (the learned rules do not:
// a 48-diagonal band misses alignments 20-30 diagonals from the predicted
// centre on a few pairs in 10^7 without its edge zones seeing them).
constexpr int kMidCells = 3;
constexpr int kBandMid = kQuadLanes * kMidCells;
constexpr int kMidEdge = kMidCells;

// One banded pair: diagonals j - r in [o_lo, o_lo + width).  In the narrow
// list items 2u and 2u+1 share a read (hap < 0 marks padding), in the wide
// list of the quad shape items 4u..4u+3.
struct alignas(8) BandItem {
  int32_t read;
  int32_t hap;
  int32_t o_lo;
  int32_t pad;
  int64_t out;
};

struct BArgs {
  // reads
  const uint8_t* rclass;
  const uint8_t* q;
  const uint8_t* ins;
  const uint8_t* del;
  const uint8_t* gcp;
  const uint8_t* gcp_const;  // per read: 1 if the gap continuation is constant along it, quality not 0
  const int64_t* read_off;
  // haplotypes
  const uint8_t* hmask;
  const int64_t* hap_off;
  // dispatch input: tasks and the per-pair band table
  const Task* tasks;
  int n_tasks;
  const int2* band;
  // banded work lists (dispatch output, band kernel input): 64-wide bands
  // (with the quad shape: groups of four pairs of one read, a group's last
  // slots possibly padding, and in items1 the rest of each read's wide
  // pairs, fewer than four, for the one-pair kernel), and 32-wide bands in
  // read-sharing pairs
  BandItem* items;
  int* item_count;
  BandItem* items1;
  int* item1_count;
  BandItem* items32;
  int* item32_count;
  // 48-wide bands (quad shape only), groups of four pairs of one read; the
  // rest of each read's mid pairs go to items1 as 64-wide bands
  BandItem* items48;
  int* item48_count;
  // pairs for the full-matrix FP32 pass (ineligible or rejected): haplotypes
  // of at most kStackHmax bases (stacked kernel), and longer ones (striped kernel)
  Task* full;
  int* full_count;
  Task* full_long;
  int* full_long_count;
  // tables
  const float* ph2pr;  // [128]
  const float* mm;     // [8256]
  // results
  double* out;
  uint8_t* form;
  float tol;                       // accept when touch mass <= tol * S
  int* accepted;                   // banded pairs accepted
  int* claims;                     // [16] chunk claim counters (kCtrStride apart): per XCD, one-pair wide then narrow kernel
  int* quad_claims;                // [8] the same for the quad kernel (kCtrStride apart)
  int* quad48_claims;              // [8] the same for the mid-band quad kernel
  unsigned long long* cells_done;  // lane-cells computed
  // long reads (FP64 first, one task per pair): the moving-band list and the
  // pairs left to the full-matrix FP64 pass (no band, or band rejected)
  const Task* tasks64;
  int n_tasks64;
  int long_on;  // 0: every long pair takes the full matrix
  int long_fp32;  // the moving band in FP32 (rescaled every 16 rows), else FP64 (every 64)
  BandItem* long_items;
  int* long_count;
  int* long_queue;
  Task* long_full;
  int* long_full_count;
  int* long_accepted;
  int long_cap;  // entries of long_items: the long kernel runs min(*long_count, long_cap) of them
  // FP32 moving band, quad shape: the long reads one task each (all their
  // haplotypes), and the units of four pairs of one read they give
  const Task* tasks64r;
  int n_tasks64r;
  BandItem* long_items4;
  int* long4_count;
  int* long4_queue;
  // Short-read pairs whose FP32 band sum fell below 1e-28 (the reference
  // would re-run them in FP64): the FP64 moving band over a 64-diagonal band
  // with the same centre, at most band_rescue_cap of them (the rest take the
  // full-matrix lists as before); null when off
  BandItem* band_rescue;
  int* band_rescue_count;
  int band_rescue_cap;
  const double* ph2pr_d;  // [128]
  const double* mm_d;     // [8256]
  double verify_above;    // log10 L at or above which FP32 might not underflow
  unsigned long long* fp64_cells;
  // short-read kernel: selector window columns per read class (dynamic LDS),
  // at least the longest banded read + kBandLanes + 2 * kBandSlide
  int sel_cols;
  // wide band shape: 1 = one pair per wavefront (one diagonal per lane),
  // kQuadPairs = four pairs of one read per wavefront (four diagonals per
  // lane) for each read's whole groups of four
  int wide_pairs;
  // the quad shape also runs 48-wide bands (three diagonals per lane); else
  // a band of 33-48 diagonals takes the wide lists (as a 64-wide band)
  int mid_on;
  // quad shape: a task's (or task piece's) last wide pairs, fewer than
  // four, fill one more unit of four (the empty slots padding) when there
  // are at least quad_fill of them, else they go to items1; 4: never
  int quad_fill;
};

// Dynamic LDS bytes of the short-read banded kernel for `pairs` pairs per
// wavefront (1 or 2): the row table, then 4 selector rows of sel_cols per pair.
constexpr int band_sel_cols(int read_max) {
  return ((read_max + kBandLanes + 2 * kBandSlide) + 7) & ~7;
}

}  // namespace bvp
