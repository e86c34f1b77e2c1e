// Banded PairHMM: shared host/device definitions.
#pragma once

#include <cstdint>

#include "pairhmm_common.h"

namespace bvp {

constexpr int kBandLanes = 64;    // diagonals evaluated per banded pair (one per lane)
constexpr int kBandEdge = 4;      // diagonals per side whose touch mass is measured
constexpr int kBandHapMax = 1024; // longest haplotype the banded kernel stages in LDS

// One banded pair: diagonals j - r in [o_lo, o_lo + 63].
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
  const int64_t* read_off;
  // haplotypes
  const uint8_t* hmask;
  const int64_t* hap_off;
  // dispatch input: tasks and the per-pair band table
  const Task* tasks;
  int n_tasks;
  const int2* band;
  // banded work list (dispatch output, band kernel input)
  BandItem* items;
  int* item_count;
  int* counter;       // band kernel work queue
  // pairs for the full-matrix FP32 pass (ineligible or rejected)
  Task* full;
  int* full_count;
  // tables
  const float* ph2pr;  // [128]
  const float* mm;     // [8256]
  // results
  double* out;
  uint8_t* form;
  float tol;                       // accept when touch mass <= tol * S
  int* accepted;                   // banded pairs accepted
  unsigned long long* cells_done;  // lane-cells computed
};

}  // namespace bvp
