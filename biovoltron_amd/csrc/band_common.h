// Shared definitions for band prediction (features, MLP).
#pragma once

#include <cstdint>

#include "pairhmm_band.h"
#include "pairhmm_common.h"

namespace bvp {

constexpr int kFeat = 36;   // features per pair
constexpr int kBins = 32;   // seed-offset histogram bins
constexpr int kHid = 64;    // hidden width
constexpr int kParams = kHid * kFeat + kHid + kHid * kHid + kHid + 2 * kHid + 2;  // 6658
constexpr int kBandMaxHap = kBandHapMax;  // longest haplotype the feature kernel indexes

constexpr int kSeedK = 12;         // seed length
constexpr int kEndSlack = 8;       // read bases the anchoring seeds may leave uncovered (narrow band)
constexpr int kEndSlackWide = 32;  // the same for the wide band
constexpr int kEndAnchorTol = 8;   // |hit offset - anchor| for an end-anchoring seed

struct FeatArgs {
  const uint8_t* rclass;
  const uint8_t* q;
  const int64_t* read_off;
  const uint8_t* hmask;
  const int64_t* hap_off;
  const unsigned long long* htab;  // per haplotype 12-mer tables (hap_table_kernel)
  const int64_t* htab_off;         // [n_haps + 1] slot offsets
  const Task* tasks;
  int n_tasks;
  float* X;      // [pairs][kFeat]
  int* anchor;   // [pairs]
  int2* ends;    // [pairs]: first seed start, last seed end (read positions) among
                 // hits within kEndAnchorTol of the anchor; (-1, -1) without hits
  int* next;     // [8] task claim counters, one per XCD (kCtrStride apart), zero at launch
  // Compact records instead of X / anchor / ends when not null: per pair
  // kRecBytes (the 32 bin counts as bytes, the anchor, the seed-end span as
  // two int16), per task (region order) the 1 / seeds factor and the mean
  // and minimum quality features.
  uint8_t* rec;
  float* trec;  // [tasks][3]
};

constexpr int kRecBytes = 48;  // 16-byte aligned: three 16-byte loads

}  // namespace bvp
