// Shared host/device definitions of the Smith-Waterman kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bvp {

constexpr int kSwMaxLen = 4096;      // longest reference or alternate sequence
constexpr int kSwMaxWeight = 65536;  // largest |weight| (scores stay far from overflow)

struct SwArgs {
  int64_t n_pairs;          // pairs of this launch (entries of order)
  const int64_t* ref_off;
  const uint8_t* ref;
  const int64_t* alt_off;
  const uint8_t* alt;
  int wm, wx, wo, we;
  const int32_t* order;     // the pairs this launch takes, in order
  int* queue;               // counter over order[0 .. n_pairs), zero at launch
  char* scratch;            // per wavefront: scratch_stride bytes
  int64_t scratch_stride;
  const int64_t* out_off;   // per pair: first CIGAR element slot
  int32_t* offset;          // per pair
  int32_t* n_elems;         // per pair
  char* ops;                // CIGAR element operations
  int32_t* lens;            // CIGAR element lengths
};

int64_t sw_scratch_bytes(int R, int H);
hipError_t launch_sw(const SwArgs& a, int grid, hipStream_t st);

}  // namespace bvp
