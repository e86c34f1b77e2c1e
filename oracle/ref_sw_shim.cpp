// Smith-Waterman reference build -- TEST INFRASTRUCTURE ONLY.
//
// A C entry point around the reference's own SmithWaterman::align
// (include/biovoltron/algo/align/inexact_match/smithwaterman.hpp), compiled
// from the header where it lies by oracle/Makefile into
// oracle/_ref/libref_sw.so.  No reference source is copied.
#include <biovoltron/algo/align/inexact_match/smithwaterman.hpp>

#include <cstdint>
#include <cstring>
#include <string>
#include <string_view>

extern "C" int ref_sw_align(int R, const char* ref, int H, const char* alt, const int32_t* params,
                            int32_t* offset, char* cigar, int cap) {
  const biovoltron::SmithWaterman::Parameters p{params[0], params[1], params[2], params[3]};
  const auto [off, c] = biovoltron::SmithWaterman::align(std::string_view(ref, R), std::string_view(alt, H), p);
  const std::string s(c);
  *offset = off;
  if (static_cast<int>(s.size()) + 1 > cap) return -1;
  std::memcpy(cigar, s.c_str(), s.size() + 1);
  return static_cast<int>(s.size());
}
