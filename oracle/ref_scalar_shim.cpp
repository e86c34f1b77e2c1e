// TEST INFRASTRUCTURE ONLY: the reference's scalar PairHMM
// (include/biovoltron/algo/align/inexact_match/pairhmm.hpp:39-141) compiled
// in place from the reference tree, behind a C ABI for the oracle's tests.
// Nothing here is product code; the product never loads this library.
#include <biovoltron/algo/align/inexact_match/pairhmm.hpp>

#include <cstdint>
#include <string>
#include <vector>

namespace {

using biovoltron::Haplotype;
using biovoltron::PairHMM;
using biovoltron::SamRecord;

std::vector<SamRecord<>> make_reads(int n_reads, const int64_t* read_off, const uint8_t* seq, const uint8_t* qual) {
  std::vector<SamRecord<>> reads(static_cast<size_t>(n_reads));
  for (int r = 0; r < n_reads; ++r) {
    const int64_t a = read_off[r], b = read_off[r + 1];
    reads[r].qname = std::to_string(r);
    reads[r].seq.assign(reinterpret_cast<const char*>(seq) + a, static_cast<size_t>(b - a));
    reads[r].qual.assign(reinterpret_cast<const char*>(qual) + a, static_cast<size_t>(b - a));
  }
  return reads;
}

}  // namespace

extern "C" {

// One region through PairHMM::compute_likelihoods (pairhmm.hpp:116-141):
// raw log10 likelihoods, normalise, filter.  out receives the kept rows in
// order (n_kept x n_haps); kept[r] = 1 for the reads that survive.  Returns
// n_kept.  trans: the 7 TrasMatrx entries.
int ref_scalar_region(int n_reads, const int64_t* read_off, const uint8_t* seq, const uint8_t* qual, int n_haps,
                      const int64_t* hap_off, const uint8_t* hap, const double* trans, double* out,
                      uint8_t* kept) {
  auto reads = make_reads(n_reads, read_off, seq, qual);
  std::vector<Haplotype> haps(static_cast<size_t>(n_haps));
  for (int h = 0; h < n_haps; ++h)
    haps[h].seq.assign(reinterpret_cast<const char*>(hap) + hap_off[h], static_cast<size_t>(hap_off[h + 1] - hap_off[h]));
  PairHMM::TrasMatrx tr{};
  for (int k = 0; k < PairHMM::TRANS_PROB_ARRAY_LENGTH; ++k) tr[k] = trans[k];
  const auto ll = PairHMM{}.compute_likelihoods(haps, reads, tr);
  for (int r = 0; r < n_reads; ++r) kept[r] = 0;
  for (size_t i = 0; i < reads.size(); ++i) {
    kept[std::stoi(reads[i].qname)] = 1;
    for (int h = 0; h < n_haps; ++h) out[i * static_cast<size_t>(n_haps) + h] = ll[i][h];
  }
  return static_cast<int>(reads.size());
}

// PairHMM::normalize_likelihoods then filter_poorly_modeled_reads
// (pairhmm.hpp:83-114) on a given raw matrix; outputs as above.
int ref_postprocess(int n_reads, int n_haps, const double* L, const int64_t* read_len, double* out,
                    uint8_t* kept) {
  std::vector<SamRecord<>> reads(static_cast<size_t>(n_reads));
  std::vector<std::vector<double>> ll(static_cast<size_t>(n_reads), std::vector<double>(static_cast<size_t>(n_haps)));
  for (int r = 0; r < n_reads; ++r) {
    reads[r].qname = std::to_string(r);
    reads[r].seq.assign(static_cast<size_t>(read_len[r]), 'A');
    for (int h = 0; h < n_haps; ++h) ll[r][h] = L[static_cast<size_t>(r) * n_haps + h];
  }
  PairHMM::normalize_likelihoods(ll);
  PairHMM::filter_poorly_modeled_reads(reads, ll);
  for (int r = 0; r < n_reads; ++r) kept[r] = 0;
  for (size_t i = 0; i < reads.size(); ++i) {
    kept[std::stoi(reads[i].qname)] = 1;
    for (int h = 0; h < n_haps; ++h) out[i * static_cast<size_t>(n_haps) + h] = ll[i][h];
  }
  return static_cast<int>(reads.size());
}

}  // extern "C"
