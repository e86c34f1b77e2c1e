// gpu_pairhmm.hpp -- header-only C++20 drop-in for Biovoltron's PairHMM classes,
// backed by the MI355X C ABI (bv_pairhmm.h, libbvpairhmm.so).
//
// Replaces (paths relative to the reference tree):
//   biovoltron::AvxPairHMM::compute_likelihoods(haps, reads)
//       include/biovoltron/algo/align/inexact_match/pairhmm_avx.hpp:160-171
//   biovoltron::PairHMM::compute_likelihoods(haps, reads, trans)
//       include/biovoltron/algo/align/inexact_match/pairhmm.hpp:116-141
//   PairHMM::normalize_likelihoods / filter_poorly_modeled_reads
//       pairhmm.hpp:83-114 (reproduced as written: reads whose best
//       likelihood is *below* the threshold are the ones kept)
//   HaplotypeCaller's one call per window (haplotypecaller.hpp:90, inside
//       the window loop of :126-151): compute_likelihoods_batch takes every
//       window at once -- one device run, normalise and filter on the device
//
// Types are duck-typed so the header needs none of the reference's headers:
// a read provides .seq and .qual (string-like) and, for the AVX model,
// insertion_gop(), deletion_gop(), overall_gcp() (string-like, at least
// .seq.size() bytes -- shorter strings are an error here, where the reference
// reads past the end of a 256-byte buffer); a haplotype provides .seq.
//
// Like the reference classes, GpuPairHMM is default-constructible, copyable
// and const-callable, so it can sit in HaplotypeCaller's `const PairHMM
// pairhmm` member (applications/haplotypecaller/haplotypecaller.hpp:29,90).
// Unlike AvxPairHMM it keeps no state between calls.
//
// Learned bands (no reference counterpart): GpuPairHMM(LearnedBand{}) runs
// the AVX model through the band network and the banded kernel, with the
// full matrix as fallback for every pair whose band is not accepted (results
// within 1e-3 in log10 of the full matrix).  The model is the trained one
// shipped next to libbvpairhmm.so unless a path is given; copies of the
// object share it.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "bv_pairhmm.h"

namespace biovoltron_amd {

// Options of the learned band (bv_plan_set_band_model): the model file
// (empty: models/band_mlp.bin beside the library), the diagonals added per
// side of the predicted span, and the widest span given a band.
struct LearnedBand {
  std::string model_path;
  int margin = 10;
  int max_width = 68;
};

struct PairHMMError : std::runtime_error {
  int status;
  PairHMMError(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

inline void check(int rc) {
  if (rc != BV_OK) throw PairHMMError(rc, std::string("bv_pairhmm: ") + bv_pairhmm_last_error());
}

// pairhmm.hpp:27-36
inline constexpr double kMaximumBestAltLikelihoodDifference = -4.5;
inline constexpr double kExpectedErrorRatePerBase = 0.02;
inline constexpr double kLog10QualityPerBase = -4.0;
inline constexpr double kMaximumExpectedErrorPerRead = 2.0;
using TransMatrix = std::array<double, 7>;  // {MM, MI, MD, IM, II, DM, DD}
inline constexpr TransMatrix kOriginalDefault{0.9998, 0.0001, 0.0001, 0.9, 0.1, 0.9, 0.1};

// pairhmm.hpp:83-93
inline void normalize_likelihoods(std::vector<std::vector<double>>& log_likelihoods) {
  for (auto& row : log_likelihoods) {
    if (row.empty()) continue;
    const double cap = *std::max_element(row.begin(), row.end()) + kMaximumBestAltLikelihoodDifference;
    for (auto& v : row)
      if (v < cap) v = cap;
  }
}

// pairhmm.hpp:95-114
template <class Read>
void filter_poorly_modeled_reads(std::vector<Read>& reads,
                                 std::vector<std::vector<double>>& log_likelihoods) {
  std::vector<Read> kept_reads;
  std::vector<std::vector<double>> kept;
  for (std::size_t i = 0; i < log_likelihoods.size(); ++i) {
    const double best = *std::max_element(log_likelihoods[i].begin(), log_likelihoods[i].end());
    const double threshold =
        std::min(kMaximumExpectedErrorPerRead,
                 std::ceil(static_cast<double>(reads[i].seq.size()) * kExpectedErrorRatePerBase)) *
        kLog10QualityPerBase;
    if (best < threshold) {
      kept_reads.push_back(std::move(reads[i]));
      kept.push_back(std::move(log_likelihoods[i]));
    }
  }
  reads.swap(kept_reads);
  log_likelihoods.swap(kept);
}

namespace detail {

template <class S>
std::string_view view(const S& s) {
  return std::string_view(s.data(), s.size());
}

struct Packed {
  std::vector<int64_t> region_read_off{0}, region_hap_off{0}, read_off{0}, hap_off{0};
  std::vector<uint8_t> rs, q, ins, del, gcp, hap;

  bv_region_batch batch() const {
    return bv_region_batch{static_cast<int64_t>(region_read_off.size()) - 1,
                           region_read_off.data(), region_hap_off.data(),
                           static_cast<int64_t>(read_off.size()) - 1, read_off.data(), rs.data(),
                           q.data(), ins.data(), del.data(), gcp.data(),
                           static_cast<int64_t>(hap_off.size()) - 1, hap_off.data(), hap.data()};
  }
};

inline void append(std::vector<uint8_t>& dst, std::string_view s, std::size_t n, const char* what) {
  if (s.size() < n)
    throw PairHMMError(BV_EINVAL, std::string(what) + " is shorter than the read sequence");
  dst.insert(dst.end(), reinterpret_cast<const uint8_t*>(s.data()),
             reinterpret_cast<const uint8_t*>(s.data()) + n);
}

// Appends one region (a window's haplotypes and reads) to a batch.
template <class Hap, class Read>
void append_region(Packed& p, const std::vector<Hap>& haps, const std::vector<Read>& reads, bool gaps) {
  for (const auto& r : reads) {
    const auto seq = view(r.seq);
    const std::size_t n = seq.size();
    append(p.rs, seq, n, "seq");
    append(p.q, view(r.qual), n, "qual");
    if (gaps) {
      append(p.ins, view(r.insertion_gop()), n, "insertion_gop");
      append(p.del, view(r.deletion_gop()), n, "deletion_gop");
      append(p.gcp, view(r.overall_gcp()), n, "overall_gcp");
    } else {
      p.ins.insert(p.ins.end(), n, 0);
      p.del.insert(p.del.end(), n, 0);
      p.gcp.insert(p.gcp.end(), n, 0);
    }
    p.read_off.push_back(static_cast<int64_t>(p.rs.size()));
  }
  for (const auto& h : haps) {
    const auto s = view(h.seq);
    p.hap.insert(p.hap.end(), reinterpret_cast<const uint8_t*>(s.data()),
                 reinterpret_cast<const uint8_t*>(s.data()) + s.size());
    p.hap_off.push_back(static_cast<int64_t>(p.hap.size()));
  }
  p.region_read_off.push_back(static_cast<int64_t>(p.read_off.size()) - 1);
  p.region_hap_off.push_back(static_cast<int64_t>(p.hap_off.size()) - 1);
}

template <class Hap, class Read>
Packed pack_region(const std::vector<Hap>& haps, const std::vector<Read>& reads, bool gaps) {
  Packed p;
  append_region(p, haps, reads, gaps);
  return p;
}

struct PlanGuard {
  bv_plan* plan = nullptr;
  ~PlanGuard() {
    if (plan) bv_plan_destroy(plan);
  }
};

// Every window as one region of one plan: one run, then normalise and
// filter on the device (bv_plan_postprocess); each window's reads are
// filtered in place and its kept rows returned, in input order.
// A loaded band model and how a plan uses it.
struct BandSetup {
  std::shared_ptr<bv_band_model> model;
  int margin = 10, max_width = 68;
};

inline void set_band(bv_plan* plan, const BandSetup* band) {
  if (band && band->model) check(bv_plan_set_band_model(plan, band->model.get(), band->margin, band->max_width));
}

template <class Hap, class Read>
std::vector<std::vector<std::vector<double>>> run_windows(
    const std::vector<std::pair<const std::vector<Hap>*, std::vector<Read>*>>& windows,
    const bv_options& opt, const BandSetup* band = nullptr) {
  Packed p;
  for (const auto& [haps, reads] : windows) {
    if (!haps || !reads) throw PairHMMError(BV_EINVAL, "a window has a null haplotype or read list");
    if (!reads->empty() && haps->empty())
      throw PairHMMError(BV_EINVAL, "a window has reads but no haplotypes");
    append_region(p, *haps, *reads, opt.mode == BV_MODE_AVX);
  }
  std::vector<std::vector<std::vector<double>>> out(windows.size());
  const auto batch = p.batch();
  if (batch.n_reads == 0) return out;
  PlanGuard g;
  check(bv_plan_create(&batch, &opt, &g.plan));
  set_band(g.plan, band);
  check(bv_plan_run(g.plan));
  bv_plan_stats st{};
  check(bv_plan_stats_get(g.plan, &st));
  std::vector<uint8_t> keep(static_cast<std::size_t>(batch.n_reads));
  std::vector<double> norm(static_cast<std::size_t>(st.n_pairs));
  check(bv_plan_postprocess(g.plan, BV_POST_NORMALIZE, nullptr, nullptr, keep.data(), norm.data(), nullptr));
  std::size_t r_glob = 0, pair = 0;
  for (std::size_t k = 0; k < windows.size(); ++k) {
    auto& reads = *windows[k].second;
    const std::size_t nh = windows[k].first->size();
    std::vector<Read> kept;
    for (std::size_t i = 0; i < reads.size(); ++i) {
      if (keep[r_glob + i]) {
        const auto row = norm.begin() + static_cast<std::ptrdiff_t>(pair + i * nh);
        out[k].emplace_back(row, row + static_cast<std::ptrdiff_t>(nh));
        kept.push_back(std::move(reads[i]));
      }
    }
    pair += reads.size() * nh;
    r_glob += reads.size();
    reads.swap(kept);
  }
  return out;
}

}  // namespace detail

// Raw log10 likelihoods [read][hap] of one region, before normalisation.
template <class Hap, class Read>
std::vector<std::vector<double>> raw_likelihoods(const std::vector<Hap>& haps,
                                                 const std::vector<Read>& reads,
                                                 const bv_options& opt,
                                                 const detail::BandSetup* band = nullptr) {
  std::vector<std::vector<double>> out(reads.size(), std::vector<double>(haps.size(), 0.0));
  if (reads.empty() || haps.empty()) return out;
  const auto packed = detail::pack_region(haps, reads, opt.mode == BV_MODE_AVX);
  const auto batch = packed.batch();
  std::vector<double> flat(reads.size() * haps.size());
  if (band && band->model) {
    detail::PlanGuard g;
    check(bv_plan_create(&batch, &opt, &g.plan));
    detail::set_band(g.plan, band);
    check(bv_plan_run(g.plan));
    check(bv_plan_results(g.plan, flat.data(), nullptr));
  } else {
    check(bv_pairhmm_batch(&batch, &opt, flat.data(), nullptr));
  }
  for (std::size_t i = 0; i < reads.size(); ++i)
    std::copy_n(flat.begin() + static_cast<std::ptrdiff_t>(i * haps.size()), haps.size(), out[i].begin());
  return out;
}

class GpuPairHMM {
 public:
  GpuPairHMM() = default;
  explicit GpuPairHMM(bool fp64_renorm) : renorm_(fp64_renorm) {}
  // The AVX model with learned bands (loads the model now; throws
  // PairHMMError when the file is missing or fails its checks).
  explicit GpuPairHMM(const LearnedBand& band, bool fp64_renorm = true) : renorm_(fp64_renorm) {
    bv_band_model* m = nullptr;
    check(bv_band_model_load(band.model_path.empty() ? nullptr : band.model_path.c_str(), &m));
    band_.model = std::shared_ptr<bv_band_model>(m, bv_band_model_destroy);
    band_.margin = band.margin;
    band_.max_width = band.max_width;
  }
  bool learned_band() const { return static_cast<bool>(band_.model); }

  // AvxPairHMM::compute_likelihoods: per-base GOP/GCP, FP32 + FP64 re-run.
  template <class Hap, class Read>
  std::vector<std::vector<double>> compute_likelihoods(const std::vector<Hap>& haps,
                                                       std::vector<Read>& reads) const {
    bv_options opt;
    bv_options_default(&opt);
    opt.mode = BV_MODE_AVX;
    opt.fp64_renorm = renorm_ ? 1 : 0;
    auto ll = raw_likelihoods(haps, reads, opt, &band_);
    normalize_likelihoods(ll);
    filter_poorly_modeled_reads(reads, ll);
    return ll;
  }

  // PairHMM::compute_likelihoods with a transition matrix (scalar model, FP64).
  template <class Hap, class Read, class Trans>
  std::vector<std::vector<double>> compute_likelihoods(const std::vector<Hap>& haps,
                                                       std::vector<Read>& reads,
                                                       const Trans& trans) const {
    bv_options opt;
    bv_options_default(&opt);
    opt.mode = BV_MODE_SCALAR;
    opt.fp64_renorm = 0;
    for (int k = 0; k < 7; ++k) opt.trans[k] = trans[k];
    auto ll = raw_likelihoods(haps, reads, opt);
    normalize_likelihoods(ll);
    filter_poorly_modeled_reads(reads, ll);
    return ll;
  }

  // Every window of a HaplotypeCaller run at once: windows[k] = {&haplotypes,
  // &reads}; returns for each window what compute_likelihoods returns for it
  // and filters its reads the same way, from one batched device run.
  template <class Hap, class Read>
  std::vector<std::vector<std::vector<double>>> compute_likelihoods_batch(
      const std::vector<std::pair<const std::vector<Hap>*, std::vector<Read>*>>& windows) const {
    bv_options opt;
    bv_options_default(&opt);
    opt.mode = BV_MODE_AVX;
    opt.fp64_renorm = renorm_ ? 1 : 0;
    return detail::run_windows(windows, opt, &band_);
  }

  // The same for the scalar model with a transition matrix.
  template <class Hap, class Read, class Trans>
  std::vector<std::vector<std::vector<double>>> compute_likelihoods_batch(
      const std::vector<std::pair<const std::vector<Hap>*, std::vector<Read>*>>& windows,
      const Trans& trans) const {
    bv_options opt;
    bv_options_default(&opt);
    opt.mode = BV_MODE_SCALAR;
    opt.fp64_renorm = 0;
    for (int k = 0; k < 7; ++k) opt.trans[k] = trans[k];
    return detail::run_windows(windows, opt);
  }

 private:
  bool renorm_ = true;
  detail::BandSetup band_;  // learned bands for the AVX model (empty: full matrix)
};

}  // namespace biovoltron_amd
