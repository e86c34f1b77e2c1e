// Exercises include/biovoltron_amd/gpu_pairhmm.hpp the way HaplotypeCaller
// uses PairHMM: a const member called per region with SamRecord-like reads.
//   test_adapter cpu   normalisation/filtering known answers, error paths
//   test_adapter gpu   prints the likelihood matrix of a small region
//   test_adapter band F  windows from file F with and without the learned band
//   test_adapter shacc F windows from file F through shacc_pairhmm::calculate
#include <biovoltron_amd/gpu_pairhmm.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

struct Read {  // the SamRecord members PairHMM reads (sam.hpp:186-259, :451-475)
  std::string seq, qual, gop, gcp;
  std::string_view insertion_gop() const { return gop; }
  std::string_view deletion_gop() const { return gop; }
  std::string_view overall_gcp() const { return gcp; }
};
struct Hap {
  std::string seq;
};

static int fails = 0;
#define EXPECT(c)                                                  \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);      \
      ++fails;                                                     \
    }                                                              \
  } while (0)

static Read make_read(const std::string& s, char q = 30, char gop = 45, char gcp = 10) {
  return Read{s, std::string(s.size(), q), std::string(s.size(), gop), std::string(s.size(), gcp)};
}

int cpu_tests() {
  using namespace biovoltron_amd;
  // normalize: cap at row best - 4.5
  std::vector<std::vector<double>> ll{{-1.0, -10.0, -5.0}, {-20.0, -30.0, -22.0}};
  normalize_likelihoods(ll);
  EXPECT(ll[0][1] == -5.5 && ll[0][2] == -5.0 && ll[0][0] == -1.0);
  EXPECT(ll[1][1] == -24.5 && ll[1][2] == -22.0);
  // filter (as the reference is written): keep rows whose best is below
  // min(2, ceil(0.02*len)) * -4
  std::vector<Read> reads{make_read(std::string(150, 'A')), make_read(std::string(150, 'C')),
                          make_read(std::string(20, 'G'))};
  std::vector<std::vector<double>> l2{{-3.0, -9.0}, {-9.0, -12.0}, {-4.5, -6.0}};
  filter_poorly_modeled_reads(reads, l2);
  EXPECT(reads.size() == 2 && l2.size() == 2);
  EXPECT(reads[0].seq[0] == 'C' && reads[1].seq[0] == 'G');  // 150bp: thr -8; 20bp: thr -4
  // error paths: short GOP string, empty read -> exceptions, not UB
  const GpuPairHMM hmm;
  std::vector<Hap> haps{{"ACGTACGT"}};
  std::vector<Read> bad{Read{"ACGT", "????", "--", "++++"}};
  bool threw = false;
  try {
    hmm.compute_likelihoods(haps, bad);
  } catch (const PairHMMError& e) {
    threw = std::strstr(e.what(), "insertion_gop") != nullptr;
  }
  EXPECT(threw);
  std::vector<Read> empty{make_read("")};
  threw = false;
  try {
    hmm.compute_likelihoods(haps, empty);
  } catch (const PairHMMError& e) {
    threw = e.status == BV_EINVAL;
  }
  EXPECT(threw);
  // batched windows: a window with reads but no haplotypes is an error
  // (raised before any device call), windows without reads give no rows
  std::vector<Hap> no_haps;
  std::vector<Read> some{make_read("ACGTACGT")}, none;
  using W = std::pair<const std::vector<Hap>*, std::vector<Read>*>;
  threw = false;
  try {
    hmm.compute_likelihoods_batch(std::vector<W>{{&haps, &none}, {&no_haps, &some}});
  } catch (const PairHMMError& e) {
    threw = e.status == BV_EINVAL && std::strstr(e.what(), "no haplotypes") != nullptr;
  }
  EXPECT(threw);
  const auto nothing = hmm.compute_likelihoods_batch(std::vector<W>{{&haps, &none}, {&no_haps, &none}});
  EXPECT(nothing.size() == 2 && nothing[0].empty() && nothing[1].empty());
  std::printf("cpu tests: %s\n", fails ? "FAILED" : "ok");
  return fails ? 1 : 0;
}

int gpu_run() {
  using namespace biovoltron_amd;
  const std::string base = "AATCGAAGGTCGTAAGGACACGGTTGAGCGTTCAGCGTTCATGTGAGTCCTCACCACTTATGGCTCCATAGCC";
  std::vector<Hap> haps{{base}, {base.substr(0, 30) + "T" + base.substr(31)}, {base.substr(5)}};
  std::vector<Read> reads{make_read(base.substr(10, 40), 35), make_read(base.substr(0, 50), 25),
                          make_read(base.substr(20, 30) + "GG", 20)};
  const GpuPairHMM hmm;
  const auto raw = raw_likelihoods(haps, reads, [] {
    bv_options o;
    bv_options_default(&o);
    return o;
  }());
  for (const auto& row : raw) {
    for (double v : row) std::printf("%.17g ", v);
    std::printf("\n");
  }
  auto copy = reads;
  const auto ll = hmm.compute_likelihoods(haps, copy);
  std::printf("kept %zu\n", ll.size());
  const auto sc = hmm.compute_likelihoods(haps, reads, kOriginalDefault);
  std::printf("scalar kept %zu\n", sc.size());

  // HaplotypeCaller batching: windows of different shapes in one device
  // run equal one compute_likelihoods call per window, reads filtered alike
  // (both models; rows compared bit for bit).  Each window's inputs and the
  // batch's kept rows are printed for the oracle check in the test driver.
  std::vector<std::vector<Hap>> whaps;
  std::vector<std::vector<Read>> wreads;
  for (int w = 0; w < 6; ++w) {
    const std::string b = base.substr(static_cast<std::size_t>(w), 60);
    std::vector<Hap> hs{{b}, {b.substr(0, 25) + "A" + b.substr(26)}, {b.substr(0, 40) + b.substr(43)}};
    hs.resize(static_cast<std::size_t>(1 + w % 3));
    std::vector<Read> rs;
    if (w != 4) {  // window 4 has haplotypes but no reads
      for (int i = 0; i < 2 + w; ++i)
        rs.push_back(make_read(b.substr(static_cast<std::size_t>(i * 3), 30 + 4 * i), static_cast<char>(15 + 5 * i)));
      rs.push_back(make_read(std::string(20, 'T'), 40));  // poorly modelled: its row is kept (filter as written)
    }
    whaps.push_back(hs);
    wreads.push_back(rs);
  }
  for (int model = 0; model < 2; ++model) {
    auto batch_reads = wreads;
    std::vector<std::pair<const std::vector<Hap>*, std::vector<Read>*>> windows;
    for (std::size_t w = 0; w < whaps.size(); ++w) windows.emplace_back(&whaps[w], &batch_reads[w]);
    const auto all = model == 0 ? hmm.compute_likelihoods_batch(windows)
                                : hmm.compute_likelihoods_batch(windows, kOriginalDefault);
    bool same = all.size() == whaps.size();
    std::size_t rows = 0;
    for (std::size_t w = 0; same && w < whaps.size(); ++w) {
      auto one_reads = wreads[w];
      const auto one = model == 0 ? hmm.compute_likelihoods(whaps[w], one_reads)
                                  : hmm.compute_likelihoods(whaps[w], one_reads, kOriginalDefault);
      same = one == all[w] && one_reads.size() == batch_reads[w].size();
      for (std::size_t i = 0; same && i < one_reads.size(); ++i) same = one_reads[i].seq == batch_reads[w][i].seq;
      rows += one.size();
    }
    std::printf("%s batch %s (%zu rows)\n", model == 0 ? "avx" : "scalar", same ? "equal" : "DIFFERENT", rows);
    for (std::size_t w = 0; w < whaps.size(); ++w) {
      std::printf("window %d %zu %zu %zu\n", model, w, whaps[w].size(), wreads[w].size());
      for (const auto& h : whaps[w]) std::printf("hap %s\n", h.seq.c_str());
      for (const auto& r : wreads[w]) std::printf("read %s %d\n", r.seq.c_str(), static_cast<int>(r.qual[0]));
      for (std::size_t i = 0; i < all[w].size(); ++i) {
        std::printf("row %s", batch_reads[w][i].seq.c_str());
        for (double v : all[w][i]) std::printf(" %.17g", v);
        std::printf("\n");
      }
    }
  }
  return 0;
}

// Windows from a text file ("W n_haps n_reads", then "H seq" per haplotype
// and "R seq q0,q1,... gop gcp" per read); false on a malformed file.
static bool read_windows(const char* path, std::vector<std::vector<Hap>>& whaps,
                         std::vector<std::vector<Read>>& wreads) {
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  bool ok = true;
  char tag[4];
  while (ok && std::fscanf(f, "%3s", tag) == 1) {
    if (tag[0] == 'W') {
      int nh = 0, nr = 0;
      ok = std::fscanf(f, "%d %d", &nh, &nr) == 2;
      whaps.emplace_back();
      wreads.emplace_back();
    } else if (tag[0] == 'H' && !whaps.empty()) {
      static char buf[1 << 16];
      ok = std::fscanf(f, "%65535s", buf) == 1;
      if (ok) whaps.back().push_back(Hap{buf});
    } else if (tag[0] == 'R' && !wreads.empty()) {
      static char seq[1 << 16], quals[1 << 19];
      int gop = 0, gcp = 0;
      ok = std::fscanf(f, "%65535s %524287s %d %d", seq, quals, &gop, &gcp) == 4;
      if (!ok) break;
      Read r = make_read(seq, 0, static_cast<char>(gop), static_cast<char>(gcp));
      std::size_t k = 0;
      for (char* t = std::strtok(quals, ","); t && k < r.qual.size(); t = std::strtok(nullptr, ","))
        r.qual[k++] = static_cast<char>(std::atoi(t));
      wreads.back().push_back(r);
    } else {
      ok = false;
    }
  }
  std::fclose(f);
  return ok;
}

// The windows of a file run with the learned band and without; prints each
// window's kept rows for both, and the plan statistics of a learned-band run
// through the C ABI on the same windows.
int band_run(const char* path) {
  using namespace biovoltron_amd;
  std::vector<std::vector<Hap>> whaps;
  std::vector<std::vector<Read>> wreads;
  if (!read_windows(path, whaps, wreads) || whaps.empty()) return 3;
  const GpuPairHMM learned(LearnedBand{}), plain;
  std::printf("learned %d\n", learned.learned_band() ? 1 : 0);
  for (int model = 0; model < 2; ++model) {
    auto reads = wreads;
    std::vector<std::pair<const std::vector<Hap>*, std::vector<Read>*>> windows;
    for (std::size_t w = 0; w < whaps.size(); ++w) windows.emplace_back(&whaps[w], &reads[w]);
    const auto all = (model == 0 ? learned : plain).compute_likelihoods_batch(windows);
    for (std::size_t w = 0; w < all.size(); ++w) {
      std::printf("bwindow %d %zu %zu\n", model, w, all[w].size());
      for (std::size_t i = 0; i < all[w].size(); ++i) {
        std::printf("brow %s", reads[w][i].seq.c_str());
        for (double v : all[w][i]) std::printf(" %.17g", v);
        std::printf("\n");
      }
    }
  }
  // one window through compute_likelihoods with the learned band
  auto one = wreads[0];
  const auto ll = learned.compute_likelihoods(whaps[0], one);
  std::printf("single %zu\n", ll.size());
  // the same windows through the C ABI: how many pairs took a band
  detail::Packed pk;
  for (std::size_t w = 0; w < whaps.size(); ++w) detail::append_region(pk, whaps[w], wreads[w], true);
  const auto batch = pk.batch();
  bv_options o;
  bv_options_default(&o);
  detail::PlanGuard g;
  check(bv_plan_create(&batch, &o, &g.plan));
  bv_band_model* m = nullptr;
  check(bv_band_model_load(nullptr, &m));
  check(bv_plan_set_band_model(g.plan, m, 10, 68));
  bv_band_model_destroy(m);
  check(bv_plan_run(g.plan));
  bv_plan_stats st{};
  check(bv_plan_stats_get(g.plan, &st));
  std::printf("bstats %lld %lld %lld\n", static_cast<long long>(st.n_pairs), static_cast<long long>(st.band_pairs),
              static_cast<long long>(st.fallback_pairs));
  return 0;
}

// The batch accelerator hook as the reference declares it
// (lib/intel_native/pairhmm/shacc_pairhmm.h:10-36): the same layout, bound
// at link time to the library's strong definition.
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
bool calculate(Batch& batch);
}  // namespace shacc_pairhmm

// Each window of a file as one shacc_pairhmm::Batch; prints the results
// (read-major log10 likelihoods) and the hook's answers on an empty batch
// and on a batch holding an empty read.
int shacc_run(const char* path) {
  std::vector<std::vector<Hap>> whaps;
  std::vector<std::vector<Read>> wreads;
  if (!read_windows(path, whaps, wreads) || whaps.empty()) return 3;
  for (std::size_t w = 0; w < whaps.size(); ++w) {
    std::vector<shacc_pairhmm::Read> rs;
    std::vector<shacc_pairhmm::Haplotype> hs;
    long cells = 0;
    for (const auto& r : wreads[w])
      rs.push_back({static_cast<int>(r.seq.size()), r.seq.data(), r.qual.data(), r.gop.data(), r.gop.data(),
                    r.gcp.data()});
    for (const auto& h : whaps[w]) hs.push_back({static_cast<int>(h.seq.size()), h.seq.data()});
    for (const auto& r : rs)
      for (const auto& h : hs) cells += static_cast<long>(r.length) * h.length;
    std::vector<float> res(rs.size() * hs.size(), 0.f);
    shacc_pairhmm::Batch b{static_cast<int>(rs.size()), static_cast<int>(hs.size()), cells, rs.data(), hs.data(),
                           res.data()};
    const bool ok = shacc_pairhmm::calculate(b);
    std::printf("swindow %zu %d %zu %zu\n", w, ok ? 1 : 0, rs.size(), hs.size());
    for (std::size_t i = 0; i < rs.size(); ++i) {
      std::printf("srow");
      for (std::size_t h = 0; h < hs.size(); ++h) std::printf(" %.9g", res[i * hs.size() + h]);
      std::printf("\n");
    }
  }
  shacc_pairhmm::Batch empty{0, 0, 0, nullptr, nullptr, nullptr};
  std::printf("sempty %d\n", shacc_pairhmm::calculate(empty) ? 1 : 0);
  const char* hb = "ACGT";
  shacc_pairhmm::Read zero{0, "", "", "", "", ""};
  shacc_pairhmm::Haplotype one{4, hb};
  float out = 0.f;
  shacc_pairhmm::Batch bad{1, 1, 0, &zero, &one, &out};
  const bool r = shacc_pairhmm::calculate(bad);
  std::printf("sbad %d %s\n", r ? 1 : 0, bv_pairhmm_last_error());
  return 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "gpu") == 0) return gpu_run();
  if (argc > 2 && std::strcmp(argv[1], "band") == 0) return band_run(argv[2]);
  if (argc > 2 && std::strcmp(argv[1], "shacc") == 0) return shacc_run(argv[2]);
  return cpu_tests();
}
