// Deterministic synthetic region generator for the PairHMM workloads.
//
// Every region draws from its own std::mt19937_64 stream, seeded from
// (seed, workload, region index) by splitmix64, so any subset of regions can
// be regenerated on its own and the output does not depend on thread count.
// Integer and real ranges use our own mappings on the raw 64-bit draws (the
// standard distributions are implementation-defined).
//
// Workloads (lengths in bp, qualities as raw Phred bytes, no +33 offset, the
// way the reference kernel reads testcase bytes: avx-pairhmm-template.h:110-125):
//   0  one region: 8 haplotypes, variants of one 300 bp base (1% SNPs, 0.2%
//      1-5 bp indels); 64 reads of 150 bp from random haplotypes at uniform
//      offsets, 1% substitutions; base qual U[20,40], GOP 45/45, GCP 10.
//   1  per region: 2-32 haplotypes of a U[100,600] base; 16-512 reads of
//      U[50,250] bp (clamped to the source haplotype), per-read substitution
//      rate U[0.5%,2%]; qualities as workload 0.
//   2  per region: 4-8 haplotypes of a U[3000,12000] base; 4-16 reads of
//      U[2000,10000] bp (clamped), per-read error rate U[5%,10%] of which 60%
//      are 1 bp indels; base qual U[5,30].
//   4  mixture: each region is workload 2 with probability 1/10 (decided by a
//      hash of the region index), else workload 1.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Rng {
  std::mt19937_64 eng;
  explicit Rng(uint64_t s) : eng(s) {}
  uint64_t next() { return eng(); }
  // uniform integer in [lo, hi] by the 128-bit multiply-high mapping
  int64_t range(int64_t lo, int64_t hi) {
    const uint64_t n = static_cast<uint64_t>(hi - lo) + 1u;
    const unsigned __int128 p = static_cast<unsigned __int128>(next()) * n;
    return lo + static_cast<int64_t>(p >> 64);
  }
  // uniform real in [0, 1) from the top 53 bits
  double unit() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

constexpr char kBases[4] = {'A', 'C', 'G', 'T'};

struct Spec {
  int n_haps_lo, n_haps_hi;
  int hap_len_lo, hap_len_hi;
  int n_reads_lo, n_reads_hi;
  int read_len_lo, read_len_hi;
  double err_lo, err_hi;   // per-read error rate range
  double indel_share;      // fraction of read errors that are 1 bp indels
  int qual_lo, qual_hi;
  double snp_rate, indel_rate;  // haplotype variant model
  int indel_len_hi;             // haplotype indel length U[1, hi]
};

Spec spec_for(int w) {
  switch (w) {
    case 0: return {8, 8, 300, 300, 64, 64, 150, 150, 0.01, 0.01, 0.0, 20, 40, 0.01, 0.002, 5};
    case 1: return {2, 32, 100, 600, 16, 512, 50, 250, 0.005, 0.02, 0.0, 20, 40, 0.01, 0.002, 5};
    case 2: return {4, 8, 3000, 12000, 4, 16, 2000, 10000, 0.05, 0.10, 0.6, 5, 30, 0.01, 0.002, 5};
    default: return spec_for(1);
  }
}

struct Region {
  int workload = 0;
  std::vector<std::string> haps;
  std::vector<std::string> reads;
  std::vector<std::string> quals;
};

char other_base(Rng& g, char b) {
  char c;
  do {
    c = kBases[g.range(0, 3)];
  } while (c == b);
  return c;
}

void make_region(int workload, uint64_t seed, int64_t index, Region& out) {
  int w = workload;
  if (workload == 4) {
    const uint64_t h = splitmix64(seed ^ (0xC0FFEEull * static_cast<uint64_t>(index + 1)));
    w = (h % 10u == 0u) ? 2 : 1;
  }
  out.workload = w;
  const Spec s = spec_for(w);
  Rng g(splitmix64(seed ^ splitmix64(static_cast<uint64_t>(w) * 0x100000001B3ull +
                                     static_cast<uint64_t>(index))));
  const int nh = static_cast<int>(g.range(s.n_haps_lo, s.n_haps_hi));
  const int blen = static_cast<int>(g.range(s.hap_len_lo, s.hap_len_hi));
  std::string base(blen, 'A');
  for (auto& c : base) c = kBases[g.range(0, 3)];
  out.haps.assign(nh, std::string());
  for (int h = 0; h < nh; ++h) {
    std::string& v = out.haps[h];
    v.reserve(blen + 16);
    for (int p = 0; p < blen;) {
      const double u = g.unit();
      if (u < s.indel_rate) {
        const int len = static_cast<int>(g.range(1, s.indel_len_hi));
        if (g.unit() < 0.5) {
          for (int k = 0; k < len; ++k) v.push_back(kBases[g.range(0, 3)]);
          v.push_back(base[p]);
          ++p;
        } else {
          p += len;
        }
      } else if (u < s.indel_rate + s.snp_rate) {
        v.push_back(other_base(g, base[p]));
        ++p;
      } else {
        v.push_back(base[p]);
        ++p;
      }
    }
    if (v.empty()) v.push_back(kBases[g.range(0, 3)]);
  }
  const int nr = static_cast<int>(g.range(s.n_reads_lo, s.n_reads_hi));
  out.reads.assign(nr, std::string());
  out.quals.assign(nr, std::string());
  for (int r = 0; r < nr; ++r) {
    const std::string& src = out.haps[g.range(0, nh - 1)];
    const int hl = static_cast<int>(src.size());
    int len = static_cast<int>(g.range(s.read_len_lo, s.read_len_hi));
    len = std::min(len, hl);
    const int off = static_cast<int>(g.range(0, hl - len));
    const double err = s.err_lo + (s.err_hi - s.err_lo) * g.unit();
    std::string& rd = out.reads[r];
    rd.reserve(len + 16);
    for (int p = 0; p < len; ++p) {
      const char b = src[off + p];
      if (g.unit() < err) {
        if (g.unit() < s.indel_share) {
          if (g.unit() < 0.5) {  // 1 bp insertion before b
            rd.push_back(kBases[g.range(0, 3)]);
            rd.push_back(b);
          }  // else 1 bp deletion of b
        } else {
          rd.push_back(other_base(g, b));
        }
      } else {
        rd.push_back(b);
      }
    }
    if (rd.empty()) rd.push_back(src[off]);
    std::string& qs = out.quals[r];
    qs.resize(rd.size());
    for (auto& c : qs) c = static_cast<char>(g.range(s.qual_lo, s.qual_hi));
  }
}

}  // namespace

extern "C" {

// Packed structure-of-arrays batch; all arrays are owned by the library and
// released by bv_synth_free.  Offsets have one more entry than items.
struct bv_synth_batch {
  int64_t n_regions, n_reads, n_haps, n_pairs, read_bytes, hap_bytes, cells;
  int64_t* region_read_off;  // [n_regions+1] read index range per region
  int64_t* region_hap_off;   // [n_regions+1] haplotype index range per region
  int64_t* region_pair_off;  // [n_regions+1] first output pair per region
  uint8_t* region_workload;  // [n_regions]
  int64_t* read_off;         // [n_reads+1] byte offsets
  uint8_t* read_bases;
  uint8_t* read_qual;
  uint8_t* read_ins;
  uint8_t* read_del;
  uint8_t* read_gcp;
  int64_t* hap_off;          // [n_haps+1]
  uint8_t* hap_bases;
};

// Generates regions [first, first+count) of a workload.  Returns NULL on
// allocation failure.
bv_synth_batch* bv_synth_generate(int workload, uint64_t seed, int64_t first,
                                  int64_t count, int n_threads) {
  if (count < 0) return nullptr;
  std::vector<Region> regs(static_cast<size_t>(count));
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads)
  for (int64_t k = 0; k < count; ++k) make_region(workload, seed, first + k, regs[k]);

  auto* b = static_cast<bv_synth_batch*>(std::calloc(1, sizeof(bv_synth_batch)));
  if (!b) return nullptr;
  b->n_regions = count;
  int64_t nreads = 0, nhaps = 0, rbytes = 0, hbytes = 0, npairs = 0, cells = 0;
  for (const auto& r : regs) {
    nreads += static_cast<int64_t>(r.reads.size());
    nhaps += static_cast<int64_t>(r.haps.size());
    int64_t hsum = 0;
    for (const auto& h : r.haps) hsum += static_cast<int64_t>(h.size());
    hbytes += hsum;
    for (const auto& rd : r.reads) {
      rbytes += static_cast<int64_t>(rd.size());
      cells += static_cast<int64_t>(rd.size()) * hsum;
    }
    npairs += static_cast<int64_t>(r.reads.size()) * static_cast<int64_t>(r.haps.size());
  }
  b->n_reads = nreads;
  b->n_haps = nhaps;
  b->n_pairs = npairs;
  b->read_bytes = rbytes;
  b->hap_bytes = hbytes;
  b->cells = cells;
  b->region_read_off = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (count + 1)));
  b->region_hap_off = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (count + 1)));
  b->region_pair_off = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (count + 1)));
  b->region_workload = static_cast<uint8_t*>(std::malloc(count + 1));
  b->read_off = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (nreads + 1)));
  b->hap_off = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (nhaps + 1)));
  b->read_bases = static_cast<uint8_t*>(std::malloc(rbytes + 1));
  b->read_qual = static_cast<uint8_t*>(std::malloc(rbytes + 1));
  b->read_ins = static_cast<uint8_t*>(std::malloc(rbytes + 1));
  b->read_del = static_cast<uint8_t*>(std::malloc(rbytes + 1));
  b->read_gcp = static_cast<uint8_t*>(std::malloc(rbytes + 1));
  b->hap_bases = static_cast<uint8_t*>(std::malloc(hbytes + 1));
  if (!b->region_read_off || !b->region_hap_off || !b->region_pair_off ||
      !b->region_workload || !b->read_off || !b->hap_off || !b->read_bases ||
      !b->read_qual || !b->read_ins || !b->read_del || !b->read_gcp || !b->hap_bases)
    return nullptr;
  int64_t ri = 0, hi = 0, ro = 0, ho = 0, po = 0;
  for (int64_t k = 0; k < count; ++k) {
    const Region& r = regs[k];
    b->region_read_off[k] = ri;
    b->region_hap_off[k] = hi;
    b->region_pair_off[k] = po;
    b->region_workload[k] = static_cast<uint8_t>(r.workload);
    for (const auto& h : r.haps) {
      b->hap_off[hi++] = ho;
      std::memcpy(b->hap_bases + ho, h.data(), h.size());
      ho += static_cast<int64_t>(h.size());
    }
    for (size_t t = 0; t < r.reads.size(); ++t) {
      const std::string& rd = r.reads[t];
      const size_t n = rd.size();
      b->read_off[ri++] = ro;
      std::memcpy(b->read_bases + ro, rd.data(), n);
      std::memcpy(b->read_qual + ro, r.quals[t].data(), n);
      std::memset(b->read_ins + ro, 45, n);
      std::memset(b->read_del + ro, 45, n);
      std::memset(b->read_gcp + ro, 10, n);
      ro += static_cast<int64_t>(n);
    }
    po += static_cast<int64_t>(r.reads.size()) * static_cast<int64_t>(r.haps.size());
  }
  b->region_read_off[count] = ri;
  b->region_hap_off[count] = hi;
  b->region_pair_off[count] = po;
  b->read_off[nreads] = ro;
  b->hap_off[nhaps] = ho;
  return b;
}

void bv_synth_free(bv_synth_batch* b) {
  if (!b) return;
  std::free(b->region_read_off);
  std::free(b->region_hap_off);
  std::free(b->region_pair_off);
  std::free(b->region_workload);
  std::free(b->read_off);
  std::free(b->hap_off);
  std::free(b->read_bases);
  std::free(b->read_qual);
  std::free(b->read_ins);
  std::free(b->read_del);
  std::free(b->read_gcp);
  std::free(b->hap_bases);
  std::free(b);
}

}  // extern "C"
