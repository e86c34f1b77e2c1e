// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only,
// built with -fsanitize=address,undefined -fno-sanitize-recover=all): the
// oracle's C restatement, the synthetic workload generator and the band
// model file reader/writer, driven over random and edge-case inputs.  Any
// out-of-bounds access, use after free, leak or undefined operation aborts
// the run with the sanitizer's report; a clean run prints "sanitize: ok".
//
// The SHA-256 of the model files is also checked here against the FIPS
// 180-4 example digests (one-block and two-block messages, and the empty
// message), which the fixed-size model body alone does not reach.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../oracle/pairhmm_oracle.h"

namespace bvp {
void sha256(const void* data, size_t n, uint8_t out[32]);
int read_model_file(const char* path, float* params, int count, std::string& err);
int write_model_file(const char* path, const float* params, int count, std::string& err);
std::string default_model_path();
}  // namespace bvp

extern "C" {
struct bv_synth_batch;
bv_synth_batch* bv_synth_generate(int workload, uint64_t seed, int64_t first, int64_t count, int n_threads);
void bv_synth_free(bv_synth_batch* b);
}

static int fails = 0;
#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++fails;                                                        \
    }                                                                 \
  } while (0)

static std::string hex(const uint8_t* d) {
  std::string s;
  char b[3];
  for (int i = 0; i < 32; ++i) {
    std::snprintf(b, sizeof(b), "%02x", d[i]);
    s += b;
  }
  return s;
}

static void sha_vectors() {
  struct V {
    std::string msg;
    const char* digest;
  };
  // FIPS 180-4 / NIST example messages
  const V vs[] = {
      {"", "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"},
      {"abc", "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"},
      {"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq",
       "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"},
      {"abcdefghbcdefghicdefghijdefghijkefghijklfghijklmghijklmnhijklmnoijklmnopjklmnopqklmnopqrlmnopqrsmnopqrstnopqrstu",
       "cf5b16a778af8380036ce59e7b0492370b249b11e8f07a51afac45037afee9d1"},
  };
  for (const auto& v : vs) {
    uint8_t d[32];
    bvp::sha256(v.msg.data(), v.msg.size(), d);
    CHECK(hex(d) == v.digest);
  }
  // one million 'a' (multi-block body)
  std::string a(1000000, 'a');
  uint8_t d[32];
  bvp::sha256(a.data(), a.size(), d);
  CHECK(hex(d) == "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0");
}

static void model_files() {
  const int n = 6658;
  std::vector<float> p(n), q(n);
  for (int i = 0; i < n; ++i) p[i] = std::sin(0.37f * i);
  std::string err;
  const char* path = "sanitize_model.bin";
  CHECK(bvp::write_model_file(path, p.data(), n, err) == 0);
  CHECK(bvp::read_model_file(path, q.data(), n, err) == 0);
  CHECK(std::memcmp(p.data(), q.data(), sizeof(float) * n) == 0);
  CHECK(bvp::read_model_file(path, q.data(), n - 1, err) != 0);  // wrong count
  CHECK(bvp::read_model_file(nullptr, q.data(), n, err) != 0);
  CHECK(bvp::read_model_file("no_such_model.bin", q.data(), n, err) != 0);
  // truncations at every header boundary and inside the body
  for (long cut : {0L, 3L, 8L, 12L, 15L, 16L, 17L, 100L, 16L + 4L * n, 16L + 4L * n + 31L}) {
    FILE* f = std::fopen(path, "rb");
    std::vector<char> buf(static_cast<size_t>(cut) + 1);
    CHECK(std::fread(buf.data(), 1, static_cast<size_t>(cut), f) == static_cast<size_t>(cut));
    std::fclose(f);
    const char* tp = "sanitize_cut.bin";
    f = std::fopen(tp, "wb");
    std::fwrite(buf.data(), 1, static_cast<size_t>(cut), f);
    std::fclose(f);
    CHECK(bvp::read_model_file(tp, q.data(), n, err) != 0);
    std::remove(tp);
  }
  std::remove(path);
  CHECK(bvp::write_model_file("/nonexistent_dir/x.bin", p.data(), n, err) != 0);
  (void)bvp::default_model_path();
}

static void synth() {
  for (int w : {0, 1, 2, 4}) {
    bv_synth_batch* b = bv_synth_generate(w, 1234 + w, 0, w == 2 ? 1 : 3, 2);
    CHECK(b != nullptr);
    bv_synth_free(b);
  }
  bv_synth_batch* e = bv_synth_generate(1, 5, 7, 0, 1);  // no regions
  bv_synth_free(e);
}

static std::vector<uint8_t> bytes(std::mt19937& g, int n, const char* alphabet) {
  std::vector<uint8_t> v(static_cast<size_t>(n));
  const int k = static_cast<int>(std::strlen(alphabet));
  for (auto& x : v) x = static_cast<uint8_t>(alphabet[g() % k]);
  return v;
}

static void oracle_pairs() {
  std::mt19937 g(7);
  const double trans[7] = {0.9998, 0.0001, 0.0001, 0.9, 0.1, 0.9, 0.1};
  for (int t = 0; t < 200; ++t) {
    const int R = 1 + static_cast<int>(g() % 70), H = 1 + static_cast<int>(g() % 90);
    auto rs = bytes(g, R, "ACGTNacgtX");
    auto hap = bytes(g, H, "ACGTN");
    std::vector<uint8_t> q(R), ins(R), del(R), gcp(R);
    for (int i = 0; i < R; ++i) {
      q[i] = static_cast<uint8_t>(g() % 128);
      ins[i] = static_cast<uint8_t>(g() % 128);
      del[i] = static_cast<uint8_t>(g() % 128);
      gcp[i] = static_cast<uint8_t>(10 + g() % 30);
    }
    (void)bvo_prob_f32(R, H, rs.data(), q.data(), ins.data(), del.data(), gcp.data(), hap.data(), 1);
    (void)bvo_prob_f64(R, H, rs.data(), q.data(), ins.data(), del.data(), gcp.data(), hap.data(), 1);
    int resc = 0;
    const double l = bvo_log10_reference(R, H, rs.data(), q.data(), ins.data(), del.data(), gcp.data(), hap.data(),
                                         1, &resc);
    const double e0 = bvo_log10_exact(R, H, rs.data(), q.data(), ins.data(), del.data(), gcp.data(), hap.data(), 0);
    (void)bvo_log10_exact(R, H, rs.data(), q.data(), ins.data(), del.data(), gcp.data(), hap.data(), 1);
    (void)bvo_scalar_log10(R, H, rs.data(), q.data(), hap.data(), trans);
    CHECK(std::isfinite(e0));
    CHECK(!(std::isfinite(l)) || std::fabs(l - e0) < 1e-2);
    std::vector<double> mass(R + H - 1), rows(R);
    (void)bvo_band_posterior(R, H, rs.data(), q.data(), ins.data(), del.data(), gcp.data(), hap.data(), mass.data(),
                             rows.data());
    double tr = 0.0, lost = 0.0;
    const int width = (g() % 2) ? 32 : 64;
    const int o_lo = -static_cast<int>(g() % (R + 8)) + static_cast<int>(g() % 16);
    (void)bvo_band_forward(R, H, rs.data(), q.data(), ins.data(), del.data(), gcp.data(), hap.data(), o_lo, width, 4,
                           &tr);
    std::vector<int> offs(R + 1), sched(R + 1);
    for (int i = 0; i <= R; ++i) sched[i] = o_lo + (i < 8 ? 8 - i : 0);
    (void)bvo_band_follow_forward(R, H, rs.data(), q.data(), ins.data(), del.data(), gcp.data(), hap.data(), o_lo, 64,
                                  4, 16, 4, 0, 0, 1e-9, 3, nullptr, &tr, &lost, offs.data());
    (void)bvo_band_follow_forward(R, H, rs.data(), q.data(), ins.data(), del.data(), gcp.data(), hap.data(), o_lo, 64,
                                  4, 1, 1, 0, 0, 1e-9, 3, sched.data(), &tr, &lost, offs.data());
    const int32_t params[4] = {200, -150, -260, -11};
    std::vector<char> ops(R + H + 3);
    std::vector<int32_t> lens(R + H + 3);
    int32_t off = 0;
    const int n = bvo_sw_align(R, rs.data(), H, hap.data(), params, &off, ops.data(), lens.data());
    CHECK(n >= 1 && n <= R + H + 3);
  }
  // a region through the threaded driver, then post-processing
  const int nr = 9, nh = 4;
  std::vector<int64_t> ro(nr + 1), ho(nh + 1);
  std::vector<uint8_t> rs, q, ins, del, gcp, hap;
  for (int r = 0; r < nr; ++r) {
    ro[r] = static_cast<int64_t>(rs.size());
    const int R = 1 + r * 13;
    auto b = bytes(g, R, "ACGT");
    rs.insert(rs.end(), b.begin(), b.end());
    q.insert(q.end(), R, 30);
    ins.insert(ins.end(), R, 45);
    del.insert(del.end(), R, 45);
    gcp.insert(gcp.end(), R, 10);
  }
  ro[nr] = static_cast<int64_t>(rs.size());
  for (int h = 0; h < nh; ++h) {
    ho[h] = static_cast<int64_t>(hap.size());
    auto b = bytes(g, 20 + 30 * h, "ACGT");
    hap.insert(hap.end(), b.begin(), b.end());
  }
  ho[nh] = static_cast<int64_t>(hap.size());
  std::vector<double> L(nr * nh), Ln(nr * nh), A(nr * 2);
  (void)bvo_region_log10(nr, ro.data(), rs.data(), q.data(), ins.data(), del.data(), gcp.data(), nh, ho.data(),
                         hap.data(), 2, L.data());
  std::vector<int64_t> rl(nr);
  for (int r = 0; r < nr; ++r) rl[r] = ro[r + 1] - ro[r];
  const int32_t ha[nh] = {0, 1, 0, 1};
  std::vector<uint8_t> keep(nr);
  bvo_postprocess_region(nr, nh, L.data(), rl.data(), 1, ha, 2, keep.data(), Ln.data(), A.data());
}

int main() {
  bvo_init();
  sha_vectors();
  model_files();
  synth();
  oracle_pairs();
  if (fails) {
    std::printf("sanitize: %d failures\n", fails);
    return 1;
  }
  std::printf("sanitize: ok\n");
  return 0;
}
