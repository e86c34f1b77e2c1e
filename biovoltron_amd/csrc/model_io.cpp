// Band model files (host code, no device calls): the trained MLP parameters
// in a small little-endian binary the C++ drop-in can load without Python.
//
//   offset  size        field
//   0       8           magic "BVBANDM" followed by a zero byte
//   8       4           uint32 format version (1)
//   12      4           uint32 parameter count (BV_BAND_PARAMS)
//   16      4 * count   float32 parameters (the bv_band_model_get_params order)
//   16+4c   32          SHA-256 of the parameter bytes
//
// SHA-256 is written from FIPS 180-4 (section 6.2); the digest of the
// parameter bytes is the same one band.py stores in its .npz files.
#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bv_pairhmm.h"

namespace bvp {

namespace {

constexpr uint32_t kK256[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void sha256_block(uint32_t h[8], const uint8_t* p) {
  uint32_t w[64];
  for (int t = 0; t < 16; ++t)
    w[t] = (uint32_t(p[4 * t]) << 24) | (uint32_t(p[4 * t + 1]) << 16) | (uint32_t(p[4 * t + 2]) << 8) |
           uint32_t(p[4 * t + 3]);
  for (int t = 16; t < 64; ++t) {
    const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
    const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
    w[t] = w[t - 16] + s0 + w[t - 7] + s1;
  }
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
  for (int t = 0; t < 64; ++t) {
    const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
    const uint32_t ch = (e & f) ^ (~e & g);
    const uint32_t t1 = k + S1 + ch + kK256[t] + w[t];
    const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
    const uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
    const uint32_t t2 = S0 + maj;
    k = g;
    g = f;
    f = e;
    e = d + t1;
    d = c;
    c = b;
    b = a;
    a = t1 + t2;
  }
  h[0] += a;
  h[1] += b;
  h[2] += c;
  h[3] += d;
  h[4] += e;
  h[5] += f;
  h[6] += g;
  h[7] += k;
}

constexpr char kMagic[8] = {'B', 'V', 'B', 'A', 'N', 'D', 'M', '\0'};
constexpr uint32_t kVersion = 1;

void put32(std::vector<uint8_t>& v, uint32_t x) {
  for (int i = 0; i < 4; ++i) v.push_back(static_cast<uint8_t>(x >> (8 * i)));
}

uint32_t get32(const uint8_t* p) {
  return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16) | (uint32_t(p[3]) << 24);
}

}  // namespace

void sha256(const void* data, size_t n, uint8_t out[32]) {
  uint32_t h[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
  const uint8_t* p = static_cast<const uint8_t*>(data);
  size_t i = 0;
  for (; i + 64 <= n; i += 64) sha256_block(h, p + i);
  uint8_t tail[128] = {0};
  const size_t rem = n - i;
  std::memcpy(tail, p + i, rem);
  tail[rem] = 0x80;
  const size_t len = rem + 9 <= 64 ? 64 : 128;
  const uint64_t bits = static_cast<uint64_t>(n) * 8;
  for (int b = 0; b < 8; ++b) tail[len - 1 - b] = static_cast<uint8_t>(bits >> (8 * b));
  sha256_block(h, tail);
  if (len == 128) sha256_block(h, tail + 64);
  for (int k = 0; k < 8; ++k)
    for (int b = 0; b < 4; ++b) out[4 * k + b] = static_cast<uint8_t>(h[k] >> (24 - 8 * b));
}

// Parses and checks a model file; on success fills params (count floats).
int read_model_file(const char* path, float* params, int count, std::string& err) {
  if (!path) {
    err = "model path is NULL";
    return BV_EINVAL;
  }
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    err = std::string("cannot open band model file ") + path;
    return BV_EINVAL;
  }
  std::vector<uint8_t> buf;
  uint8_t chunk[65536];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  std::fclose(f);
  const size_t want = 16 + 4 * static_cast<size_t>(count) + 32;
  if (buf.size() < 16 || std::memcmp(buf.data(), kMagic, 8) != 0) {
    err = std::string(path) + " is not a band model file";
    return BV_EINVAL;
  }
  if (get32(buf.data() + 8) != kVersion) {
    err = std::string(path) + ": band model format " + std::to_string(get32(buf.data() + 8)) + ", expected " +
          std::to_string(kVersion);
    return BV_EINVAL;
  }
  if (get32(buf.data() + 12) != static_cast<uint32_t>(count) || buf.size() != want) {
    err = std::string(path) + ": wrong parameter count or file size";
    return BV_EINVAL;
  }
  uint8_t digest[32];
  sha256(buf.data() + 16, 4 * static_cast<size_t>(count), digest);
  if (std::memcmp(digest, buf.data() + 16 + 4 * static_cast<size_t>(count), 32) != 0) {
    err = std::string(path) + ": parameters do not match their SHA-256";
    return BV_EINVAL;
  }
  for (int k = 0; k < count; ++k) {
    const uint32_t bits = get32(buf.data() + 16 + 4 * static_cast<size_t>(k));
    std::memcpy(params + k, &bits, 4);
  }
  return BV_OK;
}

int write_model_file(const char* path, const float* params, int count, std::string& err) {
  std::vector<uint8_t> buf(kMagic, kMagic + 8);
  put32(buf, kVersion);
  put32(buf, static_cast<uint32_t>(count));
  for (int k = 0; k < count; ++k) {
    uint32_t bits;
    std::memcpy(&bits, params + k, 4);
    put32(buf, bits);
  }
  uint8_t digest[32];
  sha256(buf.data() + 16, 4 * static_cast<size_t>(count), digest);
  buf.insert(buf.end(), digest, digest + 32);
  FILE* f = path ? std::fopen(path, "wb") : nullptr;
  if (!f) {
    err = std::string("cannot write band model file ") + (path ? path : "(null)");
    return BV_EINVAL;
  }
  const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  if (std::fclose(f) != 0 || !ok) {
    err = std::string("short write to ") + path;
    return BV_EINVAL;
  }
  return BV_OK;
}

// models/band_mlp.bin beside this library (the committed, trained model).
std::string default_model_path() {
  Dl_info info{};
  if (!dladdr(reinterpret_cast<void*>(&default_model_path), &info) || !info.dli_fname) return {};
  std::string lib = info.dli_fname;
  const auto slash = lib.rfind('/');
  const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
  return dir + "/models/band_mlp.bin";
}

}  // namespace bvp
