// XXH3-64 with seed for inputs of up to 240 bytes (xxHash 0.8.x, the version the reference links,
// conanfile.py:37). Used for the per-cell seed of Chromosome::hash (genome.cpp:317-336), which
// streams name || le64(size) || le64(cell_id); for inputs this short the streaming digest equals
// the one-shot one computed here.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>

namespace mgpu {
namespace xxh3 {

static const uint8_t kSecret[192] = {
    0xb8, 0xfe, 0x6c, 0x39, 0x23, 0xa4, 0x4b, 0xbe, 0x7c, 0x01, 0x81, 0x2c, 0xf7, 0x21, 0xad, 0x1c,
    0xde, 0xd4, 0x6d, 0xe9, 0x83, 0x90, 0x97, 0xdb, 0x72, 0x40, 0xa4, 0xa4, 0xb7, 0xb3, 0x67, 0x1f,
    0xcb, 0x79, 0xe6, 0x4e, 0xcc, 0xc0, 0xe5, 0x78, 0x82, 0x5a, 0xd0, 0x7d, 0xcc, 0xff, 0x72, 0x21,
    0xb8, 0x08, 0x46, 0x74, 0xf7, 0x43, 0x24, 0x8e, 0xe0, 0x35, 0x90, 0xe6, 0x81, 0x3a, 0x26, 0x4c,
    0x3c, 0x28, 0x52, 0xbb, 0x91, 0xc3, 0x00, 0xcb, 0x88, 0xd0, 0x65, 0x8b, 0x1b, 0x53, 0x2e, 0xa3,
    0x71, 0x64, 0x48, 0x97, 0xa2, 0x0d, 0xf9, 0x4e, 0x38, 0x19, 0xef, 0x46, 0xa9, 0xde, 0xac, 0xd8,
    0xa8, 0xfa, 0x76, 0x3f, 0xe3, 0x9c, 0x34, 0x3f, 0xf9, 0xdc, 0xbb, 0xc7, 0xc7, 0x0b, 0x4f, 0x1d,
    0x8a, 0x51, 0xe0, 0x4b, 0xcd, 0xb4, 0x59, 0x31, 0xc8, 0x9f, 0x7e, 0xc9, 0xd9, 0x78, 0x73, 0x64,
    0xea, 0xc5, 0xac, 0x83, 0x34, 0xd3, 0xeb, 0xc3, 0xc5, 0x81, 0xa0, 0xff, 0xfa, 0x13, 0x63, 0xeb,
    0x17, 0x0d, 0xdd, 0x51, 0xb7, 0xf0, 0xda, 0x49, 0xd3, 0x16, 0x55, 0x26, 0x29, 0xd4, 0x68, 0x9e,
    0x2b, 0x16, 0xbe, 0x58, 0x7d, 0x47, 0xa1, 0xfc, 0x8f, 0xf8, 0xb8, 0xd1, 0x7a, 0xd0, 0x31, 0xce,
    0x45, 0xcb, 0x3a, 0x8f, 0x95, 0x16, 0x04, 0x28, 0xaf, 0xd7, 0xfb, 0xca, 0xbb, 0x4b, 0x40, 0x7e,
};

constexpr uint64_t P64_1 = 0x9E3779B185EBCA87ULL;
constexpr uint64_t P64_2 = 0xC2B2AE3D27D4EB4FULL;
constexpr uint64_t P64_3 = 0x165667B19E3779F9ULL;
constexpr uint64_t MX1 = 0x165667919E3779F9ULL;
constexpr uint64_t MX2 = 0x9FB21C651E98DF25ULL;

inline uint64_t le64(const uint8_t* p) {
  uint64_t v;
  memcpy(&v, p, 8);
  return v;
}
inline uint32_t le32(const uint8_t* p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}
inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
inline uint64_t fold(uint64_t a, uint64_t b) {
  const unsigned __int128 m = (unsigned __int128)a * b;
  return (uint64_t)m ^ (uint64_t)(m >> 64);
}
inline uint64_t avalanche(uint64_t h) {
  h ^= h >> 37;
  h *= MX1;
  return h ^ (h >> 32);
}
inline uint64_t avalanche64(uint64_t h) {
  h ^= h >> 33;
  h *= P64_2;
  h ^= h >> 29;
  h *= P64_3;
  return h ^ (h >> 32);
}
inline uint64_t rrmxmx(uint64_t h, uint64_t len) {
  h ^= rotl(h, 49) ^ rotl(h, 24);
  h *= MX2;
  h ^= (h >> 35) + len;
  h *= MX2;
  return h ^ (h >> 28);
}
inline uint64_t mix16(const uint8_t* in, const uint8_t* sec, uint64_t seed) {
  return fold(le64(in) ^ (le64(sec) + seed), le64(in + 8) ^ (le64(sec + 8) - seed));
}

// Returns false for inputs longer than 240 bytes.
inline bool hash64(const void* data, size_t len, uint64_t seed, uint64_t* out) {
  const uint8_t* in = static_cast<const uint8_t*>(data);
  const uint8_t* s = kSecret;
  if (len == 0) {
    *out = avalanche64(seed ^ (le64(s + 56) ^ le64(s + 64)));
  } else if (len <= 3) {
    const uint32_t comb = ((uint32_t)in[0] << 16) | ((uint32_t)in[len >> 1] << 24) |
                          (uint32_t)in[len - 1] | ((uint32_t)len << 8);
    *out = avalanche64((uint64_t)comb ^ ((uint64_t)(le32(s) ^ le32(s + 4)) + seed));
  } else if (len <= 8) {
    const uint64_t sd = seed ^ ((uint64_t)__builtin_bswap32((uint32_t)seed) << 32);
    const uint64_t v = (uint64_t)le32(in + len - 4) + ((uint64_t)le32(in) << 32);
    *out = rrmxmx(v ^ ((le64(s + 8) ^ le64(s + 16)) - sd), len);
  } else if (len <= 16) {
    const uint64_t lo = le64(in) ^ ((le64(s + 24) ^ le64(s + 32)) + seed);
    const uint64_t hi = le64(in + len - 8) ^ ((le64(s + 40) ^ le64(s + 48)) - seed);
    *out = avalanche(len + __builtin_bswap64(lo) + hi + fold(lo, hi));
  } else if (len <= 128) {
    uint64_t acc = len * P64_1;
    const int pairs = len > 96 ? 4 : len > 64 ? 3 : len > 32 ? 2 : 1;
    for (int k = pairs - 1; k >= 0; --k) {
      acc += mix16(in + 16 * k, s + 32 * k, seed);
      acc += mix16(in + len - 16 * (k + 1), s + 32 * k + 16, seed);
    }
    *out = avalanche(acc);
  } else if (len <= 240) {
    uint64_t acc = len * P64_1;
    for (size_t i = 0; i < 8; ++i) acc += mix16(in + 16 * i, s + 16 * i, seed);
    uint64_t tail = mix16(in + len - 16, s + 136 - 17, seed);
    acc = avalanche(acc);
    for (size_t i = 8; i < len / 16; ++i) tail += mix16(in + 16 * i, s + 16 * (i - 8) + 3, seed);
    *out = avalanche(acc + tail);
  } else {
    return false;
  }
  return true;
}

}  // namespace xxh3
}  // namespace mgpu
