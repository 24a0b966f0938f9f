/*
 * refcpu.c: CPU oracle (serial C restatement) of MoDLE's per-cell loop-extrusion epoch loop.
 *
 * TEST INFRASTRUCTURE: the checker, never the product. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it (through oracle/refcpu.py).
 *
 * Reference paths are relative to the MoDLE 1.0.0-rc.7 source tree. Third-party algorithms the
 * reference links but that are absent here, restated from their published designs:
 *   - xoshiro-cpp 1.1 (conanfile.py:36): SplitMix64 seeding of Xoshiro256++ (random.hpp:26-30).
 *   - xxHash 0.8.1 (conanfile.py:37): XXH3_64bits_withSeed, inputs up to 240 bytes.
 *   - Boost.Random 1.79 (conanfile.py:17-40; aliases random.hpp:36-59): uniform_int_distribution
 *     (bucket rejection), bernoulli_distribution, generate_canonical<double,53>, uniform_01,
 *     normal_distribution and exponential_distribution (ziggurat, tables from gen_zig_tables.py),
 *     poisson_distribution (inversion below mean 10, PTRD above), binomial_distribution
 *     (inversion below m = 11, BTRD above).
 *   - cpp-sort 1.13 (conanfile.py:24): pdq_sort (pattern-defeating quicksort), split_sort
 *     (non-decreasing-subsequence split, pdqsort of the rest, stable merge), insertion_sort.
 * Floating point follows the reference build (ISO C++17 on x86-64 without -march): no
 * contraction (built with -ffp-contract=off), glibc log/exp/pow/sqrt.
 */
#include "refcpu.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "zig_tables.h"

#define RC_UNBOUND UINT64_MAX
#define RC_BOUND(e) ((e) != RC_UNBOUND)

/* Collision encoding, collision_encoding.hpp:54-108 / collision_encoding_impl.hpp:73-240 on a
 * 64-bit u32f: index in the low bits (INDEX_MASK = bits 0..54), event in bits 56..63. */
#define EV_COLLISION 0x10u
#define EV_CHROM 0x08u
#define EV_LEF_BAR 0x04u
#define EV_PRIMARY 0x02u
#define EV_SECONDARY 0x01u
#define COLL_INDEX_MASK ((~UINT64_C(0)) >> 9)
static inline uint64_t coll_enc(uint64_t idx, unsigned ev) {
  return (idx & COLL_INDEX_MASK) | ((uint64_t)ev << 56);
}
static inline uint64_t coll_idx(uint64_t c) { return c & COLL_INDEX_MASK; }
static inline unsigned coll_ev(uint64_t c) { return (unsigned)(c >> 56); }
static inline int coll_occurred_any(uint64_t c) { return (coll_ev(c) & EV_COLLISION) != 0; }
static inline int coll_occurred(uint64_t c, unsigned ev) { return coll_ev(c) == (ev | EV_COLLISION); }
static inline int coll_avoided(uint64_t c, unsigned ev) {
  return !coll_occurred_any(c) && coll_ev(c) == ev;
}

#define DIR_REV 1
#define DIR_FWD 2
#define BAR_INACTIVE 0
#define BAR_ACTIVE 1

static inline uint64_t umin(uint64_t a, uint64_t b) { return a < b ? a : b; }
static inline uint64_t umax(uint64_t a, uint64_t b) { return a > b ? a : b; }
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* ========================================================================================== */
/* xxHash XXH3 64-bit, seeded, short inputs (<= 240 bytes)                                    */
/* ========================================================================================== */
static const uint8_t kXXH3Secret[192] = {
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
#define XP32_1 0x9E3779B1U
#define XP32_2 0x85EBCA77U
#define XP32_3 0xC2B2AE3DU
#define XP64_1 0x9E3779B185EBCA87ULL
#define XP64_2 0xC2B2AE3D27D4EB4FULL
#define XP64_3 0x165667B19E3779F9ULL
#define XPMX1 0x165667919E3779F9ULL
#define XPMX2 0x9FB21C651E98DF25ULL

static uint64_t rd64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
  return v;
}
static uint32_t rd32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static uint64_t mul128_fold64(uint64_t a, uint64_t b) {
  unsigned __int128 p = (unsigned __int128)a * b;
  return (uint64_t)p ^ (uint64_t)(p >> 64);
}
static uint64_t xxh64_avalanche(uint64_t h) {
  h ^= h >> 33;
  h *= XP64_2;
  h ^= h >> 29;
  h *= XP64_3;
  return h ^ (h >> 32);
}
static uint64_t xxh3_avalanche(uint64_t h) {
  h ^= h >> 37;
  h *= XPMX1;
  return h ^ (h >> 32);
}
static uint64_t xxh3_rrmxmx(uint64_t h, uint64_t len) {
  h ^= rotl64(h, 49) ^ rotl64(h, 24);
  h *= XPMX2;
  h ^= (h >> 35) + len;
  h *= XPMX2;
  return h ^ (h >> 28);
}
static uint64_t xxh3_mix16(const uint8_t* in, const uint8_t* sec, uint64_t seed) {
  return mul128_fold64(rd64(in) ^ (rd64(sec) + seed), rd64(in + 8) ^ (rd64(sec + 8) - seed));
}

uint64_t rc_xxh3_64_withseed(const void* data, size_t len, uint64_t seed) {
  const uint8_t* in = (const uint8_t*)data;
  const uint8_t* sec = kXXH3Secret;
  if (len <= 16) {
    if (len > 8) {
      uint64_t bf1 = (rd64(sec + 24) ^ rd64(sec + 32)) + seed;
      uint64_t bf2 = (rd64(sec + 40) ^ rd64(sec + 48)) - seed;
      uint64_t lo = rd64(in) ^ bf1;
      uint64_t hi = rd64(in + len - 8) ^ bf2;
      uint64_t acc = len + __builtin_bswap64(lo) + hi + mul128_fold64(lo, hi);
      return xxh3_avalanche(acc);
    }
    if (len >= 4) {
      uint64_t s = seed ^ ((uint64_t)__builtin_bswap32((uint32_t)seed) << 32);
      uint32_t i1 = rd32(in), i2 = rd32(in + len - 4);
      uint64_t bf = (rd64(sec + 8) ^ rd64(sec + 16)) - s;
      uint64_t i64 = (uint64_t)i2 + ((uint64_t)i1 << 32);
      return xxh3_rrmxmx(i64 ^ bf, len);
    }
    if (len > 0) {
      uint32_t c1 = in[0], c2 = in[len >> 1], c3 = in[len - 1];
      uint32_t comb = (c1 << 16) | (c2 << 24) | c3 | ((uint32_t)len << 8);
      uint64_t bf = (uint64_t)(rd32(sec) ^ rd32(sec + 4)) + seed;
      return xxh64_avalanche((uint64_t)comb ^ bf);
    }
    return xxh64_avalanche(seed ^ (rd64(sec + 56) ^ rd64(sec + 64)));
  }
  if (len <= 128) {
    uint64_t acc = len * XP64_1;
    if (len > 32) {
      if (len > 64) {
        if (len > 96) {
          acc += xxh3_mix16(in + 48, sec + 96, seed);
          acc += xxh3_mix16(in + len - 64, sec + 112, seed);
        }
        acc += xxh3_mix16(in + 32, sec + 64, seed);
        acc += xxh3_mix16(in + len - 48, sec + 80, seed);
      }
      acc += xxh3_mix16(in + 16, sec + 32, seed);
      acc += xxh3_mix16(in + len - 32, sec + 48, seed);
    }
    acc += xxh3_mix16(in, sec, seed);
    acc += xxh3_mix16(in + len - 16, sec + 16, seed);
    return xxh3_avalanche(acc);
  }
  if (len <= 240) {
    uint64_t acc = len * XP64_1;
    size_t rounds = len / 16;
    for (size_t i = 0; i < 8; ++i) acc += xxh3_mix16(in + 16 * i, sec + 16 * i, seed);
    acc = xxh3_avalanche(acc);
    for (size_t i = 8; i < rounds; ++i) acc += xxh3_mix16(in + 16 * i, sec + 16 * (i - 8) + 3, seed);
    acc += xxh3_mix16(in + len - 16, sec + 136 - 17, seed);
    return xxh3_avalanche(acc);
  }
  return 0; /* long inputs are not needed: chromosome names are short */
}

/* Chromosome::hash, src/libmodle/internal/genome.cpp:317-336: streaming XXH3 over
 * name || le64(size) || le64(cell_id), which equals the one-shot digest for <= 240 bytes. */
uint64_t rc_cell_seed(const char* name, uint64_t name_len, uint64_t chrom_size, uint64_t cell_id,
                      uint64_t seed) {
  uint8_t buf[256];
  if (name_len > 224) return 0;
  memcpy(buf, name, name_len);
  for (int i = 0; i < 8; ++i) buf[name_len + i] = (uint8_t)(chrom_size >> (8 * i));
  for (int i = 0; i < 8; ++i) buf[name_len + 8 + i] = (uint8_t)(cell_id >> (8 * i));
  return rc_xxh3_64_withseed(buf, name_len + 16, seed);
}

/* ========================================================================================== */
/* Engine: SplitMix64 -> Xoshiro256++ (random.hpp:26-30)                                       */
/* ========================================================================================== */
void rc_rng_seed(rc_rng* g, uint64_t seed) {
  uint64_t st = seed;
  for (int i = 0; i < 4; ++i) {
    uint64_t z = (st += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    g->s[i] = z ^ (z >> 31);
  }
  g->draws = 0;
}

uint64_t rc_rng_next(rc_rng* g) {
  uint64_t* s = g->s;
  const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  g->draws++;
  return result;
}

/* ========================================================================================== */
/* Boost.Random 1.79 distributions over a full-range 64-bit engine                             */
/* ========================================================================================== */
#define TWO64 18446744073709551616.0 /* double(UINT64_MAX) rounds to 2^64 */

/* uniform_int_distribution: generate_uniform_int, brange = 2^64-1 > range. */
uint64_t rc_uniform_u64(rc_rng* g, uint64_t lo, uint64_t hi) {
  const uint64_t range = hi - lo;
  if (range == 0) return lo;
  if (range == UINT64_MAX) return rc_rng_next(g) + lo;
  uint64_t bucket = UINT64_MAX / (range + 1);
  if (UINT64_MAX % (range + 1) == range) ++bucket;
  for (;;) {
    const uint64_t r = rc_rng_next(g) / bucket;
    if (r <= range) return r + lo;
  }
}

/* bernoulli_distribution<double>: no draw when p == 0. */
int rc_bernoulli(rc_rng* g, double p) {
  if (p == 0.0) return 0;
  return (double)rc_rng_next(g) <= p * TWO64;
}

/* generate_canonical<double, 53>: one 64-bit draw, S / 2^64, 1.0 mapped below 1. */
double rc_canonical(rc_rng* g) {
  double r = (double)rc_rng_next(g) / TWO64;
  if (r == 1.0) r -= DBL_EPSILON / 2;
  return r;
}

/* uniform_01<double>: double(u) * 2^-64, redrawn when it rounds to 1. */
double rc_uniform01(rc_rng* g) {
  for (;;) {
    double r = (double)rc_rng_next(g) * (1.0 / TWO64);
    if (r < 1.0) return r;
  }
}

/* detail::generate_int_float_pair<double, 8>: low 8 bits -> bucket; bits 8..63 -> [0,1). */
static double int_float_pair(rc_rng* g, int* bucket) {
  uint64_t u = rc_rng_next(g);
  *bucket = (int)(u & 0xFFu);
  u &= ~((uint64_t)1 << 11); /* "zero out unused bits" step for m - w = 56 > 53 */
  return (double)(u >> 8) * (1.0 / 72057594037927936.0);
}

static double unit_exponential(rc_rng* g) {
  const double* X = rc_zig_exp_x;
  const double* Y = rc_zig_exp_y;
  double shift = 0.0;
  for (;;) {
    int i;
    const double r = int_float_pair(g, &i);
    const double x = r * X[i];
    if (x < X[i + 1]) return shift + x;
    if (i == 0) {
      shift += X[1];
    } else {
      const double y01 = rc_uniform01(g);
      const double y = Y[i] + y01 * (Y[i + 1] - Y[i]);
      const double above_ub = (X[i] - X[i + 1]) * y01 - (X[i] - x);
      const double above_lb = y - (Y[i] + (X[i] - x) * Y[i]);
      if (above_ub < 0 && (above_lb < 0 || y < exp(-x))) return x + shift;
    }
  }
}

double rc_exponential(rc_rng* g, double lambda) { return unit_exponential(g) / lambda; }

static double normal_tail(rc_rng* g) {
  const double tail_start = rc_zig_normal_x[1];
  for (;;) {
    const double x = unit_exponential(g) / tail_start;
    const double y = unit_exponential(g) / 1.0;
    if (2 * y > x * x) return x + tail_start;
  }
}

static double unit_normal(rc_rng* g) {
  const double* X = rc_zig_normal_x;
  const double* Y = rc_zig_normal_y;
  for (;;) {
    int i;
    const double r = int_float_pair(g, &i);
    const int sign = (i & 1) * 2 - 1;
    i >>= 1;
    const double x = r * X[i];
    if (x < X[i + 1]) return x * sign;
    if (i == 0) return normal_tail(g) * sign;
    const double y01 = rc_uniform01(g);
    const double y = Y[i] + y01 * (Y[i + 1] - Y[i]);
    double above_ub, above_lb;
    if (X[i] >= 1) {
      above_ub = (X[i] - X[i + 1]) * y01 - (X[i] - x);
      above_lb = y - (Y[i] + (X[i] - x) * Y[i] * X[i]);
    } else {
      above_lb = (X[i] - X[i + 1]) * y01 - (X[i] - x);
      above_ub = y - (Y[i] + (X[i] - x) * Y[i] * X[i]);
    }
    if (above_ub < 0 && (above_lb < 0 || y < exp(-(x * x / 2)))) return x * sign;
  }
}

double rc_normal(rc_rng* g, double mean, double sigma) { return unit_normal(g) * sigma + mean; }

/* poisson_distribution<usize, double>. */
static const double kLogFactorial[10] = {0.0,
                                         0.0,
                                         0.69314718055994529,
                                         1.7917594692280550,
                                         3.1780538303479458,
                                         4.7874917427820458,
                                         6.5792512120101012,
                                         8.5251613610654147,
                                         10.604602902745251,
                                         12.801827480081469};

uint64_t rc_poisson(rc_rng* g, double mean) {
  if (mean < 10) {
    double p = exp(-mean);
    uint64_t x = 0;
    double u = rc_uniform01(g);
    while (u > p) {
      u = u - p;
      ++x;
      p = mean * p / (double)x;
    }
    return x;
  }
  const double smu = sqrt(mean);
  const double b = 0.931 + 2.53 * smu;
  const double a = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  const double v_r = 0.9277 - 3.6224 / (b - 2);
  const double log_sqrt_2pi = 0.91893853320467267;
  for (;;) {
    double u;
    double v = rc_uniform01(g);
    if (v <= 0.86 * v_r) {
      u = v / v_r - 0.43;
      return (uint64_t)floor((2 * a / (0.5 - fabs(u)) + b) * u + mean + 0.445);
    }
    if (v >= v_r) {
      u = rc_uniform01(g) - 0.5;
    } else {
      u = v / v_r - 0.93;
      u = ((u < 0) ? -0.5 : 0.5) - u;
      v = rc_uniform01(g) * v_r;
    }
    const double us = 0.5 - fabs(u);
    if (us < 0.013 && v > us) continue;
    const double k = floor((2 * a / us + b) * u + mean + 0.445);
    v = v * inv_alpha / (a / (us * us) + b);
    if (k >= 10) {
      if (log(v * smu) <= (k + 0.5) * log(mean / k) - mean - log_sqrt_2pi + k -
                              (1 / 12. - (1 / 360. - 1 / (1260. * k * k)) / (k * k)) / k) {
        return (uint64_t)k;
      }
    } else if (k >= 0) {
      if (log(v) <= k * log(mean) - mean - kLogFactorial[(int)k]) return (uint64_t)k;
    }
  }
}

/* binomial_distribution<isize, double>. */
static const double kBinomialFc[10] = {0.08106146679532726, 0.04134069595540929,
                                       0.02767792568499834, 0.02079067210376509,
                                       0.01664469118982119, 0.01387612882307075,
                                       0.01189670994589177, 0.01041126526197209,
                                       0.009255462182712733, 0.008330563433362871};
static double binom_fc(int64_t k) {
  if (k < 10) return kBinomialFc[k];
  const double ikp1 = 1.0 / (double)(k + 1);
  return (1.0 / 12 - (1.0 / 360 - (1.0 / 1260) * (ikp1 * ikp1)) * (ikp1 * ikp1)) * ikp1;
}

static int64_t binom_invert(rc_rng* g, int64_t t, double p, double q_n) {
  const double q = 1 - p;
  const double s = p / q;
  const double a = (double)(t + 1) * s;
  double r = q_n;
  double u = rc_uniform01(g);
  int64_t x = 0;
  while (u > r) {
    u = u - r;
    ++x;
    const double r1 = ((a / (double)x) - s) * r;
    if (r1 < DBL_EPSILON && r1 < r) break;
    r = r1;
  }
  return x;
}

static int64_t binom_btrd(rc_rng* g, int64_t t, double p, int64_t m) {
  const double r = p / (1 - p);
  const double nr = (double)(t + 1) * r;
  const double npq = (double)t * p * (1 - p);
  const double sqrt_npq = sqrt(npq);
  const double b = 1.15 + 2.53 * sqrt_npq;
  const double a = -0.0873 + 0.0248 * b + 0.01 * p;
  const double c = (double)t * p + 0.5;
  const double alpha = (2.83 + 5.1 / b) * sqrt_npq;
  const double v_r = 0.92 - 4.2 / b;
  const double u_rv_r = 0.86 * v_r;
  for (;;) {
    double u;
    double v = rc_uniform01(g);
    if (v <= u_rv_r) {
      u = v / v_r - 0.43;
      return (int64_t)floor((2 * a / (0.5 - fabs(u)) + b) * u + c);
    }
    if (v >= v_r) {
      u = rc_uniform01(g) - 0.5;
    } else {
      u = v / v_r - 0.93;
      u = ((u < 0) ? -0.5 : 0.5) - u;
      v = rc_uniform01(g) * v_r;
    }
    const double us = 0.5 - fabs(u);
    const int64_t k = (int64_t)floor((2 * a / us + b) * u + c);
    if (k < 0 || k > t) continue;
    v = v * alpha / (a / (us * us) + b);
    const double km = (double)llabs(k - m);
    if (km <= 15) {
      double f = 1;
      if (m < k) {
        int64_t i = m;
        do {
          ++i;
          f = f * (nr / (double)i - r);
        } while (i != k);
      } else if (m > k) {
        int64_t i = k;
        do {
          ++i;
          v = v * (nr / (double)i - r);
        } while (i != m);
      }
      if (v <= f) return k;
      continue;
    }
    v = log(v);
    const double rho = (km / npq) * (((km / 3. + 0.625) * km + 1. / 6) / npq + 0.5);
    const double tt = -km * km / (2 * npq);
    if (v < tt - rho) return k;
    if (v > tt + rho) continue;
    const int64_t nm = t - m + 1;
    const double h =
        ((double)m + 0.5) * log((double)(m + 1) / (r * (double)nm)) + binom_fc(m) + binom_fc(t - m);
    const int64_t nk = t - k + 1;
    if (v <= h + (double)(t + 1) * log((double)nm / (double)nk) +
                 ((double)k + 0.5) * log((double)nk * r / (double)(k + 1)) - binom_fc(k) -
                 binom_fc(t - k)) {
      return k;
    }
  }
}

int64_t rc_binomial(rc_rng* g, int64_t t, double p0) {
  const double p = (0.5 < p0) ? (1 - p0) : p0;
  const int64_t m = (int64_t)((double)(t + 1) * p);
  if (m < 11) {
    const double q_n = pow(1 - p, (double)t);
    if (0.5 < p0) return t - binom_invert(g, t, 1 - p0, q_n);
    return binom_invert(g, t, p0, q_n);
  }
  if (0.5 < p0) return t - binom_btrd(g, t, p, m);
  return binom_btrd(g, t, p, m);
}

/* genextreme_value_distribution, src/common/include/modle/common/genextreme_value_distribution.hpp:87-105 */
double rc_gev(rc_rng* g, double mu, double sigma, double xi) {
  if (xi == 0.0) return (mu - sigma) * log(-log(rc_canonical(g)));
  return mu + (sigma * (1.0 - pow(-log(rc_canonical(g)), xi))) / xi;
}

void rc_draw_sequence(uint64_t seed, int kind, double a, double b, double c, uint64_t n,
                      uint64_t* out) {
  rc_rng g;
  rc_rng_seed(&g, seed);
  for (uint64_t i = 0; i < n; ++i) {
    double d = 0;
    uint64_t v = 0;
    switch (kind) {
      case 0: v = rc_rng_next(&g); break;
      case 1: d = rc_normal(&g, a, b); memcpy(&v, &d, 8); break;
      case 2: v = (uint64_t)rc_bernoulli(&g, a); break;
      case 3: d = rc_canonical(&g); memcpy(&v, &d, 8); break;
      case 4: v = rc_uniform_u64(&g, (uint64_t)a, (uint64_t)b); break;
      case 5: v = rc_poisson(&g, a); break;
      case 6: v = (uint64_t)rc_binomial(&g, (int64_t)a, b); break;
      case 7: d = rc_gev(&g, a, b, c); memcpy(&v, &d, 8); break;
      case 8: d = rc_exponential(&g, a); memcpy(&v, &d, 8); break;
      default: break;
    }
    out[i] = v;
  }
}

/* ========================================================================================== */
/* Sorters over index arrays (cpp-sort 1.13)                                                   */
/* ========================================================================================== */
#define LESS(a, b) (key[(a)] < key[(b)])
static inline void swp(size_t* a, size_t* b) {
  size_t t = *a;
  *a = *b;
  *b = t;
}

static void ins_sort(size_t* begin, size_t* end, const uint64_t* key) {
  if (begin == end) return;
  for (size_t* cur = begin + 1; cur != end; ++cur) {
    size_t* sift = cur;
    size_t* sift_1 = cur - 1;
    if (LESS(*sift, *sift_1)) {
      size_t tmp = *sift;
      do {
        *sift-- = *sift_1;
      } while (sift != begin && LESS(tmp, *--sift_1));
      *sift = tmp;
    }
  }
}

static void unguarded_ins_sort(size_t* begin, size_t* end, const uint64_t* key) {
  if (begin == end) return;
  for (size_t* cur = begin + 1; cur != end; ++cur) {
    size_t* sift = cur;
    size_t* sift_1 = cur - 1;
    if (LESS(*sift, *sift_1)) {
      size_t tmp = *sift;
      do {
        *sift-- = *sift_1;
      } while (LESS(tmp, *--sift_1));
      *sift = tmp;
    }
  }
}

static int partial_ins_sort(size_t* begin, size_t* end, const uint64_t* key) {
  if (begin == end) return 1;
  size_t limit = 0;
  for (size_t* cur = begin + 1; cur != end; ++cur) {
    size_t* sift = cur;
    size_t* sift_1 = cur - 1;
    if (LESS(*sift, *sift_1)) {
      size_t tmp = *sift;
      do {
        *sift-- = *sift_1;
      } while (sift != begin && LESS(tmp, *--sift_1));
      *sift = tmp;
      limit += (size_t)(cur - sift);
    }
    if (limit > 8) return 0;
  }
  return 1;
}

static inline void sort2(size_t* a, size_t* b, const uint64_t* key) {
  if (LESS(*b, *a)) swp(a, b);
}
static inline void sort3(size_t* a, size_t* b, size_t* c, const uint64_t* key) {
  sort2(a, b, key);
  sort2(b, c, key);
  sort2(a, b, key);
}

static size_t* partition_right(size_t* begin, size_t* end, const uint64_t* key, int* already) {
  const size_t pivot = *begin;
  size_t* first = begin;
  size_t* last = end;
  while (LESS(*++first, pivot)) {
  }
  if (first - 1 == begin) {
    while (first < last && !LESS(*--last, pivot)) {
    }
  } else {
    while (!LESS(*--last, pivot)) {
    }
  }
  *already = first >= last;
  while (first < last) {
    swp(first, last);
    while (LESS(*++first, pivot)) {
    }
    while (!LESS(*--last, pivot)) {
    }
  }
  size_t* pivot_pos = first - 1;
  *begin = *pivot_pos;
  *pivot_pos = pivot;
  return pivot_pos;
}

static size_t* partition_left(size_t* begin, size_t* end, const uint64_t* key) {
  const size_t pivot = *begin;
  size_t* first = begin;
  size_t* last = end;
  while (LESS(pivot, *--last)) {
  }
  if (last + 1 == end) {
    while (first < last && !LESS(pivot, *++first)) {
    }
  } else {
    while (!LESS(pivot, *++first)) {
    }
  }
  while (first < last) {
    swp(first, last);
    while (LESS(pivot, *--last)) {
    }
    while (!LESS(pivot, *++first)) {
    }
  }
  size_t* pivot_pos = last;
  *begin = *pivot_pos;
  *pivot_pos = pivot;
  return pivot_pos;
}

/* std::make_heap / std::sort_heap as in libstdc++ (__adjust_heap + __push_heap). */
static void adjust_heap(size_t* first, ptrdiff_t hole, ptrdiff_t len, size_t value,
                        const uint64_t* key) {
  const ptrdiff_t top = hole;
  ptrdiff_t child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (LESS(first[child], first[child - 1])) child--;
    first[hole] = first[child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    first[hole] = first[child - 1];
    hole = child - 1;
  }
  ptrdiff_t parent = (hole - 1) / 2;
  while (hole > top && LESS(first[parent], value)) {
    first[hole] = first[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  first[hole] = value;
}
static void heap_sort(size_t* first, size_t* last, const uint64_t* key) {
  const ptrdiff_t len = last - first;
  if (len >= 2) {
    ptrdiff_t parent = (len - 2) / 2;
    for (;;) {
      adjust_heap(first, parent, len, first[parent], key);
      if (parent == 0) break;
      parent--;
    }
  }
  while (last - first > 1) {
    --last;
    const size_t value = *last;
    *last = *first;
    adjust_heap(first, 0, last - first, value, key);
  }
}

static void pdq_loop(size_t* begin, size_t* end, const uint64_t* key, int bad_allowed,
                     int leftmost) {
  for (;;) {
    const ptrdiff_t size = end - begin;
    if (size < 24) {
      if (leftmost) ins_sort(begin, end, key);
      else unguarded_ins_sort(begin, end, key);
      return;
    }
    const ptrdiff_t s2 = size / 2;
    if (size > 128) {
      sort3(begin, begin + s2, end - 1, key);
      sort3(begin + 1, begin + (s2 - 1), end - 2, key);
      sort3(begin + 2, begin + (s2 + 1), end - 3, key);
      sort3(begin + (s2 - 1), begin + s2, begin + (s2 + 1), key);
      swp(begin, begin + s2);
    } else {
      sort3(begin + s2, begin, end - 1, key);
    }
    if (!leftmost && !LESS(*(begin - 1), *begin)) {
      begin = partition_left(begin, end, key) + 1;
      continue;
    }
    int already = 0;
    size_t* pivot_pos = partition_right(begin, end, key, &already);
    const ptrdiff_t l_size = pivot_pos - begin;
    const ptrdiff_t r_size = end - (pivot_pos + 1);
    const int unbalanced = l_size < size / 8 || r_size < size / 8;
    if (unbalanced) {
      if (--bad_allowed == 0) {
        heap_sort(begin, end, key);
        return;
      }
      if (l_size >= 24) {
        swp(begin, begin + l_size / 4);
        swp(pivot_pos - 1, pivot_pos - l_size / 4);
        if (l_size > 128) {
          swp(begin + 1, begin + (l_size / 4 + 1));
          swp(begin + 2, begin + (l_size / 4 + 2));
          swp(pivot_pos - 2, pivot_pos - (l_size / 4 + 1));
          swp(pivot_pos - 3, pivot_pos - (l_size / 4 + 2));
        }
      }
      if (r_size >= 24) {
        swp(pivot_pos + 1, pivot_pos + (1 + r_size / 4));
        swp(end - 1, end - r_size / 4);
        if (r_size > 128) {
          swp(pivot_pos + 2, pivot_pos + (2 + r_size / 4));
          swp(pivot_pos + 3, pivot_pos + (3 + r_size / 4));
          swp(end - 2, end - (1 + r_size / 4));
          swp(end - 3, end - (2 + r_size / 4));
        }
      }
    } else {
      if (already && partial_ins_sort(begin, pivot_pos, key) &&
          partial_ins_sort(pivot_pos + 1, end, key)) {
        return;
      }
    }
    pdq_loop(begin, pivot_pos, key, bad_allowed, leftmost);
    begin = pivot_pos + 1;
    leftmost = 0;
  }
}

void rc_pdq_sort_idx(size_t* idx, size_t n, const uint64_t* key) {
  if (n == 0) return;
  int lg = 0;
  for (size_t m = n; m >>= 1;) ++lg;
  pdq_loop(idx, idx + n, key, lg, 1);
}

/* split_sort: keep a non-decreasing subsequence in front (popping its top together with each
 * element that breaks it), pdq-sort the removed elements, then merge stably. */
void rc_split_sort_idx(size_t* a, size_t n, const uint64_t* key) {
  if (n < 2) return;
  size_t middle = 0;
  for (size_t r = 1; r < n; ++r) {
    if (LESS(a[r], a[middle])) {
      if (middle != 0) --middle;
    } else {
      ++middle;
      swp(&a[middle], &a[r]);
    }
  }
  ++middle;
  if (middle >= n) return;
  rc_pdq_sort_idx(a + middle, n - middle, key);
  /* stable merge of [0, middle) and [middle, n) */
  size_t* buf = (size_t*)malloc(middle * sizeof(size_t));
  memcpy(buf, a, middle * sizeof(size_t));
  size_t i = 0, j = middle, k = 0;
  while (i < middle && j < n) {
    if (LESS(a[j], buf[i])) a[k++] = a[j++];
    else a[k++] = buf[i++];
  }
  while (i < middle) a[k++] = buf[i++];
  free(buf);
}
#undef LESS

/* stable insertion sort by binding epoch (ascending when asc, else descending) */
static void ins_sort_by_epoch(size_t* begin, size_t* end, const uint64_t* ep, int asc) {
  if (begin == end) return;
  for (size_t* cur = begin + 1; cur != end; ++cur) {
    size_t tmp = *cur;
    size_t* sift = cur;
    while (sift != begin) {
      const size_t prev = *(sift - 1);
      const int lt = asc ? (ep[tmp] < ep[prev]) : (ep[prev] < ep[tmp]);
      if (!lt) break;
      *sift = prev;
      --sift;
    }
    *sift = tmp;
  }
}

/* ========================================================================================== */
/* Simulation steps                                                                            */
/* ========================================================================================== */

/* Simulation::rank_lefs, src/libmodle/cpu/simulation.cpp:402-487 */
static void fix_ties(size_t* ranks, size_t n, const uint64_t* pos, const uint64_t* ep, int asc) {
  for (size_t i = 1; i < n; ++i) {
    if (pos[ranks[i - 1]] == pos[ranks[i]]) {
      const size_t begin = i - 1;
      for (; i < n; ++i) {
        if (pos[ranks[i - 1]] != pos[ranks[i]]) break;
      }
      ins_sort_by_epoch(ranks + begin, ranks + i, ep, asc);
    }
  }
}

void rc_rank_lefs(const uint64_t* rev, const uint64_t* fwd, const uint64_t* epoch, size_t n,
                  size_t* rev_ranks, size_t* fwd_ranks, int partially_sorted, int init_buffers) {
  if (init_buffers) {
    for (size_t i = 0; i < n; ++i) rev_ranks[i] = fwd_ranks[i] = i;
  }
  if (partially_sorted) {
    rc_split_sort_idx(rev_ranks, n, rev);
    rc_split_sort_idx(fwd_ranks, n, fwd);
  } else {
    rc_pdq_sort_idx(rev_ranks, n, rev);
    rc_pdq_sort_idx(fwd_ranks, n, fwd);
  }
  fix_ties(rev_ranks, n, rev, epoch, 1);
  fix_ties(fwd_ranks, n, fwd, epoch, 0);
}

/* Simulation::bind_lefs, src/libmodle/cpu/simulation_impl.hpp:28-69 */
void rc_bind_lefs(uint64_t start, uint64_t end, uint64_t* rev, uint64_t* fwd, uint64_t* epoch,
                  size_t n, size_t* rev_ranks, size_t* fwd_ranks, const uint8_t* mask,
                  rc_rng* g, uint64_t current_epoch) {
  for (size_t i = 0; i < n; ++i) {
    if (!mask || mask[i]) {
      const uint64_t p = rc_uniform_u64(g, start, end - 1);
      rev[i] = fwd[i] = p;
      epoch[i] = current_epoch;
    }
  }
  rc_rank_lefs(rev, fwd, epoch, n, rev_ranks, fwd_ranks, current_epoch != 0, 0);
}

/* adjust_moves_of_consecutive_extr_units (simulation.cpp:342-399) + clamp_moves (:325-340) */
static void adjust_moves(const rc_chrom* ch, const uint64_t* rev, const uint64_t* fwd,
                         const uint64_t* ep, size_t n, const size_t* rr, const size_t* fr,
                         uint64_t* rm, uint64_t* fm) {
  for (size_t i = n - 1; i > 0; --i) {
    const size_t i1 = rr[i - 1], i2 = rr[i];
    if (RC_BOUND(ep[i1]) && RC_BOUND(ep[i2])) {
      if (rev[i1] <= ch->start + rm[i1] || rev[i2] <= ch->start + rm[i2]) continue;
      const uint64_t p1 = rev[i1] - rm[i1];
      const uint64_t p2 = rev[i2] - rm[i2];
      if (p2 <= p1) rm[i1] += (p1 - p2) + 1;
    }
  }
  for (size_t i = 1; i < n; ++i) {
    const size_t i1 = fr[i - 1], i2 = fr[i];
    if (RC_BOUND(ep[i1]) && RC_BOUND(ep[i2])) {
      if (fwd[i1] + fm[i1] > ch->end - 1 || fwd[i2] + fm[i2] > ch->end - 1) continue;
      const uint64_t p1 = fwd[i1] + fm[i1];
      const uint64_t p2 = fwd[i2] + fm[i2];
      if (p1 >= p2) fm[i2] += (p1 - p2) + 1;
    }
  }
}

static void clamp_moves(const rc_chrom* ch, const uint64_t* rev, const uint64_t* fwd,
                        const uint64_t* ep, size_t n, uint64_t* rm, uint64_t* fm) {
  for (size_t i = 0; i < n; ++i) {
    if (!RC_BOUND(ep[i])) continue;
    rm[i] = umin(rm[i], rev[i] - ch->start);
    fm[i] = umin(fm[i], ch->end - fwd[i] - 1);
  }
}

void rc_adjust_and_clamp_moves(const rc_chrom* chrom, const uint64_t* rev, const uint64_t* fwd,
                               const uint64_t* epoch, size_t n, const size_t* rev_ranks,
                               const size_t* fwd_ranks, uint64_t* rev_moves, uint64_t* fwd_moves,
                               int adjust) {
  if (adjust && n > 0) adjust_moves(chrom, rev, fwd, epoch, n, rev_ranks, fwd_ranks, rev_moves, fwd_moves);
  clamp_moves(chrom, rev, fwd, epoch, n, rev_moves, fwd_moves);
}

/* generate_moves / generate_moves_helper, simulation.cpp:265-323 */
static uint64_t gen_move(rc_rng* g, uint64_t ep, double speed, double sd) {
  if (!RC_BOUND(ep)) return 0;
  if (sd == 0.0) return (uint64_t)round(speed);
  return (uint64_t)round(fmax(0.0, rc_normal(g, speed, sd)));
}

void rc_generate_moves(const rc_config* cfg, const rc_chrom* chrom, const uint64_t* rev,
                       const uint64_t* fwd, const uint64_t* epoch, size_t n,
                       const size_t* rev_ranks, const size_t* fwd_ranks, uint64_t* rev_moves,
                       uint64_t* fwd_moves, int burnin_completed, rc_rng* g, int adjust) {
  const double rs = (double)(burnin_completed ? cfg->rev_speed : cfg->rev_speed_burnin);
  const double fs = (double)(burnin_completed ? cfg->fwd_speed : cfg->fwd_speed_burnin);
  for (size_t i = 0; i < n; ++i) rev_moves[i] = gen_move(g, epoch[i], rs, cfg->rev_speed_std);
  for (size_t i = 0; i < n; ++i) fwd_moves[i] = gen_move(g, epoch[i], fs, cfg->fwd_speed_std);
  rc_adjust_and_clamp_moves(chrom, rev, fwd, epoch, n, rev_ranks, fwd_ranks, rev_moves, fwd_moves,
                            adjust);
}

/* detect_units_at_chrom_boundaries, simulation_detect_collisions.cpp:30-125 */
static void detect_boundaries(const rc_chrom* ch, const uint64_t* rev, const uint64_t* fwd,
                              const uint64_t* ep, size_t n, const size_t* rr, const size_t* fr,
                              const uint64_t* rm, const uint64_t* fm, uint64_t* rc, uint64_t* fc,
                              size_t* n5_out, size_t* n3_out) {
  size_t n5 = 0, n3 = 0;
  const uint64_t first_fwd = fwd[fr[0]];
  size_t k = n;
  while (k > 0 && !RC_BOUND(ep[rr[k - 1]])) --k;
  const uint64_t last_rev = rev[rr[k > 0 ? k - 1 : 0]];
  for (size_t i = 0; i < n; ++i) {
    const size_t idx = rr[i];
    const uint64_t pos = rev[idx];
    if (pos == ch->start) {
      ++n5;
      rc[idx] = coll_enc(5, EV_COLLISION | EV_CHROM);
    } else if (pos > first_fwd) {
      break;
    } else if (pos - rm[idx] == ch->start) {
      rc[idx] = coll_enc(5, EV_COLLISION | EV_CHROM);
      ++n5;
      break;
    }
  }
  for (size_t i = n - 1; i > 0; --i) {
    const size_t idx = fr[i];
    if (!RC_BOUND(ep[idx])) {
      ++n3;
      continue;
    }
    const uint64_t pos = fwd[idx];
    if (pos == ch->end - 1) {
      ++n3;
      fc[idx] = coll_enc(3, EV_COLLISION | EV_CHROM);
    } else if (pos < last_rev) {
      break;
    } else if (pos + fm[idx] == ch->end - 1) {
      fc[idx] = coll_enc(3, EV_COLLISION | EV_CHROM);
      ++n3;
      break;
    }
  }
  *n5_out = n5;
  *n3_out = n3;
}

/* run_lef_bar_collision_trial / run_lef_lef_collision_trial, simulation_impl.hpp:123-131 */
static int lef_bar_trial(double pblock, rc_rng* g) { return pblock == 1.0 || rc_bernoulli(g, pblock); }
static int lef_lef_trial(const rc_config* c, rc_rng* g) {
  return c->p_bypass == 0.0 || rc_bernoulli(g, 1.0 - c->p_bypass);
}

/* detect_lef_bar_collisions, simulation_detect_collisions.cpp:128-252 */
static void detect_lef_bar(const rc_config* c, const uint64_t* rev, const uint64_t* fwd,
                           const uint64_t* ep, size_t n, const size_t* rr, const size_t* fr,
                           const uint64_t* rm, const uint64_t* fm, const rc_barriers* b,
                           const uint8_t* bst, uint64_t* rc, uint64_t* fc, rc_rng* g, size_t n5,
                           size_t n3) {
  size_t j = umin(n5, n5 - 1);
  size_t uidx = rr[j];
  uint64_t upos = rev[uidx];
  for (size_t i = 0; i < b->n; ++i) {
    if (bst[i] == BAR_INACTIVE) continue;
    const double pblock = b->dir[i] == DIR_REV ? c->pblock_major : c->pblock_minor;
    while (upos <= b->pos[i]) {
      if (++j == n) goto fwd_units;
      uidx = rr[j];
      upos = rev[uidx];
    }
    if (RC_BOUND(ep[uidx])) {
      const uint64_t delta = upos - b->pos[i];
      if (delta > 0 && delta <= rm[uidx] && lef_bar_trial(pblock, g)) {
        rc[uidx] = coll_enc(i, EV_COLLISION | EV_LEF_BAR);
      }
    }
  }
fwd_units:
  j = n - umin(n3, n3 - 1);
  uidx = fr[--j];
  upos = fwd[uidx];
  for (size_t i = b->n - 1; i != SIZE_MAX; --i) {
    if (bst[i] == BAR_INACTIVE) continue;
    const double pblock = b->dir[i] == DIR_FWD ? c->pblock_major : c->pblock_minor;
    while (upos >= b->pos[i]) {
      if (--j == SIZE_MAX) return;
      uidx = fr[j];
      upos = fwd[uidx];
    }
    if (RC_BOUND(ep[uidx])) {
      const uint64_t delta = b->pos[i] - upos;
      if (delta > 0 && delta <= fm[uidx] && lef_bar_trial(pblock, g)) {
        fc[uidx] = coll_enc(i, EV_COLLISION | EV_LEF_BAR);
      }
    }
  }
}

/* compute_lef_lef_collision_pos, simulation.cpp:514-542 */
static void lef_lef_collision_pos(uint64_t rpos, uint64_t fpos, uint64_t rmove, uint64_t fmove,
                                  uint64_t* prev, uint64_t* pfwd) {
  const uint64_t rel = rmove + fmove;
  const double ttc = (double)(rpos - fpos) / (double)rel;
  const uint64_t cpos = fpos + (uint64_t)round((double)fmove * ttc);
  if (cpos == fpos) {
    *prev = cpos + 1;
    *pfwd = cpos;
  } else {
    *prev = cpos;
    *pfwd = cpos - 1;
  }
}

/* detect_primary_lef_lef_collisions, simulation_detect_collisions.cpp:255-401 */
static void detect_primary(const rc_config* c, const uint64_t* rev, const uint64_t* fwd, size_t n,
                           const size_t* rr, const size_t* fr, const uint64_t* rm,
                           const uint64_t* fm, const rc_barriers* b, uint64_t* rc, uint64_t* fc,
                           rc_rng* g, size_t n5, size_t n3) {
  if (n5 == n || n3 == n) return;
  size_t i1 = 0, j1 = n5;
  const size_t i2 = n - umin(n3, n3 - 1);
  const size_t j2 = n;
  for (;;) {
    size_t ridx = rr[j1];
    uint64_t rpos = rev[ridx];
    size_t fidx = fr[i1];
    uint64_t fpos = fwd[fidx];
    while (rpos <= fpos) {
      if (++j1 == j2) return;
      ridx = rr[j1];
      rpos = rev[ridx];
    }
    while (fpos < rpos) {
      if (++i1 == i2) return;
      fidx = fr[i1];
      fpos = fwd[fidx];
    }
    fidx = fr[umin(i1, i1 - 1)];
    fpos = fwd[fidx];
    const uint64_t delta = rpos - fpos;
    if (delta > 0 && delta < rm[ridx] + fm[fidx] && lef_lef_trial(c, g)) {
      uint64_t cp_rev, cp_fwd;
      lef_lef_collision_pos(rpos, fpos, rm[ridx], fm[fidx], &cp_rev, &cp_fwd);
      const int rocc = coll_occurred_any(rc[ridx]);
      const int focc = coll_occurred_any(fc[fidx]);
      if (!rocc && !focc) {
        rc[ridx] = coll_enc(fidx, EV_COLLISION | EV_PRIMARY);
        fc[fidx] = coll_enc(ridx, EV_COLLISION | EV_PRIMARY);
      } else if (rocc && !focc) {
        const uint64_t bpos = b->pos[coll_idx(rc[ridx])];
        if (cp_fwd > bpos) {
          rc[ridx] = coll_enc(fidx, EV_COLLISION | EV_PRIMARY);
          fc[fidx] = coll_enc(ridx, EV_COLLISION | EV_PRIMARY);
        } else {
          fc[fidx] = coll_enc(ridx, EV_COLLISION | EV_PRIMARY);
        }
      } else if (!rocc && focc) {
        const uint64_t bpos = b->pos[coll_idx(fc[fidx])];
        rc[ridx] = coll_enc(fidx, EV_COLLISION | EV_PRIMARY);
        if (cp_rev < bpos) fc[fidx] = coll_enc(ridx, EV_COLLISION | EV_PRIMARY);
      }
    }
  }
}

/* correct_moves_for_lef_bar_collisions, simulation_correct_moves.cpp:20-51 */
static void correct_lef_bar(const uint64_t* rev, const uint64_t* fwd, size_t n,
                            const rc_barriers* b, uint64_t* rm, uint64_t* fm, const uint64_t* rc,
                            const uint64_t* fc) {
  for (size_t i = 0; i < n; ++i) {
    if (coll_occurred(rc[i], EV_LEF_BAR)) rm[i] = (rev[i] - b->pos[coll_idx(rc[i])]) - 1;
    if (coll_occurred(fc[i], EV_LEF_BAR)) fm[i] = (b->pos[coll_idx(fc[i])] - fwd[i]) - 1;
  }
}

/* correct_moves_for_primary_lef_lef_collisions, simulation_correct_moves.cpp:54-122 */
static void correct_primary(const uint64_t* rev, const uint64_t* fwd, size_t n, const size_t* rr,
                            const size_t* fr, uint64_t* rm, uint64_t* fm, const uint64_t* rc,
                            const uint64_t* fc) {
  for (size_t k = 0; k < n; ++k) {
    const size_t ridx = rr[k];
    if (!coll_occurred(rc[ridx], EV_PRIMARY)) continue;
    const size_t fidx = coll_idx(rc[ridx]);
    if (coll_occurred(fc[fidx], EV_PRIMARY)) {
      uint64_t p1, p2;
      lef_lef_collision_pos(rev[ridx], fwd[fidx], rm[ridx], fm[fidx], &p1, &p2);
      rm[ridx] = rev[ridx] - p1;
      fm[fidx] = p2 - fwd[fidx];
    } else if (coll_occurred(fc[fidx], EV_LEF_BAR)) {
      rm[ridx] = rev[ridx] - (fwd[fidx] + fm[fidx]) - 1;
    }
  }
  for (size_t k = 0; k < n; ++k) {
    const size_t fidx = fr[k];
    if (!coll_occurred(fc[fidx], EV_PRIMARY)) continue;
    const size_t ridx = coll_idx(fc[fidx]);
    if (coll_occurred(rc[ridx], EV_LEF_BAR)) {
      fm[fidx] = (rev[ridx] - rm[ridx]) - fwd[fidx] - 1;
    }
  }
}

/* process_secondary_lef_lef_collisions, simulation_detect_collisions.cpp:404-518 */
static void process_secondary(const rc_config* c, const uint64_t* rev, const uint64_t* fwd,
                              size_t n, const size_t* rr, const size_t* fr, uint64_t* rm,
                              uint64_t* fm, uint64_t* rc, uint64_t* fc, rc_rng* g, size_t n5,
                              size_t n3) {
  for (size_t i = umax(1, n5); i < n; ++i) {
    const size_t r1 = rr[i - 1];
    if (!coll_occurred_any(rc[r1])) continue;
    const size_t r2 = rr[i];
    if (coll_occurred_any(rc[r2])) continue;
    if (rev[r2] - rm[r2] <= rev[r1] - rm[r1]) {
      if (lef_lef_trial(c, g)) {
        rc[r2] = coll_enc(r1, EV_COLLISION | EV_SECONDARY);
        const uint64_t mv = rev[r2] - (rev[r1] - rm[r1]);
        rm[r2] = umin(mv, mv - 1);
      } else {
        rc[r2] = coll_enc(r1, EV_SECONDARY);
      }
    }
  }
  size_t i = n - umin(n3, n3 - 1) - 1;
  for (; i > 0; --i) {
    const size_t f2 = fr[i];
    if (!coll_occurred_any(fc[f2])) continue;
    const size_t f1 = fr[i - 1];
    if (coll_occurred_any(fc[f1])) continue;
    if (fwd[f1] + fm[f1] >= fwd[f2] + fm[f2]) {
      if (lef_lef_trial(c, g)) {
        fc[f1] = coll_enc(f2, EV_COLLISION | EV_SECONDARY);
        const uint64_t mv = (fwd[f2] + fm[f2]) - fwd[f1];
        fm[f1] = umin(mv, mv - 1);
      } else {
        fc[f1] = coll_enc(f2, EV_SECONDARY);
      }
    }
  }
}

/* fix_secondary_lef_lef_collisions, simulation_detect_collisions.cpp:520-647 */
static void fix_secondary(const rc_chrom* ch, uint64_t* rev, uint64_t* fwd, size_t n, size_t* rr,
                          size_t* fr, uint64_t* rm, uint64_t* fm, uint64_t* rc, uint64_t* fc,
                          size_t n5, size_t n3) {
  const size_t n_active_fwd = n - umin(n3, n3 - 1);
  for (size_t i = umax(1, n5); i < n; ++i) {
    const size_t r2 = rr[i];
    if (!coll_avoided(rc[r2], EV_SECONDARY)) continue;
    const size_t r1 = rr[i - 1];
    const uint64_t pos1 = rev[r1] - rm[r1];
    if (rev[r2] > pos1 + 1) rm[r2] = rev[r2] - (pos1 + 1);
    else rm[r2] = 0;
    rc[r2] = coll_enc(r1, EV_COLLISION | EV_SECONDARY);
    const uint64_t p1 = rev[r1], p2 = rev[r2];
    rev[r1] = umin(fwd[r1], p2);
    rev[r2] = umin(fwd[r2], p1);
    uint64_t t = rc[r1]; rc[r1] = rc[r2]; rc[r2] = t;
    t = rm[r1]; rm[r1] = rm[r2]; rm[r2] = t;
    rr[i - 1] = r2;
    rr[i] = r1;
    rm[rr[i - 1]] = umin(rev[rr[i - 1]] - ch->start, rm[rr[i - 1]]);
    rm[rr[i]] = umin(rev[rr[i]] - ch->start, rm[rr[i]]);
  }
  for (size_t i = 0; i < n_active_fwd - 1; ++i) {
    const size_t f1 = fr[i];
    if (!coll_avoided(fc[f1], EV_SECONDARY)) continue;
    const size_t f2 = fr[i + 1];
    const uint64_t pos2 = fwd[f2] + fm[f2];
    if (pos2 > fwd[f1] + 1) fm[f1] = pos2 - (fwd[f1] + 1);
    else fm[f1] = 0;
    fc[f1] = coll_enc(f2, EV_COLLISION | EV_SECONDARY);
    const uint64_t p1 = fwd[f1], p2 = fwd[f2];
    fwd[f1] = umax(rev[f1], p2);
    fwd[f2] = umax(rev[f2], p1);
    uint64_t t = fc[f1]; fc[f1] = fc[f2]; fc[f2] = t;
    t = fm[f1]; fm[f1] = fm[f2]; fm[f2] = t;
    fr[i] = f2;
    fr[i + 1] = f1;
    fm[fr[i]] = umin(ch->end - 1 - fwd[fr[i]], fm[fr[i]]);
    fm[fr[i + 1]] = umin(ch->end - 1 - fwd[fr[i + 1]], fm[fr[i + 1]]);
  }
}

/* process_collisions, simulation.cpp:925-955 (and the test hooks of simulation.hpp:587-651) */
void rc_process_collisions(const rc_config* cfg, const rc_chrom* chrom, uint64_t* rev,
                           uint64_t* fwd, const uint64_t* epoch, size_t n, size_t* rr, size_t* fr,
                           uint64_t* rm, uint64_t* fm, const rc_barriers* b, const uint8_t* bst,
                           uint64_t* rc, uint64_t* fc, rc_rng* g, int mode, uint64_t* n5_out,
                           uint64_t* n3_out) {
  size_t n5 = 0, n3 = 0;
  if (mode == 4) {
    fix_secondary(chrom, rev, fwd, n, rr, fr, rm, fm, rc, fc, 0, 0);
    return;
  }
  if (mode == 5) {
    detect_lef_bar(cfg, rev, fwd, epoch, n, rr, fr, rm, fm, b, bst, rc, fc, g, 0, 0);
    return;
  }
  if (mode == 6) {
    detect_primary(cfg, rev, fwd, n, rr, fr, rm, fm, b, rc, fc, g, 0, 0);
    return;
  }
  if (mode == 7) {
    correct_lef_bar(rev, fwd, n, b, rm, fm, rc, fc);
    return;
  }
  if (mode == 8) {
    detect_boundaries(chrom, rev, fwd, epoch, n, rr, fr, rm, fm, rc, fc, &n5, &n3);
    if (n5_out) *n5_out = n5;
    if (n3_out) *n3_out = n3;
    return;
  }
  if (mode == 3) {
    detect_primary(cfg, rev, fwd, n, rr, fr, rm, fm, b, rc, fc, g, 0, 0);
    correct_primary(rev, fwd, n, rr, fr, rm, fm, rc, fc);
    process_secondary(cfg, rev, fwd, n, rr, fr, rm, fm, rc, fc, g, 0, 0);
    return;
  }
  detect_boundaries(chrom, rev, fwd, epoch, n, rr, fr, rm, fm, rc, fc, &n5, &n3);
  if (n5_out) *n5_out = n5;
  if (n3_out) *n3_out = n3;
  detect_lef_bar(cfg, rev, fwd, epoch, n, rr, fr, rm, fm, b, bst, rc, fc, g, n5, n3);
  if (mode == 2) return;
  detect_primary(cfg, rev, fwd, n, rr, fr, rm, fm, b, rc, fc, g, n5, n3);
  correct_lef_bar(rev, fwd, n, b, rm, fm, rc, fc);
  correct_primary(rev, fwd, n, rr, fr, rm, fm, rc, fc);
  process_secondary(cfg, rev, fwd, n, rr, fr, rm, fm, rc, fc, g, n5, n3);
  if (mode == 1) return;
  fix_secondary(chrom, rev, fwd, n, rr, fr, rm, fm, rc, fc, n5, n3);
}

/* extrusion_barriers_impl.hpp:106-128 */
/* Collision words as the known-answer tests read them (collision_encoding_impl.hpp:150-240):
 * ev 0 asks collision_occurred() (the COLLISION bit) or collision_avoided() (no COLLISION bit on
 * a non-zero word); otherwise the event must equal ev | COLLISION, or ev without it. */
uint64_t rc_coll_encode(uint64_t idx, unsigned ev) { return coll_enc(idx, ev); }
uint64_t rc_coll_index(uint64_t c) { return coll_idx(c); }
unsigned rc_coll_event(uint64_t c) { return coll_ev(c); }
int rc_coll_occurred(uint64_t c, unsigned ev) {
  return ev ? coll_occurred(c, ev) : coll_occurred_any(c);
}
int rc_coll_avoided(uint64_t c, unsigned ev) {
  return ev ? coll_avoided(c, ev) : (!coll_occurred_any(c) && c != 0);
}

double rc_stp_active_from_occupancy(double stp_inactive, double occupancy) {
  if (occupancy == 0) return 0.0;
  const double tp_ia = 1.0 - stp_inactive;
  const double tp_ai = (tp_ia - (occupancy * tp_ia)) / occupancy;
  const double v = 1.0 - tp_ai;
  return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}
double rc_occupancy_from_stp(double stp_active, double stp_inactive) {
  if (stp_active + stp_inactive == 0) return 0.0;
  const double tp_ia = 1.0 - stp_inactive;
  const double tp_ai = 1.0 - stp_active;
  const double occ = tp_ia / (tp_ia + tp_ai);
  return occ < 0.0 ? 0.0 : (occ > 1.0 ? 1.0 : occ);
}

/* ========================================================================================== */
/* Contact matrix (contact_matrix_dense_impl.hpp:40-50, _safe_impl.hpp:53-66,                  */
/* contact_matrix_internal_impl.hpp:19-42)                                                     */
/* ========================================================================================== */
uint64_t rc_band_ncols(uint64_t bin_size, uint64_t length) { return (length + bin_size - 1) / bin_size; }
uint64_t rc_band_nrows(uint64_t diagonal_width, uint64_t bin_size, uint64_t length) {
  return umin((diagonal_width + bin_size - 1) / bin_size, rc_band_ncols(bin_size, length));
}
void rc_band_increment(rc_band* m, uint64_t row, uint64_t col) {
  const uint64_t i = row > col ? row - col : col - row;
  const uint64_t j = row > col ? row : col;
  if (i >= m->nrows) {
    __atomic_fetch_add(&m->missed, 1, __ATOMIC_RELAXED);
    return;
  }
  __atomic_fetch_add(&m->data[j * m->nrows + i], 1u, __ATOMIC_RELAXED);
}

/* ContactMatrixDense::unsafe_get, contact_matrix_dense_unsafe_impl.hpp:33-43: the pixel of
 * (row, col) transposed to the upper triangle, 0 beyond the band. */
uint32_t rc_band_get(const rc_band* m, uint64_t row, uint64_t col) {
  const uint64_t i = row > col ? row - col : col - row;
  const uint64_t j = row > col ? row : col;
  return i < m->nrows ? m->data[j * m->nrows + i] : 0u;
}

/* ContactMatrixDense::unsafe_get_block, contact_matrix_dense_unsafe_impl.hpp:89-113: the sum (in
 * the matrix's u32) of the block_size x block_size pixels around (row, col), with every
 * coordinate clamped to [0, ncols - 1]. */
uint32_t rc_band_get_block(const rc_band* m, uint64_t row, uint64_t col, uint64_t block_size) {
  if (block_size == 1) return rc_band_get(m, row, col);
  const int64_t bs = (int64_t)block_size;
  const int64_t first_row = (int64_t)row - bs / 2, first_col = (int64_t)col - bs / 2;
  const int64_t hi = (int64_t)m->ncols - 1;
  uint32_t n = 0;
  for (int64_t i = first_row; i < first_row + bs; ++i) {
    for (int64_t j = first_col; j < first_col + bs; ++j) {
      const uint64_t ii = (uint64_t)(i < 0 ? 0 : (i > hi ? hi : i));
      const uint64_t jj = (uint64_t)(j < 0 ? 0 : (j > hi ? hi : j));
      n += rc_band_get(m, ii, jj);
    }
  }
  return n;
}

/* ========================================================================================== */
/* Whole cell                                                                                  */
/* ========================================================================================== */

/* compute_num_lefs, simulation.cpp:1253-1257 */
uint64_t rc_num_lefs(double lefs_per_mbp, uint64_t simulated_size) {
  const double mbp = (double)simulated_size / 1.0e6;
  return umax(1, (uint64_t)round(lefs_per_mbp * mbp));
}

/* compute_contacts_per_epoch, simulation.cpp:1243-1251 */
uint64_t rc_contacts_per_epoch(const rc_config* cfg, uint64_t nlefs) {
  const double speed = (double)(cfg->rev_speed + cfg->fwd_speed);
  const double p = speed / (double)cfg->contact_sampling_interval;
  return (uint64_t)fmax(1.0, round((double)nlefs * p));
}

/* run_simulate target split (scheduler_simulate.cpp:131-153) and compute_tot_target_epochs
 * (simulation.cpp:1226-1241) */
uint64_t rc_plan_targets(const rc_config* cfg, uint64_t ncells, uint64_t nlefs, uint64_t npixels,
                         uint64_t* target_contacts) {
  const uint64_t tot = (uint64_t)round((double)npixels * cfg->target_contact_density);
  const uint64_t per_cell = (tot + ncells - 1) / ncells;
  uint64_t rolling = 0;
  for (uint64_t c = 0; c < ncells; ++c) {
    const uint64_t eff = umin(per_cell, tot - rolling);
    rolling += eff;
    target_contacts[c] = eff;
  }
  if (cfg->target_contact_density < 0) return ncells * cfg->target_simulation_epochs;
  const double tot_d = fmax(1.0, round(cfg->target_contact_density * (double)npixels));
  return (uint64_t)round(tot_d / (double)rc_contacts_per_epoch(cfg, nlefs));
}

typedef struct cell_state {
  size_t nlefs, nact;
  uint64_t *rev, *fwd, *ep;
  size_t *rr, *fr;
  uint64_t *rm, *fm, *rc, *fc;
  uint8_t* mask;
  uint8_t* bst;
  double *cfx, *avg; /* burn-in history, logical order [0, hlen) */
  size_t hlen;
  uint64_t epoch, burnin_epochs, contacts;
  int burnin_done;
  rc_rng g;
} cell_state;

/* stats::mean over [first, last) (descriptive_impl.hpp:21-30) */
static double mean_of(const double* v, size_t n) {
  if (n == 0) return 0.0;
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) acc = acc + v[i];
  return acc / (double)n;
}

/* compute_loop_size_stats, simulation.cpp:957-981 */
static void loop_size_stats(cell_state* s, size_t cap) {
  const size_t n = s->nact;
  if (n == 0) {
    s->hlen = 0;
    return;
  }
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) acc = acc + (double)(s->fwd[i] - s->rev[i]);
  const double mean = acc / (double)n;
  double ssd = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double d = (double)(s->fwd[i] - s->rev[i]) - mean;
    ssd = ssd + (d * d);
  }
  const double sd = sqrt(ssd / (double)n);
  if (s->hlen == cap) {
    memmove(s->avg, s->avg + 1, (cap - 1) * sizeof(double));
    memmove(s->cfx, s->cfx + 1, (cap - 1) * sizeof(double));
    s->hlen--;
  }
  s->avg[s->hlen] = mean;
  s->cfx[s->hlen] = sd / mean;
  s->hlen++;
}

/* evaluate_burnin, simulation.cpp:983-1025 */
static int dips_ratio_stable(const double* v, size_t cap, size_t w) {
  size_t n = 0;
  for (size_t k = 0; k + w + 1 <= cap - 1; ++k) {
    const double n1 = mean_of(v + k, w);
    const double n2 = mean_of(v + k + 1, w);
    n += (size_t)(n1 > n2);
  }
  const double r = (double)n / (double)(cap - w - n);
  return r >= 0.95 && r <= 1.05;
}
static int evaluate_burnin(const cell_state* s, size_t cap, size_t w) {
  if (s->hlen != cap) return 0;
  if (!dips_ratio_stable(s->cfx, cap, w)) return 0;
  return dips_ratio_stable(s->avg, cap, w);
}

/* run_burnin, simulation.cpp:1027-1054 */
static void run_burnin(const rc_config* c, cell_state* s, double rate) {
  do {
    ++s->burnin_epochs;
    if (s->nact != s->nlefs) {
      const uint64_t k = rc_poisson(&s->g, rate);
      s->nact = (size_t)umin(s->nact + k, s->nlefs);
    } else {
      loop_size_stats(s, c->burnin_history_length);
      s->burnin_done = evaluate_burnin(s, c->burnin_history_length, c->burnin_smoothing_window);
      s->burnin_done &= s->epoch > c->min_burnin_epochs;
      if (!s->burnin_done && s->epoch >= c->max_burnin_epochs) s->burnin_done = 1;
    }
  } while (s->nact == 0);
}

/* register_contacts_loop / register_contacts_tad, register_contacts.cpp:24-71,129-198 */
static void register_contacts(const rc_config* c, cell_state* s, uint64_t sp, uint64_t ep,
                              uint64_t nreg, int tad, rc_band* band) {
  const size_t n = s->nact;
  while (nreg != 0) {
    const size_t i = (size_t)rc_uniform_u64(&s->g, 0, n - 1);
    if (!RC_BOUND(s->ep[i])) continue;
    const uint64_t r = s->rev[i], f = s->fwd[i];
    if (!(r > sp && r < ep && f > sp && f < ep)) continue;
    const double a = (double)r - (c->noisify ? rc_gev(&s->g, c->gev_mu, c->gev_sigma, c->gev_xi) : 0.0);
    const double b = (double)f + (c->noisify ? rc_gev(&s->g, c->gev_mu, c->gev_sigma, c->gev_xi) : 0.0);
    const double p1 = b < a ? b : a;
    const double p2 = a < b ? b : a;
    const double spd = (double)sp, epd = (double)ep;
    if (!(p1 >= spd && p2 >= spd && p1 < epd && p2 < epd)) continue;
    uint64_t q1 = (uint64_t)p1, q2 = (uint64_t)p2;
    if (tad) {
      const uint64_t lo = q1, hi = q2;
      q1 = rc_uniform_u64(&s->g, lo, hi);
      q2 = rc_uniform_u64(&s->g, lo, hi);
    }
    rc_band_increment(band, (q1 - sp) / c->bin_size, (q2 - sp) / c->bin_size);
    --nreg;
  }
}

/* register_1d_lef_occupancy, register_contacts.cpp:200-232: nevents LEFs sampled with
 * replacement; each accepted one adds its two (noisified) unit positions to the 1D histogram. */
static void register_occupancy(const rc_config* c, cell_state* s, uint64_t sp, uint64_t ep,
                               uint64_t nevents, uint64_t* occ) {
  const size_t n = s->nact;
  while (nevents != 0) {
    const size_t i = (size_t)rc_uniform_u64(&s->g, 0, n - 1);
    if (!RC_BOUND(s->ep[i])) continue;
    const uint64_t r = s->rev[i], f = s->fwd[i];
    if (!(r > sp && r < ep && f > sp && f < ep)) continue;
    const double a = (double)r - (c->noisify ? rc_gev(&s->g, c->gev_mu, c->gev_sigma, c->gev_xi) : 0.0);
    const double b = (double)f + (c->noisify ? rc_gev(&s->g, c->gev_mu, c->gev_sigma, c->gev_xi) : 0.0);
    const double p1 = b < a ? b : a;
    const double p2 = a < b ? b : a;
    const double spd = (double)sp, epd = (double)ep;
    if (!(p1 >= spd && p2 >= spd && p1 < epd && p2 < epd)) continue;
    __atomic_fetch_add(&occ[((uint64_t)p1 - sp) / c->bin_size], 1, __ATOMIC_RELAXED);
    __atomic_fetch_add(&occ[((uint64_t)p2 - sp) / c->bin_size], 1, __ATOMIC_RELAXED);
    --nevents;
  }
}

/* sample_and_register_contacts + compute_num_contacts_loop, register_contacts.cpp:73-127 */
static void sample_contacts(const rc_config* c, const rc_chrom* ch, cell_state* s, uint64_t cpe,
                            uint64_t target, rc_band* band, uint64_t* occ) {
  uint64_t n = cpe;
  if (c->target_contact_density > 0.0) n = umin(n, target - s->contacts);
  if (n == 0) return;
  uint64_t nloop;
  if (c->tad_to_loop_ratio == 0) nloop = n;
  else if (!isfinite(c->tad_to_loop_ratio)) nloop = 0;
  else nloop = (uint64_t)rc_binomial(&s->g, (int64_t)n, 1.0 / (c->tad_to_loop_ratio + 1.0));
  const uint64_t ntad = n - nloop;
  register_contacts(c, s, ch->start + 1, ch->end - 1, nloop, 0, band);
  register_contacts(c, s, ch->start + 1, ch->end - 1, ntad, 1, band);
  if (occ) register_occupancy(c, s, ch->start + 1, ch->end - 1, n, occ);
  s->contacts += n;
}

/* release_lefs, simulation.cpp:544-592 */
static void release_lefs(const rc_config* c, cell_state* s, const rc_barriers* b) {
  const double base = s->burnin_done ? c->prob_release : c->prob_release_burnin;
  for (size_t i = 0; i < s->nact; ++i) {
    if (!RC_BOUND(s->ep[i])) continue;
    const int h = (coll_occurred(s->rc[i], EV_LEF_BAR) && b->dir[coll_idx(s->rc[i])] == DIR_REV) +
                  (coll_occurred(s->fc[i], EV_LEF_BAR) && b->dir[coll_idx(s->fc[i])] == DIR_FWD);
    const double aff = h == 0 ? 1.0 : (h == 1 ? 1.0 / c->soft_stall_mult : 1.0 / c->hard_stall_mult);
    if (rc_bernoulli(&s->g, aff * base)) {
      s->rev[i] = s->fwd[i] = RC_UNBOUND;
      s->ep[i] = RC_UNBOUND;
    }
  }
}

static uint64_t epoch_hash(const cell_state* s) {
  uint64_t h = 0x84222325cbf29ce4ULL ^ s->g.draws;
  for (size_t i = 0; i < s->nact; ++i) {
    h = (h ^ s->rev[i]) * 0x100000001b3ULL;
    h = (h ^ s->fwd[i]) * 0x100000001b3ULL;
  }
  return h ^ (s->contacts << 1);
}

static void* xalloc(size_t n) { return calloc(n ? n : 1, 1); }

/* The counts of Simulation::dump_stats (simulation.cpp:1166-1224): stalls are collisions that
 * occurred (event has the COLLISION bit), per side and on both sides of one LEF; LEF-BAR,
 * primary and secondary count occurred collisions of that type over both sides; the average
 * loop size is stats::mean over the active LEFs (descriptive_impl.hpp:21-30, sequential double
 * sum); the effective occupancy is active barriers over all barriers. */
void rc_count_state(const uint64_t* rev, const uint64_t* fwd, size_t n, const uint64_t* rcoll,
                    const uint64_t* fcoll, const uint8_t* bst, size_t nbar, rc_state_record* r) {
  r->num_active_lefs = (uint32_t)n;
  r->num_barriers = nbar;
  r->active_barriers = 0;
  for (size_t i = 0; i < nbar; ++i) r->active_barriers += bst[i] == BAR_ACTIVE;
  r->stalls_rev = r->stalls_fwd = r->stalls_both = 0;
  r->lef_bar = r->primary = r->secondary = 0;
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const int sr = coll_occurred_any(rcoll[i]), sf = coll_occurred_any(fcoll[i]);
    r->stalls_rev += (uint64_t)sr;
    r->stalls_fwd += (uint64_t)sf;
    r->stalls_both += (uint64_t)(sr && sf);
    r->lef_bar += (uint64_t)coll_occurred(rcoll[i], EV_LEF_BAR) + (uint64_t)coll_occurred(fcoll[i], EV_LEF_BAR);
    r->primary += (uint64_t)coll_occurred(rcoll[i], EV_PRIMARY) + (uint64_t)coll_occurred(fcoll[i], EV_PRIMARY);
    r->secondary +=
        (uint64_t)coll_occurred(rcoll[i], EV_SECONDARY) + (uint64_t)coll_occurred(fcoll[i], EV_SECONDARY);
    acc = acc + (double)(fwd[i] - rev[i]);
  }
  r->avg_loop_size = n == 0 ? 0.0 : acc / (double)n;
}

static void log_state(const cell_state* s, const rc_barriers* b, uint64_t cell_id, rc_outputs* out) {
  const uint64_t slot = __atomic_fetch_add(&out->log_n, 1, __ATOMIC_RELAXED);
  if (slot >= out->log_cap) return;
  rc_state_record* r = &out->log[slot];
  rc_count_state(s->rev, s->fwd, s->nact, s->rc, s->fc, s->bst, b->n, r);
  r->cell_id = cell_id;
  r->epoch = s->epoch;
  r->burnin = !s->burnin_done;
}

int rc_simulate_cell(const rc_config* c, const rc_chrom* ch, const rc_barriers* b, uint64_t nlefs,
                     uint64_t cell_id, uint64_t target_contacts, uint64_t target_epochs,
                     rc_band* band, rc_cell_stats* stats, uint64_t* trace, uint64_t max_trace) {
  return rc_simulate_cell_out(c, ch, b, nlefs, cell_id, target_contacts, target_epochs, band,
                              stats, trace, max_trace, NULL);
}

/* simulate_one_cell, simulation.cpp:1056-1152 */
int rc_simulate_cell_out(const rc_config* c, const rc_chrom* ch, const rc_barriers* b,
                         uint64_t nlefs, uint64_t cell_id, uint64_t target_contacts,
                         uint64_t target_epochs, rc_band* band, rc_cell_stats* stats,
                         uint64_t* trace, uint64_t max_trace, rc_outputs* out) {
  uint64_t* occ = out ? out->occupancy : NULL;
  if (nlefs == 0 || ch->end <= ch->start) return -1;
  cell_state s;
  memset(&s, 0, sizeof(s));
  s.nlefs = nlefs;
  s.rev = xalloc(nlefs * 8);
  s.fwd = xalloc(nlefs * 8);
  s.ep = xalloc(nlefs * 8);
  s.rr = xalloc(nlefs * sizeof(size_t));
  s.fr = xalloc(nlefs * sizeof(size_t));
  s.rm = xalloc(nlefs * 8);
  s.fm = xalloc(nlefs * 8);
  s.rc = xalloc(nlefs * 8);
  s.fc = xalloc(nlefs * 8);
  s.mask = xalloc(nlefs);
  s.bst = xalloc(b->n);
  s.cfx = xalloc((c->burnin_history_length + 1) * sizeof(double));
  s.avg = xalloc((c->burnin_history_length + 1) * sizeof(double));
  for (size_t i = 0; i < nlefs; ++i) {
    s.rev[i] = s.fwd[i] = s.ep[i] = RC_UNBOUND;
    s.rr[i] = s.fr[i] = i;
  }

  rc_rng_seed(&s.g, rc_cell_seed(ch->name, ch->name_len, ch->size, cell_id, c->seed));
  const double rate = (double)nlefs / (double)c->burnin_target_epochs_for_lef_activation;
  const uint64_t cpe = rc_contacts_per_epoch(c, nlefs);

  /* ExtrusionBarriers::init_states, extrusion_barriers.cpp:220-232 */
  for (size_t i = 0; i < b->n; ++i) {
    const double occ = rc_occupancy_from_stp(b->stp_active[i], b->stp_inactive[i]);
    s.bst[i] = rc_bernoulli(&s.g, occ) ? BAR_ACTIVE : BAR_INACTIVE;
  }
  if (c->skip_burnin) {
    s.nact = nlefs;
    s.burnin_done = 1;
  }

  for (;; ++s.epoch) {
    const int go = c->target_contact_density >= 0 ? s.contacts < target_contacts
                                                  : s.epoch - s.burnin_epochs < target_epochs;
    if (!go) break;
    if (!s.burnin_done) run_burnin(c, &s, rate);

    const size_t n = s.nact;
    for (size_t i = 0; i < n; ++i) s.mask[i] = !RC_BOUND(s.ep[i]);
    rc_bind_lefs(ch->start, ch->end, s.rev, s.fwd, s.ep, n, s.rr, s.fr, s.mask, &s.g, s.epoch);

    if (s.burnin_done) {
      sample_contacts(c, ch, &s, cpe, target_contacts, band, occ);
      if (target_contacts != 0 && s.contacts >= target_contacts) break;
    }

    rc_generate_moves(c, ch, s.rev, s.fwd, s.ep, n, s.rr, s.fr, s.rm, s.fm, s.burnin_done, &s.g, 1);

    /* ExtrusionBarriers::next_state, extrusion_barriers.cpp:145-161 */
    for (size_t i = 0; i < b->n; ++i) {
      const double u = rc_canonical(&s.g);
      if (s.bst[i] == BAR_INACTIVE && u > b->stp_inactive[i]) s.bst[i] = BAR_ACTIVE;
      else if (s.bst[i] == BAR_ACTIVE && u > b->stp_active[i]) s.bst[i] = BAR_INACTIVE;
    }

    memset(s.rc, 0, n * 8);
    memset(s.fc, 0, n * 8);
    rc_process_collisions(c, ch, s.rev, s.fwd, s.ep, n, s.rr, s.fr, s.rm, s.fm, b, s.bst, s.rc,
                          s.fc, &s.g, 0, NULL, NULL);

    /* extrude, simulation.cpp:489-512 */
    for (size_t i = 0; i < n; ++i) {
      if (!RC_BOUND(s.ep[i])) continue;
      s.rev[i] -= s.rm[i];
      s.fwd[i] += s.fm[i];
    }

    if (out && out->log) log_state(&s, b, cell_id, out);
    release_lefs(c, &s, b);
    if (trace && s.epoch < max_trace) trace[s.epoch] = epoch_hash(&s);
  }

  if (stats) {
    stats->epochs = s.epoch;
    stats->burnin_epochs = s.burnin_epochs;
    stats->contacts = s.contacts;
    stats->draws = s.g.draws;
  }
  free(s.rev); free(s.fwd); free(s.ep); free(s.rr); free(s.fr); free(s.rm); free(s.fm);
  free(s.rc); free(s.fc); free(s.mask); free(s.bst); free(s.cfx); free(s.avg);
  return 0;
}

/* Task-parallel driver: one worker per host thread pulling cell ids, as the reference's
 * simulate_worker pool does (scheduler_simulate.cpp:186-284). */
typedef struct worker_args {
  const rc_config* c;
  const rc_chrom* ch;
  const rc_barriers* b;
  uint64_t nlefs, cell_end, target_epochs;
  const uint64_t* targets;
  uint64_t cell_begin;
  uint64_t* next;
  rc_band* band;
  rc_cell_stats* stats;
  rc_outputs* out;
  int err;
} worker_args;

static void* worker_main(void* p) {
  worker_args* a = (worker_args*)p;
  for (;;) {
    const uint64_t cell = __atomic_fetch_add(a->next, 1, __ATOMIC_RELAXED);
    if (cell >= a->cell_end) break;
    rc_cell_stats st;
    if (rc_simulate_cell_out(a->c, a->ch, a->b, a->nlefs, cell, a->targets[cell - a->cell_begin],
                             a->target_epochs, a->band, &st, NULL, 0, a->out) != 0) {
      a->err = 1;
    }
    if (a->stats) a->stats[cell - a->cell_begin] = st;
  }
  return NULL;
}

int rc_simulate_cells(const rc_config* c, const rc_chrom* ch, const rc_barriers* b,
                      uint64_t nlefs, uint64_t cell_begin, uint64_t cell_end,
                      const uint64_t* targets, uint64_t target_epochs, rc_band* band,
                      int nthreads, rc_cell_stats* stats) {
  return rc_simulate_cells_out(c, ch, b, nlefs, cell_begin, cell_end, targets, target_epochs,
                               band, nthreads, stats, NULL);
}

int rc_simulate_cells_out(const rc_config* c, const rc_chrom* ch, const rc_barriers* b,
                          uint64_t nlefs, uint64_t cell_begin, uint64_t cell_end,
                          const uint64_t* targets, uint64_t target_epochs, rc_band* band,
                          int nthreads, rc_cell_stats* stats, rc_outputs* out) {
  if (nthreads < 1) nthreads = 1;
  uint64_t next = cell_begin;
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  worker_args* args = (worker_args*)calloc((size_t)nthreads, sizeof(worker_args));
  for (int t = 0; t < nthreads; ++t) {
    args[t] = (worker_args){c, ch, b, nlefs, cell_end, target_epochs, targets, cell_begin, &next,
                            band, stats, out, 0};
    pthread_create(&th[t], NULL, worker_main, &args[t]);
  }
  int err = 0;
  for (int t = 0; t < nthreads; ++t) {
    pthread_join(th[t], NULL);
    err |= args[t].err;
  }
  free(th);
  free(args);
  return err ? -1 : 0;
}
