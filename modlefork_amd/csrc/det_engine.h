// Deterministic-mode random engine and samplers (device).
//
// Reproduces the per-cell stream of the reference: Xoshiro256++ seeded through SplitMix64
// (src/common/include/modle/common/random.hpp:26-30, xoshiro-cpp 1.1) and the Boost.Random 1.79
// samplers the reference aliases (random.hpp:36-59): uniform_int (bucket rejection), bernoulli,
// generate_canonical<double,53>, uniform_01, normal/exponential (ziggurat), poisson (inversion /
// PTRD) and binomial (inversion / BTRD), plus MoDLE's generalized-extreme-value sampler
// (genextreme_value_distribution.hpp:87-105). Every function consumes engine words in the same
// order and count as the reference, so a cell's trajectory is reproduced draw for draw.
// Compiled with -ffp-contract=off: the reference build performs no FMA contraction.
#pragma once
#ifndef MGPU_EMU
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "rng_tables.h"

namespace mgpu {

struct DetRng {
  uint64_t s0, s1, s2, s3;
  uint64_t draws;
};

__device__ __host__ inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

__device__ __host__ inline void det_seed(DetRng& g, uint64_t seed) {
  uint64_t st = seed;
  uint64_t w[4];
  for (int i = 0; i < 4; ++i) {
    st += 0x9e3779b97f4a7c15ULL;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    w[i] = z ^ (z >> 31);
  }
  g.s0 = w[0];
  g.s1 = w[1];
  g.s2 = w[2];
  g.s3 = w[3];
  g.draws = 0;
}

__device__ inline uint64_t det_next(DetRng& g) {
  const uint64_t out = rotl64(g.s0 + g.s3, 23) + g.s0;
  const uint64_t t = g.s1 << 17;
  g.s2 ^= g.s0;
  g.s3 ^= g.s1;
  g.s1 ^= g.s2;
  g.s0 ^= g.s3;
  g.s2 ^= t;
  g.s3 = rotl64(g.s3, 45);
  g.draws++;
  return out;
}

#define MGPU_TWO64 18446744073709551616.0

// Normal ziggurat layer tables as a sampler reads them; a word source may supply its own copy
// (an overload for its type), e.g. one staged in LDS.
template <class G>
__device__ inline double zig_nx(const G&, int i) {
  return kZigNormalX[i];
}
template <class G>
__device__ inline double zig_ny(const G&, int i) {
  return kZigNormalY[i];
}

// uniform_int_distribution<u64>{lo, hi} over a 64-bit engine (brange = 2^64 - 1).
template <class G>
__device__ inline uint64_t det_uniform(G& g, uint64_t lo, uint64_t hi) {
  const uint64_t range = hi - lo;
  if (range == 0) return lo;
  if (range == ~0ULL) return det_next(g) + lo;
  uint64_t bucket = ~0ULL / (range + 1);
  if (~0ULL % (range + 1) == range) ++bucket;
  for (;;) {
    const uint64_t r = det_next(g) / bucket;
    if (r <= range) return r + lo;
  }
}

// Same with a precomputed bucket size (the divisor only depends on the range).
__device__ inline uint64_t bucket_of(uint64_t range) {
  uint64_t bucket = ~0ULL / (range + 1);
  if (~0ULL % (range + 1) == range) ++bucket;
  return bucket;
}
template <class G>
__device__ inline uint64_t det_uniform_b(G& g, uint64_t lo, uint64_t range, uint64_t bucket) {
  if (range == 0) return lo;
  for (;;) {
    const uint64_t r = det_next(g) / bucket;
    if (r <= range) return r + lo;
  }
}

template <class G>
__device__ inline bool det_bernoulli(G& g, double p) {
  if (p == 0.0) return false;
  return (double)det_next(g) <= p * MGPU_TWO64;
}

template <class G>
__device__ inline double det_canonical(G& g) {
  double r = (double)det_next(g) / MGPU_TWO64;
  if (r == 1.0) r -= 1.1102230246251565e-16;  // epsilon / 2
  return r;
}

template <class G>
__device__ inline double det_uniform01(G& g) {
  for (;;) {
    const double r = (double)det_next(g) * (1.0 / MGPU_TWO64);
    if (r < 1.0) return r;
  }
}

// generate_int_float_pair<double, 8>: 8-bit bucket from the low bits, 56-bit fraction above.
template <class G>
__device__ inline double det_int_float(G& g, int& bucket) {
  uint64_t u = det_next(g);
  bucket = (int)(u & 0xFFu);
  u &= ~(1ULL << 11);
  return (double)(u >> 8) * (1.0 / 72057594037927936.0);
}

template <class G>
__device__ inline double det_unit_exp(G& g) {
  double shift = 0.0;
  for (;;) {
    int i;
    const double r = det_int_float(g, i);
    const double xi = kZigExpX[i], xi1 = kZigExpX[i + 1];
    const double x = r * xi;
    if (x < xi1) return shift + x;
    if (i == 0) {
      shift += kZigExpX[1];
      continue;
    }
    const double yi = kZigExpY[i], yi1 = kZigExpY[i + 1];
    const double y01 = det_uniform01(g);
    const double y = yi + y01 * (yi1 - yi);
    const double above_ub = (xi - xi1) * y01 - (xi - x);
    const double above_lb = y - (yi + (xi - x) * yi);
    if (above_ub < 0 && (above_lb < 0 || y < exp(-x))) return x + shift;
  }
}

template <class G>
__device__ inline double det_normal_tail(G& g) {
  const double start = zig_nx(g, 1);
  for (;;) {
    const double x = det_unit_exp(g) / start;
    const double y = det_unit_exp(g) / 1.0;
    if (2 * y > x * x) return x + start;
  }
}

template <class G>
__device__ inline double det_unit_normal(G& g) {
  for (;;) {
    int i;
    const double r = det_int_float(g, i);
    const double sign = (double)((i & 1) * 2 - 1);
    i >>= 1;
    const double xi = zig_nx(g, i), xi1 = zig_nx(g, i + 1);
    const double x = r * xi;
    if (x < xi1) return x * sign;
    if (i == 0) return det_normal_tail(g) * sign;
    const double yi = zig_ny(g, i), yi1 = zig_ny(g, i + 1);
    const double y01 = det_uniform01(g);
    const double y = yi + y01 * (yi1 - yi);
    double above_ub, above_lb;
    if (xi >= 1) {
      above_ub = (xi - xi1) * y01 - (xi - x);
      above_lb = y - (yi + (xi - x) * yi * xi);
    } else {
      above_lb = (xi - xi1) * y01 - (xi - x);
      above_ub = y - (yi + (xi - x) * yi * xi);
    }
    if (above_ub < 0 && (above_lb < 0 || y < exp(-(x * x / 2)))) return x * sign;
  }
}

template <class G>
__device__ inline double det_normal(G& g, double mean, double sigma) {
  return det_unit_normal(g) * sigma + mean;
}

template <class G>
__device__ inline uint64_t det_poisson(G& g, double mean) {
  if (mean < 10) {
    double p = exp(-mean);
    uint64_t x = 0;
    double u = det_uniform01(g);
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
  // log(k!) for k < 10, as printed in Boost's table
  const double lf[10] = {0.0, 0.0, 0.69314718055994529, 1.7917594692280550, 3.1780538303479458,
                         4.7874917427820458, 6.5792512120101012, 8.5251613610654147,
                         10.604602902745251, 12.801827480081469};
  for (;;) {
    double u;
    double v = det_uniform01(g);
    if (v <= 0.86 * v_r) {
      u = v / v_r - 0.43;
      return (uint64_t)floor((2 * a / (0.5 - fabs(u)) + b) * u + mean + 0.445);
    }
    if (v >= v_r) {
      u = det_uniform01(g) - 0.5;
    } else {
      u = v / v_r - 0.93;
      u = ((u < 0) ? -0.5 : 0.5) - u;
      v = det_uniform01(g) * v_r;
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
      if (log(v) <= k * log(mean) - mean - lf[(int)k]) return (uint64_t)k;
    }
  }
}

__device__ inline double det_binom_fc(int64_t k) {
  const double t[10] = {0.08106146679532726, 0.04134069595540929, 0.02767792568499834,
                        0.02079067210376509, 0.01664469118982119, 0.01387612882307075,
                        0.01189670994589177, 0.01041126526197209, 0.009255462182712733,
                        0.008330563433362871};
  if (k < 10) return t[k];
  const double ikp1 = 1.0 / (double)(k + 1);
  return (1.0 / 12 - (1.0 / 360 - (1.0 / 1260) * (ikp1 * ikp1)) * (ikp1 * ikp1)) * ikp1;
}

// binomial_distribution<isize>{t, p}. q_n = pow(1 - min(p, 1-p), t) is passed in when the caller
// has it precomputed (pass a negative value to compute it here).
template <class G>
__device__ inline int64_t det_binomial(G& g, int64_t t, double p0, double q_n) {
  const double p = (0.5 < p0) ? (1 - p0) : p0;
  const int64_t m = (int64_t)((double)(t + 1) * p);
  if (m < 11) {
    if (q_n < 0) q_n = pow(1 - p, (double)t);
    const double pp = (0.5 < p0) ? 1 - p0 : p0;
    const double q = 1 - pp;
    const double s = pp / q;
    const double a = (double)(t + 1) * s;
    double r = q_n;
    double u = det_uniform01(g);
    int64_t x = 0;
    while (u > r) {
      u = u - r;
      ++x;
      const double r1 = ((a / (double)x) - s) * r;
      if (r1 < 2.220446049250313e-16 && r1 < r) break;
      r = r1;
    }
    return (0.5 < p0) ? t - x : x;
  }
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
  int64_t k;
  for (;;) {
    double u;
    double v = det_uniform01(g);
    if (v <= u_rv_r) {
      u = v / v_r - 0.43;
      k = (int64_t)floor((2 * a / (0.5 - fabs(u)) + b) * u + c);
      break;
    }
    if (v >= v_r) {
      u = det_uniform01(g) - 0.5;
    } else {
      u = v / v_r - 0.93;
      u = ((u < 0) ? -0.5 : 0.5) - u;
      v = det_uniform01(g) * v_r;
    }
    const double us = 0.5 - fabs(u);
    k = (int64_t)floor((2 * a / us + b) * u + c);
    if (k < 0 || k > t) continue;
    v = v * alpha / (a / (us * us) + b);
    const int64_t kmi = k - m;
    const double km = (double)(kmi < 0 ? -kmi : kmi);
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
      if (v <= f) break;
      continue;
    }
    v = log(v);
    const double rho = (km / npq) * (((km / 3. + 0.625) * km + 1. / 6) / npq + 0.5);
    const double tt = -km * km / (2 * npq);
    if (v < tt - rho) break;
    if (v > tt + rho) continue;
    const int64_t nm = t - m + 1;
    const double h = ((double)m + 0.5) * log((double)(m + 1) / (r * (double)nm)) +
                     det_binom_fc(m) + det_binom_fc(t - m);
    const int64_t nk = t - k + 1;
    if (v <= h + (double)(t + 1) * log((double)nm / (double)nk) +
                 ((double)k + 0.5) * log((double)nk * r / (double)(k + 1)) - det_binom_fc(k) -
                 det_binom_fc(t - k)) {
      break;
    }
  }
  return (0.5 < p0) ? t - k : k;
}

template <class G>
__device__ inline double det_gev(G& g, double mu, double sigma, double xi) {
  if (xi == 0.0) return (mu - sigma) * log(-log(det_canonical(g)));
  return mu + (sigma * (1.0 - pow(-log(det_canonical(g)), xi))) / xi;
}

}  // namespace mgpu
