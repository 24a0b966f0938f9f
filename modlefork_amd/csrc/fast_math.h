// Double-precision elementary functions for fast mode's epoch loop: ln, expm1 and the sine and
// cosine of a full turn, from textbook series (ln by the atanh series of the reduced mantissa,
// expm1 and sin/cos by their Taylor series after Cody-Waite / quarter-turn reduction), each to a
// few ulp.
//
// Why not the device math library: gfx950's VOP3 encoding takes no literal operands, so every
// double coefficient of a polynomial lives in a register pair. Inside the epoch loop the compiler
// hoists those materialisations out of the loop, and the ~40 coefficients of log, expm1 and
// sincospi then sit in VGPRs for the whole cell (spilled to scratch at 128 VGPRs, four waves per
// SIMD). Here every coefficient is built by a volatile two-instruction move at its use, which the
// compiler cannot hoist, so the registers are taken only while a function runs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgpu {

// A double constant materialised where it is used (two v_mov_b32 of its 32-bit halves).
template <uint64_t B>
__device__ __forceinline__ double kconst() {
  uint32_t lo, hi;
  asm volatile("v_mov_b32 %0, %2\n\tv_mov_b32 %1, %3"
               : "=v"(lo), "=v"(hi)
               : "i"((int32_t)(uint32_t)(B & 0xFFFFFFFFull)), "i"((int32_t)(uint32_t)(B >> 32)));
  return __hiloint2double((int)hi, (int)lo);
}
#define MGPU_KD(x) ::mgpu::kconst<__builtin_bit_cast(uint64_t, (double)(x))>()

// ln 2 split so that k * kLn2Hi is exact for |k| < 2^21 (its low 21 mantissa bits are zero),
// kLn2Lo the rest of ln 2 rounded to double.
#define MGPU_LN2_HI 0.6931471803691238
#define MGPU_LN2_LO 1.9082149292705877e-10

// ln x for a positive normal double. x = m 2^e with m in [sqrt(1/2), sqrt(2)); with
// s = (m - 1) / (m + 1) (|s| <= 0.1716), ln m = 2 atanh s = 2 s (1 + s^2/3 + s^4/5 + ...),
// summed through s^20 / 21 (the next term is below 2^-56 of the sum).
__device__ inline double ln_d(double x) {
  double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
  int e = __builtin_amdgcn_frexp_exp(x);
  if (m < MGPU_KD(0.70710678118654752440)) {
    m = m + m;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = MGPU_KD(1.0 / 21.0);
  p = fma(p, z, MGPU_KD(1.0 / 19.0));
  p = fma(p, z, MGPU_KD(1.0 / 17.0));
  p = fma(p, z, MGPU_KD(1.0 / 15.0));
  p = fma(p, z, MGPU_KD(1.0 / 13.0));
  p = fma(p, z, MGPU_KD(1.0 / 11.0));
  p = fma(p, z, MGPU_KD(1.0 / 9.0));
  p = fma(p, z, MGPU_KD(1.0 / 7.0));
  p = fma(p, z, MGPU_KD(1.0 / 5.0));
  p = fma(p, z, MGPU_KD(1.0 / 3.0));
  const double t = (s + s) * z * p;  // 2 atanh(s) - 2 s
  const double de = (double)e;
  return fma(de, MGPU_KD(MGPU_LN2_HI), (s + s) + fma(de, MGPU_KD(MGPU_LN2_LO), t));
}

// e^z - 1 for |z| < 700. z = k ln 2 + r with |r| <= ln(2) / 2 (Cody-Waite, two-part ln 2);
// e^r - 1 = r + r^2/2! + ... + r^14/14! (the next term is below 2^-56 of the sum), and
// e^z - 1 = 2^k (e^r - 1) + (2^k - 1) when k != 0 (both terms exact scalings or exact for
// |k| < 53, so the cancellation of 2^k e^r - 1 near z = -ln 2 loses nothing).
__device__ inline double expm1_d(double z) {
  const double k = rint(z * MGPU_KD(1.4426950408889634));
  double r = fma(-k, MGPU_KD(MGPU_LN2_HI), z);
  r = fma(-k, MGPU_KD(MGPU_LN2_LO), r);
  double p = MGPU_KD(1.0 / 87178291200.0);                        // 1/14!
  p = fma(p, r, MGPU_KD(1.0 / 6227020800.0));                     // 1/13!
  p = fma(p, r, MGPU_KD(1.0 / 479001600.0));                      // 1/12!
  p = fma(p, r, MGPU_KD(1.0 / 39916800.0));                       // 1/11!
  p = fma(p, r, MGPU_KD(1.0 / 3628800.0));                        // 1/10!
  p = fma(p, r, MGPU_KD(1.0 / 362880.0));                         // 1/9!
  p = fma(p, r, MGPU_KD(1.0 / 40320.0));                          // 1/8!
  p = fma(p, r, MGPU_KD(1.0 / 5040.0));                           // 1/7!
  p = fma(p, r, MGPU_KD(1.0 / 720.0));                            // 1/6!
  p = fma(p, r, MGPU_KD(1.0 / 120.0));                            // 1/5!
  p = fma(p, r, MGPU_KD(1.0 / 24.0));                             // 1/4!
  p = fma(p, r, MGPU_KD(1.0 / 6.0));                              // 1/3!
  p = fma(p, r, 0.5);                                             // 1/2!
  const double q = fma(p * r, r, r);                              // e^r - 1
  if (k == 0.0) return q;
  return __builtin_amdgcn_ldexp(q, (int)k) + (__builtin_amdgcn_ldexp(1.0, (int)k) - 1.0);
}

// sin(2 pi u) and cos(2 pi u) for u in [0, 1). 4u = k + r with k = rint(4u) and |r| <= 1/2
// (exact), theta = r pi / 2 (|theta| <= pi / 4); sin theta and cos theta by their Taylor series
// through theta^17 / 17! and theta^16 / 16!, then rotated by k quarter turns.
__device__ inline void sincos_2pi_d(double u, double* sn, double* cs) {
  const double t = 4.0 * u;
  const double k = rint(t);
  const double th = (t - k) * MGPU_KD(1.5707963267948966);
  const double z = th * th;
  double ps = MGPU_KD(1.0 / 355687428096000.0);                     // 1/17!
  ps = fma(ps, z, MGPU_KD(-1.0 / 1307674368000.0));                 // -1/15!
  ps = fma(ps, z, MGPU_KD(1.0 / 6227020800.0));                     // 1/13!
  ps = fma(ps, z, MGPU_KD(-1.0 / 39916800.0));                      // -1/11!
  ps = fma(ps, z, MGPU_KD(1.0 / 362880.0));                         // 1/9!
  ps = fma(ps, z, MGPU_KD(-1.0 / 5040.0));                          // -1/7!
  ps = fma(ps, z, MGPU_KD(1.0 / 120.0));                            // 1/5!
  ps = fma(ps, z, MGPU_KD(-1.0 / 6.0));                             // -1/3!
  const double s = fma(th * z, ps, th);
  double pc = MGPU_KD(1.0 / 20922789888000.0);                      // 1/16!
  pc = fma(pc, z, MGPU_KD(-1.0 / 87178291200.0));                   // -1/14!
  pc = fma(pc, z, MGPU_KD(1.0 / 479001600.0));                      // 1/12!
  pc = fma(pc, z, MGPU_KD(-1.0 / 3628800.0));                       // -1/10!
  pc = fma(pc, z, MGPU_KD(1.0 / 40320.0));                          // 1/8!
  pc = fma(pc, z, MGPU_KD(-1.0 / 720.0));                           // -1/6!
  pc = fma(pc, z, MGPU_KD(1.0 / 24.0));                             // 1/4!
  pc = fma(pc, z, -0.5);                                            // -1/2!
  const double c = fma(z, pc, 1.0);
  const int q = (int)k & 3;  // sin(theta + q pi/2), cos(theta + q pi/2)
  const double sa = (q & 1) ? c : s, ca = (q & 1) ? s : c;
  *sn = (q & 2) ? -sa : sa;
  *cs = ((q + 1) & 2) ? -ca : ca;
}

}  // namespace mgpu
