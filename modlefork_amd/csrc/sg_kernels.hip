// Surrogate network kernels: BF16 operands on MFMA with FP32 accumulation (gfx950).
//
// Layouts (channels innermost, "positions" p = (b, d, i) on a D x I grid; 1D layers use D = 1):
//   activations  bf16 [B][D][I][C]
//   conv weights bf16 / fp32 [Cout][T][Cin], T = KD x KI taps in row-major (dd, di) order,
//                centred: tap (td, ti) reads position (d + td - KD/2, i + (ti - KI/2) * dil),
//                zero outside
// Convolutions are implicit GEMMs: out[p][co] = sum_{t,ci} W[co][t][ci] * X[p + off_t][ci].
// v_mfma_f32_32x32x16_bf16: lane l supplies row (l % 32) of A and column (l % 32) of B with the
// 8 consecutive k values 8 * (l / 32) .. + 7; the accumulator holds rows
// 8 * (r / 4) + 4 * (l / 32) + (r % 4) of column l % 32 in register r (checked on the device by
// the layout probe, mgpu_debug_mfma).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "sg_kernels.h"

namespace mgpu {
namespace sg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (values here are finite)
__device__ inline uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

union Pack8 {
  uint4 u;
  bf16x8 v;
  uint16_t h[8];
};

__device__ inline int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// ---------------------------------------------------------------------------------------------
// layout probe: D (32 x 32) = A (32 x 16) * B (16 x 32), bf16 inputs row-major
// ---------------------------------------------------------------------------------------------
__global__ void mfma_probe_kernel(const uint16_t* A, const uint16_t* B, float* D) {
  const int l = threadIdx.x;
  Pack8 a, b;
  for (int s = 0; s < 8; ++s) {
    a.h[s] = A[(l & 31) * 16 + 8 * (l >> 5) + s];
    b.h[s] = B[(8 * (l >> 5) + s) * 32 + (l & 31)];
  }
  f32x16 acc = {};
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.v, b.v, acc, 0, 0, 0);
  for (int r = 0; r < 16; ++r) D[acc_row(r, l) * 32 + (l & 31)] = acc[r];
}

// ---------------------------------------------------------------------------------------------
// implicit-GEMM convolution, forward form
// ---------------------------------------------------------------------------------------------
// One workgroup: 4 waves x 64 rows = 256 positions, 64 output channels (blockIdx.y picks the
// 64-channel slice). Each wave owns a 64 x 64 tile = 2 x 2 MFMA tiles.
constexpr int kConvRowsPerWave = 64;
constexpr int kConvWaves = 4;
constexpr int kConvRows = kConvRowsPerWave * kConvWaves;

__global__ void __launch_bounds__(256)
    conv_fwd_kernel(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * kConvRows + wave * kConvRowsPerWave;
  const int n0 = blockIdx.y * 64;
  const int64_t npos = (int64_t)a.B * a.D * a.I;
  const int kh = a.KD / 2, kw = a.KI / 2;
  const int T = a.KD * a.KI;
  const int dil = a.dil > 0 ? a.dil : 1;
  // positions of this lane's two A rows
  int64_t prow[2];
  int pd[2], pi[2];
  bool pvalid[2];
  for (int r = 0; r < 2; ++r) {
    const int64_t m = m0 + 32 * r + (lane & 31);
    pvalid[r] = m < npos;
    const int64_t mm = pvalid[r] ? m : 0;
    pi[r] = (int)(mm % a.I);
    pd[r] = (int)((mm / a.I) % a.D);
    prow[r] = mm - (int64_t)pd[r] * a.I - pi[r];  // first position of this lane's (b) image
  }
  // the epilogue's inputs, read before the MFMA loop so their latency hides behind it: the
  // lane's two output channels' biases and its residual chunks (32 positions x 8 channels per
  // round r, four chunks per lane)
  float bias[2];
  for (int c = 0; c < 2; ++c) bias[c] = a.bias ? a.bias[n0 + 32 * c + (lane & 31)] : 0.0f;
  uint4 resv[2][4];
  for (int r = 0; r < 2; ++r)
    for (int u = 0; u < 4; ++u) {
      const int e = lane + 64 * u, pos = e >> 3, q = e & 7;
      const int64_t m = m0 + 32 * r + pos;
      resv[r][u] = a.res && a.out2 && m < npos
                       ? *reinterpret_cast<const uint4*>(a.res + m * a.Cout + n0 + 8 * q)
                       : make_uint4(0, 0, 0, 0);
    }
  f32x16 acc[2][2];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) acc[r][c] = f32x16{};
  const int kofs = 8 * (lane >> 5);
  // k-steps (tap t, 16 input channels from ci): the operands of step s + 1 are loaded before
  // the MFMAs of step s, so each global load overlaps a step's MFMAs
  const int ksteps = a.Cin / 16, nsteps = T * ksteps;
  auto load_step = [&](int st, Pack8* A, Pack8* Bv) {
    const int t = st / ksteps, ci = (st - t * ksteps) * 16;
    const int dd = t / a.KI - kh, di = (t % a.KI - kw) * dil;
    for (int r = 0; r < 2; ++r) {
      const int d = pd[r] + dd, i = pi[r] + di;
      const bool ok = pvalid[r] && d >= 0 && d < a.D && i >= 0 && i < a.I;
      A[r].u = ok ? *reinterpret_cast<const uint4*>(a.X + (prow[r] + (int64_t)d * a.I + i) * a.Cin +
                                                    ci + kofs)
                  : make_uint4(0, 0, 0, 0);
    }
    for (int c = 0; c < 2; ++c)
      Bv[c].u = *reinterpret_cast<const uint4*>(
          a.W + ((int64_t)(n0 + 32 * c + (lane & 31)) * T + t) * a.Cin + ci + kofs);
  };
  Pack8 A[2], Bv[2], An[2], Bn[2];
  load_step(0, A, Bv);
  for (int st = 0; st < nsteps; ++st) {
    if (st + 1 < nsteps) load_step(st + 1, An, Bn);
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c)
        acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[r].v, Bv[c].v, acc[r][c], 0, 0, 0);
    for (int k = 0; k < 2; ++k) {
      A[k] = An[k];
      Bv[k] = Bn[k];
    }
  }
  // epilogue: v = acc + bias; relu; out1 = v; out2 = res + v (all bf16), staged through LDS 32
  // positions at a time so residual reads and output writes are 16-byte vectors per lane
  constexpr int kEpStride = 64 + 4;
  __shared__ __align__(16) float stage_all[kConvWaves * 32 * kEpStride];
  float* stage = stage_all + wave * 32 * kEpStride;
  for (int r = 0; r < 2; ++r) {
    for (int c = 0; c < 2; ++c) {
      const int cl = 32 * c + (lane & 31);
      for (int k = 0; k < 16; ++k) {
        float v = acc[r][c][k] + bias[c];
        if (a.relu) v = v > 0.0f ? v : 0.0f;
        stage[acc_row(k, lane) * kEpStride + cl] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = lane + 64 * u, pos = e >> 3, q = e & 7;
      const int64_t m = m0 + 32 * r + pos;
      if (m < npos) {
        const int64_t o = m * a.Cout + n0 + 8 * q;
        const float* sv = stage + pos * kEpStride + 8 * q;
        Pack8 res;
        res.u = resv[r][u];
        Pack8 o1, o2;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          o1.h[k] = f2bf(sv[k]);
          o2.h[k] = f2bf(bf2f(res.h[k]) + sv[k]);
        }
        if (a.out1) *reinterpret_cast<uint4*>(a.out1 + o) = o1.u;
        if (a.out2) *reinterpret_cast<uint4*>(a.out2 + o) = o2.u;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------------
// 3 x 3 convolution, 64 -> 64 channels, on a D x I grid: the head's layer shape
// ---------------------------------------------------------------------------------------------
// A persistent workgroup per CU walks tiles of 4 rows (d) x 64 columns (i) of the images, all
// 64 output channels; wave w owns row w of a tile (64 positions, 2 x 2 MFMA tiles). The layer's
// weights (9 taps x 64 x 64) sit in LDS for the workgroup's lifetime, so the MFMA loop reads
// both operands from LDS and the only global loads are the input patches: the patch a tile
// reads (6 x 66 positions, zero outside the image) is staged in LDS once per tile, and the next
// tile's patch is loaded into registers while this tile computes, then stored after its
// epilogue. Patch positions and weight rows are padded to 144 bytes: 16 consecutive rows then
// start 36 banks apart and a 16-byte read per lane is free of bank conflicts.
constexpr int kT3Rows = 4, kT3Cols = 64, kT3C = 64;
constexpr int kT3PRows = kT3Rows + 2, kT3PCols = kT3Cols + 2;
constexpr int kT3Stride = kT3C + 8;  // bf16 per patch position and per weight row
constexpr int kT3PatchElems = kT3PRows * kT3PCols * kT3Stride;
constexpr int kT3WeightElems = 9 * kT3C * kT3Stride;  // [tap][co][ci]
constexpr int kT3Lds = 2 * (kT3PatchElems + kT3WeightElems);
constexpr int kT3Chunks = kT3PRows * kT3PCols * 8;  // 16-byte chunks of a patch
constexpr int kT3PerLane = (kT3Chunks + 255) / 256;

struct T3Tile {
  int b, d0, i0;
};

__device__ inline T3Tile t3_tile(int tile, int ntx, int nty) {
  return T3Tile{tile / (ntx * nty), (tile / ntx) % nty * kT3Rows, tile % ntx * kT3Cols};
}

// this lane's chunks of tile t's patch, zero outside the image
__device__ inline void t3_load(const ConvArgs& a, T3Tile t, uint4* stg) {
  const uint16_t* X = a.X + (int64_t)t.b * a.D * a.I * kT3C;
#pragma unroll
  for (int m = 0; m < kT3PerLane; ++m) {
    const int e = threadIdx.x + 256 * m;
    const int q = e & 7, pos = e >> 3;
    const int d = t.d0 + pos / kT3PCols - 1, i = t.i0 + pos % kT3PCols - 1;
    stg[m] = make_uint4(0, 0, 0, 0);
    if (e < kT3Chunks && d >= 0 && d < a.D && i >= 0 && i < a.I)
      stg[m] = *reinterpret_cast<const uint4*>(X + ((int64_t)d * a.I + i) * kT3C + 8 * q);
  }
}

__device__ inline void t3_store(uint16_t* patch, const uint4* stg) {
#pragma unroll
  for (int m = 0; m < kT3PerLane; ++m) {
    const int e = threadIdx.x + 256 * m;
    if (e < kT3Chunks)
      *reinterpret_cast<uint4*>(patch + (e >> 3) * kT3Stride + 8 * (e & 7)) = stg[m];
  }
}

__global__ void __launch_bounds__(256) conv3x3_c64_kernel(ConvArgs a, int ntx, int nty, int ntiles) {
  extern __shared__ __align__(16) uint16_t lds3[];
  uint16_t* const patch = lds3;
  uint16_t* const wl = lds3 + kT3PatchElems;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the weights, W[co][t][ci] in memory, as wl[t][co][ci]
  for (int e = threadIdx.x; e < 9 * kT3C * 8; e += 256) {
    const int q = e & 7, row = e >> 3;  // row = co * 9 + t
    const int co = row / 9, t = row % 9;
    *reinterpret_cast<uint4*>(wl + (t * kT3C + co) * kT3Stride + 8 * q) =
        *reinterpret_cast<const uint4*>(a.W + (int64_t)row * kT3C + 8 * q);
  }
  int tile = blockIdx.x;
  if (tile >= ntiles) return;
  uint4 stg[kT3PerLane];
  t3_load(a, t3_tile(tile, ntx, nty), stg);
  t3_store(patch, stg);
  __syncthreads();
  const int kofs = 8 * (lane >> 5);
  // this lane's output channels' biases (32 c + lane % 32), read once
  float bias[2];
  for (int c = 0; c < 2; ++c) bias[c] = a.bias ? a.bias[32 * c + (lane & 31)] : 0.0f;
  for (; tile < ntiles; tile += gridDim.x) {
    const T3Tile tt = t3_tile(tile, ntx, nty);
    const int next = tile + gridDim.x;
    // the next tile's patch, in flight while this one computes
    if (next < ntiles) t3_load(a, t3_tile(next, ntx, nty), stg);
    f32x16 acc[2][2];
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) acc[r][c] = f32x16{};
    // 36 k-steps (9 taps x 4 steps of 16 channels), the operands of step s + 1 read from LDS
    // while the MFMAs of step s run (one wave per SIMD: nothing else hides the LDS latency)
    const uint16_t* abase = patch + (wave * kT3PCols + (lane & 31)) * kT3Stride + kofs;
    const uint16_t* bbase = wl + (lane & 31) * kT3Stride + kofs;
    auto load_step = [&](int st, Pack8* A, Pack8* Bv) {
      const int t = st >> 2, j = st & 3;
      const int dd = t / 3, di = t % 3;  // patch offsets of tap t (the patch starts at -1, -1)
      for (int r = 0; r < 2; ++r)
        A[r].u = *reinterpret_cast<const uint4*>(abase + (dd * kT3PCols + 32 * r + di) * kT3Stride +
                                                 16 * j);
      for (int c = 0; c < 2; ++c)
        Bv[c].u = *reinterpret_cast<const uint4*>(bbase + (t * kT3C + 32 * c) * kT3Stride + 16 * j);
    };
    Pack8 A[2][2], Bv[2][2];
    load_step(0, A[0], Bv[0]);
#pragma unroll
    for (int st = 0; st < 36; ++st) {
      const int cur = st & 1;
      if (st + 1 < 36) load_step(st + 1, A[cur ^ 1], Bv[cur ^ 1]);
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c)
          acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[cur][r].v, Bv[cur][c].v,
                                                               acc[r][c], 0, 0, 0);
    }
    const int i0 = tt.i0, d0 = tt.d0, b = tt.b;
    // epilogue as in conv_fwd_kernel (v = acc + bias; relu; out1 = v; out2 = res + v), through
    // LDS: each wave stages 32 positions x 64 channels of fp32 at a time, then reads them back a
    // channel octet per lane, so the residual reads and the output writes are 16-byte vectors
    // covering whole positions. A residual that is the layer's own input is read from the patch
    // (its centre rows) before the staging overwrites it.
    const bool res_in_patch = a.res == a.X;
    uint4 rp[2][4];
    if (res_in_patch) {
      for (int r = 0; r < 2; ++r)
        for (int u = 0; u < 4; ++u) {
          const int e = lane + 64 * u, pos = e >> 3, q = e & 7;
          rp[r][u] = *reinterpret_cast<const uint4*>(
              patch + ((wave + 1) * kT3PCols + 32 * r + pos + 1) * kT3Stride + 8 * q);
        }
    }
    __syncthreads();  // every wave is done with the patch
    constexpr int kEpStride = kT3C + 4;  // floats per staged position
    float* stage = reinterpret_cast<float*>(patch) + wave * 32 * kEpStride;
    const int d = d0 + wave;
    const int64_t row = ((int64_t)b * a.D + (d < a.D ? d : 0)) * a.I;
    for (int r = 0; r < 2; ++r) {
      for (int c = 0; c < 2; ++c) {
        const int co = 32 * c + (lane & 31);
        for (int k = 0; k < 16; ++k) {
          float v = acc[r][c][k] + bias[c];
          if (a.relu) v = v > 0.0f ? v : 0.0f;
          stage[acc_row(k, lane) * kEpStride + co] = v;
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = lane + 64 * u, pos = e >> 3, q = e & 7;
        const int i = i0 + 32 * r + pos;
        if (d < a.D && i < a.I) {
          const int64_t o = (row + i) * kT3C + 8 * q;
          const float* sv = stage + pos * kEpStride + 8 * q;
          Pack8 res;
          if (!a.res || !a.out2) res.u = make_uint4(0, 0, 0, 0);
          else if (res_in_patch) res.u = rp[r][u];
          else res.u = *reinterpret_cast<const uint4*>(a.res + o);
          Pack8 o1, o2;
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            o1.h[k] = f2bf(sv[k]);
            o2.h[k] = f2bf(bf2f(res.h[k]) + sv[k]);
          }
          if (a.out1) *reinterpret_cast<uint4*>(a.out1 + o) = o1.u;
          if (a.out2) *reinterpret_cast<uint4*>(a.out2 + o) = o2.u;
        }
      }
      __syncthreads();
    }
    // the next tile's patch replaces this one's (the epilogue's last barrier ended every read)
    if (next < ntiles) t3_store(patch, stg);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------------
// weight gradient: dW[co][t][ci] += sum_p G[p][co] * X[p + off_t][ci]
// ---------------------------------------------------------------------------------------------
// Workgroup: one tap (blockIdx.y), a range of positions (blockIdx.x), all (co, ci) tiles spread
// over 4 waves. Chunks of 32 positions of G and of the shifted X are staged in LDS transposed
// ([channel][position]), so each lane reads its 8 k-values (positions) as one 16-byte word.
constexpr int kWgChunk = 32;
constexpr int kWgMaxC = 128;

__global__ void __launch_bounds__(256)
    conv_wgrad_kernel(WgradArgs a) {
  __shared__ __align__(16) uint16_t Gs[kWgMaxC][kWgChunk + 8];
  __shared__ __align__(16) uint16_t Xs[kWgMaxC][kWgChunk + 8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = blockIdx.y;
  const int dd = t / a.KI - a.KD / 2, di = (t % a.KI - a.KI / 2) * (a.dil > 0 ? a.dil : 1);
  const int T = a.KD * a.KI;
  const int64_t npos = (int64_t)a.B * a.D * a.I;
  const int64_t p_begin = (int64_t)blockIdx.x * a.pos_per_block;
  const int64_t p_end = p_begin + a.pos_per_block < npos ? p_begin + a.pos_per_block : npos;
  const int tiles_co = a.Cout / 32, tiles_ci = a.Cin / 32;
  const int ntiles = tiles_co * tiles_ci;
  // each wave owns tiles wave, wave + 4, ... (at most 4 per wave for 128 x 128)
  f32x16 acc[4];
  for (int k = 0; k < 4; ++k) acc[k] = f32x16{};
  for (int64_t p0 = p_begin; p0 < p_end; p0 += kWgChunk) {
    __syncthreads();
    // stage G[p0 .. p0+31][0 .. Cout) and X[p + off][0 .. Cin) transposed
    for (int e = threadIdx.x; e < kWgChunk * a.Cout; e += 256) {
      const int pp = e / a.Cout, ch = e % a.Cout;
      const int64_t p = p0 + pp;
      Gs[ch][pp] = p < p_end ? a.G[p * a.Cout + ch] : (uint16_t)0;
    }
    for (int e = threadIdx.x; e < kWgChunk * a.Cin; e += 256) {
      const int pp = e / a.Cin, ch = e % a.Cin;
      const int64_t p = p0 + pp;
      uint16_t v = 0;
      if (p < p_end) {
        const int i = (int)(p % a.I), d = (int)((p / a.I) % a.D);
        const int d2 = d + dd, i2 = i + di;
        if (d2 >= 0 && d2 < a.D && i2 >= 0 && i2 < a.I)
          v = a.X[(p + (int64_t)dd * a.I + di) * a.Cin + ch];
      }
      Xs[ch][pp] = v;
    }
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
      const int tile = wave + 4 * k;
      if (tile >= ntiles) break;
      const int tco = tile / tiles_ci, tci = tile % tiles_ci;
      for (int kk = 0; kk < kWgChunk; kk += 16) {
        Pack8 A, Bv;
        A.u = *reinterpret_cast<const uint4*>(&Gs[32 * tco + (lane & 31)][kk + 8 * (lane >> 5)]);
        Bv.u = *reinterpret_cast<const uint4*>(&Xs[32 * tci + (lane & 31)][kk + 8 * (lane >> 5)]);
        acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.v, Bv.v, acc[k], 0, 0, 0);
      }
    }
  }
  for (int k = 0; k < 4; ++k) {
    const int tile = wave + 4 * k;
    if (tile >= ntiles) break;
    const int tco = tile / tiles_ci, tci = tile % tiles_ci;
    const int ci = 32 * tci + (lane & 31);
    for (int r = 0; r < 16; ++r) {
      const int co = 32 * tco + acc_row(r, lane);
      atomicAdd(&a.dW[((int64_t)co * T + t) * a.Cin + ci], acc[k][r]);
    }
  }
}

// column sums of a bf16 [npos][C] matrix, added to out[C] (bias gradients)
__global__ void colsum_kernel(const uint16_t* G, int64_t npos, int C, int64_t rows_per_block,
                              float* out) {
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < npos ? r0 + rows_per_block : npos;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    float s = 0.0f;
    for (int64_t r = r0; r < r1; ++r) s += bf2f(G[r * C + c]);
    atomicAdd(&out[c], s);
  }
}

// y = x * [mask > 0] (bf16), the gradient through a ReLU whose output is mask
__global__ void relu_mask_kernel(const uint16_t* g, const uint16_t* mask, uint16_t* out, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = (mask[e] & 0x7FFFu) != 0 && (mask[e] & 0x8000u) == 0 ? g[e] : (uint16_t)0;
}

// rows [off, off + n_out) of every image of a [B][n_in][C] tensor, or the reverse with zeros
__global__ void crop_kernel(const uint16_t* in, uint16_t* out, int B, int n_in, int n_out, int off,
                            int C, int uncrop) {
  const int64_t total = (int64_t)B * (uncrop ? n_in : n_out) * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t rest = e / C;
    if (!uncrop) {
      const int l = (int)(rest % n_out);
      const int64_t b = rest / n_out;
      out[e] = in[(b * n_in + l + off) * C + c];
    } else {
      const int l = (int)(rest % n_in);
      const int64_t b = rest / n_in;
      const int lo = l - off;
      out[e] = (lo >= 0 && lo < n_out) ? in[(b * n_out + lo) * C + c] : (uint16_t)0;
    }
  }
}

// z[b][d][i][c] = relu(P[b][i][c] + Q[b][i + d][c] + E[d][c]) for i + d < L, else 0
__global__ void pair_fwd_kernel(const uint16_t* P, const uint16_t* Q, const float* E, uint16_t* Z,
                                int B, int ND, int L, int C) {
  const int64_t total = (int64_t)B * ND * L * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    int64_t rest = e / C;
    const int i = (int)(rest % L);
    rest /= L;
    const int d = (int)(rest % ND);
    const int64_t b = rest / ND;
    float v = 0.0f;
    if (i + d < L) {
      v = bf2f(P[(b * L + i) * C + c]) + bf2f(Q[(b * L + i + d) * C + c]) + E[d * C + c];
      v = v > 0.0f ? v : 0.0f;
    }
    Z[e] = f2bf(v);
  }
}

// The same with eight channels per lane (16-byte loads and stores) and 32-bit indexing:
// blockIdx.y = b * ND + d, x covers (i, channel octet); C % 8 == 0.
__global__ void pair_fwd8_kernel(const uint16_t* P, const uint16_t* Q, const float* E, uint16_t* Z,
                                 int ND, int L, int C) {
  const int bd = blockIdx.y;
  const int d = bd % ND, b = bd / ND;
  const int oct = C / 8;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L * oct) return;
  const int i = t / oct, q = t % oct;
  const int64_t zo = ((int64_t)bd * L + i) * C + 8 * q;
  Pack8 out;
  if (i + d < L) {
    Pack8 pv, qv;
    pv.u = *reinterpret_cast<const uint4*>(P + ((int64_t)b * L + i) * C + 8 * q);
    qv.u = *reinterpret_cast<const uint4*>(Q + ((int64_t)b * L + i + d) * C + 8 * q);
    const float4 e0 = *reinterpret_cast<const float4*>(E + d * C + 8 * q);
    const float4 e1 = *reinterpret_cast<const float4*>(E + d * C + 8 * q + 4);
    const float ev[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float v = bf2f(pv.h[k]) + bf2f(qv.h[k]) + ev[k];
      out.h[k] = f2bf(v > 0.0f ? v : 0.0f);
    }
  } else {
    out.u = make_uint4(0, 0, 0, 0);
  }
  *reinterpret_cast<uint4*>(Z + zo) = out.u;
}

// gs = gz * [z > 0]; dP[b][i][c] = sum_d gs[b][d][i][c]; dQ[b][j][c] = sum_d gs[b][d][j - d][c]
__global__ void pair_bwd_pq_kernel(const uint16_t* GZ, const uint16_t* Z, uint16_t* dP, uint16_t* dQ,
                                   int B, int ND, int L, int C) {
  const int64_t total = (int64_t)B * L * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t rest = e / C;
    const int i = (int)(rest % L);
    const int64_t b = rest / L;
    float sp = 0.0f, sq = 0.0f;
    for (int d = 0; d < ND; ++d) {
      if (i + d < L) {
        const int64_t o = ((b * ND + d) * L + i) * C + c;
        if (bf2f(Z[o]) > 0.0f) sp += bf2f(GZ[o]);
      }
      if (i - d >= 0) {
        const int64_t o = ((b * ND + d) * L + (i - d)) * C + c;
        if (bf2f(Z[o]) > 0.0f) sq += bf2f(GZ[o]);
      }
    }
    dP[e] = f2bf(sp);
    dQ[e] = f2bf(sq);
  }
}

// dE[d][c] += sum_{b,i} gz * [z > 0]; one block per diagonal d
__global__ void pair_bwd_e_kernel(const uint16_t* GZ, const uint16_t* Z, float* dE, int B, int ND,
                                  int L, int C) {
  const int d = blockIdx.x;
  for (int c = threadIdx.x % C; c < C; c += C) {
    float s = 0.0f;
    for (int64_t r = threadIdx.x / C; r < (int64_t)B * L; r += blockDim.x / C) {
      const int64_t b = r / L;
      const int i = (int)(r % L);
      const int64_t o = ((b * ND + d) * L + i) * C + c;
      if (bf2f(Z[o]) > 0.0f) s += bf2f(GZ[o]);
    }
    atomicAdd(&dE[d * C + c], s);
  }
}

// out[p] = c0 + sum_{t,ci} V[t][ci] * Z[p + off_t][ci] (3 x 3 taps, one output channel, fp32 out)
__global__ void out_fwd_kernel(const uint16_t* Z, const float* V, const float* c0, float* out,
                               int B, int D, int I, int C) {
  const int64_t npos = (int64_t)B * D * I;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npos;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(p % I), d = (int)((p / I) % D);
    float s = c0[0];
    for (int t = 0; t < 9; ++t) {
      const int d2 = d + t / 3 - 1, i2 = i + t % 3 - 1;
      if (d2 < 0 || d2 >= D || i2 < 0 || i2 >= I) continue;
      const uint16_t* z = Z + (p + (int64_t)(t / 3 - 1) * I + (t % 3 - 1)) * C;
      for (int ci = 0; ci < C; ci += 8) {
        Pack8 v;
        v.u = *reinterpret_cast<const uint4*>(z + ci);
        for (int s8 = 0; s8 < 8; ++s8) s += V[t * C + ci + s8] * bf2f(v.h[s8]);
      }
    }
    out[p] = s;
  }
}

// The same with eight lanes per position, each reading one channel octet (16 bytes) of every
// tap, so a wave's loads cover eight neighbouring positions contiguously; the octets' partial
// sums meet in a butterfly. Weights V [9][C] are staged in LDS. C == 64.
__global__ void __launch_bounds__(256) out_fwd8_kernel(const uint16_t* Z, const float* V,
                                                       const float* c0, float* out, int B, int D,
                                                       int I) {
  constexpr int C = 64;
  __shared__ float Vs[9 * C];
  for (int k = threadIdx.x; k < 9 * C; k += blockDim.x) Vs[k] = V[k];
  __syncthreads();
  const int64_t npos = (int64_t)B * D * I;
  const int q = threadIdx.x & 7;
  const int64_t p = (int64_t)blockIdx.x * (blockDim.x / 8) + (threadIdx.x >> 3);
  float s = 0.0f;
  if (p < npos) {
    const int i = (int)(p % I), d = (int)((p / I) % D);
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int d2 = d + t / 3 - 1, i2 = i + t % 3 - 1;
      if (d2 < 0 || d2 >= D || i2 < 0 || i2 >= I) continue;
      Pack8 v;
      v.u = *reinterpret_cast<const uint4*>(Z + (p + (int64_t)(t / 3 - 1) * I + (t % 3 - 1)) * C +
                                            8 * q);
#pragma unroll
      for (int k = 0; k < 8; ++k) s += Vs[t * C + 8 * q + k] * bf2f(v.h[k]);
    }
  }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  s += __shfl_xor(s, 4, 64);
  if (p < npos && q == 0) out[p] = s + c0[0];
}

// The output layer in two passes over memory: out[p] = c0 + sum_t Y_t[p + off_t], with
// Y_t[q] = sum_c V[t][c] Z[q][c] the projection of each position onto the nine taps' weights.
// Each position's 64 channels are read once (out_fwd8_kernel reads every position nine times);
// the planes Y_t, 36 bytes per position, are summed with coalesced shifted reads.
// out_proj_kernel: eight lanes per position, lane q holding channel octet q (a wave reads eight
// neighbouring positions contiguously); the nine partial sums meet in DPP sums over the eight
// lanes, and lane q writes plane q (lane 0 also plane 8). Each lane group handles kProjRounds
// positions 32 apart, their loads issued together so that several are in flight per lane.
// C == 64.
constexpr int kProjRounds = 4;

// A lane's value moved along its row by a DPP pattern (a VALU move, no LDS).
template <int kCtrl>
__device__ inline float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xF, 0xF, false));
}

// The sum over each aligned group of eight lanes, in every lane of the group: quad_perm
// [1,0,3,2] and [2,3,0,1] sum each quad, row_half_mirror (lane i of eight from lane 7 - i) adds
// the other quad's sum.
__device__ inline float sum8_dpp(float v) {
  v += dpp_f<0xB1>(v);
  v += dpp_f<0x4E>(v);
  v += dpp_f<0x141>(v);
  return v;
}

__global__ void __launch_bounds__(256) out_proj_kernel(const uint16_t* Z, const float* V,
                                                       float* Y, int64_t npos) {
  constexpr int C = 64;
  const int q = threadIdx.x & 7;
  const int64_t p0 = (int64_t)blockIdx.x * (kProjRounds * 32) + (threadIdx.x >> 3);
  // the lane's weights, V[t][8q .. 8q + 7] for the nine taps, held in registers
  float w[9][8];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int k = 0; k < 8; ++k) w[t][k] = V[t * C + 8 * q + k];
  // every load of the lane is issued before the first is used
  Pack8 v[kProjRounds];
#pragma unroll
  for (int r = 0; r < kProjRounds; ++r) {
    const int64_t p = p0 + 32 * r;
    v[r].u = p < npos ? *reinterpret_cast<const uint4*>(Z + p * C + 8 * q) : make_uint4(0, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < kProjRounds; ++r) {
    const int64_t p = p0 + 32 * r;
    float y[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) y[t] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float z = bf2f(v[r].h[k]);
#pragma unroll
      for (int t = 0; t < 9; ++t) y[t] += w[t][k] * z;
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) y[t] = sum8_dpp(y[t]);
    if (p < npos) {
      float mine = y[0];
#pragma unroll
      for (int t = 1; t < 8; ++t) mine = q == t ? y[t] : mine;
      Y[(int64_t)q * npos + p] = mine;
      if (q == 0) Y[8 * npos + p] = y[8];
    }
  }
}

// out[p] = c0 + sum over the taps whose neighbour lies inside the [D][I] grid of Y_t[p + off_t]
__global__ void __launch_bounds__(256) out_sum_kernel(const float* Y, const float* c0, float* out,
                                                      int B, int D, int I) {
  const int64_t npos = (int64_t)B * D * I;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npos) return;
  const int i = (int)(p % I), d = (int)((p / I) % D);
  float s = c0[0];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int dd = t / 3 - 1, di = t % 3 - 1;
    if (d + dd < 0 || d + dd >= D || i + di < 0 || i + di >= I) continue;
    s += Y[(int64_t)t * npos + p + (int64_t)dd * I + di];
  }
  out[p] = s;
}

// loss: masked squared error; g[p] = scale * (out - y) on valid pixels (i + d < I), 0 elsewhere;
// sums of (out - y)^2 over valid pixels added to loss[0]
__global__ void loss_kernel(const float* out, const float* y, float* g, float scale, float* loss,
                            int B, int D, int I) {
  const int64_t npos = (int64_t)B * D * I;
  float acc = 0.0f;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npos;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(p % I), d = (int)((p / I) % D);
    if (i + d < I) {
      const float r = out[p] - y[p];
      acc += r * r;
      g[p] = scale * r;
    } else {
      g[p] = 0.0f;
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, acc);
}

// gZ[p][ci] = sum_t V[t][ci] * g[p - off_t]  (bf16 out)
__global__ void out_bwd_data_kernel(const float* g, const float* V, uint16_t* GZ, int B, int D,
                                    int I, int C) {
  const int64_t total = (int64_t)B * D * I * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int ci = (int)(e % C);
    const int64_t p = e / C;
    const int i = (int)(p % I), d = (int)((p / I) % D);
    float s = 0.0f;
    for (int t = 0; t < 9; ++t) {
      const int d2 = d - (t / 3 - 1), i2 = i - (t % 3 - 1);
      if (d2 < 0 || d2 >= D || i2 < 0 || i2 >= I) continue;
      s += V[t * C + ci] * g[p - (int64_t)(t / 3 - 1) * I - (t % 3 - 1)];
    }
    GZ[e] = f2bf(s);
  }
}

// dV[t][ci] += sum_p Z[p + off_t][ci] * g[p]; dc += sum_p g[p]. Block: a range of positions,
// threads over (t, ci) pairs.
__global__ void out_bwd_weight_kernel(const uint16_t* Z, const float* g, float* dV, float* dc,
                                      int B, int D, int I, int C, int64_t pos_per_block) {
  const int64_t npos = (int64_t)B * D * I;
  const int64_t p0 = (int64_t)blockIdx.x * pos_per_block;
  const int64_t p1 = p0 + pos_per_block < npos ? p0 + pos_per_block : npos;
  for (int e = threadIdx.x; e < 9 * C + 1; e += blockDim.x) {
    float s = 0.0f;
    if (e == 9 * C) {
      for (int64_t p = p0; p < p1; ++p) s += g[p];
      atomicAdd(dc, s);
      continue;
    }
    const int t = e / C, ci = e % C;
    const int dd = t / 3 - 1, di = t % 3 - 1;
    for (int64_t p = p0; p < p1; ++p) {
      const int i = (int)(p % I), d = (int)((p / I) % D);
      const int d2 = d + dd, i2 = i + di;
      if (d2 < 0 || d2 >= D || i2 < 0 || i2 >= I) continue;
      s += bf2f(Z[(p + (int64_t)dd * I + di) * C + ci]) * g[p];
    }
    atomicAdd(&dV[e], s);
  }
}

// Wt[ci][T-1-t][co] = W[co][t][ci] (bf16), the weights of the backward-data convolution
__global__ void flip_weights_kernel(const uint16_t* W, uint16_t* Wt, int Cout, int T, int Cin) {
  const int64_t total = (int64_t)Cout * T * Cin;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int ci = (int)(e % Cin);
    const int t = (int)((e / Cin) % T);
    const int co = (int)(e / ((int64_t)Cin * T));
    Wt[((int64_t)ci * T + (T - 1 - t)) * Cout + co] = W[e];
  }
}

// AdamW (decoupled weight decay) on fp32 masters, then the bf16 copy
__global__ void adamw_kernel(float* p, float* m, float* v, const float* g, uint16_t* pb, int64_t n,
                             float lr, float b1, float b2, float eps, float wd, float c1, float c2) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.0f - b1) * gi;
    const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    const float upd = (mi / c1) / (sqrtf(vi / c2) + eps);
    const float pi = p[e] - lr * (upd + wd * p[e]);
    p[e] = pi;
    if (pb) pb[e] = f2bf(pi);
  }
}

__global__ void to_bf16_kernel(const float* in, uint16_t* out, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = f2bf(in[e]);
}

__global__ void fill_kernel(float* p, float v, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    p[e] = v;
}

// Stitches the windows of whole chromosomes into their bands. Window w of chromosome q (of
// npoints in the batch, nwin windows each, starting at output bin starts[w]) predicted
// out[q * nwin + w][d][j] = log1p(10 * relative count) for pixel (starts[w] + j, + d); pixel
// (i, i + d) of the chromosome comes from the last window starting at or before i, and the band
// receives the relative count expm1(y) / 10 (0 beyond the chromosome). bands[q][d][i], nbins per
// diagonal, for q of the batch starting at bands.
__global__ void stitch_bands_kernel(const float* out, int npoints, int nwin, const uint32_t* starts,
                                    int ND, int Lout, int nbins, float* bands) {
  const int64_t n = (int64_t)npoints * ND * nbins;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % nbins);
    const int d = (int)((e / nbins) % ND);
    const int q = (int)(e / ((int64_t)nbins * ND));
    float v = 0.0f;
    if (i + d < nbins) {
      int w = 0;
      while (w + 1 < nwin && (int)starts[w + 1] <= i) ++w;
      const int j = i - (int)starts[w];
      if (j >= 0 && j + d < Lout)
        v = expm1f(out[(((int64_t)q * nwin + w) * ND + d) * Lout + j]) * 0.1f;
    }
    bands[e] = v;
  }
}

// gathers batch rows idx[0, nb) of a [N][row] array into a dense [nb][row] array
__global__ void gather_rows_u16_kernel(const uint16_t* src, const uint32_t* idx, uint16_t* dst,
                                       int nb, int64_t row) {
  const int64_t total = (int64_t)nb * row;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    dst[e] = src[(int64_t)idx[e / row] * row + e % row];
}

__global__ void gather_rows_f32_kernel(const float* src, const uint32_t* idx, float* dst, int nb,
                                       int64_t row) {
  const int64_t total = (int64_t)nb * row;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    dst[e] = src[(int64_t)idx[e / row] * row + e % row];
}

inline unsigned grid_for(int64_t n, int threads = 256) {
  const int64_t g = (n + threads - 1) / threads;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace

hipError_t launch_mfma_probe(hipStream_t s, const uint16_t* A, const uint16_t* B, float* D) {
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, A, B, D);
  return hipGetLastError();
}

hipError_t launch_conv_fwd(hipStream_t s, const ConvArgs& a) {
  if (a.Cin % 16 || a.Cout % 64 || a.KD < 1 || a.KI < 1 || !a.X || !a.W) return hipErrorInvalidValue;
  if (a.KD == 3 && a.KI == 3 && a.dil <= 1 && a.Cin == kT3C && a.Cout == kT3C) {
    const int64_t ntx = (a.I + kT3Cols - 1) / kT3Cols, nty = (a.D + kT3Rows - 1) / kT3Rows;
    const int64_t ntiles = ntx * nty * a.B;
    if (ntiles > INT32_MAX) return hipErrorInvalidValue;
    // one persistent workgroup per CU (the weights and a patch take most of its LDS)
    static int cus[64] = {}, lds_set[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64) return e != hipSuccess ? e : hipErrorInvalidDevice;
    if (!lds_set[dev]) {
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(conv3x3_c64_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, kT3Lds);
      if (e == hipSuccess)
        e = hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev);
      if (e != hipSuccess) return e;
      lds_set[dev] = 1;
    }
    const int grid = (int)std::min<int64_t>(ntiles, cus[dev] > 0 ? cus[dev] : 256);
    hipLaunchKernelGGL(conv3x3_c64_kernel, dim3(grid), dim3(256), kT3Lds, s, a, (int)ntx, (int)nty,
                       (int)ntiles);
    return hipGetLastError();
  }
  const int64_t npos = (int64_t)a.B * a.D * a.I;
  const dim3 grid((unsigned)((npos + kConvRows - 1) / kConvRows), (unsigned)(a.Cout / 64));
  hipLaunchKernelGGL(conv_fwd_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_conv_wgrad(hipStream_t s, const WgradArgs& a) {
  if (a.Cin % 32 || a.Cout % 32 || a.Cin > kWgMaxC || a.Cout > kWgMaxC ||
      (a.Cout / 32) * (a.Cin / 32) > 16 || a.pos_per_block <= 0)
    return hipErrorInvalidValue;
  const int64_t npos = (int64_t)a.B * a.D * a.I;
  const dim3 grid((unsigned)((npos + a.pos_per_block - 1) / a.pos_per_block),
                  (unsigned)(a.KD * a.KI));
  hipLaunchKernelGGL(conv_wgrad_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_colsum(hipStream_t s, const uint16_t* G, int64_t npos, int C, float* out) {
  const int64_t rows = 512;
  hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)((npos + rows - 1) / rows)),
                     dim3(C < 256 ? C : 256), 0, s, G, npos, C, rows, out);
  return hipGetLastError();
}

hipError_t launch_relu_mask(hipStream_t s, const uint16_t* g, const uint16_t* mask, uint16_t* out,
                            int64_t n) {
  hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_for(n)), dim3(256), 0, s, g, mask, out, n);
  return hipGetLastError();
}

hipError_t launch_crop(hipStream_t s, const uint16_t* in, uint16_t* out, int B, int n_in, int n_out,
                       int off, int C, bool uncrop) {
  const int64_t n = (int64_t)B * (uncrop ? n_in : n_out) * C;
  hipLaunchKernelGGL(crop_kernel, dim3(grid_for(n)), dim3(256), 0, s, in, out, B, n_in, n_out, off,
                     C, uncrop ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_pair_fwd(hipStream_t s, const uint16_t* P, const uint16_t* Q, const float* E,
                           uint16_t* Z, int B, int ND, int L, int C) {
  if (C % 8 == 0 && (int64_t)B * ND <= 65535) {
    const int per = L * (C / 8);
    hipLaunchKernelGGL(pair_fwd8_kernel, dim3((per + 255) / 256, B * ND), dim3(256), 0, s, P, Q, E,
                       Z, ND, L, C);
    return hipGetLastError();
  }
  const int64_t n = (int64_t)B * ND * L * C;
  hipLaunchKernelGGL(pair_fwd_kernel, dim3(grid_for(n)), dim3(256), 0, s, P, Q, E, Z, B, ND, L, C);
  return hipGetLastError();
}

hipError_t launch_pair_bwd(hipStream_t s, const uint16_t* GZ, const uint16_t* Z, uint16_t* dP,
                           uint16_t* dQ, float* dE, int B, int ND, int L, int C) {
  if (C > 256 || 256 % C) return hipErrorInvalidValue;
  const int64_t n = (int64_t)B * L * C;
  hipLaunchKernelGGL(pair_bwd_pq_kernel, dim3(grid_for(n)), dim3(256), 0, s, GZ, Z, dP, dQ, B, ND,
                     L, C);
  hipLaunchKernelGGL(pair_bwd_e_kernel, dim3(ND), dim3(256), 0, s, GZ, Z, dE, B, ND, L, C);
  return hipGetLastError();
}

hipError_t launch_out_fwd(hipStream_t s, const uint16_t* Z, const float* V, const float* c0,
                          float* out, int B, int D, int I, int C, float* scratch) {
  if (C % 8) return hipErrorInvalidValue;
  if (C == 64 && scratch) {
    const int64_t npos = (int64_t)B * D * I;
    hipLaunchKernelGGL(out_proj_kernel,
                       dim3((unsigned)((npos + kProjRounds * 32 - 1) / (kProjRounds * 32))),
                       dim3(256), 0, s, Z, V, scratch, npos);
    hipLaunchKernelGGL(out_sum_kernel, dim3((unsigned)((npos + 255) / 256)), dim3(256), 0, s,
                       scratch, c0, out, B, D, I);
    return hipGetLastError();
  }
  if (C == 64) {
    const int64_t npos = (int64_t)B * D * I;
    hipLaunchKernelGGL(out_fwd8_kernel, dim3((unsigned)((npos + 31) / 32)), dim3(256), 0, s, Z, V,
                       c0, out, B, D, I);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(out_fwd_kernel, dim3(grid_for((int64_t)B * D * I)), dim3(256), 0, s, Z, V, c0,
                     out, B, D, I, C);
  return hipGetLastError();
}

hipError_t launch_loss(hipStream_t s, const float* out, const float* y, float* g, float scale,
                       float* loss, int B, int D, int I) {
  hipLaunchKernelGGL(loss_kernel, dim3(grid_for((int64_t)B * D * I)), dim3(256), 0, s, out, y, g,
                     scale, loss, B, D, I);
  return hipGetLastError();
}

hipError_t launch_out_bwd(hipStream_t s, const uint16_t* Z, const float* g, const float* V,
                          uint16_t* GZ, float* dV, float* dc, int B, int D, int I, int C) {
  const int64_t n = (int64_t)B * D * I * C;
  hipLaunchKernelGGL(out_bwd_data_kernel, dim3(grid_for(n)), dim3(256), 0, s, g, V, GZ, B, D, I, C);
  const int64_t npos = (int64_t)B * D * I, ppb = 2048;
  hipLaunchKernelGGL(out_bwd_weight_kernel, dim3((unsigned)((npos + ppb - 1) / ppb)), dim3(256), 0,
                     s, Z, g, dV, dc, B, D, I, C, ppb);
  return hipGetLastError();
}

hipError_t launch_flip_weights(hipStream_t s, const uint16_t* W, uint16_t* Wt, int Cout, int T,
                               int Cin) {
  const int64_t n = (int64_t)Cout * T * Cin;
  hipLaunchKernelGGL(flip_weights_kernel, dim3(grid_for(n)), dim3(256), 0, s, W, Wt, Cout, T, Cin);
  return hipGetLastError();
}

hipError_t launch_adamw(hipStream_t s, float* p, float* m, float* v, const float* g, uint16_t* pb,
                        int64_t n, float lr, float b1, float b2, float eps, float wd, int step) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adamw_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, m, v, g, pb, n, lr, b1, b2,
                     eps, wd, c1, c2);
  return hipGetLastError();
}

hipError_t launch_to_bf16(hipStream_t s, const float* in, uint16_t* out, int64_t n) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_stitch_bands(hipStream_t s, const float* out, int npoints, int nwin,
                               const uint32_t* starts, int ND, int Lout, int nbins, float* bands) {
  const int64_t n = (int64_t)npoints * ND * nbins;
  hipLaunchKernelGGL(stitch_bands_kernel, dim3(grid_for(n)), dim3(256), 0, s, out, npoints, nwin,
                     starts, ND, Lout, nbins, bands);
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, float* p, float v, int64_t n) {
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_gather_rows_u16(hipStream_t s, const uint16_t* src, const uint32_t* idx,
                                  uint16_t* dst, int nb, int64_t row) {
  hipLaunchKernelGGL(gather_rows_u16_kernel, dim3(grid_for((int64_t)nb * row)), dim3(256), 0, s, src,
                     idx, dst, nb, row);
  return hipGetLastError();
}

hipError_t launch_gather_rows_f32(hipStream_t s, const float* src, const uint32_t* idx, float* dst,
                                  int nb, int64_t row) {
  hipLaunchKernelGGL(gather_rows_f32_kernel, dim3(grid_for((int64_t)nb * row)), dim3(256), 0, s, src,
                     idx, dst, nb, row);
  return hipGetLastError();
}

}  // namespace sg
}  // namespace mgpu
