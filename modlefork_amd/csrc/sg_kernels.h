// Device kernels of the surrogate network (sg_kernels.hip). bf16 values travel as uint16_t.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgpu {
namespace sg {

// out[p][co] = act(bias[co] + sum_{t,ci} W[co][t][ci] * X[p + off_t][ci]) on a B x D x I grid;
// out1 receives the activation, out2 (if set) res + activation. Cin % 16 == 0, Cout % 64 == 0.
struct ConvArgs {
  const uint16_t* X;
  const uint16_t* W;
  const float* bias;  // may be null
  const uint16_t* res;  // may be null
  uint16_t* out1;       // may be null
  uint16_t* out2;       // may be null
  int B, D, I;
  int Cin, Cout;
  int KD, KI;
  int dil;  // dilation along I (tap ti reads i + (ti - KI/2) * dil); 0 means 1
  int relu;
};

// dW[co][t][ci] += sum_p G[p][co] * X[p + off_t][ci] (fp32 atomics). Cin, Cout multiples of 32,
// at most 128 each.
struct WgradArgs {
  const uint16_t* G;
  const uint16_t* X;
  float* dW;
  int B, D, I;
  int Cin, Cout;
  int KD, KI;
  int dil;  // as in ConvArgs
  int64_t pos_per_block;
};

hipError_t launch_mfma_probe(hipStream_t s, const uint16_t* A, const uint16_t* B, float* D);
hipError_t launch_conv_fwd(hipStream_t s, const ConvArgs& a);
hipError_t launch_conv_wgrad(hipStream_t s, const WgradArgs& a);
hipError_t launch_colsum(hipStream_t s, const uint16_t* G, int64_t npos, int C, float* out);
hipError_t launch_relu_mask(hipStream_t s, const uint16_t* g, const uint16_t* mask, uint16_t* out,
                            int64_t n);
hipError_t launch_crop(hipStream_t s, const uint16_t* in, uint16_t* out, int B, int n_in, int n_out,
                       int off, int C, bool uncrop);
hipError_t launch_pair_fwd(hipStream_t s, const uint16_t* P, const uint16_t* Q, const float* E,
                           uint16_t* Z, int B, int ND, int L, int C);
hipError_t launch_pair_bwd(hipStream_t s, const uint16_t* GZ, const uint16_t* Z, uint16_t* dP,
                           uint16_t* dQ, float* dE, int B, int ND, int L, int C);
// scratch: 9 * B * D * I floats (the tap projections) for the two-pass C == 64 path; null
// takes the one-pass kernel.
hipError_t launch_out_fwd(hipStream_t s, const uint16_t* Z, const float* V, const float* c0,
                          float* out, int B, int D, int I, int C, float* scratch = nullptr);
hipError_t launch_loss(hipStream_t s, const float* out, const float* y, float* g, float scale,
                       float* loss, int B, int D, int I);
hipError_t launch_out_bwd(hipStream_t s, const uint16_t* Z, const float* g, const float* V,
                          uint16_t* GZ, float* dV, float* dc, int B, int D, int I, int C);
hipError_t launch_flip_weights(hipStream_t s, const uint16_t* W, uint16_t* Wt, int Cout, int T,
                               int Cin);
hipError_t launch_adamw(hipStream_t s, float* p, float* m, float* v, const float* g, uint16_t* pb,
                        int64_t n, float lr, float b1, float b2, float eps, float wd, int step);
hipError_t launch_to_bf16(hipStream_t s, const float* in, uint16_t* out, int64_t n);
hipError_t launch_fill(hipStream_t s, float* p, float v, int64_t n);
hipError_t launch_stitch_bands(hipStream_t s, const float* out, int npoints, int nwin,
                               const uint32_t* starts, int ND, int Lout, int nbins, float* bands);
hipError_t launch_gather_rows_u16(hipStream_t s, const uint16_t* src, const uint32_t* idx,
                                  uint16_t* dst, int nb, int64_t row);
hipError_t launch_gather_rows_f32(hipStream_t s, const float* src, const uint32_t* idx, float* dst,
                                  int nb, int64_t row);

}  // namespace sg
}  // namespace mgpu
