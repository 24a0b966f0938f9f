// Surrogate network behind the C-ABI (mgpu_sg_*): parameters, activations and the training step.
//
// Architecture (per sample; bins along a window, channels innermost):
//   x      [Lin][in_ch]           barrier tracks + broadcast scalars (padded to 32 channels)
//   h1     = relu(conv1d_k(x))                                  enc_ch channels
//   h(l+1) = h(l) + relu(conv1d_k(h(l)))     l = 1 .. enc_layers-1 (layer l dilated by
//                                            enc_dilation[l])
//   hc     = h[out_offset, out_offset + Lout)                   crop to the output bins
//   P = hc Wp + bp, Q = hc Wq                                   head_ch channels (1 x 1)
//   z0[d][i] = relu(P[i] + Q[i + d] + E[d])  for i + d < Lout, else 0    (pairwise outer sum)
//   z(j+1) = z(j) + relu(conv3x3(z(j)))      j = 0 .. head_layers-2   on the (d, i) grid
//   out[d][i] = conv3x3(z(head_layers-1)) -> 1 channel: predicted log1p contacts of pixel
//               (i, i + d) of the output window
// Loss: mean squared error over the valid pixels (i + d < Lout).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/modle_gpu.h"
#include "sg_kernels.h"

using namespace mgpu::sg;

namespace {

constexpr int kInPad = 32;  // input channels padded for the MFMA tiles

struct Tensor {  // a span of the parameter pool
  uint64_t off = 0, n = 0;
  int cout = 0, taps = 0, cin = 0;  // conv weights: [cout][taps][cin]
};

}  // namespace

struct mgpu_sg {
  mgpu_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  mgpu_sg_config cfg{};
  int Lin = 0, Lout = 0, ND = 0, C = 0, H = 0, Le = 0, Lh = 0, K = 0;
  // parameters: fp32 masters, bf16 copies, gradients, AdamW moments (one pool each)
  std::vector<Tensor> enc_w, enc_b, head_w, head_b;
  Tensor wp, bp, wq, emb, vo, co;
  uint64_t nparam = 0;
  float *p32 = nullptr, *g32 = nullptr, *m32 = nullptr, *v32 = nullptr;
  uint16_t* p16 = nullptr;
  uint16_t* wt = nullptr;  // flipped weights of one layer
  int step = 0;
  // activations for max_batch samples
  int Bm = 0;
  uint16_t* x = nullptr;
  std::vector<uint16_t*> h, r;  // h[0..Le], r[1..Le-1]
  uint16_t *hc = nullptr, *P = nullptr, *Q = nullptr;
  std::vector<uint16_t*> z, r2;  // z[0..Lh-1], r2[0..Lh-2]
  float* out = nullptr;
  float* gout = nullptr;
  float* y = nullptr;
  float* loss = nullptr;
  uint16_t *gz = nullptr, *gz2 = nullptr, *gm = nullptr;
  uint16_t *dP = nullptr, *dQ = nullptr, *dhc = nullptr, *dhc2 = nullptr;
  uint16_t *gh = nullptr, *gh2 = nullptr, *gm1 = nullptr;
  std::vector<void*> allocs;
  // device-resident training set
  uint16_t* data_x = nullptr;
  float* data_y = nullptr;
  uint64_t ndata = 0;
  uint32_t* idx = nullptr;
  // stitched chromosome bands of mgpu_sg_predict_chromosomes (relative counts, device memory)
  float* bands = nullptr;
  uint64_t bands_points = 0, bands_nbins = 0;
  uint32_t* d_starts = nullptr;
  uint64_t starts_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
};

namespace {

thread_local std::string g_sg_err;

int sg_fail(mgpu_sg* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  g_sg_err = msg;
  return code;
}

#define SGCHK(s, expr)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return sg_fail(s, MGPU_EDEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint16_t to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

Tensor add_tensor(uint64_t& total, int cout, int taps, int cin) {
  Tensor t;
  t.off = total;
  t.cout = cout;
  t.taps = taps;
  t.cin = cin;
  t.n = (uint64_t)cout * taps * cin;
  total += (t.n + 63) & ~63ULL;  // 256-byte aligned spans
  return t;
}

void* dalloc(mgpu_sg* s, uint64_t bytes, hipError_t& e) {
  void* p = nullptr;
  if (e != hipSuccess) return nullptr;
  e = hipMalloc(&p, bytes < 256 ? 256 : bytes);
  if (e == hipSuccess) s->allocs.push_back(p);
  return p;
}

// canonical (unpadded) order of the parameters: for each tensor its [cout][taps][cin] entries,
// with the first encoder layer's cin the real input channel count
struct Span {
  uint64_t dev;    // offset in the padded pool
  uint64_t n;      // canonical entries
  int cin_real, cin_pad, rows;  // rows = cout * taps
};

std::vector<Span> canonical(const mgpu_sg* s) {
  std::vector<Span> v;
  auto add = [&](const Tensor& t, int cin_real) {
    v.push_back(Span{t.off, (uint64_t)t.cout * t.taps * cin_real, cin_real, t.cin, t.cout * t.taps});
  };
  for (int l = 0; l < s->Le; ++l) {
    add(s->enc_w[l], l == 0 ? (int)s->cfg.in_ch : s->enc_w[l].cin);
    add(s->enc_b[l], 1);
  }
  add(s->wp, s->wp.cin);
  add(s->bp, 1);
  add(s->wq, s->wq.cin);
  add(s->emb, s->H);
  for (int j = 0; j + 1 < s->Lh; ++j) {
    add(s->head_w[j], s->head_w[j].cin);
    add(s->head_b[j], 1);
  }
  add(s->vo, s->vo.cin);
  add(s->co, 1);
  return v;
}

std::vector<Tensor*> conv_tensors(mgpu_sg* s) {
  std::vector<Tensor*> v;
  for (auto& t : s->enc_w) v.push_back(&t);
  v.push_back(&s->wp);
  v.push_back(&s->wq);
  for (auto& t : s->head_w) v.push_back(&t);
  return v;
}

int dil_of(const mgpu_sg* s, int l) {
  return l < 16 && s->cfg.enc_dilation[l] > 0 ? (int)s->cfg.enc_dilation[l] : 1;
}

int64_t valid_pixels(const mgpu_sg* s) {
  int64_t n = 0;
  for (int d = 0; d < s->ND; ++d) n += std::max(0, s->Lout - d);
  return n;
}

// ---- forward over the first B samples of s->x ----
// keep_act: also store the post-ReLU activations (s->r, s->r2) that backward needs; inference
// leaves them out and saves their writes.
int forward(mgpu_sg* s, int B, bool keep_act = true) {
  hipStream_t st = s->stream;
  const uint16_t* w16 = s->p16;
  ConvArgs a{};
  a.B = B;
  a.D = 1;
  a.I = s->Lin;
  a.KD = 1;
  a.KI = s->K;
  a.relu = 1;
  for (int l = 0; l < s->Le; ++l) {
    a.dil = dil_of(s, l);
    a.X = l == 0 ? s->x : s->h[l];
    a.Cin = l == 0 ? kInPad : s->C;
    a.Cout = s->C;
    a.W = w16 + s->enc_w[l].off;
    a.bias = s->p32 + s->enc_b[l].off;
    if (l == 0) {
      a.res = nullptr;
      a.out1 = s->h[1];
      a.out2 = nullptr;
    } else {
      a.res = s->h[l];
      a.out1 = keep_act ? s->r[l] : nullptr;
      a.out2 = s->h[l + 1];
    }
    SGCHK(s, launch_conv_fwd(st, a));
  }
  SGCHK(s, launch_crop(st, s->h[s->Le], s->hc, B, s->Lin, s->Lout, (int)s->cfg.out_offset, s->C, false));
  ConvArgs pq{};
  pq.X = s->hc;
  pq.B = B;
  pq.D = 1;
  pq.I = s->Lout;
  pq.Cin = s->C;
  pq.Cout = s->H;
  pq.KD = pq.KI = 1;
  pq.relu = 0;
  pq.W = w16 + s->wp.off;
  pq.bias = s->p32 + s->bp.off;
  pq.out1 = s->P;
  SGCHK(s, launch_conv_fwd(st, pq));
  pq.W = w16 + s->wq.off;
  pq.bias = nullptr;
  pq.out1 = s->Q;
  SGCHK(s, launch_conv_fwd(st, pq));
  SGCHK(s, launch_pair_fwd(st, s->P, s->Q, s->p32 + s->emb.off, s->z[0], B, s->ND, s->Lout, s->H));
  ConvArgs hz{};
  hz.B = B;
  hz.D = s->ND;
  hz.I = s->Lout;
  hz.Cin = hz.Cout = s->H;
  hz.KD = hz.KI = 3;
  hz.relu = 1;
  for (int j = 0; j + 1 < s->Lh; ++j) {
    hz.X = s->z[j];
    hz.W = w16 + s->head_w[j].off;
    hz.bias = s->p32 + s->head_b[j].off;
    hz.res = s->z[j];
    hz.out1 = keep_act ? s->r2[j] : nullptr;
    hz.out2 = s->z[j + 1];
    SGCHK(s, launch_conv_fwd(st, hz));
  }
  // the tap projections of the output layer go to the gradient buffer, which only backward
  // (after this layer) writes: 36 bytes per position of its 2 * H
  SGCHK(s, launch_out_fwd(st, s->z[s->Lh - 1], s->p32 + s->vo.off, s->p32 + s->co.off, s->out, B,
                          s->ND, s->Lout, s->H, reinterpret_cast<float*>(s->gz)));
  return MGPU_OK;
}

// ---- loss and gradients (added to s->g32, which the caller zeroes) for targets s->y ----
int backward(mgpu_sg* s, int B) {
  hipStream_t st = s->stream;
  const uint16_t* w16 = s->p16;
  const float scale = 2.0f / (float)(valid_pixels(s) * B);
  SGCHK(s, launch_loss(st, s->out, s->y, s->gout, scale, s->loss, B, s->ND, s->Lout));
  SGCHK(s, launch_out_bwd(st, s->z[s->Lh - 1], s->gout, s->p32 + s->vo.off, s->gz,
                          s->g32 + s->vo.off, s->g32 + s->co.off, B, s->ND, s->Lout, s->H));
  const int64_t nz = (int64_t)B * s->ND * s->Lout * s->H;
  const int64_t pos2 = (int64_t)B * s->ND * s->Lout;
  for (int j = s->Lh - 2; j >= 0; --j) {
    SGCHK(s, launch_relu_mask(st, s->gz, s->r2[j], s->gm, nz));
    WgradArgs w{};
    w.G = s->gm;
    w.X = s->z[j];
    w.dW = s->g32 + s->head_w[j].off;
    w.B = B;
    w.D = s->ND;
    w.I = s->Lout;
    w.Cin = w.Cout = s->H;
    w.KD = w.KI = 3;
    w.pos_per_block = 4096;
    SGCHK(s, launch_conv_wgrad(st, w));
    SGCHK(s, launch_colsum(st, s->gm, pos2, s->H, s->g32 + s->head_b[j].off));
    SGCHK(s, launch_flip_weights(st, w16 + s->head_w[j].off, s->wt, s->H, 9, s->H));
    ConvArgs a{};
    a.X = s->gm;
    a.W = s->wt;
    a.res = s->gz;
    a.out2 = s->gz2;
    a.B = B;
    a.D = s->ND;
    a.I = s->Lout;
    a.Cin = a.Cout = s->H;
    a.KD = a.KI = 3;
    SGCHK(s, launch_conv_fwd(st, a));
    std::swap(s->gz, s->gz2);
  }
  SGCHK(s, launch_pair_bwd(st, s->gz, s->z[0], s->dP, s->dQ, s->g32 + s->emb.off, B, s->ND,
                           s->Lout, s->H));
  const int64_t pos1 = (int64_t)B * s->Lout;
  WgradArgs w{};
  w.X = s->hc;
  w.B = B;
  w.D = 1;
  w.I = s->Lout;
  w.Cin = s->C;
  w.Cout = s->H;
  w.KD = w.KI = 1;
  w.pos_per_block = 1024;
  w.G = s->dP;
  w.dW = s->g32 + s->wp.off;
  SGCHK(s, launch_conv_wgrad(st, w));
  SGCHK(s, launch_colsum(st, s->dP, pos1, s->H, s->g32 + s->bp.off));
  w.G = s->dQ;
  w.dW = s->g32 + s->wq.off;
  SGCHK(s, launch_conv_wgrad(st, w));
  ConvArgs a{};
  a.B = B;
  a.D = 1;
  a.I = s->Lout;
  a.Cin = s->H;
  a.Cout = s->C;
  a.KD = a.KI = 1;
  SGCHK(s, launch_flip_weights(st, w16 + s->wp.off, s->wt, s->H, 1, s->C));
  a.X = s->dP;
  a.W = s->wt;
  a.out1 = s->dhc;
  SGCHK(s, launch_conv_fwd(st, a));
  SGCHK(s, launch_flip_weights(st, w16 + s->wq.off, s->wt, s->H, 1, s->C));
  a.X = s->dQ;
  a.out1 = nullptr;
  a.res = s->dhc;
  a.out2 = s->dhc2;
  SGCHK(s, launch_conv_fwd(st, a));
  SGCHK(s, launch_crop(st, s->dhc2, s->gh, B, s->Lin, s->Lout, (int)s->cfg.out_offset, s->C, true));
  const int64_t nh = (int64_t)B * s->Lin * s->C;
  const int64_t pos0 = (int64_t)B * s->Lin;
  for (int l = s->Le - 1; l >= 0; --l) {
    SGCHK(s, launch_relu_mask(st, s->gh, l == 0 ? s->h[1] : s->r[l], s->gm1, nh));
    WgradArgs we{};
    we.G = s->gm1;
    we.X = l == 0 ? s->x : s->h[l];
    we.dW = s->g32 + s->enc_w[l].off;
    we.B = B;
    we.D = 1;
    we.I = s->Lin;
    we.Cin = l == 0 ? kInPad : s->C;
    we.Cout = s->C;
    we.KD = 1;
    we.KI = s->K;
    we.dil = dil_of(s, l);
    we.pos_per_block = 1024;
    SGCHK(s, launch_conv_wgrad(st, we));
    SGCHK(s, launch_colsum(st, s->gm1, pos0, s->C, s->g32 + s->enc_b[l].off));
    if (l == 0) break;
    SGCHK(s, launch_flip_weights(st, w16 + s->enc_w[l].off, s->wt, s->C, s->K, s->C));
    ConvArgs ae{};
    ae.X = s->gm1;
    ae.W = s->wt;
    ae.res = s->gh;
    ae.out2 = s->gh2;
    ae.B = B;
    ae.D = 1;
    ae.I = s->Lin;
    ae.Cin = ae.Cout = s->C;
    ae.KD = 1;
    ae.KI = s->K;
    ae.dil = dil_of(s, l);
    SGCHK(s, launch_conv_fwd(st, ae));
    std::swap(s->gh, s->gh2);
  }
  return MGPU_OK;
}

// host samples (fp32, [n][Lin][in_ch]) -> bf16 padded [n][Lin][32] in s->x
int upload_inputs(mgpu_sg* s, const float* x, int n) {
  std::vector<uint16_t> hx((size_t)n * s->Lin * kInPad, 0);
  for (int64_t b = 0; b < n; ++b)
    for (int l = 0; l < s->Lin; ++l)
      for (uint32_t c = 0; c < s->cfg.in_ch; ++c)
        hx[((size_t)b * s->Lin + l) * kInPad + c] =
            to_bf16(x[((size_t)b * s->Lin + l) * s->cfg.in_ch + c]);
  SGCHK(s, hipMemcpyAsync(s->x, hx.data(), hx.size() * 2, hipMemcpyHostToDevice, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  return MGPU_OK;
}

int check_cfg(const mgpu_sg_config* c) {
  if (!c) return sg_fail(nullptr, MGPU_EINVAL, "null config");
  if (c->in_ch == 0 || c->in_ch > (uint32_t)kInPad)
    return sg_fail(nullptr, MGPU_EINVAL, "in_ch must be in [1, 32]");
  if (c->enc_ch != 64 && c->enc_ch != 128)
    return sg_fail(nullptr, MGPU_EINVAL, "enc_ch must be 64 or 128");
  if (c->head_ch != 64) return sg_fail(nullptr, MGPU_EINVAL, "head_ch must be 64");
  if (c->enc_layers < 1 || c->head_layers < 1 || c->enc_k % 2 == 0 || c->enc_k > 9)
    return sg_fail(nullptr, MGPU_EINVAL, "invalid layer counts or encoder kernel width");
  if (c->out_len == 0 || c->out_offset + c->out_len > c->in_len || c->ndiag == 0 ||
      c->ndiag > c->out_len)
    return sg_fail(nullptr, MGPU_EINVAL, "output window outside the input or invalid band");
  if (c->max_batch == 0 || c->max_batch > 4096)
    return sg_fail(nullptr, MGPU_EINVAL, "max_batch must be in [1, 4096]");
  if (c->enc_layers > 16) return sg_fail(nullptr, MGPU_EINVAL, "at most 16 encoder layers");
  for (int l = 0; l < 16; ++l)
    if (c->enc_dilation[l] > 256) return sg_fail(nullptr, MGPU_EINVAL, "dilation above 256");
  return MGPU_OK;
}

}  // namespace

extern "C" {

int mgpu_sg_config_default(mgpu_sg_config* c) {
  if (!c) return sg_fail(nullptr, MGPU_EINVAL, "null config");
  *c = mgpu_sg_config{};
  c->in_len = 600;
  c->out_offset = 100;
  c->out_len = 400;
  c->ndiag = 100;
  c->in_ch = 6;
  c->enc_ch = 128;
  c->enc_layers = 8;
  c->enc_k = 5;
  c->head_ch = 64;
  c->head_layers = 4;
  c->max_batch = 32;
  c->seed = 1;
  const uint32_t dil[8] = {1, 1, 2, 4, 8, 16, 32, 1};
  for (int l = 0; l < 8; ++l) c->enc_dilation[l] = dil[l];
  return MGPU_OK;
}

int mgpu_sg_create(mgpu_ctx* ctx, const mgpu_sg_config* cfg, mgpu_sg** out) {
  if (!ctx || !out) return sg_fail(nullptr, MGPU_EINVAL, "null argument");
  if (int rc = check_cfg(cfg)) return rc;
  hipStream_t stream = nullptr;
  int device = 0;
  if (mgpu_context_stream(ctx, (void**)&stream, &device) != MGPU_OK)
    return sg_fail(nullptr, MGPU_EINVAL, "invalid context");
  (void)hipSetDevice(device);
  auto* s = new mgpu_sg();
  s->ctx = ctx;
  s->stream = stream;
  s->cfg = *cfg;
  s->Lin = (int)cfg->in_len;
  s->Lout = (int)cfg->out_len;
  s->ND = (int)cfg->ndiag;
  s->C = (int)cfg->enc_ch;
  s->H = (int)cfg->head_ch;
  s->Le = (int)cfg->enc_layers;
  s->Lh = (int)cfg->head_layers;
  s->K = (int)cfg->enc_k;
  s->Bm = (int)cfg->max_batch;

  uint64_t total = 0;
  for (int l = 0; l < s->Le; ++l) {
    s->enc_w.push_back(add_tensor(total, s->C, s->K, l == 0 ? kInPad : s->C));
    s->enc_b.push_back(add_tensor(total, s->C, 1, 1));
  }
  s->wp = add_tensor(total, s->H, 1, s->C);
  s->bp = add_tensor(total, s->H, 1, 1);
  s->wq = add_tensor(total, s->H, 1, s->C);
  s->emb = add_tensor(total, s->ND, 1, s->H);
  for (int j = 0; j + 1 < s->Lh; ++j) {
    s->head_w.push_back(add_tensor(total, s->H, 9, s->H));
    s->head_b.push_back(add_tensor(total, s->H, 1, 1));
  }
  s->vo = add_tensor(total, 1, 9, s->H);
  s->co = add_tensor(total, 1, 1, 1);
  s->nparam = total;

  hipError_t e = hipSuccess;
  s->p32 = static_cast<float*>(dalloc(s, total * 4, e));
  s->g32 = static_cast<float*>(dalloc(s, total * 4, e));
  s->m32 = static_cast<float*>(dalloc(s, total * 4, e));
  s->v32 = static_cast<float*>(dalloc(s, total * 4, e));
  s->p16 = static_cast<uint16_t*>(dalloc(s, total * 2, e));
  uint64_t maxw = 0;
  for (auto* t : conv_tensors(s)) maxw = std::max(maxw, t->n);
  s->wt = static_cast<uint16_t*>(dalloc(s, maxw * 2, e));
  const int64_t B = s->Bm;
  const int64_t n_in = B * s->Lin, n_out = B * s->Lout, n_z = B * s->ND * s->Lout;
  s->x = static_cast<uint16_t*>(dalloc(s, n_in * kInPad * 2, e));
  s->h.assign(s->Le + 1, nullptr);
  s->r.assign(s->Le, nullptr);
  for (int l = 1; l <= s->Le; ++l) s->h[l] = static_cast<uint16_t*>(dalloc(s, n_in * s->C * 2, e));
  for (int l = 1; l < s->Le; ++l) s->r[l] = static_cast<uint16_t*>(dalloc(s, n_in * s->C * 2, e));
  s->hc = static_cast<uint16_t*>(dalloc(s, n_out * s->C * 2, e));
  s->P = static_cast<uint16_t*>(dalloc(s, n_out * s->H * 2, e));
  s->Q = static_cast<uint16_t*>(dalloc(s, n_out * s->H * 2, e));
  s->z.assign(s->Lh, nullptr);
  s->r2.assign(s->Lh, nullptr);
  for (int j = 0; j < s->Lh; ++j) s->z[j] = static_cast<uint16_t*>(dalloc(s, n_z * s->H * 2, e));
  for (int j = 0; j + 1 < s->Lh; ++j) s->r2[j] = static_cast<uint16_t*>(dalloc(s, n_z * s->H * 2, e));
  s->out = static_cast<float*>(dalloc(s, n_z * 4, e));
  s->gout = static_cast<float*>(dalloc(s, n_z * 4, e));
  s->y = static_cast<float*>(dalloc(s, n_z * 4, e));
  s->loss = static_cast<float*>(dalloc(s, 256, e));
  s->gz = static_cast<uint16_t*>(dalloc(s, n_z * s->H * 2, e));
  s->gz2 = static_cast<uint16_t*>(dalloc(s, n_z * s->H * 2, e));
  s->gm = static_cast<uint16_t*>(dalloc(s, n_z * s->H * 2, e));
  s->dP = static_cast<uint16_t*>(dalloc(s, n_out * s->H * 2, e));
  s->dQ = static_cast<uint16_t*>(dalloc(s, n_out * s->H * 2, e));
  s->dhc = static_cast<uint16_t*>(dalloc(s, n_out * s->C * 2, e));
  s->dhc2 = static_cast<uint16_t*>(dalloc(s, n_out * s->C * 2, e));
  s->gh = static_cast<uint16_t*>(dalloc(s, n_in * s->C * 2, e));
  s->gh2 = static_cast<uint16_t*>(dalloc(s, n_in * s->C * 2, e));
  s->gm1 = static_cast<uint16_t*>(dalloc(s, n_in * s->C * 2, e));
  s->idx = static_cast<uint32_t*>(dalloc(s, B * 4, e));
  if (e == hipSuccess) e = hipEventCreate(&s->ev0);
  if (e == hipSuccess) e = hipEventCreate(&s->ev1);
  if (e != hipSuccess) {
    const std::string msg = std::string("surrogate allocation failed: ") + hipGetErrorString(e);
    mgpu_sg_destroy(s);
    return sg_fail(nullptr, MGPU_ENOMEM, msg);
  }

  // initial weights: uniform(-a, a) with a = sqrt(3 / fan_in) (residual branches scaled by 1/2,
  // the output layer by 1/10); biases and the diagonal embedding start at zero
  std::vector<float> hp(total, 0.0f);
  uint64_t st = cfg->seed;
  auto uni = [&]() { return (float)((splitmix(st) >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f; };
  auto init = [&](const Tensor& t, int cin_real, float gain) {
    const float a = gain * std::sqrt(3.0f / (float)(t.taps * cin_real));
    for (int o = 0; o < t.cout; ++o)
      for (int k = 0; k < t.taps; ++k)
        for (int c = 0; c < cin_real; ++c)
          hp[t.off + ((uint64_t)o * t.taps + k) * t.cin + c] = a * uni();
  };
  for (int l = 0; l < s->Le; ++l)
    init(s->enc_w[l], l == 0 ? (int)cfg->in_ch : s->C, l == 0 ? std::sqrt(2.0f) : 0.5f);
  init(s->wp, s->C, 1.0f);
  init(s->wq, s->C, 1.0f);
  for (int j = 0; j + 1 < s->Lh; ++j) init(s->head_w[j], s->H, 0.5f);
  init(s->vo, s->H, 0.1f);
  e = hipMemcpy(s->p32, hp.data(), total * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(s->m32, 0, total * 4);
  if (e == hipSuccess) e = hipMemset(s->v32, 0, total * 4);
  if (e == hipSuccess) e = launch_to_bf16(s->stream, s->p32, s->p16, (int64_t)total);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    const std::string msg = std::string("surrogate initialisation failed: ") + hipGetErrorString(e);
    mgpu_sg_destroy(s);
    return sg_fail(nullptr, MGPU_EDEVICE, msg);
  }
  *out = s;
  return MGPU_OK;
}

void mgpu_sg_destroy(mgpu_sg* s) {
  if (!s) return;
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->data_x) (void)hipFree(s->data_x);
  if (s->data_y) (void)hipFree(s->data_y);
  if (s->bands) (void)hipFree(s->bands);
  if (s->d_starts) (void)hipFree(s->d_starts);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  delete s;
}

const char* mgpu_sg_last_error(const mgpu_sg* s) { return s ? s->err.c_str() : g_sg_err.c_str(); }

uint64_t mgpu_sg_num_params(const mgpu_sg* s) {
  if (!s) return 0;
  uint64_t n = 0;
  for (const auto& sp : canonical(s)) n += sp.n;
  return n;
}

int mgpu_sg_get_params(mgpu_sg* s, float* out) {
  if (!s || !out) return sg_fail(s, MGPU_EINVAL, "null argument");
  std::vector<float> hp(s->nparam);
  SGCHK(s, hipMemcpy(hp.data(), s->p32, s->nparam * 4, hipMemcpyDeviceToHost));
  uint64_t k = 0;
  for (const auto& sp : canonical(s))
    for (int row = 0; row < sp.rows; ++row)
      for (int c = 0; c < sp.cin_real; ++c) out[k++] = hp[sp.dev + (uint64_t)row * sp.cin_pad + c];
  return MGPU_OK;
}

int mgpu_sg_set_params(mgpu_sg* s, const float* in) {
  if (!s || !in) return sg_fail(s, MGPU_EINVAL, "null argument");
  std::vector<float> hp(s->nparam, 0.0f);
  uint64_t k = 0;
  for (const auto& sp : canonical(s))
    for (int row = 0; row < sp.rows; ++row)
      for (int c = 0; c < sp.cin_real; ++c) hp[sp.dev + (uint64_t)row * sp.cin_pad + c] = in[k++];
  SGCHK(s, hipMemcpy(s->p32, hp.data(), s->nparam * 4, hipMemcpyHostToDevice));
  SGCHK(s, hipMemset(s->m32, 0, s->nparam * 4));
  SGCHK(s, hipMemset(s->v32, 0, s->nparam * 4));
  s->step = 0;
  SGCHK(s, launch_to_bf16(s->stream, s->p32, s->p16, (int64_t)s->nparam));
  SGCHK(s, hipStreamSynchronize(s->stream));
  return MGPU_OK;
}

int mgpu_sg_forward(mgpu_sg* s, const float* x, uint64_t n, float* y) {
  if (!s || !x || !y) return sg_fail(s, MGPU_EINVAL, "null argument");
  const int64_t per = (int64_t)s->ND * s->Lout;
  for (uint64_t b0 = 0; b0 < n; b0 += s->Bm) {
    const int B = (int)std::min<uint64_t>(s->Bm, n - b0);
    if (int rc = upload_inputs(s, x + b0 * s->Lin * s->cfg.in_ch, B)) return rc;
    if (int rc = forward(s, B, false)) return rc;
    SGCHK(s, hipMemcpyAsync(y + b0 * per, s->out, (size_t)B * per * 4, hipMemcpyDeviceToHost,
                            s->stream));
    SGCHK(s, hipStreamSynchronize(s->stream));
  }
  return MGPU_OK;
}

int mgpu_sg_gradients(mgpu_sg* s, const float* x, const float* y, uint64_t n, float* loss,
                      float* grads) {
  if (!s || !x || !y) return sg_fail(s, MGPU_EINVAL, "null argument");
  if (n == 0 || n > (uint64_t)s->Bm) return sg_fail(s, MGPU_EINVAL, "n must be in [1, max_batch]");
  const int B = (int)n;
  const int64_t per = (int64_t)s->ND * s->Lout;
  if (int rc = upload_inputs(s, x, B)) return rc;
  SGCHK(s, hipMemcpyAsync(s->y, y, (size_t)B * per * 4, hipMemcpyHostToDevice, s->stream));
  SGCHK(s, launch_fill(s->stream, s->g32, 0.0f, (int64_t)s->nparam));
  SGCHK(s, launch_fill(s->stream, s->loss, 0.0f, 1));
  if (int rc = forward(s, B)) return rc;
  if (int rc = backward(s, B)) return rc;
  float l = 0.0f;
  SGCHK(s, hipMemcpyAsync(&l, s->loss, 4, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  if (loss) *loss = l / (float)(valid_pixels(s) * B);
  if (grads) {
    std::vector<float> hg(s->nparam);
    SGCHK(s, hipMemcpy(hg.data(), s->g32, s->nparam * 4, hipMemcpyDeviceToHost));
    uint64_t k = 0;
    for (const auto& sp : canonical(s))
      for (int row = 0; row < sp.rows; ++row)
        for (int c = 0; c < sp.cin_real; ++c) grads[k++] = hg[sp.dev + (uint64_t)row * sp.cin_pad + c];
  }
  return MGPU_OK;
}

int mgpu_sg_set_data(mgpu_sg* s, const float* x, const float* y, uint64_t n) {
  if (!s || !x || !y || n == 0) return sg_fail(s, MGPU_EINVAL, "null argument or empty data");
  if (n > 0xFFFFFFFFull) return sg_fail(s, MGPU_ERANGE, "too many samples");
  if (s->data_x) (void)hipFree(s->data_x);
  if (s->data_y) (void)hipFree(s->data_y);
  s->data_x = nullptr;
  s->data_y = nullptr;
  s->ndata = 0;
  const int64_t per_x = (int64_t)s->Lin * kInPad, per_y = (int64_t)s->ND * s->Lout;
  SGCHK(s, hipMalloc(&s->data_x, n * per_x * 2));
  SGCHK(s, hipMalloc(&s->data_y, n * per_y * 4));
  constexpr uint64_t kChunk = 1024;  // samples converted and uploaded per copy
  std::vector<uint16_t> hx((size_t)(kChunk * per_x), 0);
  for (uint64_t b0 = 0; b0 < n; b0 += kChunk) {
    const uint64_t nb = std::min(kChunk, n - b0);
    for (uint64_t b = 0; b < nb; ++b)
      for (int l = 0; l < s->Lin; ++l)
        for (uint32_t c = 0; c < s->cfg.in_ch; ++c)
          hx[(size_t)(b * per_x + (uint64_t)l * kInPad + c)] =
              to_bf16(x[((b0 + b) * s->Lin + l) * s->cfg.in_ch + c]);
    SGCHK(s, hipMemcpy(s->data_x + b0 * per_x, hx.data(), nb * per_x * 2, hipMemcpyHostToDevice));
  }
  SGCHK(s, hipMemcpy(s->data_y, y, n * per_y * 4, hipMemcpyHostToDevice));
  s->ndata = n;
  return MGPU_OK;
}

int mgpu_sg_train_step(mgpu_sg* s, const uint32_t* idx, uint64_t n, float lr, float weight_decay,
                       float* loss, float* ms) {
  if (!s || !idx) return sg_fail(s, MGPU_EINVAL, "null argument");
  if (n == 0 || n > (uint64_t)s->Bm) return sg_fail(s, MGPU_EINVAL, "n must be in [1, max_batch]");
  if (!s->ndata) return sg_fail(s, MGPU_EINVAL, "no training data (mgpu_sg_set_data)");
  for (uint64_t k = 0; k < n; ++k)
    if (idx[k] >= s->ndata) return sg_fail(s, MGPU_ERANGE, "sample index out of range");
  const int B = (int)n;
  const int64_t per_x = (int64_t)s->Lin * kInPad, per_y = (int64_t)s->ND * s->Lout;
  hipStream_t st = s->stream;
  SGCHK(s, hipMemcpyAsync(s->idx, idx, n * 4, hipMemcpyHostToDevice, st));
  SGCHK(s, hipEventRecord(s->ev0, st));
  SGCHK(s, launch_gather_rows_u16(st, s->data_x, s->idx, s->x, B, per_x));
  SGCHK(s, launch_gather_rows_f32(st, s->data_y, s->idx, s->y, B, per_y));
  SGCHK(s, launch_fill(st, s->g32, 0.0f, (int64_t)s->nparam));
  SGCHK(s, launch_fill(st, s->loss, 0.0f, 1));
  if (int rc = forward(s, B)) return rc;
  if (int rc = backward(s, B)) return rc;
  ++s->step;
  SGCHK(s, launch_adamw(st, s->p32, s->m32, s->v32, s->g32, s->p16, (int64_t)s->nparam, lr, 0.9f,
                        0.999f, 1e-8f, weight_decay, s->step));
  SGCHK(s, hipEventRecord(s->ev1, st));
  float l = 0.0f;
  SGCHK(s, hipMemcpyAsync(&l, s->loss, 4, hipMemcpyDeviceToHost, st));
  SGCHK(s, hipStreamSynchronize(st));
  if (loss) *loss = l / (float)(valid_pixels(s) * B);
  if (ms) SGCHK(s, hipEventElapsedTime(ms, s->ev0, s->ev1));
  return MGPU_OK;
}

int mgpu_sg_adamw_step(mgpu_sg* s, const float* grads, float lr, float weight_decay) {
  if (!s || !grads) return sg_fail(s, MGPU_EINVAL, "null argument");
  std::vector<float> hg(s->nparam, 0.0f);
  uint64_t k = 0;
  for (const auto& sp : canonical(s))
    for (int row = 0; row < sp.rows; ++row)
      for (int c = 0; c < sp.cin_real; ++c) hg[sp.dev + (uint64_t)row * sp.cin_pad + c] = grads[k++];
  SGCHK(s, hipMemcpy(s->g32, hg.data(), s->nparam * 4, hipMemcpyHostToDevice));
  ++s->step;
  SGCHK(s, launch_adamw(s->stream, s->p32, s->m32, s->v32, s->g32, s->p16, (int64_t)s->nparam, lr,
                        0.9f, 0.999f, 1e-8f, weight_decay, s->step));
  SGCHK(s, hipStreamSynchronize(s->stream));
  return MGPU_OK;
}

int mgpu_sg_reserve_bands(mgpu_sg* s, uint64_t npoints, uint64_t nbins) {
  if (!s) return sg_fail(s, MGPU_EINVAL, "null argument");
  if (npoints == 0 || nbins == 0) return sg_fail(s, MGPU_EINVAL, "empty band reservation");
  if (s->bands) SGCHK(s, hipFree(s->bands));
  s->bands = nullptr;
  s->bands_points = s->bands_nbins = 0;
  const uint64_t bytes = npoints * (uint64_t)s->ND * nbins * 4;
  hipError_t e = hipMalloc(&s->bands, bytes);
  if (e != hipSuccess)
    return sg_fail(s, MGPU_ENOMEM, "band reservation of " + std::to_string(bytes) + " bytes: " +
                                       hipGetErrorString(e));
  s->bands_points = npoints;
  s->bands_nbins = nbins;
  return MGPU_OK;
}

int mgpu_sg_predict_chromosomes(mgpu_sg* s, const float* x, uint64_t first_point,
                                uint64_t npoints, uint64_t nwin, const uint32_t* starts) {
  if (!s || !x || !starts) return sg_fail(s, MGPU_EINVAL, "null argument");
  if (!s->bands) return sg_fail(s, MGPU_EINVAL, "no band reservation (mgpu_sg_reserve_bands)");
  if (nwin == 0 || nwin > (uint64_t)s->Bm)
    return sg_fail(s, MGPU_EINVAL, "windows per chromosome must be in [1, max_batch]");
  if (first_point + npoints > s->bands_points)
    return sg_fail(s, MGPU_ERANGE, "points beyond the band reservation");
  const uint64_t nbins = s->bands_nbins;
  for (uint64_t w = 0; w < nwin; ++w)
    if ((w && starts[w] < starts[w - 1]) || starts[w] >= nbins || starts[0] != 0)
      return sg_fail(s, MGPU_EINVAL, "window starts must rise from 0 inside the chromosome");
  if (nwin > s->starts_cap) {
    if (s->d_starts) SGCHK(s, hipFree(s->d_starts));
    s->d_starts = nullptr;
    s->starts_cap = 0;
    SGCHK(s, hipMalloc(&s->d_starts, nwin * 4));
    s->starts_cap = nwin;
  }
  SGCHK(s, hipMemcpyAsync(s->d_starts, starts, nwin * 4, hipMemcpyHostToDevice, s->stream));
  const uint64_t ppb = std::max<uint64_t>(1, (uint64_t)s->Bm / nwin);  // points per batch
  const int64_t per_win = (int64_t)s->Lin * s->cfg.in_ch;
  for (uint64_t p0 = 0; p0 < npoints; p0 += ppb) {
    const uint64_t np = std::min<uint64_t>(ppb, npoints - p0);
    const int B = (int)(np * nwin);
    if (int rc = upload_inputs(s, x + p0 * nwin * per_win, B)) return rc;
    if (int rc = forward(s, B, false)) return rc;
    SGCHK(s, launch_stitch_bands(s->stream, s->out, (int)np, (int)nwin, s->d_starts, s->ND,
                                 s->Lout, (int)nbins,
                                 s->bands + (first_point + p0) * (uint64_t)s->ND * nbins));
  }
  SGCHK(s, hipStreamSynchronize(s->stream));
  return MGPU_OK;
}

int mgpu_sg_read_bands(mgpu_sg* s, uint64_t first_point, uint64_t npoints, float* out) {
  if (!s || !out) return sg_fail(s, MGPU_EINVAL, "null argument");
  if (first_point + npoints > s->bands_points)
    return sg_fail(s, MGPU_ERANGE, "points beyond the band reservation");
  const uint64_t per = (uint64_t)s->ND * s->bands_nbins;
  SGCHK(s, hipMemcpy(out, s->bands + first_point * per, npoints * per * 4, hipMemcpyDeviceToHost));
  return MGPU_OK;
}

int mgpu_sg_predict_timed(mgpu_sg* s, uint64_t n, uint64_t repeats, float* ms) {
  if (!s || !ms) return sg_fail(s, MGPU_EINVAL, "null argument");
  if (n == 0 || n > (uint64_t)s->Bm) return sg_fail(s, MGPU_EINVAL, "n must be in [1, max_batch]");
  SGCHK(s, hipEventRecord(s->ev0, s->stream));
  for (uint64_t r = 0; r < repeats; ++r)
    if (int rc = forward(s, (int)n, false)) return rc;
  SGCHK(s, hipEventRecord(s->ev1, s->stream));
  SGCHK(s, hipEventSynchronize(s->ev1));
  SGCHK(s, hipEventElapsedTime(ms, s->ev0, s->ev1));
  return MGPU_OK;
}

int mgpu_debug_mfma(mgpu_ctx* ctx, const uint16_t* a, const uint16_t* b, float* d) {
  if (!ctx || !a || !b || !d) return MGPU_EINVAL;
  hipStream_t stream = nullptr;
  int device = 0;
  if (mgpu_context_stream(ctx, (void**)&stream, &device) != MGPU_OK) return MGPU_EINVAL;
  (void)hipSetDevice(device);
  uint8_t* buf = nullptr;
  if (hipMalloc(&buf, 1024 + 1024 + 4096) != hipSuccess) return MGPU_ENOMEM;
  hipError_t e = hipMemcpy(buf, a, 1024, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(buf + 1024, b, 1024, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_mfma_probe(stream, reinterpret_cast<uint16_t*>(buf),
                          reinterpret_cast<uint16_t*>(buf + 1024), reinterpret_cast<float*>(buf + 2048));
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e == hipSuccess) e = hipMemcpy(d, buf + 2048, 4096, hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  return e == hipSuccess ? MGPU_OK : MGPU_EDEVICE;
}

}  // extern "C"
