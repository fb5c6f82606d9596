// Learned denoiser beside plain accumulation: a 3-level U-Net (3x3
// convolutions, 32 / 64 / 96 channels, 2x2 average pooling, nearest
// upsampling with skip concatenation, ReLU), BF16 weights and activations,
// FP32 accumulation on the matrix cores, trained in place with Adam on pairs
// of low- and high-sample renders. Its output is kernel-predicting (Bako et
// al. 2017): per pixel, softmax weights of a 7x7 filter applied to the input
// colour. Input per pixel (7 + channels values): rgb, first-hit depth,
// transmittance, the hit fraction of each channel (mdwt_set_aov), the
// standard error of the pixel mean and log2 of the frame count.
// No reference counterpart; C ABI in include/mdwt_denoise.h.
#include "mdwt_denoise.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "denoise_internal.hpp"

namespace {

struct Conv {
  int cin = 0, cout = 0;          // real channel counts
  int cin_pad = 0, cout_pad = 0;  // forward layout (K multiple of 8, N multiple of 32)
  int cin_pad32 = 0;              // dgrad N (multiple of 32)
  bool relu = true;
  size_t nw = 0;                  // 9 * cout * cin
  float *w = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr;
  float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
  uint16_t *wf = nullptr, *wd = nullptr;
};

int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

struct mdwt_denoiser {
  std::string err;
  int device = 0;
  int H = 0, W = 0;
  int channels = 4, inputs = 11, in_pad = 16;  // input values per pixel, padded to 8
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Conv c[11];
  std::vector<void*> bufs;
  // activations (bf16)
  uint16_t *x0, *a0, *a1, *p1, *a2, *a3, *p2, *a4, *a5, *u2, *a6, *a7, *u1, *a8, *a9, *gm;
  // fp32: network io and gradients
  float *in32, *tgt, *out32, *pred, *err_sum;
  float* wpart;  // partial weight gradients of the row bands
  float *gout, *g9, *g8, *gu1, *g7, *g6, *gu2, *g5, *g4, *gp2, *g3, *g2, *gp1, *g1, *g0;
  int step = 0;
  bool fused = true;  // inference: fused upsample/concat, pooling and filter
};

namespace {

int fail(mdwt_denoiser* d, int code, const std::string& m) {
  d->err = m;
  return code;
}
#define DNCHK(d, expr)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail((d), MDWT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int alloc(mdwt_denoiser* d, T** p, size_t n) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return fail(d, MDWT_ERR_OOM, "denoiser allocation");
  d->bufs.push_back(q);
  DNCHK(d, hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T)));
  *p = static_cast<T*>(q);
  return MDWT_OK;
}

// layer table: (cin, cout, relu); c0 reads the input (cin set per handle),
// c10 writes the kTaps filter logits
// largest image (pixels) a denoiser handle accepts
constexpr size_t kMaxPixels = size_t(1) << 24;

const int kLayers[11][3] = {{0, 32, 1}, {32, 32, 1}, {32, 64, 1},
                            {64, 64, 1},  {64, 96, 1}, {96, 96, 1}, {160, 64, 1},
                            {64, 64, 1},  {96, 32, 1}, {32, 32, 1}, {32, mdwt::dn::kTaps, 0}};

// width in pixels of the Gaussian filter the untrained network applies
constexpr float kInitSigma = 1.5f;
// floats of partial weight gradients (16 MiB): at least 9 * cout * cin of
// every layer, and room for enough row bands to fill the GPU
constexpr size_t kWgradPart = size_t(4) << 20;

int setup(mdwt_denoiser* d, uint32_t seed) {
  std::mt19937 gen(seed);
  for (int i = 0; i < 11; ++i) {
    Conv& L = d->c[i];
    L.cin = i == 0 ? d->inputs : kLayers[i][0];
    L.cout = kLayers[i][1];
    L.relu = kLayers[i][2] != 0;
    L.cin_pad = i == 0 ? d->in_pad : round_up(L.cin, 8);
    L.cout_pad = round_up(L.cout, 32);
    L.cin_pad32 = round_up(L.cin, 32);
    L.nw = size_t(9) * L.cout * L.cin;
    int rc;
    if ((rc = alloc(d, &L.w, L.nw)) || (rc = alloc(d, &L.gw, L.nw)) || (rc = alloc(d, &L.mw, L.nw)) ||
        (rc = alloc(d, &L.vw, L.nw)) || (rc = alloc(d, &L.b, size_t(L.cout))) ||
        (rc = alloc(d, &L.gb, size_t(L.cout))) || (rc = alloc(d, &L.mb, size_t(L.cout))) ||
        (rc = alloc(d, &L.vb, size_t(L.cout))) ||
        (rc = alloc(d, &L.wf, size_t(9) * L.cout_pad * L.cin_pad)) ||
        (rc = alloc(d, &L.wd, size_t(9) * L.cin_pad32 * L.cout_pad)))
      return rc;
    // He initialisation for ReLU layers; the output layer starts small
    // with Gaussian log-weights as bias, so that the untrained network
    // filters with a Gaussian of kInitSigma pixels
    const float sd = L.relu ? std::sqrt(2.f / float(9 * L.cin)) : 1e-3f;
    std::normal_distribution<float> nd(0.f, sd);
    std::vector<float> hw(L.nw);
    for (auto& v : hw) v = nd(gen);
    DNCHK(d, hipMemcpy(L.w, hw.data(), L.nw * sizeof(float), hipMemcpyHostToDevice));
    if (!L.relu) {
      std::vector<float> hb(L.cout);
      for (int j = 0; j < L.cout; ++j) {
        const float dy = float(j / mdwt::dn::kKernel - mdwt::dn::kKernelRadius);
        const float dx = float(j % mdwt::dn::kKernel - mdwt::dn::kKernelRadius);
        hb[j] = -(dx * dx + dy * dy) / (2.f * kInitSigma * kInitSigma);
      }
      DNCHK(d, hipMemcpy(L.b, hb.data(), hb.size() * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  const size_t n = size_t(d->H) * d->W, n4 = n / 4, n16 = n / 16;
  int rc;
  if ((rc = alloc(d, &d->x0, n * d->in_pad)) || (rc = alloc(d, &d->a0, n * 32)) ||
      (rc = alloc(d, &d->a1, n * 32)) || (rc = alloc(d, &d->p1, n4 * 32)) ||
      (rc = alloc(d, &d->a2, n4 * 64)) || (rc = alloc(d, &d->a3, n4 * 64)) ||
      (rc = alloc(d, &d->p2, n16 * 64)) || (rc = alloc(d, &d->a4, n16 * 96)) ||
      (rc = alloc(d, &d->a5, n16 * 96)) || (rc = alloc(d, &d->u2, n4 * 160)) ||
      (rc = alloc(d, &d->a6, n4 * 64)) || (rc = alloc(d, &d->a7, n4 * 64)) ||
      (rc = alloc(d, &d->u1, n * 96)) || (rc = alloc(d, &d->a8, n * 32)) ||
      (rc = alloc(d, &d->a9, n * 32)) || (rc = alloc(d, &d->gm, n * mdwt::dn::kLogitStride)))
    return rc;
  if ((rc = alloc(d, &d->in32, n * d->inputs)) || (rc = alloc(d, &d->tgt, n * 3)) ||
      (rc = alloc(d, &d->out32, n * mdwt::dn::kLogitStride)) || (rc = alloc(d, &d->pred, n * 3)) ||
      (rc = alloc(d, &d->err_sum, 1)) || (rc = alloc(d, &d->wpart, kWgradPart)) ||
      (rc = alloc(d, &d->gout, n * mdwt::dn::kLogitStride)) ||
      (rc = alloc(d, &d->g9, n * 32)) || (rc = alloc(d, &d->g8, n * 32)) ||
      (rc = alloc(d, &d->gu1, n * 96)) || (rc = alloc(d, &d->g7, n4 * 64)) ||
      (rc = alloc(d, &d->g6, n4 * 64)) || (rc = alloc(d, &d->gu2, n4 * 160)) ||
      (rc = alloc(d, &d->g5, n16 * 96)) || (rc = alloc(d, &d->g4, n16 * 96)) ||
      (rc = alloc(d, &d->gp2, n16 * 64)) || (rc = alloc(d, &d->g3, n4 * 64)) ||
      (rc = alloc(d, &d->g2, n4 * 64)) || (rc = alloc(d, &d->gp1, n4 * 32)) ||
      (rc = alloc(d, &d->g1, n * 32)) || (rc = alloc(d, &d->g0, n * 32)))
    return rc;
  return MDWT_OK;
}

int refresh_weights(mdwt_denoiser* d) {
  for (Conv& L : d->c)
    DNCHK(d, mdwt::dn::weight_layouts(L.w, L.cin, L.cout, L.cin_pad, L.cout_pad, L.wf,
                                      L.cin_pad32, L.cout_pad, L.wd, d->stream));
  return MDWT_OK;
}

// conv layer i forward: x (stride xs, offset xo) -> y16 / y32 (stride ys)
int fwd(mdwt_denoiser* d, int i, const uint16_t* x, int H, int W, int xs, int xo, uint16_t* y16,
        float* y32, int ys) {
  const Conv& L = d->c[i];
  DNCHK(d, mdwt::dn::conv3x3(x, H, W, xs, xo, L.cin_pad, L.wf, L.b, L.cout, L.cout_pad, y16, y32,
                             ys, 0, L.relu, false, d->stream));
  return MDWT_OK;
}

int forward(mdwt_denoiser* d) {
  const int H = d->H, W = d->W, n = H * W;
  const int H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
  hipStream_t s = d->stream;
  int rc;
  DNCHK(d, mdwt::dn::pack_input(d->in32, n, d->inputs, d->in_pad, d->x0, s));
  if ((rc = fwd(d, 0, d->x0, H, W, d->in_pad, 0, d->a0, nullptr, 32))) return rc;
  if ((rc = fwd(d, 1, d->a0, H, W, 32, 0, d->a1, nullptr, 32))) return rc;
  DNCHK(d, mdwt::dn::pool(d->a1, H, W, 32, d->p1, s));
  if ((rc = fwd(d, 2, d->p1, H2, W2, 32, 0, d->a2, nullptr, 64))) return rc;
  if ((rc = fwd(d, 3, d->a2, H2, W2, 64, 0, d->a3, nullptr, 64))) return rc;
  DNCHK(d, mdwt::dn::pool(d->a3, H2, W2, 64, d->p2, s));
  if ((rc = fwd(d, 4, d->p2, H4, W4, 64, 0, d->a4, nullptr, 96))) return rc;
  if ((rc = fwd(d, 5, d->a4, H4, W4, 96, 0, d->a5, nullptr, 96))) return rc;
  DNCHK(d, mdwt::dn::up(d->a5, H2, W2, 96, d->u2, 160, 0, s));
  DNCHK(d, mdwt::dn::copy_slice(d->a3, 64, 0, H2 * W2, 64, d->u2, 160, 96, s));
  if ((rc = fwd(d, 6, d->u2, H2, W2, 160, 0, d->a6, nullptr, 64))) return rc;
  if ((rc = fwd(d, 7, d->a6, H2, W2, 64, 0, d->a7, nullptr, 64))) return rc;
  DNCHK(d, mdwt::dn::up(d->a7, H, W, 64, d->u1, 96, 0, s));
  DNCHK(d, mdwt::dn::copy_slice(d->a1, 32, 0, n, 32, d->u1, 96, 64, s));
  if ((rc = fwd(d, 8, d->u1, H, W, 96, 0, d->a8, nullptr, 32))) return rc;
  if ((rc = fwd(d, 9, d->a8, H, W, 32, 0, d->a9, nullptr, 32))) return rc;
  if ((rc = fwd(d, 10, d->a9, H, W, 32, 0, nullptr, d->out32, mdwt::dn::kLogitStride))) return rc;
  DNCHK(d, mdwt::dn::kpn_apply(d->in32, d->inputs, d->out32, H, W, d->pred, s));
  return MDWT_OK;
}

// conv layer i forward with a fused input source / epilogue (ConvArgs)
int fwd_fused(mdwt_denoiser* d, int i, mdwt::dn::ConvArgs a, int src, int epi) {
  const Conv& L = d->c[i];
  a.cin = L.cin_pad;
  a.w = L.wf;
  a.bias = L.b;
  a.cout = L.cout;
  a.cout_pad = L.cout_pad;
  a.relu = L.relu ? 1 : 0;
  DNCHK(d, mdwt::dn::conv(a, src, epi, d->stream));
  return MDWT_OK;
}

// Inference forward: the same values as forward() (every fused step
// computes what the kernel it replaces would have stored), without
// materialising the packed input, the pooled, upsampled and concatenated
// tensors or the logits: c0 converts the fp32 input while staging it, c1
// and c3 pool in their epilogue, c6 and c8 read their two inputs directly,
// c10 applies the filter in its epilogue.
int forward_fused(mdwt_denoiser* d) {
  using namespace mdwt::dn;
  const int H = d->H, W = d->W, n = H * W;
  const int H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
  int rc;
  auto args = [](const uint16_t* x, int xs, int Hh, int Ww, uint16_t* y16, int ys) {
    ConvArgs a{};
    a.x = x;
    a.xs = xs;
    a.H = Hh;
    a.W = Ww;
    a.y16 = y16;
    a.ys = ys;
    return a;
  };
  ConvArgs a = args(nullptr, 0, H, W, d->a0, 32);  // c0 reads the fp32 input
  a.kin = d->in32;
  a.inputs = d->inputs;
  if ((rc = fwd_fused(d, 0, a, kSrcF32, kEpiPlain))) return rc;
  a = args(d->a0, 32, H, W, d->a1, 32);
  a.pool16 = d->p1;
  a.ps = 32;
  if ((rc = fwd_fused(d, 1, a, kSrcPlain, kEpiPool))) return rc;
  if ((rc = fwd_fused(d, 2, args(d->p1, 32, H2, W2, d->a2, 64), kSrcPlain, kEpiPlain))) return rc;
  a = args(d->a2, 64, H2, W2, d->a3, 64);
  a.pool16 = d->p2;
  a.ps = 64;
  if ((rc = fwd_fused(d, 3, a, kSrcPlain, kEpiPool))) return rc;
  if ((rc = fwd_fused(d, 4, args(d->p2, 64, H4, W4, d->a4, 96), kSrcPlain, kEpiPlain))) return rc;
  if ((rc = fwd_fused(d, 5, args(d->a4, 96, H4, W4, d->a5, 96), kSrcPlain, kEpiPlain))) return rc;
  a = args(d->a3, 64, H2, W2, d->a6, 64);  // channels 0..95: up(a5), 96..159: a3
  a.xl = d->a5;
  a.cu = 96;
  if ((rc = fwd_fused(d, 6, a, kSrcUpCat, kEpiPlain))) return rc;
  if ((rc = fwd_fused(d, 7, args(d->a6, 64, H2, W2, d->a7, 64), kSrcPlain, kEpiPlain))) return rc;
  a = args(d->a1, 32, H, W, d->a8, 32);  // channels 0..63: up(a7), 64..95: a1
  a.xl = d->a7;
  a.cu = 64;
  if ((rc = fwd_fused(d, 8, a, kSrcUpCat, kEpiPlain))) return rc;
  if ((rc = fwd_fused(d, 9, args(d->a8, 32, H, W, d->a9, 32), kSrcPlain, kEpiPlain))) return rc;
  a = args(d->a9, 32, H, W, nullptr, kLogitStride);
  a.kin = d->in32;
  a.inputs = d->inputs;
  a.pred = d->pred;
  return fwd_fused(d, 10, a, kSrcPlain, kEpiKpn);
}

// backward of conv layer i: gy (fp32, stride gs) of its output y (bf16,
// stride ys); input x (stride xs, offset xo); gx (fp32, stride gxs) gets
// the input gradient (skipped when null)
int bwd(mdwt_denoiser* d, int i, const float* gy, int gs, const uint16_t* y, int ys,
        const uint16_t* x, int xs, int xo, int H, int W, float* gx, int gxs) {
  const Conv& L = d->c[i];
  hipStream_t s = d->stream;
  // the bias partial sums use the weight-gradient scratch, reduced before wgrad runs
  DNCHK(d, mdwt::dn::relu_grad(gy, gs, y, ys, 0, H * W, L.cout, L.cout_pad, L.relu, d->gm, L.gb,
                               d->wpart, kWgradPart, s));
  DNCHK(d, mdwt::dn::wgrad(d->gm, L.cout_pad, x, xs, xo, H, W, L.cin, L.cout, L.gw, d->wpart,
                           kWgradPart, s));
  if (gx)
    DNCHK(d, mdwt::dn::conv3x3(d->gm, H, W, L.cout_pad, 0, L.cout_pad, L.wd, nullptr, L.cin,
                               L.cin_pad32, nullptr, gx, gxs, 0, false, false, s));
  return MDWT_OK;
}

int backward(mdwt_denoiser* d) {
  const int H = d->H, W = d->W, n = H * W;
  const int H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
  hipStream_t s = d->stream;
  int rc;
  DNCHK(d, hipMemsetAsync(d->g1, 0, size_t(n) * 32 * sizeof(float), s));
  DNCHK(d, hipMemsetAsync(d->g3, 0, size_t(H2) * W2 * 64 * sizeof(float), s));
  if ((rc = bwd(d, 10, d->gout, mdwt::dn::kLogitStride, nullptr, 0, d->a9, 32, 0, H, W, d->g9,
                32)))
    return rc;
  if ((rc = bwd(d, 9, d->g9, 32, d->a9, 32, d->a8, 32, 0, H, W, d->g8, 32))) return rc;
  if ((rc = bwd(d, 8, d->g8, 32, d->a8, 32, d->u1, 96, 0, H, W, d->gu1, 96))) return rc;
  DNCHK(d, mdwt::dn::up_grad(d->gu1, H, W, 96, 0, 64, d->g7, s));
  DNCHK(d, mdwt::dn::add_slice(d->gu1, 96, 64, n, 32, d->g1, s));
  if ((rc = bwd(d, 7, d->g7, 64, d->a7, 64, d->a6, 64, 0, H2, W2, d->g6, 64))) return rc;
  if ((rc = bwd(d, 6, d->g6, 64, d->a6, 64, d->u2, 160, 0, H2, W2, d->gu2, 160))) return rc;
  DNCHK(d, mdwt::dn::up_grad(d->gu2, H2, W2, 160, 0, 96, d->g5, s));
  DNCHK(d, mdwt::dn::add_slice(d->gu2, 160, 96, H2 * W2, 64, d->g3, s));
  if ((rc = bwd(d, 5, d->g5, 96, d->a5, 96, d->a4, 96, 0, H4, W4, d->g4, 96))) return rc;
  if ((rc = bwd(d, 4, d->g4, 96, d->a4, 96, d->p2, 64, 0, H4, W4, d->gp2, 64))) return rc;
  DNCHK(d, mdwt::dn::pool_grad(d->gp2, H2, W2, 64, d->g3, s));
  if ((rc = bwd(d, 3, d->g3, 64, d->a3, 64, d->a2, 64, 0, H2, W2, d->g2, 64))) return rc;
  if ((rc = bwd(d, 2, d->g2, 64, d->a2, 64, d->p1, 32, 0, H2, W2, d->gp1, 32))) return rc;
  DNCHK(d, mdwt::dn::pool_grad(d->gp1, H, W, 32, d->g1, s));
  if ((rc = bwd(d, 1, d->g1, 32, d->a1, 32, d->a0, 32, 0, H, W, d->g0, 32))) return rc;
  if ((rc = bwd(d, 0, d->g0, 32, d->a0, 32, d->x0, d->in_pad, 0, H, W, nullptr, 0))) return rc;
  return MDWT_OK;
}

int zero_grads(mdwt_denoiser* d) {
  for (Conv& L : d->c) {
    DNCHK(d, hipMemsetAsync(L.gw, 0, L.nw * sizeof(float), d->stream));
    DNCHK(d, hipMemsetAsync(L.gb, 0, size_t(L.cout) * sizeof(float), d->stream));
  }
  return MDWT_OK;
}

// upload one (input, target) pair and accumulate its loss gradient
int one_pair(mdwt_denoiser* d, const float* in, const float* tgt, float inv_count) {
  const size_t n = size_t(d->H) * d->W;
  DNCHK(d, hipMemcpyAsync(d->in32, in, n * d->inputs * sizeof(float),
                          hipMemcpyHostToDevice, d->stream));
  DNCHK(d, hipMemcpyAsync(d->tgt, tgt, n * 3 * sizeof(float), hipMemcpyHostToDevice, d->stream));
  int rc = forward(d);
  if (rc) return rc;
  DNCHK(d, mdwt::dn::kpn_loss_grad(d->in32, d->inputs, d->out32, d->pred, d->tgt, d->H, d->W,
                                   inv_count, d->gout, d->err_sum, d->stream));
  return backward(d);
}

}  // namespace

extern "C" {

int mdwt_dn_create(mdwt_denoiser** out, int hip_device, int width, int height, int channels,
                   uint32_t seed) {
  if (!out) return MDWT_ERR_ARG;
  *out = nullptr;
  if (width < 4 || height < 4 || width % 4 || height % 4 || width > 8192 || height > 8192 ||
      channels < 1 || channels > 12)
    return MDWT_ERR_ARG;
  // The element-wise kernels index pixel x channel in 32-bit ints; the widest
  // full-resolution tensor has 96 channels (the level-0 concatenation), so
  // 2^24 pixels keep every index below 2^31.
  if (size_t(width) * size_t(height) > kMaxPixels) return MDWT_ERR_LIMIT;
  auto* d = new (std::nothrow) mdwt_denoiser();
  if (!d) return MDWT_ERR_OOM;
  d->device = hip_device;
  d->W = width;
  d->H = height;
  d->channels = channels;
  d->inputs = mdwt::dn::input_width(channels);
  d->in_pad = round_up(d->inputs, 8);
  int rc = MDWT_OK;
  if (hipSetDevice(hip_device) != hipSuccess ||
      hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&d->ev0) != hipSuccess || hipEventCreate(&d->ev1) != hipSuccess)
    rc = MDWT_ERR_HIP;
  if (rc == MDWT_OK) rc = setup(d, seed);
  if (rc == MDWT_OK) rc = refresh_weights(d);
  if (rc == MDWT_OK && hipStreamSynchronize(d->stream) != hipSuccess) rc = MDWT_ERR_HIP;
  if (rc != MDWT_OK) {
    mdwt_dn_destroy(d);
    return rc;
  }
  *out = d;
  return MDWT_OK;
}

int mdwt_dn_destroy(mdwt_denoiser* d) {
  if (!d) return MDWT_ERR_ARG;
  (void)hipSetDevice(d->device);
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  for (void* p : d->bufs) (void)hipFree(p);
  if (d->ev0) (void)hipEventDestroy(d->ev0);
  if (d->ev1) (void)hipEventDestroy(d->ev1);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
  return MDWT_OK;
}

const char* mdwt_dn_last_error(const mdwt_denoiser* d) { return d ? d->err.c_str() : ""; }

int mdwt_dn_num_params(const mdwt_denoiser* d, int* n) {
  if (!d || !n) return MDWT_ERR_ARG;
  size_t t = 0;
  for (const Conv& L : d->c) t += L.nw + size_t(L.cout);
  *n = int(t);
  return MDWT_OK;
}

int mdwt_dn_get_params(mdwt_denoiser* d, float* out) {
  if (!d || !out) return MDWT_ERR_ARG;
  DNCHK(d, hipSetDevice(d->device));
  DNCHK(d, hipStreamSynchronize(d->stream));
  for (const Conv& L : d->c) {
    DNCHK(d, hipMemcpy(out, L.w, L.nw * sizeof(float), hipMemcpyDeviceToHost));
    out += L.nw;
    DNCHK(d, hipMemcpy(out, L.b, size_t(L.cout) * sizeof(float), hipMemcpyDeviceToHost));
    out += L.cout;
  }
  return MDWT_OK;
}

int mdwt_dn_set_params(mdwt_denoiser* d, const float* in) {
  if (!d || !in) return MDWT_ERR_ARG;
  DNCHK(d, hipSetDevice(d->device));
  DNCHK(d, hipStreamSynchronize(d->stream));
  for (Conv& L : d->c) {
    DNCHK(d, hipMemcpy(L.w, in, L.nw * sizeof(float), hipMemcpyHostToDevice));
    in += L.nw;
    DNCHK(d, hipMemcpy(L.b, in, size_t(L.cout) * sizeof(float), hipMemcpyHostToDevice));
    in += L.cout;
  }
  int rc = refresh_weights(d);
  if (rc) return rc;
  DNCHK(d, hipStreamSynchronize(d->stream));
  return MDWT_OK;
}

int mdwt_dn_infer(mdwt_denoiser* d, const float* input, float* output, float* ms) {
  if (!d || !input || !output) return MDWT_ERR_ARG;
  DNCHK(d, hipSetDevice(d->device));
  const size_t n = size_t(d->H) * d->W;
  DNCHK(d, hipMemcpyAsync(d->in32, input, n * d->inputs * sizeof(float),
                          hipMemcpyHostToDevice, d->stream));
  DNCHK(d, hipEventRecord(d->ev0, d->stream));
  int rc = d->fused ? forward_fused(d) : forward(d);
  if (rc) return rc;
  DNCHK(d, hipEventRecord(d->ev1, d->stream));
  DNCHK(d, hipMemcpyAsync(output, d->pred, n * 3 * sizeof(float), hipMemcpyDeviceToHost, d->stream));
  DNCHK(d, hipStreamSynchronize(d->stream));
  float t = 0.f;
  DNCHK(d, hipEventElapsedTime(&t, d->ev0, d->ev1));
  if (ms) *ms = t;
  return MDWT_OK;
}

int mdwt_dn_set_fused(mdwt_denoiser* d, int on) {
  if (!d) return MDWT_ERR_ARG;
  d->fused = on != 0;
  return MDWT_OK;
}

int mdwt_dn_gradient(mdwt_denoiser* d, const float* input, const float* target, float* grad,
                     float* loss) {
  if (!d || !input || !target || !grad) return MDWT_ERR_ARG;
  DNCHK(d, hipSetDevice(d->device));
  const size_t n = size_t(d->H) * d->W;
  int rc = zero_grads(d);
  if (rc) return rc;
  DNCHK(d, hipMemsetAsync(d->err_sum, 0, sizeof(float), d->stream));
  if ((rc = one_pair(d, input, target, 1.f / float(n * 3)))) return rc;
  DNCHK(d, hipStreamSynchronize(d->stream));
  for (const Conv& L : d->c) {
    DNCHK(d, hipMemcpy(grad, L.gw, L.nw * sizeof(float), hipMemcpyDeviceToHost));
    grad += L.nw;
    DNCHK(d, hipMemcpy(grad, L.gb, size_t(L.cout) * sizeof(float), hipMemcpyDeviceToHost));
    grad += L.cout;
  }
  if (loss) {
    float e = 0.f;
    DNCHK(d, hipMemcpy(&e, d->err_sum, sizeof(float), hipMemcpyDeviceToHost));
    *loss = e / float(n * 3);
  }
  return MDWT_OK;
}

int mdwt_dn_train_step(mdwt_denoiser* d, const float* inputs, const float* targets,
                       const float* weights, int count, float lr, float* loss) {
  if (!d || !inputs || !targets || count < 1 || !(lr > 0.f)) return MDWT_ERR_ARG;
  for (int i = 0; weights && i < count; ++i)
    if (!(weights[i] >= 0.f) || std::isinf(weights[i])) return MDWT_ERR_ARG;
  DNCHK(d, hipSetDevice(d->device));
  const size_t n = size_t(d->H) * d->W;
  int rc = zero_grads(d);
  if (rc) return rc;
  DNCHK(d, hipMemsetAsync(d->err_sum, 0, sizeof(float), d->stream));
  const float inv = 1.f / float(n * 3 * size_t(count));
  for (int i = 0; i < count; ++i) {
    const float* xi = inputs + size_t(i) * n * d->inputs;
    const float wi = weights ? inv * weights[i] : inv;
    if ((rc = one_pair(d, xi, targets + size_t(i) * n * 3, wi))) return rc;
    DNCHK(d, hipStreamSynchronize(d->stream));  // host inputs are reused by the caller
  }
  ++d->step;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, float(d->step)), c2 = 1.f - std::pow(b2, float(d->step));
  for (Conv& L : d->c) {
    DNCHK(d, mdwt::dn::adam(L.w, L.gw, L.mw, L.vw, int(L.nw), lr, b1, b2, c1, c2, eps, nullptr,
                            d->stream));
    DNCHK(d, mdwt::dn::adam(L.b, L.gb, L.mb, L.vb, L.cout, lr, b1, b2, c1, c2, eps, nullptr,
                            d->stream));
  }
  rc = refresh_weights(d);
  if (rc) return rc;
  float e = 0.f;
  DNCHK(d, hipMemcpyAsync(&e, d->err_sum, sizeof(float), hipMemcpyDeviceToHost, d->stream));
  DNCHK(d, hipStreamSynchronize(d->stream));
  if (loss) *loss = e / float(n * 3 * size_t(count));
  return MDWT_OK;
}

}  // extern "C"
