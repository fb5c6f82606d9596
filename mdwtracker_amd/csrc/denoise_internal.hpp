// Host-side launchers of the denoiser kernels (kernels_denoise.hip), used by
// denoise_api.cpp. Tensors are NHWC: element (pixel p, channel c) of a tensor
// with channel stride s and channel offset o is at p * s + o + c. bf16 values
// are stored as uint16_t.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace mdwt {
namespace dn {

// values per pixel of the network input for a volume of `channels` channels:
// rgb, first-hit depth, transmittance, the hit fraction of each channel, the
// standard error of the pixel mean and log2 of the frame count (both scaled)
constexpr int input_width(int channels) { return 7 + channels; }
// kernel-predicting output: kKernel x kKernel filter weights per pixel
constexpr int kKernelRadius = 3;
constexpr int kKernel = 2 * kKernelRadius + 1;
constexpr int kTaps = kKernel * kKernel;
constexpr int kLogitStride = 64;  // logits per pixel, padded to the conv's N tiles
static_assert(kTaps <= kLogitStride, "logit stride");

// 3x3 convolution (zero padding) as an implicit GEMM on the matrix cores.
// w: bf16 [9][cout_pad][cin]; cin % 8 == 0, cout_pad % 32 == 0, xs and xo
// multiples of 8. Writes channels < cout of y16 (bf16) and/or y32 (fp32,
// added when accum).
hipError_t conv3x3(const uint16_t* x, int H, int W, int xs, int xo, int cin, const uint16_t* w,
                   const float* bias, int cout, int cout_pad, uint16_t* y16, float* y32, int ys,
                   int yo, bool relu, bool accum, hipStream_t s);
// The same convolution with a choice of input source and epilogue:
// kSrcPlain reads x; kSrcUpCat reads channel c < cu from xl (H/2 x W/2,
// stride cu) at (y / 2, x / 2) and channel c >= cu from x at c - cu (the
// U-Net's nearest upsample and concatenation); kSrcF32 reads the fp32
// network input kin (inputs values per pixel, zero up to cin) as
// pack_input would convert it. kEpiPlain stores y16 / y32;
// kEpiPool also stores the 2 x 2 mean of y16 into pool16 (stride ps), as
// pool() would; kEpiKpn (cout_pad == kLogitStride) applies the
// kernel-predicting filter to kin and writes pred instead of the logits.
// H and W even for kSrcUpCat and kEpiPool; xs, xo, cu multiples of 8.
constexpr int kSrcPlain = 0, kSrcUpCat = 1, kSrcF32 = 2;
constexpr int kEpiPlain = 0, kEpiPool = 1, kEpiKpn = 2;
struct ConvArgs {
  const uint16_t* x;
  int xs, xo;
  const uint16_t* xl;
  int cu;
  int H, W, cin;
  const uint16_t* w;
  const float* bias;
  int cout, cout_pad;
  uint16_t* y16;
  float* y32;
  int ys, yo, relu, accum;
  uint16_t* pool16;
  int ps;
  const float* kin;
  int inputs;
  float* pred;
};
hipError_t conv(const ConvArgs& a, int src, int epi, hipStream_t s);
// dw [9][cout][cin] += sum over pixels of g[p][co] * x[p + tap][ci]; part is
// scratch of part_cap floats (at least 9 * cout * cin). Strides and xo are
// multiples of 4 and the rows hold cin (cout) rounded up to 4 channels.
hipError_t wgrad(const uint16_t* g, int gs, const uint16_t* x, int xs, int xo, int H, int W,
                 int cin, int cout, float* dw, float* part, size_t part_cap, hipStream_t s);
// gm = bf16(gy * relu'(y)) [npix][cout_pad]; db += per-channel sums (partial
// sums of 4096-pixel chunks in part, at least ceil(npix / 4096) * cout of
// part_cap floats, added in chunk order)
hipError_t relu_grad(const float* gy, int gs, const uint16_t* y, int ys, int yo, int npix, int cout,
                     int cout_pad, bool relu, uint16_t* gm, float* db, float* part,
                     size_t part_cap, hipStream_t s);
hipError_t pool(const uint16_t* x, int H, int W, int C, uint16_t* y, hipStream_t s);
hipError_t pool_grad(const float* gy, int H, int W, int C, float* gx, hipStream_t s);
hipError_t up(const uint16_t* x, int H, int W, int C, uint16_t* y, int ys, int yo, hipStream_t s);
hipError_t up_grad(const float* gy, int H, int W, int ys, int yo, int C, float* gx, hipStream_t s);
hipError_t copy_slice(const uint16_t* x, int xs, int xo, int n, int C, uint16_t* y, int ys, int yo,
                      hipStream_t s);
hipError_t add_slice(const float* gy, int ys, int yo, int n, int C, float* gx, hipStream_t s);
// in: fp32 [n][inputs] -> x: bf16 [n][in_pad], channels from inputs up zero
hipError_t pack_input(const float* in, int n, int inputs, int in_pad, uint16_t* x, hipStream_t s);
// pred [H*W][3] = softmax(logits) over the kTaps neighbourhood of the input
// colour (in: fp32 [H*W][inputs], rgb first; logits: fp32 [H*W][kLogitStride])
hipError_t kpn_apply(const float* in, int inputs, const float* logits, int H, int W, float* pred,
                     hipStream_t s);
// err += sum (pred - target)^2; gz = d(inv_count * that sum) / d logits
hipError_t kpn_loss_grad(const float* in, int inputs, const float* logits, const float* pred,
                         const float* target, int H, int W, float inv_count, float* gz, float* err,
                         hipStream_t s);
hipError_t adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                float c1, float c2, float eps, uint16_t* w16, hipStream_t s);
hipError_t weight_layouts(const float* w, int cin, int cout, int cin_pad, int cout_pad,
                          uint16_t* wf, int cin_pad32, int cout_pad8, uint16_t* wd, hipStream_t s);

}  // namespace dn
}  // namespace mdwt
