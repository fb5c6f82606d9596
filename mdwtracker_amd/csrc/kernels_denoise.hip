// Kernels of the learned denoiser (a 3-level U-Net of 3x3 convolutions): BF16
// activations and weights, FP32 accumulation on the matrix cores
// (v_mfma_f32_32x32x8bf16_1k), FP32 gradients and Adam state. No reference
// counterpart (the reference renders plain accumulation only).
//
// Tensors are NHWC with a channel stride (so that a convolution can read or
// write a slice of a concatenated tensor). A 3x3 convolution is an implicit
// GEMM: M = pixels, N = output channels, K = 9 taps x input channels. One
// wave computes a 32-pixel x 32-channel tile; the MFMA lane layout is
//   A (32 x 8):  lane l holds A[l % 32][(l / 32) * 4 + j], j = 0..3
//   B (8 x 32):  lane l holds B[(l / 32) * 4 + j][l % 32]
//   D (32 x 32): lane l, register v holds D[(v / 4) * 8 + (l / 32) * 4 + v % 4][l % 32]
// (checked on the device with known matrices).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "denoise_internal.hpp"

namespace mdwt {
namespace dn {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even (finite inputs)
  const uint32_t u = __float_as_uint(f);
  return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// 4 consecutive bf16 (8 bytes, 8-byte aligned) or zeros
__device__ __forceinline__ s4 load4(const uint16_t* p, bool ok) {
  if (!ok) return s4{0, 0, 0, 0};
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  return s4{short(v.x & 0xffffu), short(v.x >> 16), short(v.y & 0xffffu), short(v.y >> 16)};
}

// y = conv3x3(x) (+ bias, ReLU): x [H*W][xs] (channels xo .. xo+cin), weights
// w [9][cout_pad][cin] (tap-major, K contiguous), output channel n of pixel p
// to y16[p * ys + yo + n] (bf16) and/or y32[p * ys + yo + n] (fp32, added to
// the existing value when accum). cin is a multiple of 8, cout_pad of 32;
// channels n >= cout are not written.
// A workgroup (4 waves) computes R image rows x TP consecutive pixels x TN
// output channels, TP x TN = 64 x 64 (NWN = 2 waves along the channels) or
// 128 x 32 (NWN = 1). Per chunk of 32 input channels it stages the input's
// R + 2 halo rows (TP + 2 pixels) and the 9 taps' weights of its TN channels
// in LDS once; each wave then runs, per tap and K step, one weight fragment
// against its R rows (R accumulators of 32 x 32). LDS rows are padded to 36
// bf16 (72 bytes): 8-byte reads stay aligned and neighbouring lanes fall in
// different banks.
constexpr int kChunk = 32;        // input channels per staged chunk
constexpr int kLdsRow = kChunk + 4;
constexpr int kConvRows = 2;      // image rows per workgroup

template <int NWN, int R>
__global__ __launch_bounds__(256) void conv3x3_kernel(const uint16_t* __restrict__ x, int H, int W,
                                                      int xs, int xo, int cin,
                                                      const uint16_t* __restrict__ w,
                                                      const float* __restrict__ bias, int cout,
                                                      int cout_pad, uint16_t* y16, float* y32,
                                                      int ys, int yo, int relu, int accum) {
  constexpr int TN = 32 * NWN, TP = 32 * (4 / NWN), XP = TP + 2;
  __shared__ __align__(16) uint16_t sA[(R + 2) * XP * kLdsRow];
  __shared__ __align__(16) uint16_t sB[9 * TN * kLdsRow];
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int tiles_x = (W + TP - 1) / TP;
  const int y0 = (blockIdx.x / tiles_x) * R, x0 = (blockIdx.x % tiles_x) * TP;
  const int n0 = blockIdx.y * TN;
  const int wp = (wv % (4 / NWN)) * 32, wn = (wv / (4 / NWN)) * 32;  // wave's pixel / channel offset
  const int r32 = l & 31, kq = (l >> 5) * 4;
  f16v acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = f16v{};
  for (int kc = 0; kc < cin; kc += kChunk) {
    const int ch = min(kChunk, cin - kc);  // multiple of 8
    const int ng = ch / 4;                 // 8-byte groups per pixel / weight row
    for (int i = t; i < (R + 2) * XP * ng; i += 256) {  // input halo rows y0-1 .. y0+R
      const int g = i % ng, q = i / ng, row = q / XP, px = q - row * XP;
      const int sy = y0 + row - 1, sx = x0 + px - 1;
      const bool in = sy >= 0 && sy < H && sx >= 0 && sx < W;
      const s4 v = load4(x + (size_t(in ? sy * W + sx : 0) * xs + xo + kc + 4 * g), in);
      *reinterpret_cast<s4*>(&sA[(row * XP + px) * kLdsRow + 4 * g]) = v;
    }
    for (int i = t; i < 9 * TN * ng; i += 256) {  // weights: 9 taps x TN channels
      const int g = i % ng, q = i / ng, tap = q / TN, n = q - tap * TN;
      const s4 v = load4(w + (size_t(tap) * cout_pad + n0 + n) * cin + kc + 4 * g, true);
      *reinterpret_cast<s4*>(&sB[(tap * TN + n) * kLdsRow + 4 * g]) = v;
    }
    __syncthreads();
#pragma unroll 1
    for (int tap = 0; tap < 9; ++tap) {
      const int dyi = tap / 3, dx = tap % 3;
      const uint16_t* ar = &sA[(dyi * XP + wp + r32 + dx) * kLdsRow + kq];
      const uint16_t* br = &sB[(tap * TN + wn + r32) * kLdsRow + kq];
      for (int k = 0; k < ch; k += 8) {
        const s4 bv = *reinterpret_cast<const s4*>(br + k);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const s4 av = *reinterpret_cast<const s4*>(ar + r * XP * kLdsRow + k);
          acc[r] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc[r], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  const int n = n0 + wn + r32;
  if (n >= cout) return;
  const float bn = bias ? bias[n] : 0.f;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int y = y0 + r;
    if (y >= H) break;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int xx = x0 + wp + (v >> 2) * 8 + kq + (v & 3);
      if (xx >= W) continue;
      float o32 = acc[r][v] + bn;
      if (relu) o32 = fmaxf(o32, 0.f);
      const size_t o = (size_t(y) * W + xx) * ys + yo + n;
      if (y16) y16[o] = f2bf(o32);
      if (y32) y32[o] = accum ? y32[o] + o32 : o32;
    }
  }
}

// Weight gradient, GEMM M = output channels, N = input channels, K =
// pixels:
//   part[band][tap][co][ci] = sum over the band's pixels p of g[p][co] * x[p + tap][ci]
// g: bf16 [H*W][gs] (the masked output gradient), x: bf16 [H*W][xs] channels
// xo .. xo+cin. One wave per (ci tile, co tile, band of image rows, tap row
// dy): the 3 taps of one row share the staged input. Each 32-pixel row
// segment is staged through LDS once: the gradient tile transposed ([co][px],
// so a lane's 4 consecutive pixels are one 8-byte read) and the input row
// y + dy with a one-pixel halo ([ci][px + 1]); the 3 taps read shifted
// windows of it. Bands write partial sums, which wgrad_reduce_kernel adds
// into dw in a fixed order (no atomics).
constexpr int kSeg = 32;         // pixels per row segment: 4 MFMA K steps
constexpr int kRow = kSeg + 4;   // LDS row (u16): kSeg + 2 pixels, 8-byte multiple

__global__ __launch_bounds__(64) void wgrad_band_kernel(const uint16_t* __restrict__ g, int gs,
                                                        const uint16_t* __restrict__ x, int xs,
                                                        int xo, int H, int W, int cin, int cout,
                                                        int rows, float* __restrict__ part) {
  __shared__ __align__(16) uint16_t gT[32 * kRow];
  __shared__ __align__(16) uint16_t xT[32 * kRow];
  const int l = threadIdx.x;
  const int ci0 = blockIdx.x * 32, co0 = blockIdx.y * 32;
  const int band = blockIdx.z / 3, dyi = blockIdx.z % 3;
  const int y0 = band * rows, y1 = min(H, y0 + rows);
  const int r32 = l & 31, kq = (l >> 5) * 4;
  f16v acc[3] = {f16v{}, f16v{}, f16v{}};
  for (int y = y0; y < y1; ++y) {
    const int sy = y + dyi - 1;
    for (int x0 = 0; x0 < W; x0 += kSeg) {
      for (int i = l; i < kSeg * 8; i += 64) {  // gradient: 32 px x 8 groups of 4 channels
        const int px = i >> 3, grp = i & 7, c = co0 + grp * 4, xx = x0 + px;
        const s4 v = load4(g + (size_t(y) * W + xx) * gs + c, xx < W && c < cout);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          gT[(grp * 4 + j) * kRow + px] = c + j < cout ? uint16_t(v[j]) : uint16_t(0);
      }
      for (int i = l; i < (kSeg + 2) * 8; i += 64) {  // input row sy: 34 px x 8 groups
        const int grp = i & 7, px = i >> 3;
        const int sx = x0 + px - 1, c = ci0 + grp * 4;
        const bool in = sy >= 0 && sy < H && sx >= 0 && sx < W && c < cin;
        const s4 v = load4(x + (size_t(in ? sy * W + sx : 0) * xs + xo + (in ? c : 0)), in);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          xT[(grp * 4 + j) * kRow + px] = c + j < cin ? uint16_t(v[j]) : uint16_t(0);
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < kSeg / 8; ++ks) {
        const int kb = ks * 8 + kq;  // this lane's 4 pixels of the K step
        const s4 a = *reinterpret_cast<const s4*>(&gT[r32 * kRow + kb]);
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          const uint16_t* xr = &xT[r32 * kRow + kb + dx];
          const s4 b = s4{short(xr[0]), short(xr[1]), short(xr[2]), short(xr[3])};
          acc[dx] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[dx], 0, 0, 0);
        }
      }
      __syncthreads();
    }
  }
  const int ci = ci0 + r32;
  if (ci >= cin) return;
#pragma unroll
  for (int dx = 0; dx < 3; ++dx) {
    const int tap = dyi * 3 + dx;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int co = co0 + (v >> 2) * 8 + kq + (v & 3);
      if (co < cout) part[((size_t(band) * 9 + tap) * cout + co) * cin + ci] = acc[dx][v];
    }
  }
}

// dw[i] += sum over bands of part[band][i]. Block: 16 consecutive outputs x
// 16 band groups (group g sums bands g, g + 16, ...); the groups are then
// added in order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float* __restrict__ part,
                                                           int bands, int n,
                                                           float* __restrict__ dw) {
  __shared__ float sums[16][17];
  const int t = threadIdx.x, o = t & 15, grp = t >> 4;
  const int i = blockIdx.x * 16 + o;
  float sum = 0.f;
  if (i < n)
    for (int b = grp; b < bands; b += 16) sum += part[size_t(b) * n + i];
  sums[grp][o] = sum;
  __syncthreads();
  if (t < 16 && i < n) {
    float acc = 0.f;
    for (int g2 = 0; g2 < 16; ++g2) acc += sums[g2][t];
    dw[i] += acc;
  }
}

// gm = bf16(gy * (y > 0)) for ReLU layers (gy * 1 otherwise).
// gy: fp32 [n][gs]; y: bf16 [n][ys] at offset yo; gm: bf16 [n][cout_pad]
// (channels >= cout zeroed). One thread per (pixel, channel).
__global__ void relu_grad_kernel(const float* __restrict__ gy, int gs,
                                 const uint16_t* __restrict__ y, int ys, int yo, int npix,
                                 int cout, int cout_pad, int relu, uint16_t* gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix * cout_pad) return;
  const int p = i / cout_pad, c = i - p * cout_pad;
  float v = 0.f;
  if (c < cout) {
    v = gy[size_t(p) * gs + c];
    if (relu && !(bf2f(y[size_t(p) * ys + yo + c]) > 0.f)) v = 0.f;
  }
  gm[size_t(p) * cout_pad + c] = f2bf(v);
}

// part[chunk][c] = sum over a chunk of pixels of gm[p][c]: block (c, chunk),
// 256 threads, a fixed-order reduction inside the block
__global__ __launch_bounds__(256) void bias_grad_kernel(const uint16_t* __restrict__ gm,
                                                        int cout_pad, int npix, int chunk,
                                                        float* __restrict__ part) {
  __shared__ float sums[4];
  const int c = blockIdx.x;
  const int p0 = blockIdx.y * chunk, p1 = min(npix, p0 + chunk);
  float sum = 0.f;
  for (int p = p0 + int(threadIdx.x); p < p1; p += blockDim.x) sum += bf2f(gm[size_t(p) * cout_pad + c]);
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_down(sum, o, 64);
  if ((threadIdx.x & 63) == 0) sums[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) part[size_t(blockIdx.y) * gridDim.x + c] = sums[0] + sums[1] + sums[2] + sums[3];
}
// db[c] += sum over chunks of part[chunk][c], chunks in order (no atomics)
__global__ void bias_reduce_kernel(const float* __restrict__ part, int chunks, int cout,
                                   float* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cout) return;
  float sum = 0.f;
  for (int k = 0; k < chunks; ++k) sum += part[size_t(k) * cout + c];
  db[c] += sum;
}

// 2x2 average pool: y[(py, px)] = mean of x over the 2x2 block (bf16, C channels)
__global__ void pool_kernel(const uint16_t* x, int H, int W, int C, uint16_t* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = H / 2, w = W / 2;
  if (i >= h * w * C) return;
  const int c = i % C, q = i / C, qy = q / w, qx = q - qy * w;
  const size_t b = (size_t(2 * qy) * W + 2 * qx) * C + c;
  const float s = bf2f(x[b]) + bf2f(x[b + C]) + bf2f(x[b + size_t(W) * C]) + bf2f(x[b + size_t(W) * C + C]);
  y[i] = f2bf(0.25f * s);
}
// gx[p] += gy[p / 2] / 4 (fp32; gx channel stride C)
__global__ void pool_grad_kernel(const float* gy, int H, int W, int C, float* gx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= H * W * C) return;
  const int c = i % C, p = i / C, py = p / W, px = p - py * W;
  gx[i] += 0.25f * gy[(size_t(py / 2) * (W / 2) + px / 2) * C + c];
}
// nearest upsample into a channel slice: y[(p)][yo + c] = x[p / 2][c]
__global__ void up_kernel(const uint16_t* x, int H, int W, int C, uint16_t* y, int ys, int yo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= H * W * C) return;
  const int c = i % C, p = i / C, py = p / W, px = p - py * W;
  y[size_t(p) * ys + yo + c] = x[(size_t(py / 2) * (W / 2) + px / 2) * C + c];
}
// gx[q][c] = sum of gy over the 2x2 block of q (gy: slice yo of stride ys)
__global__ void up_grad_kernel(const float* gy, int H, int W, int ys, int yo, int C, float* gx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = H / 2, w = W / 2;
  if (i >= h * w * C) return;
  const int c = i % C, q = i / C, qy = q / w, qx = q - qy * w;
  const size_t b = (size_t(2 * qy) * W + 2 * qx) * ys + yo + c;
  gx[i] = gy[b] + gy[b + ys] + gy[b + size_t(W) * ys] + gy[b + size_t(W) * ys + ys];
}
// copy a channel slice (bf16): y[p][yo + c] = x[p][xo + c]
__global__ void copy_slice_kernel(const uint16_t* x, int xs, int xo, int n, int C, uint16_t* y,
                                  int ys, int yo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * C) return;
  const int c = i % C, p = i / C;
  y[size_t(p) * ys + yo + c] = x[size_t(p) * xs + xo + c];
}
// gx[p][c] += gy[p][yo + c] (fp32 slices)
__global__ void add_slice_kernel(const float* gy, int ys, int yo, int n, int C, float* gx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * C) return;
  const int c = i % C, p = i / C;
  gx[i] += gy[size_t(p) * ys + yo + c];
}

// network input: in (fp32 [n][inputs]) -> bf16 [n][in_pad] (channels from
// inputs up zero)
__global__ void pack_input_kernel(const float* in, int n, int inputs, int in_pad, uint16_t* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * in_pad) return;
  const int p = i / in_pad, c = i - p * in_pad;
  x[i] = c < inputs ? f2bf(in[size_t(p) * inputs + c]) : uint16_t(0);
}
// Kernel-predicting output (Bako et al. 2017): per pixel, the softmax of
// the kTaps logits (fp32 [n][kLogitStride]) weighs the input colour of the
// kKernel x kKernel neighbourhood (coordinates clamped to the image).
__device__ __forceinline__ int kpn_neighbour(int py, int px, int j, int H, int W) {
  const int sy = min(max(py + j / kKernel - kKernelRadius, 0), H - 1);
  const int sx = min(max(px + j % kKernel - kKernelRadius, 0), W - 1);
  return sy * W + sx;
}
// softmax weights of pixel p into w (unnormalised) and their sum
__device__ __forceinline__ float kpn_weights(const float* logits, int p, float* w) {
  const float* z = logits + size_t(p) * kLogitStride;
  float m = z[0];
#pragma unroll
  for (int j = 1; j < kTaps; ++j) m = fmaxf(m, z[j]);
  float sum = 0.f;
#pragma unroll
  for (int j = 0; j < kTaps; ++j) {
    w[j] = __expf(z[j] - m);
    sum += w[j];
  }
  return sum;
}
// pred [n][3] = sum_j softmax(z)_j * rgb(neighbour j)
__global__ void kpn_apply_kernel(const float* __restrict__ in, int inputs,
                                 const float* __restrict__ logits, int H, int W,
                                 float* __restrict__ pred) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= H * W) return;
  const int py = p / W, px = p - py * W;
  float w[kTaps];
  const float inv = 1.f / kpn_weights(logits, p, w);
  float r = 0.f, g = 0.f, b = 0.f;
#pragma unroll
  for (int j = 0; j < kTaps; ++j) {
    const float* c = in + size_t(kpn_neighbour(py, px, j, H, W)) * inputs;
    r += w[j] * c[0];
    g += w[j] * c[1];
    b += w[j] * c[2];
  }
  pred[size_t(p) * 3 + 0] = r * inv;
  pred[size_t(p) * 3 + 1] = g * inv;
  pred[size_t(p) * 3 + 2] = b * inv;
}
// Squared error of pred against target: err += sum (pred - target)^2, and
// the gradient of inv_count * that sum with respect to the logits,
// dz_j = w_j * (c_j - pred) . gpred with gpred = 2 inv_count (pred - target),
// into gz [n][kLogitStride] (channels < kTaps).
__global__ void kpn_loss_grad_kernel(const float* __restrict__ in, int inputs,
                                     const float* __restrict__ logits,
                                     const float* __restrict__ pred, const float* __restrict__ target,
                                     int H, int W, float inv_count, float* __restrict__ gz,
                                     float* err) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (p < H * W) {
    const int py = p / W, px = p - py * W;
    float w[kTaps];
    const float inv = 1.f / kpn_weights(logits, p, w);
    const float pr = pred[size_t(p) * 3], pg = pred[size_t(p) * 3 + 1], pb = pred[size_t(p) * 3 + 2];
    const float dr = pr - target[size_t(p) * 3], dg = pg - target[size_t(p) * 3 + 1],
                db = pb - target[size_t(p) * 3 + 2];
    e = dr * dr + dg * dg + db * db;
    const float s = 2.f * inv_count;
    const float gr = s * dr, gg = s * dg, gb = s * db;
    const float base = pr * gr + pg * gg + pb * gb;
#pragma unroll
    for (int j = 0; j < kTaps; ++j) {
      const float* c = in + size_t(kpn_neighbour(py, px, j, H, W)) * inputs;
      gz[size_t(p) * kLogitStride + j] = w[j] * inv * (c[0] * gr + c[1] * gg + c[2] * gb - base);
    }
  }
  for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(err, e);
}

// Adam step on fp32 masters (Kingma & Ba 2015), then the bf16 forward copy.
__global__ void adam_kernel(float* w, const float* g, float* m, float* v, int n, float lr,
                            float b1, float b2, float c1, float c2, float eps, uint16_t* w16) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = f2bf(wi);
}

// forward weights [9][cout_pad][cin_pad] (bf16) and dgrad weights
// [9][cin_pad32][cout_pad8] = w[8 - tap][co][ci] from the fp32 master
// [9][cout][cin]; padding zero.
__global__ void weight_layouts_kernel(const float* w, int cin, int cout, int cin_pad, int cout_pad,
                                      uint16_t* wf, int cin_pad32, int cout_pad8, uint16_t* wd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nf = 9 * cout_pad * cin_pad, nd = 9 * cin_pad32 * cout_pad8;
  if (i < nf) {
    const int ci = i % cin_pad, r = i / cin_pad, co = r % cout_pad, tap = r / cout_pad;
    wf[i] = (ci < cin && co < cout) ? f2bf(w[(size_t(tap) * cout + co) * cin + ci]) : uint16_t(0);
  }
  if (i < nd) {
    const int co = i % cout_pad8, r = i / cout_pad8, ci = r % cin_pad32, tap = r / cin_pad32;
    wd[i] = (ci < cin && co < cout) ? f2bf(w[(size_t(8 - tap) * cout + co) * cin + ci])
                                    : uint16_t(0);
  }
}

inline unsigned nb(size_t n, unsigned t = 256) { return unsigned((n + t - 1) / t); }

}  // namespace

hipError_t conv3x3(const uint16_t* x, int H, int W, int xs, int xo, int cin, const uint16_t* w,
                   const float* bias, int cout, int cout_pad, uint16_t* y16, float* y32, int ys,
                   int yo, bool relu, bool accum, hipStream_t s) {
  // 8-byte input and weight groups
  if (cin % 8 || cout_pad % 32 || xs % 4 || xo % 4 || xo + cin > xs) return hipErrorInvalidValue;
  constexpr int R = kConvRows;
  const unsigned rows = unsigned((H + R - 1) / R);
  if (cout_pad % 64 == 0) {
    const dim3 grid(unsigned((W + 63) / 64) * rows, unsigned(cout_pad / 64));
    hipLaunchKernelGGL((conv3x3_kernel<2, R>), grid, dim3(256), 0, s, x, H, W, xs, xo, cin, w,
                       bias, cout, cout_pad, y16, y32, ys, yo, relu ? 1 : 0, accum ? 1 : 0);
  } else {
    const dim3 grid(unsigned((W + 127) / 128) * rows, unsigned(cout_pad / 32));
    hipLaunchKernelGGL((conv3x3_kernel<1, R>), grid, dim3(256), 0, s, x, H, W, xs, xo, cin, w,
                       bias, cout, cout_pad, y16, y32, ys, yo, relu ? 1 : 0, accum ? 1 : 0);
  }
  return hipGetLastError();
}

hipError_t wgrad(const uint16_t* g, int gs, const uint16_t* x, int xs, int xo, int H, int W,
                 int cin, int cout, float* dw, float* part, size_t part_cap, hipStream_t s) {
  // 8-byte channel groups: strides and offsets in whole groups, and the
  // groups that hold the last channels inside the row
  if (gs % 4 || xs % 4 || xo % 4 || (cout + 3) / 4 * 4 > gs || xo + (cin + 3) / 4 * 4 > xs)
    return hipErrorInvalidValue;
  const int tiles = ((cin + 31) / 32) * ((cout + 31) / 32);
  const size_t per_band = size_t(9) * cout * cin;
  if (per_band > part_cap) return hipErrorInvalidValue;
  // about 2048 waves (3 tap rows per band), within the partial-sum buffer
  int bands = std::max(1, std::min(H, std::min((683 + tiles - 1) / tiles, int(part_cap / per_band))));
  const int rows = (H + bands - 1) / bands;
  bands = (H + rows - 1) / rows;
  const dim3 grid(unsigned((cin + 31) / 32), unsigned((cout + 31) / 32), unsigned(bands * 3));
  hipLaunchKernelGGL(wgrad_band_kernel, grid, dim3(64), 0, s, g, gs, x, xs, xo, H, W, cin, cout,
                     rows, part);
  hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(nb(per_band, 16)), dim3(256), 0, s, part, bands,
                     int(per_band), dw);
  return hipGetLastError();
}

hipError_t relu_grad(const float* gy, int gs, const uint16_t* y, int ys, int yo, int npix, int cout,
                     int cout_pad, bool relu, uint16_t* gm, float* db, float* part,
                     size_t part_cap, hipStream_t s) {
  const int chunk = 4096;
  const int chunks = (npix + chunk - 1) / chunk;
  if (size_t(chunks) * cout > part_cap) return hipErrorInvalidValue;
  hipLaunchKernelGGL(relu_grad_kernel, dim3(nb(size_t(npix) * cout_pad)), dim3(256), 0, s, gy, gs,
                     y, ys, yo, npix, cout, cout_pad, relu ? 1 : 0, gm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(bias_grad_kernel, dim3(unsigned(cout), unsigned(chunks)), dim3(256), 0, s, gm,
                     cout_pad, npix, chunk, part);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(bias_reduce_kernel, dim3(nb(size_t(cout))), dim3(256), 0, s, part, chunks,
                     cout, db);
  return hipGetLastError();
}

hipError_t pool(const uint16_t* x, int H, int W, int C, uint16_t* y, hipStream_t s) {
  hipLaunchKernelGGL(pool_kernel, dim3(nb(size_t(H / 2) * (W / 2) * C)), dim3(256), 0, s, x, H, W,
                     C, y);
  return hipGetLastError();
}
hipError_t pool_grad(const float* gy, int H, int W, int C, float* gx, hipStream_t s) {
  hipLaunchKernelGGL(pool_grad_kernel, dim3(nb(size_t(H) * W * C)), dim3(256), 0, s, gy, H, W, C,
                     gx);
  return hipGetLastError();
}
hipError_t up(const uint16_t* x, int H, int W, int C, uint16_t* y, int ys, int yo, hipStream_t s) {
  hipLaunchKernelGGL(up_kernel, dim3(nb(size_t(H) * W * C)), dim3(256), 0, s, x, H, W, C, y, ys,
                     yo);
  return hipGetLastError();
}
hipError_t up_grad(const float* gy, int H, int W, int ys, int yo, int C, float* gx, hipStream_t s) {
  hipLaunchKernelGGL(up_grad_kernel, dim3(nb(size_t(H / 2) * (W / 2) * C)), dim3(256), 0, s, gy, H,
                     W, ys, yo, C, gx);
  return hipGetLastError();
}
hipError_t copy_slice(const uint16_t* x, int xs, int xo, int n, int C, uint16_t* y, int ys, int yo,
                      hipStream_t s) {
  hipLaunchKernelGGL(copy_slice_kernel, dim3(nb(size_t(n) * C)), dim3(256), 0, s, x, xs, xo, n, C,
                     y, ys, yo);
  return hipGetLastError();
}
hipError_t add_slice(const float* gy, int ys, int yo, int n, int C, float* gx, hipStream_t s) {
  hipLaunchKernelGGL(add_slice_kernel, dim3(nb(size_t(n) * C)), dim3(256), 0, s, gy, ys, yo, n, C,
                     gx);
  return hipGetLastError();
}
hipError_t pack_input(const float* in, int n, int inputs, int in_pad, uint16_t* x, hipStream_t s) {
  hipLaunchKernelGGL(pack_input_kernel, dim3(nb(size_t(n) * in_pad)), dim3(256), 0, s, in, n,
                     inputs, in_pad, x);
  return hipGetLastError();
}
hipError_t kpn_apply(const float* in, int inputs, const float* logits, int H, int W, float* pred,
                     hipStream_t s) {
  hipLaunchKernelGGL(kpn_apply_kernel, dim3(nb(size_t(H) * W)), dim3(256), 0, s, in, inputs,
                     logits, H, W, pred);
  return hipGetLastError();
}
hipError_t kpn_loss_grad(const float* in, int inputs, const float* logits, const float* pred,
                         const float* target, int H, int W, float inv_count, float* gz, float* err,
                         hipStream_t s) {
  hipLaunchKernelGGL(kpn_loss_grad_kernel, dim3(nb(size_t(H) * W)), dim3(256), 0, s, in, inputs,
                     logits, pred, target, H, W, inv_count, gz, err);
  return hipGetLastError();
}
hipError_t adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                float c1, float c2, float eps, uint16_t* w16, hipStream_t s) {
  hipLaunchKernelGGL(adam_kernel, dim3(nb(size_t(n))), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2,
                     c1, c2, eps, w16);
  return hipGetLastError();
}
hipError_t weight_layouts(const float* w, int cin, int cout, int cin_pad, int cout_pad,
                          uint16_t* wf, int cin_pad32, int cout_pad8, uint16_t* wd, hipStream_t s) {
  const size_t n = std::max(size_t(9) * cout_pad * cin_pad, size_t(9) * cin_pad32 * cout_pad8);
  hipLaunchKernelGGL(weight_layouts_kernel, dim3(nb(n)), dim3(256), 0, s, w, cin, cout, cin_pad,
                     cout_pad, wf, cin_pad32, cout_pad8, wd);
  return hipGetLastError();
}

}  // namespace dn
}  // namespace mdwt
