// Kernels of the learned denoiser (a 3-level U-Net of 3x3 convolutions): BF16
// activations and weights, FP32 accumulation on the matrix cores
// (v_mfma_f32_32x32x16_bf16 for the convolutions, v_mfma_f32_32x32x8bf16_1k
// for the weight gradients), FP32 gradients and Adam state. No reference
// counterpart (the reference renders plain accumulation only).
//
// Tensors are NHWC with a channel stride (so that a convolution can read or
// write a slice of a concatenated tensor). A 3x3 convolution is an implicit
// GEMM: M = pixels, N = output channels, K = 9 taps x input channels. One
// wave computes 32-pixel x 32-channel tiles; the MFMA lane layouts are
//   A (32 x 8):  lane l holds A[l % 32][(l / 32) * 4 + j], j = 0..3
//   B (8 x 32):  lane l holds B[(l / 32) * 4 + j][l % 32]
//   D (32 x 32): lane l, register v holds D[(v / 4) * 8 + (l / 32) * 4 + v % 4][l % 32]
// (checked on the device with known matrices), and for the 16-deep K step
// lane l holds K entries (l / 32) * 8 + j, j = 0..7, of its A row and B
// column (the sum over K does not depend on which entries a lane holds, as
// long as A and B agree), with the same D layout.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "denoise_internal.hpp"

namespace mdwt {
namespace dn {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
// fp32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32: two per instruction)
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
  const bf2v v = {__bf16(lo), __bf16(hi)};
  return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ uint16_t f2bf(float f) { return uint16_t(f2bf2(f, 0.f)); }

// 4 consecutive bf16 (8 bytes, 8-byte aligned) or zeros
__device__ __forceinline__ s4 load4(const uint16_t* p, bool ok) {
  if (!ok) return s4{0, 0, 0, 0};
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  return s4{short(v.x & 0xffffu), short(v.x >> 16), short(v.y & 0xffffu), short(v.y >> 16)};
}

// Kernel-predicting output (Bako et al. 2017): per pixel, the softmax of
// the kTaps logits (fp32 [n][kLogitStride]) weighs the input colour of the
// kKernel x kKernel neighbourhood (coordinates clamped to the image).
__device__ __forceinline__ int kpn_neighbour(int py, int px, int j, int H, int W) {
  const int sy = min(max(py + j / kKernel - kKernelRadius, 0), H - 1);
  const int sx = min(max(px + j % kKernel - kKernelRadius, 0), W - 1);
  return sy * W + sx;
}
// softmax weights of one pixel's logits z into w (unnormalised) and their sum
__device__ __forceinline__ float kpn_weights_z(const float* z, float* w) {
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
__device__ __forceinline__ float kpn_weights(const float* logits, int p, float* w) {
  return kpn_weights_z(logits + size_t(p) * kLogitStride, w);
}
// out[0..2] = sum_j softmax(z)_j * rgb(neighbour j) for pixel (py, px)
__device__ __forceinline__ void kpn_filter(const float* __restrict__ in, int inputs, const float* z,
                                           int py, int px, int H, int W, float* out) {
  float w[kTaps];
  const float inv = 1.f / kpn_weights_z(z, w);
  float r = 0.f, g = 0.f, b = 0.f;
#pragma unroll
  for (int j = 0; j < kTaps; ++j) {
    const float* c = in + size_t(kpn_neighbour(py, px, j, H, W)) * inputs;
    r += w[j] * c[0];
    g += w[j] * c[1];
    b += w[j] * c[2];
  }
  out[0] = r * inv;
  out[1] = g * inv;
  out[2] = b * inv;
}

// y = conv3x3(x) (+ bias, ReLU) on v_mfma_f32_32x32x16_bf16: weights w
// [9][cout_pad][cin] (tap-major, K contiguous); output channel n of pixel p
// to y16[p * ys + yo + n] (bf16) and/or y32[p * ys + yo + n] (fp32, added to
// the existing value when accum); channels n >= cout are not written.
//
// A workgroup (4 waves, WX along x by WR along rows) computes R = 2 WR image
// rows x TP = 32 WX pixels x TN = 32 NT output channels; each wave 2 rows x
// 32 pixels x TN channels (2 NT accumulators of 32 x 32). Per chunk of 32
// input channels the workgroup stages the R + 2 halo rows (TP + 2 pixels) and
// the 9 taps' weights of its TN channels in LDS, one 16-byte granule (8
// channels) per load; the next chunk's granules are loaded into registers
// while the current one is multiplied (one barrier pair per chunk). LDS rows
// hold 32 channels padded to 40 bf16 (80 bytes): 16-byte reads stay aligned
// and any 8 consecutive rows cover the 32 banks once. With 16-channel K
// steps, lane l supplies channels (l / 32) * 8 .. + 7 of its pixel (A) and of
// its output channel's weights (B); D is the 32 x 32 layout above.
//
// Input sources: kSrcPlain reads x (stride xs, offset xo). kSrcUpCat reads
// channel c < cu from xl, the tensor at half resolution (stride cu), at the
// pixel (y / 2, x / 2) (nearest upsampling), and channel c >= cu from x at
// c - cu: the U-Net's upsample-and-concatenate without materialising it.
// Epilogues: kEpiPlain stores as above. kEpiPool also writes the 2 x 2 mean of
// the stored bf16 output to pool16 (stride ps), as pool_kernel computes it.
// kEpiKpn stages the fp32 logits of the tile in LDS and applies the
// kernel-predicting filter (as kpn_apply_kernel) to write pred directly.
constexpr int kChunk = 32;   // input channels per staged chunk
constexpr int kPitch = 40;   // LDS row (bf16): 32 channels + 8 padding
constexpr int kZPitch = kLogitStride + 1;  // LDS row (fp32) of one pixel's logits
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf8 lds_frag(const uint16_t* p) {
  return __builtin_bit_cast(bf8, *reinterpret_cast<const uint4*>(p));
}

template <int WX, int WR, int NT>
struct ConvShape {
  static constexpr int RW = 2, R = RW * WR, TP = 32 * WX, XP = TP + 2, TN = 32 * NT;
  static constexpr int kARows = (R + 2) * XP;             // staged pixels
  static constexpr int kAGr = kARows * 4, kBGr = 9 * TN * 4;  // 16-byte granules
  static constexpr int kAPer = (kAGr + 255) / 256, kBPer = (kBGr + 255) / 256;
  static constexpr int kLds = (kARows + 9 * TN) * kPitch;  // bf16
};

// kpn_filter with the colours read from a staged tile: rgb points at the
// pixel's top-left neighbour in a [rows][pitch][3] block of clamped input
// colours (the same values kpn_neighbour would fetch), same arithmetic
__device__ __forceinline__ void kpn_filter_tile(const float* z, const float* rgb, int pitch,
                                                float* out) {
  float w[kTaps];
  const float inv = 1.f / kpn_weights_z(z, w);
  float r = 0.f, g = 0.f, b = 0.f;
#pragma unroll
  for (int j = 0; j < kTaps; ++j) {
    const float* c = rgb + ((j / kKernel) * pitch + j % kKernel) * 3;
    r += w[j] * c[0];
    g += w[j] * c[1];
    b += w[j] * c[2];
  }
  out[0] = r * inv;
  out[1] = g * inv;
  out[2] = b * inv;
}

// The MFMAs of one staged chunk for waves of one 32-channel tile: 9 taps x
// STEPS K16 steps, fully unrolled, with the next step's LDS fragments read
// before the current step's MFMAs are issued (two fragment sets in
// registers). Waves of two or three tiles have no registers to spare for
// the second set and run the tap loop in the kernel.
template <int WX, int WR, int NT, int STEPS>
__device__ __forceinline__ void mma_chunk(const uint16_t* sA, const uint16_t* sB, int wr, int wx,
                                          int r32, int kh, f16v (&acc)[2][NT]) {
  using S = ConvShape<WX, WR, NT>;
  constexpr int RW = S::RW, XP = S::XP, TN = S::TN, NS = 9 * STEPS;
  const uint16_t* ab = &sA[((wr * RW) * XP + wx * 32 + r32) * kPitch + kh];
  const uint16_t* bb = &sB[r32 * kPitch + kh];
  bf8 fa[2][RW], fb[2][NT];
  auto load = [&](int i, int buf) {
    const int tap = i / STEPS, st = i % STEPS, dyi = tap / 3, dx = tap % 3;
#pragma unroll
    for (int j = 0; j < NT; ++j) fb[buf][j] = lds_frag(bb + (tap * TN + j * 32) * kPitch + 16 * st);
#pragma unroll
    for (int r = 0; r < RW; ++r) fa[buf][r] = lds_frag(ab + ((r + dyi) * XP + dx) * kPitch + 16 * st);
  };
  load(0, 0);
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    if (i + 1 < NS) load(i + 1, (i + 1) & 1);
#pragma unroll
    for (int r = 0; r < RW; ++r)
      acc[r][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fb[i & 1][0], fa[i & 1][r], acc[r][0], 0, 0, 0);
    // keep the schedule to one step ahead (hoisting every read would need
    // all fragments in registers at once)
    __builtin_amdgcn_sched_barrier(0);
  }
}

// bias of channels n .. n + 3 (0 past cout or without bias); n % 4 == 0
__device__ __forceinline__ float4 bias4(const float* bias, int n, int cout) {
  if (!bias || n >= cout) return make_float4(0.f, 0.f, 0.f, 0.f);
  if (n + 3 < cout) return *reinterpret_cast<const float4*>(bias + n);
  return make_float4(bias[n], n + 1 < cout ? bias[n + 1] : 0.f, n + 2 < cout ? bias[n + 2] : 0.f, 0.f);
}

template <int WX, int WR, int NT, int SRC, int EPI>
__global__ __launch_bounds__(256, 2) void conv_kernel(ConvArgs a) {
  using S = ConvShape<WX, WR, NT>;
  constexpr int RW = S::RW, R = S::R, TP = S::TP, XP = S::XP, TN = S::TN;
  static_assert(EPI != kEpiKpn || (R * TP * kZPitch + (R + 2 * kKernelRadius) *
                                                      (TP + 2 * kKernelRadius) * 3) * 4 <= S::kLds * 2,
                "filter tile");
  __shared__ __align__(16) uint16_t lds[S::kLds];
  uint16_t* sA = lds;
  uint16_t* sB = lds + S::kARows * kPitch;
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int tiles_x = (a.W + TP - 1) / TP, ntiles = tiles_x * ((a.H + R - 1) / R);
  const int n0 = blockIdx.y * TN;
  const int wx = wv % WX, wr = wv / WX;
  const int r32 = l & 31, kh = (l >> 5) * 8, kq = (l >> 5) * 4;
  const uint4 zero = {0u, 0u, 0u, 0u};
  // Tiles of this workgroup: with a grid that is a multiple of the 8 XCDs
  // (workgroup i runs on XCD i % 8), XCD x takes the x-th eighth of the
  // tiles (row bands that share halo rows stay in one L2) and its workgroups
  // stride through it; otherwise a plain grid stride.
  int first, stride, last;
  if (gridDim.x % 8 == 0) {
    const int per = int(gridDim.x) / 8, span = (ntiles + 7) / 8, xcd = int(blockIdx.x) % 8;
    first = xcd * span + int(blockIdx.x) / 8;
    stride = per;
    last = min(ntiles, (xcd + 1) * span);
  } else {
    first = blockIdx.x;
    stride = gridDim.x;
    last = ntiles;
  }
  if (first >= last) return;
  // one chunk of input channels: the weights stay in LDS for every tile
  // (not with the filter epilogue, whose logits overwrite them)
  const bool b_once = EPI != kEpiKpn && a.cin <= kChunk;

  uint4 pa[S::kAPer], pb[S::kBPer];
  // kSrcF32: the fp32 network input, 8 values per granule, converted to
  // bf16 when stored to LDS (as pack_input_kernel converts them)
  constexpr int kF32 = SRC == kSrcF32 ? S::kAPer : 1;
  float pf[kF32][8];
  auto fetch = [&](int tile, int kc, bool with_b) {
    const int y0 = (tile / tiles_x) * R, x0 = (tile % tiles_x) * TP;
#pragma unroll
    for (int i = 0; i < S::kAPer; ++i) {
      const int f = t + 256 * i;
      pa[i] = zero;
      if constexpr (SRC == kSrcF32)
#pragma unroll
        for (int e = 0; e < 8; ++e) pf[i][e] = 0.f;
      if (f < S::kAGr) {
        const int g = f & 3, pp = f >> 2, row = pp / XP, px = pp - row * XP;
        const int sy = y0 + row - 1, sx = x0 + px - 1, c = kc + 8 * g;
        if (sy >= 0 && sy < a.H && sx >= 0 && sx < a.W && c < a.cin) {
          if constexpr (SRC == kSrcF32) {
            const float* q = a.kin + (uint32_t(sy) * uint32_t(a.W) + uint32_t(sx)) * uint32_t(a.inputs);
#pragma unroll
            for (int e = 0; e < 8; ++e)
              if (c + e < a.inputs) pf[i][e] = q[c + e];
            continue;
          }
          const uint16_t* p;
          if (SRC == kSrcUpCat && c < a.cu)
            p = a.xl + (uint32_t(sy >> 1) * uint32_t(a.W >> 1) + uint32_t(sx >> 1)) * uint32_t(a.cu) +
                uint32_t(c);
          else
            p = a.x + (uint32_t(sy) * uint32_t(a.W) + uint32_t(sx)) * uint32_t(a.xs) +
                uint32_t(a.xo + (SRC == kSrcUpCat ? c - a.cu : c));
          pa[i] = *reinterpret_cast<const uint4*>(p);
        }
      }
    }
    if (!with_b) return;
#pragma unroll
    for (int i = 0; i < S::kBPer; ++i) {
      const int f = t + 256 * i;
      pb[i] = zero;
      if (f < S::kBGr) {
        const int g = f & 3, q = f >> 2, tap = q / TN, n = q - tap * TN, c = kc + 8 * g;
        if (c < a.cin)
          pb[i] = *reinterpret_cast<const uint4*>(
              a.w + (uint32_t(tap) * uint32_t(a.cout_pad) + uint32_t(n0 + n)) * uint32_t(a.cin) + uint32_t(c));
      }
    }
  };
  auto store = [&](bool with_b) {
#pragma unroll
    for (int i = 0; i < S::kAPer; ++i) {
      const int f = t + 256 * i;
      if constexpr (SRC == kSrcF32) {
        uint32_t w[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
          w[e] = f2bf2(pf[i][2 * e], pf[i][2 * e + 1]);
        pa[i] = make_uint4(w[0], w[1], w[2], w[3]);
      }
      if (f < S::kAGr) *reinterpret_cast<uint4*>(&sA[(f >> 2) * kPitch + 8 * (f & 3)]) = pa[i];
    }
    if (!with_b) return;
#pragma unroll
    for (int i = 0; i < S::kBPer; ++i) {
      const int f = t + 256 * i;
      if (f < S::kBGr) *reinterpret_cast<uint4*>(&sB[(f >> 2) * kPitch + 8 * (f & 3)]) = pb[i];
    }
  };

  f16v acc[RW][NT];
#pragma unroll
  for (int r = 0; r < RW; ++r)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[r][j] = f16v{};
  // stages (tile, chunk) in order; the next stage's granules are loaded while
  // the current one is multiplied and while the finished tile is stored
  int tile = first, kc = 0;
  bool first_stage = true;
  fetch(tile, 0, true);
  while (true) {
    store(first_stage || !b_once);
    first_stage = false;
    __syncthreads();
    int ntile = tile, nkc = kc + kChunk;
    if (nkc >= a.cin) {
      nkc = 0;
      ntile += stride;
    }
    const bool more = ntile < last;
    if (more) fetch(ntile, nkc, !b_once);
    // K16 steps of the chunk (granules past cin are zero): 9 taps x 1 or 2
    if constexpr (NT == 1) {
      if (a.cin - kc > 16)
        mma_chunk<WX, WR, NT, 2>(sA, sB, wr, wx, r32, kh, acc);
      else
        mma_chunk<WX, WR, NT, 1>(sA, sB, wr, wx, r32, kh, acc);
    } else {
      const int steps = a.cin - kc > 16 ? 2 : 1;
#pragma unroll 1
      for (int tap = 0; tap < 9; ++tap) {
        const int dyi = tap / 3, dx = tap - dyi * 3;
        const uint16_t* ar = &sA[((wr * RW + dyi) * XP + wx * 32 + r32 + dx) * kPitch + kh];
        const uint16_t* br = &sB[(tap * TN + r32) * kPitch + kh];
        for (int st = 0; st < steps; ++st) {
          bf8 b[NT];
#pragma unroll
          for (int j = 0; j < NT; ++j) b[j] = lds_frag(br + j * 32 * kPitch + 16 * st);
#pragma unroll
          for (int r = 0; r < RW; ++r) {
            const bf8 av = lds_frag(ar + r * XP * kPitch + 16 * st);
#pragma unroll
            for (int j = 0; j < NT; ++j)
              acc[r][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], av, acc[r][j], 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();
    if (nkc == 0) {  // the tile is complete
      const int y0 = (tile / tiles_x) * R, x0 = (tile % tiles_x) * TP;
      // acc[r][j]: D = weights x pixels, so lane l holds pixel r32 of row
      // y0 + wr * RW + r and, in register v, output channel
      // j * 32 + (v / 4) * 8 + kq + v % 4: four consecutive channels per
      // register group
      const int px = x0 + wx * 32 + r32;
      if constexpr (EPI == kEpiKpn) {
        // logits (fp32, kLogitStride per pixel; n0 == 0, TN == kLogitStride)
        // of the tile into LDS, and after them the input colours of the tile
        // with a kKernelRadius halo (coordinates clamped to the image); then
        // one thread per pixel applies the filter
        float* sz = reinterpret_cast<float*>(lds);
        float* srgb = sz + R * TP * kZPitch;
        constexpr int kRgbCols = TP + 2 * kKernelRadius, kRgbPx = (R + 2 * kKernelRadius) * kRgbCols;
        constexpr int kRgbPer = (kRgbPx + 255) / 256;
        float cr[kRgbPer][3];
#pragma unroll
        for (int i = 0; i < kRgbPer; ++i) {
          const int q = t + 256 * i;
          if (q < kRgbPx) {
            const int iy = q / kRgbCols, ix = q - iy * kRgbCols;
            const int sy = min(max(y0 - kKernelRadius + iy, 0), a.H - 1);
            const int sx = min(max(x0 - kKernelRadius + ix, 0), a.W - 1);
            const float* c = a.kin + (size_t(sy) * a.W + sx) * a.inputs;
            cr[i][0] = c[0];
            cr[i][1] = c[1];
            cr[i][2] = c[2];
          }
        }
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
          for (int gq = 0; gq < 4; ++gq) {
            const int n = j * 32 + gq * 8 + kq;
            const float4 b4 = bias4(a.bias, n, a.cout);
            const float bn[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
            for (int r = 0; r < RW; ++r) {
              const int pl = (wr * RW + r) * TP + wx * 32 + r32;
#pragma unroll
              for (int e = 0; e < 4; ++e) sz[pl * kZPitch + n + e] = acc[r][j][gq * 4 + e] + bn[e];
            }
          }
#pragma unroll
        for (int i = 0; i < kRgbPer; ++i) {
          const int q = t + 256 * i;
          if (q < kRgbPx) {
            srgb[q * 3 + 0] = cr[i][0];
            srgb[q * 3 + 1] = cr[i][1];
            srgb[q * 3 + 2] = cr[i][2];
          }
        }
        __syncthreads();
        for (int pl = t; pl < R * TP; pl += 256) {
          const int ly = pl / TP, lx = pl % TP, py = y0 + ly, qx = x0 + lx;
          if (py >= a.H || qx >= a.W) continue;
          kpn_filter_tile(sz + pl * kZPitch, srgb + (ly * kRgbCols + lx) * 3, kRgbCols,
                          a.pred + (size_t(py) * a.W + qx) * 3);
        }
        __syncthreads();  // before the next stage overwrites the logits
      } else {
        const int y = y0 + wr * RW;  // even: R and RW are
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
          for (int gq = 0; gq < 4; ++gq) {
            const int n = n0 + j * 32 + gq * 8 + kq;  // channels n .. n + 3
            const float4 b4 = bias4(a.bias, n, a.cout);
            const float bn[4] = {b4.x, b4.y, b4.z, b4.w};
            uint32_t o16[RW][2];  // bf16 pairs (n, n + 1), (n + 2, n + 3)
#pragma unroll
            for (int r = 0; r < RW; ++r) {
              float o32[4];
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                o32[e] = acc[r][j][gq * 4 + e] + bn[e];
                if (a.relu) o32[e] = fmaxf(o32[e], 0.f);
              }
              o16[r][0] = f2bf2(o32[0], o32[1]);
              o16[r][1] = f2bf2(o32[2], o32[3]);
              if (y + r >= a.H || px >= a.W || n >= a.cout) continue;
              // element offsets stay below 2^31 (at most 2^24 pixels x 96
              // channels at full resolution, 160 at half)
              const uint32_t o = (uint32_t(y + r) * uint32_t(a.W) + uint32_t(px)) * uint32_t(a.ys) +
                                 uint32_t(a.yo + n);
              if (n + 3 < a.cout) {  // 8-byte bf16 / 16-byte fp32 stores
                if (a.y16) *reinterpret_cast<uint2*>(a.y16 + o) = make_uint2(o16[r][0], o16[r][1]);
                if (a.y32) {
                  float4* q = reinterpret_cast<float4*>(a.y32 + o);
                  float4 v4 = make_float4(o32[0], o32[1], o32[2], o32[3]);
                  if (a.accum) {
                    const float4 old = *q;
                    v4 = make_float4(old.x + v4.x, old.y + v4.y, old.z + v4.z, old.w + v4.w);
                  }
                  *q = v4;
                }
              } else {
                for (int e = 0; n + e < a.cout; ++e) {
                  if (a.y16) a.y16[o + e] = uint16_t(o16[r][e >> 1] >> ((e & 1) * 16));
                  if (a.y32) a.y32[o + e] = a.accum ? a.y32[o + e] + o32[e] : o32[e];
                }
              }
            }
            if constexpr (EPI == kEpiPool) {
              // pixel px + 1 is lane l + 1's; even lanes (px even) write the
              // mean of rows y, y + 1 and pixels px, px + 1
              uint32_t nb16[RW][2];
#pragma unroll
              for (int r = 0; r < RW; ++r)
#pragma unroll
                for (int d = 0; d < 2; ++d)  // DPP quad_perm [1, 0, 3, 2]: lane l ^ 1
                  nb16[r][d] = uint32_t(__builtin_amdgcn_mov_dpp(int(o16[r][d]), 0xB1, 0xF, 0xF, false));
              if ((r32 & 1) == 0 && y < a.H && px < a.W && n < a.cout) {
                float mf[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                  const int d = e >> 1, sh = (e & 1) * 16;
                  const float s = bf2f(uint16_t(o16[0][d] >> sh)) + bf2f(uint16_t(nb16[0][d] >> sh)) +
                                  bf2f(uint16_t(o16[1][d] >> sh)) + bf2f(uint16_t(nb16[1][d] >> sh));
                  mf[e] = 0.25f * s;
                }
                const uint32_t m01 = f2bf2(mf[0], mf[1]), m23 = f2bf2(mf[2], mf[3]);
                const uint32_t o = (uint32_t(y >> 1) * uint32_t(a.W >> 1) + uint32_t(px >> 1)) *
                                       uint32_t(a.ps) + uint32_t(n);
                if (n + 3 < a.cout)
                  *reinterpret_cast<uint2*>(a.pool16 + o) = make_uint2(m01, m23);
                else
                  for (int e = 0; n + e < a.cout; ++e)
                    a.pool16[o + e] = uint16_t((e < 2 ? m01 : m23) >> ((e & 1) * 16));
              }
            }
          }
      }
#pragma unroll
      for (int r = 0; r < RW; ++r)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[r][j] = f16v{};
    }
    if (!more) break;
    tile = ntile;
    kc = nkc;
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
  // stages: (row y, 32-pixel segment) in row order; the next stage's
  // gradient (32 px x 8 groups of 4 channels) and input row y + dy - 1
  // (34 px x 8 groups) are loaded into registers while this one is multiplied
  constexpr int kGPer = kSeg * 8 / 64, kXPer = ((kSeg + 2) * 8 + 63) / 64;
  const int segs = (W + kSeg - 1) / kSeg;
  const int nst = (y1 - y0) * segs;
  s4 pg[kGPer], pxv[kXPer];
  auto fetch = [&](int st) {
    const int y = y0 + st / segs, x0 = (st % segs) * kSeg, sy = y + dyi - 1;
#pragma unroll
    for (int i = 0; i < kGPer; ++i) {
      const int idx = l + 64 * i, px = idx >> 3, grp = idx & 7, c = co0 + grp * 4, xx = x0 + px;
      pg[i] = load4(g + (size_t(y) * W + xx) * gs + c, xx < W && c < cout);
    }
#pragma unroll
    for (int i = 0; i < kXPer; ++i) {
      const int idx = l + 64 * i, grp = idx & 7, px = idx >> 3;
      const int sx = x0 + px - 1, c = ci0 + grp * 4;
      const bool in = idx < (kSeg + 2) * 8 && sy >= 0 && sy < H && sx >= 0 && sx < W && c < cin;
      pxv[i] = load4(x + (size_t(in ? sy * W + sx : 0) * xs + xo + (in ? c : 0)), in);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < kGPer; ++i) {
      const int idx = l + 64 * i, px = idx >> 3, grp = idx & 7, c = co0 + grp * 4;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        gT[(grp * 4 + j) * kRow + px] = c + j < cout ? uint16_t(pg[i][j]) : uint16_t(0);
    }
#pragma unroll
    for (int i = 0; i < kXPer; ++i) {
      const int idx = l + 64 * i, grp = idx & 7, px = idx >> 3, c = ci0 + grp * 4;
      if (idx >= (kSeg + 2) * 8) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        xT[(grp * 4 + j) * kRow + px] = c + j < cin ? uint16_t(pxv[i][j]) : uint16_t(0);
    }
  };
  if (nst > 0) fetch(0);
  for (int st = 0; st < nst; ++st) {
    store();
    __syncthreads();
    if (st + 1 < nst) fetch(st + 1);
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
  // one thread per 8-channel granule (16-byte store); in_pad % 8 == 0
  const int groups = in_pad / 8;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * groups) return;
  const int p = i / groups, c0 = (i - p * groups) * 8;
  const float* src = in + size_t(p) * inputs;
  uint32_t w[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int c = c0 + 2 * e;
    const uint32_t lo = c < inputs ? f2bf(src[c]) : 0u, hi = c + 1 < inputs ? f2bf(src[c + 1]) : 0u;
    w[e] = lo | (hi << 16);
  }
  *reinterpret_cast<uint4*>(x + size_t(p) * in_pad + c0) = make_uint4(w[0], w[1], w[2], w[3]);
}
// pred [n][3] = sum_j softmax(z)_j * rgb(neighbour j)
__global__ void kpn_apply_kernel(const float* __restrict__ in, int inputs,
                                 const float* __restrict__ logits, int H, int W,
                                 float* __restrict__ pred) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= H * W) return;
  const int py = p / W, px = p - py * W;
  kpn_filter(in, inputs, logits + size_t(p) * kLogitStride, py, px, H, W, pred + size_t(p) * 3);
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

// compute units of the current device (cached per device)
int cu_count() {
  static std::atomic<int> counts[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = counts[dev].load(std::memory_order_relaxed);
  if (n == 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    counts[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

}  // namespace

hipError_t conv(const ConvArgs& a, int src, int epi, hipStream_t s) {
  // 16-byte granules of input and weights
  if (a.cin % 8 || a.cout_pad % 32 || a.xs % 8 || a.xo % 8 || a.ys % 4 || a.yo % 4)
    return hipErrorInvalidValue;
  if (src == kSrcPlain && a.xo + a.cin > a.xs) return hipErrorInvalidValue;
  if (src == kSrcUpCat && (a.cu % 8 || a.cu >= a.cin || a.xo + a.cin - a.cu > a.xs || a.H % 2 ||
                           a.W % 2 || !a.xl))
    return hipErrorInvalidValue;
  if (src == kSrcF32 && (!a.kin || a.inputs > a.cin)) return hipErrorInvalidValue;
  if (epi == kEpiPool && (a.H % 2 || a.W % 2 || !a.pool16 || !a.y16)) return hipErrorInvalidValue;
  if (epi == kEpiKpn && (a.cout_pad != kLogitStride || !a.kin || !a.pred)) return hipErrorInvalidValue;
  // workgroup shape: 2 x 2 waves of 32 pixels x 2 rows (64 pixels x 4 rows),
  // each with 64 output channels, 96 (the 96-channel level) or 32
  const int shape = a.cout_pad % 64 == 0 ? 2 : a.cout_pad % 96 == 0 ? 3 : 1;
  const int TP = 64, R = 4, TN = 32 * shape;
  if (epi == kEpiKpn && shape != 2) return hipErrorInvalidValue;
  // a resident grid (two workgroups per CU, one for the 96-channel shape)
  // whose workgroups loop over the tiles
  const int tiles = ((a.W + TP - 1) / TP) * ((a.H + R - 1) / R), ny = a.cout_pad / TN;
  const int resident = std::max(1, cu_count() * (shape == 3 ? 1 : 2) / ny);
  const dim3 grid(unsigned(std::min(tiles, resident)), unsigned(ny));
  const int key = shape * 100 + src * 10 + epi;
  switch (key) {
    case 100: hipLaunchKernelGGL((conv_kernel<2, 2, 1, kSrcPlain, kEpiPlain>), grid, dim3(256), 0, s, a); break;
    case 101: hipLaunchKernelGGL((conv_kernel<2, 2, 1, kSrcPlain, kEpiPool>), grid, dim3(256), 0, s, a); break;
    case 110: hipLaunchKernelGGL((conv_kernel<2, 2, 1, kSrcUpCat, kEpiPlain>), grid, dim3(256), 0, s, a); break;
    case 120: hipLaunchKernelGGL((conv_kernel<2, 2, 1, kSrcF32, kEpiPlain>), grid, dim3(256), 0, s, a); break;
    case 200: hipLaunchKernelGGL((conv_kernel<2, 2, 2, kSrcPlain, kEpiPlain>), grid, dim3(256), 0, s, a); break;
    case 201: hipLaunchKernelGGL((conv_kernel<2, 2, 2, kSrcPlain, kEpiPool>), grid, dim3(256), 0, s, a); break;
    case 202: hipLaunchKernelGGL((conv_kernel<2, 2, 2, kSrcPlain, kEpiKpn>), grid, dim3(256), 0, s, a); break;
    case 210: hipLaunchKernelGGL((conv_kernel<2, 2, 2, kSrcUpCat, kEpiPlain>), grid, dim3(256), 0, s, a); break;
    case 300: hipLaunchKernelGGL((conv_kernel<2, 2, 3, kSrcPlain, kEpiPlain>), grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t conv3x3(const uint16_t* x, int H, int W, int xs, int xo, int cin, const uint16_t* w,
                   const float* bias, int cout, int cout_pad, uint16_t* y16, float* y32, int ys,
                   int yo, bool relu, bool accum, hipStream_t s) {
  ConvArgs a{};
  a.x = x;
  a.xs = xs;
  a.xo = xo;
  a.H = H;
  a.W = W;
  a.cin = cin;
  a.w = w;
  a.bias = bias;
  a.cout = cout;
  a.cout_pad = cout_pad;
  a.y16 = y16;
  a.y32 = y32;
  a.ys = ys;
  a.yo = yo;
  a.relu = relu ? 1 : 0;
  a.accum = accum ? 1 : 0;
  return conv(a, kSrcPlain, kEpiPlain, s);
}

// waves of one weight-gradient launch (8192 measured no faster: the band
// reduction grows with them)
constexpr int kWgradWaves = 2048;

hipError_t wgrad(const uint16_t* g, int gs, const uint16_t* x, int xs, int xo, int H, int W,
                 int cin, int cout, float* dw, float* part, size_t part_cap, hipStream_t s) {
  // 8-byte channel groups: strides and offsets in whole groups, and the
  // groups that hold the last channels inside the row
  if (gs % 4 || xs % 4 || xo % 4 || (cout + 3) / 4 * 4 > gs || xo + (cin + 3) / 4 * 4 > xs)
    return hipErrorInvalidValue;
  const int tiles = ((cin + 31) / 32) * ((cout + 31) / 32);
  const size_t per_band = size_t(9) * cout * cin;
  if (per_band > part_cap) return hipErrorInvalidValue;
  // about kWgradWaves waves (3 tap rows per band), within the partial-sum buffer
  int bands = std::max(1, std::min(H, std::min((kWgradWaves / 3 + tiles - 1) / tiles, int(part_cap / per_band))));
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
  if (in_pad % 8) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pack_input_kernel, dim3(nb(size_t(n) * (in_pad / 8))), dim3(256), 0, s, in, n,
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
