// Launch-parameter structs shared by the host library and the kernels. They
// take the place of the reference's LaunchParams (deviceCode.h:61-136): one
// by-value kernel argument per launch instead of OWL named variables.
#pragma once
#include <stdint.h>

namespace mdwt {

enum { MDWT_DT_U8 = 0, MDWT_DT_U16 = 1, MDWT_DT_F16 = 2, MDWT_DT_F32 = 3 };
constexpr int kMaxChannels = 12;   // deviceCode.h:13
// auxiliary buffers: per pixel first-hit distance, transmittance and the hit
// fraction of every channel, packed into float4 groups
constexpr int aov_groups(int channels) { return (2 + channels + 3) / 4; }
constexpr int kBlockPixels = 16;   // a frame workgroup covers 16x16 pixels
constexpr int kThreads = 256;      // 4 wave64, each an 8x8 pixel quad
constexpr int kMaxTexelsLds = 10240;  // C * stride float4 texels (160 KiB)

struct CameraParams {
  float org[3], llc[3], hor[3], ver[3];  // LaunchParams::camera
};

// Volume in HBM, two copies: channel c's voxels in planar order (x
// fastest, native dtype) at vol + c * chan_stride bytes, read by the
// macrocell builds, and the copy the samplers read, in 4 x 4 x 4 bricks
// (bricks in x-fastest order, 64 voxels each, x fastest inside; channel c at
// bricks + c * brick_stride; nbx, nby bricks along x, y), where a trilinear
// footprint touches about 1.8 cache lines instead of 4.
// value = raw * scale + offset (rawFile.cpp:100-123).
struct VolumeParams {
  const uint8_t* vol;
  uint64_t chan_stride;
  const uint8_t* bricks;
  uint64_t brick_stride;
  int nbx, nby;
  int nx, ny, nz;
  int channels;
  float scale, offset;
  float blo[3], bhi[3];  // world box (root box = union of channel bounds)
};

// Per-channel transfer function (LaunchParams::transferFunction): texels of
// channel c at tab + c * stride (alpha already cubed).
struct TfParams {
  const void* tab;  // float4
  int stride;
  int n[kMaxChannels];
  float opacity[kMaxChannels];
  float xf_lo[kMaxChannels], xf_hi[kMaxChannels];
  float dom_lo[kMaxChannels], dom_hi[kMaxChannels];
};

// Lighting and per-channel materials (LaunchParams lightDir, enableShadows,
// enableGradientShading, material[], deviceCode.h:72-79; renderer.h:121-142).
struct ShadeParams {
  float light_dir[3];
  int shadows, gradient;
  float diffuse[kMaxChannels], specular[kMaxChannels];
  float shininess[kMaxChannels], grad_step[kMaxChannels];
  float spec[kMaxChannels][3];
};

// One view of a batched launch: camera, size, samples, RNG seed and its own
// accumulators (a batched view's image equals a single-view render with the
// same camera, size, spp and seed).
struct ViewParams {
  CameraParams cam;
  int width, height, spp;
  uint32_t seed;
  void* accum;  // float4 per pixel
  void* m2;     // float4 per pixel
};

struct FrameParams {
  int width, height, spp, accumulate;
  float bg[3];
  float light_intensity, ambient;
  ShadeParams sh;
  uint32_t seed, frame_id, accum_id;
  // 1: the y jitter is word 1 of the jitter block; 0: both axes take word 0,
  // the reference's one scalar (deviceCode.cu:298)
  int jitter2d;
  void* accum;  // float4 per pixel
  void* m2;     // float4 per pixel
  unsigned long long* counters;  // samples, visits, rejections, rays
  const uint32_t* block_order;   // (bx | by << 16) per 16x16 pixel block, Hilbert order
  int n_blocks;
  int schedule;  // 0: one thread per pixel, 1: wave-coherent (-m 0 to -m 5, no lighting)
  // cost-sorted order: wall-clock ticks (100 MHz) of each wave, indexed by
  // waves_per_block * (position in block_order) + wave; nullptr = not recorded
  uint32_t* wave_cost;
  int bw, bh;  // pixel block of a workgroup: 8x8, 16x8 or 16x16 (bw*bh/64 waves of 8x8 pixels)
  // batched views (views == nullptr: the single view described above). An
  // item is one pixel block of one view (view << 20 | by << 10 | bx); launch
  // form 2 reads the frame and accumulation ids from work[1] / work[2], which
  // the device advances so that a captured graph replays frame after frame.
  const ViewParams* views;
  const uint32_t* items;
  int n_items;
  unsigned* work;
  unsigned long long* item_ticks;  // start / end wall clock per item, nullptr = off
  // auxiliary per-pixel buffers for the denoiser: aov_groups(C) float4 per
  // pixel, running means over samples of (first-hit distance, transmittance,
  // hit fraction of each channel); nullptr = off
  void* aov;
};

struct RenderParams {
  VolumeParams vol;
  TfParams tf;
  CameraParams cam;
  int mc[3];
  const float* maj;
  float dt;
  FrameParams fr;
};

// Range-grid build: per axis and macrocell index m, the inclusive voxel
// window [lo, hi] whose values bound the cell (lo > hi: no voxel covers it).
struct RangeParams {
  const uint8_t* vol;
  uint64_t chan_stride;
  int nx, ny, nz;
  int channels;
  float scale, offset;
  int mc[3];
  const int* win;  // 2*(mx + my + mz) ints: x windows, then y, then z
  void* ranges;    // float2 per (cell, channel)
};

struct MajorantParams {
  const void* ranges;  // float2 per (cell, channel)
  TfParams tf;
  int channels;
  int n_cells;
  int mode;
  int conservative;  // 1: channel_majorant_conservative (an option), 0: the reference's
  float* maj;
};

}  // namespace mdwt
