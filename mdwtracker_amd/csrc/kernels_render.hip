// Frame kernel of one view, one thread per pixel (kernels (b)+(c)): per-pixel
// ray generation, root clip, 3D-DDA over the majorant grid and the
// per-channel Woodcock tracking of every -m mode (render_paths.hpp), plus
// accumulation; also the test-hook kernels and the frame launcher (which also
// dispatches the wave-coherent schedule of kernels_wave.hip).
//
// Reference: mainRG (deviceCode.cu:283-479) launched as one OptiX thread per
// pixel and volumeIntersection (deviceCode.cu:1676-1726). Here OptiX is gone:
// a workgroup of one to four wave64 owns an 8x8 / 16x8 / 16x16 pixel block,
// each wave an 8x8 quad (rays of a wave start close together and walk the
// same macrocells); blocks are issued in Hilbert or cost-sorted order, dealt
// round-robin over the eight XCDs; transfer functions sit in LDS; voxels are
// read in their native dtype from planar channels in HBM.
#include <mutex>

#include "mdwt_internal.hpp"
#include "render_paths.hpp"

namespace mdwt {

// Staging area of the allocation-free frame-buffer read
// (mdwt_read_fb_noalloc): RGBA8 of up to kFbScratchPixels pixels. Every GPU
// the library runs on holds its own copy until the process ends; calls that
// overlap on one GPU share it and race. Host code never passes `scratch`
// itself (in host code the name is a stand-in, not a device address): it
// uses the per-device address that fb_scratch_setup takes from
// hipGetSymbolAddress.
__device__ __align__(256) uint32_t scratch[kFbScratchPixels];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[kMaxDevices] = {};

// LIT: shadow rays, gradient shading and the denoiser's auxiliary outputs
// compiled in. The other instances keep the register budget of the plain
// method (the extra code needs ~15 more VGPRs, which would cost a wave per
// SIMD). Workgroup i renders pixel block
// block_order[i] of the single view.
template <int MODE, int DT, bool COUNT, bool LIT>
__device__ __forceinline__ void render_body(const RenderParams& P) {
  extern __shared__ float4 s_tf[];
  __shared__ ChanTf s_ch[kMaxChannels];
  const uint64_t t_begin = P.fr.wave_cost ? wall_clock64() : 0ull;
  load_tf_lds(P, s_tf, s_ch);
  __syncthreads();

  Ctx<DT> X;
  X.init(P, s_tf, s_ch);
  Counts cnt = {0, 0, 0, 0};

  const uint32_t slot = block_slot(blockIdx.x, uint32_t(P.fr.n_blocks));
  const uint32_t bxy = P.fr.block_order[slot];
  ViewCtx V;
  V.cam = &P.cam;
  V.W = P.fr.width;
  V.H = P.fr.height;
  V.spp = P.fr.spp;
  V.seed = P.fr.seed;
  V.frame_id = P.fr.frame_id;
  V.accum_id = P.fr.accum_id;
  V.accum = P.fr.accum;
  V.m2 = P.fr.m2;
  render_block<MODE, DT, COUNT, LIT>(P, X, V, int(bxy & 0xffffu), int(bxy >> 16), cnt);
  if (COUNT) flush_counts(P, cnt);
  // per-wave duration for the cost-sorted block order of later frames
  if (!COUNT && P.fr.wave_cost && (threadIdx.x & 63) == 0)
    P.fr.wave_cost[(blockDim.x >> 6) * slot + (threadIdx.x >> 6)] = uint32_t(wall_clock64() - t_begin);
}

template <int MODE, int DT, bool COUNT, bool LIT>
__global__ __launch_bounds__(256) void render_kernel(RenderParams P) {
  render_body<MODE, DT, COUNT, LIT>(P);
}

// -m 0 and -m 2 without lighting at six waves per SIMD (80 VGPRs; the
// compiler spills the rest to scratch): more waves hide the latency of the
// sample's dependent loads and the cell walk better than the spills cost
// (-m 2 0.334 -> 0.317 ms, -m 0 0.370 -> 0.345 ms per 1024^2 frame on four
// 256^3 u8 channels). -m 3 and the lit kernels spill too much and stay at
// their natural budget.
template <int MODE, int DT, bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void render_kernel_w6(
    RenderParams P) {
  render_body<MODE, DT, COUNT, false>(P);
}

// ---------------------------------------------------------- test hooks -----
__global__ void gen_rays_kernel(CameraParams cam, int W, int H, uint32_t seed, uint32_t frame,
                                uint32_t sample, int jitter2d, float* org, float* dir) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  const int px = i % W, py = i / W;
  Rng rng;
  rng.init(seed, uint32_t(i), frame, sample, 0u);
  float jx, jy;
  rng.jitter2(jitter2d != 0, jx, jy);
  V3 o, d;
  camera_ray(cam, px, py, W, H, jx, jy, o, d);
  org[3 * i] = o.x;
  org[3 * i + 1] = o.y;
  org[3 * i + 2] = o.z;
  dir[3 * i] = d.x;
  dir[3 * i + 1] = d.y;
  dir[3 * i + 2] = d.z;
}

__global__ void clip_kernel(VolumeParams vol, const float* org, const float* dir, int n, float* t01,
                            int* hit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float t0 = 0.f, t1 = 0.f;
  const bool h = clip_root(load3(org + 3 * i), load3(dir + 3 * i), load3(vol.blo), load3(vol.bhi),
                           t0, t1);
  t01[2 * i] = t0;
  t01[2 * i + 1] = t1;
  hit[i] = h ? 1 : 0;
}

__global__ void trace_cells_kernel(VolumeParams vol, int mx, int my, int mz, const float* org,
                                   const float* dir, int n, int max_cells, int* cells, float* t01,
                                   int* counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const GridRay g = make_grid_ray(load3(org + 3 * i), load3(dir + 3 * i), load3(vol.blo),
                                  load3(vol.bhi), mx, my, mz);
  Dda d;
  int k = 0;
  if (d.begin(g.org, g.dir, 1e20f, mx, my, mz)) {
    while (true) {
      float t0, t1;
      d.range(t0, t1);
      if (k < max_cells) {
        const size_t o = size_t(i) * max_cells + k;
        cells[o] = d.cx + d.cy * mx + d.cz * mx * my;
        t01[2 * o] = t0;
        t01[2 * o + 1] = t1;
      }
      ++k;
      if (!d.advance()) break;
    }
  }
  for (int j = k; j < max_cells; ++j) {  // unused slots: cell -1, t = 0
    const size_t o = size_t(i) * max_cells + j;
    cells[o] = -1;
    t01[2 * o] = 0.f;
    t01[2 * o + 1] = 0.f;
  }
  counts[i] = k;
}

// make_rgba (OWL) of the accumulated colour: min(255, max(0, int(256 c)))
__device__ __forceinline__ uint32_t to8(float f) {
  const float c = fminf(fmaxf(f * 256.f, -1.f), 256.f);
  const int i = int(c);
  return uint32_t(i < 0 ? 0 : (i > 255 ? 255 : i));
}
__global__ void rgba8_kernel(const float4* acc, uint32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 c = acc[i];
  out[i] = to8(c.x) | (to8(c.y) << 8) | (to8(c.z) << 16) | (to8(c.w) << 24);
}

__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ src,
                                                   float4* __restrict__ dst, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

// Read-only stream over n float4 (the HBM bandwidth a read-dominated kernel
// can reach): each thread sums four independent float4 loads per iteration of
// a grid-stride loop, and each workgroup writes one partial sum, so the only
// writes are kReadGrid floats per launch.
__global__ __launch_bounds__(256) void read_kernel(const float4* __restrict__ src, size_t n,
                                                   float* __restrict__ sums) {
  __shared__ float s_part[4];
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a, c = a, d = a;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const float4 x0 = src[i], x1 = src[i + stride], x2 = src[i + 2 * stride],
                 x3 = src[i + 3 * stride];
    a.x += x0.x; a.y += x0.y; a.z += x0.z; a.w += x0.w;
    b.x += x1.x; b.y += x1.y; b.z += x1.z; b.w += x1.w;
    c.x += x2.x; c.y += x2.y; c.z += x2.z; c.w += x2.w;
    d.x += x3.x; d.y += x3.y; d.z += x3.z; d.w += x3.w;
  }
  for (; i < n; i += stride) {
    const float4 x0 = src[i];
    a.x += x0.x; a.y += x0.y; a.z += x0.z; a.w += x0.w;
  }
  float v = (a.x + a.y + a.z + a.w) + (b.x + b.y + b.z + b.w) + (c.x + c.y + c.z + c.w) +
            (d.x + d.y + d.z + d.w);
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) sums[blockIdx.x] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
}

// Blend tracking (-m 4 / 5) needs more than 128 VGPRs, so it runs at three
// waves per SIMD whatever it is given. Without that stated, the scheduler
// squeezed it towards four waves by issuing the blended sample's voxel loads
// one or two at a time with a full wait in between (0.41 -> 0.9 ms per frame
// from an unrelated four-register change); with the target stated it keeps
// them in flight together.
template <int MODE, int DT, bool COUNT, bool LIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 3))) void render_kernel_w3(
    RenderParams P) {
  render_body<MODE, DT, COUNT, LIT>(P);
}

// wide: the six-wave entry for -m 0 / -m 2 without lighting (see
// render_kernel_w6; chosen per launch by six_waves()).
template <int MODE, int DT, bool COUNT, bool LIT>
RenderFn static_fn(bool wide) {
  if constexpr ((MODE == 0 || MODE == 2) && !LIT) {
    if (wide) return render_kernel_w6<MODE, DT, COUNT>;
    return render_kernel<MODE, DT, COUNT, LIT>;
  } else if constexpr (MODE == 4 || MODE == 5) {
    return render_kernel_w3<MODE, DT, COUNT, LIT>;
  } else {
    return render_kernel<MODE, DT, COUNT, LIT>;
  }
}
template <int MODE, bool COUNT, bool LIT>
RenderFn static_fn_dt(int dtype, bool wide) {
  switch (dtype) {
    case MDWT_DT_U8: return static_fn<MODE, MDWT_DT_U8, COUNT, LIT>(wide);
    case MDWT_DT_U16: return static_fn<MODE, MDWT_DT_U16, COUNT, LIT>(wide);
    case MDWT_DT_F16: return static_fn<MODE, MDWT_DT_F16, COUNT, LIT>(wide);
    default: return static_fn<MODE, MDWT_DT_F32, COUNT, LIT>(wide);
  }
}

#define MDWT_STATIC(M)                                                                       \
  (count ? (lit ? static_fn_dt<M, true, true>(dtype, wide) : static_fn_dt<M, true, false>(dtype, wide)) \
         : (lit ? static_fn_dt<M, false, true>(dtype, wide) : static_fn_dt<M, false, false>(dtype, wide)))
// (-m 9 has no lighting: its shadow code is commented out in the reference;
// its LIT instance only adds the auxiliary outputs.)
RenderFn select_render(int mode, int dtype, bool count, bool lit, bool wide) {
  switch (mode) {  // modes with the thread-per-pixel kernel only
    case 1: return MDWT_STATIC(1);
    case 4: return MDWT_STATIC(4);
    case 5: return MDWT_STATIC(5);
    case 6: return MDWT_STATIC(6);
    case 7: return MDWT_STATIC(7);
    case 8: return MDWT_STATIC(8);
    case 9: return MDWT_STATIC(9);
    default: break;
  }
  if (mode == 0) return MDWT_STATIC(0);
  if (mode == 2) return MDWT_STATIC(2);
  return MDWT_STATIC(3);
}
#undef MDWT_STATIC

}  // namespace

size_t render_lds_bytes(const RenderParams& p) {
  return size_t(p.vol.channels) * size_t(p.tf.stride) * 16;
}

namespace {
// the extended kernel: shadow rays, gradient shading or denoiser inputs
bool is_lit(const RenderParams& p) {
  return p.fr.sh.shadows != 0 || p.fr.sh.gradient != 0 || p.fr.aov != nullptr;
}
// Six waves per SIMD (with scratch spills) pay on at most four channels of
// 1- or 2-byte voxels: -m 2 0.334 -> 0.317 ms on four 256^3 u8 channels. On
// four 512^3 f16 channels (0.303 -> 0.329 ms) and on ten f32 channels the
// natural four waves are faster.
bool six_waves(const RenderParams& p, int dtype) {
  return p.vol.channels <= 4 && (dtype == MDWT_DT_U8 || dtype == MDWT_DT_U16);
}
// Kernels whose static plus dynamic LDS passes 64 KiB must say so first.
hipError_t prepare(RenderFn f, size_t lds) {
  hipFuncAttributes a{};
  hipError_t e = hipFuncGetAttributes(&a, reinterpret_cast<const void*>(f));
  if (e != hipSuccess) return e;
  if (a.sharedSizeBytes + lds <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(f),
                             hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
}
}  // namespace

hipError_t launch_render(const RenderParams& p, int mode, int dtype, bool count, hipStream_t s) {
  const bool lit = is_lit(p);
  const bool wave = p.fr.schedule == 1 && wave_schedule_applies(p, mode, count, lit);
  const size_t lds = wave ? wave_lds_bytes(p) : render_lds_bytes(p);
  const RenderFn f = wave ? select_wave_kernel(p, mode, dtype, 0)
                          : select_render(mode, dtype, count, lit, six_waves(p, dtype));
  const hipError_t e = prepare(f, lds);
  if (e != hipSuccess) return e;
  // one workgroup per pixel block
  hipLaunchKernelGGL(f, dim3(unsigned(p.fr.n_blocks)), dim3(unsigned(p.fr.bw * p.fr.bh)), lds, s, p);
  return hipGetLastError();
}

hipError_t launch_gen_rays(const CameraParams& cam, int W, int H, uint32_t seed, uint32_t frame,
                           uint32_t sample, int jitter2d, float* org, float* dir, hipStream_t s) {
  const int n = W * H;
  hipLaunchKernelGGL(gen_rays_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, cam, W, H,
                     seed, frame, sample, jitter2d, org, dir);
  return hipGetLastError();
}

hipError_t launch_clip(const VolumeParams& vol, const float* org, const float* dir, int n,
                       float* t01, int* hit, hipStream_t s) {
  hipLaunchKernelGGL(clip_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, vol, org, dir,
                     n, t01, hit);
  return hipGetLastError();
}

hipError_t launch_trace_cells(const VolumeParams& vol, const int mc[3], const float* org,
                              const float* dir, int n, int max_cells, int* cells, float* t01,
                              int* counts, hipStream_t s) {
  hipLaunchKernelGGL(trace_cells_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, vol,
                     mc[0], mc[1], mc[2], org, dir, n, max_cells, cells, t01, counts);
  return hipGetLastError();
}

hipError_t fb_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = static_cast<uint32_t*>(p);
  return hipSuccess;
}

uint32_t* fb_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch[device];
}

hipError_t launch_rgba8(const void* accum, uint32_t* out, int n, hipStream_t s) {
  hipLaunchKernelGGL(rgba8_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(accum), out, n);
  return hipGetLastError();
}

hipError_t launch_read(const void* src, size_t n16, float* sums, hipStream_t s) {
  hipLaunchKernelGGL(read_kernel, dim3(unsigned(kReadGrid)), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(src), n16, sums);
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t n16, hipStream_t s) {
  hipLaunchKernelGGL(copy_kernel, dim3(2048u), dim3(256), 0, s, reinterpret_cast<const float4*>(src),
                     reinterpret_cast<float4*>(dst), n16);
  return hipGetLastError();
}

}  // namespace mdwt
