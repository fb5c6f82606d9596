// Batched views of uneven size and spp.
// render_views_kernel (launch form 1): one ordinary launch per frame, one
// workgroup per (view, pixel block) item, every view of the frame in a single
// grid, so small views share the machine with large ones without a launch
// each.
// render_views_persistent (launch form 2): one persistent launch per frame,
// captured once in a HIP graph with the kernel that advances the frame and
// accumulation ids it reads from device memory, and replayed frame after
// frame. Workgroup g renders items g, g + G, g + 2G, ... (G = the resident
// grid): a static split with no device work counter, no atomics and no
// barrier inside the item loop, so every workgroup's loop ends after a fixed
// number of items.
// Under the wave-coherent schedule, -m 0 to -m 5 views without lighting run
// kernels_wave.hip's render_views_wave and render_views_wave_persistent
// instead, with the same items and ids.
// No reference counterpart: the reference renders one camera per owlLaunch2D
// (renderer.cpp:949).
#include "mdwt_internal.hpp"
#include "render_paths.hpp"

namespace mdwt {
namespace {

// One launch for all views, one workgroup per work item: workgroup i renders
// item i (view << 20 | by << 10 | bx). Items run heaviest view first, each
// view's blocks in Hilbert order, so consecutive workgroups (dealt
// round-robin over the XCDs) cover neighbouring blocks of one view. Frame and
// accumulation ids are launch arguments shared by every view. With
// item_ticks set, each workgroup records its start and end wall clock.
template <int MODE, int DT, bool COUNT, bool LIT>
__device__ __forceinline__ void render_views_body(const RenderParams& P) {
  extern __shared__ float4 s_tf[];
  __shared__ ChanTf s_ch[kMaxChannels];
  const uint64_t t_begin = P.fr.item_ticks ? wall_clock64() : 0ull;
  load_tf_lds(P, s_tf, s_ch);
  __syncthreads();

  Ctx<DT> X;
  X.init(P, s_tf, s_ch);
  Counts cnt = {0, 0, 0, 0};
  const uint32_t code = P.fr.items[blockIdx.x];
  const ViewParams* vp = P.fr.views + (code >> 20);
  ViewCtx V;
  V.cam = &vp->cam;
  V.W = vp->width;
  V.H = vp->height;
  V.spp = vp->spp;
  V.seed = vp->seed;
  V.frame_id = P.fr.frame_id;
  V.accum_id = P.fr.accum_id;
  V.accum = vp->accum;
  V.m2 = vp->m2;
  render_block<MODE, DT, COUNT, LIT>(P, X, V, int(code & 0x3ffu), int((code >> 10) & 0x3ffu),
                                     cnt);
  if (COUNT) flush_counts(P, cnt);
  if (!COUNT && P.fr.item_ticks) {
    __syncthreads();
    if (threadIdx.x == 0) {
      P.fr.item_ticks[2 * blockIdx.x] = t_begin;
      P.fr.item_ticks[2 * blockIdx.x + 1] = wall_clock64();
    }
  }
}

// Launch form 2: workgroup g renders items g, g + G, g + 2G, ... at the frame
// and accumulation ids held in work[1] / work[2] (device memory, advanced by
// views_advance_kernel after each frame, so a captured graph replays frame
// after frame). No barrier, atomic or device counter inside the item loop.
// Wave 0 records each item's start and end wall clock (the view latencies).
template <int MODE, int DT, bool LIT>
__device__ __forceinline__ void render_views_persistent_body(const RenderParams& P) {
  extern __shared__ float4 s_tf[];
  __shared__ ChanTf s_ch[kMaxChannels];
  load_tf_lds(P, s_tf, s_ch);
  __syncthreads();

  Ctx<DT> X;
  X.init(P, s_tf, s_ch);
  Counts cnt = {0, 0, 0, 0};
  const uint32_t fid = P.fr.work[1], aid = P.fr.work[2];
  const uint32_t n = uint32_t(P.fr.n_items);
  for (uint32_t it = blockIdx.x; it < n; it += gridDim.x) {
    const uint64_t t_item = P.fr.item_ticks ? wall_clock64() : 0ull;
    const uint32_t code = P.fr.items[it];
    const ViewParams* vp = P.fr.views + (code >> 20);
    ViewCtx V;
    V.cam = &vp->cam;
    V.W = vp->width;
    V.H = vp->height;
    V.spp = vp->spp;
    V.seed = vp->seed;
    V.frame_id = fid;
    V.accum_id = aid;
    V.accum = vp->accum;
    V.m2 = vp->m2;
    render_block<MODE, DT, false, LIT>(P, X, V, int(code & 0x3ffu), int((code >> 10) & 0x3ffu),
                                       cnt);
    if (P.fr.item_ticks && threadIdx.x == 0) {
      P.fr.item_ticks[2 * it] = t_item;
      P.fr.item_ticks[2 * it + 1] = wall_clock64();
    }
  }
}

// Occupancy targets as for the single-view frame kernel (kernels_render.hip):
// six waves per SIMD for -m 0 / -m 2 without lighting on at most four
// channels of 1- or 2-byte voxels (wide), a stated three for blend tracking,
// the natural budget otherwise.
template <int MODE, int DT, bool COUNT, bool LIT>
__global__ __launch_bounds__(256) void render_views_kernel(RenderParams P) {
  render_views_body<MODE, DT, COUNT, LIT>(P);
}
template <int MODE, int DT, bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void render_views_kernel_w6(
    RenderParams P) {
  render_views_body<MODE, DT, COUNT, false>(P);
}
template <int MODE, int DT, bool COUNT, bool LIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 3))) void render_views_kernel_w3(
    RenderParams P) {
  render_views_body<MODE, DT, COUNT, LIT>(P);
}
template <int MODE, int DT, bool LIT>
__global__ __launch_bounds__(256) void render_views_persistent(RenderParams P) {
  render_views_persistent_body<MODE, DT, LIT>(P);
}
template <int MODE, int DT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void
render_views_persistent_w6(RenderParams P) {
  render_views_persistent_body<MODE, DT, false>(P);
}
template <int MODE, int DT, bool LIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 3))) void
render_views_persistent_w3(RenderParams P) {
  render_views_persistent_body<MODE, DT, LIT>(P);
}
template <int MODE, int DT, bool LIT>
RenderFn persistent_fn(bool wide) {
  if constexpr ((MODE == 0 || MODE == 2) && !LIT) {
    if (wide) return render_views_persistent_w6<MODE, DT>;
    return render_views_persistent<MODE, DT, LIT>;
  } else if constexpr (MODE == 4 || MODE == 5) {
    return render_views_persistent_w3<MODE, DT, LIT>;
  } else {
    return render_views_persistent<MODE, DT, LIT>;
  }
}
template <int MODE, bool LIT>
RenderFn persistent_fn_dt(int dtype, bool wide) {
  switch (dtype) {
    case MDWT_DT_U8: return persistent_fn<MODE, MDWT_DT_U8, LIT>(wide);
    case MDWT_DT_U16: return persistent_fn<MODE, MDWT_DT_U16, LIT>(wide);
    case MDWT_DT_F16: return persistent_fn<MODE, MDWT_DT_F16, LIT>(wide);
    default: return persistent_fn<MODE, MDWT_DT_F32, LIT>(wide);
  }
}
#define MDWT_PERSISTENT(M) \
  (lit ? persistent_fn_dt<M, true>(dtype, wide) : persistent_fn_dt<M, false>(dtype, wide))
RenderFn select_persistent(int mode, int dtype, bool lit, bool wide) {
  switch (mode) {
    case 0: return MDWT_PERSISTENT(0);
    case 1: return MDWT_PERSISTENT(1);
    case 2: return MDWT_PERSISTENT(2);
    case 3: return MDWT_PERSISTENT(3);
    case 4: return MDWT_PERSISTENT(4);
    case 5: return MDWT_PERSISTENT(5);
    case 6: return MDWT_PERSISTENT(6);
    case 7: return MDWT_PERSISTENT(7);
    case 8: return MDWT_PERSISTENT(8);
    default: return MDWT_PERSISTENT(9);
  }
}
#undef MDWT_PERSISTENT

// frame id + 1, accumulation id + 1 (0 without accumulation): the last node
// of a captured batched frame
__global__ void views_advance_kernel(unsigned* work, int accumulate) {
  if (threadIdx.x == 0) {
    work[1] = work[1] + 1u;
    work[2] = accumulate ? work[2] + 1u : 0u;
  }
}

template <int MODE, int DT, bool COUNT, bool LIT>
RenderFn views_fn(bool wide) {
  if constexpr ((MODE == 0 || MODE == 2) && !LIT) {
    if (wide) return render_views_kernel_w6<MODE, DT, COUNT>;
    return render_views_kernel<MODE, DT, COUNT, LIT>;
  } else if constexpr (MODE == 4 || MODE == 5) {
    return render_views_kernel_w3<MODE, DT, COUNT, LIT>;
  } else {
    return render_views_kernel<MODE, DT, COUNT, LIT>;
  }
}

template <int MODE, bool COUNT, bool LIT>
RenderFn views_fn_dt(int dtype, bool wide) {
  switch (dtype) {
    case MDWT_DT_U8: return views_fn<MODE, MDWT_DT_U8, COUNT, LIT>(wide);
    case MDWT_DT_U16: return views_fn<MODE, MDWT_DT_U16, COUNT, LIT>(wide);
    case MDWT_DT_F16: return views_fn<MODE, MDWT_DT_F16, COUNT, LIT>(wide);
    default: return views_fn<MODE, MDWT_DT_F32, COUNT, LIT>(wide);
  }
}

#define MDWT_VIEWS(M)                                                                 \
  (count ? (lit ? views_fn_dt<M, true, true>(dtype, wide) : views_fn_dt<M, true, false>(dtype, wide)) \
         : (lit ? views_fn_dt<M, false, true>(dtype, wide) : views_fn_dt<M, false, false>(dtype, wide)))
RenderFn select_views(int mode, int dtype, bool count, bool lit, bool wide) {
  switch (mode) {
    case 0: return MDWT_VIEWS(0);
    case 1: return MDWT_VIEWS(1);
    case 2: return MDWT_VIEWS(2);
    case 3: return MDWT_VIEWS(3);
    case 4: return MDWT_VIEWS(4);
    case 5: return MDWT_VIEWS(5);
    case 6: return MDWT_VIEWS(6);
    case 7: return MDWT_VIEWS(7);
    case 8: return MDWT_VIEWS(8);
    default: return MDWT_VIEWS(9);
  }
}
#undef MDWT_VIEWS

// the extended kernel: shadow rays, gradient shading or denoiser inputs
bool is_lit(const RenderParams& p) {
  return p.fr.sh.shadows != 0 || p.fr.sh.gradient != 0 || p.fr.aov != nullptr;
}

hipError_t prepare(RenderFn f, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(f),
                             hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
}

}  // namespace

namespace {
bool wide_views(const RenderParams& p, int dtype) {
  return p.vol.channels <= 4 && (dtype == MDWT_DT_U8 || dtype == MDWT_DT_U16);
}
// batched views on the wave-coherent kernels (schedule 1 where it applies)
bool wave_views(const RenderParams& p, int mode, bool count) {
  return p.fr.schedule == 1 && wave_schedule_applies(p, mode, count, is_lit(p));
}
}  // namespace

hipError_t views_persistent_blocks_per_cu(const RenderParams& p, int mode, int dtype, int* out) {
  const bool wave = wave_views(p, mode, false);
  const size_t lds = wave ? wave_lds_bytes(p) : render_lds_bytes(p);
  const RenderFn f = wave ? select_wave_kernel(p, mode, dtype, 2)
                          : select_persistent(mode, dtype, is_lit(p), wide_views(p, dtype));
  const hipError_t e = prepare(f, lds);
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, f, p.fr.bw * p.fr.bh, lds);
}

hipError_t launch_views_persistent(const RenderParams& p, int mode, int dtype, int grid,
                                   hipStream_t s) {
  if (p.fr.n_items <= 0 || grid <= 0 || !p.fr.work) return hipErrorInvalidValue;
  const bool wave = wave_views(p, mode, false);
  const size_t lds = wave ? wave_lds_bytes(p) : render_lds_bytes(p);
  const RenderFn f = wave ? select_wave_kernel(p, mode, dtype, 2)
                          : select_persistent(mode, dtype, is_lit(p), wide_views(p, dtype));
  hipError_t e = prepare(f, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(f, dim3(unsigned(grid)), dim3(unsigned(p.fr.bw * p.fr.bh)), lds, s, p);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(views_advance_kernel, dim3(1), dim3(64), 0, s, p.fr.work, p.fr.accumulate);
  return hipGetLastError();
}

hipError_t launch_views(const RenderParams& p, int mode, int dtype, bool count, hipStream_t s) {
  if (p.fr.n_items <= 0) return hipSuccess;
  const bool wave = wave_views(p, mode, count);
  const size_t lds = wave ? wave_lds_bytes(p) : render_lds_bytes(p);
  const RenderFn f = wave ? select_wave_kernel(p, mode, dtype, 1)
                          : select_views(mode, dtype, count, is_lit(p), wide_views(p, dtype));
  const hipError_t e = prepare(f, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(f, dim3(unsigned(p.fr.n_items)), dim3(unsigned(p.fr.bw * p.fr.bh)), lds, s, p);
  return hipGetLastError();
}

}  // namespace mdwt
