// Host-side launchers of the MDWT kernels (defined in the .hip files, called
// by the C ABI in mdwt_api.cpp). Each returns the launch status.
#pragma once
#include <hip/hip_runtime.h>

#include "mdwt_types.hpp"

namespace mdwt {

// host stub of a frame kernel
using RenderFn = void (*)(RenderParams);
// wave-coherent frame kernels (kernels_wave.hip, schedule 1): -m 0 to -m 5 on
// at most twelve channels, no lighting, not the counting pass; form 0 one view,
// 1 batched views in one launch, 2 batched views in a persistent launch
bool wave_schedule_applies(const RenderParams& p, int mode, bool count, bool lit);
RenderFn select_wave_kernel(const RenderParams& p, int mode, int dtype, int form);
// dynamic LDS of a wave-coherent launch (transfer-function opacities)
size_t wave_lds_bytes(const RenderParams& p);

hipError_t launch_ranges(const RangeParams& p, int dtype, hipStream_t s);
hipError_t launch_majorants(const MajorantParams& p, hipStream_t s);
// v.bricks (brick_stride bytes per channel) = v.vol in 4 x 4 x 4 bricks
hipError_t launch_bricks(const VolumeParams& v, int dtype, hipStream_t s);
hipError_t launch_render(const RenderParams& p, int mode, int dtype, bool count, hipStream_t s);
// batched views, launch form 2: the persistent frame kernel over p.fr.items
// (grid workgroups, workgroup g takes items g, g + grid, ...) at the ids in
// p.fr.work[1..2], then the kernel that advances those ids; the pair is what a
// graph captures for one frame
hipError_t launch_views_persistent(const RenderParams& p, int mode, int dtype, int grid,
                                   hipStream_t s);
// resident workgroups per CU of that persistent kernel
hipError_t views_persistent_blocks_per_cu(const RenderParams& p, int mode, int dtype, int* out);
// batched views in one launch: one workgroup per item of p.fr.items (view,
// pixel block), frame and accumulation ids from p.fr.frame_id / accum_id
hipError_t launch_views(const RenderParams& p, int mode, int dtype, bool count, hipStream_t s);
// dynamic LDS of a frame launch (transfer-function texels)
size_t render_lds_bytes(const RenderParams& p);
hipError_t launch_gen_rays(const CameraParams& cam, int W, int H, uint32_t seed, uint32_t frame,
                           uint32_t sample, int jitter2d, float* org, float* dir, hipStream_t s);
hipError_t launch_clip(const VolumeParams& vol, const float* org, const float* dir, int n,
                       float* t01, int* hit, hipStream_t s);
hipError_t launch_trace_cells(const VolumeParams& vol, const int mc[3], const float* org,
                              const float* dir, int n, int max_cells, int* cells, float* t01,
                              int* counts, hipStream_t s);
hipError_t launch_rgba8(const void* accum, uint32_t* out, int n, hipStream_t s);
// Allocation-free frame-buffer read: a file-scope device array of
// kFbScratchPixels RGBA8 pixels per GPU. fb_scratch_setup (once per device,
// with that device current; called by mdwt_init) takes its address with
// hipGetSymbolAddress; fb_scratch returns it (nullptr before the set-up).
constexpr size_t kFbScratchPixels = size_t(1) << 20;
hipError_t fb_scratch_setup(int device);
uint32_t* fb_scratch(int device);
hipError_t launch_copy(const void* src, void* dst, size_t n16, hipStream_t s);
// read-only reduction over n16 float4 (bandwidth probe): kReadGrid workgroups,
// one partial sum each into sums
constexpr unsigned kReadGrid = 4096;
hipError_t launch_read(const void* src, size_t n16, float* sums, hipStream_t s);

}  // namespace mdwt
