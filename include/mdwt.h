/*
 * mdwt.h — C ABI of the MI355X multi-density Woodcock tracking (MDWT) library
 * (libmdwt.so, built from mdwtracker_amd/csrc).
 *
 * Drop-in boundary for the RAW-volume render path of alpers-git/MDWTracker.
 * Each entry point replaces one member of dtracker::Renderer (renderer.h:52-222)
 * or one step of its OWL launch; the reference interface it replaces is cited
 * beside it. Conventions:
 *   - every function returns an int status (MDWT_OK = 0); nothing throws or
 *     exits across the ABI; mdwt_last_error() describes the last failure;
 *   - plain pointers and sizes only; all device memory is owned by the handle;
 *     caller buffers are host memory and are copied, never retained;
 *   - one handle = one HIP device and one HIP stream; every call is
 *     synchronous with respect to the host; a handle is not thread-safe
 *     (as the reference Renderer is not).
 */
#ifndef MDWT_H
#define MDWT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDWT_MAX_CHANNELS 12 /* deviceCode.h:13 MAX_CHANNELS */

enum mdwt_status {
  MDWT_OK = 0,
  MDWT_ERR_ARG = 1,   /* invalid argument (bad channel id, size, dtype, NaN ...) */
  MDWT_ERR_STATE = 2, /* call not valid in the current state (e.g. render before init) */
  MDWT_ERR_HIP = 3,   /* HIP runtime error (no device, launch failure ...) */
  MDWT_ERR_LIMIT = 4, /* a documented size limit was exceeded */
  MDWT_ERR_OOM = 5    /* device allocation failed */
};

/* voxel types; value conversion follows rawFile.cpp:100-123 (u8 -> v*255,
 * u16 -> v*65535, f16/f32 as is) */
enum mdwt_dtype { MDWT_U8 = 0, MDWT_U16 = 1, MDWT_F16 = 2, MDWT_F32 = 3 };

/* -m values (renderer.h:24-36 Mode): 0 N-DDA, 1 multi-majorant (one
 * exponential clock per channel), 2 syncDDA, 3 cumulative majorant; 4/5
 * Woodcock tracking of the max-opacity / mixed blend of the channels on the
 * summed majorant; 6/7/8 ray marchers over the density / max / mix blend;
 * 9 per-channel ray marcher composited by opacity */
enum mdwt_mode {
  MDWT_MODE_NDDA = 0,
  MDWT_MODE_MULTI = 1,
  MDWT_MODE_SYNCDDA = 2,
  MDWT_MODE_CUMULATIVE = 3,
  MDWT_MODE_MAX = 4,
  MDWT_MODE_MIX = 5,
  MDWT_MODE_MARCHER_MULTI = 6,
  MDWT_MODE_MARCHER_MAX = 7,
  MDWT_MODE_MARCHER_MIX = 8,
  MDWT_MODE_MARCHER_COMP = 9
};

typedef struct mdwt_renderer mdwt_renderer;

/* Renderer::Renderer (renderer.cpp:316) — binds HIP device `hip_device`.
 * Touches no GPU until mdwt_init. */
int mdwt_create(mdwt_renderer** out, int hip_device);
/* Renderer::Terminate (renderer.cpp:1010-1014) + destructor */
int mdwt_destroy(mdwt_renderer* h);
/* text of the last error on this handle ("" if none); h may be NULL for
 * failures of mdwt_create */
const char* mdwt_last_error(const mdwt_renderer* h);

/* Renderer::PushMesh(shared_ptr<RawR>) (renderer.cpp:1031-1044) with the RawR
 * load (rawFile.cpp:31-128): copies nx*ny*nz voxels (x fastest). At most
 * MDWT_MAX_CHANNELS channels (the reference's off-by-one 13th is rejected);
 * all channels must share dims and dtype. */
int mdwt_push_channel(mdwt_renderer* h, const void* voxels, int dtype, int nx, int ny, int nz);
/* RawR::reshapeBounds(remap) (rawFile.cpp:228-237, the -cb flag,
 * viewer.cpp:305-317): world box [0, remap] instead of [0,1]^3. Before init. */
int mdwt_set_bounds(mdwt_renderer* h, const float remap[3]);
/* Renderer::Init(mode, autoSetCamera) (renderer.cpp:324-942, RAW 638-903):
 * uploads the volume, builds the value-range grid (renderer.cu:1116-1179).
 * mc = macrocells per side (renderer.h:146 macrocellDims, -mc); 0 entries
 * take CalculateMCGridDims(4) (renderer.cpp:1317-1361). */
int mdwt_init(mdwt_renderer* h, int mode, const int mc[3], int auto_camera);
/* Switch -m without re-uploading (no reference counterpart: the reference
 * re-runs Init); rebuilds the majorant grid. */
int mdwt_set_mode(mdwt_renderer* h, int mode);

/* SetXFColormap + SetXFOpacityScale + SetXFRange in one call
 * (renderer.cpp:1094-1193): rgba = n_texels float4 (alpha is cubed on upload,
 * renderer.cpp:1098-1101); rebuilds majorants (renderer.cu:206-274) once all
 * channels have a transfer function. */
int mdwt_set_tf(mdwt_renderer* h, int ch, const float* rgba, int n_texels, float opacity_scale,
                float xf_lo, float xf_hi);
/* the three reference setters individually */
int mdwt_set_xf_colormap(mdwt_renderer* h, int ch, const float* rgba, int n_texels);
int mdwt_set_xf_opacity_scale(mdwt_renderer* h, int ch, float opacity_scale);
int mdwt_set_xf_range(mdwt_renderer* h, int ch, float xf_lo, float xf_hi);

/* camera.setOrientation(from, at, up, fovy) + UpdateCamera()
 * (camera.h:83-110, renderer.cpp:1058-1092) */
int mdwt_set_camera(mdwt_renderer* h, const float from[3], const float at[3], const float up[3],
                    float fovy_deg);
/* Setdt (renderer.cpp:1460-1466) / Resetdt (renderer.cpp:1444-1453) */
int mdwt_set_dt(mdwt_renderer* h, float dt);
int mdwt_reset_dt(mdwt_renderer* h);
/* fbSize/Resize (renderer.cpp:1046-1056), spp, enableAccumulation, bgColor;
 * at most 65536 pixels per side and 2^30 pixels per frame (MDWT_ERR_LIMIT) */
int mdwt_set_frame(mdwt_renderer* h, int width, int height, int spp, int accumulate,
                   const float bg[3]);
/* SetLightIntensity / SetAmbient (renderer.cpp:1476-1486) */
int mdwt_set_lighting(mdwt_renderer* h, float light_intensity, float ambient);
/* SetLightDirection (renderer.cpp:1468-1474) and the enableShadows /
 * enableGradientShading fields (renderer.h:121-124; the -l flag,
 * viewer.cpp:271-281). Shadows trace one ray per hit towards -light_dir
 * through the same method (deviceCode.cu:383-403; the ray marchers march
 * their own, 1117-1154); gradient shading is Blinn-Phong on the TF-opacity
 * gradient of the hit channel (deviceCode.cu:326-381). Resets accumulation. */
int mdwt_set_light(mdwt_renderer* h, const float light_dir[3], int shadows, int gradient_shading);
/* SetMaterialDiffuse / Specular / Shininess / GradientStep / SpecularColor
 * (renderer.cpp:1195-1248) for one channel (0..MDWT_MAX_CHANNELS-1; the
 * reference ignores unknown ids, this returns MDWT_ERR_ARG). Resets
 * accumulation. */
int mdwt_set_material(mdwt_renderer* h, int channel, float diffuse, float specular,
                      float shininess, float gradient_step, const float specular_color[3]);
/* ApplyMaterialPreset's table (renderer.cpp:1250-1301): out = {diffuse,
 * specular, shininess, gradient_step, specular r, g, b} for "Plastic",
 * "Iron", "Chrome", "Emerald"; any other name gives the default material. */
int mdwt_material_preset(const char* name, float out[7]);
/* seed of the counter-based RNG (replaces the LCG seeding, deviceCode.cu:296-297) */
int mdwt_set_seed(mdwt_renderer* h, uint32_t seed);

/* Render() (renderer.cpp:944-973) n_frames times, synchronously. ms_out
 * (may be NULL) receives the GPU time of the n_frames launches (HIP events). */
int mdwt_render(mdwt_renderer* h, int n_frames, float* ms_out);
/* ResetAccumulation (renderer.cpp:1488-1495) */
int mdwt_reset_accumulation(mdwt_renderer* h);
/* Second-moment buffers on (1) or off (0, the default): when on, every frame
 * also keeps the per-pixel running mean of the squared colour beside the
 * running mean (an addition to the reference, which keeps only the mean,
 * deviceCode.cu:463-469), for confidence intervals and the error estimate of
 * time-to-RMSE; it costs 32 more bytes of HBM traffic per pixel and frame.
 * Applies to the single view and to batched views; restarts the
 * accumulation. No reference counterpart. */
int mdwt_set_moments(mdwt_renderer* h, int on);
/* accumBuffer (float4 per pixel, device-resident in the reference) and the
 * per-pixel running mean of squares (only with mdwt_set_moments on, else
 * MDWT_ERR_STATE when m2 is not NULL); either pointer may be NULL. */
int mdwt_read_accum(mdwt_renderer* h, float* rgba, float* m2);
/* frameBuffer (uint32 RGBA8, renderer.cpp:908; make_rgba of the accumulated
 * colour, deviceCode.cu:463-469) */
int mdwt_read_fb(mdwt_renderer* h, uint32_t* rgba8);
/* mdwt_read_fb without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under graph capture,
 * tight serving loops). The RGBA8 conversion is staged in a static device
 * array of 2^20 pixels per GPU (4 MiB, held until the process ends): frames
 * of more pixels return MDWT_ERR_LIMIT. Calls that overlap on one GPU share
 * that array and race. Same output as mdwt_read_fb. */
int mdwt_read_fb_noalloc(mdwt_renderer* h, uint32_t* rgba8);

/* ---- inspection hooks (tests, bench) ---- */
/* value-range grid: float2 (min, max) per (cell, channel), index cell*C + ch */
int mdwt_read_ranges(mdwt_renderer* h, float* out);
/* majorants: float per (cell, channel) for modes 0-2, float per cell (the
 * channel sum) for modes 3-9 */
int mdwt_read_majorants(mdwt_renderer* h, float* out);
/* out = {mc[3], dims[3], channels, mode, frame_id, accum_id} */
int mdwt_get_info(mdwt_renderer* h, int out[10]);
int mdwt_get_dt(mdwt_renderer* h, float* dt);
/* origin, lower_left_corner, horizontal, vertical (LaunchParams camera) */
int mdwt_get_camera(mdwt_renderer* h, float out[12]);
/* per-channel value domain (renderer.cpp:743-745), 2 floats per channel */
int mdwt_get_domains(mdwt_renderer* h, float* out);
/* DDA over the majorant grid for n world rays (grid transform of the
 * closest-hit programs + dda.h:82-203): up to max_cells (cell id, t0, t1) per
 * ray, counts[r] = number of cells the DDA emits; slots past a ray's count
 * hold cell -1 and t0 = t1 = 0. */
int mdwt_trace_cells(mdwt_renderer* h, const float* org, const float* dir, int n, int max_cells,
                     int32_t* cells, float* t01, int32_t* counts);
/* camera rays (origin, direction) of every pixel for (frame, sample)
 * (deviceCode.cu:293-300 jitter + generateRay 61-76) */
int mdwt_gen_rays(mdwt_renderer* h, uint32_t frame, uint32_t sample, float* org, float* dir);
/* root-box clip (volumeIntersection, deviceCode.cu:1676-1726) of n rays */
int mdwt_clip_rays(mdwt_renderer* h, const float* org, const float* dir, int n, float* t01,
                   int32_t* hit);
/* Counting pass: renders the next frame's work without writing the image or
 * advancing the frame; out = {samples, cell visits, rejections, rays}
 * (prd.samples / prd.rejections, deviceCode.cu:590,611; cell visits added). */
int mdwt_count_frame(mdwt_renderer* h, uint64_t out[4]);
/* Average duration in ms of the last mdwt_render's frame kernel launches. */
int mdwt_last_kernel_ms(mdwt_renderer* h, float* ms);
/* Block order of the frame kernel: 0 = row-major, 1 = Hilbert curve over
 * tile x tile-pixel tiles (tile 8, 16 or 32; blocks row-major inside a tile),
 * 2 = cost-sorted (default): starts as the Hilbert order, records each wave's
 * wall-clock time and, at the next mdwt_render, issues the blocks heaviest
 * first (one-thread-per-pixel schedule only). Images do not depend on the
 * order. No reference counterpart (the reference sorts umesh elements along a
 * Hilbert curve, renderer.cu:1364-1411; hilbert.h). */
int mdwt_set_tile_order(mdwt_renderer* h, int order, int tile);
/* Workgroup of the thread-per-pixel frame kernel: 64 (8x8 pixels), 128 (16x8)
 * or 256 (16x16, default) threads, one 8x8-pixel quad per wave. Images do not
 * depend on it. No reference counterpart (OptiX picks the launch shape). */
int mdwt_set_block(mdwt_renderer* h, int threads);
/* Rebuilds the range and majorant grids for a new macrocell count per side
 * (the -mc value, viewer.cpp:204-224) on the resident volume, without a new
 * Init; resets accumulation (also of the batched views). */
int mdwt_set_macrocells(mdwt_renderer* h, const int mc[3]);

/* Auxiliary buffers for the denoiser (no reference counterpart): when on,
 * every frame also accumulates per pixel the running means of the first-hit
 * distance (0 on a miss), the transmittance (1 - hit fraction) and the hit
 * fraction of every channel. Switching restarts the accumulation. */
int mdwt_set_aov(mdwt_renderer* h, int on);
/* Accumulate the auxiliary buffers during the first `frames` frames of each
 * accumulation only (0, the default: every frame); later frames render as
 * with the buffers off and the buffers keep the means over those frames.
 * Changing it restarts the accumulation. */
int mdwt_set_aov_frames(mdwt_renderer* h, int frames);
/* width * height * (2 + channels) floats per pixel in that order */
int mdwt_read_aov(mdwt_renderer* h, float* out);

/* ---- batched views (no reference counterpart: the reference renders one
 * camera per launch). Many views of uneven size and spp, each with its own
 * accumulators, rendered frame by frame: with one frame-kernel launch per
 * view, the views spread round-robin over four HIP streams so that small
 * views run side by side (launch form 0, the default); with one launch per
 * frame whose grid holds every view's pixel blocks (launch form 1); or with
 * one persistent launch per frame (a resident grid, each workgroup taking
 * every G-th pixel block of the batch) that is captured once in a HIP graph
 * together with the kernel advancing the frame ids, and replayed every frame
 * (launch form 2). Calls are ordered with the handle's other work and return
 * when the frames are done. A view's image equals a single-view render with
 * the same camera, size, spp and seed (the RNG is keyed by pixel), so views
 * never depend on the batch they are in. */
typedef struct mdwt_view {
  float from[3], at[3], up[3]; /* camera.h setOrientation arguments */
  float fovy;                  /* degrees */
  int width, height, spp;
  uint32_t seed;
} mdwt_view;
/* Replaces the batch (1..4096 views); each view gets its own accumulators. */
int mdwt_batch_set_views(mdwt_renderer* h, const mdwt_view* views, int n);
/* Renders n_frames frames of every view (ms = total GPU time). */
int mdwt_batch_render(mdwt_renderer* h, int n_frames, float* ms);
/* Counting pass over the batch: {samples, cell visits, rejections, rays}. */
int mdwt_batch_count(mdwt_renderer* h, uint64_t out[4]);
/* Restarts the views' accumulation (frame ids keep counting). */
int mdwt_batch_reset(mdwt_renderer* h);
/* Launch form of the batch: 0 = one launch per view over four streams (each
 * view finishes as soon as its own blocks do: low per-view latency), 1 = one
 * launch for all views per frame (every view finishes with the frame: highest
 * throughput when views are small), 2 = one persistent launch per frame
 * replayed from a captured HIP graph (captured again when the launch's
 * parameters change). Images and counters do not depend on it. */
int mdwt_batch_set_launch(mdwt_renderer* h, int launch);
/* width * height float4 running means (and second moments) of one view. */
int mdwt_batch_read_view(mdwt_renderer* h, int view, float* rgba, float* m2);
/* Per view: GPU time in microseconds of the view in the last batched frame.
 * Launch form 0: HIP events around its launch on its stream (so it includes
 * time shared with views on the other streams); form 1: from the launch's
 * first workgroup start to the view's last workgroup end (device wall
 * clock). MDWT_ERR_STATE before the first batched frame. */
int mdwt_batch_view_times(mdwt_renderer* h, float* us);
/* Autotuner: times `frames` frames (after one warm-up frame) for every
 * combination of macrocell size s (voxels per macrocell side; grid = dims /
 * s), workgroup threads (64/128/256) and Hilbert tile (8/16/32), on the batch
 * if one is set, else on the single view, with the tile order running (no
 * cost sort). table gets 6 floats per combination: s, threads, tile,
 * ms/frame, samples/s, cell visits/s. Leaves the fastest combination set
 * (best = {s, threads, tile}) and accumulation reset. */
int mdwt_autotune(mdwt_renderer* h, const int* mc_sizes, int n_mc, const int* threads,
                  int n_threads, const int* tiles, int n_tiles, int frames, float* table,
                  int best[3]);
/* Frame-kernel schedule: 0 = one thread per pixel, nested cell/channel/step
 * loops; 1 = wave-coherent tracking (a wave traces 64 rays and compacts their
 * free-flight steps and null-collision tests into full 64-lane batches), the
 * default. Schedule 1 covers the tracking modes -m 0 to -m 5 without
 * lighting; the ray marchers (-m 6-9), lighting and the counting pass run
 * schedule 0. Images do not depend on it.
 * Returns MDWT_ERR_ARG for other values. No reference counterpart (OptiX
 * schedules the reference's launch). */
int mdwt_set_schedule(mdwt_renderer* h, int schedule);
/* Pixel jitter: 0 (the default) jitters x and y by one uniform, as the
 * reference's raygen does (deviceCode.cu:297-298: one random() added to both
 * coordinates); 1 draws y's jitter from the next word of the same Philox
 * block, a box filter over the pixel. The estimator then converges to the
 * pixel-area average instead of the reference's diagonal average. No
 * reference counterpart (an option beside the reference's behaviour). */
int mdwt_set_jitter_2d(mdwt_renderer* h, int on);
/* Majorants: 0 (the default) as the reference computes them (renderer.cu:
 * 159-203), which can fall below the opacities a TF lookup interpolates by up
 * to half a texel of lerp; 1 the largest opacity of every texel a lookup of
 * the cell's value range reads (the range widened by 2^-12 (1 + |r|), the
 * maximum by 1 + 2^-20): an upper bound, so Woodcock tracking is unbiased.
 * Where the reference's interval runs a texel past the lookups it is also
 * lower than the reference's majorant.
 * Rebuilds the majorant grid and restarts the accumulation when the setting
 * changes. No reference counterpart (an option beside its behaviour). */
int mdwt_set_conservative_majorants(mdwt_renderer* h, int on);
/* Device copy of `bytes` bytes, `iters` times (one read and one write per
 * byte); GB/s of (read + write) bytes. No reference counterpart. */
int mdwt_stream_copy(mdwt_renderer* h, uint64_t bytes, int iters, float* gbps);
/* Read-only stream of `bytes` bytes (a reduction whose only writes are one
 * float per workgroup), `iters` times; GB/s of bytes read. With at least
 * 4 GiB the 256 MB last-level cache cannot serve it, so it measures the HBM
 * bandwidth a read-dominated kernel reaches: the denominator of the frame
 * kernel's roofline. No reference counterpart. */
int mdwt_stream_read(mdwt_renderer* h, uint64_t bytes, int iters, float* gbps);
/* 2-D Hilbert curve of side 2^bits: index d -> (x, y) (host function, no GPU) */
int mdwt_hilbert_d2xy(int bits, uint32_t d, uint32_t* x, uint32_t* y);

#ifdef __cplusplus
}
#endif
#endif /* MDWT_H */
