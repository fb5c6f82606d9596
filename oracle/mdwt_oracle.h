/*
 * CPU oracle for multi-density Woodcock tracking (MDWT).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product in
 * mdwtracker_amd/. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it, and only to check or to time it beside the GPU
 * path. Nothing in the product links, loads or falls back to it.
 *
 * It is a scalar C++ restatement of the reference's RAW-volume render path
 * (alpers-git/MDWTracker): deviceCode.cu (mainRG, volumeIntersection,
 * baseLineDTCH, altMultiMajDTCH, cummilativeDTCH, transferFunction,
 * generateRay, sampleVolumeTexture), dda.h (dda3), renderer.cu (range grid
 * and majorant kernels), renderer.cpp (Init, UpdateCamera, SetXF*, Resetdt,
 * CalculateMCGridDims), rawFile.cpp (value conversion) and camera.h. The
 * per-function file:line citations are in mdwt_oracle.cpp.
 *
 * Definitions that the reference leaves to hardware or to its un-vendored OWL
 * dependency are fixed here once and followed by the HIP kernels:
 *   - RNG: Philox4x32-7 keyed by (seed, pixel), counter (frame, sample, block,
 *     stream) instead of OWL's LCG<4> keyed by threadIdx;
 *   - log(1-u): a portable fp32 log (Cody-Waite reduction + atanh series);
 *   - texture filtering: fp32-weight trilinear / linear lerp a + f*(b-a)
 *     instead of 9-bit fixed-point hardware weights;
 *   - OWL vector helpers: normalize(v) = v * (1/sqrt(dot(v,v))), component-wise
 *     division, fminf/fmaxf for min/max; make_rgba: min(255,max(0,int(256*c))).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_renderer oracle_renderer;

/* dtype codes, shared with include/mdwt.h */
enum { ORACLE_U8 = 0, ORACLE_U16 = 1, ORACLE_F16 = 2, ORACLE_F32 = 3 };

oracle_renderer* oracle_create(void);
void oracle_destroy(oracle_renderer* o);
const char* oracle_last_error(oracle_renderer* o);

int oracle_push_channel(oracle_renderer* o, const void* vox, int dtype, int nx, int ny, int nz);
int oracle_set_bounds(oracle_renderer* o, const float remap[3]);
int oracle_init(oracle_renderer* o, int mode, const int mc[3], int auto_camera);
int oracle_set_tf(oracle_renderer* o, int ch, const float* rgba, int n_texels,
                  float opacity_scale, float xf_lo, float xf_hi);
int oracle_set_camera(oracle_renderer* o, const float from[3], const float at[3],
                      const float up[3], float fovy_deg);
int oracle_set_dt(oracle_renderer* o, float dt);
int oracle_reset_dt(oracle_renderer* o);
int oracle_set_frame(oracle_renderer* o, int w, int h, int spp, int accumulate, const float bg[3]);
int oracle_set_seed(oracle_renderer* o, uint32_t seed);
int oracle_set_lighting(oracle_renderer* o, float light_intensity, float ambient);
int oracle_set_light(oracle_renderer* o, const float dir[3], int shadows, int gradient);
int oracle_set_material(oracle_renderer* o, int ch, float diffuse, float specular,
                        float shininess, float gradient_step, const float spec[3]);
/* Reference-faithful tracking on (1) / off (0, default): libm logf for the
 * free-flight step (deviceCode.cu:563,805,913), the two divisions of the TF
 * remap (deviceCode.cu:36-37) and one sequential RNG stream per ray in the
 * reference's draw order (prd.rng(), deviceCode.cu:296-298,797,830) instead
 * of the GPU-shared fixed word layout. Statistically the same estimator: it
 * checks the default path against the reference's semantics (the default
 * path's images are what the GPU must match bit for bit). */
int oracle_set_faithful(oracle_renderer* o, int on);
/* Pixel jitter of y from word 1 of the jitter block instead of the one
 * scalar both axes take by default (deviceCode.cu:297-298); mirrors
 * mdwt_set_jitter_2d. */
int oracle_set_jitter_2d(oracle_renderer* o, int on);
/* Majorants bounding every opacity a TF lookup of the cell's value range
 * interpolates, instead of renderer.cu:159-203's (half a texel low at
 * worst); mirrors mdwt_set_conservative_majorants. Rebuilds them when the
 * setting changes and restarts the accumulation. */
int oracle_set_conservative_majorants(oracle_renderer* o, int on);
/* OpenMP threads for the range-grid build (result does not depend on it) */
int oracle_set_threads(oracle_renderer* o, int n);
/* Render n_frames frames; rows [y0, y1) only (y1 <= 0 means all rows).
 * nthreads > 1 uses OpenMP over rows (results do not depend on it). */
int oracle_render(oracle_renderer* o, int n_frames, int y0, int y1, int nthreads);
int oracle_reset_accumulation(oracle_renderer* o);
int oracle_read_accum(oracle_renderer* o, float* rgba, float* m2);
int oracle_read_fb(oracle_renderer* o, uint32_t* rgba8);
int oracle_read_ranges(oracle_renderer* o, float* out);
int oracle_read_majorants(oracle_renderer* o, float* out);
/* counters of the last oracle_render call: samples, cell visits, rejections, rays */
int oracle_read_counters(oracle_renderer* o, uint64_t out[4]);
/* info: mc[3], dims[3], n_channels, mode, frame_id, accum_id */
int oracle_get_info(oracle_renderer* o, int out[10]);
float oracle_get_dt(oracle_renderer* o);
/* camera: origin, lower_left_corner, horizontal, vertical (12 floats);
 * volume domain per channel (2*C floats) */
int oracle_get_camera(oracle_renderer* o, float out[12]);
int oracle_get_domains(oracle_renderer* o, float* out);

/* Per-ray DDA over the majorant grid for world-space rays (grid transform of
 * the closest-hit programs + dda3). Writes up to max_cells (cellID, t0, t1). */
int oracle_trace_cells(oracle_renderer* o, const float* org, const float* dir, int n,
                       int max_cells, int32_t* cells, float* t01, int32_t* counts);
/* Camera rays of (frame, sample) for every pixel: org[3], dir[3], jitter. */
int oracle_gen_rays(oracle_renderer* o, uint32_t frame, uint32_t sample, float* org,
                    float* dir);
/* Root-box clip (volumeIntersection): per ray t0, t1, hit flag. */
int oracle_clip_rays(oracle_renderer* o, const float* org, const float* dir, int n,
                     float* t01, int32_t* hit);

/* unit hooks */
/* Philox4x32 with `rounds` rounds (the renderer's streams use
 * oracle_philox_rounds(), 7; 10 checks the published known answers) */
void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], int rounds, uint32_t out[4]);
int oracle_philox_rounds(void);
float oracle_logf(float x);
float oracle_logf_poly(float x);
float oracle_expf(float x);
float oracle_powf(float x, float y);
float oracle_sample(oracle_renderer* o, int ch, const float p[3]);
void oracle_tf(oracle_renderer* o, int ch, float value, float out[4]);

#ifdef __cplusplus
}
#endif
