/* C ABI of the learned denoiser beside plain accumulation (libmdwt.so).
 *
 * No reference counterpart: the reference accumulates plain samples only
 * (deviceCode.cu:463-469). The denoiser is a 3-level U-Net of 3x3
 * convolutions (32 / 64 / 96 channels, 2x2 average pooling, nearest
 * upsampling with skip concatenation, ReLU) with BF16 weights and
 * activations and FP32 accumulation on the matrix cores. Its output is
 * kernel-predicting (Bako et al. 2017): per pixel, the softmax of 49 logits
 * weighs the input colour of the 7x7 neighbourhood (coordinates clamped to
 * the image), so every output is a convex combination of input colours; the
 * untrained network applies a Gaussian of 1.5 pixels. Input per pixel,
 * 7 + channels floats: r, g, b (the accumulated image), first-hit depth,
 * transmittance (1 - hit fraction), the hit fraction of each channel
 * (mdwt_set_aov / mdwt_read_aov), the standard error of the pixel mean (from
 * the second moment of mdwt_read_accum) and log2 of the accumulated frame
 * count, the last two in the scaling the caller trained with. Images are
 * row-major, pixel-interleaved float arrays.
 * Width and height must be multiples of 4 (two pooling levels). One HIP
 * stream per handle; not thread-safe. Status codes are those of mdwt.h. */
#ifndef MDWT_DENOISE_H
#define MDWT_DENOISE_H

#include <stdint.h>

#include "mdwt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdwt_denoiser mdwt_denoiser;

/* A denoiser for width x height images of a volume with `channels` channels
 * (1..12) on HIP device hip_device; weights initialised from seed (He
 * initialisation; output layer near zero with the Gaussian's log-weights as
 * bias). width and height: multiples of 4 up to 8192 (MDWT_ERR_ARG), at most
 * 2^24 pixels in all (MDWT_ERR_LIMIT; e.g. 4096 x 4096). */
int mdwt_dn_create(mdwt_denoiser** out, int hip_device, int width, int height, int channels,
                   uint32_t seed);
int mdwt_dn_destroy(mdwt_denoiser* d);
const char* mdwt_dn_last_error(const mdwt_denoiser* d);
/* number of trainable parameters (weights [9][cout][cin] then bias, per layer) */
int mdwt_dn_num_params(const mdwt_denoiser* d, int* n);
int mdwt_dn_get_params(mdwt_denoiser* d, float* out);
int mdwt_dn_set_params(mdwt_denoiser* d, const float* in);
/* Denoise one image: input [h][w][7 + channels] -> output [h][w][3]; ms = GPU time of
 * the network (uploads and downloads excluded). */
int mdwt_dn_infer(mdwt_denoiser* d, const float* input, float* output, float* ms);
/* Inference kernel sequence: on (default) fuses the upsample-and-concatenate
 * inputs, the pooling and the output filter into the convolutions; off runs
 * them as separate kernels (the training forward). Both give identical
 * outputs; the switch exists for testing and profiling. */
int mdwt_dn_set_fused(mdwt_denoiser* d, int on);
/* Gradient of the mean squared error against target [h][w][3] for one
 * image, in the parameter order of mdwt_dn_get_params; loss = that error. */
int mdwt_dn_gradient(mdwt_denoiser* d, const float* input, const float* target, float* grad,
                     float* loss);
/* One Adam step (beta 0.9 / 0.999) on count (input, target) pairs stored
 * back to back. The gradient is that of sum_i weights[i] * MSE_i / count
 * (weights may be NULL: all 1; each finite and >= 0); loss = the unweighted
 * mean squared error of the pairs before the step. */
int mdwt_dn_train_step(mdwt_denoiser* d, const float* inputs, const float* targets,
                       const float* weights, int count, float lr, float* loss);

#ifdef __cplusplus
}
#endif
#endif /* MDWT_DENOISE_H */
