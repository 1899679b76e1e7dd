/*
 * rtrace_denoise.h — learned shadow denoiser on gfx950 (extension; no reference
 * counterpart). The reference estimates each area light with 16 shadow rays per hit
 * (src/lighting.rs:145-217). Here a frame is rendered with few light samples plus the
 * per-pixel features of rt_render_aux, a small U-Net predicts the many-sample colour, and
 * tiles whose visibility-variance estimate is high are re-rendered with many samples.
 *
 * Network (all 3x3 convolutions, zero padding, bf16 weights and activations, FP32
 * accumulation on MFMA; frames are padded to a multiple of 8):
 *   layer  name    input                          cin  cout  level  ReLU
 *   0      enc0a   features (13 used of 32)        32    32    1/1   yes
 *   1      enc0b   enc0a                           32    32    1/1   yes
 *   2      enc1a   avgpool2(enc0b)                 32    64    1/2   yes
 *   3      enc1b   enc1a                           64    64    1/2   yes
 *   4      enc2a   avgpool2(enc1b)                 64   128    1/4   yes
 *   5      enc2b   enc2a                          128   128    1/4   yes
 *   6      bota    avgpool2(enc2b)                128   128    1/8   yes
 *   7      botb    bota                           128   128    1/8   yes
 *   8      dec2a   [up2(botb), enc2b]             256   128    1/4   yes
 *   9      dec2b   dec2a                          128   128    1/4   yes
 *   10     dec1a   [up2(dec2b), enc1b]            192    64    1/2   yes
 *   11     dec1b   dec1a                           64    64    1/2   yes
 *   12     dec0a   [up2(dec1b), enc0b]             96    32    1/1   yes
 *   13     dec0b   dec0a                           32    32    1/1   yes
 *   14     out     dec0b                           32     3    1/1   no
 * The output is added to the noisy radiance (features 0-2). up2 = nearest upsampling.
 * Input normalisation: radiance clamped to [0, 4], depth -> log1p(depth)/4, visibility
 * variance -> 4*sqrt(variance); the rest as given.
 *
 * Parameters (f32, rt_denoiser_param_count() of them): per layer in order, the weights
 * [9 taps (row-major 3x3)][cin][coutp] then the bias [coutp], coutp = cout rounded up to 32.
 * Training: mean squared error of (noisy + output) against a target radiance, Adam.
 */
#ifndef RTRACE_DENOISE_H
#define RTRACE_DENOISE_H

#include <stdint.h>

#include "rtrace_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rt_denoiser rt_denoiser;

enum { RT_DN_LAYERS = 15 };

/* Create with He-normal weights from `seed` (output layer scaled by 0.1, biases 0). */
int rt_denoiser_create(rt_ctx* ctx, uint64_t seed, rt_denoiser** out);
void rt_denoiser_destroy(rt_denoiser* dn);
int64_t rt_denoiser_param_count(void);
/* Layer geometry: input channels, padded output channels, offsets of weights and bias. */
int rt_denoiser_layer_info(int32_t layer, int32_t* cin, int32_t* coutp, int64_t* w_off, int64_t* b_off);
int rt_denoiser_get_params(rt_denoiser* dn, float* host_out, int64_t n);
/* Replaces the parameters and resets the optimizer state. */
int rt_denoiser_set_params(rt_denoiser* dn, const float* host_in, int64_t n);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on a batch. Device inputs:
 * features [batch][h][w][RT_FEAT_COUNT] f32, target [batch][h][w][3] f32. */
int rt_denoiser_train_step(rt_denoiser* dn, const float* d_features, const float* d_target, int32_t batch,
                           int32_t h, int32_t w, float lr, float* loss_out);
/* Loss and parameter gradients of a batch without an update (host_grads: param_count, nullable). */
int rt_denoiser_gradients(rt_denoiser* dn, const float* d_features, const float* d_target, int32_t batch,
                          int32_t h, int32_t w, float* host_grads, float* loss_out);
/* Output activation of `layer` (0..13 bf16 post-ReLU, 14 = f32 network output) of the last
 * forward pass, as f32 [batch][hp][wp][cout] on the host; n = capacity in floats. */
int rt_denoiser_activation(rt_denoiser* dn, int32_t layer, float* host_out, int64_t n);

/* Inference on one frame: d_features [h][w][RT_FEAT_COUNT] -> noisy + output as f32
 * [h][w][3] (nullable) and/or quantised RGB u8 (nullable). ms_out (nullable) = device time. */
int rt_denoise(rt_denoiser* dn, const float* d_features, int32_t h, int32_t w, float* d_rgb_f32, uint8_t* d_rgb_u8,
               float* ms_out);

typedef struct {
    double ms_render;    /* few-sample render with features */
    double ms_denoise;   /* tile decision + network + compose */
    double ms_fallback;  /* many-sample re-render of the flagged tiles, on a second stream beside the network */
    double ms_total;     /* the whole frame, first launch to last (the fallback overlaps the network) */
    int32_t tiles, tiles_fallback;
    uint64_t rays;       /* all rays traced, both renders */
} rt_denoise_stats;

/* Denoised frame into d_rgb_out (device, w*h*3 u8): render with opts (its light_samples is
 * the low count; 0 = each light's own), then, side by side, denoise and re-render with
 * `fallback_samples` light samples the tiles (tile x tile pixels) whose mean visibility
 * variance exceeds `threshold` (a negative threshold re-renders every tile, one above 1 none):
 * the network's output goes to the other tiles. Needs an opts without the outline pass. */
int rt_render_denoised(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, rt_denoiser* dn,
                       int32_t fallback_samples, float threshold, int32_t tile, void* d_rgb_out,
                       rt_denoise_stats* stats);

#ifdef __cplusplus
}
#endif

#endif
