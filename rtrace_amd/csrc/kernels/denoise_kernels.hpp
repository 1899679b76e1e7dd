// denoise_kernels.hpp — argument blocks and launchers of the denoiser kernels (denoise.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rt {
namespace dn {

constexpr int kFeatIn = 13;  // rt_render_aux feature channels

// Input of a 3x3 conv: concat(src0, src1) along channels. src0 may be stored at half
// resolution (nearest 2x upsampling on read). Channel counts are multiples of 32; bf16 NHWC.
struct ConvIn {
    const uint16_t* src0;
    int c0;
    int up0;
    const uint16_t* src1;  // same resolution as the output, or null with c1 = 0
    int c1;
};

struct ConvArgs {
    ConvIn in;
    const uint16_t* wt;  // bf16 [9][coutp][cin]
    const float* bias;   // [coutp] or null
    int coutp;           // output channels computed (multiple of 32)
    int cout;            // output channels written (<= coutp)
    int out_pitch;       // channels per output pixel
    uint16_t* out_bf16;  // bf16 NHWC output (+bias, optional ReLU) ...
    float* out_f32;      // ... or f32 NHWC output (+bias), when set
    int relu;
    int B, H, W;  // output resolution
    uint16_t* pool_out;  // optional: bf16 [B][H/2][W/2][out_pitch], the 2x2 average of the bf16 outputs
                         // of each 2x2 block; H and W even, bf16 output only
};

struct WgradArgs {
    ConvIn in;
    const uint16_t* dz;  // bf16 [B][H][W][coutp]
    int coutp;
    int B, H, W;
    float* part;         // [splits][9][cin][coutp]
};

hipError_t launch_conv(const ConvArgs& a, hipStream_t st);
hipError_t launch_wgrad(const WgradArgs& a, int splits, hipStream_t st);
hipError_t launch_bias_grad(const uint16_t* dz, long long n_px, int c, float* part, int splits, hipStream_t st);
hipError_t launch_reduce(const float* part, int splits, long long n, float* out, bool accumulate, hipStream_t st);
hipError_t launch_prep(const float* feat, int B, int H, int W, int Hp, int Wp, uint16_t* x0, hipStream_t st);
hipError_t launch_pool_bwd(const float* gh, int B, int H, int W, int C, float* gf, hipStream_t st);
hipError_t launch_split(const float* dx, int B, int H, int W, int c0, int c1, float* g_low, float* g_skip,
                        hipStream_t st);
hipError_t launch_relu_bwd(const float* g, const uint16_t* a, long long n, uint16_t* dz, hipStream_t st);
hipError_t launch_loss(const float* y, const float* feat, const float* target, int B, int H, int W, int Hp, int Wp,
                       uint16_t* dz, float* part, int blocks, hipStream_t st);
// skip (nullable): T x T tile flags whose 8-bit output is left to the fallback render
hipError_t launch_compose(const float* y, const float* feat, int H, int W, int Wp, float* out_f32, uint8_t* out_u8,
                          const uint8_t* skip, int T, hipStream_t st);
hipError_t launch_tile_mask(const float* feat, int H, int W, int T, float threshold, uint8_t* mask, int* n_flagged,
                            hipStream_t st);
hipError_t launch_adam(float* w, const float* grad, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, hipStream_t st);
hipError_t launch_pack(const float* w, int cin, int coutp, uint16_t* wf, uint16_t* wb, hipStream_t st);

}  // namespace dn
}  // namespace rt
