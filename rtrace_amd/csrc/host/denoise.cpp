// denoise.cpp — host side of the shadow denoiser: the U-Net's layer table, its buffers in
// HBM, the forward / backward / Adam launch sequences, and the denoised-render pipeline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../kernels/denoise_kernels.hpp"
#include "context.hpp"
#include "frame.hpp"
#include "../../../include/rtrace_denoise.h"
#include "scene_host.hpp"

namespace rt {
namespace {

#define DN_TRY(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw Error(e_ == hipErrorOutOfMemory ? RT_E_OOM : RT_E_DEVICE,                       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                       \
    } while (0)

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void reserve(size_t n) {
        if (p && n <= bytes) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        DN_TRY(hipMalloc(&p, n ? n : 16));
        bytes = n ? n : 16;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Layer table (include/rtrace_denoise.h). in0 / in1: activation index feeding the layer
// (-1 = network input, -2/-3/-4 = pool of activation 1/3/5); up0: in0 is upsampled.
struct LayerDef {
    int level, in0, up0, in1, cin, cout, relu;
};
constexpr int kPool1 = -2, kPool3 = -3, kPool5 = -4, kInput = -1;
const LayerDef kLayers[RT_DN_LAYERS] = {
    {0, kInput, 0, -9, 32, 32, 1},  // enc0a
    {0, 0, 0, -9, 32, 32, 1},       // enc0b
    {1, kPool1, 0, -9, 32, 64, 1},  // enc1a
    {1, 2, 0, -9, 64, 64, 1},       // enc1b
    {2, kPool3, 0, -9, 64, 128, 1}, // enc2a
    {2, 4, 0, -9, 128, 128, 1},     // enc2b
    {3, kPool5, 0, -9, 128, 128, 1},// bota
    {3, 6, 0, -9, 128, 128, 1},     // botb
    {2, 7, 1, 5, 256, 128, 1},      // dec2a
    {2, 8, 0, -9, 128, 128, 1},     // dec2b
    {1, 9, 1, 3, 192, 64, 1},       // dec1a
    {1, 10, 0, -9, 64, 64, 1},      // dec1b
    {0, 11, 1, 1, 96, 32, 1},       // dec0a
    {0, 12, 0, -9, 32, 32, 1},      // dec0b
    {0, 13, 0, -9, 32, 3, 0},       // out
};
constexpr int kOut = RT_DN_LAYERS - 1;

int coutp_of(int l) { return (kLayers[l].cout + 31) / 32 * 32; }
int c0_of(int l) {
    const LayerDef& d = kLayers[l];
    if (d.in1 == -9) return d.cin;
    return kLayers[d.in0].cout;
}

struct Offsets {
    long long w[RT_DN_LAYERS], b[RT_DN_LAYERS], total;
    Offsets() {
        long long o = 0;
        for (int l = 0; l < RT_DN_LAYERS; ++l) {
            w[l] = o;
            o += 9ll * kLayers[l].cin * coutp_of(l);
            b[l] = o;
            o += coutp_of(l);
        }
        total = o;
    }
};
const Offsets& offsets() {
    static Offsets o;
    return o;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace
}  // namespace rt

struct rt_denoiser {
    rt_ctx* ctx = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    long long n_params = 0;
    rt::Buf params, grads, adam_m, adam_v;
    rt::Buf wf[RT_DN_LAYERS], wb[RT_DN_LAYERS];
    bool packed = false;
    long long step = 0;
    // activations for the current shape
    int B = 0, Hp = 0, Wp = 0;
    rt::Buf x0, act[RT_DN_LAYERS], pool1, pool3, pool5;
    // training buffers
    rt::Buf g[RT_DN_LAYERS - 1], gpool, dz, dx, part, bpart, lpart;
    // denoised frames: features, tile mask and flag count, the fallback render's frame buffers
    // and the stream it runs on beside the network (default priority, like every stream here)
    rt::Buf feat, mask, flagged, fb_arena;
    hipStream_t st_fb = nullptr;
    hipEvent_t ev[7] = {};
};

namespace rt {
namespace {

int level_h(const rt_denoiser* d, int lv) { return d->Hp >> lv; }
int level_w(const rt_denoiser* d, int lv) { return d->Wp >> lv; }
long long level_px(const rt_denoiser* d, int lv) { return (long long)d->B * level_h(d, lv) * level_w(d, lv); }

void pack_weights(rt_denoiser* d, bool with_backward) {
    const Offsets& o = offsets();
    for (int l = 0; l < RT_DN_LAYERS; ++l) {
        const int cin = kLayers[l].cin, cp = coutp_of(l);
        d->wf[l].reserve(9ull * cin * cp * 2);
        if (with_backward) d->wb[l].reserve(9ull * cin * cp * 2);
        DN_TRY(dn::launch_pack(d->params.as<float>() + o.w[l], cin, cp, d->wf[l].as<uint16_t>(),
                               with_backward ? d->wb[l].as<uint16_t>() : nullptr, d->st));
    }
    d->packed = true;
}

void ensure_shape(rt_denoiser* d, int B, int H, int W, bool training) {
    const int Hp = (H + 7) / 8 * 8, Wp = (W + 7) / 8 * 8;
    d->B = B;
    d->Hp = Hp;
    d->Wp = Wp;
    d->x0.reserve((size_t)level_px(d, 0) * 32 * 2);
    for (int l = 0; l < RT_DN_LAYERS; ++l) {
        const size_t px = (size_t)level_px(d, kLayers[l].level);
        if (l == kOut) d->act[l].reserve(px * 3 * sizeof(float));
        else d->act[l].reserve(px * kLayers[l].cout * 2);
    }
    d->pool1.reserve((size_t)level_px(d, 1) * 32 * 2);
    d->pool3.reserve((size_t)level_px(d, 2) * 64 * 2);
    d->pool5.reserve((size_t)level_px(d, 3) * 128 * 2);
    if (!training) return;
    size_t dz_max = 0, dx_max = 0;
    for (int l = 0; l < RT_DN_LAYERS; ++l) {
        const size_t px = (size_t)level_px(d, kLayers[l].level);
        if (l < kOut) d->g[l].reserve(px * kLayers[l].cout * sizeof(float));
        dz_max = std::max(dz_max, px * coutp_of(l));
        dx_max = std::max(dx_max, px * kLayers[l].cin);
    }
    d->gpool.reserve((size_t)level_px(d, 1) * 32 * sizeof(float) + 64);
    d->dz.reserve(dz_max * 2);
    d->dx.reserve(dx_max * sizeof(float));
    d->part.reserve((size_t)1024 * 9 * 32 * 32 * sizeof(float) + 64);
    d->bpart.reserve((size_t)256 * 128 * sizeof(float));
    d->lpart.reserve(256 * sizeof(float));
}

dn::ConvIn conv_input(rt_denoiser* d, int l) {
    const LayerDef& L = kLayers[l];
    dn::ConvIn in{};
    switch (L.in0) {
        case kInput: in.src0 = d->x0.as<uint16_t>(); break;
        case kPool1: in.src0 = d->pool1.as<uint16_t>(); break;
        case kPool3: in.src0 = d->pool3.as<uint16_t>(); break;
        case kPool5: in.src0 = d->pool5.as<uint16_t>(); break;
        default: in.src0 = d->act[L.in0].as<uint16_t>(); break;
    }
    in.c0 = c0_of(l);
    in.up0 = L.up0;
    if (L.in1 != -9) {
        in.src1 = d->act[L.in1].as<uint16_t>();
        in.c1 = L.cin - in.c0;
    }
    return in;
}

void forward(rt_denoiser* d, const float* feat, int H, int W) {
    const Offsets& o = offsets();
    DN_TRY(dn::launch_prep(feat, d->B, H, W, d->Hp, d->Wp, d->x0.as<uint16_t>(), d->st));
    for (int l = 0; l < RT_DN_LAYERS; ++l) {
        const LayerDef& L = kLayers[l];
        dn::ConvArgs a{};
        a.in = conv_input(d, l);
        a.wt = d->wf[l].as<uint16_t>();
        a.bias = d->params.as<float>() + o.b[l];
        a.coutp = coutp_of(l);
        a.cout = L.cout;
        a.out_pitch = L.cout;
        if (l == kOut) a.out_f32 = d->act[l].as<float>();
        else a.out_bf16 = d->act[l].as<uint16_t>();
        a.relu = L.relu;
        a.B = d->B;
        a.H = level_h(d, L.level);
        a.W = level_w(d, L.level);
        // the encoders' second convolutions also write their 2x2 average pools (the next
        // level's input) from the outputs they hold
        if (l == 1) a.pool_out = d->pool1.as<uint16_t>();
        if (l == 3) a.pool_out = d->pool3.as<uint16_t>();
        if (l == 5) a.pool_out = d->pool5.as<uint16_t>();
        DN_TRY(dn::launch_conv(a, d->st));
    }
}

// Parameter gradients into d->grads; returns the loss.
float backward(rt_denoiser* d, const float* feat, const float* target, int H, int W) {
    const Offsets& o = offsets();
    const int loss_blocks = 256;
    DN_TRY(dn::launch_loss(d->act[kOut].as<float>(), feat, target, d->B, H, W, d->Hp, d->Wp, d->dz.as<uint16_t>(),
                           d->lpart.as<float>(), loss_blocks, d->st));
    for (int l = kOut; l >= 0; --l) {
        const LayerDef& L = kLayers[l];
        const int lh = level_h(d, L.level), lw = level_w(d, L.level);
        const long long px = level_px(d, L.level);
        const int cp = coutp_of(l);
        if (l != kOut)
            DN_TRY(dn::launch_relu_bwd(d->g[l].as<float>(), d->act[l].as<uint16_t>(), px * L.cout,
                                       d->dz.as<uint16_t>(), d->st));
        // weights: split-K partial sums over pixel tiles, then an ordered reduction
        dn::WgradArgs wa{};
        wa.in = conv_input(d, l);
        wa.dz = d->dz.as<uint16_t>();
        wa.coutp = cp;
        wa.B = d->B;
        wa.H = lh;
        wa.W = lw;
        wa.part = d->part.as<float>();
        const int gy = (L.cin / 32) * (cp / 32);
        const long long n_tiles = (long long)d->B * ((lh + 7) / 8) * ((lw + 15) / 16);
        const int splits = (int)std::max(1ll, std::min(n_tiles, (long long)std::max(1, 1024 / gy)));
        DN_TRY(dn::launch_wgrad(wa, splits, d->st));
        DN_TRY(dn::launch_reduce(d->part.as<float>(), splits, 9ll * L.cin * cp, d->grads.as<float>() + o.w[l], false,
                                 d->st));
        const int bsplits = (int)std::min(256ll, std::max(1ll, px / 64));
        DN_TRY(dn::launch_bias_grad(d->dz.as<uint16_t>(), px, cp, d->bpart.as<float>(), bsplits, d->st));
        DN_TRY(dn::launch_reduce(d->bpart.as<float>(), bsplits, cp, d->grads.as<float>() + o.b[l], false, d->st));
        if (l == 0) break;
        // input gradient: the conv with flipped, transposed weights
        dn::ConvArgs a{};
        a.in.src0 = d->dz.as<uint16_t>();
        a.in.c0 = cp;
        a.wt = d->wb[l].as<uint16_t>();
        a.coutp = L.cin;
        a.cout = L.cin;
        a.out_pitch = L.cin;
        a.B = d->B;
        a.H = lh;
        a.W = lw;
        const bool concat = L.in1 != -9;
        const bool pooled = L.in0 == kPool1 || L.in0 == kPool3 || L.in0 == kPool5;
        if (concat) a.out_f32 = d->dx.as<float>();
        else if (pooled) a.out_f32 = d->gpool.as<float>();
        else a.out_f32 = d->g[L.in0].as<float>();
        DN_TRY(dn::launch_conv(a, d->st));
        if (concat) {
            const int c0 = c0_of(l);
            DN_TRY(dn::launch_split(d->dx.as<float>(), d->B, lh, lw, c0, L.cin - c0, d->g[L.in0].as<float>(),
                                    d->g[L.in1].as<float>(), d->st));
        } else if (pooled) {
            const int src = L.in0 == kPool1 ? 1 : L.in0 == kPool3 ? 3 : 5;
            DN_TRY(dn::launch_pool_bwd(d->gpool.as<float>(), d->B, level_h(d, L.level - 1), level_w(d, L.level - 1),
                                       kLayers[src].cout, d->g[src].as<float>(), d->st));
        }
    }
    std::vector<float> lp(loss_blocks);
    DN_TRY(hipMemcpyAsync(lp.data(), d->lpart.p, loss_blocks * sizeof(float), hipMemcpyDeviceToHost, d->st));
    DN_TRY(hipStreamSynchronize(d->st));
    double s = 0;
    for (float v : lp) s += v;
    return (float)(s / ((double)d->B * H * W * 3));
}

void check_shape(int B, int H, int W) {
    if (B <= 0 || H <= 0 || W <= 0) throw Error(RT_E_INVALID_ARG, "batch, height and width must be positive");
    if ((long long)B * ((H + 7) / 8 * 8) * ((W + 7) / 8 * 8) > (1ll << 28))
        throw Error(RT_E_INVALID_ARG, "denoiser input too large");
}

template <class F>
int dn_guard(rt_ctx* ctx, F fn) {
    try {
        fn();
        return RT_OK;
    } catch (const Error& e) {
        return ctx_fail(ctx, e.code, e.what());
    } catch (const std::bad_alloc&) {
        return ctx_fail(ctx, RT_E_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return ctx_fail(ctx, RT_E_INTERNAL, e.what());
    }
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" {

int64_t rt_denoiser_param_count(void) { return offsets().total; }

int rt_denoiser_layer_info(int32_t layer, int32_t* cin, int32_t* coutp, int64_t* w_off, int64_t* b_off) {
    if (layer < 0 || layer >= RT_DN_LAYERS) return RT_E_INVALID_ARG;
    if (cin) *cin = kLayers[layer].cin;
    if (coutp) *coutp = coutp_of(layer);
    if (w_off) *w_off = offsets().w[layer];
    if (b_off) *b_off = offsets().b[layer];
    return RT_OK;
}

int rt_denoiser_create(rt_ctx* ctx, uint64_t seed, rt_denoiser** out) {
    if (!ctx || !out) return ctx_fail(ctx, RT_E_INVALID_ARG, "null argument");
    *out = nullptr;
    auto d = std::make_unique<rt_denoiser>();
    int rc = dn_guard(ctx, [&] {
        d->ctx = ctx;
        d->st = ctx_stream(ctx);
        DN_TRY(hipSetDevice(ctx_device(ctx)));
        DN_TRY(hipEventCreate(&d->e0));
        DN_TRY(hipEventCreate(&d->e1));
        for (auto& e : d->ev) DN_TRY(hipEventCreate(&e));
        DN_TRY(hipStreamCreateWithFlags(&d->st_fb, hipStreamNonBlocking));
        const Offsets& o = offsets();
        d->n_params = o.total;
        // He-normal weights over the real fan-in (13 input features for layer 0); padded
        // input / output channels start at zero and, having zero gradient, stay there
        std::vector<float> w(o.total, 0.0f);
        uint64_t s = seed;
        for (int l = 0; l < RT_DN_LAYERS; ++l) {
            const int cin = kLayers[l].cin, cp = coutp_of(l), cout = kLayers[l].cout;
            const int fan_in_ch = l == 0 ? dn::kFeatIn : cin;
            const double std_dev = std::sqrt(2.0 / (9.0 * fan_in_ch)) * (l == kOut ? 0.1 : 1.0);
            for (int tap = 0; tap < 9; ++tap)
                for (int ci = 0; ci < cin; ++ci)
                    for (int co = 0; co < cp; ++co) {
                        double u1 = ((splitmix(s) >> 11) + 1) * 0x1.0p-53, u2 = (splitmix(s) >> 11) * 0x1.0p-53;
                        double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                        bool live = co < cout && ci < fan_in_ch;
                        w[o.w[l] + ((size_t)tap * cin + ci) * cp + co] = live ? (float)(z * std_dev) : 0.0f;
                    }
        }
        d->params.reserve(o.total * sizeof(float));
        d->grads.reserve(o.total * sizeof(float));
        d->adam_m.reserve(o.total * sizeof(float));
        d->adam_v.reserve(o.total * sizeof(float));
        DN_TRY(hipMemcpyAsync(d->params.p, w.data(), o.total * sizeof(float), hipMemcpyHostToDevice, d->st));
        DN_TRY(hipMemsetAsync(d->adam_m.p, 0, o.total * sizeof(float), d->st));
        DN_TRY(hipMemsetAsync(d->adam_v.p, 0, o.total * sizeof(float), d->st));
        DN_TRY(hipMemsetAsync(d->grads.p, 0, o.total * sizeof(float), d->st));
        pack_weights(d.get(), true);
        DN_TRY(hipStreamSynchronize(d->st));
    });
    if (rc == RT_OK) *out = d.release();
    return rc;
}

void rt_denoiser_destroy(rt_denoiser* dn) {
    if (!dn) return;
    if (dn->st) (void)hipStreamSynchronize(dn->st);
    if (dn->st_fb) {
        (void)hipStreamSynchronize(dn->st_fb);
        (void)hipStreamDestroy(dn->st_fb);
    }
    if (dn->e0) (void)hipEventDestroy(dn->e0);
    if (dn->e1) (void)hipEventDestroy(dn->e1);
    for (auto e : dn->ev)
        if (e) (void)hipEventDestroy(e);
    delete dn;
}

int rt_denoiser_get_params(rt_denoiser* dn, float* host_out, int64_t n) {
    if (!dn || !host_out) return ctx_fail(dn ? dn->ctx : nullptr, RT_E_INVALID_ARG, "null argument");
    return dn_guard(dn->ctx, [&] {
        if (n < dn->n_params) throw Error(RT_E_INVALID_ARG, "parameter buffer too small");
        DN_TRY(hipMemcpyAsync(host_out, dn->params.p, dn->n_params * sizeof(float), hipMemcpyDeviceToHost, dn->st));
        DN_TRY(hipStreamSynchronize(dn->st));
    });
}

int rt_denoiser_set_params(rt_denoiser* dn, const float* host_in, int64_t n) {
    if (!dn || !host_in) return ctx_fail(dn ? dn->ctx : nullptr, RT_E_INVALID_ARG, "null argument");
    return dn_guard(dn->ctx, [&] {
        if (n != dn->n_params) throw Error(RT_E_INVALID_ARG, "parameter count mismatch");
        DN_TRY(hipMemcpyAsync(dn->params.p, host_in, n * sizeof(float), hipMemcpyHostToDevice, dn->st));
        DN_TRY(hipMemsetAsync(dn->adam_m.p, 0, n * sizeof(float), dn->st));
        DN_TRY(hipMemsetAsync(dn->adam_v.p, 0, n * sizeof(float), dn->st));
        dn->step = 0;
        pack_weights(dn, true);
        DN_TRY(hipStreamSynchronize(dn->st));
    });
}

int rt_denoiser_gradients(rt_denoiser* dn, const float* d_features, const float* d_target, int32_t batch, int32_t h,
                          int32_t w, float* host_grads, float* loss_out) {
    if (!dn || !d_features || !d_target) return ctx_fail(dn ? dn->ctx : nullptr, RT_E_INVALID_ARG, "null argument");
    return dn_guard(dn->ctx, [&] {
        check_shape(batch, h, w);
        ensure_shape(dn, batch, h, w, true);
        forward(dn, d_features, h, w);
        float loss = backward(dn, d_features, d_target, h, w);
        if (loss_out) *loss_out = loss;
        if (host_grads) {
            DN_TRY(hipMemcpyAsync(host_grads, dn->grads.p, dn->n_params * sizeof(float), hipMemcpyDeviceToHost, dn->st));
            DN_TRY(hipStreamSynchronize(dn->st));
        }
    });
}

int rt_denoiser_train_step(rt_denoiser* dn, const float* d_features, const float* d_target, int32_t batch, int32_t h,
                           int32_t w, float lr, float* loss_out) {
    if (!dn || !d_features || !d_target) return ctx_fail(dn ? dn->ctx : nullptr, RT_E_INVALID_ARG, "null argument");
    return dn_guard(dn->ctx, [&] {
        check_shape(batch, h, w);
        ensure_shape(dn, batch, h, w, true);
        forward(dn, d_features, h, w);
        float loss = backward(dn, d_features, d_target, h, w);
        ++dn->step;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.0f - (float)std::pow(b1, (double)dn->step);
        const float bc2 = 1.0f - (float)std::pow(b2, (double)dn->step);
        DN_TRY(dn::launch_adam(dn->params.as<float>(), dn->grads.as<float>(), dn->adam_m.as<float>(),
                               dn->adam_v.as<float>(), dn->n_params, lr, b1, b2, 1e-8f, bc1, bc2, dn->st));
        pack_weights(dn, true);
        if (loss_out) *loss_out = loss;
    });
}

int rt_denoiser_activation(rt_denoiser* dn, int32_t layer, float* host_out, int64_t n) {
    if (!dn || !host_out) return ctx_fail(dn ? dn->ctx : nullptr, RT_E_INVALID_ARG, "null argument");
    return dn_guard(dn->ctx, [&] {
        if (layer < 0 || layer >= RT_DN_LAYERS) throw Error(RT_E_INVALID_ARG, "layer out of range");
        if (dn->B == 0) throw Error(RT_E_INVALID_ARG, "no forward pass yet");
        const long long px = level_px(dn, kLayers[layer].level);
        if (layer == kOut) {
            if (n < px * 3) throw Error(RT_E_INVALID_ARG, "buffer too small");
            DN_TRY(hipMemcpyAsync(host_out, dn->act[layer].p, px * 3 * sizeof(float), hipMemcpyDeviceToHost, dn->st));
            DN_TRY(hipStreamSynchronize(dn->st));
            return;
        }
        const long long cnt = px * kLayers[layer].cout;
        if (n < cnt) throw Error(RT_E_INVALID_ARG, "buffer too small");
        std::vector<uint16_t> h(cnt);
        DN_TRY(hipMemcpyAsync(h.data(), dn->act[layer].p, cnt * 2, hipMemcpyDeviceToHost, dn->st));
        DN_TRY(hipStreamSynchronize(dn->st));
        for (long long i = 0; i < cnt; ++i) {
            uint32_t u = (uint32_t)h[i] << 16;
            std::memcpy(&host_out[i], &u, 4);
        }
    });
}

int rt_denoise(rt_denoiser* dn, const float* d_features, int32_t h, int32_t w, float* d_rgb_f32, uint8_t* d_rgb_u8,
               float* ms_out) {
    if (!dn || !d_features) return ctx_fail(dn ? dn->ctx : nullptr, RT_E_INVALID_ARG, "null argument");
    return dn_guard(dn->ctx, [&] {
        check_shape(1, h, w);
        ensure_shape(dn, 1, h, w, false);
        DN_TRY(hipEventRecord(dn->e0, dn->st));
        forward(dn, d_features, h, w);
        DN_TRY(dn::launch_compose(dn->act[kOut].as<float>(), d_features, h, w, dn->Wp, d_rgb_f32, d_rgb_u8, nullptr, 0,
                                  dn->st));
        DN_TRY(hipEventRecord(dn->e1, dn->st));
        DN_TRY(hipEventSynchronize(dn->e1));
        if (ms_out) DN_TRY(hipEventElapsedTime(ms_out, dn->e0, dn->e1));
    });
}

int rt_render_denoised(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, rt_denoiser* dn,
                       int32_t fallback_samples, float threshold, int32_t tile, void* d_rgb_out,
                       rt_denoise_stats* stats) {
    if (!ctx || !scene || !opts || !dn || !d_rgb_out) return ctx_fail(ctx, RT_E_INVALID_ARG, "null argument");
    return dn_guard(ctx, [&] {
        // the network runs on the denoiser's stream and the compose on ctx's: one context
        if (dn->ctx != ctx) throw Error(RT_E_INVALID_ARG, "the denoiser belongs to another context");
        if (scene_ctx(scene) != ctx) throw Error(RT_E_INVALID_ARG, "the scene belongs to another context");
        if (tile <= 0) throw Error(RT_E_INVALID_ARG, "tile must be positive");
        if (fallback_samples < 1) throw Error(RT_E_INVALID_ARG, "fallback_samples must be positive");
        if (opts->outline_enabled && opts->aa_mode != RT_AA_QUINCUNX)
            throw Error(RT_E_INVALID_ARG, "the denoised render does not take the outline pass");
        const int W = opts->width, H = opts->height;
        const int tiles = ((W + tile - 1) / tile) * ((H + tile - 1) / tile);
        dn->feat.reserve((size_t)W * H * RT_FEAT_COUNT * sizeof(float));
        dn->mask.reserve(tiles);
        dn->flagged.reserve(sizeof(int));
        // The few-sample render uses the context's frame buffers (so a repeated frame replays its
        // captured graph); the masked many-sample re-render uses the denoiser's own, so both
        // frames' counters survive until the one host wait at the end.
        rt_render_opts o1 = *opts;
        o1.profile = 0;
        rt_render_opts o2 = *opts;
        o2.profile = 0;
        o2.light_samples = fallback_samples;
        dn->fb_arena.reserve(ctx_frame_bytes(scene, &o2));
        const FixedScratch fb_arena{dn->fb_arena.as<char>(), dn->fb_arena.bytes};
        hipStream_t st = ctx_stream(ctx);
        rt_render_aux aux{};
        aux.features = dn->feat.as<float>();
        rt_render_aux a2{};
        a2.tile_mask = dn->mask.as<uint8_t>();
        a2.mask_tile = tile;
        PendingFrame p1{}, p2{};
        hipStream_t sf = dn->st_fb;
        DN_TRY(hipEventRecord(dn->ev[0], st));
        ctx_render_launch(ctx, scene, &o1, d_rgb_out, &aux, nullptr, &p1);
        DN_TRY(hipEventRecord(dn->ev[1], st));
        // the tile decision needs only the features: the flagged tiles' many-sample re-render
        // starts on the second stream while the network runs on this one; the compose leaves
        // the flagged tiles' pixels to it, so the two write disjoint pixels of d_rgb_out
        DN_TRY(hipMemsetAsync(dn->flagged.p, 0, sizeof(int), st));
        DN_TRY(dn::launch_tile_mask(dn->feat.as<float>(), H, W, tile, threshold, dn->mask.as<uint8_t>(),
                                    dn->flagged.as<int>(), st));
        DN_TRY(hipEventRecord(dn->ev[2], st));
        DN_TRY(hipStreamWaitEvent(sf, dn->ev[2], 0));
        DN_TRY(hipEventRecord(dn->ev[3], sf));
        // every flagged tile re-rendered with the many-sample light; with none flagged the
        // launches find empty queues (no host round trip to find out)
        ctx_render_launch(ctx, scene, &o2, d_rgb_out, &a2, &fb_arena, &p2, sf);
        DN_TRY(hipEventRecord(dn->ev[4], sf));
        ensure_shape(dn, 1, H, W, false);
        forward(dn, dn->feat.as<float>(), H, W);
        DN_TRY(dn::launch_compose(dn->act[kOut].as<float>(), dn->feat.as<float>(), H, W, dn->Wp, nullptr,
                                  static_cast<uint8_t*>(d_rgb_out), dn->mask.as<uint8_t>(), tile, st));
        DN_TRY(hipEventRecord(dn->ev[5], st));
        DN_TRY(hipStreamWaitEvent(st, dn->ev[4], 0));  // the frame ends with both
        DN_TRY(hipEventRecord(dn->ev[6], st));
        int n_flagged = 0;
        DN_TRY(hipMemcpyAsync(&n_flagged, dn->flagged.p, sizeof(int), hipMemcpyDeviceToHost, st));
        DN_TRY(hipStreamSynchronize(st));
        if (stats) {
            rt_render_stats s1{}, s2{};
            frame_stats(ctx, p1, &s1, true);
            frame_stats(ctx, p2, &s2, false);
            float a = 0, b = 0, c = 0, t = 0;
            DN_TRY(hipEventElapsedTime(&a, dn->ev[0], dn->ev[1]));
            DN_TRY(hipEventElapsedTime(&b, dn->ev[1], dn->ev[5]));
            DN_TRY(hipEventElapsedTime(&c, dn->ev[3], dn->ev[4]));
            DN_TRY(hipEventElapsedTime(&t, dn->ev[0], dn->ev[6]));
            stats->ms_render = a;
            stats->ms_denoise = b;
            stats->ms_fallback = c;
            stats->ms_total = t;
            stats->tiles = tiles;
            stats->tiles_fallback = n_flagged;
            stats->rays = s1.rays_primary + s1.rays_reflect + s1.rays_shadow + s2.rays_primary + s2.rays_reflect +
                          s2.rays_shadow;
        } else {
            frame_stats(ctx, p1, nullptr, true);
        }
    });
}

}  // extern "C"
