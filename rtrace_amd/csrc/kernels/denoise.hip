// denoise.hip — BF16 MFMA kernels of the shadow denoiser (a 4-level U-Net of 3x3 convs).
//
// Tensors are NHWC: [batch][height][width][channels] with channels a multiple of 32, so a
// 16-byte vector is 8 consecutive channels of one pixel. Activations are bf16, gradients
// f32, and every product runs on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// A 3x3 convolution is an implicit GEMM: M = pixels, N = output channels, K = 9 taps x
// input channels. A workgroup owns an 8x16-pixel tile and 32 output channels; per K chunk
// of 32 input channels it stages the 10x18 input halo and the 9x32x32 weight block in LDS
// and each of its four waves runs 9 taps x (2 pixel rows x 2 channel halves) MFMAs. The
// same kernel computes backward-data (input = dZ, weights flipped and transposed, f32
// output). Weight gradients stage the halo and dZ transposed (channel-major, one copy per
// x shift) so that both MFMA operands are 16-byte aligned 8-pixel runs. Convolutions with
// 32 input channels (forward and backward-data) run on a kernel without LDS: one wavefront
// per 16-pixel column strip, the weights in registers, the pixels read straight from memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "denoise_kernels.hpp"

namespace rt {
namespace dn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// two floats rounded to bf16 (nearest even) and packed, low one first: gfx950's
// v_cvt_pk_bf16_f32, the same rounding as f2bf
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
    const bf16x2 v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}

constexpr int kTH = 8, kTW = 16;                   // output tile (pixels)

// 2x2 average pooling of four bf16 outputs (top-left, top-right, bottom-left, bottom-right):
// the two row sums added, times 1/4, rounded to bf16.
__device__ __forceinline__ uint16_t pool4(uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
    return f2bf(((bf2f(a) + bf2f(b)) + (bf2f(c) + bf2f(d))) * 0.25f);
}
constexpr int kHH = kTH + 2, kHW = kTW + 2;        // input halo
constexpr int kCK = 32;                            // input channels per K chunk
constexpr int kPS = 40;                            // LDS pixel stride in bf16 (80 B)
constexpr int kNT = 32;                            // output channels per workgroup

// The same vector without branches and without the zeroing: the address is clamped into
// the image and *ok says whether the pixel is inside it (outside pixels read as zero once
// the caller applies *ok). The load's value is not touched here, so a prefetch is not
// waited for until the value is stored.
__device__ __forceinline__ uint4 gather8_raw(const ConvIn& in, int b, int y, int x, int H, int W, int c, bool* ok) {
    *ok = y >= 0 && x >= 0 && y < H && x < W;
    const int yc = min(max(y, 0), H - 1), xc = min(max(x, 0), W - 1);
    const int h2 = H >> 1, w2 = W >> 1;
    const size_t i0 = in.up0 ? (((size_t)b * h2 + (yc >> 1)) * w2 + (xc >> 1)) : (((size_t)b * H + yc) * W + xc);
    const size_t i1 = ((size_t)b * H + yc) * W + xc;
    const bool first = c < in.c0;
    const uint16_t* p = first ? in.src0 + i0 * in.c0 + c : in.src1 + i1 * in.c1 + (c - in.c0);
    return *reinterpret_cast<const uint4*>(p);
}

// The same vector without branches: the address is selected and clamped into the image,
// and a pixel outside it reads as zero. Staging loops issue all their loads before the
// first wait this way (a load under a divergent branch is waited for at the branch's end).
__device__ __forceinline__ uint4 gather8_nb(const ConvIn& in, int b, int y, int x, int H, int W, int c) {
    const bool ok = y >= 0 && x >= 0 && y < H && x < W;
    const int yc = min(max(y, 0), H - 1), xc = min(max(x, 0), W - 1);
    const int h2 = H >> 1, w2 = W >> 1;
    const size_t i0 = in.up0 ? (((size_t)b * h2 + (yc >> 1)) * w2 + (xc >> 1)) : (((size_t)b * H + yc) * W + xc);
    const size_t i1 = ((size_t)b * H + yc) * W + xc;
    const bool first = c < in.c0;
    const uint16_t* p = first ? in.src0 + i0 * in.c0 + c : in.src1 + i1 * in.c1 + (c - in.c0);
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return ok ? v : make_uint4(0, 0, 0, 0);
}

// 16-byte vector of channels [c, c+8) of the (virtual) input pixel (b, y, x) of a conv whose
// input is concat(src0 (optionally at half resolution, nearest upsampling), src1).
__device__ __forceinline__ uint4 gather8(const ConvIn& in, int b, int y, int x, int H, int W, int c) {
    if (y < 0 || x < 0 || y >= H || x >= W) return make_uint4(0, 0, 0, 0);
    if (c < in.c0) {
        const uint16_t* p;
        if (in.up0) {
            const int h2 = H >> 1, w2 = W >> 1;
            p = in.src0 + (((size_t)b * h2 + (y >> 1)) * w2 + (x >> 1)) * in.c0;
        } else {
            p = in.src0 + (((size_t)b * H + y) * W + x) * in.c0;
        }
        return *reinterpret_cast<const uint4*>(p + c);
    }
    const uint16_t* p = in.src1 + (((size_t)b * H + y) * W + x) * in.c1;
    return *reinterpret_cast<const uint4*>(p + (c - in.c0));
}

// One pixel tile per workgroup, all K chunks: layers with several K chunks (one tile's
// chunks in flight per workgroup and many workgroups per CU measured faster there than the
// persistent kernel below: 1080p dec0a (96 -> 32) 229 us against 424 us).
__global__ __launch_bounds__(256) void k_conv3x3_tile(ConvArgs a) {
    // one spare pixel slot each: the staging threads with no vector to store write there, so
    // the staging has no branches and all of its loads are in flight together
    __shared__ __attribute__((aligned(16))) uint16_t s_in[(kHH * kHW + 1) * kPS];
    __shared__ __attribute__((aligned(16))) uint16_t s_w[(9 * kNT + 1) * kPS];
    const int tiles_x = (a.W + kTW - 1) / kTW, tiles_y = (a.H + kTH - 1) / kTH;
    int t = blockIdx.x;
    const int b = t / (tiles_x * tiles_y);
    t -= b * tiles_x * tiles_y;
    const int y0 = (t / tiles_x) * kTH, x0 = (t % tiles_x) * kTW;
    const int co0 = blockIdx.y * kNT;
    const int cin = a.in.c0 + a.in.c1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    constexpr int kHaloVec = kHH * kHW * 4, kWVec = 9 * kNT * 4;  // 16-byte vectors per K chunk
    constexpr int kHaloPer = (kHaloVec + 255) / 256, kWPer = (kWVec + 255) / 256;
    for (int ci0 = 0; ci0 < cin; ci0 += kCK) {
        // every load of the chunk first (indices clamped, no branches), then the stores
        uint4 hv[kHaloPer], wv[kWPer];
#pragma unroll
        for (int r = 0; r < kHaloPer; ++r) {
            const int v = min((int)threadIdx.x + r * 256, kHaloVec - 1);
            const int px = v >> 2, q = v & 3;
            const int hy = px / kHW, hx = px - hy * kHW;
            hv[r] = gather8_nb(a.in, b, y0 + hy - 1, x0 + hx - 1, a.H, a.W, ci0 + q * 8);
        }
#pragma unroll
        for (int r = 0; r < kWPer; ++r) {
            const int v = min((int)threadIdx.x + r * 256, kWVec - 1);
            const int q = v & 3, n = (v >> 2) % kNT, tap = (v >> 2) / kNT;
            wv[r] = *reinterpret_cast<const uint4*>(a.wt + ((size_t)tap * a.coutp + co0 + n) * cin + ci0 + q * 8);
        }
#pragma unroll
        for (int r = 0; r < kHaloPer; ++r) {
            const int v = (int)threadIdx.x + r * 256;
            const int slot = v < kHaloVec ? (v >> 2) * kPS + (v & 3) * 8 : kHH * kHW * kPS;
            *reinterpret_cast<uint4*>(s_in + slot) = hv[r];
        }
#pragma unroll
        for (int r = 0; r < kWPer; ++r) {
            const int v = (int)threadIdx.x + r * 256;
            const int q = v & 3, n = (v >> 2) % kNT, tap = (v >> 2) / kNT;
            const int slot = v < kWVec ? (tap * kNT + n) * kPS + q * 8 : 9 * kNT * kPS;
            *reinterpret_cast<uint4*>(s_w + slot) = wv[r];
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap - dy * 3;
            bf16x8 bfr[2];
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
                bfr[nt] = *reinterpret_cast<const bf16x8*>(s_w + (tap * kNT + nt * 16 + r16) * kPS + g * 8);
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                const int row = wave * 2 + mt;
                bf16x8 afr = *reinterpret_cast<const bf16x8*>(s_in + ((row + dy) * kHW + r16 + dx) * kPS + g * 8);
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[nt], acc[mt][nt], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D layout: lane holds rows (pixels) 4g..4g+3 of column (channel) r16
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const int y = y0 + wave * 2 + mt;
        if (y >= a.H) continue;
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int co = co0 + nt * 16 + r16;
            if (co >= a.cout) continue;
            const float bias = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int x = x0 + 4 * g + i;
                if (x >= a.W) continue;
                float v = acc[mt][nt][i] + bias;
                const size_t o = (((size_t)b * a.H + y) * a.W + x) * a.out_pitch + co;
                if (a.out_f32) {
                    a.out_f32[o] = v;
                } else {
                    if (a.relu) v = fmaxf(v, 0.0f);
                    a.out_bf16[o] = (uint16_t)f2bf2(v, 0.0f);
                }
            }
        }
    }
    if (a.pool_out) {
        // this lane holds output rows y, y + 1 (y even) and columns x .. x + 3 (x even) of
        // channel co: two pooled pixels
        const int y = y0 + wave * 2;
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int co = co0 + nt * 16 + r16;
            const float bias = a.bias && co < a.cout ? a.bias[co] : 0.0f;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int x = x0 + 4 * g + 2 * p;
                if (y + 1 < a.H && x + 1 < a.W && co < a.cout) {
                    auto q = [&](int mt, int i) {
                        float v = acc[mt][nt][i] + bias;
                        if (a.relu) v = fmaxf(v, 0.0f);
                        return (uint16_t)f2bf2(v, 0.0f);
                    };
                    const size_t o = (((size_t)b * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1)) * a.out_pitch + co;
                    a.pool_out[o] = pool4(q(0, 2 * p), q(0, 2 * p + 1), q(1, 2 * p), q(1, 2 * p + 1));
                }
            }
        }
    }
}

// Staging of one step (a pixel tile's halo and, optionally, a weight block for one
// 32-channel K chunk) for the persistent kernel: loads into registers, later stores into
// LDS. Halo loads are raw buffer loads when both sources are under 2 GiB: a pixel outside
// the image gets an offset past the buffer's end and reads as zero, so no value is
// selected on before it is stored.
constexpr int kHaloVec = kHH * kHW * 4, kWVec = 9 * kNT * 4;  // 16-byte vectors per K chunk
constexpr int kHaloPer = (kHaloVec + 255) / 256, kWPer = (kWVec + 255) / 256;

struct StepGeom {
    int tiles_x, tiles_y, h2, w2, cin, co0;
    unsigned long long bytes0, bytes1;
    bool use_buf;
};

__device__ __forceinline__ void tile_origin(const StepGeom& g, int t, int& b, int& y0, int& x0) {
    b = t / (g.tiles_x * g.tiles_y);
    t -= b * g.tiles_x * g.tiles_y;
    y0 = (t / g.tiles_x) * kTH;
    x0 = (t % g.tiles_x) * kTW;
}

__device__ __forceinline__ void load_halo(const ConvArgs& a, const StepGeom& g, int t, int ci0, uint4 (&hv)[kHaloPer],
                                          uint32_t& hok) {
    int b, y0, x0;
    tile_origin(g, t, b, y0, x0);
    hok = 0;
    if (g.use_buf) {
        // a K chunk lies within one source (both channel counts are multiples of 32)
        const bool first = ci0 < a.in.c0;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(first ? a.in.src0 : a.in.src1), (short)0, (int)(first ? g.bytes0 : g.bytes1), 0x00020000);
#pragma unroll
        for (int r = 0; r < kHaloPer; ++r) {
            const int v = min((int)threadIdx.x + r * 256, kHaloVec - 1);
            const int px = v >> 2, q = v & 3;
            const int hy = px / kHW, hx = px - hy * kHW;
            const int y = y0 + hy - 1, x = x0 + hx - 1, c = ci0 + q * 8;
            const bool ok = y >= 0 && x >= 0 && y < a.H && x < a.W;
            const uint32_t e = first ? (uint32_t)(a.in.up0 ? ((b * g.h2 + (y >> 1)) * g.w2 + (x >> 1))
                                                          : ((b * a.H + y) * a.W + x)) * (uint32_t)a.in.c0 + (uint32_t)c
                                     : (uint32_t)((b * a.H + y) * a.W + x) * (uint32_t)a.in.c1 + (uint32_t)(c - a.in.c0);
            const auto w4 = __builtin_amdgcn_raw_buffer_load_b128(rs, ok ? e * 2u : 0x80000000u, 0, 0);
            hv[r] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
        }
        hok = (1u << kHaloPer) - 1u;
    } else {
#pragma unroll
        for (int r = 0; r < kHaloPer; ++r) {
            const int v = min((int)threadIdx.x + r * 256, kHaloVec - 1);
            const int px = v >> 2, q = v & 3;
            const int hy = px / kHW, hx = px - hy * kHW;
            bool ok;
            hv[r] = gather8_raw(a.in, b, y0 + hy - 1, x0 + hx - 1, a.H, a.W, ci0 + q * 8, &ok);
            hok |= ok ? 1u << r : 0u;
        }
    }
}

__device__ __forceinline__ void load_w(const ConvArgs& a, const StepGeom& g, int ci0, uint4 (&wv)[kWPer]) {
#pragma unroll
    for (int r = 0; r < kWPer; ++r) {
        const int v = min((int)threadIdx.x + r * 256, kWVec - 1);
        const int q = v & 3, n = (v >> 2) % kNT, tap = (v >> 2) / kNT;
        wv[r] = *reinterpret_cast<const uint4*>(a.wt + ((size_t)tap * a.coutp + g.co0 + n) * g.cin + ci0 + q * 8);
    }
}

__device__ __forceinline__ void store_halo(uint16_t* s_in, const uint4 (&hv)[kHaloPer], uint32_t hok) {
#pragma unroll
    for (int r = 0; r < kHaloPer; ++r) {
        const int v = (int)threadIdx.x + r * 256;
        const int slot = v < kHaloVec ? (v >> 2) * kPS + (v & 3) * 8 : kHH * kHW * kPS;
        *reinterpret_cast<uint4*>(s_in + slot) = ((hok >> r) & 1u) ? hv[r] : make_uint4(0, 0, 0, 0);
    }
}

__device__ __forceinline__ void store_w(uint16_t* s_w, const uint4 (&wv)[kWPer]) {
#pragma unroll
    for (int r = 0; r < kWPer; ++r) {
        const int v = (int)threadIdx.x + r * 256;
        const int q = v & 3, n = (v >> 2) % kNT, tap = (v >> 2) / kNT;
        const int slot = v < kWVec ? (tap * kNT + n) * kPS + q * 8 : 9 * kNT * kPS;
        *reinterpret_cast<uint4*>(s_w + slot) = wv[r];
    }
}

// Layers with a single 32-channel K chunk (32 input channels), persistent and
// software-pipelined: workgroup (bx, by) stages the weights of output channels
// [32 by, 32 by + 32) once, then computes the pixel tiles bx, bx + gridDim.x, ... The halo
// loads of the tiles kHaloDepth steps ahead are issued into registers before the current
// tile's MFMAs, so they are in flight while the matrix cores work, then stored into LDS
// between two barriers. The MFMAs accumulate in tap order, as the one-tile kernel does.
// Pixel tiles whose halo loads a workgroup keeps in flight: 1080p forward pass 1.44 ms with
// one, 1.43 with two, 1.40 with three held to 4 wavefronts per SIMD (124 VGPRs; left to the
// compiler, three take 134 VGPRs, 3 wavefronts per SIMD and 1.51 ms).
constexpr int kHaloDepth = 3;
template <int kDepth>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_conv3x3(ConvArgs a) {
    // one spare pixel slot each: the staging threads with no vector to store write there, so
    // the staging has no branches and all of its loads are in flight together
    __shared__ __attribute__((aligned(16))) uint16_t s_in[(kHH * kHW + 1) * kPS];
    __shared__ __attribute__((aligned(16))) uint16_t s_w[(9 * kNT + 1) * kPS];
    StepGeom g;
    g.tiles_x = (a.W + kTW - 1) / kTW;
    g.tiles_y = (a.H + kTH - 1) / kTH;
    g.h2 = a.H >> 1;
    g.w2 = a.W >> 1;
    g.cin = a.in.c0 + a.in.c1;
    g.co0 = blockIdx.y * kNT;
    g.bytes0 = (unsigned long long)a.B * (a.in.up0 ? (unsigned long long)g.h2 * g.w2 : (unsigned long long)a.H * a.W) *
               a.in.c0 * 2;
    g.bytes1 = (unsigned long long)a.B * a.H * a.W * a.in.c1 * 2;
    g.use_buf = g.bytes0 < (1ull << 31) && g.bytes1 < (1ull << 31);
    const int n_tiles = a.B * g.tiles_x * g.tiles_y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int gq = lane >> 4, r16 = lane & 15;

    // kDepth register sets of halo loads in flight: the tile staged now is followed by the
    // kDepth - 1 next ones of this workgroup, so a tile's loads have kDepth tiles of MFMAs to land
    uint4 hv[kDepth][kHaloPer];
    uint32_t hok[kDepth] = {};
    if ((int)blockIdx.x >= n_tiles) return;
#pragma unroll
    for (int d = 0; d < kDepth; ++d) {
        const int t = (int)blockIdx.x + d * (int)gridDim.x;
        if (t < n_tiles) load_halo(a, g, t, 0, hv[d], hok[d]);
    }
    {
        uint4 wv[kWPer];
        load_w(a, g, 0, wv);
        store_w(s_w, wv);  // ordered before the first MFMA by the loop's second barrier
    }
    f32x4 acc00 = f32x4{0.f, 0.f, 0.f, 0.f}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
    for (int base = (int)blockIdx.x; base < n_tiles; base += kDepth * (int)gridDim.x) {
#pragma unroll
        for (int d = 0; d < kDepth; ++d) {
            const int tile = base + d * (int)gridDim.x;
            if (tile >= n_tiles) break;  // uniform across the workgroup
            __syncthreads();  // the previous tile's MFMAs are done with the LDS
            store_halo(s_in, hv[d], hok[d]);
            __syncthreads();
            // this register set now loads the tile kDepth steps ahead
            const int ahead = tile + kDepth * (int)gridDim.x;
            if (ahead < n_tiles) load_halo(a, g, ahead, 0, hv[d], hok[d]);
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) {
                const int dy = tap / 3, dx = tap - dy * 3;
                const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(s_w + (tap * kNT + r16) * kPS + gq * 8);
                const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(s_w + (tap * kNT + 16 + r16) * kPS + gq * 8);
                const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(s_in + ((wave * 2 + dy) * kHW + r16 + dx) * kPS + gq * 8);
                const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(s_in + ((wave * 2 + 1 + dy) * kHW + r16 + dx) * kPS + gq * 8);
                acc00 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc00, 0, 0, 0);
                acc01 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc01, 0, 0, 0);
                acc10 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc10, 0, 0, 0);
                acc11 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc11, 0, 0, 0);
            }
            {
                // D layout: lane holds rows (pixels) 4gq..4gq+3 of column (channel) r16
                int b, y0, x0;
                tile_origin(g, tile, b, y0, x0);
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) {
                    const int y = y0 + wave * 2 + mt;
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
                        const f32x4 acc = mt == 0 ? (nt == 0 ? acc00 : acc01) : (nt == 0 ? acc10 : acc11);
                        const int co = g.co0 + nt * 16 + r16;
                        const float bias = a.bias && co < a.cout ? a.bias[co] : 0.0f;
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int x = x0 + 4 * gq + i;
                            if (y < a.H && co < a.cout && x < a.W) {
                                float v = acc[i] + bias;
                                const size_t o = (((size_t)b * a.H + y) * a.W + x) * a.out_pitch + co;
                                if (a.out_f32) {
                                    a.out_f32[o] = v;
                                } else {
                                    if (a.relu) v = fmaxf(v, 0.0f);
                                    a.out_bf16[o] = f2bf(v);
                                }
                            }
                        }
                    }
                }
                if (a.pool_out) {
                    // rows y, y + 1 (y even), columns x .. x + 3 (x even) of channel co: two pooled pixels
                    const int y = y0 + wave * 2;
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
                        const f32x4 top = nt == 0 ? acc00 : acc01, bot = nt == 0 ? acc10 : acc11;
                        const int co = g.co0 + nt * 16 + r16;
                        const float bias = a.bias && co < a.cout ? a.bias[co] : 0.0f;
                        auto q = [&](float s) {
                            float v = s + bias;
                            if (a.relu) v = fmaxf(v, 0.0f);
                            return f2bf(v);
                        };
#pragma unroll
                        for (int p = 0; p < 2; ++p) {
                            const int x = x0 + 4 * gq + 2 * p;
                            if (y + 1 < a.H && x + 1 < a.W && co < a.cout) {
                                const size_t o =
                                    (((size_t)b * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1)) * a.out_pitch + co;
                                a.pool_out[o] = pool4(q(top[2 * p]), q(top[2 * p + 1]), q(bot[2 * p]), q(bot[2 * p + 1]));
                            }
                        }
                    }
                }
                acc00 = acc01 = acc10 = acc11 = f32x4{0.f, 0.f, 0.f, 0.f};
            }
        }
    }
}

// Layers with 32 input channels, one wavefront per column strip (16 pixels wide, kRows
// rows tall) and no LDS or barrier: the wavefront keeps its 32 output channels' weights in
// registers (9 taps x 2 halves of 16 channels) and walks the strip from top to bottom. The
// pixel operand of a tap, 8 channels of one input pixel per lane, is read straight from
// memory (16 pixels of a row are 1 KB in a row), once per input row and pixel shift; the
// rows in use stay in kSlots register sets, and the loads of the row kSlots - 3 rows ahead
// are in flight during the current row's 18 MFMAs. The products are the transposed ones of
// the LDS kernels (weights as the A operand, pixels as B), so each lane ends with 4
// consecutive channels of one pixel and stores them as one vector; every output is the
// same 32-term dot product per tap, accumulated in tap order, as in those kernels. Sources
// and outputs are under 2 GiB: raw buffer loads read an offset past the end as zero (the
// convolution's zero padding) and buffer stores there are dropped (pixels past the edge).
constexpr int kStripW = 16;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kOff = 0x80000000u;  // a buffer offset past every buffer's end

// 2x2 average pools of two channel pairs (top-left, top-right, bottom-left, bottom-right
// packed bf16), as pool4, packed
__device__ __forceinline__ uint32_t pool4x2(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    auto lo = [](uint32_t v) { return __uint_as_float(v << 16); };
    auto hi = [](uint32_t v) { return __uint_as_float(v & 0xffff0000u); };
    return f2bf2(((lo(a) + lo(b)) + (lo(c) + lo(d))) * 0.25f, ((hi(a) + hi(b)) + (hi(c) + hi(d))) * 0.25f);
}
// the value of lane ^ 1 (DPP quad_perm [1, 0, 3, 2])
__device__ __forceinline__ uint32_t lane_xor1(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
}

template <int kRows, int kSlots>
__global__ __launch_bounds__(256) void k_conv3x3_c32(ConvArgs a, int strips_x,
                                                                                            int strips_y,
                                                                                            unsigned src_bytes,
                                                                                            unsigned out_bytes,
                                                                                            unsigned pool_bytes) {
    static_assert(kSlots >= 4 && kRows % 2 == 0, "three rows in use plus one loading; strips start on even rows");
    const int lane = threadIdx.x & 63;
    const int gq = lane >> 4, r16 = lane & 15;
    const int co0 = blockIdx.y * kNT;
    const bool first = a.in.c0 > 0;  // the 32 channels come from one source
    const bool up = first && a.in.up0;
    const int sh = up ? a.H >> 1 : a.H, sw = up ? a.W >> 1 : a.W;  // the source's resolution
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(first ? a.in.src0 : a.in.src1), (short)0, (int)src_bytes, 0x00020000);
    const bool f32 = a.out_f32 != nullptr;
    const __amdgpu_buffer_rsrc_t ro = __builtin_amdgcn_make_buffer_rsrc(
        f32 ? (void*)a.out_f32 : (void*)a.out_bf16, (short)0, (int)out_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rp =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.pool_out, (short)0, (int)pool_bytes, 0x00020000);

    // weights of tap t, half nt: output channel co0 + 16 nt + r16, input channels 8 gq .. 8 gq + 7
    bf16x8 w[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
            w[t][nt] = *reinterpret_cast<const bf16x8*>(a.wt + ((size_t)t * a.coutp + co0 + nt * 16 + r16) * kCK + gq * 8);
    // this lane's outputs: channels co0 + 16 nt + 4 gq + k of pixel x0 + r16
    float bias[2][4];
    bool full[2];  // all four channels written
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int c = co0 + nt * 16 + 4 * gq;
        full[nt] = c + 3 < a.cout;
#pragma unroll
        for (int k = 0; k < 4; ++k) bias[nt][k] = a.bias && c + k < a.cout ? a.bias[c + k] : 0.0f;
    }

    const int n_strips = a.B * strips_x * strips_y;
    for (int s = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6); s < n_strips; s += (int)gridDim.x * 4) {
        const int b = s / (strips_x * strips_y);
        const int t = s - b * strips_x * strips_y;
        const int y0 = (t / strips_x) * kRows, x0 = (t % strips_x) * kStripW;
        const int y_end = min(y0 + kRows, a.H);
        const int x = x0 + r16;
        // input row yy at the three pixel shifts of lane r16 (columns x - 1 .. x + 1)
        auto load_row = [&](int yy, bf16x8(&f)[3]) {
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int xx = x + dx - 1;
                const bool ok = yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
                const uint32_t e = up ? (uint32_t)((b * sh + (yy >> 1)) * sw + (xx >> 1))
                                      : (uint32_t)((b * a.H + yy) * a.W + xx);
                const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, ok ? (e * kCK + gq * 8) * 2u : kOff, 0, 0);
                f[dx] = __builtin_bit_cast(bf16x8, v);
            }
        };
        // slot k % kSlots holds input row y0 - 1 + k
        bf16x8 F[kSlots][3];
#pragma unroll
        for (int k = 0; k < kSlots - 1; ++k) load_row(y0 - 1 + k, F[k]);
        uint32_t top[2][2];  // the even row's bf16 outputs (two per dword), for the 2x2 pool
        for (int yb = y0; yb < y_end; yb += kSlots) {
#pragma unroll
            for (int i = 0; i < kSlots; ++i) {
                const int y = yb + i;
                if (y >= y_end) break;
                // input row y + kSlots - 2 into the slot row y - 2 held
                const int ya = y + kSlots - 2;
                if (ya <= y_end) load_row(ya, F[(i + kSlots - 1) % kSlots]);
                f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
                for (int tap = 0; tap < 9; ++tap) {
                    const int dy = tap / 3, dx = tap - dy * 3;
                    const bf16x8 pv = F[(i + dy) % kSlots][dx];
                    acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[tap][0], pv, acc[0], 0, 0, 0);
                    acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[tap][1], pv, acc[1], 0, 0, 0);
                }
                // D layout: lane holds channels 4 gq .. 4 gq + 3 (of each half) of pixel r16
                const uint32_t px = (uint32_t)((b * a.H + y) * a.W + x);
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) {
                    const int c = co0 + nt * 16 + 4 * gq;
                    const uint32_t el = px * (uint32_t)a.out_pitch + (uint32_t)c;
                    float v[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) v[k] = acc[nt][k] + bias[nt][k];
                    if (f32) {
                        if (full[nt]) {
                            const u32x4 q = {__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]),
                                             __float_as_uint(v[3])};
                            __builtin_amdgcn_raw_buffer_store_b128(q, ro, x < a.W ? el * 4u : kOff, 0, 0);
                        } else {
#pragma unroll
                            for (int k = 0; k < 4; ++k)
                                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[k]), ro,
                                                                      x < a.W && c + k < a.cout ? (el + k) * 4u : kOff,
                                                                      0, 0);
                        }
                        continue;
                    }
                    if (a.relu) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) v[k] = fmaxf(v[k], 0.0f);
                    }
                    const uint32_t lo = f2bf2(v[0], v[1]), hi = f2bf2(v[2], v[3]);
                    if (full[nt]) {
                        __builtin_amdgcn_raw_buffer_store_b64(u32x2{lo, hi}, ro, x < a.W ? el * 2u : kOff, 0, 0);
                    } else {
                        const uint32_t qw[2] = {lo, hi};
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(qw[k >> 1] >> ((k & 1) * 16)), ro,
                                                                  x < a.W && c + k < a.cout ? (el + k) * 2u : kOff, 0, 0);
                    }
                    if (a.pool_out) {
                        if ((y & 1) == 0) {
                            top[nt][0] = lo;
                            top[nt][1] = hi;
                        } else {
                            // the right-hand pixel's rows come from lane r16 ^ 1; even lanes store
                            const uint32_t p0 = pool4x2(top[nt][0], lane_xor1(top[nt][0]), lo, lane_xor1(lo));
                            const uint32_t p1 = pool4x2(top[nt][1], lane_xor1(top[nt][1]), hi, lane_xor1(hi));
                            const bool st = (r16 & 1) == 0 && x + 1 < a.W;
                            const uint32_t pe =
                                ((uint32_t)((b * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1))) * (uint32_t)a.out_pitch +
                                (uint32_t)c;
                            if (full[nt]) {
                                __builtin_amdgcn_raw_buffer_store_b64(u32x2{p0, p1}, rp, st ? pe * 2u : kOff, 0, 0);
                            } else {
                                const uint32_t pw[2] = {p0, p1};
#pragma unroll
                                for (int k = 0; k < 4; ++k)
                                    __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(pw[k >> 1] >> ((k & 1) * 16)), rp,
                                                                          st && c + k < a.cout ? (pe + k) * 2u : kOff, 0, 0);
                            }
                        }
                    }
                }
            }
        }
    }
}

// Layers with several 32-channel K chunks, one wavefront per 16x4-pixel step and the
// weights staged once: a workgroup of kRowsWaves wavefronts (one per CU, three per SIMD)
// copies all weights of its 32 output channels into LDS (18 KB per K chunk, fragment-major
// so a wavefront's read of one fragment is 1 KB in a row), then each wavefront takes steps
// of 16 pixels x 4 rows on its own. Per K chunk it reads the 6 input rows x 3 pixel shifts
// the step needs straight from memory (the pixel operand, 8 channels of one pixel per lane;
// nearest upsampling of the first source is an index shift), and each weight fragment it
// reads from LDS feeds the 4 rows' MFMAs. Products are the transposed ones of the LDS tile
// kernel (weights as A, pixels as B) and accumulate in the same order, K chunk by K chunk
// and tap by tap, so the outputs are the same bit for bit; each lane ends with 4
// consecutive channels of one pixel. Sources and outputs are under 2 GiB (buffer loads past
// the end read zero, stores there are dropped).
constexpr size_t kRowsMaxLds = 8 * 9 * kNT * kCK * 2;  // 8 K chunks: 144 KiB

template <int kStepRows, int kRowsWaves>
__global__ __launch_bounds__(kRowsWaves * 64) void k_conv3x3_rows(ConvArgs a, int strips_x, int steps_y,
                                                                  unsigned bytes0, unsigned bytes1, unsigned out_bytes,
                                                                  unsigned pool_bytes) {
    extern __shared__ __attribute__((aligned(16))) uint16_t s_wt[];  // [chunk][tap][half][lane] x 8 bf16
    const int cin = a.in.c0 + a.in.c1, nc = cin / kCK, nc0 = a.in.c0 / kCK;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int gq = lane >> 4, r16 = lane & 15;
    const int co0 = blockIdx.y * kNT;
    {
        const int units = nc * 9 * 2 * 64;
        for (int u = threadIdx.x; u < units; u += blockDim.x) {
            const int l = u & 63, nt = (u >> 6) & 1, tc = u >> 7, tap = tc % 9, c = tc / 9;
            const uint4 v = *reinterpret_cast<const uint4*>(
                a.wt + ((size_t)tap * a.coutp + co0 + nt * 16 + (l & 15)) * cin + c * kCK + (l >> 4) * 8);
            *reinterpret_cast<uint4*>(s_wt + (size_t)u * 8) = v;
        }
    }
    __syncthreads();  // the only barrier: from here on every wavefront works alone
    const __amdgpu_buffer_rsrc_t rs0 =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.in.src0, (short)0, (int)bytes0, 0x00020000);
    const __amdgpu_buffer_rsrc_t rs1 =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.in.src1, (short)0, (int)bytes1, 0x00020000);
    const bool f32 = a.out_f32 != nullptr;
    const __amdgpu_buffer_rsrc_t ro = __builtin_amdgcn_make_buffer_rsrc(
        f32 ? (void*)a.out_f32 : (void*)a.out_bf16, (short)0, (int)out_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rp =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.pool_out, (short)0, (int)pool_bytes, 0x00020000);
    const int h2 = a.H >> 1, w2 = a.W >> 1;

    const int n_steps = a.B * strips_x * steps_y;
    for (int s = (int)blockIdx.x * kRowsWaves + wave; s < n_steps; s += (int)gridDim.x * kRowsWaves) {
        const int b = s / (strips_x * steps_y);
        const int t = s - b * strips_x * steps_y;
        const int y0 = (t / strips_x) * kStepRows, x = (t % strips_x) * kStripW + r16;
        f32x4 acc[kStepRows][2];
#pragma unroll
        for (int r = 0; r < kStepRows; ++r) acc[r][0] = acc[r][1] = f32x4{0.f, 0.f, 0.f, 0.f};
        // input rows y0 - 1 .. y0 + kStepRows of K chunk c at the three pixel shifts
        auto load_chunk = [&](int c, bf16x8(&f)[kStepRows + 2][3]) {
            const bool first = c < nc0;
            const bool up = first && a.in.up0;
            const int cc = first ? c : c - nc0, cs = first ? a.in.c0 : a.in.c1;
#pragma unroll
            for (int r = 0; r < kStepRows + 2; ++r) {
                const int yy = y0 + r - 1;
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const int xx = x + dx - 1;
                    const bool ok = yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
                    const uint32_t e = up ? (uint32_t)((b * h2 + (yy >> 1)) * w2 + (xx >> 1))
                                          : (uint32_t)((b * a.H + yy) * a.W + xx);
                    const uint32_t off = ok ? (e * (uint32_t)cs + (uint32_t)(cc * kCK + gq * 8)) * 2u : kOff;
                    const u32x4 v = first ? __builtin_amdgcn_raw_buffer_load_b128(rs0, off, 0, 0)
                                          : __builtin_amdgcn_raw_buffer_load_b128(rs1, off, 0, 0);
                    f[r][dx] = __builtin_bit_cast(bf16x8, v);
                }
            }
        };
        auto mfma_chunk = [&](int c, const bf16x8(&f)[kStepRows + 2][3]) {
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) {
                const int dy = tap / 3, dx = tap - dy * 3;
                const uint16_t* wp = s_wt + ((size_t)((c * 9 + tap) * 2) * 64 + lane) * 8;
                const bf16x8 w0 = *reinterpret_cast<const bf16x8*>(wp);
                const bf16x8 w1 = *reinterpret_cast<const bf16x8*>(wp + 64 * 8);
#pragma unroll
                for (int r = 0; r < kStepRows; ++r) {
                    acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0, f[r + dy][dx], acc[r][0], 0, 0, 0);
                    acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, f[r + dy][dx], acc[r][1], 0, 0, 0);
                }
            }
        };
        for (int c = 0; c < nc; ++c) {
            bf16x8 f[kStepRows + 2][3];
            load_chunk(c, f);
            mfma_chunk(c, f);
        }
        // D layout: lane holds channels 4 gq .. 4 gq + 3 (of each half) of pixel x
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int c = co0 + nt * 16 + 4 * gq;
            const bool full_nt = c + 3 < a.cout;  // all four channels written
            float bias[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) bias[k] = a.bias && c + k < a.cout ? a.bias[c + k] : 0.0f;
            uint32_t top[2] = {0u, 0u};
#pragma unroll
            for (int r = 0; r < kStepRows; ++r) {
                const int y = y0 + r;
                const bool in = y < a.H && x < a.W;
                const uint32_t el = (uint32_t)((b * a.H + y) * a.W + x) * (uint32_t)a.out_pitch + (uint32_t)c;
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) v[k] = acc[r][nt][k] + bias[k];
                if (f32) {
                    if (full_nt) {
                        const u32x4 q = {__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]),
                                         __float_as_uint(v[3])};
                        __builtin_amdgcn_raw_buffer_store_b128(q, ro, in ? el * 4u : kOff, 0, 0);
                    } else {
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[k]), ro,
                                                                  in && c + k < a.cout ? (el + k) * 4u : kOff, 0, 0);
                    }
                    continue;
                }
                if (a.relu) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) v[k] = fmaxf(v[k], 0.0f);
                }
                const uint32_t lo = f2bf2(v[0], v[1]), hi = f2bf2(v[2], v[3]);
                if (full_nt) {
                    __builtin_amdgcn_raw_buffer_store_b64(u32x2{lo, hi}, ro, in ? el * 2u : kOff, 0, 0);
                } else {
                    const uint32_t qw[2] = {lo, hi};
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(qw[k >> 1] >> ((k & 1) * 16)), ro,
                                                              in && c + k < a.cout ? (el + k) * 2u : kOff, 0, 0);
                }
                if (a.pool_out) {
                    if ((r & 1) == 0) {
                        top[0] = lo;
                        top[1] = hi;
                    } else {
                        // the right-hand pixel's rows come from lane r16 ^ 1; even lanes store
                        const uint32_t p0 = pool4x2(top[0], lane_xor1(top[0]), lo, lane_xor1(lo));
                        const uint32_t p1 = pool4x2(top[1], lane_xor1(top[1]), hi, lane_xor1(hi));
                        const bool st = (r16 & 1) == 0 && y < a.H && x + 1 < a.W;
                        const uint32_t pe =
                            (uint32_t)((b * h2 + (y >> 1)) * w2 + (x >> 1)) * (uint32_t)a.out_pitch + (uint32_t)c;
                        if (full_nt) {
                            __builtin_amdgcn_raw_buffer_store_b64(u32x2{p0, p1}, rp, st ? pe * 2u : kOff, 0, 0);
                        } else {
                            const uint32_t pw[2] = {p0, p1};
#pragma unroll
                            for (int k = 0; k < 4; ++k)
                                __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(pw[k >> 1] >> ((k & 1) * 16)), rp,
                                                                      st && c + k < a.cout ? (pe + k) * 2u : kOff, 0, 0);
                        }
                    }
                }
            }
        }
    }
}

// Weight-gradient partial sums: part[split][tap][ci][co] = sum over this split's tiles of
// X(p + tap)[ci] * dZ(p)[co]. grid.y enumerates (32-ci block, 32-co block).
__global__ __launch_bounds__(256) void k_conv_wgrad(WgradArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t s_x[3][kCK][kHH][kTW];   // x-shifted, channel-major
    __shared__ __attribute__((aligned(16))) uint16_t s_z[kNT][kTH][kTW];      // channel-major
    const int cin = a.in.c0 + a.in.c1;
    const int nci = cin / kCK;
    const int ci0 = (blockIdx.y % nci) * kCK, co0 = (blockIdx.y / nci) * kNT;
    const int tiles_x = (a.W + kTW - 1) / kTW, tiles_y = (a.H + kTH - 1) / kTH;
    const int n_tiles = a.B * tiles_x * tiles_y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int cit = wave >> 1, cot = wave & 1;  // this wave's 16x16 (ci, co) block, all 9 taps
    f32x4 acc[9];
    for (int j = 0; j < 9; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        int t = tile;
        const int b = t / (tiles_x * tiles_y);
        t -= b * tiles_x * tiles_y;
        const int y0 = (t / tiles_x) * kTH, x0 = (t % tiles_x) * kTW;
        for (int v = threadIdx.x; v < kHH * kHW * 4; v += 256) {
            const int px = v >> 2, q = v & 3;
            const int hy = px / kHW, hx = px - hy * kHW;
            uint4 val = gather8(a.in, b, y0 + hy - 1, x0 + hx - 1, a.H, a.W, ci0 + q * 8);
            const uint16_t* h = reinterpret_cast<const uint16_t*>(&val);
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int x = hx - dx;
                if (x < 0 || x >= kTW) continue;
#pragma unroll
                for (int k = 0; k < 8; ++k) s_x[dx][q * 8 + k][hy][x] = h[k];
            }
        }
        for (int v = threadIdx.x; v < kTH * kTW * 4; v += 256) {
            const int px = v >> 2, q = v & 3;
            const int ty = px / kTW, tx = px - ty * kTW;
            const int y = y0 + ty, x = x0 + tx;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (y < a.H && x < a.W)
                val = *reinterpret_cast<const uint4*>(a.dz + (((size_t)b * a.H + y) * a.W + x) * a.coutp + co0 + q * 8);
            const uint16_t* h = reinterpret_cast<const uint16_t*>(&val);
#pragma unroll
            for (int k = 0; k < 8; ++k) s_z[q * 8 + k][ty][tx] = h[k];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int rowk = 2 * kk + (g >> 1), xk = 8 * (g & 1);
            const bf16x8 bfr = *reinterpret_cast<const bf16x8*>(&s_z[cot * 16 + r16][rowk][xk]);
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int dy = j / 3, dx = j - dy * 3;
                const bf16x8 afr = *reinterpret_cast<const bf16x8*>(&s_x[dx][cit * 16 + r16][rowk + dy][xk]);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr, acc[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float* out = a.part + (size_t)blockIdx.x * 9 * cin * a.coutp;
#pragma unroll
    for (int j = 0; j < 9; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            out[((size_t)j * cin + ci0 + cit * 16 + 4 * g + i) * a.coutp + co0 + cot * 16 + r16] = acc[j][i];
}

// Bias-gradient partial sums: part[split][c] = sum of dz over this split's pixels.
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* dz, long long n_px, int c, float* part) {
    __shared__ float s[256];
    const int lanes = 256 / c;  // c in {32, 64, 128}
    const int ch = threadIdx.x % c, pl = threadIdx.x / c;
    float sum = 0.0f;
    for (long long p = (long long)blockIdx.x * lanes + pl; p < n_px; p += (long long)gridDim.x * lanes)
        sum += bf2f(dz[p * c + ch]);
    s[threadIdx.x] = sum;
    __syncthreads();
    if (pl == 0) {
        for (int k = 1; k < lanes; ++k) sum += s[k * c + ch];
        part[(size_t)blockIdx.x * c + ch] = sum;
    }
}

// out[i] (=|+=) sum over s of part[s][i], in split order.
__global__ void k_reduce_splits(const float* part, int splits, long long n, float* out, int accumulate) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int k = 0; k < splits; ++k) s += part[(size_t)k * n + i];
        out[i] = accumulate ? out[i] + s : s;
    }
}

// Input features -> 32-channel bf16 network input (channels 13..31 zero), zero outside the
// frame (the network runs on a size rounded up to a multiple of 8).
// One padded pixel per thread: the 13 features (normalised), written as 32 bf16 channels
// with four 16-byte stores.
__global__ void k_prep(const float* feat, int B, int H, int W, int Hp, int Wp, uint16_t* x0) {
    const long long n = (long long)B * Hp * Wp;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % Wp);
        const int y = (int)((i / Wp) % Hp);
        const int b = (int)(i / ((long long)Wp * Hp));
        float v[kFeatIn];
        if (y < H && x < W) {
            const float* f = feat + (((size_t)b * H + y) * W + x) * kFeatIn;
#pragma unroll
            for (int c = 0; c < kFeatIn; ++c) v[c] = f[c];
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = fminf(fmaxf(v[c], 0.0f), 4.0f);  // radiance
            v[9] = log1pf(fmaxf(v[9], 0.0f)) * 0.25f;                             // depth
            v[11] = sqrtf(fmaxf(v[11], 0.0f)) * 4.0f;                             // visibility std
        } else {
#pragma unroll
            for (int c = 0; c < kFeatIn; ++c) v[c] = 0.0f;
        }
        uint32_t w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float lo = 2 * k < kFeatIn ? v[2 * k] : 0.0f, hi = 2 * k + 1 < kFeatIn ? v[2 * k + 1] : 0.0f;
            w[k] = (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
        }
        uint4* o = reinterpret_cast<uint4*>(x0 + i * 32);
        o[0] = make_uint4(w[0], w[1], w[2], w[3]);
        o[1] = make_uint4(w[4], w[5], w[6], w[7]);
        o[2] = make_uint4(0, 0, 0, 0);
        o[3] = make_uint4(0, 0, 0, 0);
    }
}

// 2x2 average pooling, 8 channels (one 16-byte vector) of one output pixel per thread.
__global__ void k_pool_bwd(const float* gh, int B, int H, int W, int C, float* gf) {
    const long long n = (long long)B * H * W * C;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long long p = i / C;
        const int x = (int)(p % W);
        const int y = (int)((p / W) % H);
        const int b = (int)(p / ((long long)W * H));
        gf[i] += 0.25f * gh[(((size_t)b * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * C + c];
    }
}

// Gradient of concat(up(low), skip): g_low = sum over each 2x2 block of dx[..., :c0],
// g_skip = dx[..., c0:] (written, it is the skip's first contribution).
__global__ void k_split(const float* dx, int B, int H, int W, int c0, int c1, float* g_low, float* g_skip) {
    const int C = c0 + c1;
    const long long n_skip = (long long)B * H * W * c1;
    const int h2 = H >> 1, w2 = W >> 1;
    const long long n_low = (long long)B * h2 * w2 * c0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_skip + n_low;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < n_skip) {
            const long long p = i / c1;
            g_skip[i] = dx[p * C + c0 + (i % c1)];
        } else {
            const long long j = i - n_skip;
            const int c = (int)(j % c0);
            long long p = j / c0;
            const int x = (int)(p % w2);
            const int y = (int)((p / w2) % h2);
            const int b = (int)(p / ((long long)w2 * h2));
            const float* q = dx + (((size_t)b * H + 2 * y) * W + 2 * x) * C + c;
            g_low[j] = (q[0] + q[C]) + (q[(size_t)W * C] + q[(size_t)W * C + C]);
        }
    }
}

// dz = a > 0 ? bf16(g) : 0 (ReLU backward; `a` is the stored post-ReLU activation)
__global__ void k_relu_bwd(const float* g, const uint16_t* a, long long n, uint16_t* dz) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dz[i] = bf2f(a[i]) > 0.0f ? f2bf(g[i]) : (uint16_t)0;
}

// Residual output, loss and its gradient. y: [B][Hp][Wp][3] network output; the result is
// noisy + y on the H x W frame. dz: [B][Hp][Wp][32] bf16 = dLoss/dy in channels 0..2
// (loss = mean squared error over B*H*W*3 values); part[block] = partial sum of squares.
__global__ __launch_bounds__(256) void k_loss(const float* y, const float* feat, const float* target, int B, int H,
                                              int W, int Hp, int Wp, uint16_t* dz, float* part) {
    __shared__ float s[256];
    const long long n = (long long)B * Hp * Wp;
    const float scale = 2.0f / (float)((long long)B * H * W * 3);
    float acc = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % Wp);
        const int yy = (int)((i / Wp) % Hp);
        const int b = (int)(i / ((long long)Wp * Hp));
        uint16_t* d = dz + i * 32;
        for (int c = 3; c < 32; ++c) d[c] = 0;
        if (yy < H && x < W) {
            const size_t p = ((size_t)b * H + yy) * W + x;
            for (int c = 0; c < 3; ++c) {
                const float diff = (feat[p * kFeatIn + c] + y[i * 3 + c]) - target[p * 3 + c];
                acc += diff * diff;
                d[c] = f2bf(scale * diff);
            }
        } else {
            d[0] = d[1] = d[2] = 0;
        }
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Denoised frame: rgb = noisy + y (f32 and/or quantised like create_image_from_data).
// skip (optional): T x T tile flags; the 8-bit output of a flagged tile is left alone (the
// tile-masked fallback render writes it).
__global__ void k_compose(const float* y, const float* feat, int H, int W, int Wp, float* out_f32, uint8_t* out_u8,
                          const uint8_t* skip, int T) {
    const long long n = (long long)H * W;
    const int tx = skip ? (W + T - 1) / T : 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % W), yy = (int)(i / W);
        const size_t q = ((size_t)yy * Wp + x) * 3;
        const bool keep = skip && skip[(yy / T) * tx + x / T];
        for (int c = 0; c < 3; ++c) {
            const float v = feat[i * kFeatIn + c] + y[q + c];
            if (out_f32) out_f32[i * 3 + c] = v;
            if (out_u8 && !keep) {
                const float cl = fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f;
                out_u8[i * 3 + c] = (uint8_t)(cl >= 255.0f ? 255 : (int)cl);
            }
        }
    }
}

// Per-tile fallback decision: a tile is re-rendered when the mean over its pixels of the
// binomial variance of the visibility estimate exceeds `threshold`.
__global__ __launch_bounds__(256) void k_tile_mask(const float* feat, int H, int W, int T, float threshold,
                                                   uint8_t* mask, int* n_flagged) {
    __shared__ float s[256];
    const int tx = (W + T - 1) / T;
    const int x0 = (blockIdx.x % tx) * T, y0 = (blockIdx.x / tx) * T;
    float acc = 0.0f;
    for (int k = threadIdx.x; k < T * T; k += 256) {
        const int x = x0 + k % T, y = y0 + k / T;
        if (x < W && y < H) acc += feat[((size_t)y * W + x) * kFeatIn + 11];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int w = min(T, W - x0), h = min(T, H - y0);
        const bool f = s[0] / (float)(w * h) > threshold;
        mask[blockIdx.x] = f ? 1 : 0;
        if (f) atomicAdd(n_flagged, 1);
    }
}

// Adam (Kingma & Ba 2015) on the f32 master parameters.
__global__ void k_adam(float* w, const float* grad, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = grad[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

// Master [9][cin][coutp] f32 -> forward bf16 [9][coutp][cin] and backward-data bf16
// [9][cin][coutp] with the taps flipped (tap t reads master tap 8 - t).
__global__ void k_pack(const float* w, int cin, int coutp, uint16_t* wf, uint16_t* wb) {
    const long long n = 9ll * cin * coutp;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int co = (int)(i % coutp);
        const int ci = (int)((i / coutp) % cin);
        const int tap = (int)(i / ((long long)coutp * cin));
        const uint16_t h = f2bf(w[i]);
        wf[((size_t)tap * coutp + co) * cin + ci] = h;
        if (wb) wb[((size_t)(8 - tap) * cin + ci) * coutp + co] = h;
    }
}

// ------------------------------------------------------------------ launchers
namespace {
int grid1(long long n, int block, int cap) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (int)(g < cap ? g : cap);
}
}  // namespace

// Workgroups of k_conv3x3 resident on the device at once (every CU at the kernel's
// occupancy), queried once per device.
int conv_resident() {
    static int cached[64] = {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) dev = 0;
    if (cached[dev] == 0) {
        int cus = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3x3<kHaloDepth>, 256, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        cached[dev] = cus * per_cu;
    }
    return cached[dev];
}
// Rows per strip and row register sets of k_conv3x3_c32 (217 VGPRs, 2 wavefronts per SIMD).
// 1080p forward pass: 1.34 ms with 4 sets (rows 1 ahead), 1.33 with 6, 8 or 10 sets or 32
// rows per strip; 1.46 ms on the LDS kernel. Held to 3 wavefronts per SIMD, 4 sets spilled
// 20 registers and took 1.62 ms.
constexpr int kC32Rows = 16, kC32Slots = 8;
// Workgroups of k_conv3x3_c32 resident on the device at once, queried once per device.
int c32_resident() {
    static int cached[64] = {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) dev = 0;
    if (cached[dev] == 0) {
        int cus = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3x3_c32<kC32Rows, kC32Slots>, 256, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        cached[dev] = cus * per_cu;
    }
    return cached[dev];
}

// Row-step kernel: rows per step and wavefronts per workgroup (167 VGPRs, 3 per SIMD), and
// the pixels a launch needs for it. It runs the full-resolution layers with several K chunks:
// 1080p dec0a (96 -> 32) 196 us on the tile kernel, 153 us here. At half resolution and
// below it was slower (1080p dec1a 152 -> 183 us, dec2a 101 -> 132 us, bota 20 -> 33 us),
// also with the next K chunk's rows loaded during the current chunk's MFMAs (245 VGPRs, 8
// wavefronts: dec1a 170 us, dec0a 160 us) and with 2-row steps (forward pass 1.51-1.60 ms
// instead of 1.39).
constexpr int kRowsStep = 4, kRowsWaveCount = 12;
constexpr unsigned long long kRowsMinPixels = 1ull << 20;
// Compute units of the current device, queried once per device.
int cu_count() {
    static int cached[64] = {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) dev = 0;
    if (cached[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
        // the row-step kernel's weights take up to 8 K chunks of LDS
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_conv3x3_rows<kRowsStep, kRowsWaveCount>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRowsMaxLds);
        cached[dev] = cus;
    }
    return cached[dev];
}

hipError_t launch_conv(const ConvArgs& a, hipStream_t st) {
    const int tiles = a.B * ((a.H + kTH - 1) / kTH) * ((a.W + kTW - 1) / kTW);
    const int ny = a.coutp / kNT;
    // Several K chunks: the row-step kernel for launches of 2^20 pixels and more, else one
    // tile per workgroup. One K chunk: the strip kernel, or, with sources or outputs of
    // 2 GiB and more (beyond a buffer load's 32-bit offsets), the persistent LDS kernel.
    const unsigned long long px = (unsigned long long)a.B * a.H * a.W;
    const unsigned long long out_bytes = px * a.out_pitch * (a.out_f32 ? 4 : 2);
    const unsigned long long pool_bytes = a.pool_out ? px / 4 * a.out_pitch * 2 : 0;
    const int cin = a.in.c0 + a.in.c1;
    if (cin != kCK) {
        const unsigned long long bytes0 =
            (unsigned long long)a.B * (a.in.up0 ? (unsigned long long)(a.H >> 1) * (a.W >> 1) : px / a.B) * a.in.c0 * 2;
        const unsigned long long bytes1 = px * a.in.c1 * 2;
        const size_t lds = (size_t)(cin / kCK) * 9 * kNT * kCK * 2;
        if (px >= kRowsMinPixels && lds <= kRowsMaxLds && bytes0 < (1ull << 31) && bytes1 < (1ull << 31) &&
            out_bytes < (1ull << 31) && pool_bytes < (1ull << 31)) {
            const int sx = (a.W + kStripW - 1) / kStripW, sy = (a.H + kRowsStep - 1) / kRowsStep;
            const long long steps = (long long)a.B * sx * sy;
            const int nx = (int)std::max(1LL, std::min((steps + kRowsWaveCount - 1) / kRowsWaveCount,
                                                       (long long)std::max(1, cu_count() / ny)));
            hipLaunchKernelGGL((k_conv3x3_rows<kRowsStep, kRowsWaveCount>), dim3(nx, ny),
                               dim3(kRowsWaveCount * 64), lds, st, a, sx, sy, (unsigned)bytes0, (unsigned)bytes1,
                               (unsigned)out_bytes, (unsigned)pool_bytes);
            return hipGetLastError();
        }
        hipLaunchKernelGGL(k_conv3x3_tile, dim3(tiles, ny), dim3(256), 0, st, a);
        return hipGetLastError();
    }
    const bool first = a.in.c0 > 0;
    const unsigned long long src_bytes =
        (first && a.in.up0 ? (unsigned long long)a.B * (a.H >> 1) * (a.W >> 1) : px) * kCK * 2;
    if (src_bytes < (1ull << 31) && out_bytes < (1ull << 31) && pool_bytes < (1ull << 31)) {
        const int sx = (a.W + kStripW - 1) / kStripW, sy = (a.H + kC32Rows - 1) / kC32Rows;
        const long long strips = (long long)a.B * sx * sy;
        const int nx = (int)std::max(1LL, std::min((strips + 3) / 4, (long long)std::max(1, c32_resident() / ny)));
        hipLaunchKernelGGL((k_conv3x3_c32<kC32Rows, kC32Slots>), dim3(nx, ny), dim3(256), 0, st, a, sx, sy,
                           (unsigned)src_bytes, (unsigned)out_bytes, (unsigned)pool_bytes);
        return hipGetLastError();
    }
    const int nx = std::max(1, std::min(tiles, conv_resident() / ny));
    hipLaunchKernelGGL(k_conv3x3<kHaloDepth>, dim3(nx, ny), dim3(256), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_wgrad(const WgradArgs& a, int splits, hipStream_t st) {
    const int cin = a.in.c0 + a.in.c1;
    hipLaunchKernelGGL(k_conv_wgrad, dim3(splits, (cin / kCK) * (a.coutp / kNT)), dim3(256), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_bias_grad(const uint16_t* dz, long long n_px, int c, float* part, int splits, hipStream_t st) {
    hipLaunchKernelGGL(k_bias_grad, dim3(splits), dim3(256), 0, st, dz, n_px, c, part);
    return hipGetLastError();
}
hipError_t launch_reduce(const float* part, int splits, long long n, float* out, bool acc, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_splits, dim3(grid1(n, 256, 4096)), dim3(256), 0, st, part, splits, n, out, acc ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_prep(const float* feat, int B, int H, int W, int Hp, int Wp, uint16_t* x0, hipStream_t st) {
    hipLaunchKernelGGL(k_prep, dim3(grid1((long long)B * Hp * Wp, 256, 8192)), dim3(256), 0, st, feat, B, H, W, Hp, Wp,
                       x0);
    return hipGetLastError();
}
hipError_t launch_pool_bwd(const float* gh, int B, int H, int W, int C, float* gf, hipStream_t st) {
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid1((long long)B * H * W * C, 256, 8192)), dim3(256), 0, st, gh, B, H, W, C,
                       gf);
    return hipGetLastError();
}
hipError_t launch_split(const float* dx, int B, int H, int W, int c0, int c1, float* g_low, float* g_skip,
                        hipStream_t st) {
    const long long n = (long long)B * H * W * c1 + (long long)B * (H / 2) * (W / 2) * c0;
    hipLaunchKernelGGL(k_split, dim3(grid1(n, 256, 8192)), dim3(256), 0, st, dx, B, H, W, c0, c1, g_low, g_skip);
    return hipGetLastError();
}
hipError_t launch_relu_bwd(const float* g, const uint16_t* a, long long n, uint16_t* dz, hipStream_t st) {
    hipLaunchKernelGGL(k_relu_bwd, dim3(grid1(n, 256, 8192)), dim3(256), 0, st, g, a, n, dz);
    return hipGetLastError();
}
hipError_t launch_loss(const float* y, const float* feat, const float* target, int B, int H, int W, int Hp, int Wp,
                       uint16_t* dz, float* part, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_loss, dim3(blocks), dim3(256), 0, st, y, feat, target, B, H, W, Hp, Wp, dz, part);
    return hipGetLastError();
}
hipError_t launch_compose(const float* y, const float* feat, int H, int W, int Wp, float* out_f32, uint8_t* out_u8,
                          const uint8_t* skip, int T, hipStream_t st) {
    if (skip && T <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_compose, dim3(grid1((long long)H * W, 256, 8192)), dim3(256), 0, st, y, feat, H, W, Wp,
                       out_f32, out_u8, skip, T);
    return hipGetLastError();
}
hipError_t launch_tile_mask(const float* feat, int H, int W, int T, float threshold, uint8_t* mask, int* n_flagged,
                            hipStream_t st) {
    const int tiles = ((W + T - 1) / T) * ((H + T - 1) / T);
    hipLaunchKernelGGL(k_tile_mask, dim3(tiles), dim3(256), 0, st, feat, H, W, T, threshold, mask, n_flagged);
    return hipGetLastError();
}
hipError_t launch_adam(float* w, const float* grad, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, hipStream_t st) {
    hipLaunchKernelGGL(k_adam, dim3(grid1(n, 256, 4096)), dim3(256), 0, st, w, grad, m, v, n, lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError();
}
hipError_t launch_pack(const float* w, int cin, int coutp, uint16_t* wf, uint16_t* wb, hipStream_t st) {
    hipLaunchKernelGGL(k_pack, dim3(grid1(9ll * cin * coutp, 256, 4096)), dim3(256), 0, st, w, cin, coutp, wf, wb);
    return hipGetLastError();
}

}  // namespace dn
}  // namespace rt
