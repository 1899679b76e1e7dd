// device_math.hpp — device-side building blocks of the trace path:
//   * the reference's RNG (rand 0.8 StdRng = ChaCha12, seeded by rand_core 0.6's PCG32
//     seed_from_u64) as a counter-based generator with random access by word index;
//   * exact f64 primitive tests with the reference's operation order (used for the
//     analytic objects and to re-test FP32 triangle survivors);
//   * FP32 prefilters (conservative) for spheres, triangles and boxes.
// Compiled with -ffp-contract=off: every f64 op is rounded separately, as in Rust.
#pragma once

#include <hip/hip_runtime.h>

#include "../common/layout.hpp"
#include "../common/vec.hpp"

namespace rt {

// ---------------------------------------------------------------- bit casts
RT_HD double bits_to_f64(uint64_t b) {
    union { uint64_t u; double d; } x;
    x.u = b;
    return x.d;
}
RT_HD uint64_t f64_to_bits(double d) {
    union { uint64_t u; double d; } x;
    x.d = d;
    return x.u;
}
RT_HD float bits_to_f32(uint32_t b) {
    union { uint32_t u; float f; } x;
    x.u = b;
    return x.f;
}
RT_HD uint32_t f32_to_bits(float f) {
    union { uint32_t u; float f; } x;
    x.f = f;
    return x.u;
}

// ---------------------------------------------------------------- RNG
RT_HD uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

#define RT_QR(a, b, c, d)           \
    a += b; d ^= a; d = rotl32(d, 16); \
    c += d; b ^= c; b = rotl32(b, 12); \
    a += b; d ^= a; d = rotl32(d, 8);  \
    c += d; b ^= c; b = rotl32(b, 7);

// ChaCha12 block `blk` of key k (words 12-13 = 64-bit block counter, 14-15 = stream 0),
// as rand_chacha 0.3 lays out its buffer (src/lighting.rs:166 and renderer seeds use it).
RT_HD void chacha12_block(const uint32_t k[8], uint64_t blk, uint32_t out[16]) {
    uint32_t x0 = 0x61707865u, x1 = 0x3320646eu, x2 = 0x79622d32u, x3 = 0x6b206574u;
    uint32_t x4 = k[0], x5 = k[1], x6 = k[2], x7 = k[3], x8 = k[4], x9 = k[5], x10 = k[6], x11 = k[7];
    uint32_t x12 = (uint32_t)blk, x13 = (uint32_t)(blk >> 32), x14 = 0, x15 = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        RT_QR(x0, x4, x8, x12);
        RT_QR(x1, x5, x9, x13);
        RT_QR(x2, x6, x10, x14);
        RT_QR(x3, x7, x11, x15);
        RT_QR(x0, x5, x10, x15);
        RT_QR(x1, x6, x11, x12);
        RT_QR(x2, x7, x8, x13);
        RT_QR(x3, x4, x9, x14);
    }
    out[0] = x0 + 0x61707865u; out[1] = x1 + 0x3320646eu; out[2] = x2 + 0x79622d32u; out[3] = x3 + 0x6b206574u;
    out[4] = x4 + k[0]; out[5] = x5 + k[1]; out[6] = x6 + k[2]; out[7] = x7 + k[3];
    out[8] = x8 + k[4]; out[9] = x9 + k[5]; out[10] = x10 + k[6]; out[11] = x11 + k[7];
    out[12] = x12 + (uint32_t)blk; out[13] = x13 + (uint32_t)(blk >> 32); out[14] = x14; out[15] = x15;
}

// rand_core 0.6 seed_from_u64: 8 PCG32 (XSH-RR) outputs, state advanced first.
RT_HD void stdrng_key(uint64_t state, uint32_t k[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        state = state * 6364136223846793005ull + 11634580027462260723ull;
        uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        k[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
}

// Sequential StdRng reader over the counter-based block function.
struct StdRngDev {
    uint32_t key[8];
    uint32_t buf[16];
    uint64_t blk;  // next block to generate
    int idx;       // next word in buf (16 = empty)

    RT_HD void init(uint64_t seed, uint64_t first_word = 0) {
        stdrng_key(seed, key);
        blk = first_word >> 4;
        idx = 16;
        if (first_word & 15) {
            chacha12_block(key, blk, buf);
            ++blk;
            idx = (int)(first_word & 15);
        }
    }
    RT_HD uint32_t next_u32() {
        if (idx >= 16) {
            chacha12_block(key, blk, buf);
            ++blk;
            idx = 0;
        }
        // select from the register buffer without dynamic VGPR indexing
        uint32_t v = buf[0];
#pragma unroll
        for (int i = 1; i < 16; ++i) v = (idx == i) ? buf[i] : v;
        ++idx;
        return v;
    }
    RT_HD uint64_t next_u64() {
        uint64_t lo = next_u32();
        uint64_t hi = next_u32();
        return lo | (hi << 32);
    }
    // rand 0.8 Standard f64: (u64 >> 11) * 2^-53
    RT_HD double gen_f64() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
    // rand 0.8 UniformFloat::<f64>::sample_single(low, high)
    RT_HD double gen_range(double low, double high) {
        double scale = high - low;
        for (;;) {
            uint64_t bits = (next_u64() >> 12) | (1023ull << 52);
            double v01 = bits_to_f64(bits) - 1.0;
            double res = v01 * scale + low;
            if (res < high) return res;
            scale = bits_to_f64(f64_to_bits(scale) - 1);
        }
    }
};

// ---------------------------------------------------------------- rays
struct RayD {
    d3 o, d;  // d is unit (Ray::new normalizes, src/ray.rs:13-18)
};

// Cold f64 paths stay out of line on the device so the traversal loops keep their
// register budget (they run for a small fraction of the tests).
#if defined(__HIP_DEVICE_COMPILE__)
#define RT_COLD __host__ __device__ __attribute__((noinline))
#elif defined(__HIPCC__)
#define RT_COLD __host__ __device__ inline
#else
#define RT_COLD inline
#endif
// Hot traversal steps are inlined into their kernels so per-ray state stays in registers.
#if defined(__HIPCC__)
#define RT_HOT __host__ __device__ inline __attribute__((always_inline))
#else
#define RT_HOT inline
#endif

// Loads from device memory through a pointer the compiler knows only as generic (one read
// from memory, such as a mesh's node and triangle arrays): plain dereferences compile to
// flat loads, which also count against the LDS wait counter and take 64-bit addresses;
// these load through the global address space instead. On the host, plain loads.
#if defined(__HIP_DEVICE_COMPILE__)
#define RT_AS_GLOBAL __attribute__((address_space(1)))
typedef float rt_f4v __attribute__((ext_vector_type(4)));
typedef uint32_t rt_u4v __attribute__((ext_vector_type(4)));
RT_HOT float4 ldg(const float4* p) {
    const rt_f4v v = *(const RT_AS_GLOBAL rt_f4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
RT_HOT uint4 ldg(const uint4* p) {
    const rt_u4v v = *(const RT_AS_GLOBAL rt_u4v*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
RT_HOT double ldg(const double* p) { return *(const RT_AS_GLOBAL double*)p; }
RT_HOT uint32_t ldg(const uint32_t* p) { return *(const RT_AS_GLOBAL uint32_t*)p; }
#else
template <class T>
RT_HOT T ldg(const T* p) {
    return *p;
}
#endif

// ---------------------------------------------------------------- exact f64 tests
// Sphere::hit (src/ray.rs:89-120); t_max inclusive.
RT_HD bool sphere_exact(const ObjDev& s, const RayD& r, double tmin, double tmax, double& t) {
    d3 oc = sub(r.o, mk(s.c[0], s.c[1], s.c[2]));
    double a = dot(r.d, r.d);
    double hb = dot(oc, r.d);
    double c = dot(oc, oc) - s.r * s.r;
    double disc = hb * hb - a * c;
    if (disc < 0.0) return false;
    double sq = sqrt(disc);
    double root = (-hb - sq) / a;
    if (root < tmin || tmax < root) {
        root = (-hb + sq) / a;
        if (root < tmin || tmax < root) return false;
    }
    t = root;
    return true;
}
// Plane::hit (src/ray.rs:136-176) without the uv part.
RT_HD bool plane_exact(const ObjDev& p, const RayD& r, double tmin, double tmax, double& t) {
    d3 n = mk(p.n[0], p.n[1], p.n[2]);
    double denom = dot(n, r.d);
    if (fabs(denom) < 1e-8) return false;
    double tt = dot(sub(mk(p.c[0], p.c[1], p.c[2]), r.o), n) / denom;
    if (tt < tmin || tt > tmax) return false;
    t = tt;
    return true;
}
// Cube::hit (src/ray.rs:410-491). Returns the local-space t (compared against world t's
// by World::hit, as the reference does), the world hit point and the world normal
// before face orientation.
RT_HD bool cube_exact(const ObjDev& cb, const RayD& r, double tmin, double tmax, double& t, d3& wp,
                                  d3& wn) {
    d3 c = mk(cb.c[0], cb.c[1], cb.c[2]);
    d3 lo = m4_apply(cb.w2l, sub(r.o, c), 0.0);
    d3 ld0 = m4_apply(cb.w2l, r.d, 0.0);
    if (length(ld0) < 1e-8) return false;
    d3 ld = unit(ld0);
    double tmin_hit = tmin, tmax_hit = tmax;
    d3 normal = mk(0.0, 0.0, 0.0);
    for (int axis = 0; axis < 3; ++axis) {
        double da = axis == 0 ? ld.x : (axis == 1 ? ld.y : ld.z);
        double oa = axis == 0 ? lo.x : (axis == 1 ? lo.y : lo.z);
        double h = cb.half[axis];
        double inv = 1.0 / da;
        double t0 = (-h - oa) * inv;
        double t1 = (h - oa) * inv;
        double an = inv < 0.0 ? 1.0 : -1.0;
        if (inv < 0.0) {
            double tmp = t0;
            t0 = t1;
            t1 = tmp;
            an = -an;
        }
        if (t0 > tmin_hit) {
            tmin_hit = t0;
            normal = mk(axis == 0 ? an : 0.0, axis == 1 ? an : 0.0, axis == 2 ? an : 0.0);
        }
        if (t1 < tmax_hit) tmax_hit = t1;
        if (tmin_hit > tmax_hit) return false;
    }
    double tt = tmin_hit > tmin ? tmin_hit : tmax_hit;
    if (tt < tmin || tt > tmax) return false;
    d3 lp = add(lo, scale(tt, ld));
    wp = add(c, m4_apply(cb.l2w, lp, 1.0));
    if (cb.t_identity) {
        wn = normal;
    } else {
        const double* m = cb.ntf;  // row-major 3x3, gemv order
        double x[3] = {normal.x, normal.y, normal.z}, y[3];
        for (int row = 0; row < 3; ++row) {
            double acc = m[row * 3 + 0] * x[0];
            acc = m[row * 3 + 1] * x[1] + acc;
            acc = m[row * 3 + 2] * x[2] + acc;
            y[row] = acc;
        }
        wn = mk(y[0], y[1], y[2]);
    }
    t = tt;
    return true;
}
// intersect_triangle (src/ray.rs:527-583): strict t_min < t < t_max.
RT_HD bool tri_exact(const double* v, const RayD& r, double tmin, double tmax, double& t, d3& n, double& u,
                                 double& vv) {
    d3 v0 = mk(ldg(v + 0), ldg(v + 1), ldg(v + 2));
    d3 e1 = sub(mk(ldg(v + 3), ldg(v + 4), ldg(v + 5)), v0);
    d3 e2 = sub(mk(ldg(v + 6), ldg(v + 7), ldg(v + 8)), v0);
    d3 h = cross(r.d, e2);
    double a = dot(e1, h);
    if (a > -1e-8 && a < 1e-8) return false;
    double f = 1.0 / a;
    d3 s = sub(r.o, v0);
    double uu = f * dot(s, h);
    if (!(uu >= 0.0 && uu <= 1.0)) return false;
    d3 q = cross(s, e1);
    double w = f * dot(r.d, q);
    if (w < 0.0 || uu + w > 1.0) return false;
    double tt = f * dot(e2, q);
    if (!(tt > tmin && tt < tmax)) return false;
    d3 nn = cross(e1, e2);
    if (length(nn) < 1e-8) return false;
    if (a < 0.0) nn = neg(nn);
    n = unit(nn);
    t = tt;
    u = uu;
    vv = w;
    return true;
}

// ---------------------------------------------------------------- FP32 prefilters
struct RayF {
    float ox, oy, oz;  // origin (absolute; meshes subtract their centre)
    float dx, dy, dz;
    float ix, iy, iz;  // 1/d
};

RT_HD RayF make_rayf(const RayD& r) {
    RayF f;
    f.ox = (float)r.o.x; f.oy = (float)r.o.y; f.oz = (float)r.o.z;
    f.dx = (float)r.d.x; f.dy = (float)r.d.y; f.dz = (float)r.d.z;
    f.ix = 1.0f / f.dx; f.iy = 1.0f / f.dy; f.iz = 1.0f / f.dz;
    return f;
}

// Conservative: false only when the f64 Sphere::hit cannot return a root in [tmin, tmax].
RT_HD bool sphere_may_hit(const ObjDev& s, const RayD& r, const RayF& f, double tmin, double tmax) {
    float ox = (float)(r.o.x - s.c[0]), oy = (float)(r.o.y - s.c[1]), oz = (float)(r.o.z - s.c[2]);
    float b = ox * f.dx + oy * f.dy + oz * f.dz;
    float qx = ox - b * f.dx, qy = oy - b * f.dy, qz = oz - b * f.dz;
    float d2 = qx * qx + qy * qy + qz * qz;
    float oo = ox * ox + oy * oy + oz * oz;
    float r2 = s.rf * s.rf;
    float tol2 = 1e-5f * (oo + r2) + 1e-20f;
    if (d2 > r2 + tol2) return false;
    float tol = 1e-5f * (fabsf(b) + s.rf) + 1e-6f;
    if (-b + s.rf < (float)tmin - tol) return false;
    if (-b - s.rf > (float)tmax + tol) return false;
    return true;
}

// Conservative FP32 prefilter of plane_exact: false only when Plane::hit rejects for certain,
// i.e. |n.d| < 1e-8, or t = n.(c - o) / n.d lies below tmin (0 <= tmin: numerator and
// denominator of opposite signs) or above tmax for every value the f32 inputs allow. The f32
// copies of the point, normal, origin and direction carry 2^-24 of their magnitudes, and three
// products and two sums add a few more rounding units: 1e-6 of the magnitudes (about 16 units)
// bounds the error of the numerator and the denominator.
RT_HD bool plane_may_hit(const ObjDev& p, const RayF& f, double tmin, double tmax) {
    const float nx = p.pn[0], ny = p.pn[1], nz = p.pn[2];
    const float den = nx * f.dx + ny * f.dy + nz * f.dz;
    const float eden = 1e-6f * (fabsf(nx * f.dx) + fabsf(ny * f.dy) + fabsf(nz * f.dz)) + 1e-30f;
    if (fabsf(den) + eden < 1e-8f) return false;  // 1e-8f is below the f64 1e-8
    const float vx = p.pc[0] - f.ox, vy = p.pc[1] - f.oy, vz = p.pc[2] - f.oz;
    const float num = nx * vx + ny * vy + nz * vz;
    const float enm = 1e-6f * (fabsf(nx) * (fabsf(p.pc[0]) + fabsf(f.ox)) + fabsf(ny) * (fabsf(p.pc[1]) + fabsf(f.oy)) +
                               fabsf(nz) * (fabsf(p.pc[2]) + fabsf(f.oz))) +
                      1e-30f;
    if (tmin >= 0.0) {
        if (num + enm < 0.0f && den - eden > 0.0f) return false;
        if (num - enm > 0.0f && den + eden < 0.0f) return false;
    }
    if (tmax < 1e30) {
        const bool same = (num - enm > 0.0f && den - eden > 0.0f) || (num + enm < 0.0f && den + eden < 0.0f);
        const float lo = fabsf(num) - enm, hi = (fabsf(den) + eden) * (float)tmax * (1.0f + 1e-6f);
        if (same && lo > hi) return false;
    }
    return true;
}

}  // namespace rt
