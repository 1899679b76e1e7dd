// vec.hpp — f64 vector algebra shared by the host scene builder and the device
// kernels. Operation order follows nalgebra as the reference uses it, so that host
// and device produce the same bits as the Rust renderer: dot = (x*x' + y*y') + z*z',
// norm = sqrt(dot), new_normalize divides each component by the norm
// (nalgebra Unit::new_normalize), cross = (ay*bz - az*by, az*bx - ax*bz, ax*by - ay*bx).
// Every file that includes this is compiled with -ffp-contract=off.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define RT_HD __host__ __device__ inline
#else
#define RT_HD inline
#endif

namespace rt {

struct d3 {
    double x, y, z;
};

RT_HD d3 mk(double x, double y, double z) { return d3{x, y, z}; }
RT_HD d3 add(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RT_HD d3 sub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RT_HD d3 neg(d3 a) { return {-a.x, -a.y, -a.z}; }
RT_HD d3 scale(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
RT_HD d3 divs(d3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
RT_HD d3 hadamard(d3 a, d3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
RT_HD double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RT_HD d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
RT_HD double length(d3 a) { return sqrt(dot(a, a)); }
RT_HD d3 unit(d3 a) {
    double n = length(a);
    return {a.x / n, a.y / n, a.z / n};
}
RT_HD double get(d3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// f64::max / f64::min: a NaN operand is ignored (IEEE maxNum/minNum)
RT_HD double fmax_r(double a, double b) { return fmax(a, b); }
RT_HD double fmin_r(double a, double b) { return fmin(a, b); }
// f64::clamp keeps NaN
RT_HD double clamp_r(double x, double lo, double hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
// `as u8` / `as u64` / `as i32` casts saturate and map NaN to 0
RT_HD uint8_t to_u8_sat(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 255.0) return 255;
    return (uint8_t)x;
}
RT_HD uint64_t to_u64_sat(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 18446744073709551616.0) return ~0ull;
    return (uint64_t)x;
}
RT_HD int32_t to_i32_sat(double x) {
    if (x != x) return 0;
    if (x <= -2147483648.0) return (int32_t)0x80000000u;
    if (x >= 2147483647.0) return 2147483647;
    return (int32_t)x;
}
// f64::fract
RT_HD double fract_r(double x) { return x - trunc(x); }

// 4x4 column-major (m[col*4+row]) times (v, w), in nalgebra's gemv order:
// y_r = ((M_r0 v0 + M_r1 v1) + M_r2 v2) + M_r3 w
RT_HD d3 m4_apply(const double* m, d3 v, double w) {
    double in[4] = {v.x, v.y, v.z, w};
    double out[3];
    for (int r = 0; r < 3; ++r) {
        double acc = m[0 * 4 + r] * in[0];
        acc = m[1 * 4 + r] * in[1] + acc;
        acc = m[2 * 4 + r] * in[2] + acc;
        acc = m[3 * 4 + r] * in[3] + acc;
        out[r] = acc;
    }
    return {out[0], out[1], out[2]};
}

}  // namespace rt
