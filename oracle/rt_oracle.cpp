// rt_oracle.cpp — CPU f64 restatement of rtrace's per-pixel trace path.
//
// TEST INFRASTRUCTURE ONLY (the parity oracle): tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg are the only users. Nothing in rtrace_amd/
// links or calls this file.
//
// Built with -O2 -ffp-contract=off -fno-fast-math so that every f64 operation
// is a single IEEE-rounded op in the same order as the Rust reference (which
// never contracts to FMA). glibc libm provides sin/cos/tan/exp/pow, the same
// library Rust's std calls on linux-gnu.
//
// Citations are to the reference repository SeaRyanC/rtrace.

#include "rt_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

using u64 = uint64_t;
using u32 = uint32_t;

struct Err : std::runtime_error {
    int code;
    Err(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---------------------------------------------------------------------------
// Vector algebra with nalgebra's evaluation order (dot = (x*x' + y*y') + z*z';
// norm = sqrt(dot(v,v)); new_normalize divides each component by the norm).
// ---------------------------------------------------------------------------
struct V3 {
    double x, y, z;
};
inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 cmul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
    double n = norm(a);
    return {a.x / n, a.y / n, a.z / n};
}
inline double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline void setc(V3& a, int i, double v) { (i == 0 ? a.x : (i == 1 ? a.y : a.z)) = v; }
// f64::max / f64::min ignore a NaN operand (IEEE maxNum), like fmax/fmin.
inline double rmax(double a, double b) { return std::fmax(a, b); }
inline double rmin(double a, double b) { return std::fmin(a, b); }
// f64::clamp: NaN stays NaN.
inline double rclamp(double x, double lo, double hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
// Rust `as u8` / `as u64` / `as i32` saturate, NaN -> 0.
inline uint8_t sat_u8(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 255.0) return 255;
    return (uint8_t)x;
}
inline u64 sat_u64(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 18446744073709551616.0) return ~0ull;
    return (u64)x;
}
inline int32_t sat_i32(double x) {
    if (std::isnan(x)) return 0;
    if (x <= -2147483648.0) return INT32_MIN;
    if (x >= 2147483647.0) return INT32_MAX;
    return (int32_t)x;
}
// f64::fract = x - trunc(x)
inline double rfract(double x) { return x - std::trunc(x); }

// 4x4 matrices, column-major storage m[c*4 + r] as nalgebra. Products follow
// nalgebra's gemv: y_r = ((M_r0 x0 + M_r1 x1) + M_r2 x2) + M_r3 x3.
struct M4 {
    double m[16];
    double& at(int r, int c) { return m[c * 4 + r]; }
    double at(int r, int c) const { return m[c * 4 + r]; }
};
M4 m4_identity() {
    M4 a{};
    for (int i = 0; i < 4; ++i) a.at(i, i) = 1.0;
    return a;
}
bool m4_is_identity(const M4& a) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            if (a.at(r, c) != (r == c ? 1.0 : 0.0)) return false;
    return true;
}
void m4_mulv(const M4& a, const double x[4], double y[4]) {
    for (int r = 0; r < 4; ++r) {
        double acc = a.at(r, 0) * x[0];
        acc = a.at(r, 1) * x[1] + acc;
        acc = a.at(r, 2) * x[2] + acc;
        acc = a.at(r, 3) * x[3] + acc;
        y[r] = acc;
    }
}
M4 m4_mul(const M4& a, const M4& b) {
    M4 c{};
    for (int j = 0; j < 4; ++j) {
        double x[4] = {b.at(0, j), b.at(1, j), b.at(2, j), b.at(3, j)}, y[4];
        m4_mulv(a, x, y);
        for (int r = 0; r < 4; ++r) c.at(r, j) = y[r];
    }
    return c;
}
V3 m4_point(const M4& a, V3 p) {  // M * p.to_homogeneous() (w = 1), xyz
    double x[4] = {p.x, p.y, p.z, 1.0}, y[4];
    m4_mulv(a, x, y);
    return {y[0], y[1], y[2]};
}
V3 m4_vec(const M4& a, V3 v) {  // M * v.to_homogeneous() (w = 0), xyz
    double x[4] = {v.x, v.y, v.z, 0.0}, y[4];
    m4_mulv(a, x, y);
    return {y[0], y[1], y[2]};
}
// Cofactor 4x4 inverse (the gluInvertMatrix formulas used by nalgebra's
// try_inverse for 4x4). Bit order vs nalgebra is not pinned by any reference test.
bool m4_inverse(const M4& M, M4& out) {
    const double* m = M.m;
    double inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0) return false;
    double inv_det = 1.0 / det;
    for (int i = 0; i < 16; ++i) out.m[i] = inv[i] * inv_det;
    return true;
}
// nalgebra do_inverse3 on the upper-left 3x3 (src/ray.rs:478-479), then transpose
// applied to n: returns (inv(A))^T * n.
V3 normal_transform(const M4& T, V3 n) {
    auto a = [&](int r, int c) { return T.at(r, c); };
    double m12_23 = a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2);
    double m11_23 = a(1, 0) * a(2, 2) - a(2, 0) * a(1, 2);
    double m11_22 = a(1, 0) * a(2, 1) - a(2, 0) * a(1, 1);
    double det = a(0, 0) * m12_23 - a(0, 1) * m11_23 + a(0, 2) * m11_22;
    double inv[3][3];
    if (det == 0.0) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) inv[r][c] = (r == c) ? 1.0 : 0.0;
    } else {
        inv[0][0] = m12_23 / det;
        inv[0][1] = (a(0, 2) * a(2, 1) - a(2, 2) * a(0, 1)) / det;
        inv[0][2] = (a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1)) / det;
        inv[1][0] = -m11_23 / det;
        inv[1][1] = (a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0)) / det;
        inv[1][2] = (a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2)) / det;
        inv[2][0] = m11_22 / det;
        inv[2][1] = (a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1)) / det;
        inv[2][2] = (a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0)) / det;
    }
    // (inv^T) * n, 3x3 gemv order: y_r = (T_r0 n0 + T_r1 n1) + T_r2 n2 with T = inv^T
    double x[3] = {n.x, n.y, n.z}, y[3];
    for (int r = 0; r < 3; ++r) {
        double acc = inv[0][r] * x[0];
        acc = inv[1][r] * x[1] + acc;
        acc = inv[2][r] * x[2] + acc;
        y[r] = acc;
    }
    return {y[0], y[1], y[2]};
}

// ---------------------------------------------------------------------------
// Minimal JSON reader (serde_json semantics for the fields the scene uses).
// ---------------------------------------------------------------------------
struct J {
    enum T { Null, Bool, Num, Str, Arr, Obj } t = Null;
    bool b = false;
    double n = 0;
    std::string s;
    std::vector<J> a;
    std::vector<std::pair<std::string, J>> o;
    const J* get(const char* k) const {
        for (auto& kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};
struct JP {
    const char* p;
    const char* e;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    [[noreturn]] void fail(const char* m) { throw Err(-2, std::string("JSON parse error: ") + m); }
    J parse() {
        ws();
        if (p >= e) fail("unexpected end");
        char c = *p;
        J v;
        if (c == '{') {
            ++p;
            v.t = J::Obj;
            ws();
            if (p < e && *p == '}') { ++p; return v; }
            for (;;) {
                ws();
                if (p >= e || *p != '"') fail("expected key");
                std::string k = str();
                ws();
                if (p >= e || *p != ':') fail("expected ':'");
                ++p;
                J x = parse();
                v.o.emplace_back(k, std::move(x));
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == '}') { ++p; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            ++p;
            v.t = J::Arr;
            ws();
            if (p < e && *p == ']') { ++p; return v; }
            for (;;) {
                v.a.push_back(parse());
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == ']') { ++p; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.t = J::Str;
            v.s = str();
        } else if (c == 't' && e - p >= 4 && !strncmp(p, "true", 4)) {
            p += 4; v.t = J::Bool; v.b = true;
        } else if (c == 'f' && e - p >= 5 && !strncmp(p, "false", 5)) {
            p += 5; v.t = J::Bool; v.b = false;
        } else if (c == 'n' && e - p >= 4 && !strncmp(p, "null", 4)) {
            p += 4; v.t = J::Null;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            const char* s0 = p;
            if (*p == '-') ++p;
            while (p < e && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-')) ++p;
            std::string num(s0, p);
            char* end = nullptr;
            v.n = strtod(num.c_str(), &end);  // glibc strtod rounds correctly
            if (!end || *end) fail("bad number");
            v.t = J::Num;
        } else {
            fail("unexpected character");
        }
        return v;
    }
    std::string str() {
        ++p;
        std::string r;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= e) fail("bad escape");
                char c = *p++;
                switch (c) {
                    case 'n': r += '\n'; break;
                    case 't': r += '\t'; break;
                    case 'r': r += '\r'; break;
                    case 'b': r += '\b'; break;
                    case 'f': r += '\f'; break;
                    case 'u': {
                        if (e - p < 4) fail("bad \\u");
                        unsigned cp = (unsigned)strtoul(std::string(p, p + 4).c_str(), nullptr, 16);
                        p += 4;
                        if (cp < 0x80) r += (char)cp;
                        else if (cp < 0x800) { r += (char)(0xC0 | (cp >> 6)); r += (char)(0x80 | (cp & 0x3F)); }
                        else { r += (char)(0xE0 | (cp >> 12)); r += (char)(0x80 | ((cp >> 6) & 0x3F)); r += (char)(0x80 | (cp & 0x3F)); }
                        break;
                    }
                    default: r += c;
                }
            } else {
                r += *p++;
            }
        }
        if (p >= e) fail("unterminated string");
        ++p;
        return r;
    }
};

const J& req(const J& o, const char* k) {
    const J* v = o.get(k);
    if (!v) throw Err(-2, std::string("missing field `") + k + "`");
    return *v;
}
double num(const J& v, const char* what) {
    if (v.t != J::Num) throw Err(-2, std::string("invalid type for `") + what + "`, expected f64");
    return v.n;
}
std::string strv(const J& v, const char* what) {
    if (v.t != J::Str) throw Err(-2, std::string("invalid type for `") + what + "`, expected a string");
    return v.s;
}
void arr3(const J& v, double out[3], const char* what) {
    if (v.t != J::Arr || v.a.size() != 3) throw Err(-2, std::string("invalid length for `") + what + "`, expected an array of length 3");
    for (int i = 0; i < 3; ++i) out[i] = num(v.a[i], what);
}
bool optnum(const J& o, const char* k, double& out) {
    const J* v = o.get(k);
    if (!v || v->t == J::Null) return false;
    out = num(*v, k);
    return true;
}

// hex_to_color (src/scene.rs:15-30): strips every leading '#', needs 6 bytes,
// u8::from_str_radix(_, 16) per pair (which accepts one leading '+'), /255.0.
bool hex_to_color(const std::string& hex_in, V3& out) {
    size_t i = 0;
    while (i < hex_in.size() && hex_in[i] == '#') ++i;
    std::string h = hex_in.substr(i);
    if (h.size() != 6) return false;
    int v[3];
    for (int k = 0; k < 3; ++k) {
        const char* s = h.c_str() + 2 * k;
        int len = 2;
        if (s[0] == '+') { ++s; --len; }
        if (len == 0) return false;
        int acc = 0;
        for (int j = 0; j < len; ++j) {
            char c = s[j];
            int d;
            if (c >= '0' && c <= '9') d = c - '0';
            else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
            else if (c >= 'A' && c <= 'F') d = c - 'A' + 10;
            else return false;
            acc = acc * 16 + d;
        }
        v[k] = acc;
    }
    out = v3(v[0] / 255.0, v[1] / 255.0, v[2] / 255.0);
    return true;
}
V3 hex_or(const std::string& s, V3 dflt) {
    V3 c;
    return hex_to_color(s, c) ? c : dflt;
}

// Rust `str::parse::<f64>()` accepts [+-]digits[.digits][e[+-]digits], "inf",
// "infinity", "nan" (case-insensitive) and nothing else (no whitespace, no hex).
bool rust_parse_f64(const std::string& s, double& out) {
    if (s.empty()) return false;
    std::string l;
    for (char c : s) l += (char)tolower((unsigned char)c);
    size_t i = 0;
    bool neg = false;
    if (l[0] == '+' || l[0] == '-') { neg = l[0] == '-'; i = 1; }
    std::string body = l.substr(i);
    if (body == "inf" || body == "infinity") { out = neg ? -INFINITY : INFINITY; return true; }
    if (body == "nan") { out = NAN; return true; }
    size_t j = 0;
    bool digits = false;
    while (j < body.size() && isdigit((unsigned char)body[j])) { ++j; digits = true; }
    if (j < body.size() && body[j] == '.') {
        ++j;
        while (j < body.size() && isdigit((unsigned char)body[j])) { ++j; digits = true; }
    }
    if (!digits) return false;
    if (j < body.size() && body[j] == 'e') {
        ++j;
        if (j < body.size() && (body[j] == '+' || body[j] == '-')) ++j;
        bool ed = false;
        while (j < body.size() && isdigit((unsigned char)body[j])) { ++j; ed = true; }
        if (!ed) return false;
    }
    if (j != body.size()) return false;
    out = strtod(s.c_str(), nullptr);
    return true;
}
// char::is_whitespace: Unicode White_Space.
bool is_white(uint32_t cp) {
    return cp == 0x20 || (cp >= 0x09 && cp <= 0x0D) || cp == 0x85 || cp == 0xA0 || cp == 0x1680 ||
           (cp >= 0x2000 && cp <= 0x200A) || cp == 0x2028 || cp == 0x2029 || cp == 0x202F || cp == 0x205F ||
           cp == 0x3000;
}
// Decodes the code point at s[i] (lossy: an ill-formed byte reads as U+FFFD, length 1).
uint32_t decode_at(const std::string& s, size_t i, size_t& len) {
    const unsigned char c = (unsigned char)s[i];
    len = 1;
    if (c < 0x80) return c;
    int extra = (c & 0xE0) == 0xC0 ? 1 : (c & 0xF0) == 0xE0 ? 2 : (c & 0xF8) == 0xF0 ? 3 : -1;
    if (extra < 0 || i + extra >= s.size()) return 0xFFFD;
    uint32_t cp = c & (0x7F >> (extra + 1));
    for (int k = 1; k <= extra; ++k) {
        const unsigned char d = (unsigned char)s[i + k];
        if ((d & 0xC0) != 0x80) return 0xFFFD;
        cp = (cp << 6) | (d & 0x3F);
    }
    const uint32_t min_cp = extra == 1 ? 0x80 : extra == 2 ? 0x800 : 0x10000;
    if (cp < min_cp || cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) return 0xFFFD;
    len = (size_t)extra + 1;
    return cp;
}
// (start, length) of each code point of s.
std::vector<std::pair<size_t, size_t>> code_points(const std::string& s) {
    std::vector<std::pair<size_t, size_t>> r;
    for (size_t i = 0, len; i < s.size(); i += len) {
        decode_at(s, i, len);
        r.push_back({i, len});
    }
    return r;
}
bool white_at(const std::string& s, const std::pair<size_t, size_t>& c) {
    size_t len;
    return is_white(decode_at(s, c.first, len));
}
// str::trim
std::string trim(const std::string& s) {
    auto cps = code_points(s);
    size_t a = 0, b = cps.size();
    while (a < b && white_at(s, cps[a])) ++a;
    while (b > a && white_at(s, cps[b - 1])) --b;
    if (a == b) return std::string();
    return s.substr(cps[a].first, cps[b - 1].first + cps[b - 1].second - cps[a].first);
}

// Transform::from_str + to_matrix + parse_transforms (src/scene.rs:117-203).
M4 euler(double roll, double pitch, double yaw) {
    // Rotation3::from_euler_angles: rows as nalgebra writes them.
    double sr = std::sin(roll), cr = std::cos(roll);
    double sp = std::sin(pitch), cp = std::cos(pitch);
    double sy = std::sin(yaw), cy = std::cos(yaw);
    M4 r = m4_identity();
    r.at(0, 0) = cy * cp;
    r.at(0, 1) = cy * sp * sr - sy * cr;
    r.at(0, 2) = cy * sp * cr + sy * sr;
    r.at(1, 0) = sy * cp;
    r.at(1, 1) = sy * sp * sr + cy * cr;
    r.at(1, 2) = sy * sp * cr - cy * sr;
    r.at(2, 0) = -sp;
    r.at(2, 1) = cp * sr;
    r.at(2, 2) = cp * cr;
    return r;
}
bool parse_one_transform(const std::string& s0, M4& out) {
    std::string s = trim(s0);
    int kind = -1;
    const char* pre[3] = {"rotate(", "translate(", "scale("};
    std::string params;
    for (int k = 0; k < 3; ++k) {
        size_t L = strlen(pre[k]);
        if (s.compare(0, L, pre[k]) == 0) {
            kind = k;
            params = s.substr(L);
            break;
        }
    }
    if (kind < 0) return false;
    if (params.empty() || params.back() != ')') return false;
    params.pop_back();
    std::vector<double> vals;
    size_t st = 0;
    for (;;) {
        size_t c = params.find(',', st);
        std::string tok = trim(params.substr(st, c == std::string::npos ? std::string::npos : c - st));
        double v;
        if (!rust_parse_f64(tok, v)) return false;
        vals.push_back(v);
        if (c == std::string::npos) break;
        st = c + 1;
    }
    if (vals.size() != 3) return false;
    if (kind == 0) {
        const double k = M_PI / 180.0;  // f64::to_radians
        M4 rx = euler(vals[0] * k, 0.0, 0.0);
        M4 ry = euler(0.0, vals[1] * k, 0.0);
        M4 rz = euler(0.0, 0.0, vals[2] * k);
        out = m4_mul(m4_mul(rz, ry), rx);
    } else if (kind == 1) {
        out = m4_identity();
        out.at(0, 3) = vals[0];
        out.at(1, 3) = vals[1];
        out.at(2, 3) = vals[2];
    } else {
        out = m4_identity();
        out.at(0, 0) = vals[0];
        out.at(1, 1) = vals[1];
        out.at(2, 2) = vals[2];
    }
    return true;
}
bool parse_transforms(const std::vector<std::string>& list, M4& out) {
    M4 c = m4_identity();
    for (auto& s : list) {
        M4 m;
        if (!parse_one_transform(s, m)) return false;
        c = m4_mul(m, c);
    }
    out = c;
    return true;
}

// ---------------------------------------------------------------------------
// Scene model (src/scene.rs:34-446)
// ---------------------------------------------------------------------------
struct Material {
    std::string color;
    double ambient, diffuse, specular, shininess;
    bool has_refl = false;
    double reflectivity = 0;
    int tex = 0;  // 0 none, 1 grid, 2 checkerboard
    std::string line_color;
    double line_width = 0, cell_size = 0;
    std::shared_ptr<Material> material_b;
};
Material parse_material(const J& m) {
    if (m.t != J::Obj) throw Err(-2, "invalid type for `material`");
    Material r;
    r.color = strv(req(m, "color"), "color");
    r.ambient = num(req(m, "ambient"), "ambient");
    r.diffuse = num(req(m, "diffuse"), "diffuse");
    r.specular = num(req(m, "specular"), "specular");
    r.shininess = num(req(m, "shininess"), "shininess");
    r.has_refl = optnum(m, "reflectivity", r.reflectivity);
    const J* t = m.get("texture");
    if (t && t->t != J::Null) {
        std::string ty = strv(req(*t, "type"), "type");
        if (ty == "grid") {
            r.tex = 1;
            r.line_color = strv(req(*t, "line_color"), "line_color");
            r.line_width = num(req(*t, "line_width"), "line_width");
            r.cell_size = num(req(*t, "cell_size"), "cell_size");
        } else if (ty == "checkerboard") {
            r.tex = 2;
            r.material_b = std::make_shared<Material>(parse_material(req(*t, "material_b")));
        } else {
            throw Err(-2, "unknown variant `" + ty + "`, expected `grid` or `checkerboard`");
        }
    }
    return r;
}

struct Tri {
    V3 v[3];
};

// ---- STL loader (src/mesh.rs:751-1003) ----
std::vector<std::string> rust_lines(const std::string& content) {
    std::vector<std::string> out;
    size_t st = 0;
    while (st < content.size()) {
        size_t nl = content.find('\n', st);
        std::string line = content.substr(st, nl == std::string::npos ? std::string::npos : nl - st);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        out.push_back(line);
        if (nl == std::string::npos) break;
        st = nl + 1;
    }
    return out;
}
// str::split_whitespace
std::vector<std::string> split_ws(const std::string& s) {
    std::vector<std::string> r;
    std::string cur;
    for (const auto& c : code_points(s)) {
        if (white_at(s, c)) {
            if (!cur.empty()) r.push_back(cur);
            cur.clear();
        } else {
            cur += s.substr(c.first, c.second);
        }
    }
    if (!cur.empty()) r.push_back(cur);
    return r;
}
// std::str::from_utf8 acceptance: shortest forms only, no surrogates, at most U+10FFFF.
bool utf8_ok(const std::string& s) {
    size_t i = 0;
    const size_t n = s.size();
    while (i < n) {
        const unsigned char c = (unsigned char)s[i];
        if (c < 0x80) {
            ++i;
            continue;
        }
        uint32_t cp;
        int extra;
        if ((c & 0xE0) == 0xC0) { cp = c & 0x1F; extra = 1; }
        else if ((c & 0xF0) == 0xE0) { cp = c & 0x0F; extra = 2; }
        else if ((c & 0xF8) == 0xF0) { cp = c & 0x07; extra = 3; }
        else return false;
        if (i + extra >= n) return false;  // truncated sequence
        for (int k = 1; k <= extra; ++k) {
            const unsigned char d = (unsigned char)s[i + k];
            if ((d & 0xC0) != 0x80) return false;
            cp = (cp << 6) | (d & 0x3F);
        }
        const uint32_t min_cp = extra == 1 ? 0x80 : extra == 2 ? 0x800 : 0x10000;
        if (cp < min_cp || cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) return false;
        i += extra + 1;
    }
    return true;
}
// is_ascii_stl (src/mesh.rs:787-809): BufRead::lines over the file; each line read must
// be UTF-8 (an io error otherwise); the loop stops after reading an 11th line.
bool is_ascii_stl(const std::string& bytes) {
    auto lines = rust_lines(bytes);
    int count = 0;
    for (const std::string& l : lines) {
        if (!utf8_ok(l)) throw Err(-3, "stream did not contain valid UTF-8");
        if (++count > 10) break;
        std::string t = trim(l);
        if (t.compare(0, 12, "facet normal") == 0 || t == "outer loop" || t == "endloop") return true;
    }
    return false;
}
std::vector<Tri> load_stl(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Err(-3, "No such file or directory (os error 2): " + path);
    std::string bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (bytes.size() < 80) throw Err(-3, "failed to fill whole buffer");
    std::string header = trim(bytes.substr(0, 80));  // from_utf8_lossy(header).trim_start()
    std::vector<Tri> tris;
    if (header.compare(0, 5, "solid") == 0 && is_ascii_stl(bytes)) {
        if (!utf8_ok(bytes)) throw Err(-3, "stream did not contain valid UTF-8");  // read_to_string (:831-835)
        auto lines = rust_lines(bytes);
        size_t i = 0;
        while (i < lines.size()) {
            std::string line = trim(lines[i]);
            if (line.compare(0, 12, "facet normal") == 0) {
                auto parts = split_ws(line);
                if (parts.size() != 5) throw Err(-3, "Invalid facet normal format");
                double nrm[3];  // Triangle::normal, never read by the path (:855-858)
                for (int k = 0; k < 3; ++k)
                    if (!rust_parse_f64(parts[k + 2], nrm[k])) throw Err(-3, "invalid float literal");
                ++i;
                if (i >= lines.size() || trim(lines[i]) != "outer loop") throw Err(-3, "Expected 'outer loop' after facet normal");
                Tri t;
                for (int j = 0; j < 3; ++j) {
                    ++i;
                    if (i >= lines.size()) throw Err(-3, "Unexpected end of file while reading vertex");
                    std::string vl = trim(lines[i]);
                    if (vl.compare(0, 6, "vertex") != 0) throw Err(-3, "Expected vertex line");
                    auto p = split_ws(vl);
                    if (p.size() != 4) throw Err(-3, "Invalid vertex format");
                    double c[3];
                    for (int k = 0; k < 3; ++k)
                        if (!rust_parse_f64(p[k + 1], c[k])) throw Err(-3, "invalid float literal");
                    t.v[j] = v3(c[0], c[1], c[2]);
                }
                ++i;
                if (i >= lines.size() || trim(lines[i]) != "endloop") throw Err(-3, "Expected 'endloop'");
                ++i;
                if (i >= lines.size() || trim(lines[i]) != "endfacet") throw Err(-3, "Expected 'endfacet'");
                tris.push_back(t);
            }
            ++i;
        }
        return tris;
    }
    if (bytes.size() < 84) throw Err(-3, "Binary STL too short");
    u32 n;
    memcpy(&n, bytes.data() + 80, 4);
    size_t expected = 84 + (size_t)n * 50;
    if (bytes.size() < expected)
        throw Err(-3, "Binary STL size mismatch: expected " + std::to_string(expected) + ", got " + std::to_string(bytes.size()));
    size_t off = 84;
    tris.resize(n);
    for (u32 k = 0; k < n; ++k) {
        off += 12;  // facet normal is read but never used by the path
        for (int j = 0; j < 3; ++j) {
            float c[3];
            memcpy(c, bytes.data() + off, 12);
            tris[k].v[j] = v3((double)c[0], (double)c[1], (double)c[2]);
            off += 12;
        }
        off += 2;
    }
    return tris;
}

// ---- k-d tree (src/mesh.rs:75-465) ----
struct KdNode {
    int leaf;
    int axis;
    double split;
    int left, right;
    std::vector<int> tris;
    V3 bmin, bmax;
};
struct KdTree {
    std::vector<KdNode> nodes;  // nodes[0] = root when non-empty
    int max_depth = 16, max_leaf = 10;
};
V3 tri_center(const Tri& t) {  // (v0 + v1 + v2) / 3.0 (src/mesh.rs:21-23)
    V3 s = t.v[0] + t.v[1];
    s = s + t.v[2];
    return s / 3.0;
}
void tri_bounds(const Tri& t, V3& mn, V3& mx) {  // src/mesh.rs:26-36
    mn = t.v[0];
    mx = t.v[0];
    for (int k = 1; k < 3; ++k) {
        mn = v3(rmin(mn.x, t.v[k].x), rmin(mn.y, t.v[k].y), rmin(mn.z, t.v[k].z));
        mx = v3(rmax(mx.x, t.v[k].x), rmax(mx.y, t.v[k].y), rmax(mx.z, t.v[k].z));
    }
}
void bounds_of(const std::vector<Tri>& T, const std::vector<int>& idx, V3& mn, V3& mx) {  // :200-215
    if (idx.empty()) { mn = mx = v3(0, 0, 0); return; }
    mn = v3(INFINITY, INFINITY, INFINITY);
    mx = v3(-INFINITY, -INFINITY, -INFINITY);
    for (int i : idx) {
        V3 a, b;
        tri_bounds(T[i], a, b);
        mn = v3(rmin(mn.x, a.x), rmin(mn.y, a.y), rmin(mn.z, a.z));
        mx = v3(rmax(mx.x, b.x), rmax(mx.y, b.y), rmax(mx.z, b.z));
    }
}
int kd_build(KdTree& kt, const std::vector<Tri>& T, std::vector<int> idx, V3 bmin, V3 bmax, int depth) {  // :125-197
    int me = (int)kt.nodes.size();
    kt.nodes.push_back(KdNode{});
    if (depth >= kt.max_depth || (int)idx.size() <= kt.max_leaf) {
        KdNode& n = kt.nodes[me];
        n.leaf = 1;
        n.tris = std::move(idx);
        n.bmin = bmin;
        n.bmax = bmax;
        return me;
    }
    int axis = depth % 3;
    std::vector<std::pair<double, int>> pos;
    pos.reserve(idx.size());
    for (int i : idx) pos.emplace_back(comp(tri_center(T[i]), axis), i);
    for (auto& p : pos)
        if (std::isnan(p.first)) throw Err(-3, "called `Option::unwrap()` on a `None` value");
    std::stable_sort(pos.begin(), pos.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
        return a.first < b.first;
    });
    double split = pos[pos.size() / 2].first;
    std::vector<int> L, R;
    for (auto& p : pos) {
        V3 mn, mx;
        tri_bounds(T[p.second], mn, mx);
        if (comp(mn, axis) <= split) L.push_back(p.second);
        if (comp(mx, axis) > split) R.push_back(p.second);
    }
    if (L.empty()) {
        L.push_back(R.back());
        R.pop_back();
    } else if (R.empty()) {
        R.push_back(L.back());
        L.pop_back();
    }
    V3 lmin, lmax, rmin_, rmax_;
    bounds_of(T, L, lmin, lmax);
    bounds_of(T, R, rmin_, rmax_);
    int l = kd_build(kt, T, std::move(L), lmin, lmax, depth + 1);
    int r = kd_build(kt, T, std::move(R), rmin_, rmax_, depth + 1);
    KdNode& n = kt.nodes[me];
    n.leaf = 0;
    n.axis = axis;
    n.split = split;
    n.left = l;
    n.right = r;
    n.bmin = bmin;
    n.bmax = bmax;
    return me;
}
KdTree kd_new(const std::vector<Tri>& T) {  // :75-101
    KdTree kt;
    if (!T.empty()) {
        std::vector<int> idx(T.size());
        for (size_t i = 0; i < T.size(); ++i) idx[i] = (int)i;
        V3 mn, mx;
        bounds_of(T, idx, mn, mx);
        kd_build(kt, T, std::move(idx), mn, mx, 0);
    }
    return kt;
}
bool ray_intersects_bounds(V3 o, V3 d, V3 mn, V3 mx) {  // :218-255
    double tmin = -INFINITY, tmax = INFINITY;
    for (int a = 0; a < 3; ++a) {
        double da = comp(d, a), oa = comp(o, a);
        if (std::fabs(da) < 1e-9) {
            if (oa < comp(mn, a) || oa > comp(mx, a)) return false;
        } else {
            double inv = 1.0 / da;
            double t0 = (comp(mn, a) - oa) * inv;
            double t1 = (comp(mx, a) - oa) * inv;
            if (t0 > t1) std::swap(t0, t1);
            tmin = rmax(tmin, t0);
            tmax = rmin(tmax, t1);
            if (tmin > tmax) return false;
        }
    }
    return tmax >= 0.0;
}

struct Counters {
    u64 primary = 0, reflect = 0, shadow = 0, kd_inner = 0, kd_leaf = 0, tri = 0, sphere = 0;
    void add(const Counters& c) {
        primary += c.primary; reflect += c.reflect; shadow += c.shadow;
        kd_inner += c.kd_inner; kd_leaf += c.kd_leaf; tri += c.tri; sphere += c.sphere;
    }
};

template <class F>
void kd_traverse(const KdTree& kt, int ni, V3 o, V3 d, F& cb, Counters& C) {  // :399-465
    const KdNode& n = kt.nodes[ni];
    if (n.leaf) {
        C.kd_leaf++;
        if (ray_intersects_bounds(o, d, n.bmin, n.bmax)) cb(n.tris);
        return;
    }
    C.kd_inner++;
    double op = comp(o, n.axis), dir = comp(d, n.axis);
    if (std::fabs(dir) < 1e-9) {
        if (op <= n.split) kd_traverse(kt, n.left, o, d, cb, C);
        else kd_traverse(kt, n.right, o, d, cb, C);
        return;
    }
    double t_split = (n.split - op) / dir;
    if (op <= n.split) {
        kd_traverse(kt, n.left, o, d, cb, C);
        if (t_split >= 0.0) kd_traverse(kt, n.right, o, d, cb, C);
    } else {
        kd_traverse(kt, n.right, o, d, cb, C);
        if (t_split >= 0.0) kd_traverse(kt, n.left, o, d, cb, C);
    }
}

// ---------------------------------------------------------------------------
// Geometry (src/ray.rs)
// ---------------------------------------------------------------------------
struct Ray {
    V3 o, d;  // d is unit: Ray::new normalizes (src/ray.rs:13-18)
};
Ray make_ray(V3 o, V3 d) { return Ray{o, normalize(d)}; }

struct Hit {
    V3 p, n;
    double t;
    bool front;
    int obj;
    int tri = -1;
    bool has_uv = false;
    double u = 0, v = 0;
};
// HitRecord::new (src/ray.rs:39-63)
void set_face(Hit& h, const Ray& r, V3 outward) {
    h.front = dot(r.d, outward) < 0.0;
    h.n = h.front ? normalize(outward) : normalize(-outward);
}

enum ObjKind { SPHERE, PLANE, CUBE, MESH };
struct Obj {
    ObjKind kind;
    int index;  // scene object index == material index
    V3 c;       // sphere/cube centre, plane point
    double r = 0;
    V3 n;       // plane unit normal
    V3 half;    // cube half size
    M4 T, Inv;  // cube: world->local (M^-1) and local->world (M)
    bool t_identity = true;
    std::vector<Tri> tris;
    V3 bmin, bmax;
    KdTree kd;
    bool use_kd = true;
};

bool sphere_hit(const Obj& s, const Ray& r, double tmin, double tmax, Hit& h) {  // :89-120
    V3 oc = r.o - s.c;
    double a = dot(r.d, r.d);
    double half_b = dot(oc, r.d);
    double c = dot(oc, oc) - s.r * s.r;
    double disc = half_b * half_b - a * c;
    if (disc < 0.0) return false;
    double sq = std::sqrt(disc);
    double root = (-half_b - sq) / a;
    if (root < tmin || tmax < root) {
        root = (-half_b + sq) / a;
        if (root < tmin || tmax < root) return false;
    }
    h.t = root;
    h.p = r.o + root * r.d;
    V3 outward = (h.p - s.c) / s.r;
    set_face(h, r, outward);
    h.obj = s.index;
    h.tri = -1;
    h.has_uv = false;
    return true;
}
bool plane_hit(const Obj& pl, const Ray& r, double tmin, double tmax, Hit& h) {  // :136-176
    double denom = dot(pl.n, r.d);
    if (std::fabs(denom) < 1e-8) return false;
    double t = dot(pl.c - r.o, pl.n) / denom;
    if (t < tmin || t > tmax) return false;
    h.t = t;
    h.p = r.o + t * r.d;
    set_face(h, r, pl.n);
    h.obj = pl.index;
    h.tri = -1;
    V3 ua = std::fabs(pl.n.x) > 0.9 ? v3(0, 1, 0) : v3(1, 0, 0);
    ua = normalize(cross(ua, pl.n));
    V3 va = normalize(cross(pl.n, ua));
    V3 rel = h.p - pl.c;
    h.has_uv = true;
    h.u = dot(rel, ua);
    h.v = dot(rel, va);
    return true;
}
bool cube_hit(const Obj& cb, const Ray& r, double tmin, double tmax, Hit& h) {  // :410-491
    V3 lo = m4_vec(cb.T, r.o - cb.c);
    V3 ld = m4_vec(cb.T, r.d);
    if (norm(ld) < 1e-8) return false;
    Ray lr = make_ray(lo, ld);
    double tmin_hit = tmin, tmax_hit = tmax;
    V3 normal = v3(0, 0, 0);
    for (int axis = 0; axis < 3; ++axis) {
        double inv = 1.0 / comp(lr.d, axis);
        double h_ = comp(cb.half, axis), oa = comp(lr.o, axis);
        double t0 = (-h_ - oa) * inv;
        double t1 = (h_ - oa) * inv;
        V3 an = v3(0, 0, 0);
        setc(an, axis, inv < 0.0 ? 1.0 : -1.0);
        if (inv < 0.0) {
            std::swap(t0, t1);
            setc(an, axis, -comp(an, axis));
        }
        if (t0 > tmin_hit) {
            tmin_hit = t0;
            normal = an;
        }
        if (t1 < tmax_hit) tmax_hit = t1;
        if (tmin_hit > tmax_hit) return false;
    }
    double t = tmin_hit > tmin ? tmin_hit : tmax_hit;
    if (t < tmin || t > tmax) return false;
    V3 lp = lr.o + t * lr.d;
    V3 wp = cb.c + m4_point(cb.Inv, lp);
    V3 wn = cb.t_identity ? normal : normal_transform(cb.T, normal);
    h.t = t;
    h.p = wp;
    set_face(h, r, wn);
    h.obj = cb.index;
    h.tri = -1;
    h.has_uv = false;
    return true;
}
// intersect_triangle (src/ray.rs:527-583)
bool tri_hit(const Tri& T, const Ray& r, double tmin, double tmax, double& t_out, V3& n_out, double& u_out, double& v_out) {
    V3 e1 = T.v[1] - T.v[0];
    V3 e2 = T.v[2] - T.v[0];
    V3 hh = cross(r.d, e2);
    double a = dot(e1, hh);
    if (a > -1e-8 && a < 1e-8) return false;
    double f = 1.0 / a;
    V3 s = r.o - T.v[0];
    double u = f * dot(s, hh);
    if (!(u >= 0.0 && u <= 1.0)) return false;
    V3 q = cross(s, e1);
    double v = f * dot(r.d, q);
    if (v < 0.0 || u + v > 1.0) return false;
    double t = f * dot(e2, q);
    if (t > tmin && t < tmax) {
        V3 n = cross(e1, e2);
        if (norm(n) < 1e-8) return false;
        if (a < 0.0) n = -n;
        n = normalize(n);
        t_out = t;
        n_out = n;
        u_out = u;
        v_out = v;
        return true;
    }
    return false;
}
bool mesh_bounds(const Obj& m, const Ray& r, double tmin, double tmax) {  // :586-610
    double a0 = tmin, a1 = tmax;
    for (int axis = 0; axis < 3; ++axis) {
        double inv = 1.0 / comp(r.d, axis);
        double t0 = (comp(m.bmin, axis) - comp(r.o, axis)) * inv;
        double t1 = (comp(m.bmax, axis) - comp(r.o, axis)) * inv;
        if (inv < 0.0) std::swap(t0, t1);
        a0 = rmax(a0, t0);
        a1 = rmin(a1, t1);
        if (a0 > a1) return false;
    }
    return true;
}
bool mesh_hit(const Obj& m, const Ray& r, double tmin, double tmax, Hit& h, Counters& C) {  // :614-675
    if (!mesh_bounds(m, r, tmin, tmax)) return false;
    bool found = false;
    double closest = tmax;
    auto test = [&](int ti) {
        double t, u, v;
        V3 n;
        C.tri++;
        if (tri_hit(m.tris[ti], r, tmin, closest, t, n, u, v)) {
            if (t < closest) {
                closest = t;
                h.t = t;
                h.p = r.o + t * r.d;
                set_face(h, r, n);
                h.obj = m.index;
                h.tri = ti;
                h.has_uv = true;
                h.u = u;
                h.v = v;
                found = true;
            }
        }
    };
    if (m.use_kd) {
        if (!m.kd.nodes.empty()) {
            auto cb = [&](const std::vector<int>& ids) {
                for (int ti : ids) test(ti);
            };
            kd_traverse(m.kd, 0, r.o, r.d, cb, C);
        }
    } else {
        for (size_t i = 0; i < m.tris.size(); ++i) test((int)i);
    }
    return found;
}
struct World {
    std::vector<Obj> objs;
    bool hit(const Ray& r, double tmin, double tmax, Hit& out, Counters& C) const {  // :699-711
        bool any = false;
        double closest = tmax;
        Hit h;
        for (const Obj& o : objs) {
            bool ok = false;
            switch (o.kind) {
                case SPHERE: C.sphere++; ok = sphere_hit(o, r, tmin, closest, h); break;
                case PLANE: ok = plane_hit(o, r, tmin, closest, h); break;
                case CUBE: ok = cube_hit(o, r, tmin, closest, h); break;
                case MESH: ok = mesh_hit(o, r, tmin, closest, h, C); break;
            }
            if (ok) {
                closest = h.t;
                out = h;
                any = true;
            }
        }
        return any;
    }
};

// ---------------------------------------------------------------------------
// RNG: rand 0.8 StdRng = ChaCha12 (rand_chacha 0.3) seeded by rand_core 0.6
// seed_from_u64 (8 PCG32 outputs form the 256-bit key); block counter in
// words 12-13, stream id words 14-15 = 0; next_u64 = lo | hi << 32 from
// consecutive words of consecutive blocks.
// ---------------------------------------------------------------------------
inline u32 rotl(u32 x, int k) { return (x << k) | (x >> (32 - k)); }
#define QR(a, b, c, d)                  \
    a += b; d ^= a; d = rotl(d, 16);    \
    c += d; b ^= c; b = rotl(b, 12);    \
    a += b; d ^= a; d = rotl(d, 8);     \
    c += d; b ^= c; b = rotl(b, 7);
void chacha_block(const u32 in[16], int drounds, u32 out[16]) {
    u32 x[16];
    memcpy(x, in, sizeof(x));
    for (int i = 0; i < drounds; ++i) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}
void seed_key(u64 state, u32 key[8]) {  // rand_core 0.6 SeedableRng::seed_from_u64
    const u64 MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
    for (int i = 0; i < 8; ++i) {
        state = state * MUL + INC;
        u32 xs = (u32)(((state >> 18) ^ state) >> 27);
        u32 rot = (u32)(state >> 59);
        key[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
}
struct StdRng {
    u32 in[16];
    u32 buf[16];
    int idx = 16;
    u64 block = 0;
    explicit StdRng(u64 seed) {
        u32 key[8];
        seed_key(seed, key);
        init(key);
    }
    explicit StdRng(const u32 key[8]) { init(key); }  // SeedableRng::from_seed, key words little-endian
    void init(const u32 key[8]) {
        in[0] = 0x61707865u; in[1] = 0x3320646eu; in[2] = 0x79622d32u; in[3] = 0x6b206574u;
        memcpy(in + 4, key, 8 * sizeof(u32));
        in[12] = in[13] = in[14] = in[15] = 0;
    }
    u32 next_u32() {
        if (idx >= 16) {
            in[12] = (u32)block;
            in[13] = (u32)(block >> 32);
            chacha_block(in, 6, buf);
            ++block;
            idx = 0;
        }
        return buf[idx++];
    }
    u64 next_u64() {
        u64 lo = next_u32();
        u64 hi = next_u32();
        return lo | (hi << 32);
    }
    // BlockRng::fill_bytes for whole words: consecutive words, each little-endian.
    void fill_words(u32* out, int n) {
        for (int i = 0; i < n; ++i) out[i] = next_u32();
    }
    double gen_f64();
    double gen_range(double low, double high);
};

// rand_pcg 0.3 Pcg32 (Lcg64Xsh32): only the published value-stability vectors of rand's
// float distributions use it, so it pins gen_f64 / gen_range below independently of ChaCha.
struct Pcg32 {
    u64 state, inc;
    Pcg32(u64 st, u64 stream) : state(st), inc((stream << 1) | 1) {  // Pcg32::new
        state += inc;
        state = state * 6364136223846793005ull + inc;
    }
    u32 next_u32() {
        u64 old = state;
        state = old * 6364136223846793005ull + inc;
        u32 xs = (u32)(((old >> 18) ^ old) >> 27);
        u32 rot = (u32)(old >> 59);
        return (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    u64 next_u64() {  // rand_core impls::next_u64_via_u32: low word first
        u64 lo = next_u32();
        u64 hi = next_u32();
        return lo | (hi << 32);
    }
};

// rand 0.8 Standard for f64: (next_u64 >> 11) * 2^-53.
template <class R>
double gen_f64_of(R& r) {
    return (double)(r.next_u64() >> 11) * (1.0 / 9007199254740992.0);
}
// rand 0.8 UniformFloat::<f64>::sample_single (what gen_range(low..high) calls): a value in
// [1, 2) from the top 52 bits, minus 1, times the scale plus low; a result reaching `high`
// retries with the scale lowered by one ulp.
template <class R>
double gen_range_of(R& r, double low, double high) {
    double scale = high - low;
    for (;;) {
        u64 bits = (r.next_u64() >> 12) | (1023ull << 52);
        double v12;
        memcpy(&v12, &bits, 8);
        double v01 = v12 - 1.0;
        double res = v01 * scale + low;
        if (res < high) return res;
        u64 sb;
        memcpy(&sb, &scale, 8);
        sb -= 1;
        memcpy(&scale, &sb, 8);
    }
}
double StdRng::gen_f64() { return gen_f64_of(*this); }
double StdRng::gen_range(double low, double high) { return gen_range_of(*this, low, high); }

// ---------------------------------------------------------------------------
// Shading (src/lighting.rs)
// ---------------------------------------------------------------------------
struct Light {
    V3 pos;
    std::string color;
    double intensity;
    bool has_d = false;
    double diameter = 0;
    // extension keys (the reference's serde model ignores them): "shape": "square",
    // "size", "samples"
    bool square = false;
    double size = 0;
    int samples = 16;
};
struct Fog {
    std::string color;
    double density, start, end;
};
struct Camera {  // src/camera.rs:7-19
    V3 origin, horizontal, vertical, llc, view_dir;
    bool persp;
    bool has_grid = false;
    double pitch = 0, thick = 0;
    V3 grid_color;
};
struct Ctx {
    const World* world;
    std::vector<Light> lights;
    std::vector<Material> mats;
    std::string amb_color;
    double amb_intensity;
    bool has_fog = false;
    Fog fog;
    V3 cam_pos;
    V3 background;
    int light_samples = 0;  // > 0 overrides every area light's sample count (extension)
};

V3 reflect(V3 inc, V3 n) {  // :291-294
    double d2 = 2.0 * dot(inc, n);
    return normalize(inc - d2 * n);
}
// apply_texture (:9-54): returns the effective material
const Material* apply_texture(const Material& m, double u, double v, Material& scratch) {
    if (m.tex == 1) {
        V3 gc = hex_or(m.line_color, v3(0, 0, 0));
        double hw = m.line_width / 2.0;
        double um = std::fabs(rfract(u / m.cell_size));
        double vm = std::fabs(rfract(v / m.cell_size));
        bool on_u = um <= hw || um >= (1.0 - hw);
        bool on_v = vm <= hw || vm >= (1.0 - hw);
        if (on_u || on_v) {
            scratch = m;
            char buf[16];
            snprintf(buf, sizeof(buf), "#%02X%02X%02X", sat_u8(gc.x * 255.0), sat_u8(gc.y * 255.0), sat_u8(gc.z * 255.0));
            scratch.color = buf;
            return &scratch;
        }
        return &m;
    }
    if (m.tex == 2) {
        int32_t cu = sat_i32(std::floor(u));
        int32_t cv = sat_i32(std::floor(v));
        int32_t sum = (int32_t)((uint32_t)cu + (uint32_t)cv);
        if (sum % 2 == 0) return &m;
        return m.material_b.get();
    }
    return &m;
}
// First-hit feature probe (extension outputs for the denoiser; no reference counterpart):
// shadow-sample visibility counts, effective material colour and fog factor. A light none of
// whose samples faces the hit point (diffuse strength max(n.l, 0) = 0 for every sample; only
// the first 32 lights of a scene are judged so) counts its samples as taken and not visible:
// its contribution is +0 whatever they see, so they need not be traced.
struct Probe {
    int nv = 0, nt = 0;  // area-light samples visible / taken
    int pv = 0, pt = 0;  // point lights visible / taken
    int light = 0;       // index of the light being shaded
    V3 mc{};
    double fog = 0.0;
};
V3 point_light(const Ctx& X, const Hit& h, const Material& mat, const Light& L, V3 lc, V3 mc, Counters& C,
               Probe* pr = nullptr) {  // :100-142
    V3 ld = normalize(L.pos - h.p);
    Ray sr = make_ray(h.p + 0.001 * h.n, ld);
    double dist = norm(L.pos - h.p);
    Hit tmp;
    C.shadow++;
    if (pr) pr->pt++;
    if (X.world->hit(sr, 0.001, dist, tmp, C)) return v3(0, 0, 0);
    double ds = rmax(dot(h.n, ld), 0.0);
    if (pr && (ds > 0.0 || pr->light >= 32)) pr->pv++;
    V3 diffuse = ((mat.diffuse * ds) * L.intensity) * cmul(lc, mc);
    V3 spec = v3(0, 0, 0);
    if (ds > 0.0) {
        V3 vd = normalize(X.cam_pos - h.p);
        V3 rd = reflect(-ld, h.n);
        double ss = std::pow(rmax(dot(vd, rd), 0.0), mat.shininess);
        spec = ((mat.specular * ss) * L.intensity) * lc;
    }
    return diffuse + spec;
}
// Square-light extension (no reference counterpart): n samples on a jittered grid of
// cols x n/cols cells over the side x side square centred on the light, in the disk's
// light-facing basis; cell (s % cols, s / cols) is jittered by SplitMix64 outputs 2s+1 and
// 2s+2 of the stream seeded with the light seed, each turned into [0,1) as (x >> 11) * 2^-53.
u64 splitmix64_nth(u64 seed, u64 k) {  // k-th output (k >= 1) of SplitMix64 from `seed`
    u64 z = seed + k * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
int jitter_cols(int n) {  // largest divisor of n that is <= sqrt(n)
    int c = 1;
    for (int d = 1; (long long)d * d <= n; ++d)
        if (n % d == 0) c = d;
    return c;
}
u64 light_seed(u64 seed, V3 p) {  // src/lighting.rs:161-165
    u64 ls = seed * 0x9E3779B97F4A7C15ull;
    ls = ls + sat_u64(p.x * 1000.0) * 0x85EBCA6Bull;
    ls = ls + sat_u64(p.y * 1000.0) * 0xC2B2AE35ull;
    ls = ls + sat_u64(p.z * 1000.0) * 0x6C8E9CF5ull;
    return ls;
}
// Sample s of an area light seen from hp: sample_disk_light_point (:69-97) drawing from
// rng, or the square extension's jittered cell s.
V3 area_sample(const Light& L, V3 hp, StdRng& rng, u64 ls, int s, int cols, int rows) {
    double radius = L.diameter / 2.0;
    V3 ldir = normalize(L.pos - hp);
    V3 up = std::fabs(ldir.x) < 0.9 ? v3(1, 0, 0) : v3(0, 1, 0);
    V3 u = normalize(cross(up, ldir));
    V3 v = normalize(cross(ldir, u));
    double dx, dy;
    if (L.square) {
        double jx = (double)(splitmix64_nth(ls, 2 * (u64)s + 1) >> 11) * 0x1.0p-53;
        double jy = (double)(splitmix64_nth(ls, 2 * (u64)s + 2) >> 11) * 0x1.0p-53;
        dx = (((double)(s % cols) + jx) / (double)cols - 0.5) * L.size;
        dy = (((double)(s / cols) + jy) / (double)rows - 0.5) * L.size;
    } else {
        for (;;) {  // sample_disk_point (:57-66)
            dx = rng.gen_range(-radius, radius);
            dy = rng.gen_range(-radius, radius);
            if (dx * dx + dy * dy <= radius * radius) break;
        }
    }
    return (L.pos + dx * u) + dy * v;
}
V3 area_light(const Ctx& X, const Hit& h, const Material& mat, const Light& L, V3 lc, V3 mc, u64 seed, Counters& C,
              Probe* pr = nullptr) {  // :145-217
    const int SAMPLES = X.light_samples > 0 ? X.light_samples : L.samples;  // reference: 16 (:158)
    u64 ls = light_seed(seed, h.p);
    StdRng rng(ls);
    const int cols = jitter_cols(SAMPLES), rows = SAMPLES / cols;
    V3 total = v3(0, 0, 0);
    int visible = 0;
    bool faces = false;  // some sample has diffuse strength > 0 (the feature probe's rule)
    for (int s = 0; s < SAMPLES; ++s) {
        V3 sp = area_sample(L, h.p, rng, ls, s, cols, rows);
        V3 ld = normalize(sp - h.p);
        double dist = norm(sp - h.p);
        if (rmax(dot(h.n, ld), 0.0) > 0.0) faces = true;
        Ray sr = make_ray(h.p + 0.001 * h.n, ld);
        Hit tmp;
        C.shadow++;
        if (X.world->hit(sr, 0.001, dist, tmp, C)) continue;
        visible++;
        double ds = rmax(dot(h.n, ld), 0.0);
        V3 diffuse = ((mat.diffuse * ds) * L.intensity) * cmul(lc, mc);
        V3 spec = v3(0, 0, 0);
        if (ds > 0.0) {
            V3 vd = normalize(X.cam_pos - h.p);
            V3 rd = reflect(-ld, h.n);
            double ss = std::pow(rmax(dot(vd, rd), 0.0), mat.shininess);
            spec = ((mat.specular * ss) * L.intensity) * lc;
        }
        total = total + (diffuse + spec);
    }
    if (pr) {
        pr->nt += SAMPLES;
        if (faces || pr->light >= 32) pr->nv += visible;
    }
    return (total / (double)SAMPLES) * ((double)visible / (double)SAMPLES);
}
V3 phong(const Ctx& X, const Hit& h, const Material& mat, u64 seed, Counters& C, Probe* pr = nullptr) {  // :220-288
    Material scratch;
    const Material* eff = &mat;
    if (mat.tex != 0 && h.has_uv) eff = apply_texture(mat, h.u, h.v, scratch);
    V3 mc = hex_or(eff->color, v3(1, 1, 1));
    V3 ac = hex_or(X.amb_color, v3(1, 1, 1));
    V3 color = (eff->ambient * X.amb_intensity) * cmul(ac, mc);
    for (size_t li = 0; li < X.lights.size(); ++li) {
        const Light& L = X.lights[li];
        V3 lc = hex_or(L.color, v3(1, 1, 1));
        if (pr) pr->light = (int)li;
        V3 contrib = (L.has_d || L.square) ? area_light(X, h, *eff, L, lc, mc, seed, C, pr)
                                           : point_light(X, h, *eff, L, lc, mc, C, pr);
        color = color + contrib;
    }
    if (pr) pr->mc = mc;
    return color;
}
V3 apply_fog(const Ctx& X, V3 c, double dist, double* f_out = nullptr) {  // :297-319
    if (!X.has_fog) return c;
    V3 fc = hex_or(X.fog.color, v3(0.5, 0.5, 0.5));
    double f;
    if (dist <= X.fog.start) f = 0.0;
    else if (dist >= X.fog.end) f = 1.0;
    else f = (dist - X.fog.start) / (X.fog.end - X.fog.start);
    f = 1.0 - std::exp(-X.fog.density * f);
    f = rclamp(f, 0.0, 1.0);
    if (f_out) *f_out = f;
    return c * (1.0 - f) + fc * f;
}
bool grid_color(const Camera& cam, const Ray& r, V3& out) {  // src/camera.rs:171-253
    if (cam.persp) return false;
    if (!cam.has_grid) return false;
    double ht = cam.thick / 2.0, p = cam.pitch;
    auto check = [&](double a, double b) {
        double am = std::fabs(rfract(a / p)), bm = std::fabs(rfract(b / p));
        double ad = rmin(am * p, (1.0 - am) * p);
        double bd = rmin(bm * p, (1.0 - bm) * p);
        return ad <= ht || bd <= ht;
    };
    if (std::fabs(r.d.z) > 1e-10) {
        double t = -r.o.z / r.d.z;
        if (t > 0.0) {
            V3 ip = r.o + t * r.d;
            if (check(ip.x, ip.y)) { out = cam.grid_color; return true; }
        }
    }
    if (std::fabs(r.d.y) > 1e-10) {
        double t = -r.o.y / r.d.y;
        if (t > 0.0) {
            V3 ip = r.o + t * r.d;
            if (check(ip.x, ip.z)) { out = cam.grid_color; return true; }
        }
    }
    if (std::fabs(r.d.x) > 1e-10) {
        double t = -r.o.x / r.d.x;
        if (t > 0.0) {
            V3 ip = r.o + t * r.d;
            if (check(ip.y, ip.z)) { out = cam.grid_color; return true; }
        }
    }
    return false;
}
struct PrimRec {
    int obj = -1, tri = -1;
    double t = 0;
};
// ray_color_with_camera / ray_color / ray_color_with_data (:323-506). `cam` is
// null for ray_color (quincunx). depth/normal are the first hit's (with_data).
// Features of the first hit (extension; zeros on a miss), in the device's channel order.
struct Feat {
    float v[13] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};
V3 ray_color(const Ctx& X, const Ray& r, int max_depth, const Camera* cam, u64 seed, Counters& C, bool primary,
             PrimRec* rec, bool* has_dn, double* depth_out, V3* normal_out, Feat* feat = nullptr) {
    if (max_depth <= 0) return v3(0, 0, 0);
    Hit h;
    if (primary) C.primary++; else C.reflect++;
    if (X.world->hit(r, 0.001, INFINITY, h, C)) {
        if (rec) { rec->obj = h.obj; rec->tri = h.tri; rec->t = h.t; }
        if (has_dn) {
            *has_dn = true;
            *depth_out = norm(h.p - X.cam_pos);
            *normal_out = h.n;
        }
        const Material& mat = X.mats[h.obj];
        Probe pr;
        V3 color = phong(X, h, mat, seed, C, feat ? &pr : nullptr);
        double dist = norm(h.p - X.cam_pos);
        color = apply_fog(X, color, dist, &pr.fog);
        if (feat) {
            float* f = feat->v;
            f[3] = (float)pr.mc.x; f[4] = (float)pr.mc.y; f[5] = (float)pr.mc.z;
            f[6] = (float)h.n.x; f[7] = (float)h.n.y; f[8] = (float)h.n.z;
            f[9] = (float)dist;
            int nv = pr.nv, nt = pr.nt;
            if (nt == 0) { nv = pr.pv; nt = pr.pt; }
            float mean = nt > 0 ? (float)nv / (float)nt : 1.0f;
            f[10] = mean;
            f[11] = nt > 0 ? mean * (1.0f - mean) / (float)nt : 0.0f;
            f[12] = (float)pr.fog;
        }
        if (mat.has_refl && mat.reflectivity > 0.0 && max_depth > 1) {
            V3 vd = normalize(X.cam_pos - h.p);
            V3 rd = reflect(-vd, h.n);
            Ray rr = make_ray(h.p + 0.001 * h.n, rd);
            V3 rc = ray_color(X, rr, max_depth - 1, cam, seed, C, false, nullptr, nullptr, nullptr, nullptr);
            color = color * (1.0 - mat.reflectivity) + rc * mat.reflectivity;
        }
        return color;
    }
    if (rec) { rec->obj = -1; rec->tri = -1; rec->t = 0; }
    if (has_dn) *has_dn = false;
    V3 g;
    if (cam && grid_color(*cam, r, g)) return g;
    return X.background;
}

// Camera::from_config (src/camera.rs:23-152)
Camera make_camera(const J& cj, double aspect) {
    double pos[3], tgt[3], up[3];
    std::string kind = strv(req(cj, "kind"), "kind");
    arr3(req(cj, "position"), pos, "position");
    arr3(req(cj, "target"), tgt, "target");
    arr3(req(cj, "up"), up, "up");
    double cw = num(req(cj, "width"), "width"), ch = num(req(cj, "height"), "height");
    double fov = 45.0;
    bool has_fov = optnum(cj, "fov", fov);
    if (!has_fov) fov = 45.0;
    double pitch = 0, thick = 0;
    bool hp = optnum(cj, "grid_pitch", pitch);
    bool ht = optnum(cj, "grid_thickness", thick);
    const J* gcj = cj.get("grid_color");
    bool hc = gcj && gcj->t != J::Null;
    V3 gcol{};
    if (hc && !hex_to_color(strv(*gcj, "grid_color"), gcol)) throw Err(-4, "Invalid hex color format");
    Camera cam;
    V3 origin = v3(pos[0], pos[1], pos[2]), target = v3(tgt[0], tgt[1], tgt[2]), upv = v3(up[0], up[1], up[2]);
    V3 w = normalize(origin - target);
    V3 u = normalize(cross(upv, w));
    V3 v = cross(w, u);
    V3 view = normalize(-w);
    cam.origin = origin;
    cam.view_dir = view;
    if (kind == "ortho") {
        double vh = ch;
        double vw = rmax(cw, vh * aspect);
        cam.horizontal = vw * u;
        cam.vertical = vh * v;
        cam.llc = (origin - cam.horizontal / 2.0) - cam.vertical / 2.0;
        cam.persp = false;
    } else if (kind == "perspective") {
        if (fov <= 0.0 || fov >= 180.0) throw Err(-4, "Field of view must be between 0 and 180 degrees");
        double focal = 1.0;
        double theta = fov * (M_PI / 180.0);
        double half_h = std::tan(theta / 2.0);
        double half_w = aspect * half_h;
        double vph = (2.0 * half_h) * focal;
        double vpw = (2.0 * half_w) * focal;
        cam.horizontal = vpw * u;
        cam.vertical = vph * v;
        V3 center = origin + focal * view;
        cam.llc = (center - cam.horizontal / 2.0) - cam.vertical / 2.0;
        cam.persp = true;
    } else {
        throw Err(-4, "Unsupported camera type: " + kind);
    }
    if (hp && hc && ht && pitch > 0.0 && thick > 0.0) {
        cam.has_grid = true;
        cam.pitch = pitch;
        cam.thick = thick;
        cam.grid_color = gcol;
    }
    return cam;
}
Ray get_ray(const Camera& cam, double u, double v) {  // :155-167
    V3 vp = (cam.llc + u * cam.horizontal) + v * cam.vertical;
    if (cam.persp) {
        V3 d = normalize(vp - cam.origin);
        return make_ray(cam.origin, d);
    }
    return make_ray(vp, cam.view_dir);
}

std::string dir_of(const std::string& p) {
    size_t k = p.find_last_of('/');
    if (k == std::string::npos) return ".";
    if (k == 0) return "/";
    return p.substr(0, k);
}

struct Loaded {
    World world;
    Ctx ctx;
    const J* camj;
};

// Renderer::render world assembly (src/renderer.rs:137-317)
void build_world(const J& root, const std::string& scene_path, bool use_kd, Loaded& L) {
    const J& objs = req(root, "objects");
    if (objs.t != J::Arr) throw Err(-2, "invalid type for `objects`");
    const J& lights = req(root, "lights");
    if (lights.t != J::Arr) throw Err(-2, "invalid type for `lights`");
    const J& ss = req(root, "scene_settings");
    L.camj = &req(root, "camera");
    std::string base = dir_of(scene_path);
    int index = 0;
    for (const J& o : objs.a) {
        std::string kind = strv(req(o, "kind"), "kind");
        Material mat = parse_material(req(o, "material"));
        const J* tj = o.get("transform");
        bool has_t = tj && tj->t != J::Null;
        std::vector<std::string> tl;
        if (has_t) {
            if (tj->t != J::Arr) throw Err(-2, "invalid type for `transform`");
            for (auto& s : tj->a) tl.push_back(strv(s, "transform"));
        }
        M4 M;
        bool tok = has_t && parse_transforms(tl, M);
        V3 col;
        Obj ob;
        ob.index = index;
        if (kind == "sphere") {
            double c[3];
            arr3(req(o, "center"), c, "center");
            double r = num(req(o, "radius"), "radius");
            V3 cp = v3(c[0], c[1], c[2]);
            if (tok) {
                cp = m4_point(M, cp);
                auto colnorm = [&](int k) { return norm(v3(M.at(0, k), M.at(1, k), M.at(2, k))); };
                double sx = colnorm(0), sy = colnorm(1), sz = colnorm(2);
                r = r * rmax(rmax(sx, sy), sz);
            }
            if (!hex_to_color(mat.color, col)) throw Err(-4, "Invalid hex color format");
            ob.kind = SPHERE;
            ob.c = cp;
            ob.r = r;
        } else if (kind == "plane") {
            double p[3], n[3];
            arr3(req(o, "point"), p, "point");
            arr3(req(o, "normal"), n, "normal");
            V3 pp = v3(p[0], p[1], p[2]), nn = v3(n[0], n[1], n[2]);
            if (tok) {
                pp = m4_point(M, pp);
                M4 inv;
                if (m4_inverse(M, inv)) {
                    M4 it;
                    for (int r_ = 0; r_ < 4; ++r_)
                        for (int c_ = 0; c_ < 4; ++c_) it.at(r_, c_) = inv.at(c_, r_);
                    nn = m4_vec(it, nn);
                }
            }
            if (!hex_to_color(mat.color, col)) throw Err(-4, "Invalid hex color format");
            ob.kind = PLANE;
            ob.c = pp;
            ob.n = normalize(nn);
        } else if (kind == "cube") {
            double c[3], s[3];
            arr3(req(o, "center"), c, "center");
            arr3(req(o, "size"), s, "size");
            if (!hex_to_color(mat.color, col)) throw Err(-4, "Invalid hex color format");
            ob.kind = CUBE;
            ob.c = v3(c[0], c[1], c[2]);
            ob.half = v3(s[0], s[1], s[2]) / 2.0;
            if (tok) {
                M4 inv;
                if (!m4_inverse(M, inv)) inv = m4_identity();
                ob.T = inv;
                ob.Inv = M;
            } else {
                ob.T = m4_identity();
                ob.Inv = m4_identity();
            }
            ob.t_identity = m4_is_identity(ob.T);
        } else if (kind == "mesh") {
            std::string fn = strv(req(o, "filename"), "filename");
            std::string path = (fn.size() && fn[0] == '/') ? fn : base + "/" + fn;
            ob.kind = MESH;
            ob.tris = load_stl(path);
            if (tok) {
                for (Tri& t : ob.tris)
                    for (int k = 0; k < 3; ++k) t.v[k] = m4_point(M, t.v[k]);
            }
            // Mesh::compute_bounds (src/mesh.rs:1017-1033)
            if (ob.tris.empty()) {
                ob.bmin = ob.bmax = v3(0, 0, 0);
            } else {
                ob.bmin = v3(INFINITY, INFINITY, INFINITY);
                ob.bmax = v3(-INFINITY, -INFINITY, -INFINITY);
                for (Tri& t : ob.tris)
                    for (int k = 0; k < 3; ++k) {
                        ob.bmin = v3(rmin(ob.bmin.x, t.v[k].x), rmin(ob.bmin.y, t.v[k].y), rmin(ob.bmin.z, t.v[k].z));
                        ob.bmax = v3(rmax(ob.bmax.x, t.v[k].x), rmax(ob.bmax.y, t.v[k].y), rmax(ob.bmax.z, t.v[k].z));
                    }
            }
            ob.kd = kd_new(ob.tris);
            ob.use_kd = use_kd;
            if (!hex_to_color(mat.color, col)) throw Err(-4, "Invalid hex color format");
        } else {
            throw Err(-2, "unknown variant `" + kind + "`");
        }
        L.world.objs.push_back(std::move(ob));
        L.ctx.mats.push_back(mat);
        ++index;
    }
    for (const J& l : lights.a) {
        Light li;
        double p[3];
        arr3(req(l, "position"), p, "position");
        li.pos = v3(p[0], p[1], p[2]);
        li.color = strv(req(l, "color"), "color");
        li.intensity = num(req(l, "intensity"), "intensity");
        li.has_d = optnum(l, "diameter", li.diameter);
        const J* shp = l.get("shape");
        if (shp && shp->t != J::Null) {
            std::string k = strv(*shp, "shape");
            if (k == "square") li.square = true;
            else if (k != "disk") throw Err(-2, "unknown variant `" + k + "`, expected `disk` or `square`");
        }
        if (li.square && !optnum(l, "size", li.size)) {
            if (!li.has_d) throw Err(-2, "missing field `size`");
            li.size = li.diameter;
        }
        double ns = 0;
        if (optnum(l, "samples", ns)) {
            if (!(ns >= 1.0 && ns <= 65536.0) || ns != std::floor(ns))
                throw Err(-2, "invalid value for `samples`, expected an integer in 1..=65536");
            li.samples = (int)ns;
        }
        L.ctx.lights.push_back(li);
    }
    const J& amb = req(ss, "ambient_illumination");
    L.ctx.amb_color = strv(req(amb, "color"), "color");
    L.ctx.amb_intensity = num(req(amb, "intensity"), "intensity");
    const J* fj = ss.get("fog");
    if (fj && fj->t != J::Null) {
        L.ctx.has_fog = true;
        L.ctx.fog.color = strv(req(*fj, "color"), "color");
        L.ctx.fog.density = num(req(*fj, "density"), "density");
        L.ctx.fog.start = num(req(*fj, "start"), "start");
        L.ctx.fog.end = num(req(*fj, "end"), "end");
    }
    const J* bg = ss.get("background_color");
    L.ctx.background = v3(0, 0, 0);
    if (bg && bg->t != J::Null) {
        if (!hex_to_color(strv(*bg, "background_color"), L.ctx.background)) throw Err(-4, "Invalid hex color format");
    }
    double cp[3];
    arr3(req(*L.camj, "position"), cp, "position");
    L.ctx.cam_pos = v3(cp[0], cp[1], cp[2]);
}

// Outline (src/outline.rs:86-278). depth NaN / normal NaN mark None.
void outline_apply(int W, int H, const std::vector<double>& depth, const std::vector<V3>& nrm,
                   const std::vector<unsigned char>& fg, const oracle_opts& o, std::vector<V3>& color) {
    auto get = [&](long x, long y, double& d, V3& n) {
        if (x < 0 || y < 0 || x >= W || y >= H) return false;
        size_t i = (size_t)y * W + x;
        if (!fg[i]) return false;
        d = depth[i];
        n = nrm[i];
        return true;
    };
    std::vector<double> mask((size_t)W * H, 0.0);
    for (long y = 0; y < H; ++y)
        for (long x = 0; x < W; ++x) {
            std::vector<std::pair<long, long>> nb;
            if (o.outline_use_8_neighbors) {
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (!dx && !dy) continue;
                        long nx = x + dx, ny = y + dy;
                        if (nx >= 0 && ny >= 0) nb.emplace_back(nx, ny);
                    }
            } else {
                if (x > 0) nb.emplace_back(x - 1, y);
                nb.emplace_back(x + 1, y);
                if (y > 0) nb.emplace_back(x, y - 1);
                nb.emplace_back(x, y + 1);
            }
            double cd, nd;
            V3 cn, nn;
            bool cf = get(x, y, cd, cn);
            double mx = 0.0;
            for (auto& p : nb) {
                bool nf = get(p.first, p.second, nd, nn);
                double e = 0.0;
                if (cf && nf) {
                    double dd = std::fabs(cd - nd);
                    double ndd = dd / rmax(cd * 0.1, 0.1);
                    double dp = rclamp(dot(cn, nn), -1.0, 1.0);
                    double ndiff = 1.0 - dp;
                    e = o.outline_depth_weight * ndd + o.outline_normal_weight * ndiff;
                } else if (cf != nf) {
                    e = o.outline_depth_weight * 2.0 + o.outline_normal_weight * 2.0;
                }
                mx = rmax(mx, e);
            }
            if (mx > o.outline_threshold)
                mask[(size_t)y * W + x] = (mx - o.outline_threshold) / rmax(1.0 - o.outline_threshold, 0.001);
        }
    if (o.outline_thickness > 1.0) {
        int radius = (int)std::ceil(o.outline_thickness - 1.0);
        if (radius > 0) {
            std::vector<double> dil(mask.size());
            for (long y = 0; y < H; ++y)
                for (long x = 0; x < W; ++x) {
                    double m = mask[(size_t)y * W + x];
                    for (int dy = -radius; dy <= radius; ++dy)
                        for (int dx = -radius; dx <= radius; ++dx) {
                            long nx = x + dx, ny = y + dy;
                            if (nx >= 0 && ny >= 0 && nx < W && ny < H) {
                                double dist = std::sqrt((double)(dx * dx + dy * dy));
                                if (dist <= o.outline_thickness) {
                                    double fall = 1.0 - (dist / o.outline_thickness);
                                    m = rmax(m, mask[(size_t)ny * W + nx] * fall);
                                }
                            }
                        }
                    dil[(size_t)y * W + x] = m;
                }
            mask.swap(dil);
        }
    }
    V3 ec = v3(o.outline_color[0], o.outline_color[1], o.outline_color[2]);
    for (size_t i = 0; i < mask.size(); ++i) {
        if (mask[i] > 0.0) {
            double w = rclamp(rmin(mask[i], 1.0), 0.0, 1.0);
            V3 c = color[i];
            color[i] = v3(c.x * (1.0 - w) + ec.x * w, c.y * (1.0 - w) + ec.y * w, c.z * (1.0 - w) + ec.z * w);
        }
    }
}

template <class F>
void parallel_rows(int rows, int threads, F fn) {
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (;;) {
            int r = next.fetch_add(1);
            if (r >= rows) break;
            fn(r);
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < threads; ++i) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
}

int do_render(const char* scene_path, const oracle_opts* op, uint8_t* rgb, int32_t* hobj, int32_t* htri, double* ht,
              oracle_stats* st, float* feat_out = nullptr, float* rad_out = nullptr) {
    if (op->samples == 0) throw Err(-1, "Samples must be greater than 0");
    std::ifstream f(scene_path, std::ios::binary);
    if (!f) throw Err(-3, std::string("No such file or directory (os error 2): ") + scene_path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    JP jp{text.data(), text.data() + text.size()};
    J root = jp.parse();
    if (root.t != J::Obj) throw Err(-2, "expected a scene object");
    Loaded L;
    build_world(root, scene_path, op->use_kdtree != 0, L);
    L.ctx.world = &L.world;
    L.ctx.light_samples = op->light_samples;
    const int W = op->width, H = op->height;
    double aspect = (double)W / (double)H;
    Camera cam = make_camera(*L.camj, aspect);
    int threads = op->threads > 0 ? op->threads : (int)std::max(1u, std::thread::hardware_concurrency());
    u64 seed = op->has_seed ? op->seed : 0;
    const Ctx& X = L.ctx;
    std::vector<V3> color((size_t)W * H);
    std::vector<Counters> cnt(H + 1);
    auto t0 = std::chrono::steady_clock::now();
    if (op->aa_mode == 1) {
        // render_quincunx (src/renderer.rs:752-897): corner grid, then centres.
        std::vector<V3> corners((size_t)(W + 1) * (H + 1));
        double pw = 1.0 / (double)W, ph = 1.0 / (double)H;
        std::vector<Counters> cc(H + 1);
        parallel_rows(H + 1, threads, [&](int cy) {
            for (int cx = 0; cx <= W; ++cx) {
                double u = rclamp((double)cx * pw, 0.0, 1.0);
                double v = rclamp(1.0 - (double)cy * ph, 0.0, 1.0);
                Ray r = get_ray(cam, u, v);
                u64 s = seed * 0x9E3779B97F4A7C15ull + (u64)cx + (u64)cy * 0x85EBCA6Bull;
                PrimRec rec;
                size_t k = (size_t)cy * (W + 1) + cx;
                corners[k] = ray_color(X, r, op->max_depth, nullptr, s, cc[cy], true, &rec, nullptr, nullptr, nullptr);
                if (hobj) {
                    size_t o = (size_t)W * H + k;
                    hobj[o] = rec.obj; htri[o] = rec.tri; ht[o] = rec.t;
                }
            }
        });
        parallel_rows(H, threads, [&](int y) {
            for (int x = 0; x < W; ++x) {
                double u = ((double)x + 0.5) * pw;
                double v = 1.0 - ((double)y + 0.5) * ph;
                Ray r = get_ray(cam, u, v);
                u64 s = seed * 0x9E3779B97F4A7C15ull + (u64)x * 0x85EBCA6Bull + (u64)y * 0xC2B2AE35ull + 0x12345678ull;
                PrimRec rec;
                Feat fe;
                V3 c = ray_color(X, r, op->max_depth, nullptr, s, cnt[y], true, &rec, nullptr, nullptr, nullptr,
                                 feat_out ? &fe : nullptr);
                size_t k = (size_t)y * W + x;
                if (feat_out) memcpy(feat_out + k * 13, fe.v, sizeof(fe.v));
                if (hobj) { hobj[k] = rec.obj; htri[k] = rec.tri; ht[k] = rec.t; }
                V3 tot = c + corners[(size_t)y * (W + 1) + x];
                tot = tot + corners[(size_t)y * (W + 1) + x + 1];
                tot = tot + corners[(size_t)(y + 1) * (W + 1) + x];
                tot = tot + corners[(size_t)(y + 1) * (W + 1) + x + 1];
                color[k] = tot / 5.0;
            }
        });
        for (auto& c : cc) cnt[H].add(c);
    } else {
        // render_standard / render_standard_with_outline (src/renderer.rs:445-749)
        bool outline = op->outline_enabled != 0;
        std::vector<double> depth;
        std::vector<V3> nrm;
        std::vector<unsigned char> fg;
        if (outline) {
            depth.assign((size_t)W * H, 0.0);
            nrm.assign((size_t)W * H, v3(0, 0, 0));
            fg.assign((size_t)W * H, 0);
        }
        const int N = op->samples;
        parallel_rows(H, threads, [&](int y) {
            for (int x = 0; x < W; ++x) {
                double pu = (double)x / (double)(W - 1);
                double pv = (double)(H - 1 - y) / (double)(H - 1);
                double pw = 1.0 / (double)(W - 1);
                double ph = 1.0 / (double)(H - 1);
                V3 total = v3(0, 0, 0);
                u64 ps = seed * 0x9E3779B97F4A7C15ull + (u64)x * 0x85EBCA6Bull + (u64)y * 0xC2B2AE35ull;
                StdRng rng(ps);
                bool have = false;
                double pdepth = 0;
                V3 pnorm{};
                for (int s = 0; s < N; ++s) {
                    double su = pu, sv = pv;
                    if (op->aa_mode == 2) {
                        if (N == 1) {
                            double ju = rng.gen_f64() - 0.5;
                            double jv = rng.gen_f64() - 0.5;
                            su = pu + ju * pw;
                            sv = pv + jv * ph;
                        } else {
                            double angle = 2.0 * M_PI * (double)s / (double)N;
                            double phase = rng.gen_f64() * 2.0 * M_PI;
                            double rot = angle + phase;
                            double rad = 0.5 * rng.gen_f64();
                            double ju = rad * std::cos(rot);
                            double jv = rad * std::sin(rot);
                            su = pu + ju * pw;
                            sv = pv + jv * ph;
                        }
                    }
                    Ray r = get_ray(cam, su, sv);
                    u64 ss = ps + (u64)s * 0x1F845FEDull;
                    PrimRec rec;
                    bool hd = false;
                    double dd = 0;
                    V3 dn{};
                    Feat fe;
                    V3 c = ray_color(X, r, op->max_depth, &cam, ss, cnt[y], true, &rec, outline ? &hd : nullptr, &dd, &dn,
                                     (feat_out && s == 0) ? &fe : nullptr);
                    if (feat_out && s == 0) memcpy(feat_out + ((size_t)y * W + x) * 13, fe.v, sizeof(fe.v));
                    if (hobj) {
                        size_t o = ((size_t)y * W + x) * N + s;
                        hobj[o] = rec.obj; htri[o] = rec.tri; ht[o] = rec.t;
                    }
                    total = total + c;
                    if (outline && hd) {
                        if (!have || dd < pdepth) {
                            have = true;
                            pdepth = dd;
                            pnorm = dn;
                        }
                    }
                }
                size_t k = (size_t)y * W + x;
                color[k] = total / (double)N;
                if (outline && have) {
                    fg[k] = 1;
                    depth[k] = pdepth;
                    nrm[k] = pnorm;
                }
            }
        });
        if (outline) outline_apply(W, H, depth, nrm, fg, *op, color);
    }
    auto t1 = std::chrono::steady_clock::now();
    for (size_t i = 0; i < color.size(); ++i) {
        float c3[3] = {(float)color[i].x, (float)color[i].y, (float)color[i].z};
        if (rad_out) memcpy(rad_out + 3 * i, c3, sizeof(c3));
        if (feat_out) memcpy(feat_out + 13 * i, c3, sizeof(c3));
    }
    // create_image_from_data (src/renderer.rs:899-912)
    for (size_t i = 0; i < color.size(); ++i) {
        rgb[3 * i + 0] = sat_u8(rclamp(color[i].x, 0.0, 1.0) * 255.0);
        rgb[3 * i + 1] = sat_u8(rclamp(color[i].y, 0.0, 1.0) * 255.0);
        rgb[3 * i + 2] = sat_u8(rclamp(color[i].z, 0.0, 1.0) * 255.0);
    }
    if (st) {
        Counters tot;
        for (auto& c : cnt) tot.add(c);
        st->rays_primary = tot.primary;
        st->rays_reflect = tot.reflect;
        st->rays_shadow = tot.shadow;
        st->kd_inner_visits = tot.kd_inner;
        st->kd_leaf_visits = tot.kd_leaf;
        st->tri_tests = tot.tri;
        st->sphere_tests = tot.sphere;
        st->render_seconds = std::chrono::duration<double>(t1 - t0).count();
        st->threads_used = threads;
    }
    return 0;
}

void set_err(char* err, int errlen, const std::string& m) {
    if (err && errlen > 0) {
        snprintf(err, (size_t)errlen, "%s", m.c_str());
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// AutoCamera (src/auto_camera.rs:9-130) over Scene::compute_finite_bounds
// (src/scene.rs:538-660), straight from the scene JSON.
// ---------------------------------------------------------------------------
struct CamOut {
    bool persp;
    V3 pos, target, up;
    double width, height, fov;
};
void auto_cameras(const std::string& scene_path, CamOut out[4]) {
    std::ifstream f(scene_path, std::ios::binary);
    if (!f) throw Err(-3, std::string("No such file or directory (os error 2): ") + scene_path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    JP jp{text.data(), text.data() + text.size()};
    J root = jp.parse();
    const J& objs = req(root, "objects");
    std::string base = dir_of(scene_path);
    bool any = false;
    V3 mn{}, mx{};
    auto merge = [&](V3 a, V3 b) {
        if (!any) {
            mn = a;
            mx = b;
            any = true;
            return;
        }
        mn = v3(rmin(mn.x, a.x), rmin(mn.y, a.y), rmin(mn.z, a.z));
        mx = v3(rmax(mx.x, b.x), rmax(mx.y, b.y), rmax(mx.z, b.z));
    };
    for (const J& o : objs.a) {
        std::string kind = strv(req(o, "kind"), "kind");
        const J* tj = o.get("transform");
        std::vector<std::string> tl;
        bool has_t = tj && tj->t == J::Arr;
        if (has_t)
            for (auto& s_ : tj->a) tl.push_back(strv(s_, "transform"));
        M4 M;
        bool tok = has_t && parse_transforms(tl, M);
        if (kind == "sphere") {
            double c[3];
            arr3(req(o, "center"), c, "center");
            V3 cp = v3(c[0], c[1], c[2]);
            double r = num(req(o, "radius"), "radius");
            if (tok) {
                cp = m4_point(M, cp);
                double sx = norm(v3(M.at(0, 0), M.at(1, 0), M.at(2, 0)));
                double sy = norm(v3(M.at(0, 1), M.at(1, 1), M.at(2, 1)));
                double sz = norm(v3(M.at(0, 2), M.at(1, 2), M.at(2, 2)));
                r *= rmax(rmax(sx, sy), sz);
            }
            merge(cp - v3(r, r, r), cp + v3(r, r, r));
        } else if (kind == "cube") {
            double c[3], sz[3];
            arr3(req(o, "center"), c, "center");
            arr3(req(o, "size"), sz, "size");
            V3 cp = v3(c[0], c[1], c[2]), h = v3(sz[0], sz[1], sz[2]) / 2.0;
            M4 inv;
            if (tok && !m4_inverse(M, inv)) inv = m4_identity();
            if (!tok || m4_is_identity(inv)) {
                merge(cp - h, cp + h);
            } else {
                V3 a = v3(INFINITY, INFINITY, INFINITY), b = v3(-INFINITY, -INFINITY, -INFINITY);
                for (int k = 0; k < 8; ++k) {
                    V3 corner = v3((k & 4) ? h.x : -h.x, (k & 2) ? h.y : -h.y, (k & 1) ? h.z : -h.z);
                    V3 w = cp + m4_point(M, corner);
                    a = v3(rmin(a.x, w.x), rmin(a.y, w.y), rmin(a.z, w.z));
                    b = v3(rmax(b.x, w.x), rmax(b.y, w.y), rmax(b.z, w.z));
                }
                merge(a, b);
            }
        } else if (kind == "mesh") {
            std::string fn = strv(req(o, "filename"), "filename");
            std::string path = (fn.size() && fn[0] == '/') ? fn : base + "/" + fn;
            std::vector<Tri> tris = load_stl(path);
            V3 a = tris[0].v[0], b = tris[0].v[0];
            for (const Tri& t : tris)
                for (const V3& v : t.v) {
                    a = v3(rmin(a.x, v.x), rmin(a.y, v.y), rmin(a.z, v.z));
                    b = v3(rmax(b.x, v.x), rmax(b.y, v.y), rmax(b.z, v.z));
                }
            if (tok) {
                V3 cs[8] = {v3(a.x, a.y, a.z), v3(a.x, a.y, b.z), v3(a.x, b.y, a.z), v3(a.x, b.y, b.z),
                            v3(b.x, a.y, a.z), v3(b.x, a.y, b.z), v3(b.x, b.y, a.z), v3(b.x, b.y, b.z)};
                V3 na = m4_point(M, cs[0]), nb = na;
                for (int k = 1; k < 8; ++k) {
                    V3 w = m4_point(M, cs[k]);
                    na = v3(rmin(na.x, w.x), rmin(na.y, w.y), rmin(na.z, w.z));
                    nb = v3(rmax(nb.x, w.x), rmax(nb.y, w.y), rmax(nb.z, w.z));
                }
                a = na;
                b = nb;
            }
            merge(a, b);
        }
    }
    if (!any) throw Err(-2, "Scene has no finite objects to compute bounds");
    V3 center = v3((mn.x + mx.x) / 2.0, (mn.y + mx.y) / 2.0, (mn.z + mx.z) / 2.0);
    V3 size = mx - mn;
    double vw = rmax(rmax(size.x, size.y), size.z) * 1.15;
    double d = vw * 2.0;
    out[0] = {false, v3(center.x, center.y - d, center.z), center, v3(0, 0, 1), vw, vw, 0};
    out[1] = {false, v3(center.x - d, center.y, center.z), center, v3(0, 0, 1), vw, vw, 0};
    out[2] = {false, v3(center.x, center.y, center.z + d), center, v3(0, 1, 0), vw, vw, 0};
    double dist = vw / std::tan(50.0 * (M_PI / 180.0) / 2.0);
    double hd = dist * std::cos(35.0 * (M_PI / 180.0));
    V3 pp = v3(center.x + hd * std::cos(45.0 * (M_PI / 180.0)), center.y + hd * std::sin(45.0 * (M_PI / 180.0)),
               center.z + dist * std::sin(35.0 * (M_PI / 180.0)));
    out[3] = {true, pp, center, v3(0, 0, 1), 1.0, 1.0, 50.0};
}

extern "C" {

int oracle_render_aux(const char* scene_path, const oracle_opts* opts, uint8_t* rgb_out, float* features,
                      float* radiance, oracle_stats* stats, char* err, int errlen) {
    try {
        return do_render(scene_path, opts, rgb_out, nullptr, nullptr, nullptr, stats, features, radiance);
    } catch (const Err& e) {
        set_err(err, errlen, e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return -99;
    }
}

int oracle_render(const char* scene_path, const oracle_opts* opts, uint8_t* rgb_out, int32_t* hit_obj,
                  int32_t* hit_tri, double* hit_t, oracle_stats* stats, char* err, int errlen) {
    try {
        return do_render(scene_path, opts, rgb_out, hit_obj, hit_tri, hit_t, stats);
    } catch (const Err& e) {
        set_err(err, errlen, e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return -99;
    }
}

int oracle_load_stl(const char* path, double* out, int64_t capacity_tris, int64_t* n_tris, char* err, int errlen) {
    try {
        std::vector<Tri> t = load_stl(path);
        *n_tris = (int64_t)t.size();
        if (out)
            for (size_t k = 0; k < t.size() && (int64_t)k < capacity_tris; ++k)
                for (int j = 0; j < 3; ++j) {
                    out[k * 9 + j * 3 + 0] = t[k].v[j].x;
                    out[k * 9 + j * 3 + 1] = t[k].v[j].y;
                    out[k * 9 + j * 3 + 2] = t[k].v[j].z;
                }
        return 0;
    } catch (const Err& e) {
        set_err(err, errlen, e.what());
        return e.code;
    }
}

int oracle_auto_cameras(const char* scene_path, double* out, char* err, int errlen) {
    try {
        CamOut c[4];
        auto_cameras(scene_path, c);
        for (int k = 0; k < 4; ++k) {
            double* o = out + 14 * k;
            o[0] = c[k].persp ? 1 : 0;
            o[1] = c[k].pos.x; o[2] = c[k].pos.y; o[3] = c[k].pos.z;
            o[4] = c[k].target.x; o[5] = c[k].target.y; o[6] = c[k].target.z;
            o[7] = c[k].up.x; o[8] = c[k].up.y; o[9] = c[k].up.z;
            o[10] = c[k].width; o[11] = c[k].height; o[12] = c[k].fov; o[13] = 0;
        }
        return 0;
    } catch (const Err& e) {
        set_err(err, errlen, e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return -99;
    }
}

void oracle_chacha_block(const uint32_t in[16], int double_rounds, uint32_t out[16]) {
    chacha_block(in, double_rounds, out);
}
void oracle_seed_key(uint64_t seed, uint32_t key_out[8]) { seed_key(seed, key_out); }
void oracle_stdrng_u64(uint64_t seed, int n, uint64_t* out) {
    StdRng r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next_u64();
}
void oracle_stdrng_f64(uint64_t seed, int n, double* out) {
    StdRng r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.gen_f64();
}
void oracle_stdrng_construction(const uint8_t seed[32], uint64_t out[2]) {
    u32 key[8];
    for (int i = 0; i < 8; ++i)
        key[i] = (u32)seed[4 * i] | (u32)seed[4 * i + 1] << 8 | (u32)seed[4 * i + 2] << 16 | (u32)seed[4 * i + 3] << 24;
    StdRng r0(key);
    out[0] = r0.next_u64();
    u32 key1[8];
    r0.fill_words(key1, 8);  // StdRng::from_rng(r0): fill_bytes of a 32-byte seed
    StdRng r1(key1);
    out[1] = r1.next_u64();
}
void oracle_pcg32_f64(uint64_t state, uint64_t stream, int kind, double low, double high, int n, double* out) {
    Pcg32 r(state, stream);
    for (int i = 0; i < n; ++i) out[i] = kind == 0 ? gen_f64_of(r) : gen_range_of(r, low, high);
}
void oracle_area_samples(uint64_t path_seed, const double p[3], const double light_pos[3], int square,
                         double extent, int n, double* out) {
    Light L;
    L.pos = v3(light_pos[0], light_pos[1], light_pos[2]);
    L.square = square != 0;
    L.has_d = !L.square;
    L.diameter = extent;
    L.size = extent;
    V3 hp = v3(p[0], p[1], p[2]);
    u64 ls = light_seed(path_seed, hp);
    StdRng rng(ls);
    const int cols = jitter_cols(n), rows = n / cols;
    for (int s = 0; s < n; ++s) {
        V3 q = area_sample(L, hp, rng, ls, s, cols, rows);
        out[3 * s] = q.x;
        out[3 * s + 1] = q.y;
        out[3 * s + 2] = q.z;
    }
}
uint64_t oracle_splitmix64_nth(uint64_t seed, uint64_t k) { return splitmix64_nth(seed, k); }
int oracle_jitter_cols(int n) { return jitter_cols(n); }
void oracle_sample_disk(uint64_t seed, double radius, int n, double* out) {
    StdRng r(seed);
    for (int i = 0; i < n; ++i) {
        double x, y;
        for (;;) {
            x = r.gen_range(-radius, radius);
            y = r.gen_range(-radius, radius);
            if (x * x + y * y <= radius * radius) break;
        }
        out[2 * i] = x;
        out[2 * i + 1] = y;
    }
}
int oracle_parse_transforms(const char* joined, double m_out[16]) {
    std::vector<std::string> list;
    std::string s(joined);
    size_t st = 0;
    while (st <= s.size() && !s.empty()) {
        size_t k = s.find('\n', st);
        list.push_back(s.substr(st, k == std::string::npos ? std::string::npos : k - st));
        if (k == std::string::npos) break;
        st = k + 1;
    }
    M4 m;
    if (!parse_transforms(list, m)) return -1;
    memcpy(m_out, m.m, sizeof(m.m));
    return 0;
}
int oracle_hex_to_color(const char* hex, double out[3]) {
    V3 c;
    if (!hex_to_color(hex, c)) return -1;
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
    return 0;
}
int oracle_kd_stats(const char* stl_path, int32_t* n_tris, int32_t* n_leaves, int32_t* max_leaf, int32_t* n_nodes,
                    char* err, int errlen) {
    try {
        auto tris = load_stl(stl_path);
        KdTree kt = kd_new(tris);
        int leaves = 0, mx = 0;
        for (auto& n : kt.nodes)
            if (n.leaf) { ++leaves; mx = std::max(mx, (int)n.tris.size()); }
        *n_tris = (int32_t)tris.size();
        *n_leaves = leaves;
        *max_leaf = mx;
        *n_nodes = (int32_t)kt.nodes.size();
        return 0;
    } catch (const Err& e) {
        set_err(err, errlen, e.what());
        return e.code;
    }
}
int oracle_kd_tree(const double* verts, int64_t n_tris, int32_t* leaf, int32_t* axis, double* split, int32_t* right,
                   int32_t* first, int32_t* count, double* bounds, int64_t node_cap, int32_t* tri_ids, int64_t id_cap,
                   int64_t* n_nodes, int64_t* n_ids) {
    // KdTree::new (src/mesh.rs:75-101) on the given triangles; kd_build pushes a node before
    // its children, left subtree first, so nodes[] is in depth-first pre-order
    std::vector<Tri> T((size_t)n_tris);
    for (int64_t i = 0; i < n_tris; ++i)
        for (int k = 0; k < 3; ++k) T[i].v[k] = v3(verts[9 * i + 3 * k], verts[9 * i + 3 * k + 1], verts[9 * i + 3 * k + 2]);
    KdTree kt = kd_new(T);
    int64_t ids = 0;
    for (size_t i = 0; i < kt.nodes.size(); ++i) {
        const KdNode& n = kt.nodes[i];
        if ((int64_t)i < node_cap) {
            leaf[i] = n.leaf;
            axis[i] = n.leaf ? -1 : n.axis;
            split[i] = n.leaf ? 0.0 : n.split;
            right[i] = n.leaf ? -1 : n.right;
            first[i] = n.leaf ? (int32_t)ids : -1;
            count[i] = n.leaf ? (int32_t)n.tris.size() : 0;
            const double b[6] = {n.bmin.x, n.bmin.y, n.bmin.z, n.bmax.x, n.bmax.y, n.bmax.z};
            for (int k = 0; k < 6; ++k) bounds[6 * i + k] = b[k];
        }
        if (n.leaf)
            for (int t : n.tris) {
                if (ids < id_cap) tri_ids[ids] = t;
                ++ids;
            }
    }
    *n_nodes = (int64_t)kt.nodes.size();
    *n_ids = ids;
    return 0;
}
int oracle_mesh_hit(const char* stl_path, int use_kdtree, const double origin[3], const double dir[3], double t_min,
                    double t_max, double* t_out) {
    static std::map<std::string, std::shared_ptr<Obj>> cache;
    static std::mutex mu;
    std::shared_ptr<Obj> ob;
    {
        std::lock_guard<std::mutex> g(mu);
        auto& slot = cache[stl_path];
        if (!slot) {
            slot = std::make_shared<Obj>();
            slot->kind = MESH;
            slot->index = 0;
            slot->tris = load_stl(stl_path);
            slot->bmin = v3(INFINITY, INFINITY, INFINITY);
            slot->bmax = v3(-INFINITY, -INFINITY, -INFINITY);
            for (Tri& t : slot->tris)
                for (int k = 0; k < 3; ++k) {
                    slot->bmin = v3(rmin(slot->bmin.x, t.v[k].x), rmin(slot->bmin.y, t.v[k].y), rmin(slot->bmin.z, t.v[k].z));
                    slot->bmax = v3(rmax(slot->bmax.x, t.v[k].x), rmax(slot->bmax.y, t.v[k].y), rmax(slot->bmax.z, t.v[k].z));
                }
            slot->kd = kd_new(slot->tris);
        }
        ob = slot;
    }
    Obj m = *ob;
    m.use_kd = use_kdtree != 0;
    Ray r = make_ray(v3(origin[0], origin[1], origin[2]), v3(dir[0], dir[1], dir[2]));
    Hit h;
    Counters C;
    if (!mesh_hit(m, r, t_min, t_max, h, C)) return -1;
    *t_out = h.t;
    return h.tri;
}
void oracle_outline(int32_t w, int32_t h, const double* depth, const double* normal, const oracle_opts* opts,
                    double* color_inout) {
    size_t n = (size_t)w * h;
    std::vector<double> d(n);
    std::vector<V3> nr(n), col(n);
    std::vector<unsigned char> fg(n);
    for (size_t i = 0; i < n; ++i) {
        fg[i] = !std::isnan(depth[i]);
        d[i] = depth[i];
        nr[i] = v3(normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]);
        col[i] = v3(color_inout[3 * i], color_inout[3 * i + 1], color_inout[3 * i + 2]);
    }
    outline_apply(w, h, d, nr, fg, *opts, col);
    for (size_t i = 0; i < n; ++i) {
        color_inout[3 * i] = col[i].x;
        color_inout[3 * i + 1] = col[i].y;
        color_inout[3 * i + 2] = col[i].z;
    }
}

}  // extern "C"
