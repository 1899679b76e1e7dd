// scene_load.cpp — Scene JSON / STL -> HostScene (host glue feeding the trace path).
//
// Follows the reference's serde model (src/scene.rs:34-446), its World assembly
// (src/renderer.rs:137-317: transform baking per object kind, object index ==
// material index) and its STL loader (src/mesh.rs:751-1003). Colours are parsed
// once here; the shade-time fallbacks of the reference (unwrap_or white for
// material/ambient/light colours, grey for fog, black for grid lines) are applied
// at the same points.

#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../common/vec.hpp"
#include "json.hpp"
#include "scene_host.hpp"

namespace rt {

namespace {

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(RT_E_IO, "No such file or directory (os error 2): " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// Byte length of the code point at p when it has Unicode's White_Space property (what
// Rust's str::trim and split_whitespace remove), else 0.
size_t ws_len(const unsigned char* p, const unsigned char* e) {
    const unsigned c = p[0];
    const ptrdiff_t left = e - p;
    if (c == ' ' || (c >= 0x09 && c <= 0x0D)) return 1;
    if (c == 0xC2) return left >= 2 && (p[1] == 0x85 || p[1] == 0xA0) ? 2 : 0;  // U+0085, U+00A0
    if (left < 3) return 0;
    if (c == 0xE1) return p[1] == 0x9A && p[2] == 0x80 ? 3 : 0;  // U+1680
    if (c == 0xE3) return p[1] == 0x80 && p[2] == 0x80 ? 3 : 0;  // U+3000
    if (c != 0xE2) return 0;
    if (p[1] == 0x81) return p[2] == 0x9F ? 3 : 0;  // U+205F
    if (p[1] != 0x80) return 0;
    const unsigned d = p[2];  // U+2000..U+200A, U+2028, U+2029, U+202F
    return (d >= 0x80 && d <= 0x8A) || d == 0xA8 || d == 0xA9 || d == 0xAF ? 3 : 0;
}

// str::trim
std::string strip(const std::string& s) {
    const unsigned char* p = reinterpret_cast<const unsigned char*>(s.data());
    size_t a = 0, b = s.size();
    for (size_t k; a < b && (k = ws_len(p + a, p + b)) != 0;) a += k;
    for (bool more = true; more && b > a;) {
        more = false;
        for (size_t k = 1; k <= 3 && k <= b - a; ++k)
            if (ws_len(p + b - k, p + b) == k) {
                b -= k;
                more = true;
                break;
            }
    }
    return s.substr(a, b - a);
}

// str::split_whitespace
std::vector<std::string> split_ws(const std::string& s) {
    const unsigned char* p = reinterpret_cast<const unsigned char*>(s.data());
    const unsigned char* e = p + s.size();
    std::vector<std::string> w;
    size_t i = 0, start = std::string::npos;
    while (i < s.size()) {
        const size_t k = ws_len(p + i, e);
        if (k == 0) {
            if (start == std::string::npos) start = i;
            ++i;
            continue;
        }
        if (start != std::string::npos) w.push_back(s.substr(start, i - start));
        start = std::string::npos;
        i += k;
    }
    if (start != std::string::npos) w.push_back(s.substr(start));
    return w;
}

// Rust's f64 FromStr grammar: [+-] (inf | infinity | nan | digits[.digits][e[+-]digits]).
bool parse_rust_f64(const std::string& s, double& out) {
    if (s.empty()) return false;
    size_t i = (s[0] == '+' || s[0] == '-') ? 1 : 0;
    std::string body;
    for (size_t k = i; k < s.size(); ++k) body += (char)std::tolower((unsigned char)s[k]);
    bool neg = s[0] == '-';
    if (body == "inf" || body == "infinity") { out = neg ? -INFINITY : INFINITY; return true; }
    if (body == "nan") { out = NAN; return true; }
    size_t j = 0, digits = 0;
    while (j < body.size() && std::isdigit((unsigned char)body[j])) { ++j; ++digits; }
    if (j < body.size() && body[j] == '.') {
        ++j;
        while (j < body.size() && std::isdigit((unsigned char)body[j])) { ++j; ++digits; }
    }
    if (!digits) return false;
    if (j < body.size() && body[j] == 'e') {
        ++j;
        if (j < body.size() && (body[j] == '+' || body[j] == '-')) ++j;
        size_t e0 = j;
        while (j < body.size() && std::isdigit((unsigned char)body[j])) ++j;
        if (j == e0) return false;
    }
    if (j != body.size()) return false;
    out = std::strtod(s.c_str(), nullptr);
    return true;
}

using rtj::Value;

const Value& field(const Value& o, const char* k) {
    if (o.kind != Value::kObject) throw Error(RT_E_SCENE, std::string("invalid type: expected struct containing `") + k + "`");
    const Value* v = o.find(k);
    if (!v) throw Error(RT_E_SCENE, std::string("missing field `") + k + "`");
    return *v;
}
const Value* opt_field(const Value& o, const char* k) {
    const Value* v = o.find(k);
    return (v && !v->is_null()) ? v : nullptr;
}
double as_f64(const Value& v, const char* k) {
    if (v.kind != Value::kNumber) throw Error(RT_E_SCENE, std::string("invalid type for `") + k + "`, expected f64");
    return v.number;
}
std::string as_str(const Value& v, const char* k) {
    if (v.kind != Value::kString) throw Error(RT_E_SCENE, std::string("invalid type for `") + k + "`, expected a string");
    return v.text;
}
void as_vec3(const Value& v, const char* k, double out[3]) {
    if (v.kind != Value::kArray || v.items.size() != 3)
        throw Error(RT_E_SCENE, std::string("invalid length for `") + k + "`, expected an array of length 3");
    for (int i = 0; i < 3; ++i) out[i] = as_f64(v.items[i], k);
}

// Rotation3::from_euler_angles(roll, pitch, yaw) as a 4x4 (column-major).
void euler4(double roll, double pitch, double yaw, double m[16]) {
    double sr = std::sin(roll), cr = std::cos(roll);
    double sp = std::sin(pitch), cp = std::cos(pitch);
    double sy = std::sin(yaw), cy = std::cos(yaw);
    double r[3][3] = {{cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr},
                      {sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr},
                      {-sp, cp * sr, cp * cr}};
    for (int i = 0; i < 16; ++i) m[i] = 0.0;
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) m[col * 4 + row] = r[row][col];
    m[15] = 1.0;
}
void ident4(double m[16]) {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0 : 0.0;
}
// C = A * B (column-major), each column by nalgebra's gemv order.
void mul4(const double A[16], const double B[16], double C[16]) {
    double out[16];
    for (int j = 0; j < 4; ++j) {
        for (int r = 0; r < 4; ++r) {
            double acc = A[0 * 4 + r] * B[j * 4 + 0];
            acc = A[1 * 4 + r] * B[j * 4 + 1] + acc;
            acc = A[2 * 4 + r] * B[j * 4 + 2] + acc;
            acc = A[3 * 4 + r] * B[j * 4 + 3] + acc;
            out[j * 4 + r] = acc;
        }
    }
    memcpy(C, out, sizeof(out));
}

bool parse_transform(const std::string& raw, double m[16]) {
    std::string s = strip(raw);
    static const char* names[3] = {"rotate(", "translate(", "scale("};
    int kind = -1;
    std::string args;
    for (int k = 0; k < 3; ++k) {
        size_t n = strlen(names[k]);
        if (s.size() >= n && s.compare(0, n, names[k]) == 0) {
            kind = k;
            args = s.substr(n);
            break;
        }
    }
    if (kind < 0 || args.empty() || args.back() != ')') return false;
    args.pop_back();
    std::vector<double> v;
    size_t pos = 0;
    for (;;) {
        size_t comma = args.find(',', pos);
        double x;
        if (!parse_rust_f64(strip(args.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos)), x))
            return false;
        v.push_back(x);
        if (comma == std::string::npos) break;
        pos = comma + 1;
    }
    if (v.size() != 3) return false;
    if (kind == 0) {
        const double k = M_PI / 180.0;  // f64::to_radians
        double rx[16], ry[16], rz[16], t[16];
        euler4(v[0] * k, 0.0, 0.0, rx);
        euler4(0.0, v[1] * k, 0.0, ry);
        euler4(0.0, 0.0, v[2] * k, rz);
        mul4(rz, ry, t);
        mul4(t, rx, m);
    } else if (kind == 1) {
        ident4(m);
        m[12] = v[0];
        m[13] = v[1];
        m[14] = v[2];
    } else {
        ident4(m);
        m[0] = v[0];
        m[5] = v[1];
        m[10] = v[2];
    }
    return true;
}

struct MaterialIn {
    std::string color;
    double ambient, diffuse, specular, shininess;
    bool has_refl = false;
    double refl = 0;
    int texture = 0;
    std::string line_color;
    double line_width = 0, cell_size = 0;
    std::vector<MaterialIn> alt;  // checkerboard material_b (0 or 1 entries)
};

MaterialIn read_material(const Value& m) {
    MaterialIn r;
    r.color = as_str(field(m, "color"), "color");
    r.ambient = as_f64(field(m, "ambient"), "ambient");
    r.diffuse = as_f64(field(m, "diffuse"), "diffuse");
    r.specular = as_f64(field(m, "specular"), "specular");
    r.shininess = as_f64(field(m, "shininess"), "shininess");
    if (const Value* x = opt_field(m, "reflectivity")) {
        r.has_refl = true;
        r.refl = as_f64(*x, "reflectivity");
    }
    if (const Value* t = opt_field(m, "texture")) {
        std::string type = as_str(field(*t, "type"), "type");
        if (type == "grid") {
            r.texture = 1;
            r.line_color = as_str(field(*t, "line_color"), "line_color");
            r.line_width = as_f64(field(*t, "line_width"), "line_width");
            r.cell_size = as_f64(field(*t, "cell_size"), "cell_size");
        } else if (type == "checkerboard") {
            r.texture = 2;
            r.alt.push_back(read_material(field(*t, "material_b")));
        } else {
            throw Error(RT_E_SCENE, "unknown variant `" + type + "`, expected `grid` or `checkerboard`");
        }
    }
    return r;
}

rt_material_desc lower_material(const MaterialIn& m, bool base) {
    rt_material_desc d{};
    if (!hex_color(m.color, d.color)) {
        // base colours fail the World build (src/renderer.rs:175,222,240,299);
        // a checkerboard alternate falls back to white at shade time (src/lighting.rs:241-242)
        if (base) throw Error(RT_E_CAMERA, "Invalid hex color format");
        d.color[0] = d.color[1] = d.color[2] = 1.0;
    }
    d.ambient = m.ambient;
    d.diffuse = m.diffuse;
    d.specular = m.specular;
    d.shininess = m.shininess;
    d.has_reflectivity = m.has_refl;
    d.reflectivity = m.refl;
    d.texture = m.texture;
    d.line_width = m.line_width;
    d.cell_size = m.cell_size;
    d.material_b = -1;
    if (m.texture == 1) {
        double g[3];
        if (!hex_color(m.line_color, g)) g[0] = g[1] = g[2] = 0.0;  // unwrap_or black (src/lighting.rs:16)
        // "#{:02X}" of (c * 255) as u8, parsed back (src/lighting.rs:28-36)
        for (int i = 0; i < 3; ++i) d.grid_line_color[i] = (double)to_u8_sat(g[i] * 255.0) / 255.0;
    }
    return d;
}

std::string parent_dir(const std::string& path) {
    size_t k = path.find_last_of('/');
    if (k == std::string::npos) return ".";
    if (k == 0) return "/";
    return path.substr(0, k);
}

}  // namespace

bool hex_color(const std::string& s, double out[3]) {
    size_t i = 0;
    while (i < s.size() && s[i] == '#') ++i;  // trim_start_matches('#')
    std::string h = s.substr(i);
    if (h.size() != 6) return false;
    for (int k = 0; k < 3; ++k) {
        std::string pair = h.substr(2 * k, 2);
        size_t p = 0;
        if (pair[0] == '+') p = 1;  // u8::from_str_radix accepts a leading '+'
        if (p >= pair.size()) return false;
        unsigned v = 0;
        for (; p < pair.size(); ++p) {
            char c = pair[p];
            unsigned d;
            if (c >= '0' && c <= '9') d = (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') d = (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') d = (unsigned)(c - 'A' + 10);
            else return false;
            v = v * 16 + d;
        }
        out[k] = (double)v / 255.0;
    }
    return true;
}

bool parse_transform_list(const std::vector<std::string>& list, double out[16]) {
    double acc[16];
    ident4(acc);
    for (const auto& s : list) {
        double m[16];
        if (!parse_transform(s, m)) return false;
        mul4(m, acc, acc);
    }
    memcpy(out, acc, sizeof(acc));
    return true;
}

bool invert4(const double m[16], double out[16]) {
    double c[16];
    c[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    c[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    c[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    c[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    c[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    c[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    c[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    c[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    c[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    c[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    c[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    c[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    c[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    c[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    c[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    c[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * c[0] + m[1] * c[4] + m[2] * c[8] + m[3] * c[12];
    if (det == 0.0) return false;
    double inv = 1.0 / det;
    for (int i = 0; i < 16; ++i) out[i] = c[i] * inv;
    return true;
}

// Whether bytes [b, e) are well-formed UTF-8 (what Rust's String::from_utf8 accepts).
bool valid_utf8(const unsigned char* b, const unsigned char* e) {
    while (b < e) {
        unsigned c = *b;
        int n;
        unsigned lo = 0x80, hi = 0xBF;
        if (c < 0x80) { ++b; continue; }
        if (c >= 0xC2 && c <= 0xDF) n = 1;
        else if (c == 0xE0) { n = 2; lo = 0xA0; }
        else if (c == 0xED) { n = 2; hi = 0x9F; }
        else if (c >= 0xE1 && c <= 0xEF) n = 2;
        else if (c == 0xF0) { n = 3; lo = 0x90; }
        else if (c >= 0xF1 && c <= 0xF3) n = 3;
        else if (c == 0xF4) { n = 3; hi = 0x8F; }
        else return false;
        if (e - b <= n) return false;
        if (b[1] < lo || b[1] > hi) return false;
        for (int k = 2; k <= n; ++k)
            if (b[k] < 0x80 || b[k] > 0xBF) return false;
        b += n + 1;
    }
    return true;
}

std::vector<double> load_stl_file(const std::string& path) {
    std::string bytes = read_file(path);
    if (bytes.size() < 80) throw Error(RT_E_IO, "failed to fill whole buffer");
    std::vector<double> out;
    const unsigned char* data = reinterpret_cast<const unsigned char*>(bytes.data());
    // ASCII when the 80-byte header starts with "solid" (after leading whitespace) and
    // one of the first 10 lines is a facet/loop marker (src/mesh.rs:754-828). The probe
    // reads whole lines as UTF-8 strings (BufRead::lines, up to an 11th line before it
    // stops), so a binary file whose header starts with "solid" usually fails there.
    bool ascii = false;
    {
        // String::from_utf8_lossy(header).trim_start(): an ill-formed byte becomes U+FFFD,
        // which is not whitespace, so trimming the raw bytes gives the same prefix
        size_t a = 0;
        for (size_t k; a < 80 && (k = ws_len(data + a, data + 80)) != 0;) a += k;
        if (bytes.compare(a, 5, "solid") == 0 && a + 5 <= 80) {
            size_t pos = 0;
            for (int line = 0; pos < bytes.size(); ++line) {
                size_t nl = bytes.find('\n', pos);
                size_t end = nl == std::string::npos ? bytes.size() : nl;
                if (!valid_utf8(data + pos, data + end)) throw Error(RT_E_IO, "stream did not contain valid UTF-8");
                if (line == 10) break;
                std::string l = strip(bytes.substr(pos, end - pos));
                if (l.compare(0, 12, "facet normal") == 0 || l == "outer loop" || l == "endloop") { ascii = true; break; }
                if (nl == std::string::npos) break;
                pos = nl + 1;
            }
        }
    }
    if (ascii) {
        // read_to_string: the whole file must be UTF-8
        if (!valid_utf8(data, data + bytes.size())) throw Error(RT_E_IO, "stream did not contain valid UTF-8");
        std::vector<std::string> lines;
        size_t pos = 0;
        while (pos < bytes.size()) {
            size_t nl = bytes.find('\n', pos);
            std::string l = bytes.substr(pos, nl == std::string::npos ? std::string::npos : nl - pos);
            if (!l.empty() && l.back() == '\r') l.pop_back();
            lines.push_back(strip(l));
            if (nl == std::string::npos) break;
            pos = nl + 1;
        }
        for (size_t i = 0; i < lines.size(); ++i) {
            if (lines[i].compare(0, 12, "facet normal") != 0) continue;
            const auto nw = split_ws(lines[i]);
            if (nw.size() != 5) throw Error(RT_E_IO, "Invalid facet normal format");
            double normal;  // parsed for its errors only: the path never reads it
            for (int k = 2; k < 5; ++k)
                if (!parse_rust_f64(nw[k], normal)) throw Error(RT_E_IO, "invalid float literal");
            if (++i >= lines.size() || lines[i] != "outer loop") throw Error(RT_E_IO, "Expected 'outer loop' after facet normal");
            double tri[9];
            for (int v = 0; v < 3; ++v) {
                if (++i >= lines.size()) throw Error(RT_E_IO, "Unexpected end of file while reading vertex");
                if (lines[i].compare(0, 6, "vertex") != 0) throw Error(RT_E_IO, "Expected vertex line");
                auto w = split_ws(lines[i]);
                if (w.size() != 4) throw Error(RT_E_IO, "Invalid vertex format");
                for (int k = 0; k < 3; ++k)
                    if (!parse_rust_f64(w[k + 1], tri[3 * v + k])) throw Error(RT_E_IO, "invalid float literal");
            }
            if (++i >= lines.size() || lines[i] != "endloop") throw Error(RT_E_IO, "Expected 'endloop'");
            if (++i >= lines.size() || lines[i] != "endfacet") throw Error(RT_E_IO, "Expected 'endfacet'");
            out.insert(out.end(), tri, tri + 9);
        }
        return out;
    }
    if (bytes.size() < 84) throw Error(RT_E_IO, "Binary STL too short");
    uint32_t n;
    memcpy(&n, bytes.data() + 80, 4);
    size_t need = 84 + (size_t)n * 50;
    if (bytes.size() < need)
        throw Error(RT_E_IO, "Binary STL size mismatch: expected " + std::to_string(need) + ", got " + std::to_string(bytes.size()));
    out.resize((size_t)n * 9);
    const char* p = bytes.data() + 84;
    for (uint32_t t = 0; t < n; ++t, p += 50) {
        float v[9];
        memcpy(v, p + 12, 36);  // skip the facet normal, which the path never reads
        for (int k = 0; k < 9; ++k) out[(size_t)t * 9 + k] = (double)v[k];
    }
    return out;
}

// Cube::bounds (src/ray.rs:228-262): the plain box when untransformed, else the box of
// center + M (corner, 1) over the 8 corners.
void cube_bounds(const rt_object_desc& d, ObjBounds& ob) {
    bool ident = true;
    for (int i = 0; i < 16; ++i)
        if (d.world_to_local[i] != ((i % 5 == 0) ? 1.0 : 0.0)) ident = false;
    ob.finite = true;
    if (ident) {
        for (int i = 0; i < 3; ++i) {
            ob.lo[i] = d.center[i] - d.half_size[i];
            ob.hi[i] = d.center[i] + d.half_size[i];
        }
        return;
    }
    for (int i = 0; i < 3; ++i) {
        ob.lo[i] = INFINITY;
        ob.hi[i] = -INFINITY;
    }
    const double* h = d.half_size;
    const double cs[8][3] = {{-h[0], -h[1], -h[2]}, {-h[0], -h[1], h[2]}, {-h[0], h[1], -h[2]}, {-h[0], h[1], h[2]},
                             {h[0], -h[1], -h[2]},  {h[0], -h[1], h[2]},  {h[0], h[1], -h[2]},  {h[0], h[1], h[2]}};
    for (const auto& c : cs) {
        d3 w = m4_apply(d.local_to_world, mk(c[0], c[1], c[2]), 1.0);
        const double wc[3] = {d.center[0] + w.x, d.center[1] + w.y, d.center[2] + w.z};
        for (int i = 0; i < 3; ++i) {
            ob.lo[i] = fmin_r(ob.lo[i], wc[i]);
            ob.hi[i] = fmax_r(ob.hi[i], wc[i]);
        }
    }
}

// Mesh::bounds (src/mesh.rs:26-36) of the loaded vertices; with a transform, the box of
// its 8 corners (in the order of src/scene.rs:612-621) through M (src/scene.rs:599-648).
void mesh_bounds(const std::vector<double>& v, const double* M, ObjBounds& ob) {
    if (v.empty()) return;  // an empty STL is rejected by the loader before this
    for (int i = 0; i < 3; ++i) ob.lo[i] = ob.hi[i] = v[i];
    for (size_t k = 3; k < v.size(); k += 3)
        for (int i = 0; i < 3; ++i) {
            ob.lo[i] = fmin_r(ob.lo[i], v[k + i]);
            ob.hi[i] = fmax_r(ob.hi[i], v[k + i]);
        }
    ob.finite = true;
    if (!M) return;
    const double* a = ob.lo;
    const double* b = ob.hi;
    const double cs[8][3] = {{a[0], a[1], a[2]}, {a[0], a[1], b[2]}, {a[0], b[1], a[2]}, {a[0], b[1], b[2]},
                             {b[0], a[1], a[2]}, {b[0], a[1], b[2]}, {b[0], b[1], a[2]}, {b[0], b[1], b[2]}};
    double lo[3], hi[3];
    for (int k = 0; k < 8; ++k) {
        d3 w = m4_apply(M, mk(cs[k][0], cs[k][1], cs[k][2]), 1.0);
        const double wc[3] = {w.x, w.y, w.z};
        for (int i = 0; i < 3; ++i) {
            lo[i] = k == 0 ? wc[i] : fmin_r(lo[i], wc[i]);
            hi[i] = k == 0 ? wc[i] : fmax_r(hi[i], wc[i]);
        }
    }
    for (int i = 0; i < 3; ++i) {
        ob.lo[i] = lo[i];
        ob.hi[i] = hi[i];
    }
}

HostScene load_scene_json(const std::string& text, const std::string& base_dir) {
    Value root;
    try {
        root = rtj::parse(text);
    } catch (const rtj::ParseError& e) {
        throw Error(RT_E_SCENE, e.what());
    }
    HostScene S;
    const Value& cam = field(root, "camera");
    const Value& objs = field(root, "objects");
    const Value& lights = field(root, "lights");
    const Value& ss = field(root, "scene_settings");
    if (objs.kind != Value::kArray) throw Error(RT_E_SCENE, "invalid type for `objects`, expected a sequence");
    if (lights.kind != Value::kArray) throw Error(RT_E_SCENE, "invalid type for `lights`, expected a sequence");

    // camera (src/scene.rs:34-46)
    std::string kind = as_str(field(cam, "kind"), "kind");
    as_vec3(field(cam, "position"), "position", S.camera.position);
    as_vec3(field(cam, "target"), "target", S.camera.target);
    as_vec3(field(cam, "up"), "up", S.camera.up);
    S.camera.width = as_f64(field(cam, "width"), "width");
    S.camera.height = as_f64(field(cam, "height"), "height");
    if (const Value* f = opt_field(cam, "fov")) {
        S.camera.has_fov = 1;
        S.camera.fov = as_f64(*f, "fov");
    }
    const Value* gp = opt_field(cam, "grid_pitch");
    const Value* gc = opt_field(cam, "grid_color");
    const Value* gt = opt_field(cam, "grid_thickness");
    if (gc && !hex_color(as_str(*gc, "grid_color"), S.camera.grid_color))
        throw Error(RT_E_CAMERA, "Invalid hex color format");  // Camera::from_config (src/camera.rs:35-39)
    if (kind == "ortho") S.camera.kind = 0;
    else if (kind == "perspective") S.camera.kind = 1;
    else throw Error(RT_E_CAMERA, "Unsupported camera type: " + kind);
    if (gp && gc && gt) {
        S.camera.grid_pitch = as_f64(*gp, "grid_pitch");
        S.camera.grid_thickness = as_f64(*gt, "grid_thickness");
        S.camera.has_grid = S.camera.grid_pitch > 0.0 && S.camera.grid_thickness > 0.0;
    }

    // objects, in scene order (src/renderer.rs:141-310)
    std::vector<MaterialIn> mats;
    for (size_t idx = 0; idx < objs.items.size(); ++idx) {
        const Value& o = objs.items[idx];
        std::string k = as_str(field(o, "kind"), "kind");
        MaterialIn mat = read_material(field(o, "material"));
        double M[16];
        bool have_m = false;
        if (const Value* tr = opt_field(o, "transform")) {
            if (tr->kind != Value::kArray) throw Error(RT_E_SCENE, "invalid type for `transform`, expected a sequence");
            std::vector<std::string> list;
            for (auto& s : tr->items) list.push_back(as_str(s, "transform"));
            have_m = parse_transform_list(list, M);  // invalid strings are ignored (src/renderer.rs:153-155)
        }
        rt_object_desc d{};
        ObjBounds ob;
        d.material = (int32_t)idx;
        d.mesh = -1;
        ident4(d.world_to_local);
        ident4(d.local_to_world);
        if (k == "sphere") {
            d.kind = RT_OBJ_SPHERE;
            as_vec3(field(o, "center"), "center", d.center);
            d.radius = as_f64(field(o, "radius"), "radius");
            if (have_m) {
                d3 c = m4_apply(M, mk(d.center[0], d.center[1], d.center[2]), 1.0);
                d.center[0] = c.x; d.center[1] = c.y; d.center[2] = c.z;
                double s0 = length(mk(M[0], M[1], M[2])), s1 = length(mk(M[4], M[5], M[6])), s2 = length(mk(M[8], M[9], M[10]));
                d.radius *= fmax_r(fmax_r(s0, s1), s2);
            }
            for (int i = 0; i < 3; ++i) {  // center -/+ (r, r, r)
                ob.lo[i] = d.center[i] - d.radius;
                ob.hi[i] = d.center[i] + d.radius;
            }
            ob.finite = true;
        } else if (k == "plane") {
            d.kind = RT_OBJ_PLANE;
            as_vec3(field(o, "point"), "point", d.point);
            double n[3];
            as_vec3(field(o, "normal"), "normal", n);
            d3 nn = mk(n[0], n[1], n[2]);
            if (have_m) {
                d3 p = m4_apply(M, mk(d.point[0], d.point[1], d.point[2]), 1.0);
                d.point[0] = p.x; d.point[1] = p.y; d.point[2] = p.z;
                double inv[16], it[16];
                if (invert4(M, inv)) {
                    for (int r = 0; r < 4; ++r)
                        for (int c = 0; c < 4; ++c) it[c * 4 + r] = inv[r * 4 + c];
                    nn = m4_apply(it, nn, 0.0);
                }
            }
            d3 u = unit(nn);
            d.normal[0] = u.x; d.normal[1] = u.y; d.normal[2] = u.z;
        } else if (k == "cube") {
            d.kind = RT_OBJ_CUBE;
            as_vec3(field(o, "center"), "center", d.center);
            double size[3];
            as_vec3(field(o, "size"), "size", size);
            for (int i = 0; i < 3; ++i) d.half_size[i] = size[i] / 2.0;
            if (have_m) {  // Cube::new_with_transform (src/ray.rs:208-225)
                if (!invert4(M, d.world_to_local)) ident4(d.world_to_local);
                memcpy(d.local_to_world, M, sizeof(M));
            }
            cube_bounds(d, ob);
        } else if (k == "mesh") {
            d.kind = RT_OBJ_MESH;
            std::string fn = as_str(field(o, "filename"), "filename");
            std::string path = (!fn.empty() && fn[0] == '/') ? fn : base_dir + "/" + fn;
            std::vector<double> v = load_stl_file(path);
            mesh_bounds(v, have_m ? M : nullptr, ob);
            if (have_m) {
                for (size_t i = 0; i < v.size(); i += 3) {
                    d3 p = m4_apply(M, mk(v[i], v[i + 1], v[i + 2]), 1.0);
                    v[i] = p.x; v[i + 1] = p.y; v[i + 2] = p.z;
                }
            }
            d.mesh = (int32_t)S.meshes.size();
            S.meshes.push_back(std::move(v));
        } else {
            throw Error(RT_E_SCENE, "unknown variant `" + k + "`, expected one of `sphere`, `plane`, `cube`, `mesh`");
        }
        S.objects.push_back(d);
        S.bounds.push_back(ob);
        mats.push_back(std::move(mat));
    }
    for (auto& m : mats) S.materials.push_back(lower_material(m, true));
    for (size_t i = 0; i < mats.size(); ++i) {
        if (mats[i].texture == 2) {
            S.materials[i].material_b = (int32_t)S.materials.size();
            S.materials.push_back(lower_material(mats[i].alt[0], false));
        }
    }

    // lights (src/scene.rs:348-353); colours fall back to white at shade time
    for (const Value& l : lights.items) {
        rt_light_desc d{};
        as_vec3(field(l, "position"), "position", d.position);
        if (!hex_color(as_str(field(l, "color"), "color"), d.color)) d.color[0] = d.color[1] = d.color[2] = 1.0;
        d.intensity = as_f64(field(l, "intensity"), "intensity");
        if (const Value* dm = opt_field(l, "diameter")) {
            d.is_area = 1;
            d.diameter = as_f64(*dm, "diameter");
        }
        // extension keys (ignored by the reference's serde model)
        d.shape = RT_LIGHT_DISK;
        d.samples = 16;
        if (const Value* sh = opt_field(l, "shape")) {
            std::string k = as_str(*sh, "shape");
            if (k == "square") d.shape = RT_LIGHT_SQUARE;
            else if (k != "disk") throw Error(RT_E_SCENE, "unknown variant `" + k + "`, expected `disk` or `square`");
        }
        if (d.shape == RT_LIGHT_SQUARE) {
            const Value* sz = opt_field(l, "size");
            if (!sz) sz = opt_field(l, "diameter");
            if (!sz) throw Error(RT_E_SCENE, "missing field `size`");
            d.size = as_f64(*sz, "size");
            d.is_area = 1;
        }
        if (const Value* ns = opt_field(l, "samples")) {
            double v = as_f64(*ns, "samples");
            if (!(v >= 1.0 && v <= 65536.0) || v != std::floor(v))
                throw Error(RT_E_SCENE, "invalid value for `samples`, expected an integer in 1..=65536");
            d.samples = (int32_t)v;
        }
        S.lights.push_back(d);
    }

    // scene settings (src/scene.rs:421-437)
    const Value& amb = field(ss, "ambient_illumination");
    if (!hex_color(as_str(field(amb, "color"), "color"), S.ambient_color))
        S.ambient_color[0] = S.ambient_color[1] = S.ambient_color[2] = 1.0;
    S.ambient_intensity = as_f64(field(amb, "intensity"), "intensity");
    if (const Value* f = opt_field(ss, "fog")) {
        S.has_fog = true;
        if (!hex_color(as_str(field(*f, "color"), "color"), S.fog_color)) S.fog_color[0] = S.fog_color[1] = S.fog_color[2] = 0.5;
        S.fog_density = as_f64(field(*f, "density"), "density");
        S.fog_start = as_f64(field(*f, "start"), "start");
        S.fog_end = as_f64(field(*f, "end"), "end");
    }
    if (const Value* bg = opt_field(ss, "background_color")) {
        if (!hex_color(as_str(*bg, "background_color"), S.background)) throw Error(RT_E_CAMERA, "Invalid hex color format");
    }
    if (const Value* ol = opt_field(ss, "outline")) {
        const Value& en = field(*ol, "enabled");
        if (en.kind != Value::kBool) throw Error(RT_E_SCENE, "invalid type for `enabled`, expected a boolean");
        auto num_or = [&](const char* k, double dflt) {
            const Value* v = ol->find(k);
            return v ? as_f64(*v, k) : dflt;
        };
        rt_render_opts& o = S.outline;
        o.outline_depth_weight = num_or("depth_weight", 1.0);
        o.outline_normal_weight = num_or("normal_weight", 1.0);
        o.outline_threshold = num_or("threshold", 0.1);
        o.outline_thickness = num_or("thickness", 1.0);
        const Value* u8n = ol->find("use_8_neighbors");
        o.outline_use_8_neighbors = u8n ? (u8n->kind == Value::kBool && u8n->boolean) : 0;
        const Value* c = ol->find("color");
        std::string col = c ? as_str(*c, "color") : std::string("#000000");
        S.outline_enabled = en.boolean;
        if (S.outline_enabled) {
            o.outline_enabled = 1;
            S.outline_color_ok = hex_color(col, o.outline_color);
        }
    }
    return S;
}

HostScene load_scene_file(const std::string& path) {
    return load_scene_json(read_file(path), parent_dir(path));
}

rt_scene_desc HostScene::desc(std::vector<rt_mesh_desc>& ms) const {
    ms.clear();
    for (const auto& v : meshes) ms.push_back(rt_mesh_desc{(int32_t)(v.size() / 9), v.data()});
    rt_scene_desc d{};
    d.camera = camera;
    d.n_objects = (int32_t)objects.size();
    d.objects = objects.data();
    d.n_materials = (int32_t)materials.size();
    d.materials = materials.data();
    d.n_lights = (int32_t)lights.size();
    d.lights = lights.data();
    d.n_meshes = (int32_t)ms.size();
    d.meshes = ms.data();
    memcpy(d.ambient_color, ambient_color, sizeof(ambient_color));
    d.ambient_intensity = ambient_intensity;
    d.has_fog = has_fog;
    memcpy(d.fog_color, fog_color, sizeof(fog_color));
    d.fog_density = fog_density;
    d.fog_start = fog_start;
    d.fog_end = fog_end;
    memcpy(d.background, background, sizeof(background));
    return d;
}

HostScene from_desc(const rt_scene_desc& d) {
    HostScene S;
    S.camera = d.camera;
    if (d.n_objects < 0 || d.n_materials < 0 || d.n_lights < 0 || d.n_meshes < 0)
        throw Error(RT_E_INVALID_ARG, "negative count in rt_scene_desc");
    S.objects.assign(d.objects, d.objects + d.n_objects);
    S.materials.assign(d.materials, d.materials + d.n_materials);
    S.lights.assign(d.lights, d.lights + d.n_lights);
    for (auto& l : S.lights) {
        if (l.samples == 0) l.samples = 16;  // zero-initialised descriptors get the reference count
        if (l.samples < 0 || l.samples > 65536) throw Error(RT_E_INVALID_ARG, "light samples out of range");
        if (l.shape != RT_LIGHT_DISK && l.shape != RT_LIGHT_SQUARE) throw Error(RT_E_INVALID_ARG, "unknown light shape");
        if (l.shape == RT_LIGHT_SQUARE) l.is_area = 1;
    }
    for (int i = 0; i < d.n_meshes; ++i) {
        if (d.meshes[i].n_tris < 0 || (d.meshes[i].n_tris > 0 && !d.meshes[i].vertices))
            throw Error(RT_E_INVALID_ARG, "mesh without vertices");
        S.meshes.emplace_back(d.meshes[i].vertices, d.meshes[i].vertices + (size_t)d.meshes[i].n_tris * 9);
    }
    for (const auto& o : S.objects) {
        ObjBounds ob;
        if (o.kind == RT_OBJ_SPHERE) {
            for (int i = 0; i < 3; ++i) {
                ob.lo[i] = o.center[i] - o.radius;
                ob.hi[i] = o.center[i] + o.radius;
            }
            ob.finite = true;
        } else if (o.kind == RT_OBJ_CUBE) {
            cube_bounds(o, ob);
        } else if (o.kind == RT_OBJ_MESH && o.mesh >= 0 && o.mesh < d.n_meshes) {
            mesh_bounds(S.meshes[o.mesh], nullptr, ob);  // vertices arrive transformed
        }
        S.bounds.push_back(ob);
    }
    for (auto& o : S.objects) {
        if (o.material < 0 || o.material >= d.n_materials) throw Error(RT_E_INVALID_ARG, "object material out of range");
        if (o.kind == RT_OBJ_MESH && (o.mesh < 0 || o.mesh >= d.n_meshes)) throw Error(RT_E_INVALID_ARG, "object mesh out of range");
        if (o.kind < RT_OBJ_SPHERE || o.kind > RT_OBJ_MESH) throw Error(RT_E_INVALID_ARG, "unknown object kind");
    }
    for (auto& m : S.materials)
        if (m.texture == 2 && (m.material_b < 0 || m.material_b >= d.n_materials))
            throw Error(RT_E_INVALID_ARG, "checkerboard material_b out of range");
    memcpy(S.ambient_color, d.ambient_color, sizeof(S.ambient_color));
    S.ambient_intensity = d.ambient_intensity;
    S.has_fog = d.has_fog != 0;
    memcpy(S.fog_color, d.fog_color, sizeof(S.fog_color));
    S.fog_density = d.fog_density;
    S.fog_start = d.fog_start;
    S.fog_end = d.fog_end;
    memcpy(S.background, d.background, sizeof(S.background));
    return S;
}

void auto_cameras(const HostScene& S, rt_camera_desc out[4]) {
    bool any = false;
    double lo[3], hi[3];
    for (const ObjBounds& b : S.bounds) {
        if (!b.finite) continue;
        for (int i = 0; i < 3; ++i) {
            lo[i] = any ? fmin_r(lo[i], b.lo[i]) : b.lo[i];
            hi[i] = any ? fmax_r(hi[i], b.hi[i]) : b.hi[i];
        }
        any = true;
    }
    if (!any) throw Error(RT_E_SCENE, "Scene has no finite objects to compute bounds");
    const double c[3] = {(lo[0] + hi[0]) / 2.0, (lo[1] + hi[1]) / 2.0, (lo[2] + hi[2]) / 2.0};
    const double size[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    const double margin = 1.15;
    const double vw = fmax_r(fmax_r(size[0], size[1]), size[2]) * margin;  // square viewports
    const double dist = vw * 2.0;
    auto ortho = [&](rt_camera_desc& cam, double px, double py, double pz, double ux, double uy, double uz) {
        cam = rt_camera_desc{};
        cam.kind = 0;
        cam.position[0] = px;
        cam.position[1] = py;
        cam.position[2] = pz;
        for (int i = 0; i < 3; ++i) cam.target[i] = c[i];
        cam.up[0] = ux;
        cam.up[1] = uy;
        cam.up[2] = uz;
        cam.width = vw;
        cam.height = vw;
    };
    ortho(out[0], c[0], c[1] - dist, c[2], 0, 0, 1);  // left: from -y
    ortho(out[1], c[0] - dist, c[1], c[2], 0, 0, 1);  // front: from -x
    ortho(out[2], c[0], c[1], c[2] + dist, 0, 1, 0);  // top: from +z, y forward
    // perspective (src/auto_camera.rs:97-130): 50 degree fov, 35 degrees down, 45 degrees in x-y
    const double fov = 50.0, down = 35.0;
    const double max_dim = fmax_r(fmax_r(size[0], size[1]), size[2]) * margin;
    const double fov_rad = fov * (3.141592653589793 / 180.0);
    const double d_fov = max_dim / std::tan(fov_rad / 2.0);
    const double xy = 45.0 * (3.141592653589793 / 180.0), dr = down * (3.141592653589793 / 180.0);
    const double horiz = d_fov * std::cos(dr);
    rt_camera_desc& p = out[3];
    p = rt_camera_desc{};
    p.kind = 1;
    p.position[0] = c[0] + horiz * std::cos(xy);
    p.position[1] = c[1] + horiz * std::sin(xy);
    p.position[2] = c[2] + d_fov * std::sin(dr);
    for (int i = 0; i < 3; ++i) p.target[i] = c[i];
    p.up[2] = 1.0;
    p.width = 1.0;
    p.height = 1.0;
    p.has_fov = 1;
    p.fov = fov;
}

}  // namespace rt
