// rtrace CLI: the reference's flags (cli/src/main.rs:8-32) plus --backend, the same
// validation order, sizing (:79-89), outline/anti-aliasing rules (:103-124) and messages;
// renders through librtrace_hip and writes a PNG.
//
// --backend defaults to the reference's own renderer, "cpu", which is the Rust crate and
// is not part of this build: that choice stops with a message naming --backend hip.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/rtrace_hip.h"
#include "png.hpp"

namespace {

void usage(FILE* f) {
    fprintf(f,
            "Usage: rtrace -i <INPUT> -o <OUTPUT> [-s <SIZE>] [--max-depth <N>] [--samples <N>]\n"
            "              [--anti-aliasing quincunx|stochastic|no-jitter] [--backend cpu|hip]\n"
            "  --backend  cpu (default) is the reference's Rust renderer, which this build does not\n"
            "             contain; hip renders on an MI355X (gfx950) through librtrace_hip\n");
}

// clap's message for a value that does not parse (exit status 2)
[[noreturn]] void bad_value(const std::string& v, const char* arg, const char* why) {
    fprintf(stderr, "error: invalid value '%s' for '%s': %s\n\nFor more information, try '--help'.\n", v.c_str(), arg,
            why);
    exit(2);
}

// Rust's u32 / i32 FromStr: optional sign ('-' only for signed), decimal digits, no spaces
long long parse_int(const std::string& v, const char* arg, bool is_signed) {
    if (v.empty()) bad_value(v, arg, "cannot parse integer from empty string");
    size_t i = 0;
    bool neg = false;
    if (v[0] == '+' || (is_signed && v[0] == '-')) {
        neg = v[0] == '-';
        i = 1;
        if (v.size() == 1) bad_value(v, arg, "invalid digit found in string");
    }
    long long x = 0;
    const long long hi = is_signed ? (neg ? 2147483648ll : 2147483647ll) : 4294967295ll;
    for (; i < v.size(); ++i) {
        if (v[i] < '0' || v[i] > '9') bad_value(v, arg, "invalid digit found in string");
        x = x * 10 + (v[i] - '0');
        if (x > hi) bad_value(v, arg, neg ? "number too small to fit in target type" : "number too large to fit in target type");
    }
    return neg ? -x : x;
}

// format_duration (src/renderer.rs:1348-1378)
std::string format_duration(double seconds) {
    if (seconds < 0.0) return "0s";
    const unsigned long long total = (unsigned long long)std::llround(seconds);
    if (total == 0) return "0s";
    const unsigned long long h = total / 3600, m = (total % 3600) / 60, s = total % 60;
    char buf[64];
    if (h > 0) {
        if (m > 0) snprintf(buf, sizeof buf, "%lluh%llum", h, m);
        else snprintf(buf, sizeof buf, "%lluh", h);
    } else if (m > 0) {
        if (s > 0) snprintf(buf, sizeof buf, "%llum%llus", m, s);
        else snprintf(buf, sizeof buf, "%llum", m);
    } else {
        snprintf(buf, sizeof buf, "%llus", s);
    }
    return buf;
}

}  // namespace

int main(int argc, char** argv) {
    std::string input, output, aa = "quincunx", backend = "cpu";
    bool have_input = false, have_output = false;
    long long size = 1000, max_depth = 10, samples = -1;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&](const char* name) -> std::string {
            if (i + 1 >= argc) {
                fprintf(stderr, "error: a value is required for '%s' but none was supplied\n\nFor more information, try '--help'.\n", name);
                exit(2);
            }
            return argv[++i];
        };
        if (a == "-i" || a == "--input") input = val("--input <INPUT>"), have_input = true;
        else if (a == "-o" || a == "--output") output = val("--output <OUTPUT>"), have_output = true;
        else if (a == "-s" || a == "--size") size = parse_int(val("--size <SIZE>"), "--size <SIZE>", false);
        else if (a == "--max-depth") max_depth = parse_int(val("--max-depth <MAX_DEPTH>"), "--max-depth <MAX_DEPTH>", true);
        else if (a == "--samples") samples = parse_int(val("--samples <SAMPLES>"), "--samples <SAMPLES>", false);
        else if (a == "--anti-aliasing") aa = val("--anti-aliasing <ANTI_ALIASING>");
        else if (a == "--backend") backend = val("--backend <BACKEND>");
        else if (a == "-h" || a == "--help") { usage(stdout); return 0; }
        else {
            fprintf(stderr, "error: unexpected argument '%s' found\n\n", a.c_str());
            usage(stderr);
            return 2;
        }
    }
    if (!have_input || !have_output) {
        fprintf(stderr, "error: the following required arguments were not provided:\n%s%s\n",
                have_input ? "" : "  --input <INPUT>\n", have_output ? "" : "  --output <OUTPUT>\n");
        usage(stderr);
        return 2;
    }
    FILE* probe = fopen(input.c_str(), "rb");
    if (!probe) {
        fprintf(stderr, "Error: Input file '%s' does not exist\n", input.c_str());
        return 1;
    }
    fclose(probe);
    int mode;
    if (aa == "quincunx") mode = RT_AA_QUINCUNX;
    else if (aa == "stochastic") mode = RT_AA_STOCHASTIC;
    else if (aa == "no-jitter") mode = RT_AA_NO_JITTER;
    else {
        fprintf(stderr, "Error: Invalid anti-aliasing mode '%s'. Valid options are: quincunx, stochastic, no-jitter\n", aa.c_str());
        return 1;
    }
    const long long nsamp = samples < 0 ? 1 : samples;
    if (nsamp == 0) {
        fprintf(stderr, "Error: Samples must be greater than 0\n");
        return 1;
    }
    if (backend == "cpu") {
        fprintf(stderr,
                "Error: the cpu backend is the reference's Rust renderer, which this build does not contain; "
                "use --backend hip\n");
        return 1;
    }
    if (backend != "hip") {
        fprintf(stderr, "Error: Invalid backend '%s'. Valid options are: cpu, hip\n", backend.c_str());
        return 1;
    }

    rt_ctx* ctx = nullptr;
    rt_scene* scene = nullptr;
    auto finish = [&](int code) {
        if (scene) rt_scene_free(scene);
        if (ctx) rt_ctx_destroy(ctx);
        return code;
    };
    if (rt_ctx_create(0, &ctx) != RT_OK) {
        fprintf(stderr, "Error: %s\n", rt_last_error(nullptr));
        return finish(1);
    }
    if (rt_scene_load_json_file(ctx, input.c_str(), &scene) != RT_OK) {
        fprintf(stderr, "Error loading scene from '%s': %s\n", input.c_str(), rt_last_error(ctx));
        return finish(1);
    }
    int32_t n_obj = 0, n_lights = 0;
    rt_scene_counts(scene, &n_obj, &n_lights);
    printf("Loaded scene with %d objects and %d lights\n", n_obj, n_lights);
    double cw = 0, ch = 0;
    int32_t outline_on = 0;
    rt_render_opts opts;
    rt_render_opts_default(&opts, 1, 1);
    const int info = rt_scene_info(scene, &cw, &ch, &outline_on, &opts);
    const double r = cw / ch;
    const double hf = (double)size / std::sqrt(r * r + 1.0);
    const double wf = r * hf;
    // f64::round (half away from zero), then the saturating `as u32`
    auto to_u32 = [](double x) -> long long {
        x = std::round(x);
        if (!(x > 0.0)) return 0;
        return x >= 4294967295.0 ? 4294967295ll : (long long)x;
    };
    const long long W = to_u32(wf), H = to_u32(hf);
    printf("Using camera aspect ratio %.3f to compute %lld×%lld pixels from diagonal %lld\n", r, W, H, size);
    if (W <= 0 || H <= 0 || W > INT32_MAX || H > INT32_MAX) {
        fprintf(stderr, "Error rendering image: invalid image size %lld×%lld\n", W, H);
        return finish(1);
    }
    opts.width = (int32_t)W;
    opts.height = (int32_t)H;
    opts.max_depth = (int32_t)max_depth;
    opts.samples = (int32_t)std::min<long long>(nsamp, INT32_MAX);
    opts.has_seed = 1;
    opts.seed = 0;
    if (info != RT_OK) {
        fprintf(stderr, "Error: Invalid outline color in scene: %s\n", rt_last_error(ctx));
        return finish(1);
    }
    if (outline_on) {
        printf("Outline detection enabled from scene configuration\n");
        if (mode == RT_AA_QUINCUNX) {
            printf("Warning: Quincunx anti-aliasing is not compatible with outline detection. Switching to no-jitter mode.\n");
            mode = RT_AA_NO_JITTER;
        }
    }
    opts.aa_mode = mode;
    const char* names[3] = {"no-jitter", "quincunx", "stochastic"};
    printf("Rendering %lld×%lld image (diagonal %lld) with %s anti-aliasing (%lld samples)...\n", W, H, size,
           names[mode], nsamp);
    fflush(stdout);
    std::vector<uint8_t> rgb((size_t)W * H * 3);
    rt_render_stats st;
    if (rt_render(ctx, scene, &opts, rgb.data(), &st) != RT_OK) {
        fprintf(stderr, "Error rendering image: %s\n", rt_last_error(ctx));
        return finish(1);
    }
    printf("Total rendering time: %s\n", format_duration(st.ms_render / 1000.0).c_str());
    if (!rtpng::write_png(output.c_str(), rgb.data(), (int)W, (int)H)) {
        fprintf(stderr, "Error rendering image: cannot write %s\n", output.c_str());
        return finish(1);
    }
    printf("Image saved to: %s\n", output.c_str());
    printf("Successfully rendered to '%s'\n", output.c_str());
    return finish(0);
}
