// rtrace CLI on the HIP backend: the reference's flags (cli/src/main.rs:8-32) plus
// --backend hip, the same sizing (:79-89), outline/anti-aliasing rules (:103-124) and
// messages; renders through librtrace_hip and writes a PNG.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/rtrace_hip.h"
#include "png.hpp"

namespace {

// f64::round: half away from zero
double rust_round(double x) { return std::round(x); }

void usage() {
    fprintf(stderr,
            "Usage: rtrace -i <INPUT> -o <OUTPUT> [-s <SIZE>] [--max-depth <N>] [--samples <N>]\n"
            "              [--anti-aliasing quincunx|stochastic|no-jitter] [--backend hip]\n");
}

}  // namespace

int main(int argc, char** argv) {
    std::string input, output, aa = "quincunx", backend = "hip";
    long size = 1000, max_depth = 10, samples = -1;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&](const char* name) -> std::string {
            if (i + 1 >= argc) {
                fprintf(stderr, "error: a value is required for '%s'\n", name);
                exit(2);
            }
            return argv[++i];
        };
        if (a == "-i" || a == "--input") input = val("--input");
        else if (a == "-o" || a == "--output") output = val("--output");
        else if (a == "-s" || a == "--size") size = strtol(val("--size").c_str(), nullptr, 10);
        else if (a == "--max-depth") max_depth = strtol(val("--max-depth").c_str(), nullptr, 10);
        else if (a == "--samples") samples = strtol(val("--samples").c_str(), nullptr, 10);
        else if (a == "--anti-aliasing") aa = val("--anti-aliasing");
        else if (a == "--backend") backend = val("--backend");
        else if (a == "-h" || a == "--help") { usage(); return 0; }
        else { fprintf(stderr, "error: unexpected argument '%s'\n", a.c_str()); usage(); return 2; }
    }
    if (input.empty() || output.empty()) { usage(); return 2; }
    if (backend != "hip") {
        fprintf(stderr, "Error: this build only provides --backend hip\n");
        return 1;
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
    long nsamp = samples < 0 ? 1 : samples;
    if (nsamp == 0) {
        fprintf(stderr, "Error: Samples must be greater than 0\n");
        return 1;
    }
    rt_ctx* ctx = nullptr;
    if (rt_ctx_create(0, &ctx) != RT_OK) {
        fprintf(stderr, "Error: %s\n", rt_last_error(nullptr));
        return 1;
    }
    rt_scene* scene = nullptr;
    if (rt_scene_load_json_file(ctx, input.c_str(), &scene) != RT_OK) {
        fprintf(stderr, "Error loading scene from '%s': %s\n", input.c_str(), rt_last_error(ctx));
        rt_ctx_destroy(ctx);
        return 1;
    }
    double cw = 0, ch = 0;
    int32_t outline_on = 0;
    rt_render_opts opts;
    rt_render_opts_default(&opts, 1, 1);
    int info = rt_scene_info(scene, &cw, &ch, &outline_on, &opts);
    double r = cw / ch;
    double hf = (double)size / std::sqrt(r * r + 1.0);
    double wf = r * hf;
    opts.width = (int32_t)rust_round(wf);
    opts.height = (int32_t)rust_round(hf);
    opts.max_depth = (int32_t)max_depth;
    opts.samples = (int32_t)nsamp;
    opts.has_seed = 1;
    opts.seed = 0;
    printf("Using camera aspect ratio %.3f to compute %d×%d pixels from diagonal %ld\n", r, opts.width, opts.height, size);
    if (info != RT_OK) {
        fprintf(stderr, "Error: Invalid outline color in scene: %s\n", rt_last_error(ctx));
        return 1;
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
    printf("Rendering %d×%d image (diagonal %ld) with %s anti-aliasing (%ld samples)...\n", opts.width, opts.height,
           size, names[mode], nsamp);
    std::vector<uint8_t> rgb((size_t)opts.width * opts.height * 3);
    rt_render_stats st;
    if (rt_render(ctx, scene, &opts, rgb.data(), &st) != RT_OK) {
        fprintf(stderr, "Error rendering image: %s\n", rt_last_error(ctx));
        return 1;
    }
    printf("Total rendering time: %.3f ms (device)\n", st.ms_render);
    if (!rtpng::write_png(output.c_str(), rgb.data(), opts.width, opts.height)) {
        fprintf(stderr, "Error rendering image: cannot write %s\n", output.c_str());
        return 1;
    }
    printf("Image saved to: %s\n", output.c_str());
    printf("Successfully rendered to '%s'\n", output.c_str());
    rt_scene_free(scene);
    rt_ctx_destroy(ctx);
    return 0;
}
