// shadow_task (kernels/trace_core.hpp) maps the shadow stage's queue items onto the tasks
// hit * slots + slot: checks that it is a permutation of [0, hits * slots) for mixes of
// point, disk and square lights (with and without the per-render sample override) and hit
// counts around the 64-hit blocks, and that within a full block the 64 consecutive items
// of a point light belong to 64 different hits and that light's slot. Prints "ok <cases>".
#include <cstdio>
#include <vector>

#include "../../rtrace_amd/csrc/kernels/trace_core.hpp"

using namespace rt;

int main() {
    struct Mix {
        std::vector<int> area_samples;  // 0: a point light, else an area light with that many samples
        int override_samples;
    };
    const std::vector<Mix> mixes = {{{16}, 0}, {{0, 0, 0, 0}, 0}, {{0, 8, 0, 5}, 0}, {{0, 16}, 4}, {{7}, 0}, {{0}, 0}};
    const int hit_counts[] = {0, 1, 63, 64, 65, 128, 1000, 4097};
    int cases = 0;
    for (const Mix& m : mixes) {
        std::vector<LightDev> lights(m.area_samples.size());
        int slots = 0;
        for (size_t i = 0; i < lights.size(); ++i) {
            lights[i] = LightDev{};
            lights[i].is_area = m.area_samples[i] > 0;
            lights[i].samples = m.area_samples[i];
        }
        SceneView S{};
        S.lights = lights.data();
        S.n_lights = (int)lights.size();
        S.light_samples = m.override_samples;
        for (const LightDev& L : lights) slots += light_count(S, L);
        S.shadow_per_hit = slots;
        for (int nh : hit_counts) {
            const int n = nh * slots;
            std::vector<int> seen(n, 0);
            for (int i = 0; i < n; ++i) {
                const int t = shadow_task(S, nh, i);
                if (t < 0 || t >= n || seen[t]++) {
                    std::printf("FAIL not a permutation: hits %d slots %d item %d task %d\n", nh, slots, i, t);
                    return 1;
                }
            }
            // a point light's 64 items of a full block: 64 hits, one slot
            int off = 0;
            for (int li = 0; li < S.n_lights && nh >= 64; ++li) {
                const int cnt = light_count(S, lights[li]);
                if (cnt == 1) {
                    for (int h = 0; h < 64; ++h) {
                        const int t = shadow_task(S, nh, 64 * off + h);
                        if (t != h * slots + off) {
                            std::printf("FAIL point light %d: item %d task %d\n", li, 64 * off + h, t);
                            return 1;
                        }
                    }
                }
                off += cnt;
            }
            ++cases;
        }
    }
    std::printf("ok %d\n", cases);
    return 0;
}
