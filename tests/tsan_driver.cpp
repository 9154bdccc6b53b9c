// ThreadSanitizer driver for the host code's thread pools: the scene library's generators,
// extrusion and BVH build (specrel_amd/csrc/scene/scene.cpp) and the oracle's row-parallel renderer
// (oracle/oracle.c), linked in directly and run with several threads on small scenes. Built and run
// by tests/run_sanitized.sh; host code only.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "specrel_scene.h"

extern "C" {
#include "../oracle/oracle.h"
}

static int fail(const char *what) {
    std::fprintf(stderr, "tsan driver: %s failed: %s\n", what, specrel_scene_last_error());
    return 1;
}

int main() {
    const char *names[] = {"twin", "barn", "carousel", "large", "accel_clocks"};
    for (const char *name : names) {
        specrel_gen_params p;
        if (specrel_gen_defaults(name, &p)) return fail("gen_defaults");
        if (!std::strcmp(name, "large")) {  // a reduced large scene: the same code, seconds not minutes
            p.n_objects = 6;
            p.slices = 12;
            p.n_stars = 200;
        }
        specrel_host_scene *sc = nullptr;
        if (specrel_scene_generate(name, &p, &sc)) return fail("generate");
        specrel_scene_arrays a;
        if (specrel_scene_get_arrays(sc, &a)) return fail("get_arrays");
        specrel_camera cam;
        if (specrel_scene_default_camera(name, &p, &cam)) return fail("default_camera");
        specrel_render_config cfg;
        std::memset(&cfg, 0, sizeof cfg);
        cfg.width = 96;
        cfg.height = 60;
        specrel_frame f;
        if (specrel_frame_build(&cam, &cfg, nullptr, &f)) return fail("frame_build");
        std::vector<uint32_t> rgb(96 * 60), hit(96 * 60);
        oracle_counts c;
        if (oracle_render(&a, &f, 4, rgb.data(), hit.data(), nullptr, nullptr, nullptr, &c)) return fail("render");
        std::printf("%s: %u nodes, %u prims, %llu rays, %llu hits\n", name, a.num_nodes, a.num_prims,
                    (unsigned long long)c.rays, (unsigned long long)c.hits);
        specrel_scene_free(sc);
    }
    return 0;
}
