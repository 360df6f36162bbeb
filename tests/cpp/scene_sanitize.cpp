// TEST ONLY: the host-side scene generators (simplefluid_amd/csrc/sph_scene.cpp)
// built with AddressSanitizer and UndefinedBehaviorSanitizer: every reference
// scene at a few resolutions (count query, fill, a too-small buffer), and the
// ball sampler on a dense case (checked for spacing and containment by brute
// force), a too-full ball and invalid arguments.
#include <cmath>
#include <cstdio>
#include <vector>

#include "sph_engine.h"

static int fail(int code, const char* what) {
    std::printf("FAIL %s\n", what);
    return code;
}

int main() {
    for (int scene = 0; scene < 4; ++scene)
        for (int res : {8, 16, 24}) {
            uint32_t n = 0;
            if (sph_scene_generate(scene, res, nullptr, 0, &n) != SPH_OK || n == 0) return fail(1, "scene count");
            std::vector<float> x(3 * (size_t)n);
            uint32_t m = 0;
            if (sph_scene_generate(scene, res, x.data(), n, &m) != SPH_OK || m != n) return fail(2, "scene fill");
            if (sph_scene_generate(scene, res, x.data(), n - 1, &m) != SPH_E_CAPACITY) return fail(3, "scene capacity");
            for (float v : x)
                if (!(std::fabs(v) <= 1.0f)) return fail(4, "scene inside [-1, 1]");
        }
    if (sph_scene_generate(7, 24, nullptr, 0, nullptr) != SPH_E_INVALID) return fail(5, "scene args");

    const float c[3] = {0.1f, 0.2f, 0.3f};
    const float R = 0.05f, s = 0.004f;
    const uint32_t want = 5000;
    std::vector<float> b(3 * 20000);
    uint32_t n = 0;
    if (sph_scene_sample_ball(5, c, R, s, want, b.data(), &n) != SPH_OK || n != want) return fail(6, "ball sample");
    for (uint32_t i = 0; i < n; ++i) {
        double r2 = 0;
        for (int k = 0; k < 3; ++k) r2 += ((double)b[3 * i + k] - c[k]) * ((double)b[3 * i + k] - c[k]);
        if (!(r2 < (double)R * R * (1 + 1e-5))) return fail(7, "ball containment");
        for (uint32_t j = 0; j < i; ++j) {
            double d2 = 0;
            for (int k = 0; k < 3; ++k) d2 += ((double)b[3 * i + k] - b[3 * j + k]) * ((double)b[3 * i + k] - b[3 * j + k]);
            if (!(d2 >= (double)s * s)) return fail(8, "ball spacing");
        }
    }
    if (sph_scene_sample_ball(5, c, 0.02f, 0.01f, 20000, b.data(), &n) != SPH_E_CAPACITY || n == 0 || n >= 20000)
        return fail(9, "too-full ball");
    if (sph_scene_sample_ball(5, nullptr, R, s, 10, b.data(), &n) != SPH_E_INVALID) return fail(10, "null centre");
    if (sph_scene_sample_ball(5, c, -1.0f, s, 10, b.data(), &n) != SPH_E_INVALID) return fail(11, "radius");
    if (sph_scene_sample_ball(5, c, R, NAN, 10, b.data(), &n) != SPH_E_INVALID) return fail(12, "spacing");
    std::printf("ok\n");
    return 0;
}
