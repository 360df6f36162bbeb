// Drives the engine through include/sph_solver.hpp the way the reference's
// Simulator drives QtSPHSolver (Source/Simulator.cpp:33-63, 85-103):
// parameters from the GUI defaults at a resolution, a reference scene written
// into getParticles(), makeReady(), then frames of advanceFrame() summed to
// >= 1/30 s.  Writes n, the frame times and the final positions and
// velocities (original order, float32) to the output file.
//
//   qt_adapter_demo SCENE RES FRAMES OUT.bin
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "sph_solver.hpp"

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s SCENE RES FRAMES OUT.bin\n", argv[0]);
        return 2;
    }
    const int scene = std::atoi(argv[1]), res = std::atoi(argv[2]), frames = std::atoi(argv[3]);
    try {
        auto params = std::make_shared<sph::SPHParameters>();
        params->setResolution(res);
        sph::QtSPHSolver<> solver(params);

        // SceneManager::setupScene (Source/SceneManager.cpp:21-38) fills the arrays.
        uint32_t n = 0;
        if (sph_scene_generate(scene, res, nullptr, 0, &n) != SPH_OK) return 3;
        auto& x = solver.getParticles();
        x.resize(n);
        if (sph_scene_generate(scene, res, &x[0].x, n, &n) != SPH_OK) return 3;
        solver.getVelocity().assign(n, sph::Vec3f{0.f, 0.f, 0.f});

        solver.makeReady();
        std::vector<float> frame_times;
        float sim_time = 0.f;
        for (int f = 0; f < frames && sim_time < params->stopTime; ++f) {
            float frame_time = 0.f;
            while (frame_time < 0.0333333333f) frame_time += solver.advanceFrame();
            sim_time += frame_time;
            frame_times.push_back(frame_time);
        }
        // The renderer's hand-off: a snapshot taken now, collected later.
        solver.snapshotBegin();
        std::vector<sph::Vec3f> snap;
        solver.snapshotWait(snap);
        const auto& xp = solver.getParticles();
        const auto& vp = solver.getVelocity();
        for (size_t i = 0; i < snap.size(); ++i)
            if (snap[i].x != xp[i].x || snap[i].y != xp[i].y || snap[i].z != xp[i].z) return 5;
        std::FILE* out = std::fopen(argv[4], "wb");
        if (!out) return 4;
        const uint32_t nf = (uint32_t)frame_times.size(), np = solver.getNumParticles();
        std::fwrite(&np, 4, 1, out);
        std::fwrite(&nf, 4, 1, out);
        std::fwrite(frame_times.data(), 4, nf, out);
        std::fwrite(xp.data(), 12, np, out);
        std::fwrite(vp.data(), 12, np, out);
        std::fclose(out);
        std::printf("particles %u frames %u sim time %.6f\n", np, nf, sim_time);
    } catch (const sph::Error& e) {
        std::fprintf(stderr, "engine error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
