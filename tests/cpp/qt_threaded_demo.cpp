// The reference's thread layout without its two races
// (Include/Simulator.h:67 `volatile bool m_bStop` shared across threads;
// Source/Simulator.cpp:50 -> Source/FluidRenderWidget.cpp:211, the GUI thread
// reading the solver's position array while the worker writes it):
//   - a worker thread owns the solver and runs frames (Source/Simulator.cpp:44-51)
//     until a std::atomic<bool> stop flag is set;
//   - after each frame it takes a position snapshot (snapshotBegin), enqueues
//     the next frame, collects the snapshot (snapshotWait) and publishes it
//     under a mutex;
//   - the GUI thread (main) reads published frames until it has seen FRAMES of
//     them, checks each, then sets the stop flag and joins the worker.
// Then the same solver type replays, on this thread, exactly as many frames as
// the worker ran, and the two final states must agree bit for bit.
//
//   qt_threaded_demo SCENE RES FRAMES
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "sph_solver.hpp"

namespace {

using Solver = sph::QtSPHSolver<>;

void setup(Solver& solver, int scene, int res) {
    uint32_t n = 0;
    if (sph_scene_generate(scene, res, nullptr, 0, &n) != SPH_OK) throw sph::Error(SPH_E_INVALID, "scene count");
    auto& x = solver.getParticles();
    x.resize(n);
    if (sph_scene_generate(scene, res, &x[0].x, n, &n) != SPH_OK) throw sph::Error(SPH_E_INVALID, "scene");
    solver.getVelocity().assign(n, sph::Vec3f{0.f, 0.f, 0.f});
    solver.makeReady();
}

struct Published {
    std::mutex m;
    std::condition_variable cv;
    std::vector<sph::Vec3f> positions;
    uint32_t frame = 0;        // frames whose snapshot has been published
    bool failed = false;
    std::string error;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s SCENE RES FRAMES\n", argv[0]);
        return 2;
    }
    const int scene = std::atoi(argv[1]), res = std::atoi(argv[2]);
    const uint32_t want = (uint32_t)std::atoi(argv[3]);
    auto params = std::make_shared<sph::SPHParameters>();
    params->setResolution(res);

    std::atomic<bool> stop{false};
    Published pub;
    std::vector<sph::Vec3f> worker_final;
    uint32_t worker_frames = 0;

    std::thread worker([&] {
        try {
            Solver solver(params);
            setup(solver, scene, res);
            std::vector<sph::Vec3f> snap;
            uint32_t frames = 0;
            bool pending = false;
            while (!stop.load(std::memory_order_acquire)) {
                solver.advanceFrameBatch();
                ++frames;
                if (pending) {
                    // The previous frame's snapshot copied while this frame ran.
                    solver.snapshotWait(snap);
                    std::lock_guard<std::mutex> lk(pub.m);
                    pub.positions = snap;
                    pub.frame = frames - 1;
                    pub.cv.notify_all();
                }
                solver.snapshotBegin();
                pending = true;
            }
            if (pending) solver.snapshotWait(snap);
            worker_final = solver.getParticles();
            worker_frames = frames;
            for (size_t i = 0; i < snap.size(); ++i)
                if (snap[i].x != worker_final[i].x || snap[i].y != worker_final[i].y || snap[i].z != worker_final[i].z)
                    throw sph::Error(SPH_E_INVALID, "last snapshot differs from the final state");
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(pub.m);
            pub.failed = true;
            pub.error = e.what();
            pub.cv.notify_all();
        }
    });

    // GUI thread: read frames as they are published, as the renderer would.
    const float lo = -1.0f + params->particleRadius, hi = 1.0f - params->particleRadius;
    uint32_t seen = 0, last = 0;
    bool bad = false;
    while (seen < want) {
        std::vector<sph::Vec3f> frame;
        {
            std::unique_lock<std::mutex> lk(pub.m);
            if (!pub.cv.wait_for(lk, std::chrono::seconds(60), [&] { return pub.failed || pub.frame > last; })) {
                std::fprintf(stderr, "no frame published within 60 s\n");
                bad = true;
                break;
            }
            if (pub.failed) break;
            frame = pub.positions;
            last = pub.frame;
        }
        for (const auto& p : frame)
            if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z) || p.x < lo || p.x > hi ||
                p.y < lo || p.y > hi || p.z < lo || p.z > hi)
                bad = true;
        ++seen;
    }
    stop.store(true, std::memory_order_release);
    worker.join();
    if (pub.failed) {
        std::fprintf(stderr, "worker failed: %s\n", pub.error.c_str());
        return 1;
    }
    if (bad) {
        std::fprintf(stderr, "a published frame was not finite or left the box\n");
        return 6;
    }

    // Same frame count on one thread: identical state.
    try {
        Solver solver(params);
        setup(solver, scene, res);
        for (uint32_t f = 0; f < worker_frames; ++f) solver.advanceFrameBatch();
        const auto& x = solver.getParticles();
        if (x.size() != worker_final.size()) return 7;
        for (size_t i = 0; i < x.size(); ++i)
            if (x[i].x != worker_final[i].x || x[i].y != worker_final[i].y || x[i].z != worker_final[i].z) {
                std::fprintf(stderr, "threaded run differs from the single-threaded replay at particle %zu\n", i);
                return 8;
            }
    } catch (const sph::Error& e) {
        std::fprintf(stderr, "engine error %d: %s\n", e.code(), e.what());
        return 1;
    }
    std::printf("frames read by the GUI thread %u, worker frames %u, particles %zu\n", seen, worker_frames,
                worker_final.size());
    return 0;
}
