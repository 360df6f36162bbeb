/*
 * sph_solver.hpp - header-only C++ adapter with the method names of the
 * reference's QtSPHSolver, over the C ABI in sph_engine.h.
 *
 * The reference application drives its solver like this:
 *   Include/Simulator.h:42        m_SPHSolver = make_shared<QtSPHSolver>(particleData, m_SimParams)
 *   Source/SceneManager.cpp:21-173 setupScene* fill getParticles() / getVelocity()
 *   Source/Controller.cpp:52-64   kernelRadius = 2/res ... m_SimParams->updateParams()
 *   Source/Simulator.cpp:42       m_SPHSolver->makeReady()
 *   Source/Simulator.cpp:46-51    while (frameTime < 1/30) frameTime += m_SPHSolver->advanceFrame()
 *   Source/Simulator.cpp:95-96,102 getParticles(), getVelocity(), getNumParticles()
 *   Source/FluidRenderWidget.cpp:211 renderer uploads the position array
 * QtSPHSolver (Include/QtSPHSolver.h:27-36) exposes the particle arrays as
 * Vec_Vec3<float>& (contiguous float triples).  This class keeps those names
 * and that layout: the arrays are host vectors the caller fills before
 * makeReady(), and getParticles() / getVelocity() return them refreshed from
 * the device (in the original particle order) whenever steps have run since
 * the last read.
 *
 * Vec3 is any 12-byte type of three floats (glm::vec3 in the reference; the
 * default is sph::Vec3f).  Errors from the engine throw sph::Error; the
 * reference has no error channel (void methods, Q_ASSERT only).
 * One thread at a time, as the reference's single worker thread.
 */
#ifndef SPH_SOLVER_HPP
#define SPH_SOLVER_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sph_engine.h"

namespace sph {

struct Vec3f {
    float x, y, z;
};

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

/* Field names of SPHParameters<float> as the reference sets them
 * (Source/Controller.cpp:52-64, Source/SceneManager.cpp:45, Source/Simulator.cpp:39,44),
 * plus the engine's own formulation constants. */
struct SPHParameters {
    float kernelRadius;
    float particleRadius;
    float pressureStiffness;
    float viscosity;
    float boundaryRestitution;
    bool  bUseAttractivePressure;
    float attractivePressureRatio;
    float stopTime;
    int   numThreads;
    float restDensity;
    float gravity[3];
    float timestep;
    float boxMin[3];
    float boxMax[3];
    float cflNumber;   // 0: fixed timestep; > 0: advanceFrame() takes the CFL step

    /* The GUI defaults (resolution 24 and the panel values). */
    SPHParameters() {
        sph_params p;
        sph_params_default(&p);
        from_c(p);
    }
    /* Reference: updateParams() after kernelRadius changes; the engine's rule
     * is particleRadius = kernelRadius / 4 (lattice spacing = h / 2). */
    void updateParams() { particleRadius = kernelRadius * 0.25f; }
    void setResolution(int res) {
        kernelRadius = 2.0f / (float)res;   // Source/Controller.cpp:55
        updateParams();
    }

    sph_params to_c() const {
        sph_params p;
        p.particle_radius = particleRadius;
        p.kernel_radius = kernelRadius;
        p.rest_density = restDensity;
        p.pressure_stiffness = pressureStiffness;
        p.viscosity = viscosity;
        p.boundary_restitution = boundaryRestitution;
        p.use_attractive_pressure = bUseAttractivePressure ? 1 : 0;
        p.attractive_pressure_ratio = attractivePressureRatio;
        for (int k = 0; k < 3; ++k) {
            p.gravity[k] = gravity[k];
            p.box_min[k] = boxMin[k];
            p.box_max[k] = boxMax[k];
        }
        p.dt = timestep;
        p.stop_time = stopTime;
        p.num_threads = numThreads;
        p.cfl_number = cflNumber;
        return p;
    }
    void from_c(const sph_params& p) {
        particleRadius = p.particle_radius;
        kernelRadius = p.kernel_radius;
        restDensity = p.rest_density;
        pressureStiffness = p.pressure_stiffness;
        viscosity = p.viscosity;
        boundaryRestitution = p.boundary_restitution;
        bUseAttractivePressure = p.use_attractive_pressure != 0;
        attractivePressureRatio = p.attractive_pressure_ratio;
        for (int k = 0; k < 3; ++k) {
            gravity[k] = p.gravity[k];
            boxMin[k] = p.box_min[k];
            boxMax[k] = p.box_max[k];
        }
        timestep = p.dt;
        stopTime = p.stop_time;
        numThreads = p.num_threads;
        cflNumber = p.cfl_number;
    }
};

template <class Vec3 = Vec3f>
class QtSPHSolver {
    static_assert(sizeof(Vec3) == 3 * sizeof(float), "Vec3 must be three packed floats");

public:
    using Vec_Vec3 = std::vector<Vec3>;

    explicit QtSPHSolver(std::shared_ptr<SPHParameters> params) : params_(std::move(params)) {
        if (!params_) throw Error(SPH_E_INVALID, "null SPHParameters");
        const sph_params p = params_->to_c();
        sph_handle* h = nullptr;
        const int rc = sph_create(&p, &h);
        if (rc != SPH_OK) throw Error(rc, std::string("sph_create: ") + sph_last_error(nullptr));
        h_ = h;
    }
    ~QtSPHSolver() { sph_destroy(h_); }
    QtSPHSolver(const QtSPHSolver&) = delete;
    QtSPHSolver& operator=(const QtSPHSolver&) = delete;

    /* Source/Simulator.cpp:42: parameters and the host arrays go to the device;
     * then block sizes are picked for this size (sph_autotune: about 10 ms the
     * first time, cached per process; no result depends on them).  A failed
     * autotune (e.g. no room under a device memory cap for its copy of the
     * state) leaves the state as it was and keeps the default block sizes. */
    void makeReady() {
        const sph_params p = params_->to_c();
        check(sph_set_params(h_, &p));
        if (!velocities_.empty() && velocities_.size() != particles_.size())
            throw Error(SPH_E_INVALID, "velocity array size differs from the particle array");
        const float* v = velocities_.empty() ? nullptr : reinterpret_cast<const float*>(velocities_.data());
        check(sph_set_particles(h_, reinterpret_cast<const float*>(particles_.data()), v,
                                (uint32_t)particles_.size()));
        check(sph_make_ready(h_));
        (void)sph_autotune(h_);
        velocities_.resize(particles_.size(), Vec3{0.f, 0.f, 0.f});
        host_stale_ = false;
    }

    /* Source/Simulator.cpp:49: one step; returns the time it advanced. */
    float advanceFrame() {
        float dt = 0.f;
        check(sph_advance_frame(h_, &dt));
        host_stale_ = true;
        return dt;
    }

    /* The inner frame loop of Source/Simulator.cpp:46-51 on the device:
     * steps until the accumulated float time reaches 1/30 s. */
    float advanceFrameBatch(uint32_t* steps = nullptr) {
        float t = 0.f;
        uint32_t k = 0;
        check(sph_run_frame(h_, &t, &k));
        host_stale_ = true;
        if (steps) *steps = k;
        return t;
    }

    unsigned int getNumParticles() const {
        return h_ && sph_num_particles(h_) ? sph_num_particles(h_) : (unsigned int)particles_.size();
    }
    /* Include/QtSPHSolver.h:34-35.  Writing into these arrays changes the
     * state only at the next makeReady() (the reference's setupScene path). */
    Vec_Vec3& getParticles() {
        refresh();
        return particles_;
    }
    Vec_Vec3& getVelocity() {
        refresh();
        return velocities_;
    }
    /* Frame-cadence hand-off to a renderer: snapshotBegin() captures the
     * positions now and returns; later advanceFrame() calls overlap the copy
     * to the host; snapshotWait() fills `out` with the captured positions. */
    void snapshotBegin() { check(sph_snapshot_begin(h_)); }
    void snapshotWait(Vec_Vec3& out) {
        out.resize(sph_num_particles(h_));
        check(sph_snapshot_wait(h_, reinterpret_cast<float*>(out.data())));
    }
    /* State dump of the current frame (.npy, float32 (n, 6): x, v in original order). */
    void exportFrame(const std::string& path) { check(sph_export_frame(h_, path.c_str())); }
    std::shared_ptr<SPHParameters> getParameters() const { return params_; }
    sph_handle* handle() const { return h_; }

private:
    void check(int rc) const {
        if (rc != SPH_OK) throw Error(rc, sph_last_error(h_));
    }
    void refresh() {
        if (!host_stale_) return;
        const uint32_t n = sph_num_particles(h_);
        particles_.resize(n);
        velocities_.resize(n);
        check(sph_get_particles(h_, reinterpret_cast<float*>(particles_.data())));
        check(sph_get_velocity(h_, reinterpret_cast<float*>(velocities_.data())));
        host_stale_ = false;
    }

    std::shared_ptr<SPHParameters> params_;
    sph_handle* h_ = nullptr;
    Vec_Vec3 particles_, velocities_;
    bool host_stale_ = false;
};

}  // namespace sph

#endif /* SPH_SOLVER_HPP */
