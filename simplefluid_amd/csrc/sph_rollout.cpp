// Differentiable rollout: forward with state checkpoints, reverse pass that
// recomputes one segment at a time onto a tape of step intermediates and
// runs the hand-written adjoint kernels over it (sph_adjoint.hip).
//
// Memory: (steps/every + 1) state checkpoints of 36 B per particle, plus
// `every` taped steps of ~116 B per particle and 8 B per cell.
#include <algorithm>
#include <cstring>

#include "sph_engine.hpp"

using namespace sph;

namespace {

hipError_t copy_state(sph_handle* h, float4* dp, float4* dv, uint32_t* di, const float4* sp, const float4* sv,
                      const uint32_t* si) {
    const size_t n = h->n;
    hipError_t e = hipMemcpyAsync(dp, sp, n * sizeof(float4), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dv, sv, n * sizeof(float4), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(di, si, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, h->stream);
    return e;
}

hipError_t save_checkpoint(sph_handle* h, RolloutState& R, uint32_t slot) {
    const size_t n = h->n;
    return copy_state(h, R.ck_pos.p + slot * n, R.ck_vel.p + slot * n, R.ck_id.p + slot * n, h->pos.p, h->vel.p,
                      h->id.p);
}

hipError_t load_checkpoint(sph_handle* h, RolloutState& R, uint32_t slot) {
    const size_t n = h->n;
    h->keys_valid = false;
    return copy_state(h, h->pos.p, h->vel.p, h->id.p, R.ck_pos.p + slot * n, R.ck_vel.p + slot * n,
                      R.ck_id.p + slot * n);
}

}  // namespace

extern "C" {

int sph_rollout_forward(sph_handle* h, uint32_t steps, uint32_t every) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (!h->ready) return fail(h, SPH_E_NOT_READY, "call sph_make_ready first");
    if (steps == 0 || every == 0) return fail(h, SPH_E_INVALID, "steps and checkpoint interval must be > 0");
    if (h->cfl) return fail(h, SPH_E_INVALID, "the differentiable rollout needs a fixed dt (cfl_number = 0)");
    if (!h->rollout) h->rollout.reset(new (std::nothrow) RolloutState());
    if (!h->rollout) return fail(h, SPH_E_OOM, "rollout state");
    RolloutState& R = *h->rollout;
    R.valid = false;
    R.steps = steps;
    R.every = std::min(every, steps);
    R.nseg = (steps + R.every - 1) / R.every;
    R.last_taped = false;
    const size_t n = h->n;
    hipError_t e = R.ck_pos.ensure((R.nseg + 1) * n);
    if (e == hipSuccess) e = R.ck_vel.ensure((R.nseg + 1) * n);
    if (e == hipSuccess) e = R.ck_id.ensure((R.nseg + 1) * n);
    if (e != hipSuccess) return hip_fail(h, e, "checkpoint allocation");
    // One segment's tape (each step's sorted state, cells, density, masks,
    // acceleration); the last segment is recorded on it as it runs, so the
    // reverse pass starts without recomputing that segment.
    while (R.tape.size() < R.every) R.tape.emplace_back(new StepBufs());
    for (uint32_t t = 0; t < R.every; ++t) {
        const int rc = alloc_step_bufs(h, *R.tape[t]);
        if (rc != SPH_OK) return rc;
        e = R.tape[t]->acc.ensure(n);
        if (e != hipSuccess) return hip_fail(h, e, "tape allocation");
    }
    for (uint32_t s = 0; s < R.nseg; ++s) {
        e = save_checkpoint(h, R, s);
        if (e != hipSuccess) return hip_fail(h, e, "checkpoint copy");
        const uint32_t len = std::min(R.every, steps - s * R.every);
        if (s + 1 < R.nseg) {
            const int rc = sph_advance_steps(h, len);
            if (rc != SPH_OK) return rc;
        } else {
            for (uint32_t t = 0; t < len; ++t) {
                const int rc = enqueue_step(h, *R.tape[t], false);
                if (rc != SPH_OK) return rc;
            }
        }
    }
    e = save_checkpoint(h, R, R.nseg);
    if (e != hipSuccess) return hip_fail(h, e, "checkpoint copy");
    R.valid = true;
    R.last_taped = true;
    R.n = h->n;
    R.gen = h->config_gen;
    return SPH_OK;
}

int sph_rollout_backward(sph_handle* h, const float* dL_dx_final, float* dL_dB, float* dL_dmu, float* dL_dv0) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !dL_dx_final) return SPH_E_INVALID;
    if (!h->rollout || !h->rollout->valid) return fail(h, SPH_E_NOT_READY, "run sph_rollout_forward first");
    RolloutState& R = *h->rollout;
    if (R.n != h->n || R.gen != h->config_gen)
        return fail(h, SPH_E_NOT_READY,
                    "particles, parameters or grid changed since sph_rollout_forward; run it again");
    const size_t n = h->n;
    const uint32_t ns = (uint32_t)h->sc_h.size();
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
    while (R.tape.size() < R.every) R.tape.emplace_back(new StepBufs());
    for (uint32_t t = 0; t < R.every; ++t) {
        const int rc = alloc_step_bufs(h, *R.tape[t]);
        if (rc != SPH_OK) return rc;
        chk(R.tape[t]->acc.ensure(n));
    }
    chk(R.xbar.ensure(n)); chk(R.vbar.ensure(n));
    chk(R.xs.ensure(n)); chk(R.vs.ensure(n)); chk(R.abar.ensure(n));
    chk(R.rhobar.ensure(n));
    chk(R.gB.ensure(ns)); chk(R.gmu.ensure(ns));
    chk(R.gbuf.ensure(3 * n + 1));
    uint32_t ntiles = 0;
    const uint2* tiles = tiles_for(h, 256, ntiles);  // tiles of 256 particles, scene by scene
    if (!tiles) return fail(h, SPH_E_OOM, "tile table");
    // First tile of each scene, as the tile table lists them.
    std::vector<uint32_t> first(ns + 1, 0);
    for (uint32_t s = 0; s < ns; ++s)
        first[s + 1] = first[s] + (h->sc_h[s].p_end - h->sc_h[s].p_begin + 255u) / 256u;
    if (first[ns] != ntiles) return fail(h, SPH_E_INVALID, "tile table does not match the scenes");
    chk(R.tile_acc.ensure(std::max<uint32_t>(ntiles, 1)));
    chk(R.tile_first.ensure(ns + 1));
    if (e != hipSuccess) return hip_fail(h, e, "adjoint allocation");

    hipStream_t st = h->stream;
    chk(hipMemcpy(R.tile_first.p, first.data(), (ns + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    chk(hipMemsetAsync(R.tile_acc.p, 0, std::max<uint32_t>(ntiles, 1) * sizeof(double2), st));
    chk(hipMemcpyAsync(R.gbuf.p, dL_dx_final, 3 * n * sizeof(float), hipMemcpyHostToDevice, st));
    chk(launch_gather3(st, (uint32_t)n, R.gbuf.p, R.ck_id.p + (size_t)R.nseg * n, R.xbar.p));
    chk(launch_pack3(st, (uint32_t)n, nullptr, R.vbar.p));
    if (e != hipSuccess) return hip_fail(h, e, "adjoint seed");

    const uint32_t grid = ntiles;  // one block per tile
    for (int seg = (int)R.nseg - 1; seg >= 0; --seg) {
        const uint32_t len = std::min(R.every, R.steps - (uint32_t)seg * R.every);
        if (seg == (int)R.nseg - 1 && R.last_taped) {
            R.last_taped = false;  // the tape is about to hold earlier segments
        } else {
            chk(load_checkpoint(h, R, (uint32_t)seg));
            if (e != hipSuccess) return hip_fail(h, e, "checkpoint restore");
            for (uint32_t t = 0; t < len; ++t) {
                const int rc = enqueue_step(h, *R.tape[t], false);
                if (rc != SPH_OK) return rc;
            }
        }
        for (int t = (int)len - 1; t >= 0; --t) {
            StepBufs& w = *R.tape[t];
            AdjArgs a;
            std::memset(&a, 0, sizeof(a));
            a.pos_s = w.pos_s.p; a.vel_s = w.vel_s.p; a.acc = w.acc.p; a.rho = w.rho.p;
            a.mask = w.mask.p; a.cstart = w.cstart.p; a.perm = w.perm.p;
            a.xbar = R.xbar.p; a.vbar = R.vbar.p; a.xs = R.xs.p; a.vs = R.vs.p; a.abar = R.abar.p;
            a.rhobar = R.rhobar.p; a.tile_acc = R.tile_acc.p;
            a.scenes = h->sc_d.p; a.tiles = tiles; a.ntiles = ntiles; a.n = (uint32_t)n; a.ncells = h->ncells;
            chk(launch_adj_integrate(st, 256, grid, a));
            // The reverse walk takes the forward force pass's form.
            chk(launch_adj_force(st, 256, grid, a, h->tune.force_staged));
            chk(launch_adj_density(st, 256, grid, a, h->tune.force_staged));
            if (e != hipSuccess) return hip_fail(h, e, "adjoint launch");
        }
    }
    chk(launch_scene_tile_sums(st, ns, R.tile_first.p, R.tile_acc.p, R.gB.p, R.gmu.p));
    // vbar is now in the initial state's order (checkpoint 0's ids).
    if (dL_dv0) {
        chk(launch_unpermute3(st, (uint32_t)n, R.vbar.p, R.ck_id.p, R.gbuf.p));
        chk(hipMemcpyAsync(dL_dv0, R.gbuf.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    std::vector<double> gb(ns), gm(ns);
    chk(hipMemcpyAsync(gb.data(), R.gB.p, ns * sizeof(double), hipMemcpyDeviceToHost, st));
    chk(hipMemcpyAsync(gm.data(), R.gmu.p, ns * sizeof(double), hipMemcpyDeviceToHost, st));
    chk(load_checkpoint(h, R, R.nseg));
    if (e != hipSuccess) return hip_fail(h, e, "adjoint readback");
    const int rc = sync(h);
    if (rc != SPH_OK) return rc;
    for (uint32_t s = 0; s < ns; ++s) {
        if (dL_dB) dL_dB[s] = (float)gb[s];
        if (dL_dmu) dL_dmu[s] = (float)gm[s];
    }
    return SPH_OK;
}

}  // extern "C"
