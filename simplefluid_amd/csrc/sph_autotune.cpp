// Autotuner: picks the block size of the density, force and reorder passes,
// the keys per thread of the sort's digit passes, whether density and force
// run as persistent blocks over tile queues and whether the force and
// density passes stage their rows in LDS, by timing each candidate
// with HIP events on the handle's own state.  The state is saved first and
// restored before every trial, so all candidates run the same steps.
// Results are cached per process, keyed by device arch, particle-count
// bucket (power of two) and scene count.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>

#include "sph_engine.hpp"

using namespace sph;

namespace {

std::mutex g_mu;
std::map<std::string, Tuning> g_cache;

std::string cache_key(const sph_handle* h) {
    int bucket = 0;
    while ((1u << bucket) < h->n && bucket < 31) ++bucket;
    return h->arch + ":" + std::to_string(bucket) + ":" + std::to_string(h->sc_h.size()) + ":" +
           std::to_string(h->xsub);
}

struct Saved {
    DevBuf<float4> pos, vel;
    DevBuf<uint32_t> id;
};

}  // namespace

extern "C" int sph_autotune(sph_handle* h) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (!h->ready) return fail(h, SPH_E_NOT_READY, "call sph_make_ready first");
    const std::string key = cache_key(h);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
            h->tune = it->second;
            invalidate_graph(h);
            return SPH_OK;
        }
    }
    const size_t n = h->n;
    Saved sv;
    hipError_t e = sv.pos.ensure(n);
    if (e == hipSuccess) e = sv.vel.ensure(n);
    if (e == hipSuccess) e = sv.id.ensure(n);
    if (e == hipSuccess) e = hipMemcpyAsync(sv.pos.p, h->pos.p, n * sizeof(float4), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sv.vel.p, h->vel.p, n * sizeof(float4), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sv.id.p, h->id.p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "autotune save");

    const bool prof_was = h->prof;
    auto restore = [&]() -> hipError_t {
        hipError_t x = hipMemcpyAsync(h->pos.p, sv.pos.p, n * sizeof(float4), hipMemcpyDeviceToDevice, h->stream);
        if (x == hipSuccess) x = hipMemcpyAsync(h->vel.p, sv.vel.p, n * sizeof(float4), hipMemcpyDeviceToDevice, h->stream);
        if (x == hipSuccess) x = hipMemcpyAsync(h->id.p, sv.id.p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, h->stream);
        h->keys_valid = false;
        return x;
    };
    // Mean milliseconds per pass over `steps` profiled steps from the saved state.
    auto trial = [&](double* ms) -> int {
        if (restore() != hipSuccess) return SPH_E_HIP;
        h->prof = false;
        int rc = enqueue_step(h, h->work, false);  // warm-up (hash, caches)
        if (rc != SPH_OK) return rc;
        double tmp[SPH_NUM_PASSES];
        uint32_t cnt = 0;
        rc = sph_profile_read(h, tmp, &cnt);
        if (rc != SPH_OK) return rc;
        const int steps = 4;
        for (int s = 0; s < steps; ++s) {
            rc = enqueue_step(h, h->work, true);
            if (rc != SPH_OK) return rc;
        }
        rc = sph_profile_read(h, ms, &cnt);
        if (rc != SPH_OK) return rc;
        for (int p = 0; p < SPH_NUM_PASSES; ++p) ms[p] /= std::max(1u, cnt);
        return check_device_error(h);
    };

    const int blocks[] = {64, 128, 256, 512};
    const int items[] = {4, 8, 12, 16};
    Tuning best = h->tune;
    double ms[SPH_NUM_PASSES];
    int rc = SPH_OK;
    // Density and force blocks, then reorder block, then sort items.
    struct Axis { int* field; const int* cand; int ncand; int pass; };
    Axis axes[] = {{&best.density_block, blocks, 4, SPH_PASS_DENSITY},
                   {&best.force_block, blocks, 4, SPH_PASS_FORCE},
                   {&best.reorder_block, blocks, 4, SPH_PASS_REORDER},
                   {&best.sort_items, items, 4, SPH_PASS_SORT}};
    for (const Axis& ax : axes) {
        // Keys per thread only shape the radix sort's digit passes.
        if (ax.field == &best.sort_items && best.sort_mode != kSortRadix) continue;
        double best_ms = 1e30;
        int best_v = *ax.field;
        for (int c = 0; c < ax.ncand; ++c) {
            *ax.field = ax.cand[c];
            h->tune = best;
            rc = trial(ms);
            if (rc != SPH_OK) break;
            if (ms[ax.pass] < best_ms) { best_ms = ms[ax.pass]; best_v = ax.cand[c]; }
        }
        *ax.field = best_v;
        if (rc != SPH_OK) break;
    }
    // Launch form of density and force (persistent or one block per tile;
    // force rows staged in LDS or gathered), judged on their summed time.
    if (rc == SPH_OK) {
        double best_ms = 1e30;
        int best_v = 0;
        for (int c = 0; c < 4; ++c) {
            best.persistent = c & 1;
            best.force_staged = c >> 1;
            h->tune = best;
            rc = trial(ms);
            if (rc != SPH_OK) break;
            const double t = ms[SPH_PASS_DENSITY] + ms[SPH_PASS_FORCE];
            if (t < best_ms) { best_ms = t; best_v = c; }
        }
        best.persistent = best_v & 1;
        best.force_staged = best_v >> 1;
    }
    // Density rows staged in LDS or read from global memory.
    if (rc == SPH_OK) {
        double best_ms = 1e30;
        int best_v = 0;
        for (int c = 0; c < 2; ++c) {
            best.density_staged = c;
            h->tune = best;
            rc = trial(ms);
            if (rc != SPH_OK) break;
            if (ms[SPH_PASS_DENSITY] < best_ms) { best_ms = ms[SPH_PASS_DENSITY]; best_v = c; }
        }
        best.density_staged = best_v;
    }
    h->tune = best;
    h->prof = prof_was;
    const hipError_t re = restore();
    invalidate_graph(h);
    if (rc != SPH_OK) return rc;
    if (re != hipSuccess) return hip_fail(h, re, "autotune restore");
    rc = sync(h);
    if (rc != SPH_OK) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best;
    return SPH_OK;
}
