// Host side of the engine: handle lifecycle, step enqueue, HIP graph replay,
// per-pass profiling and the C ABI declared in include/sph_engine.h.
#include "sph_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

using namespace sph;

sph_handle::~sph_handle() {
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    if (snap_host) (void)hipHostFree(snap_host);
    if (snap_ready) (void)hipEventDestroy(snap_ready);
    if (snap_done) (void)hipEventDestroy(snap_done);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (host_pinned) (void)hipHostFree(host_pinned);
    sph::invalidate_graph(this);
    for (auto e : prof_events) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace sph {

// Message of the last failure on this thread that had no handle to carry it
// (sph_create, sph_batch_create); sph_last_error(NULL) returns it.
// A fixed buffer, so that recording a failure allocates nothing.
thread_local char t_last_error[512] = "no error";

void set_thread_error(const char* msg) { std::snprintf(t_last_error, sizeof(t_last_error), "%s", msg); }
thread_local MemBudget* t_budget = nullptr;

int fail(sph_handle* h, int code, const std::string& msg) {
    if (h && code == SPH_E_OOM && h->mem.refused) {
        h->last_error = msg + " (an allocation of " + std::to_string(h->mem.refused) +
                        " bytes would exceed the handle's device memory limit of " + std::to_string(h->mem.limit) +
                        " bytes; " + std::to_string(h->mem.used) + " in use)";
        h->mem.refused = 0;
    } else if (h) {
        h->last_error = msg;
    } else {
        set_thread_error(msg.c_str());
    }
    return code;
}

int hip_fail(sph_handle* h, hipError_t e, const char* what) {
    if (e == hipSuccess) return SPH_OK;
    const int code = (e == hipErrorOutOfMemory) ? SPH_E_OOM : SPH_E_HIP;
    return fail(h, code, std::string(what) + ": " + hipGetErrorString(e));
}

static bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// Derived constants, formed exactly as the CPU restatement forms them:
// normalisations in double from the float inputs, rounded once to float.
// The passes read the C + 1 cell-start words through buffer resources,
// whose byte range and offsets are 32-bit: 4 (C + 1) must stay below 2^32.
constexpr uint64_t kMaxCells = (1ull << 30) - 2;

// err receives a static message on failure.
int derive_scene_core(const sph_params& p, SceneConst& sc, const char** err, int xsub) {
    if (!(p.particle_radius > 0.f) || !std::isfinite(p.particle_radius)) { *err = "particle_radius must be > 0"; return SPH_E_INVALID; }
    if (!(p.kernel_radius > 0.f) || !std::isfinite(p.kernel_radius)) { *err = "kernel_radius must be > 0"; return SPH_E_INVALID; }
    if (!(p.rest_density > 0.f) || !std::isfinite(p.rest_density)) { *err = "rest_density must be > 0"; return SPH_E_INVALID; }
    if (!(p.pressure_stiffness >= 0.f) || !std::isfinite(p.pressure_stiffness)) { *err = "pressure_stiffness must be >= 0"; return SPH_E_INVALID; }
    if (!(p.viscosity >= 0.f) || !std::isfinite(p.viscosity)) { *err = "viscosity must be >= 0"; return SPH_E_INVALID; }
    if (!(p.boundary_restitution >= 0.f) || !std::isfinite(p.boundary_restitution)) { *err = "boundary_restitution must be >= 0"; return SPH_E_INVALID; }
    if (!(p.dt > 0.f) || !std::isfinite(p.dt)) { *err = "dt must be > 0"; return SPH_E_INVALID; }
    if (!finite3(p.gravity) || !finite3(p.box_min) || !finite3(p.box_max)) { *err = "gravity and box must be finite"; return SPH_E_INVALID; }
    if (!std::isfinite(p.attractive_pressure_ratio)) { *err = "attractive_pressure_ratio must be finite"; return SPH_E_INVALID; }
    std::memset(&sc, 0, sizeof(sc));
    const float r = p.particle_radius;
    const float sp = 2.0f * r;
    const float mass = p.rest_density * (sp * sp * sp);
    sc.h = p.kernel_radius;
    sc.h2 = sc.h * sc.h;
    const double hd = (double)p.kernel_radius;
    const double h3 = hd * hd * hd, h6 = h3 * h3, h9 = h6 * h3;
    const double poly6 = 315.0 / (64.0 * M_PI * h9);
    const double spiky = 45.0 / (M_PI * h6);
    sc.cD = (float)((double)mass * poly6);
    sc.cP = (float)((double)mass * spiky);
    sc.cV = (float)((double)p.viscosity * (double)mass * spiky);
    sc.inv_rho0 = (float)(1.0 / (double)p.rest_density);
    sc.inv_cs = (float)(1.0 / (hd * (1.0 + 1.0 / 1024.0)));
    if (xsub != 1 && xsub != 2 && xsub != 4) { *err = "x subdivision must be 1, 2 or 4"; return SPH_E_INVALID; }
    sc.inv_csx = sc.inv_cs * (float)xsub;  // exact: power-of-two scale
    sc.xsubf = (float)xsub;
    {
        const float hh = sc.h * sc.inv_cs;
        sc.hc2 = hh * hh;
    }
    sc.wmargin = 1.0f / 256.0f;
    sc.B = p.pressure_stiffness;
    sc.attr = p.use_attractive_pressure ? p.attractive_pressure_ratio : 0.0f;
    sc.attractive = p.use_attractive_pressure ? 1 : 0;
    sc.e = p.boundary_restitution;
    sc.dt = p.dt;
    sc.dt_max = p.dt;
    if (!(p.cfl_number >= 0.f) || !std::isfinite(p.cfl_number)) { *err = "cfl_number must be >= 0"; return SPH_E_INVALID; }
    sc.cfl = p.cfl_number;
    sc.c0 = (float)std::sqrt(7.0 * (double)p.pressure_stiffness / (double)p.rest_density);
    sc.mu = p.viscosity;
    sc.rho0 = p.rest_density;
    sc.inv_h = (float)(1.0 / (double)p.kernel_radius);
    sc.cW = (float)(90.0 * (double)p.rest_density / (double)p.kernel_radius);
    uint64_t cells = 1;
    for (int k = 0; k < 3; ++k) {
        sc.g[k] = p.gravity[k];
        sc.bmin[k] = p.box_min[k];
        sc.wmax[k] = p.box_max[k];
        const float ext = p.box_max[k] - p.box_min[k];
        if (!(ext > 2.0f * r)) { *err = "box must be wider than 2*particle_radius on every axis"; return SPH_E_INVALID; }
        int G = (int)std::ceil(ext * (k == 0 ? sc.inv_csx : sc.inv_cs));
        if (G < 1) G = 1;
        if (G > (k == 0 ? 1024 * xsub : 1024)) {
            *err = "grid exceeds 1024 cells on an axis (box / kernel_radius too large)";
            return SPH_E_INVALID;
        }
        sc.G[k] = G;
        sc.lo[k] = p.box_min[k] + r;
        sc.hi[k] = p.box_max[k] - r;
        cells *= (uint64_t)G;
    }
    if (cells > kMaxCells) { *err = "too many grid cells (limit 2^30 - 2)"; return SPH_E_INVALID; }
    return SPH_OK;
}

int derive_scene(const sph_params& p, SceneConst& sc, std::string& err, int xsub) {
    const char* msg = nullptr;
    const int rc = derive_scene_core(p, sc, &msg, xsub);
    if (rc != SPH_OK) err = msg;
    return rc;
}

int derive_scene_noalloc(const sph_params& p, SceneConst& sc, int xsub) {
    const char* msg = nullptr;
    return derive_scene_core(p, sc, &msg, xsub);
}

void invalidate_graph(sph_handle* h) {
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    if (h->gexec_multi) (void)hipGraphExecDestroy(h->gexec_multi);
    if (h->graph_multi) (void)hipGraphDestroy(h->graph_multi);
    if (h->gexec_frame) (void)hipGraphExecDestroy(h->gexec_frame);
    if (h->graph_frame) (void)hipGraphDestroy(h->graph_frame);
    h->gexec = nullptr;
    h->graph = nullptr;
    h->gexec_multi = nullptr;
    h->graph_multi = nullptr;
    h->gexec_frame = nullptr;
    h->graph_frame = nullptr;
}

uint2* tiles_for(sph_handle* h, int block, uint32_t& ntiles) {
    auto it = h->tiles.find(block);
    if (it != h->tiles.end() && it->second.p) {
        ntiles = h->ntiles[block];
        return it->second.p;
    }
    std::vector<uint2> t;
    for (uint32_t s = 0; s < (uint32_t)h->sc_h.size(); ++s)
        for (uint32_t b = h->sc_h[s].p_begin; b < h->sc_h[s].p_end; b += (uint32_t)block)
            t.push_back(make_uint2(s, b));
    DevBuf<uint2>& d = h->tiles[block];
    if (d.ensure(std::max<size_t>(t.size(), 1)) != hipSuccess) { ntiles = 0; return nullptr; }
    if (!t.empty()) (void)hipMemcpy(d.p, t.data(), t.size() * sizeof(uint2), hipMemcpyHostToDevice);
    h->ntiles[block] = (uint32_t)t.size();
    ntiles = (uint32_t)t.size();
    return d.p;
}

// Persistent launches keep as many blocks resident as the kernel's
// occupancy allows (2048 lanes per CU, or what the staged force pass's LDS
// and registers leave) and loop over tile queues; they are used only when
// there are more tiles than that.
static uint32_t resident_blocks(const sph_handle* h, int block, int blocks_per_cu) {
    const int per_cu = blocks_per_cu > 0 ? blocks_per_cu : std::max(1, 2048 / block);
    return (uint32_t)std::max(8, h->cu_count) * (uint32_t)per_cu;
}

uint32_t grid_for(sph_handle* h, int block, uint32_t ntiles, uint32_t** queue, int blocks_per_cu) {
    if (queue) *queue = nullptr;
    const uint32_t resident = resident_blocks(h, block, blocks_per_cu);
    if (!h->tune.persistent || !queue || ntiles <= resident) return ntiles;
    *queue = h->queue.p;
    return resident;
}

StepArgs step_args(sph_handle* h, StepBufs& w, int block) {
    StepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.pos_s = w.pos_s.p;
    a.soa = w.soa.p;
    a.vel_s = w.vel_s.p;
    a.pos_o = h->pos.p;
    a.vel_o = h->vel.p;
    a.id_s = w.id_s.p;
    a.id_o = h->id.p;
    a.key0 = h->key0.p;
    a.rank0 = h->rank0.p;
    a.count = h->count.p;
    a.cstart = w.cstart.p;
    a.rho = w.rho.p;
    a.mask = w.mask.p;
    a.acc = (h->keep_acc || w.acc.p) ? w.acc.p : nullptr;
    a.scenes = h->sc_d.p;
    a.tiles = tiles_for(h, block, a.ntiles);
    a.n = h->n;
    a.ncells = h->ncells;
    return a;
}

int alloc_step_bufs(sph_handle* h, StepBufs& w) {
    const size_t n = h->n;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
    chk(w.pos_s.ensure(n));
    chk(w.soa.ensure(3 * n));
    chk(w.vel_s.ensure(n));
    chk(w.id_s.ensure(n));
    chk(w.skey.ensure(n));
    chk(w.perm.ensure(n));
    chk(w.cstart.ensure((size_t)h->ncells + 1));
    // The table's contents are for another grid (or none) now: every tile
    // is rewritten at the next scan.
    chk(w.tile_pre.ensure(scan_tiles(h->ncells)));
    if (e == hipSuccess) chk(launch_fill_u32(h->stream, w.tile_pre.p, 0xffffffffu, scan_tiles(h->ncells)));
    if (h->tune.sort_mode == kSortCellCount) {
        chk(w.tmp.ensure(n));
    } else {
        for (int i = 0; i < 2; ++i) { chk(w.kbuf[i].ensure(n)); chk(w.vbuf[i].ensure(n)); }
    }
    chk(w.rho.ensure(n));
    chk(w.mask.ensure((size_t)kRows * n));
    if (h->keep_acc) chk(w.acc.ensure(n));
    return hip_fail(h, e, "device allocation");
}

static int ensure_buffers(sph_handle* h) {
    const size_t n = h->n;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
    chk(h->pos.ensure(n));
    chk(h->vel.ensure(n));
    chk(h->id.ensure(n));
    chk(h->key0.ensure(n));
    chk(h->rank0.ensure(n));
    const size_t count_words = (size_t)h->ncells + 2 * (size_t)scan_tiles(h->ncells);
    const bool fresh_count = h->count.count < count_words || !h->count.p;
    chk(h->count.ensure(count_words));
    chk(h->hist.ensure(kMaxPasses * 256));
    chk(h->status.ensure((size_t)kMaxPasses * sort_tiles((uint32_t)n, 4) * 256 + 256));
    chk(h->tickets.ensure(kMaxPasses));
    chk(h->err.ensure(1));
    chk(h->queue.ensure(kQueueWords));
    chk(h->vmax2.ensure(h->sc_h.size()));
    chk(h->dt_dev.ensure(1));
    chk(h->frame.ensure(kFrameWords));
    chk(h->staging.ensure(3 * n + 1));
    chk(h->sc_d.ensure(h->sc_h.size()));
    if (e != hipSuccess) return hip_fail(h, e, "device allocation");
    chk(launch_fill_u32(h->stream, h->hist.p, 0u, kMaxPasses * 256));
    chk(launch_fill_u32(h->stream, h->err.p, 0u, 1));
    chk(launch_fill_u32(h->stream, h->queue.p, 0u, kQueueWords));
    chk(launch_fill_u32(h->stream, h->vmax2.p, 0u, h->sc_h.size()));
    if (fresh_count) chk(launch_fill_u32(h->stream, h->count.p, 0u, count_words));
    chk(hipMemcpyAsync(h->sc_d.p, h->sc_h.data(), h->sc_h.size() * sizeof(SceneConst), hipMemcpyHostToDevice, h->stream));
    if (e != hipSuccess) return hip_fail(h, e, "buffer init");
    return alloc_step_bufs(h, h->work);
}

static void prof_record(sph_handle* h, int slot) {
    const size_t idx = h->prof_used + slot;
    while (h->prof_events.size() <= idx) {
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) return;
        h->prof_events.push_back(ev);
    }
    (void)hipEventRecord(h->prof_events[idx], h->stream);
}

int enqueue_step(sph_handle* h, StepBufs& w, bool record_prof, uint32_t* frame) {
    hipStream_t st = h->stream;
    if (frame) {
        if (!h->keys_valid || h->tune.sort_mode != kSortCellCount)
            return fail(h, SPH_E_INVALID, "the device frame loop needs a steady cell-count step");
        const hipError_t ge = launch_frame_gate(st, frame);
        if (ge != hipSuccess) return hip_fail(h, ge, "frame gate launch");
    }
    const uint8_t hashed = h->keys_valid ? 0 : 1;
    if (record_prof) prof_record(h, 0);
    if (!h->keys_valid) {
        // New state: counts start from zero, then every particle is hashed.
        hipError_t e = launch_fill_u32(st, h->count.p, 0u, (size_t)h->ncells + 2 * (size_t)scan_tiles(h->ncells));
        if (e == hipSuccess && h->cfl) e = launch_fill_u32(st, h->vmax2.p, 0u, h->sc_h.size());
        uint32_t nt = 0;
        const uint2* tl = tiles_for(h, 256, nt);
        if (e == hipSuccess)
            e = launch_hash(st, 256, nt, h->pos.p, h->vel.p, h->key0.p, h->rank0.p, h->count.p, h->ncells,
                            h->cfl ? h->vmax2.p : nullptr, h->sc_d.p, tl, nt);
        if (e != hipSuccess) return hip_fail(h, e, "hash launch");
        h->keys_valid = true;
    }
    if (record_prof) prof_record(h, 1);
    SortPlan sp;
    sp.n = h->n;
    sp.ncells = h->ncells;
    sp.mode = h->tune.sort_mode;
    sp.passes = h->passes;
    sp.items = h->tune.sort_items;
    sp.reorder_block = h->tune.reorder_block;
    sp.key0 = h->key0.p;
    sp.rank0 = h->rank0.p;
    sp.count = h->count.p;
    sp.cstart = w.cstart.p;
    sp.tile_pre = w.tile_pre.p;
    sp.tmp = w.tmp.p;
    for (int i = 0; i < 2; ++i) { sp.kbuf[i] = w.kbuf[i].p; sp.vbuf[i] = w.vbuf[i].p; }
    sp.skey = w.skey.p;
    sp.perm = w.perm.p;
    sp.hist = h->hist.p;
    sp.status = h->status.p;
    sp.tickets = h->tickets.p;
    sp.err = h->err.p;
    sp.queue = h->queue.p;
    sp.pos_in = h->pos.p;
    sp.vel_in = h->vel.p;
    sp.id_in = h->id.p;
    sp.pos_out = w.pos_s.p;
    sp.soa_out = h->tune.density_staged ? nullptr : w.soa.p;
    sp.halt = frame;
    sp.vel_out = w.vel_s.p;
    sp.id_out = w.id_s.p;
    hipError_t e = launch_sort_keys(st, sp);
    if (e != hipSuccess) return hip_fail(h, e, "sort launch");
    if (record_prof) prof_record(h, 2);
    e = launch_sort_gather(st, sp);
    if (e != hipSuccess) return hip_fail(h, e, "reorder launch");
    if (record_prof) prof_record(h, 3);
    {
        StepArgs a = step_args(h, w, h->tune.density_block);
        a.halt = frame;
        const uint32_t grid = grid_for(h, h->tune.density_block, a.ntiles, &a.queue);
        e = launch_density(st, h->tune.density_block, grid, a, h->tune.density_staged);
        if (e != hipSuccess) return hip_fail(h, e, "density launch");
    }
    if (record_prof) prof_record(h, 4);
    if (h->cfl) {
        // Step length from the largest speed of the state this step starts from.
        e = launch_cfl(st, (uint32_t)h->sc_h.size(), h->sc_d.p, h->vmax2.p, h->dt_dev.p, frame);
        if (e != hipSuccess) return hip_fail(h, e, "cfl launch");
    }
    {
        StepArgs a = step_args(h, w, h->tune.force_block);
        a.vmax2 = h->cfl ? h->vmax2.p : nullptr;
        a.halt = frame;
        const uint32_t grid = grid_for(h, h->tune.force_block, a.ntiles, &a.queue,
                                       force_blocks_per_cu(h->tune.force_block, h->tune.force_staged));
        e = launch_force(st, h->tune.force_block, grid, a, h->tune.force_staged);
        if (e != hipSuccess) return hip_fail(h, e, "force launch");
    }
    if (record_prof) {
        prof_record(h, 5);
        h->prof_used += 6;
        h->prof_hashed.push_back(hashed);
    }
    h->has_step = true;
    return SPH_OK;
}

int check_device_error(sph_handle* h) {
    uint32_t flag = 0;
    hipError_t e = hipMemcpy(&flag, h->err.p, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(h, e, "error flag read");
    if (flag) {
        std::string what;
        if (flag & 1u) what += " sort look-back wait exceeded its bound;";
        if (flag & 2u) what += " sort scatter index out of range;";
        if (flag & 4u) what += " sort gather index out of range;";
        if (flag & 8u)
            what += " a cell held more than 1024 particles (collapsed or non-finite positions), so its order "
                    "was the arrival order, not the stable one (SPH_SORT_RADIX has no such limit);";
        return fail(h, SPH_E_DEVICE, "device error flag set:" + what);
    }
    return SPH_OK;
}

int sync(sph_handle* h) {
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "stream synchronize");
    return check_device_error(h);
}

}  // namespace sph

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// x subdivision of the cells: the largest of 4, 2, 1 whose cell count stays
// within max(4N, 2^20), so the per-step cell scan stays small next to the
// particle traffic; a caller override (sph_set_cell_subdivision) wins.
static int pick_xsub(const std::vector<sph_params>& ps, uint64_t n, int override_xsub) {
    if (override_xsub) return override_xsub;
    uint64_t cells1 = 0;
    for (const sph_params& p : ps) {
        SceneConst sc;
        std::string err;
        if (derive_scene(p, sc, err, 1) != SPH_OK) return 1;
        cells1 += (uint64_t)sc.G[0] * sc.G[1] * sc.G[2];
    }
    // Finer x cells shorten the neighbour rows; the cost is a scan over C
    // cells per step.  Up to ~12 cells per particle the shorter rows win.
    const uint64_t budget = std::max<uint64_t>(12 * n, 1u << 20);
    for (int x : {4, 2}) if (cells1 * (uint64_t)x <= budget) return x;
    return 1;
}

static int configure(sph_handle* h, const std::vector<sph_params>& ps, const std::vector<uint32_t>& ns) {
    ++h->config_gen;
    std::vector<SceneConst> sc(ps.size());
    uint64_t cell_off = 0, p_off = 0;
    std::vector<uint32_t> off(ps.size() + 1, 0);
    uint64_t ntot = 0;
    for (uint32_t v : ns) ntot += v;
    const int xsub = pick_xsub(ps, ntot, h->xsub_override);
    for (size_t s = 0; s < ps.size(); ++s) {
        std::string err;
        const int rc = derive_scene(ps[s], sc[s], err, xsub);
        if (rc != SPH_OK) return fail(h, rc, ps.size() > 1 ? "scene " + std::to_string(s) + ": " + err : err);
        if (ps[s].dt != ps[0].dt) return fail(h, SPH_E_INVALID, "all scenes of a batch must share dt");
        if (ps.size() > 1 && ps[s].cfl_number > 0.f)
            return fail(h, SPH_E_INVALID, "the CFL step is for single-scene handles (a batch shares dt)");
        sc[s].cell_offset = (uint32_t)cell_off;
        cell_off += (uint64_t)sc[s].G[0] * sc[s].G[1] * sc[s].G[2];
        sc[s].p_begin = (uint32_t)p_off;
        p_off += ns[s];
        sc[s].p_end = (uint32_t)p_off;
        off[s + 1] = (uint32_t)p_off;
    }
    if (cell_off > kMaxCells) return fail(h, SPH_E_INVALID, "too many grid cells in total (limit 2^30 - 2)");
    if (p_off >= (1ull << 28)) return fail(h, SPH_E_INVALID, "too many particles (limit 2^28)");
    h->params = ps;
    h->scene_n = ns;
    h->scene_off = off;
    h->sc_h = sc;
    h->n = (uint32_t)p_off;
    h->ncells = (uint32_t)cell_off;
    h->passes = sort_key_passes(h->ncells);
    h->xsub = xsub;
    h->cfl = ps.size() == 1 && ps[0].cfl_number > 0.f;
    h->tiles.clear();
    h->ntiles.clear();
    invalidate_graph(h);
    return SPH_OK;
}

static int new_handle(sph_handle** out) {
    if (!out) return SPH_E_INVALID;
    *out = nullptr;
    sph_handle* h = new (std::nothrow) sph_handle();
    if (!h) return SPH_E_OOM;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, dev);
        if (e == hipSuccess) {
            h->cu_count = prop.multiProcessorCount;
            h->arch = prop.gcnArchName;
        }
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return fail(nullptr, SPH_E_HIP, std::string("no usable HIP device: ") + hipGetErrorString(e));
    }
    *out = h;
    return SPH_OK;
}

extern "C" {

int sph_params_default(sph_params* p) {
    if (!p) return SPH_E_INVALID;
    std::memset(p, 0, sizeof(*p));
    p->rest_density = 1000.0f;
    p->pressure_stiffness = 50000.0f;
    p->viscosity = 0.05f;
    p->boundary_restitution = 0.1f;
    p->use_attractive_pressure = 0;
    p->attractive_pressure_ratio = 0.1f;
    p->gravity[0] = 0.0f; p->gravity[1] = -9.81f; p->gravity[2] = 0.0f;
    p->dt = 1e-4f;
    for (int k = 0; k < 3; ++k) { p->box_min[k] = -1.0f; p->box_max[k] = 1.0f; }
    p->stop_time = 5.0f;
    p->num_threads = 0;
    return sph_params_set_resolution(p, 24);
}

int sph_params_set_resolution(sph_params* p, int res) {
    if (!p || res <= 0) return SPH_E_INVALID;
    p->kernel_radius = 2.0f / (float)res;
    p->particle_radius = p->kernel_radius / 4.0f;
    return SPH_OK;
}

int sph_create(const sph_params* p, sph_handle** out) {
    if (!p || !out) return SPH_E_INVALID;
    sph_handle* h = nullptr;
    int rc = new_handle(&h);
    if (rc != SPH_OK) return rc;
    SceneConst tmp;
    std::string err;
    rc = derive_scene(*p, tmp, err, 1);
    if (rc != SPH_OK) {
        delete h;
        return fail(nullptr, rc, err);
    }
    h->params.assign(1, *p);
    h->scene_n.assign(1, 0);
    *out = h;
    return SPH_OK;
}

void sph_destroy(sph_handle* h) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
}

int sph_set_params(sph_handle* h, const sph_params* p) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !p) return SPH_E_INVALID;
    if (h->batched) return fail(h, SPH_E_INVALID, "use sph_set_scene_params on a batch handle");
    SceneConst tmp;
    std::string err;
    int rc = derive_scene(*p, tmp, err, 1);
    if (rc != SPH_OK) return fail(h, rc, err);
    h->params[0] = *p;
    if (h->has_particles) {
        rc = configure(h, h->params, h->scene_n);
        if (rc != SPH_OK) return rc;
        h->ready = false;
        h->keys_valid = false;
    }
    return SPH_OK;
}

int sph_get_params(const sph_handle* h, sph_params* out) {
    if (!h || !out) return SPH_E_INVALID;
    *out = h->params[0];
    return SPH_OK;
}

static int upload_state(sph_handle* h, const float* xyz, const float* vel, const uint32_t* ids) {
    ++h->config_gen;
    const size_t n = h->n;
    int rc = ensure_buffers(h);
    if (rc != SPH_OK) return rc;
    if (n == 0) return SPH_OK;
    // Stage through the handle's 3n scratch: positions, then velocities.
    hipError_t e = hipMemcpyAsync(h->staging.p, xyz, 3 * n * sizeof(float), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = launch_pack3(h->stream, (uint32_t)n, h->staging.p, h->pos.p);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && vel) {
        e = hipMemcpyAsync(h->staging.p, vel, 3 * n * sizeof(float), hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = launch_pack3(h->stream, (uint32_t)n, h->staging.p, h->vel.p);
    } else if (e == hipSuccess) {
        e = launch_pack3(h->stream, (uint32_t)n, nullptr, h->vel.p);
    }
    if (e == hipSuccess) {
        if (ids) {
            e = hipMemcpyAsync(h->id.p, ids, n * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream);
        } else {
            std::vector<uint32_t> iota(n);
            for (size_t i = 0; i < n; ++i) iota[i] = (uint32_t)i;
            e = hipMemcpyAsync(h->id.p, iota.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "state upload");
    h->keys_valid = false;
    invalidate_graph(h);
    return SPH_OK;
}

int sph_set_particles(sph_handle* h, const float* xyz, const float* vel, uint32_t n) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (n > 0 && !xyz) return fail(h, SPH_E_INVALID, "xyz is NULL");
    if (h->batched) {
        if (n != h->n) return fail(h, SPH_E_INVALID, "a batch takes exactly the sum of its scene counts");
    } else {
        int rc = configure(h, h->params, std::vector<uint32_t>(1, n));
        if (rc != SPH_OK) return rc;
    }
    h->has_particles = true;
    h->ready = false;
    h->has_step = false;
    return upload_state(h, xyz, vel, nullptr);
}

int sph_make_ready(sph_handle* h) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (!h->has_particles) return fail(h, SPH_E_NOT_READY, "no particles set");
    if (h->n == 0) return fail(h, SPH_E_NOT_READY, "zero particles");
    int rc = ensure_buffers(h);
    if (rc != SPH_OK) return rc;
    h->ready = true;
    return SPH_OK;
}

static int step_once(sph_handle* h) {
    if (!h->ready) return fail(h, SPH_E_NOT_READY, "call sph_make_ready first");
    return enqueue_step(h, h->work, h->prof);
}

// Captures one steady step; gated: the device frame loop's step.
// Steps in the multi-step graph.
constexpr uint32_t kGraphSteps = 8;

enum GraphKind { kGraphOne, kGraphMulti, kGraphFrame };

static int capture_graph(sph_handle* h, GraphKind kind = kGraphOne) {
    const bool gated = kind == kGraphFrame;
    hipGraphExec_t& gx = gated ? h->gexec_frame : kind == kGraphMulti ? h->gexec_multi : h->gexec;
    hipGraph_t& gr = gated ? h->graph_frame : kind == kGraphMulti ? h->graph_multi : h->graph;
    if (gx) (void)hipGraphExecDestroy(gx);
    if (gr) (void)hipGraphDestroy(gr);
    gx = nullptr;
    gr = nullptr;
    // Tile tables are built on first use with a synchronous copy, which a
    // capturing stream does not allow: build the ones the step uses first
    // (a block size set after steps have run has none yet).
    const int blocks[] = {h->tune.density_block, h->tune.force_block};
    for (int b : blocks) {
        uint32_t nt = 0;
        if (!tiles_for(h, b, nt)) return fail(h, SPH_E_OOM, "tile table");
    }
    (void)force_blocks_per_cu(h->tune.force_block, h->tune.force_staged);  // runtime query, outside capture
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(h, e, "begin capture");
    int rc = SPH_OK;
    const uint32_t steps = kind == kGraphMulti ? kGraphSteps : 1u;
    for (uint32_t s = 0; s < steps && rc == SPH_OK; ++s)
        rc = enqueue_step(h, h->work, false, gated ? h->frame.p : nullptr);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(h->stream, &g);
    if (rc != SPH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(h, e, "end capture");
    gr = g;
    e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        invalidate_graph(h);
        return hip_fail(h, e, "graph instantiate");
    }
    return SPH_OK;
}

int sph_advance_frame(sph_handle* h, float* dt_out) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    int rc = sph_advance_steps(h, 1);
    if (rc != SPH_OK) return rc;
    float dt = h->params[0].dt;
    if (h->cfl) {
        const hipError_t e = hipMemcpyAsync(&dt, h->dt_dev.p, sizeof(float), hipMemcpyDeviceToHost, h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "step length read");
        rc = sync(h);
        if (rc != SPH_OK) return rc;
    }
    if (dt_out) *dt_out = dt;
    return SPH_OK;
}

int sph_advance_steps(sph_handle* h, uint32_t k) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (!h->ready) return fail(h, SPH_E_NOT_READY, "call sph_make_ready first");
    for (uint32_t i = 0; i < k;) {
        // The graphs hold steady-state steps: keys from the last integrate,
        // a cell table to clear sparsely.  Earlier steps run eagerly.
        const bool steady = h->keys_valid && !h->keep_acc;
        if (h->use_graph && steady && !h->prof) {
            const bool multi = k - i >= kGraphSteps;
            hipGraphExec_t& gx = multi ? h->gexec_multi : h->gexec;
            if (!gx) {
                const int rc = capture_graph(h, multi ? kGraphMulti : kGraphOne);
                if (rc != SPH_OK) return rc;
            }
            const hipError_t e = hipGraphLaunch(gx, h->stream);
            if (e != hipSuccess) return hip_fail(h, e, "graph launch");
            i += multi ? kGraphSteps : 1u;
        } else {
            const int rc = step_once(h);
            if (rc != SPH_OK) return rc;
            ++i;
        }
    }
    return SPH_OK;
}

// Frame loop of the adaptive step on the device: the gated step graph is
// replayed in batches (sized from the last frame) and the frame state read
// once per batch, instead of the step length after every step.  Steps
// launched after the frame is complete do nothing.  The frame time is the
// same float sum, compared with the same bound, as the host loop's.
static int run_frame_on_device(sph_handle* h, float& frame_time, uint32_t& steps) {
    hipStream_t st = h->stream;
    hipError_t e = launch_frame_reset(st, h->frame.p);
    if (e != hipSuccess) return hip_fail(h, e, "frame reset");
    if (!h->gexec_frame) {
        const int rc = capture_graph(h, kGraphFrame);
        if (rc != SPH_OK) return rc;
    }
    uint32_t batch = h->frame_steps_hint ? h->frame_steps_hint + 1 : 16;
    uint64_t launched = 0;
    uint32_t fs[kFrameWords] = {0u, 0u, 0u, 0u};
    for (;;) {
        for (uint32_t i = 0; i < batch; ++i) {
            e = hipGraphLaunch(h->gexec_frame, st);
            if (e != hipSuccess) return hip_fail(h, e, "graph launch");
        }
        launched += batch;
        e = hipMemcpyAsync(fs, h->frame.p, sizeof(fs), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(h, e, "frame state read");
        const int rc = sync(h);
        if (rc != SPH_OK) return rc;
        if (fs[1]) break;  // the frame reached 1/30 s
        float ft;
        std::memcpy(&ft, &fs[3], sizeof(ft));
        const double per = fs[2] ? (double)ft / (double)fs[2] : 1e-4;
        const double need = (0.0333333333 - (double)ft) / std::max(per, 1e-12);
        batch = (uint32_t)std::min(std::max(need + 1.0, 1.0), 65536.0);
        if (launched > (1ull << 24)) return fail(h, SPH_E_DEVICE, "the frame did not reach 1/30 s");
    }
    std::memcpy(&frame_time, &fs[3], sizeof(frame_time));
    steps = fs[2];
    h->frame_steps_hint = steps;
    return SPH_OK;
}

int sph_run_frame(sph_handle* h, float* frame_time_out, uint32_t* steps_out) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    float frame_time = 0.0f;
    uint32_t steps = 0;
    if (h->cfl && h->ready && h->use_graph && !h->prof && h->keys_valid && !h->keep_acc &&
        h->tune.sort_mode == kSortCellCount) {
        const int rc = run_frame_on_device(h, frame_time, steps);
        if (rc != SPH_OK) return rc;
        if (frame_time_out) *frame_time_out = frame_time;
        if (steps_out) *steps_out = steps;
        return SPH_OK;
    }
    if (!h->cfl) {
        // Fixed step: the frame's step count follows from the same float sum
        // on the host alone, and the steps run as one call (multi-step graphs).
        if (!h->ready) return fail(h, SPH_E_NOT_READY, "call sph_make_ready first");
        const float dt = h->params[0].dt;
        if (!(dt > 0.0f)) return fail(h, SPH_E_INVALID, "dt must be positive");
        while (frame_time < 0.0333333333) {
            frame_time += dt;
            // A step too short to move the float sum would never end the frame.
            if (++steps > (1u << 24)) return fail(h, SPH_E_INVALID, "dt is too small to reach a 1/30 s frame");
        }
        const int rc = sph_advance_steps(h, steps);
        if (rc != SPH_OK) return rc;
    } else {
        while (frame_time < 0.0333333333) {
            float dt = 0.0f;
            int rc = sph_advance_frame(h, &dt);
            if (rc != SPH_OK) return rc;
            frame_time += dt;
            ++steps;
        }
    }
    if (frame_time_out) *frame_time_out = frame_time;
    if (steps_out) *steps_out = steps;
    return SPH_OK;
}

uint32_t sph_num_particles(const sph_handle* h) { return h ? h->n : 0u; }

// .npy (format 1.0): magic, version, little-endian header length, then a
// Python dict literal padded with spaces to a 64-byte boundary ending in '\n'.
static bool write_npy_f32(const char* path, const float* data, uint32_t rows, uint32_t cols) {
    std::string dict = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + std::to_string(rows) + ", " +
                       std::to_string(cols) + "), }";
    const size_t pre = 10;
    size_t total = pre + dict.size() + 1;
    const size_t pad = (64 - total % 64) % 64;
    dict.append(pad, ' ');
    dict.push_back('\n');
    const uint16_t hlen = (uint16_t)dict.size();
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const unsigned char head[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    bool ok = std::fwrite(head, 1, 8, f) == 8;
    const unsigned char hl[2] = {(unsigned char)(hlen & 0xff), (unsigned char)(hlen >> 8)};
    ok = ok && std::fwrite(hl, 1, 2, f) == 2;
    ok = ok && std::fwrite(dict.data(), 1, dict.size(), f) == dict.size();
    ok = ok && std::fwrite(data, sizeof(float), (size_t)rows * cols, f) == (size_t)rows * cols;
    ok = (std::fclose(f) == 0) && ok;
    return ok;
}

static int fetch3(sph_handle* h, const float4* src, float* out) {
    if (!h || !out) return SPH_E_INVALID;
    if (!h->has_particles) return fail(h, SPH_E_NOT_READY, "no particles set");
    if (h->n == 0) return SPH_OK;
    const size_t cnt = 3 * (size_t)h->n;
    // Download through a page-locked buffer of the handle (a DMA copy at full
    // link rate), then one host copy into the caller's array.
    if (h->host_pinned_n < cnt) {
        if (h->host_pinned) (void)hipHostFree(h->host_pinned);
        h->host_pinned = nullptr;
        h->host_pinned_n = 0;
        const hipError_t ae = hipHostMalloc((void**)&h->host_pinned, cnt * sizeof(float), hipHostMallocDefault);
        if (ae != hipSuccess) {
            h->host_pinned = nullptr;
            return hip_fail(h, ae, "pinned host buffer");
        }
        h->host_pinned_n = cnt;
    }
    hipError_t e = launch_unpermute3(h->stream, h->n, src, h->id.p, h->staging.p);
    if (e == hipSuccess) e = hipMemcpyAsync(h->host_pinned, h->staging.p, cnt * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "particle download");
    const int rc = sync(h);
    if (rc != SPH_OK) return rc;
    std::memcpy(out, h->host_pinned, cnt * sizeof(float));
    return SPH_OK;
}

int sph_get_particles(sph_handle* h, float* xyz_out) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    return fetch3(h, h ? h->pos.p : nullptr, xyz_out);
}
int sph_get_velocity(sph_handle* h, float* vel_out) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    return fetch3(h, h ? h->vel.p : nullptr, vel_out);
}

int sph_get_density(sph_handle* h, float* rho_out) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !rho_out) return SPH_E_INVALID;
    if (!h->has_step) return fail(h, SPH_E_NOT_READY, "no step has run");
    // Density is in the last step's sorted order, which is the state order.
    hipError_t e = launch_unpermute1(h->stream, h->n, h->work.rho.p, h->id.p, h->staging.p);
    if (e == hipSuccess) e = hipMemcpyAsync(rho_out, h->staging.p, (size_t)h->n * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "density download");
    return sync(h);
}

int sph_snapshot_begin(sph_handle* h) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (!h->has_particles || h->n == 0) return fail(h, SPH_E_NOT_READY, "no particles set");
    const size_t cnt = 3 * (size_t)h->n;
    hipError_t e = hipSuccess;
    if (!h->copy_stream) e = hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess && !h->snap_ready) e = hipEventCreateWithFlags(&h->snap_ready, hipEventDisableTiming);
    if (e == hipSuccess && !h->snap_done) e = hipEventCreateWithFlags(&h->snap_done, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(h, e, "snapshot stream");
    if (h->snap_n < cnt) {
        // A resize waits for any copy still reading the old buffers.
        if (h->snap_pending) (void)hipEventSynchronize(h->snap_done);
        h->snap_pending = false;
        if (h->snap_host) (void)hipHostFree(h->snap_host);
        h->snap_host = nullptr;
        h->snap_n = 0;
        e = h->snap_d.ensure(cnt);
        if (e == hipSuccess) e = hipHostMalloc((void**)&h->snap_host, cnt * sizeof(float), hipHostMallocDefault);
        if (e != hipSuccess) {
            h->snap_host = nullptr;
            return hip_fail(h, e, "snapshot buffers");
        }
        h->snap_n = cnt;
    }
    // The previous snapshot's copy must have left snap_d before it is rewritten.
    if (h->snap_pending) e = hipStreamWaitEvent(h->stream, h->snap_done, 0);
    if (e == hipSuccess) e = launch_unpermute3(h->stream, h->n, h->pos.p, h->id.p, h->snap_d.p);
    if (e == hipSuccess) e = hipEventRecord(h->snap_ready, h->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(h->copy_stream, h->snap_ready, 0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h->snap_host, h->snap_d.p, cnt * sizeof(float), hipMemcpyDeviceToHost, h->copy_stream);
    if (e == hipSuccess) e = hipEventRecord(h->snap_done, h->copy_stream);
    if (e != hipSuccess) return hip_fail(h, e, "snapshot enqueue");
    h->snap_pending = true;
    h->snap_count = h->n;
    return SPH_OK;
}

int sph_snapshot_wait(sph_handle* h, float* xyz_out) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !xyz_out) return SPH_E_INVALID;
    if (!h->snap_pending) return fail(h, SPH_E_NOT_READY, "no snapshot in flight (call sph_snapshot_begin)");
    const hipError_t e = hipEventSynchronize(h->snap_done);
    if (e != hipSuccess) return hip_fail(h, e, "snapshot wait");
    h->snap_pending = false;
    if (h->snap_count != h->n)
        return fail(h, SPH_E_INVALID, "particle count changed since sph_snapshot_begin (" +
                                          std::to_string(h->snap_count) + " then, " + std::to_string(h->n) +
                                          " now); the snapshot is dropped");
    std::memcpy(xyz_out, h->snap_host, 3 * (size_t)h->snap_count * sizeof(float));
    return check_device_error(h);
}

int sph_export_frame(sph_handle* h, const char* path) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !path) return SPH_E_INVALID;
    if (!h->has_particles) return fail(h, SPH_E_NOT_READY, "no particles set");
    const size_t n = h->n;
    std::vector<float> x(3 * n), v(3 * n), xv(6 * n);
    int rc = sph_get_particles(h, x.data());
    if (rc == SPH_OK) rc = sph_get_velocity(h, v.data());
    if (rc != SPH_OK) return rc;
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            xv[6 * i + c] = x[3 * i + c];
            xv[6 * i + 3 + c] = v[3 * i + c];
        }
    if (!write_npy_f32(path, xv.data(), (uint32_t)n, 6u))
        return fail(h, SPH_E_INVALID, std::string("cannot write ") + path);
    return SPH_OK;
}

int sph_synchronize(sph_handle* h) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    return sync(h);
}

const char* sph_last_error(const sph_handle* h) { return h ? h->last_error.c_str() : t_last_error; }

int sph_set_use_graph(sph_handle* h, int on) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    h->use_graph = on != 0;
    if (!h->use_graph) invalidate_graph(h);
    return SPH_OK;
}

// ---- batched ----

int sph_batch_create(const sph_scene_desc* scenes, uint32_t nscenes, sph_handle** out) {
    if (!scenes || nscenes == 0 || !out) return SPH_E_INVALID;
    if (nscenes > 65535) return SPH_E_INVALID;
    sph_handle* h = nullptr;
    int rc = new_handle(&h);
    if (rc != SPH_OK) return rc;
    std::vector<sph_params> ps(nscenes);
    std::vector<uint32_t> ns(nscenes);
    for (uint32_t s = 0; s < nscenes; ++s) { ps[s] = scenes[s].params; ns[s] = scenes[s].num_particles; }
    h->batched = true;
    BudgetScope budget_scope(&h->mem);
    rc = configure(h, ps, ns);
    if (rc == SPH_OK) rc = ensure_buffers(h);
    if (rc != SPH_OK) {
        const std::string msg = h->last_error;
        delete h;
        return fail(nullptr, rc, msg);
    }
    *out = h;
    return SPH_OK;
}

uint32_t sph_num_scenes(const sph_handle* h) { return h ? (uint32_t)h->sc_h.size() : 0u; }

int sph_scene_offsets(const sph_handle* h, uint32_t* offsets) {
    if (!h || !offsets) return SPH_E_INVALID;
    std::copy(h->scene_off.begin(), h->scene_off.end(), offsets);
    return SPH_OK;
}

int sph_set_scene_params(sph_handle* h, uint32_t scene, const sph_params* p) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !p) return SPH_E_INVALID;
    if (scene >= h->sc_h.size()) return fail(h, SPH_E_INVALID, "scene index out of range");
    ++h->config_gen;
    const sph_params& o = h->params[scene];
    const bool same_geom = p->particle_radius == o.particle_radius && p->kernel_radius == o.kernel_radius &&
                           p->dt == o.dt && p->rest_density == o.rest_density &&
                           std::memcmp(p->box_min, o.box_min, sizeof(o.box_min)) == 0 &&
                           std::memcmp(p->box_max, o.box_max, sizeof(o.box_max)) == 0;
    if (!same_geom) return fail(h, SPH_E_INVALID, "box, radii, rest density and dt are fixed at create time");
    if (h->batched && p->cfl_number > 0.f)
        return fail(h, SPH_E_INVALID, "the CFL step is for single-scene handles (a batch shares dt)");
    SceneConst sc;
    std::string err;
    int rc = derive_scene(*p, sc, err, h->xsub);
    if (rc != SPH_OK) return fail(h, rc, err);
    sc.cell_offset = h->sc_h[scene].cell_offset;
    sc.p_begin = h->sc_h[scene].p_begin;
    sc.p_end = h->sc_h[scene].p_end;
    h->params[scene] = *p;
    h->sc_h[scene] = sc;
    hipError_t e = hipMemcpyAsync(h->sc_d.p + scene, &h->sc_h[scene], sizeof(SceneConst), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return hip_fail(h, e, "scene table upload");
}

// ---- device memory ----

int sph_set_device_memory_limit(sph_handle* h, uint64_t bytes) {
    if (!h) return SPH_E_INVALID;
    h->mem.limit = bytes;
    return SPH_OK;
}

int sph_device_memory_used(const sph_handle* h, uint64_t* bytes) {
    if (!h || !bytes) return SPH_E_INVALID;
    *bytes = h->mem.used;
    return SPH_OK;
}

// ---- measurement ----

int sph_profile_enable(sph_handle* h, int on) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    h->prof = on != 0;
    return SPH_OK;
}

int sph_profile_read(sph_handle* h, double* ms, uint32_t* steps) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "stream synchronize");
    for (size_t b = 0; b + 6 <= h->prof_used; b += 6) {
        // A step that launched no hash kernel reports 0 for that pass (its two
        // events are adjacent; their gap is event overhead, not work).
        const bool hashed = b / 6 < h->prof_hashed.size() && h->prof_hashed[b / 6];
        for (int p = hashed ? 0 : 1; p < SPH_NUM_PASSES; ++p) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, h->prof_events[b + p], h->prof_events[b + p + 1]) == hipSuccess)
                h->prof_ms[p] += t;
        }
        h->prof_steps++;
    }
    h->prof_used = 0;
    h->prof_hashed.clear();
    if (ms) for (int p = 0; p < SPH_NUM_PASSES; ++p) ms[p] = h->prof_ms[p];
    if (steps) *steps = h->prof_steps;
    for (int p = 0; p < SPH_NUM_PASSES; ++p) h->prof_ms[p] = 0;
    h->prof_steps = 0;
    return SPH_OK;
}

int sph_get_tuning(const sph_handle* h, int* cfg) {
    if (!h || !cfg) return SPH_E_INVALID;
    cfg[0] = h->tune.density_block;
    cfg[1] = h->tune.force_block;
    cfg[2] = h->tune.reorder_block;
    cfg[3] = h->tune.sort_items;
    cfg[4] = h->tune.persistent | (h->tune.force_staged << 1) | (h->tune.density_staged << 2);
    return SPH_OK;
}

static bool valid_block(int b) { return b == 64 || b == 128 || b == 256 || b == 512; }

int sph_set_tuning(sph_handle* h, const int* cfg) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !cfg) return SPH_E_INVALID;
    if (!valid_block(cfg[0]) || !valid_block(cfg[1]) || !valid_block(cfg[2]))
        return fail(h, SPH_E_INVALID, "block sizes must be 64, 128, 256 or 512");
    if (cfg[3] != 4 && cfg[3] != 8 && cfg[3] != 12 && cfg[3] != 16)
        return fail(h, SPH_E_INVALID, "sort items must be 4, 8, 12 or 16");
    if (cfg[4] < 0 || cfg[4] > 7)
        return fail(h, SPH_E_INVALID,
                    "launch form must be 0 to 7 (bit 0 persistent, bit 1 staged force, bit 2 staged density)");
    h->tune.density_block = cfg[0];
    h->tune.force_block = cfg[1];
    h->tune.reorder_block = cfg[2];
    h->tune.sort_items = cfg[3];
    h->tune.persistent = cfg[4] & 1;
    h->tune.force_staged = (cfg[4] >> 1) & 1;
    h->tune.density_staged = (cfg[4] >> 2) & 1;
    invalidate_graph(h);
    return SPH_OK;
}

int sph_set_sort_mode(sph_handle* h, int mode) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (mode != SPH_SORT_CELL_COUNT && mode != SPH_SORT_RADIX) return fail(h, SPH_E_INVALID, "unknown sort mode");
    if (mode == SPH_SORT_RADIX && h->passes > kMaxPasses) return fail(h, SPH_E_INVALID, "key space too large");
    h->tune.sort_mode = mode;
    invalidate_graph(h);
    if (h->n) {
        const int rc = alloc_step_bufs(h, h->work);
        if (rc != SPH_OK) return rc;
    }
    return SPH_OK;
}

int sph_get_sort_mode(const sph_handle* h) { return h ? h->tune.sort_mode : SPH_E_INVALID; }

int sph_set_cell_subdivision(sph_handle* h, int xsub) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (xsub != 0 && xsub != 1 && xsub != 2 && xsub != 4) return fail(h, SPH_E_INVALID, "x subdivision must be 0, 1, 2 or 4");
    h->xsub_override = xsub;
    if (!h->has_particles && !h->batched) return SPH_OK;  // applied when particles are set
    // Re-grid keeping the state (it stays in its current order; keys are rebuilt).
    int rc = configure(h, h->params, h->scene_n);
    if (rc != SPH_OK) return rc;
    h->keys_valid = false;
    rc = ensure_buffers(h);
    if (rc != SPH_OK) return rc;
    if (!h->has_particles) return SPH_OK;
    h->ready = false;
    return sph_make_ready(h);
}

// ---- test-only views ----

int sph_debug_constants(const sph_handle* h, uint32_t scene, float* out14, int* G4) {
    if (!h || !out14 || !G4) return SPH_E_INVALID;
    SceneConst sc;
    if (scene < h->sc_h.size()) {
        sc = h->sc_h[scene];
    } else if (scene == 0) {
        std::string err;
        if (derive_scene(h->params[0], sc, err, 1) != SPH_OK) return SPH_E_INVALID;
    } else {
        return SPH_E_INVALID;
    }
    const sph_params& p = h->params[scene < h->params.size() ? scene : 0];
    const float sp = 2.0f * p.particle_radius;
    const float mass = p.rest_density * (sp * sp * sp);
    const float v[14] = {sc.h, sc.h2, sc.inv_cs, mass, sc.cD, sc.cP, sc.cV,
                         sc.inv_rho0, sc.lo[0], sc.lo[1], sc.lo[2], sc.hi[0], sc.hi[1], sc.hi[2]};
    std::copy(v, v + 14, out14);
    for (int k = 0; k < 3; ++k) G4[k] = sc.G[k];
    G4[3] = (int)sc.xsubf;
    return SPH_OK;
}

int sph_debug_get_state(sph_handle* h, float* xyz, float* vel, uint32_t* ids) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (!h->has_particles) return fail(h, SPH_E_NOT_READY, "no particles set");
    const size_t n = h->n;
    std::vector<float4> tmp(n);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && xyz) {
        e = hipMemcpy(tmp.data(), h->pos.p, n * sizeof(float4), hipMemcpyDeviceToHost);
        for (size_t i = 0; i < n && e == hipSuccess; ++i) { xyz[3 * i] = tmp[i].x; xyz[3 * i + 1] = tmp[i].y; xyz[3 * i + 2] = tmp[i].z; }
    }
    if (e == hipSuccess && vel) {
        e = hipMemcpy(tmp.data(), h->vel.p, n * sizeof(float4), hipMemcpyDeviceToHost);
        for (size_t i = 0; i < n && e == hipSuccess; ++i) { vel[3 * i] = tmp[i].x; vel[3 * i + 1] = tmp[i].y; vel[3 * i + 2] = tmp[i].z; }
    }
    if (e == hipSuccess && ids) e = hipMemcpy(ids, h->id.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(h, e, "state download");
    return check_device_error(h);
}

int sph_debug_set_state(sph_handle* h, const float* xyz, const float* vel, const uint32_t* ids) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !xyz) return SPH_E_INVALID;
    if (ids) {  // ids index the original-order outputs: they must be a permutation of 0..n-1
        std::vector<uint8_t> seen(h->n, 0);
        for (uint32_t k = 0; k < h->n; ++k) {
            if (ids[k] >= h->n || seen[ids[k]]) return fail(h, SPH_E_INVALID, "ids must be a permutation of 0..n-1");
            seen[ids[k]] = 1;
        }
    }
    if (!h->has_particles) return fail(h, SPH_E_NOT_READY, "no particles set");
    return upload_state(h, xyz, vel, ids);
}

int sph_debug_keep_acc(sph_handle* h, int on) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    h->keep_acc = on != 0;
    invalidate_graph(h);
    if (h->keep_acc && h->n) {
        if (h->work.acc.ensure(h->n) != hipSuccess) return fail(h, SPH_E_OOM, "acc buffer");
    }
    return SPH_OK;
}

int sph_debug_step_outputs(sph_handle* h, uint32_t* keys_in, uint32_t* perm, uint32_t* sorted_keys,
                           float* rho, float* acc3) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h) return SPH_E_INVALID;
    if (!h->has_step) return fail(h, SPH_E_NOT_READY, "no step has run");
    int rc = sync(h);
    if (rc != SPH_OK) return rc;
    const size_t n = h->n;
    hipError_t e = hipSuccess;
    if (perm) e = hipMemcpy(perm, h->work.perm.p, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && sorted_keys) e = hipMemcpy(sorted_keys, h->work.skey.p, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && rho) e = hipMemcpy(rho, h->work.rho.p, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && acc3) {
        if (!h->work.acc.p) return fail(h, SPH_E_NOT_READY, "enable sph_debug_keep_acc before stepping");
        std::vector<float4> tmp(n);
        e = hipMemcpy(tmp.data(), h->work.acc.p, n * sizeof(float4), hipMemcpyDeviceToHost);
        for (size_t i = 0; i < n; ++i) { acc3[3 * i] = tmp[i].x; acc3[3 * i + 1] = tmp[i].y; acc3[3 * i + 2] = tmp[i].z; }
    }
    if (e == hipSuccess && keys_in) {
        // Unsorted keys of the last step's input: keys_in[perm[k]] = sorted_keys[k].
        std::vector<uint32_t> pm(n), sk(n);
        e = hipMemcpy(pm.data(), h->work.perm.p, n * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(sk.data(), h->work.skey.p, n * 4, hipMemcpyDeviceToHost);
        for (size_t k = 0; k < n && e == hipSuccess; ++k) keys_in[pm[k]] = sk[k];
    }
    return hip_fail(h, e, "debug download");
}

uint32_t sph_debug_num_cells(const sph_handle* h) { return h ? h->ncells : 0u; }

int sph_debug_cells(sph_handle* h, uint32_t* start, uint32_t* end) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !start || !end) return SPH_E_INVALID;
    if (!h->has_step) return fail(h, SPH_E_NOT_READY, "no step has run");
    int rc = sync(h);
    if (rc != SPH_OK) return rc;
    std::vector<uint32_t> c((size_t)h->ncells + 1);
    hipError_t e = hipMemcpy(c.data(), h->work.cstart.p, c.size() * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(h, e, "cell download");
    // Report (start, end) per cell; empty cells as (~0, 0).
    for (size_t i = 0; i < h->ncells; ++i) {
        const bool empty = c[i] == c[i + 1];
        start[i] = empty ? 0xffffffffu : c[i];
        end[i] = empty ? 0u : c[i + 1];
    }
    return SPH_OK;
}

int sph_debug_neighbors(sph_handle* h, uint32_t* offsets, uint32_t* list, uint64_t capacity, uint64_t* count) {
    BudgetScope budget_scope(h ? &h->mem : nullptr);
    if (!h || !offsets) return SPH_E_INVALID;
    if (!h->has_step) return fail(h, SPH_E_NOT_READY, "no step has run");
    const size_t n = h->n;
    DevBuf<uint32_t> cnt, off, lst;
    if (cnt.ensure(n + 1) != hipSuccess || off.ensure(n + 1) != hipSuccess) return fail(h, SPH_E_OOM, "neighbour buffers");
    StepArgs a = step_args(h, h->work, 256);
    hipError_t e = launch_neighbors(h->stream, 256, a.ntiles, a, nullptr, cnt.p, 0);
    std::vector<uint32_t> c(n);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(c.data(), cnt.p, n * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(h, e, "neighbour count");
    uint64_t total = 0;
    offsets[0] = 0;
    for (size_t i = 0; i < n; ++i) { total += c[i]; offsets[i + 1] = (uint32_t)total; }
    if (count) *count = total;
    if (!list) return SPH_OK;
    if (total > capacity) return fail(h, SPH_E_CAPACITY, "neighbour list capacity too small");
    if (lst.ensure(std::max<uint64_t>(total, 1)) != hipSuccess) return fail(h, SPH_E_OOM, "neighbour list");
    e = hipMemcpy(off.p, offsets, (n + 1) * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_neighbors(h->stream, 256, a.ntiles, a, off.p, lst.p, 1);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && total) e = hipMemcpy(list, lst.p, total * 4, hipMemcpyDeviceToHost);
    return hip_fail(h, e, "neighbour fill");
}

}  // extern "C"
