// Engine handle shared by the C-ABI translation units.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "sph_engine.h"
#include "sph_internal.hpp"

namespace sph {

// Device bytes a handle's buffers hold, against an optional cap.
struct MemBudget {
    uint64_t limit = 0;     // 0: no cap
    uint64_t used = 0;
    uint64_t refused = 0;   // size of the last allocation the cap refused
};

// The budget that buffers allocated on this thread are charged to: each C-ABI
// entry point sets it to its handle's for the duration of the call (a handle
// is used by one thread at a time).  A buffer remembers the budget it was
// charged to and credits it back when freed.
extern thread_local MemBudget* t_budget;

struct BudgetScope {
    MemBudget* prev;
    explicit BudgetScope(MemBudget* b) : prev(t_budget) { t_budget = b; }
    ~BudgetScope() { t_budget = prev; }
    BudgetScope(const BudgetScope&) = delete;
    BudgetScope& operator=(const BudgetScope&) = delete;
};

// Device buffer owned by the handle.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t count = 0;
    MemBudget* owner = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;             // owns its allocation
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            (void)hipFree(p);
            if (owner) owner->used -= (uint64_t)count * sizeof(T);
        }
        p = nullptr;
        count = 0;
        owner = nullptr;
    }
    hipError_t ensure(size_t n) {
        if (n <= count && p) return hipSuccess;
        release();
        if (n == 0) return hipSuccess;
        const uint64_t bytes = (uint64_t)n * sizeof(T);
        MemBudget* b = t_budget;
        if (b && b->limit && b->used + bytes > b->limit) {
            b->refused = bytes;
            return hipErrorOutOfMemory;
        }
        hipError_t e = hipMalloc((void**)&p, bytes);
        if (e != hipSuccess) { p = nullptr; return e; }
        count = n;
        owner = b;
        if (b) b->used += bytes;
        return hipSuccess;
    }
};

// One step's intermediate buffers; the rollout keeps one per step of a
// checkpoint segment (its tape), the plain step uses slot 0.
struct StepBufs {
    DevBuf<float4> pos_s, vel_s;   // sorted input state
    DevBuf<float> soa;             // sorted positions as x, y, z planes (density candidates)
    DevBuf<uint32_t> id_s;
    DevBuf<uint32_t> skey, perm;   // sorted keys, sorted index -> state index
    DevBuf<uint32_t> cstart;       // C + 1 lower-bound cell-start table
    DevBuf<uint32_t> tile_pre;     // per scan tile of cstart: prefix held by an empty tile, or ~0u
    DevBuf<uint32_t> tmp;          // cell-count sort slots
    DevBuf<uint32_t> kbuf[2], vbuf[2];  // radix ping-pong (radix mode only)
    DevBuf<float> rho;
    DevBuf<uint32_t> mask;
    DevBuf<float4> acc;
};

struct Tuning {
    int density_block = 256;
    int force_block = 256;
    int reorder_block = 256;
    int sort_items = 8;
    int persistent = 0;   // density and force: 0 one block per tile, 1 persistent blocks over tile queues
    int force_staged = 0; // force: 1 each row's candidates staged in LDS, 0 gathered from global memory
    int density_staged = 0;  // density: the same choice
    int sort_mode = kSortCellCount;
};

// Differentiable rollout: a checkpoint of the state every `every` steps and
// a tape of one segment's step intermediates, rebuilt segment by segment in
// the reverse pass.
struct RolloutState {
    uint32_t steps = 0;
    uint32_t every = 1;
    uint32_t nseg = 0;
    bool valid = false;
    uint32_t n = 0;                                // particles of the rollout
    uint64_t gen = 0;                              // handle configuration it was run on
    bool last_taped = false;                       // the tape holds the last segment (forward recorded it)
    std::vector<std::unique_ptr<StepBufs>> tape;   // `every` slots
    DevBuf<float4> ck_pos, ck_vel;                 // (nseg + 1) x n; slot nseg = final state
    DevBuf<uint32_t> ck_id;
    DevBuf<float4> xbar, vbar;                     // adjoint of the state (state order)
    DevBuf<float4> xs, vs, abar;                   // per-step adjoint temporaries (sorted order)
    DevBuf<float> rhobar;
    DevBuf<double> gB, gmu;                        // per-scene totals
    DevBuf<double2> tile_acc;                      // per-tile running (dL/dB, dL/dmu)
    DevBuf<uint32_t> tile_first;                   // first tile of each scene (nscenes + 1)
    DevBuf<float> gbuf;                            // 3n staging for gradients
};

}  // namespace sph

struct sph_handle {
    // Declared first so that it outlives every buffer charged to it.
    sph::MemBudget mem;
    // Configuration.
    std::vector<sph_params> params;        // per scene
    std::vector<uint32_t> scene_n;         // particles per scene
    std::vector<uint32_t> scene_off;       // nscenes + 1
    std::vector<sph::SceneConst> sc_h;
    bool batched = false;
    uint32_t n = 0;
    uint32_t ncells = 0;
    int passes = 1;
    int xsub = 1;                          // x cells per y/z cell size
    int xsub_override = 0;                 // 0 = automatic
    bool ready = false;
    bool has_particles = false;
    bool has_step = false;
    bool keys_valid = false;
    bool keep_acc = false;
    int cu_count = 256;
    std::string arch;
    sph::Tuning tune;

    hipStream_t stream = nullptr;
    // State (state order) and per-step buffers.
    sph::DevBuf<float4> pos, vel;
    sph::DevBuf<uint32_t> id, key0, rank0, count;
    sph::StepBufs work;
    sph::DevBuf<uint32_t> hist, status, tickets, err, queue;
    bool cfl = false;                      // adaptive step (single scene)
    sph::DevBuf<uint32_t> vmax2;           // per-scene max |v|^2 bits for the next step
    sph::DevBuf<float> dt_dev;             // step length of the last step (scene 0)
    sph::DevBuf<sph::SceneConst> sc_d;
    std::map<int, sph::DevBuf<uint2>> tiles;
    std::map<int, uint32_t> ntiles;
    sph::DevBuf<float> staging;
    float* host_pinned = nullptr;          // 3n floats of page-locked host memory for downloads
    size_t host_pinned_n = 0;
    // Asynchronous frame snapshot: positions un-permuted on the step stream
    // into snap_d, copied to snap_host on copy_stream while steps continue.
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_ready = nullptr, snap_done = nullptr;
    sph::DevBuf<float> snap_d;
    float* snap_host = nullptr;
    size_t snap_n = 0;                     // floats snap_d / snap_host hold
    uint32_t snap_count = 0;               // particles when the pending snapshot began
    bool snap_pending = false;

    // Graph of one step, and of kGraphSteps steps in a row (replayed while at
    // least that many steps remain: one launch instead of several).
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec_multi = nullptr;
    hipGraph_t graph_multi = nullptr;
    // Device frame loop (CFL): a step graph whose kernels skip once the
    // frame state says the frame is complete; frame holds that state.
    hipGraphExec_t gexec_frame = nullptr;
    hipGraph_t graph_frame = nullptr;
    sph::DevBuf<uint32_t> frame;
    uint32_t frame_steps_hint = 0;         // steps of the last device-loop frame

    // Profiling.
    bool prof = false;
    std::vector<hipEvent_t> prof_events;   // 6 per recorded step
    std::vector<uint8_t> prof_hashed;      // per recorded step: the hash pass ran
    size_t prof_used = 0;
    double prof_ms[SPH_NUM_PASSES] = {0, 0, 0, 0, 0};
    uint32_t prof_steps = 0;

    // Bumped by every change of particles, parameters or grid: a rollout
    // recorded under another configuration cannot be reversed.
    uint64_t config_gen = 0;
    std::unique_ptr<sph::RolloutState> rollout;
    std::string last_error;

    ~sph_handle();
};

namespace sph {

int fail(sph_handle* h, int code, const std::string& msg);
int hip_fail(sph_handle* h, hipError_t e, const char* what);
int derive_scene(const sph_params& p, SceneConst& sc, std::string& err, int xsub);
// The same checks and constants without building a message (allocates nothing).
int derive_scene_noalloc(const sph_params& p, SceneConst& sc, int xsub);
// Records the failure message of a call that has no handle (sph_last_error(NULL)).
void set_thread_error(const char* msg);
int alloc_step_bufs(sph_handle* h, StepBufs& w);
void invalidate_graph(sph_handle* h);
uint2* tiles_for(sph_handle* h, int block, uint32_t& ntiles);
// Grid of a density or force launch; *queue receives the tile counters when
// the launch is persistent, else null.
// blocks_per_cu: the kernel's resident blocks per CU (0: 2048 lanes' worth).
uint32_t grid_for(sph_handle* h, int block, uint32_t ntiles, uint32_t** queue, int blocks_per_cu = 0);
StepArgs step_args(sph_handle* h, StepBufs& w, int block);
// Enqueue one step reading/writing the state buffers with intermediates in w.
// frame: device frame state (every kernel of the step skips once it is
// halted; cell-count sort only), or null.
int enqueue_step(sph_handle* h, StepBufs& w, bool record_prof, uint32_t* frame = nullptr);
int check_device_error(sph_handle* h);
int sync(sph_handle* h);

}  // namespace sph
