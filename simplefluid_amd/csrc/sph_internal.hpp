// Host-side declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_device.hpp"

namespace sph {

constexpr int kMaxPasses = 4;   // 8-bit digits, keys up to 32 bits

constexpr int kSortCellCount = 0;   // one-digit (counting) radix sort, default
constexpr int kSortRadix = 1;       // multi-pass 8-bit LSD radix sort

struct SortPlan {
    uint32_t n = 0;
    uint32_t ncells = 0;
    int mode = kSortCellCount;
    int passes = 1;                   // radix: 8-bit digits
    int items = 8;                    // radix: keys per thread of a digit-pass tile
    int reorder_block = 256;
    const uint32_t* key0 = nullptr;   // unsorted keys (state order)
    const uint32_t* rank0 = nullptr;  // slot each particle took in its cell's count
    uint32_t* count = nullptr;        // C cell counts (zeroed by the scan) + per-tile sums
    uint32_t* cstart = nullptr;       // C + 1, written by the scan
    uint32_t* tile_pre = nullptr;     // per scan tile: the prefix this cstart holds in all its cells
                                      // when the tile was empty at its last write, else 0xffffffff
    uint32_t* tmp = nullptr;          // cell-count: slot -> particle
    uint32_t* kbuf[2] = {nullptr, nullptr};  // radix ping-pong
    uint32_t* vbuf[2] = {nullptr, nullptr};
    uint32_t* skey = nullptr;         // sorted keys
    uint32_t* perm = nullptr;         // sorted index -> state index
    uint32_t* hist = nullptr;         // radix: kMaxPasses x 256, zero on entry
    uint32_t* status = nullptr;       // radix: passes x tiles x 256
    uint32_t* tickets = nullptr;      // radix: kMaxPasses
    uint32_t* err = nullptr;
    uint32_t* queue = nullptr;        // kQueueWords tile counters, zeroed by the scan
    const uint32_t* halt = nullptr;   // device frame loop: non-zero word skips the step (cell-count mode)
    const float4* pos_in = nullptr;
    const float4* vel_in = nullptr;
    const uint32_t* id_in = nullptr;
    float4* pos_out = nullptr;
    float* soa_out = nullptr;         // sorted x, y, z planes (3 N floats) for the gathered density
                                      // pass; null when the density pass is staged (it reads pos_out)
    float4* vel_out = nullptr;
    uint32_t* id_out = nullptr;
};

uint32_t sort_tiles(uint32_t n, int items);
// Tiles of 4096 cells.  count holds C cell counts, then T = scan_tiles(C)
// per-tile sums, then their exclusive scan (T words).
uint32_t scan_tiles(uint32_t ncells);
int sort_key_passes(uint32_t ncells);
// Cell-start table plus the sort proper (keys and permutation) ...
hipError_t launch_sort_keys(hipStream_t st, const SortPlan& sp);
// ... then the gather of the state into sorted order.
hipError_t launch_sort_gather(hipStream_t st, const SortPlan& sp);

// Arguments of the per-particle passes.  Work is split in tiles of blockDim
// particles that never straddle two scenes (tiles[t] = (scene, first)).
struct StepArgs {
    float4* pos_s;           // sorted positions; w <- p/rho^2 (density pass)
    const float* soa;        // the same positions as x, y, z planes (3 N floats)
    float4* vel_s;           // sorted velocities; w <- 1/rho (density pass)
    float4* pos_o;           // state out (integrate)
    float4* vel_o;
    const uint32_t* id_s;
    uint32_t* id_o;
    uint32_t* key0;          // next step's unsorted keys
    uint32_t* rank0;         // next step's slot in its cell count
    uint32_t* count;         // next step's per-cell counts, then per-scan-tile sums (atomic)
    const uint32_t* cstart;  // cell-start table of this step
    float* rho;
    uint32_t* mask;
    float4* acc;             // optional: acceleration (sorted order)
    const SceneConst* scenes;
    const uint2* tiles;
    uint32_t ntiles;
    uint32_t n;
    uint32_t ncells;
    uint32_t* queue;         // persistent launch: 8 tile counters (zeroed by the scan), else null
    uint32_t* vmax2;         // CFL: per-scene max |v|^2 (float bits) of the new state, else null
    const uint32_t* halt;    // device frame loop: non-zero word skips the step, else null
};

// Tile counters of the persistent density and force launches (8 each),
// zeroed by the cell-start scan of every step.
constexpr int kQueueWords = 16;

hipError_t launch_hash(hipStream_t st, int block, uint32_t grid, const float4* pos, const float4* vel,
                       uint32_t* key0, uint32_t* rank0, uint32_t* count, uint32_t ncells, uint32_t* vmax2,
                       const SceneConst* scenes, const uint2* tiles, uint32_t ntiles);
// CFL step of every scene from its max |v|^2 (scenes with cfl = 0 keep dt);
// clears vmax2 and stores scene 0's step in *dt_out.  With a frame state
// (below) the kernel does nothing once the frame is halted, and otherwise
// adds scene 0's step to the frame time.
hipError_t launch_cfl(hipStream_t st, uint32_t nscenes, SceneConst* scenes, uint32_t* vmax2, float* dt_out,
                      uint32_t* frame);
// Device frame loop state, 4 words: halt (steps are skipped), pending (the
// frame reached 1/30 s in the current step), steps taken, frame time (float
// bits).  reset zeroes them; gate, run first in each step, turns pending
// into halt, so the step that completes the frame runs to its end.
constexpr int kFrameWords = 4;
hipError_t launch_frame_reset(hipStream_t st, uint32_t* frame);
hipError_t launch_frame_gate(hipStream_t st, uint32_t* frame);
// staged: the density pass reads each row's candidates from LDS (same results).
hipError_t launch_density(hipStream_t st, int block, uint32_t grid, const StepArgs& a, int staged = 0);
// staged: the force pass reads each row's candidates from LDS (same results).
hipError_t launch_force(hipStream_t st, int block, uint32_t grid, const StepArgs& a, int staged = 0);
// Resident blocks per CU of the force kernel at this block size and form
// (the runtime's occupancy figure; 0 if it cannot say).
int force_blocks_per_cu(int block, int staged);
hipError_t launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t value, size_t count);
// Debug: neighbour CSR as the force pass sees it.
hipError_t launch_neighbors(hipStream_t st, int block, uint32_t grid, const StepArgs& a,
                            const uint32_t* offsets, uint32_t* counts_or_list, int fill);
// Original-order gather for get_*: out[3*id[k]+c] = src[k].c
hipError_t launch_unpermute3(hipStream_t st, uint32_t n, const float4* src, const uint32_t* id,
                             float* out);
hipError_t launch_unpermute1(hipStream_t st, uint32_t n, const float* src, const uint32_t* id,
                             float* out);
// dst[k] = (src[3k], src[3k+1], src[3k+2], 0), or zero when src is null.
hipError_t launch_pack3(hipStream_t st, uint32_t n, const float* src, float4* dst);

// Reverse pass of one step over a taped step (sorted order of that step).
struct AdjArgs {
    const float4* pos_s;
    const float4* vel_s;
    const float4* acc;
    const float* rho;
    const uint32_t* mask;
    const uint32_t* cstart;
    const uint32_t* perm;
    float4* xbar;            // in: adjoint of the step output; out: of its input
    float4* vbar;
    float4* xs;
    float4* vs;
    float4* abar;
    float* rhobar;
    double2* tile_acc;       // per tile: running (dL/dB, dL/dmu) sums over the reverse steps
    const SceneConst* scenes;
    const uint2* tiles;
    uint32_t ntiles;
    uint32_t n;
    uint32_t ncells;
};

hipError_t launch_adj_integrate(hipStream_t st, int block, uint32_t grid, const AdjArgs& a);
// staged: each row's candidates read from LDS (same results; whole waves only).
hipError_t launch_adj_force(hipStream_t st, int block, uint32_t grid, const AdjArgs& a, int staged = 0);
hipError_t launch_adj_density(hipStream_t st, int block, uint32_t grid, const AdjArgs& a, int staged = 0);
// Per-scene totals of the tile accumulators in a fixed order: scene s sums
// tiles [tile_first[s], tile_first[s + 1]) into acc_a[s], acc_b[s].
hipError_t launch_scene_tile_sums(hipStream_t st, uint32_t nscenes, const uint32_t* tile_first,
                                  const double2* tile_acc, double* acc_a, double* acc_b);
// out[k] = (g[3 id[k]], g[3 id[k]+1], g[3 id[k]+2], 0)
hipError_t launch_gather3(hipStream_t st, uint32_t n, const float* g, const uint32_t* id, float4* out);

}  // namespace sph
