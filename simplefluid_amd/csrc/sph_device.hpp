// Shared device-side definitions of the gfx950 SPH engine.
//
// Data layout in HBM (N particles, C cells over all scenes):
//   pos[2], vel[2]  float4 x N   state order / sorted order (w unused, 16 B aligned loads)
//   id[2]           uint32 x N   original particle id, state / sorted order
//   key0            uint32 x N   cell key of the state positions (written by integrate)
//   sort ping-pong  2 x (uint32 key + uint32 value) x N
//   count           uint32 x C   particles per cell (atomic, built by the integrate pass)
//   cstart          uint32 x C+1 lower-bound cell-start table: cstart[c] = #particles with key < c
//   rho             float  x N   density (sorted order)
//   aux             float2 x N   (p/rho^2, 1/rho) read by the force pass
//   mask            uint32 x 9N  neighbour bits, row-major [row][particle]: bit b of row r
//                                marks candidate lo_r + b as a neighbour (b < 32)
// All kernels are written for wave64 and 256 CUs in 8 XCDs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

// Per-scene constants.  The single-scene path passes one by value (kernel
// arguments live in SGPRs); the batched path indexes a table per lane.
struct SceneConst {
    float bmin[3];
    float inv_cs;      // 1 / (h (1 + 2^-10)), the cell size's reciprocal
    float lo[3];       // clamp bounds: box_min + r
    float h;
    float hi[3];       // box_max - r
    float h2;
    int   G[3];        // grid cells per axis
    uint32_t cell_offset;  // first cell key of this scene
    float cD;          // m * 315/(64 pi h^9)
    float cP;          // m * 45/(pi h^6)
    float cV;          // mu * m * 45/(pi h^6)
    float inv_rho0;
    float B;
    float attr;        // attractive-pressure ratio (0 when off)
    float e;           // restitution
    float dt;
    float g[3];
    int   attractive;
    uint32_t p_begin;  // particle range of this scene (state and sorted order)
    uint32_t p_end;
    float mu;
    float pad;
};
static_assert(sizeof(SceneConst) == 128, "SceneConst layout");

constexpr int kRows = 9;          // 3x3 (dz, dy) rows of three x-adjacent cells
constexpr int kMaskBits = 32;     // candidates per row tracked in the mask

__device__ __forceinline__ int cell_coord(float x, float bmin, float inv_cs, int G) {
    float f = (x - bmin) * inv_cs;
    f = fminf(fmaxf(f, 0.0f), (float)(G - 1));
    return (int)f;
}

__device__ __forceinline__ float dist2(float dx, float dy, float dz) {
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Workgroups go round-robin over the 8 XCDs; give each XCD one contiguous run
// of logical blocks so the neighbour data its blocks share stays in its L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblocks) {
    const uint32_t X = 8;
    const uint32_t q = nblocks / X, r = nblocks % X;
    const uint32_t x = bid % X, i = bid / X;
    return x * q + (x < r ? x : r) + i;
}

// Sorted-index range covering cells cx-1..cx+1 of grid row (y, z).  Cells of
// a row are contiguous in sorted order and cstart holds a lower bound for
// every cell (empty ones included), so the row is [cstart[first],
// cstart[last + 1]).  Rows outside the grid are empty (lo = hi = 0).
__device__ __forceinline__ void row_range(const uint32_t* __restrict__ cstart, const SceneConst& sc,
                                          int cx, int y, int z, uint32_t& lo, uint32_t& hi) {
    lo = 0u;
    hi = 0u;
    if (y < 0 || y >= sc.G[1] || z < 0 || z >= sc.G[2]) return;
    const int x0 = cx > 0 ? cx - 1 : 0;
    const int x1 = cx + 1 < sc.G[0] ? cx + 1 : sc.G[0] - 1;
    const uint32_t base = sc.cell_offset + ((uint32_t)z * (uint32_t)sc.G[1] + (uint32_t)y) * (uint32_t)sc.G[0];
    lo = cstart[base + (uint32_t)x0];
    hi = cstart[base + (uint32_t)x1 + 1u];
}

__device__ __forceinline__ uint32_t cell_key(const SceneConst& sc, float x, float y, float z) {
    const int cx = cell_coord(x, sc.bmin[0], sc.inv_cs, sc.G[0]);
    const int cy = cell_coord(y, sc.bmin[1], sc.inv_cs, sc.G[1]);
    const int cz = cell_coord(z, sc.bmin[2], sc.inv_cs, sc.G[2]);
    return sc.cell_offset + ((uint32_t)cz * (uint32_t)sc.G[1] + (uint32_t)cy) * (uint32_t)sc.G[0] + (uint32_t)cx;
}

// Tait EOS, exponent 7 by the chain x^2, x^4, (x^4 x^2) x; p < 0 is clamped
// to 0 or scaled by the attractive ratio.
__device__ __forceinline__ float eos(const SceneConst& sc, float rho) {
    const float x = rho * sc.inv_rho0;
    const float x2 = x * x;
    const float x4 = x2 * x2;
    const float x7 = (x4 * x2) * x;
    float p = sc.B * (x7 - 1.0f);
    if (p < 0.0f) p = sc.attractive ? p * sc.attr : 0.0f;
    return p;
}

}  // namespace sph
