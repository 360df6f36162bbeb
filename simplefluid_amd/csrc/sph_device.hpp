// Shared device-side definitions of the gfx950 SPH engine.
//
// Data layout in HBM (N particles, C cells over all scenes):
//   pos[2], vel[2]  float4 x N   state order / sorted order (16 B aligned loads); in
//                                sorted order pos.w = p/rho^2 and vel.w = 1/rho
//                                (written by the density pass for the force pass)
//   id[2]           uint32 x N   original particle id, state / sorted order
//   key0, rank0     uint32 x N   cell key of the state positions and the slot taken
//                                in its cell count (written by integrate)
//   sort ping-pong  2 x (uint32 key + uint32 value) x N (radix mode)
//   count           uint32 x C+2T particles per cell, then per scan tile of 4096 cells
//                                (atomic, built by the integrate pass), then the
//                                exclusive scan of the tile sums
//   cstart          uint32 x C+1 lower-bound cell-start table: cstart[c] = #particles with key < c
//   rho             float  x N   density (sorted order)
//   mask            uint32 x 9N  neighbour bits, row-major [row][particle] (layout at
//                                mask_span below)
// All kernels are written for wave64 and 256 CUs in 8 XCDs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

// Per-scene constants, one table entry per scene (a single-scene handle has
// a table of one).  Kernels read a tile's entry through scene_at.
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
    float inv_csx;     // x cells are cs/xsub wide: xsub * inv_cs
    float xsubf;       // x cells per cs (1, 2 or 4)
    float hc2;         // (h * inv_cs)^2: the support radius squared in y/z cell units
    float wmargin;     // x-window safety margin in x cells
    float cfl;         // CFL number (0: fixed dt)
    float c0;          // sqrt(7 B / rho0), the EOS sound speed
    float dt_max;      // the parameter dt: the step, or its cap under CFL
    float rho0;        // rest density (the walls' half-space fluid)
    float inv_h;       // 1 / h
    float cW;          // 2 * 45 rho0 / h: the walls' pressure term
    float wmax[3];     // box_max (box_min is bmin)
    uint32_t pad[4];
};
static_assert(sizeof(SceneConst) == 192, "SceneConst layout");

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

// A tile's (scene, first particle) and its scene's constants.  Tile and
// scene are the same for the whole block, and neither table is written
// while a pass runs, so both are read as constant memory: with a uniform
// index the loads are scalar and the constants sit in SGPRs, instead of
// per-lane loads that fences between LDS phases make the compiler repeat.
typedef __attribute__((address_space(4))) const uint32_t ConstWord;
__device__ __forceinline__ uint2 tile_at(const uint2* tiles, uint32_t t) {
    ConstWord* p = (ConstWord*)(tiles + t);
    return make_uint2(p[0], p[1]);
}
__device__ __forceinline__ SceneConst scene_at(const SceneConst* scenes, uint32_t s) {
    ConstWord* p = (ConstWord*)(scenes + s);
    SceneConst c;
    uint32_t* d = reinterpret_cast<uint32_t*>(&c);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(SceneConst) / 4; ++i) d[i] = p[i];
    return c;
}

// Work over tiles.  A plain launch has one block per tile, in XCD-aware
// order.  A persistent launch (queue != null, gridDim >= 8) gives each XCD an
// equal contiguous share of the tiles; the XCD's blocks (blockIdx % 8) take
// tiles from that share through an atomic counter, so costly tiles do not
// hold up a fixed split.  The eight counters are zeroed before the launch.
template <class Body>
__device__ __forceinline__ void for_each_tile(uint32_t ntiles, uint32_t* __restrict__ queue, Body&& body) {
    if (!queue) {
        const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
        if (t < ntiles) body(t);
        return;
    }
    __shared__ uint32_t s_next;
    const uint32_t x = blockIdx.x & 7u;
    const uint32_t t_beg = (uint32_t)(((uint64_t)ntiles * x) >> 3);
    const uint32_t t_end = (uint32_t)(((uint64_t)ntiles * (x + 1u)) >> 3);
    for (;;) {
        if (threadIdx.x == 0) s_next = t_beg + atomicAdd(queue + x, 1u);
        __syncthreads();
        const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_next);  // the same in every lane
        __syncthreads();
        if (t >= t_end) break;
        body(t);
    }
}

// Neighbour rows.  Cells are cs = h(1 + 2^-10) in y and z and cs/xsub in x.
// The candidates of a particle lie in the 3x3 rows (dy, dz) of y/z cells
// around its own; within a row, only x cells inside the chord of the support
// sphere at that row's nearest y/z distance can hold neighbours.  A row's x
// cells are contiguous in sorted order, and cstart holds a lower bound for
// every cell (empty ones included), so a row is the single range
// [cstart[first x cell], cstart[last x cell + 1]).  The window is a superset
// of the neighbours (distances are under-estimated by 2^-10 cells and the
// chord is widened by wmargin); the predicate d^2 < h^2 decides membership.
struct RowGeom {
    float fx;            // x coordinate in x cells
    uint32_t row0;       // key of cell (x = 0, cy, cz)
    float qy[3], qz[3];  // squared y / z distance to rows -1, 0, +1 (cell units);
                         // 4 for a row outside the grid, which empties its window
};

__device__ __forceinline__ RowGeom row_geom(const SceneConst& sc, float x, float y, float z) {
    RowGeom g;
    g.fx = (x - sc.bmin[0]) * sc.inv_csx;
    const float fy = (y - sc.bmin[1]) * sc.inv_cs;
    const float fz = (z - sc.bmin[2]) * sc.inv_cs;
    const int cy = (int)fminf(fmaxf(fy, 0.0f), (float)(sc.G[1] - 1));
    const int cz = (int)fminf(fmaxf(fz, 0.0f), (float)(sc.G[2] - 1));
    const float m = 0.0009765625f;
    const float ry = fminf(fmaxf(fy - (float)cy, 0.0f), 1.0f);
    const float rz = fminf(fmaxf(fz - (float)cz, 0.0f), 1.0f);
    const float ly = fmaxf(ry - m, 0.0f), uy = fmaxf(1.0f - ry - m, 0.0f);
    const float lz = fmaxf(rz - m, 0.0f), uz = fmaxf(1.0f - rz - m, 0.0f);
    g.qy[0] = cy > 0 ? ly * ly : 4.0f;
    g.qy[1] = 0.0f;
    g.qy[2] = cy < sc.G[1] - 1 ? uy * uy : 4.0f;
    g.qz[0] = cz > 0 ? lz * lz : 4.0f;
    g.qz[1] = 0.0f;
    g.qz[2] = cz < sc.G[2] - 1 ? uz * uz : 4.0f;
    // A particle with a non-finite coordinate has no neighbours (every d^2
    // with it is NaN, which fails d^2 < h^2): all its windows are empty, so
    // a pile of such particles in one clamped cell costs nothing here.
    if (!(__builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z))) {
#pragma unroll
        for (int i = 0; i < 3; ++i) g.qy[i] = g.qz[i] = 4.0f;
    }
    g.row0 = sc.cell_offset + ((uint32_t)cz * (uint32_t)sc.G[1] + (uint32_t)cy) * (uint32_t)sc.G[0];
    return g;
}

// Raw buffer resource over n elements of T: loads past the end return 0.
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const T* p, uint32_t n) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * (uint32_t)sizeof(T)), 0x00020000);
}

// Window [lo, hi) of row (dy, dz); rcs is a buffer over cstart.
__device__ __forceinline__ void row_window(__amdgpu_buffer_rsrc_t rcs, const SceneConst& sc, const RowGeom& g,
                                           int dy, int dz, uint32_t& lo, uint32_t& hi) {
    lo = 0u;
    hi = 0u;
    const float w2 = sc.hc2 - (g.qy[dy + 1] + g.qz[dz + 1]);
    if (!(w2 > 0.0f)) return;
    const float wx = __builtin_amdgcn_sqrtf(w2) * sc.xsubf + sc.wmargin;
    const float gmax = (float)(sc.G[0] - 1);
    const uint32_t x0 = (uint32_t)(int)fminf(fmaxf(g.fx - wx, 0.0f), gmax);
    const uint32_t x1 = (uint32_t)(int)fminf(fmaxf(g.fx + wx, 0.0f), gmax);
    const uint32_t c = g.row0 + (uint32_t)(dz * sc.G[1] + dy) * (uint32_t)sc.G[0];
    lo = __builtin_amdgcn_raw_buffer_load_b32(rcs, (int)((c + x0) << 2), 0, 0);
    hi = __builtin_amdgcn_raw_buffer_load_b32(rcs, (int)((c + x1 + 1u) << 2), 0, 0);
}

// Row masks.  The density pass tests a row's candidates [lo, hi) four at a
// time and shifts one bit per slot into the row's mask (slots past hi shift
// in zeros), so candidate c of a row with T groups sits at bit 4T-1-c and
// only the last 32 slots survive.  Rows of more than 32 candidates re-test
// the ones before with the predicate.  row_neighbors visits a row's
// neighbours in ascending order: first the re-tested head, then the bits
// from the highest (the lowest candidate) down.
__device__ __forceinline__ uint32_t mask_span(uint32_t lo, uint32_t hi) { return (hi - lo + 3u) & ~3u; }

// Bit of candidate j of a row (>= 32 when j precedes the mask's slots).
__device__ __forceinline__ uint32_t mask_bit(uint32_t lo, uint32_t hi, uint32_t j) {
    return mask_span(lo, hi) - 1u - (j - lo);
}

template <class Retest, class F>
__device__ __forceinline__ void row_neighbors(uint32_t lo, uint32_t hi, uint32_t m, Retest&& retest, F&& f) {
    const uint32_t span = mask_span(lo, hi);
    const uint32_t jb = lo + span - 32u;  // candidate of bit 31 (wraps when span < 32)
    if (span > 32u)
        for (uint32_t j = lo; j < jb; ++j)
            if (retest(j)) f(j);
    // Bit 31 - i of m is candidate jb + i; reversed, the lowest set bit is
    // the next candidate in ascending order, and r & (r - 1) clears it.
    for (uint32_t r = __builtin_bitreverse32(m); r; r &= r - 1u) f(jb + (uint32_t)__builtin_ctz(r));
}

// The first candidate the mask covers: lo, or for a row of more than 32
// candidates the first of its last 32.
__device__ __forceinline__ uint32_t row_first(uint32_t lo, uint32_t hi) {
    const uint32_t span = mask_span(lo, hi);
    return span > 32u ? lo + span - 32u : lo;
}

// row_neighbors with the mask's visits given as offsets: g(i) for candidate
// row_first(lo, hi) + i (f(j) for the re-tested head, as before).  A caller
// reading the row from a staged copy then adds the row's offset once per
// row instead of once per neighbour.  Same neighbours, same order.
template <class Retest, class F, class G>
__device__ __forceinline__ void row_neighbors_from(uint32_t lo, uint32_t hi, uint32_t m, Retest&& retest, F&& f,
                                                   G&& g) {
    const uint32_t span = mask_span(lo, hi);
    const uint32_t j0 = row_first(lo, hi);
    for (uint32_t j = lo; j < j0; ++j)
        if (retest(j)) f(j);
    // Reversed, bit i is candidate lo + span - 32 + i; a row of fewer than
    // 32 slots drops the 32 - span empty low bits, so bit i is lo + i.
    uint32_t r = __builtin_bitreverse32(m);
    if (span - 1u < 31u) r >>= 32u - span;
    for (; r; r &= r - 1u) g((uint32_t)__builtin_ctz(r));
}

// Wave-wide minimum / maximum of one value per lane (all 64 lanes active):
// DPP within each row of 16 lanes, then row broadcasts fold row 0 into 1 and
// row 2 into 3 (lane 15 and 47 to the odd rows), then rows 0-1 into 2-3
// (lane 31 to rows 2 and 3), so lane 63 holds the result.  The operation's
// identity is the "old" operand: rows the row mask leaves out take it and
// keep their own value, and the compiler folds each move into the min / max
// as its DPP source.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_min(uint32_t v) {
    return min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_max(uint32_t v) {
    return max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false));
}
__device__ __forceinline__ uint32_t wave_umin(uint32_t v) {
    v = dpp_min<0xB1, 0xF>(v);   // quad_perm 1,0,3,2
    v = dpp_min<0x4E, 0xF>(v);   // quad_perm 2,3,0,1
    v = dpp_min<0x141, 0xF>(v);  // row_half_mirror
    v = dpp_min<0x140, 0xF>(v);  // row_mirror
    v = dpp_min<0x142, 0xA>(v);  // row_bcast:15
    v = dpp_min<0x143, 0xC>(v);  // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ __forceinline__ uint32_t wave_umax(uint32_t v) {
    v = dpp_max<0xB1, 0xF>(v);
    v = dpp_max<0x4E, 0xF>(v);
    v = dpp_max<0x141, 0xF>(v);
    v = dpp_max<0x140, 0xF>(v);
    v = dpp_max<0x142, 0xA>(v);
    v = dpp_max<0x143, 0xC>(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Direct copy of 16 bytes at byte offset off of a buffer into LDS: lane i
// writes dst[i] (the wave's active lanes; dst is the same for all lanes).
// No registers hold the data; lds_copy_wait() waits for the wave's copies.
typedef __attribute__((address_space(3))) void lds_void;
__device__ __forceinline__ void lds_copy16(__amdgpu_buffer_rsrc_t r, float4* dst, uint32_t off) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void*)dst, 16, (int)off, 0, 0, 0);
}
__device__ __forceinline__ void lds_copy_wait() {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt unconstrained
    __builtin_amdgcn_wave_barrier();
}

// LDS order between a wave's own stores and its later loads of other
// lanes' entries (and back).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Components go through named vector members: reading elements of the
// builtin's result by subscript inside __builtin_bit_cast loses all but one
// dword of the load with this compiler.
__device__ __forceinline__ float4 load4(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)byte_off, 0, 0);
    float4 f;
    f.x = __uint_as_float(v.x);
    f.y = __uint_as_float(v.y);
    f.z = __uint_as_float(v.z);
    f.w = __uint_as_float(v.w);
    return f;
}

__device__ __forceinline__ float2 load2(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)byte_off, 0, 0);
    float2 f;
    f.x = __uint_as_float(v.x);
    f.y = __uint_as_float(v.y);
    return f;
}

__device__ __forceinline__ uint32_t cell_key(const SceneConst& sc, float x, float y, float z) {
    const int cx = cell_coord(x, sc.bmin[0], sc.inv_csx, sc.G[0]);
    const int cy = cell_coord(y, sc.bmin[1], sc.inv_cs, sc.G[1]);
    const int cz = cell_coord(z, sc.bmin[2], sc.inv_cs, sc.G[2]);
    return sc.cell_offset + ((uint32_t)cz * (uint32_t)sc.G[1] + (uint32_t)cy) * (uint32_t)sc.G[0] + (uint32_t)cx;
}

// Atomic add of each active lane's 1 to arr[idx], one atomic per run of
// equal idx in consecutive lanes.  Returns the slot this lane took (the old
// value plus its place in the run).  All lanes of the wave must call it;
// `active` marks the lanes that hold a particle.
__device__ __forceinline__ uint32_t run_add(uint32_t* __restrict__ arr, uint32_t idx, bool active) {
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t kk = active ? idx : 0xffffffffu;
    const uint32_t prev = __shfl_up(kk, 1, 64);
    const bool start = active && (lane == 0 || prev != kk);
    const uint64_t starts = __ballot(start);
    const uint64_t act = __ballot(active);
    // Leader of this lane's run: the highest run start at or below the lane.
    const uint64_t le = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const int leader = le ? 63 - __clzll(le) : lane;
    // Run length: up to the next start or the first inactive lane above.
    const uint64_t above = (starts | ~act) & ~((2ull << lane) - 1ull) & (lane == 63 ? 0ull : ~0ull);
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    uint32_t base = 0;
    if (start) base = atomicAdd(&arr[kk], (uint32_t)(next - lane));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)(lane - leader);
}

// Cells per tile of the cell-start scan (4096): the integrate pass also sums
// particles per tile, so every scan block finds its prefix without waiting.
constexpr uint32_t kScanTileShift = 12;

// Per-cell count (lanes follow sorted order, so most neighbours share a
// cell) and the per-tile sums that follow the C cell counts.  Returns the
// slot this lane took in its cell.
__device__ __forceinline__ uint32_t count_slot(uint32_t* __restrict__ count, uint32_t* __restrict__ tsum,
                                               uint32_t key, bool active) {
    const uint32_t slot = run_add(count, key, active);
    (void)run_add(tsum, key >> kScanTileShift, active);
    return slot;
}

// Adds the wave's largest |v|^2 to arr[idx] (float bits; atomic max on the
// bit patterns orders non-negative floats).  NaN counts as 0.  All lanes of
// the wave must call it, with idx uniform.
__device__ __forceinline__ void wave_max_v2(uint32_t* __restrict__ arr, uint32_t idx, float vx, float vy, float vz,
                                            bool active) {
    float v2 = active ? fmaxf(fmaf(vz, vz, fmaf(vy, vy, vx * vx)), 0.0f) : 0.0f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v2 = fmaxf(v2, __shfl_xor(v2, o, 64));
    if ((threadIdx.x & 63u) == 0u) atomicMax(arr + idx, __float_as_uint(v2));
}

// Walls.  Each box face bounds a half-space of fluid at rest density; a
// particle at distance u h (u < 1) from a face gets that fluid's poly6
// density rho0 wall_rho(u) and its spiky pressure term, taken at the
// particle's own p/rho^2, cW (p/rho^2) wall_j(u) along the face's inward
// normal:
//   wall_rho(u) = (1-u)^5 (128 + 325u + 345u^2 + 175u^3 + 35u^4) / 256
//               = (315/256) int_u^1 (1 - t^2)^4 dt          (rho0/2 at u = 0)
//   wall_j(u)   = (1-u)^4 (1 + 4u) / 30 = int_u^1 (1-s)^2 (s^2 - u^2) ds
// with the derivatives the reverse pass needs.
__device__ __forceinline__ float wall_rho(float u) {
    if (!(u < 1.0f)) return 0.0f;
    u = fmaxf(u, 0.0f);
    const float w = 1.0f - u;
    const float w2 = w * w;
    const float w5 = (w2 * w2) * w;
    const float poly = 128.0f + u * (325.0f + u * (345.0f + u * (175.0f + u * 35.0f)));
    return (w5 * poly) * 0.00390625f;
}
__device__ __forceinline__ float wall_j(float u) {
    if (!(u < 1.0f)) return 0.0f;
    u = fmaxf(u, 0.0f);
    const float w = 1.0f - u;
    const float w2 = w * w;
    return ((w2 * w2) * (1.0f + 4.0f * u)) * (float)(1.0 / 30.0);
}
__device__ __forceinline__ float wall_rho_du(float u) {
    if (!(u < 1.0f)) return 0.0f;
    u = fmaxf(u, 0.0f);
    const float s = 1.0f - u * u;
    const float s2 = s * s;
    return -1.23046875f * (s2 * s2);
}
__device__ __forceinline__ float wall_j_du(float u) {
    if (!(u < 1.0f)) return 0.0f;
    u = fmaxf(u, 0.0f);
    const float w = 1.0f - u;
    return -(float)(2.0 / 3.0) * u * ((w * w) * w);
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
