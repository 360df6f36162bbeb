// Per-particle passes of one SPH step on gfx950:
//   hash     cell key and per-cell count of every state position (first step
//            only; afterwards the integrate pass produces the next keys)
//   density  rho, EOS, neighbour bit masks (candidate tests happen here once)
//   force    pressure + viscosity from the masks, integrate, clamp, next key
// (sorting and the gather into sorted order live in sph_sort.hip).
// Neighbour order is canonical: rows (dz, dy) = (-1..1, -1..1), dz outer; a
// row is the contiguous sorted range of the x cells inside the support
// sphere's chord at that row; ascending index.  Density sums and force sums
// run in that order, one particle per lane.  Density and force each come in
// two forms with the same results: candidates read from global memory, or
// each row's candidates first copied by the wave into LDS ("staged").
#include <atomic>

#include "sph_internal.hpp"

namespace sph {

namespace {

// Tiles handled by this block (hash and the debug neighbour dump): a
// contiguous run per block when gridDim < ntiles, XCD-aware order.
__device__ __forceinline__ void tile_span(uint32_t ntiles, uint32_t& t0, uint32_t& t1) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    t0 = lb * per;
    t1 = t0 + per < ntiles ? t0 + per : ntiles;
}

// m = (m << 1) | (sign bit of c): one funnel shift (v_alignbit_b32).
__device__ __forceinline__ uint32_t push_sign(uint32_t m, float c) {
    return __builtin_amdgcn_alignbit(m, __float_as_uint(c), 31u);
}

// Candidates come from the x, y, z planes: one 16-byte load per plane gives
// four consecutive candidates (three loads per trip instead of four, which
// relieves the texture address units this pass otherwise saturates).
// Trips of four full candidates run without slot checks; the row's last,
// partial trip masks the slots past its end.
template <bool FULL>
__device__ __forceinline__ void density_trip(const float4& pi, float h2, const float4& X, const float4& Y,
                                             const float4& Z, int rem, float& sum, uint32_t& m) {
    const float xs[4] = {X.x, X.y, X.z, X.w}, ys[4] = {Y.x, Y.y, Y.z, Y.w}, zs[4] = {Z.x, Z.y, Z.z, Z.w};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        // q = d2 - h2 is negative exactly when d2 < h2 (an exact tie rounds
        // to +0), and -q = h2 - d2 (rounding is symmetric) is the kernel's t;
        // slots past the row end get q = 1.  c = min(q, 0) is q where q < 0
        // and +0 elsewhere, a NaN q included (the minimum of a quiet NaN and
        // a number is the number): a NaN distance (a candidate with a
        // non-finite coordinate) may carry a set sign bit, and it is no
        // neighbour, as d2 < h2 says.  So c's sign bit is d2 < h2, and -c is
        // t = max(h2 - d2, 0): c * c * -c adds t^3, and -0 where t is 0.
        float q = dist2(pi.x - xs[u], pi.y - ys[u], pi.z - zs[u]) - h2;
        if (!FULL && u != 0 && u >= rem) q = 1.0f;
        const float c = fminf(q, 0.0f);
        sum = fmaf(c * c, -c, sum);
        m = push_sign(m, c);
    }
}

__device__ __forceinline__ void density_one(const StepArgs& a, const SceneConst& sc, uint32_t k,
                                            __amdgpu_buffer_rsrc_t rpos, __amdgpu_buffer_rsrc_t rx,
                                            __amdgpu_buffer_rsrc_t ry, __amdgpu_buffer_rsrc_t rz,
                                            __amdgpu_buffer_rsrc_t rcs) {
    const float4 pi = load4(rpos, k << 4);
    const RowGeom g = row_geom(sc, pi.x, pi.y, pi.z);
    uint32_t lo[kRows], hi[kRows];
#pragma unroll
    for (int row = 0; row < kRows; ++row) row_window(rcs, sc, g, row % 3 - 1, row / 3 - 1, lo[row], hi[row]);
    const float h2 = sc.h2;
    float sum = 0.0f;
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        uint32_t m = 0u;
        const uint32_t cnt = hi[row] - lo[row];
        uint32_t off = lo[row] << 2;
        const uint32_t full_end = off + ((cnt & ~3u) << 2);
        // Slots past the row end shift in zero bits (past the array end the
        // buffer returns 0).  Each trip's planes are loaded while the trip
        // before is tested, so a lane has two trips' loads in flight.
        if (off != full_end) {
            float4 X = load4(rx, off), Y = load4(ry, off), Z = load4(rz, off);
            for (off += 16u; off != full_end; off += 16u) {
                const float4 Xn = load4(rx, off), Yn = load4(ry, off), Zn = load4(rz, off);
                density_trip<true>(pi, h2, X, Y, Z, 4, sum, m);
                X = Xn;
                Y = Yn;
                Z = Zn;
            }
            density_trip<true>(pi, h2, X, Y, Z, 4, sum, m);
        }
        const int rem = (int)(cnt & 3u);
        if (rem > 0) density_trip<false>(pi, h2, load4(rx, off), load4(ry, off), load4(rz, off), rem, sum, m);
        a.mask[(size_t)row * a.n + k] = m;
    }
    // The walls' half-space fluid, faces in the order x-, x+, y-, y+, z-, z+.
    float wr = 0.0f;
    const float pk[3] = {pi.x, pi.y, pi.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        wr += wall_rho((pk[d] - sc.bmin[d]) * sc.inv_h);
        wr += wall_rho((sc.wmax[d] - pk[d]) * sc.inv_h);
    }
    const float rho = sc.cD * sum + sc.rho0 * wr;
    const float p = eos(sc, rho);
    const float ir = 1.0f / rho;
    a.rho[k] = rho;
    // p/rho^2 and 1/rho ride in the unused w of the sorted position and
    // velocity, so the force pass gets them with the same loads.  Neighbour
    // reads of pos_s in this pass take x, y, z only.
    reinterpret_cast<float*>(a.pos_s + k)[3] = (p * ir) * ir;
    reinterpret_cast<float*>(a.vel_s + k)[3] = ir;
}

// Candidates a wave stages per row in the density pass, plus the slots a
// row's last, partial trip reads past its end (masked out).  With 124
// floats per plane, a trip's reads of all three planes fit the 8-bit dword
// offsets of ds_read2_b32 from one address (z at 248..251).
constexpr uint32_t kStageD = 120;
static_assert(kStageD <= 128, "the copy into a slot takes two records per lane");
constexpr uint32_t kPlaneD = kStageD + 4;
struct DensitySlot {
    float v[3 * kPlaneD];  // the x plane, then y, then z
};

// Four consecutive slot entries of plane d (0 x, 1 y, 2 z) from the x
// plane's address of the first.
__device__ __forceinline__ float4 slot_quad(const float* p, int d) {
    const float* q = p + d * kPlaneD;
    return make_float4(q[0], q[1], q[2], q[3]);
}

// Four consecutive sorted positions as x, y, z quadruples (the gathered
// trip's planes, read from the position records instead).
typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ void load_quad(__amdgpu_buffer_rsrc_t rpos, uint32_t j, float4& X, float4& Y, float4& Z) {
    // x, y, z of each record (12-byte loads: the w lane is not needed).
    const u32x3 p0 = __builtin_amdgcn_raw_buffer_load_b96(rpos, (int)(j << 4), 0, 0);
    const u32x3 p1 = __builtin_amdgcn_raw_buffer_load_b96(rpos, (int)((j + 1u) << 4), 0, 0);
    const u32x3 p2 = __builtin_amdgcn_raw_buffer_load_b96(rpos, (int)((j + 2u) << 4), 0, 0);
    const u32x3 p3 = __builtin_amdgcn_raw_buffer_load_b96(rpos, (int)((j + 3u) << 4), 0, 0);
    X = make_float4(__uint_as_float(p0.x), __uint_as_float(p1.x), __uint_as_float(p2.x), __uint_as_float(p3.x));
    Y = make_float4(__uint_as_float(p0.y), __uint_as_float(p1.y), __uint_as_float(p2.y), __uint_as_float(p3.y));
    Z = make_float4(__uint_as_float(p0.z), __uint_as_float(p1.z), __uint_as_float(p2.z), __uint_as_float(p3.z));
}

// The sorted range the staged density pass copies for one row: the union of
// the wave's windows of that row.  A wave whose active lanes (a prefix of the
// wave) share one y/z column holds their x cells in ascending order, so every
// lane's window of a row lies in the row's x cells [cx0 - wmax, cx1 + 1 +
// wmax]: cx0 and cx1 the first and last lanes' cells, wmax above every
// lane's half chord in x cells.  Such a wave bounds each row's union by two
// scalar reads of the cell-start table (a superset of the exact union, copied
// whole) instead of a wave-wide minimum and maximum, which the other waves,
// and rows whose superset does not fit the slot, still take.
struct WaveColumn {
    bool one_col;
    uint32_t row0, x0c, x1c;
};

__device__ __forceinline__ WaveColumn wave_column(const SceneConst& sc, const RowGeom& g, bool active) {
    WaveColumn wc;
    const uint64_t act = __ballot(active);
    wc.row0 = __builtin_amdgcn_readfirstlane(g.row0);
    wc.one_col = act != 0u && __ballot(active && g.row0 != wc.row0) == 0u;
    wc.x0c = wc.x1c = 0u;
    if (wc.one_col) {
        const float gmax = (float)(sc.G[0] - 1);
        const int cx = (int)fminf(fmaxf(g.fx, 0.0f), gmax);
        const int cx0 = __builtin_amdgcn_readfirstlane(cx);
        const int cx1 = __builtin_amdgcn_readlane(cx, 63 - __builtin_clzll(act));
        // sqrt(hc2) bounds every lane's sqrt(w2) up to rounding, which the
        // added 1/16 cell covers.
        const float wmax = __builtin_amdgcn_sqrtf(sc.hc2) * sc.xsubf + sc.wmargin + 0.0625f;
        wc.x0c = (uint32_t)(int)fminf(fmaxf((float)cx0 - wmax, 0.0f), gmax);
        wc.x1c = (uint32_t)(int)fminf(fmaxf((float)cx1 + 1.0f + wmax, 0.0f), gmax);
    }
    return wc;
}

// [nlo, nhi) of row `row` for a slot of CAP entries (empty when no lane has
// a window in the row); false when it is empty or does not fit (the row is
// then read from global memory).  Every lane of the wave calls it.
template <uint32_t CAP>
__device__ __forceinline__ bool row_union(const StepArgs& a, const SceneConst& sc, const WaveColumn& wc, int row,
                                          uint32_t lo, uint32_t hi, uint32_t& nlo, uint32_t& nhi) {
    const bool has = hi > lo;
    nlo = nhi = 0u;
    if (__ballot(has) == 0u) return false;
    if (wc.one_col) {
        // Some lane's window of this row is not empty, so the row's column
        // lies inside the grid and both reads inside the table.
        const uint32_t c = wc.row0 + (uint32_t)((row / 3 - 1) * sc.G[1] + (row % 3 - 1)) * (uint32_t)sc.G[0];
        nlo = *(ConstWord*)(a.cstart + (c + wc.x0c));
        nhi = *(ConstWord*)(a.cstart + (c + wc.x1c + 1u));
        if (nlo < nhi && nhi - nlo <= CAP) return true;
    }
    nlo = wave_umin(has ? lo : 0xffffffffu);
    nhi = wave_umax(has ? hi : 0u);
    return nlo < nhi && nhi - nlo <= CAP;
}

// density_one with each row's candidates read from LDS: per row the wave
// copies the positions of the union of its lanes' windows into its slot as
// x, y, z planes (coalesced record loads), then every lane runs its own
// four-candidate trips over the slot.  It reads the sorted position records
// only, so the reorder pass need not write the planes for it.  Same
// candidates, same order, same arithmetic: density_one's results bit for
// bit.  Every lane of the wave calls it; inactive lanes have empty windows
// and store nothing.
__device__ __forceinline__ void density_one_staged(const StepArgs& a, const SceneConst& sc, uint32_t k, bool active,
                                                   __amdgpu_buffer_rsrc_t rpos, __amdgpu_buffer_rsrc_t rcs,
                                                   DensitySlot& slot) {
    const float4 pi = load4(rpos, k << 4);  // zeros past the buffer end
    const RowGeom g = row_geom(sc, pi.x, pi.y, pi.z);
    uint32_t lo[kRows], hi[kRows];
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        row_window(rcs, sc, g, row % 3 - 1, row / 3 - 1, lo[row], hi[row]);
        if (!active) lo[row] = hi[row] = 0u;
    }
    const uint32_t lane = threadIdx.x & 63u;
    const float h2 = sc.h2;
    float sum = 0.0f;
    const WaveColumn wc = wave_column(sc, g, active);
    // Row r + 1's records are loaded into registers while row r's trips run
    // from the slot, and stored into the slot once those trips are done, so
    // each copy's latency hides behind the previous row's work.  A copy holds
    // at most 2 x 64 records: lanes take entries lane and lane + 64.
    uint32_t nlo = 0u, nhi = 0u;
    bool nfit;
    float4 n0 = make_float4(0.f, 0.f, 0.f, 0.f), n1 = n0;
    auto fetch = [&](int row) {
        nfit = row_union<kStageD>(a, sc, wc, row, lo[row], hi[row], nlo, nhi);
        if (nfit) {
            if (lane < nhi - nlo) n0 = load4(rpos, (nlo + lane) << 4);
            if (lane + 64u < nhi - nlo) n1 = load4(rpos, (nlo + lane + 64u) << 4);
        }
    };
    fetch(0);
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        uint32_t m = 0u;
        const uint32_t cnt = hi[row] - lo[row];
        const bool has = cnt > 0u;
        const uint32_t ulo = nlo, uhi = nhi;
        const bool fit = nfit;
        if (fit) {
            const uint32_t span = uhi - ulo;
            if (lane < span) {
                slot.v[lane] = n0.x;
                slot.v[kPlaneD + lane] = n0.y;
                slot.v[2 * kPlaneD + lane] = n0.z;
            }
            if (lane + 64u < span) {
                slot.v[lane + 64u] = n1.x;
                slot.v[kPlaneD + lane + 64u] = n1.y;
                slot.v[2 * kPlaneD + lane + 64u] = n1.z;
            }
            wave_lds_sync();
        }
        if (row + 1 < kRows) fetch(row + 1);
        if (fit) {
            // One induction variable, the trip's slot address, is also the
            // loop's exit test.
            const float* p = slot.v + (lo[row] - ulo);
            const float* const full_end = p + (cnt & ~3u);
            for (; p != full_end; p += 4)
                density_trip<true>(pi, h2, slot_quad(p, 0), slot_quad(p, 1), slot_quad(p, 2), 4, sum, m);
            const int rem = (int)(cnt & 3u);
            if (rem > 0)
                density_trip<false>(pi, h2, slot_quad(p, 0), slot_quad(p, 1), slot_quad(p, 2), rem, sum, m);
            wave_lds_sync();
        } else if (has) {
            // Past the buffer end the records read as zeros, as the planes do.
            uint32_t j = lo[row];
            const uint32_t full_end = j + (cnt & ~3u);
            float4 X, Y, Z;
            for (; j != full_end; j += 4u) {
                load_quad(rpos, j, X, Y, Z);
                density_trip<true>(pi, h2, X, Y, Z, 4, sum, m);
            }
            const int rem = (int)(cnt & 3u);
            if (rem > 0) {
                load_quad(rpos, j, X, Y, Z);
                density_trip<false>(pi, h2, X, Y, Z, rem, sum, m);
            }
        }
        if (active) a.mask[(size_t)row * a.n + k] = m;
    }
    if (!active) return;
    float wr = 0.0f;
    const float pk[3] = {pi.x, pi.y, pi.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        wr += wall_rho((pk[d] - sc.bmin[d]) * sc.inv_h);
        wr += wall_rho((sc.wmax[d] - pk[d]) * sc.inv_h);
    }
    const float rho = sc.cD * sum + sc.rho0 * wr;
    const float p = eos(sc, rho);
    const float ir = 1.0f / rho;
    a.rho[k] = rho;
    reinterpret_cast<float*>(a.pos_s + k)[3] = (p * ir) * ir;
    reinterpret_cast<float*>(a.vel_s + k)[3] = ir;
}

struct ForceAcc {
    float px, py, pz, vx, vy, vz;
};

// pos.w = p/rho^2 and vel.w = 1/rho of each particle.
__device__ __forceinline__ void force_pair(const SceneConst& sc, const float4& pi, const float4& vi,
                                           const float4& pj, const float4& vj, ForceAcc& f) {
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float d2 = dist2(dx, dy, dz);
    // Hardware reciprocal square root.  d2 of two distinct particles is never
    // denormal; coincident particles (d2 = 0) get a finite 1/r times a zero
    // offset, so their pressure term vanishes as it does with 1/r := 0.
    const float inv_r = __builtin_amdgcn_rsqf(fmaxf(d2, 1e-30f));
    const float r = d2 * inv_r;
    const float hr = sc.h - r;
    const float w = hr * hr;
    const float s = ((pi.w + pj.w) * w) * inv_r;
    f.px = fmaf(s, dx, f.px);
    f.py = fmaf(s, dy, f.py);
    f.pz = fmaf(s, dz, f.pz);
    const float vv = hr * vj.w;
    f.vx = fmaf(vv, vj.x - vi.x, f.vx);
    f.vy = fmaf(vv, vj.y - vi.y, f.vy);
    f.vz = fmaf(vv, vj.z - vi.z, f.vz);
}

__device__ __forceinline__ uint32_t force_finish(const StepArgs& a, const SceneConst& sc, uint32_t k,
                                                 const float4& pi, const float4& vi, const ForceAcc& f);

__device__ __forceinline__ uint32_t force_one(const StepArgs& a, const SceneConst& sc, uint32_t k,
                                              __amdgpu_buffer_rsrc_t rpos, __amdgpu_buffer_rsrc_t rvel,
                                              __amdgpu_buffer_rsrc_t rcs) {
    const float4 pi = load4(rpos, k << 4);
    const float4 vi = load4(rvel, k << 4);
    const RowGeom g = row_geom(sc, pi.x, pi.y, pi.z);
    uint32_t lo[kRows], hi[kRows], mk[kRows];
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        row_window(rcs, sc, g, row % 3 - 1, row / 3 - 1, lo[row], hi[row]);
        mk[row] = a.mask[(size_t)row * a.n + k];
    }
    {
        const uint32_t b = mask_bit(lo[4], hi[4], k);  // own bit, in the centre row
        if (b < (uint32_t)kMaskBits) mk[4] &= ~(1u << b);
    }
    ForceAcc f = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        const uint32_t span = mask_span(lo[row], hi[row]);
        const uint32_t j0 = row_first(lo[row], hi[row]);
        for (uint32_t j = lo[row]; j < j0; ++j) {
            const float4 pj = load4(rpos, j << 4);
            if (j != k && dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < sc.h2)
                force_pair(sc, pi, vi, pj, load4(rvel, j << 4), f);
        }
        // The mask's neighbours in ascending order (bit i: candidate j0 + i),
        // each one's records loaded while the one before is summed, so a
        // lane has two neighbours' loads in flight; unrolled twice, the two
        // register sets swap roles instead of being copied.
        uint32_t r = __builtin_bitreverse32(mk[row]);
        if (span - 1u < 31u) r >>= 32u - span;
        if (r) {
            uint32_t j = j0 + (uint32_t)__builtin_ctz(r);
            float4 pj = load4(rpos, j << 4), vj = load4(rvel, j << 4);
#pragma unroll 2
            for (r &= r - 1u; r; r &= r - 1u) {
                j = j0 + (uint32_t)__builtin_ctz(r);
                const float4 pn = load4(rpos, j << 4), vn = load4(rvel, j << 4);
                force_pair(sc, pi, vi, pj, vj, f);
                pj = pn;
                vj = vn;
            }
            force_pair(sc, pi, vi, pj, vj, f);
        }
    }
    return force_finish(a, sc, k, pi, vi, f);
}

// Walls, integrate, clamp and stores of particle k; returns its next key.
__device__ __forceinline__ uint32_t force_finish(const StepArgs& a, const SceneConst& sc, uint32_t k,
                                                 const float4& pi, const float4& vi, const ForceAcc& f) {
    const float cvi = sc.cV * vi.w;
    // The walls' pressure term at this particle's own p/rho^2.
    const float cwi = sc.cW * pi.w;
    const float pk[3] = {pi.x, pi.y, pi.z};
    float aw[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        aw[d] = cwi * (wall_j((pk[d] - sc.bmin[d]) * sc.inv_h) - wall_j((sc.wmax[d] - pk[d]) * sc.inv_h));
    const float a0 = ((sc.cP * f.px + cvi * f.vx) + sc.g[0]) + aw[0];
    const float a1 = ((sc.cP * f.py + cvi * f.vy) + sc.g[1]) + aw[1];
    const float a2 = ((sc.cP * f.pz + cvi * f.vz) + sc.g[2]) + aw[2];
    if (a.acc) a.acc[k] = make_float4(a0, a1, a2, 0.0f);
    float v[3] = {vi.x + sc.dt * a0, vi.y + sc.dt * a1, vi.z + sc.dt * a2};
    float p[3] = {pi.x + sc.dt * v[0], pi.y + sc.dt * v[1], pi.z + sc.dt * v[2]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (p[d] < sc.lo[d]) {
            p[d] = sc.lo[d];
            if (v[d] < 0.0f) v[d] = -sc.e * v[d];
        } else if (p[d] > sc.hi[d]) {
            p[d] = sc.hi[d];
            if (v[d] > 0.0f) v[d] = -sc.e * v[d];
        }
    }
    a.pos_o[k] = make_float4(p[0], p[1], p[2], 0.0f);
    a.vel_o[k] = make_float4(v[0], v[1], v[2], 0.0f);
    a.id_o[k] = a.id_s[k];
    return cell_key(sc, p[0], p[1], p[2]);
}

// Candidates a wave stages per row: its lanes' windows of one row lie in
// one sorted range of about 64 + 2 chords of particles when the wave sits in
// one y/z column (73 on average on the 1M dam break, 99 at the 99th
// percentile of the 4.4M sphere drop); a wider range (a wave across two
// columns) is walked from global memory as in force_one.  A wave has two
// slots, one filled while the other is walked.
constexpr uint32_t kStage = 96;
static_assert(kStage <= 128, "the copy into a slot takes two records per lane");
struct StageSlot {
    float4 pos[kStage];
    float4 vel[kStage];
};

// force_one with each row's candidates read from LDS: per row, the wave
// copies the union of its lanes' windows (contiguous, so the copy is
// coalesced) into a slot, then every lane walks its own mask over the
// slot.  Same pairs, same order, same arithmetic: the results are
// force_one's bit for bit.  Every lane of the wave calls it (the copy and
// the wave-wide bounds need them all); inactive lanes have empty windows.
__device__ __forceinline__ uint32_t force_one_staged(const StepArgs& a, const SceneConst& sc, uint32_t k,
                                                     bool active, __amdgpu_buffer_rsrc_t rpos,
                                                     __amdgpu_buffer_rsrc_t rvel, __amdgpu_buffer_rsrc_t rcs,
                                                     StageSlot& slot_a, StageSlot& slot_b) {
    // Loads past the buffer end return zeros.
    const float4 pi = load4(rpos, k << 4);
    const float4 vi = load4(rvel, k << 4);
    const RowGeom g = row_geom(sc, pi.x, pi.y, pi.z);
    uint32_t lo[kRows], hi[kRows], mk[kRows];
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        row_window(rcs, sc, g, row % 3 - 1, row / 3 - 1, lo[row], hi[row]);
        mk[row] = active ? a.mask[(size_t)row * a.n + k] : 0u;
        if (!active) lo[row] = hi[row] = 0u;
    }
    {
        const uint32_t b = mask_bit(lo[4], hi[4], k);
        if (active && b < (uint32_t)kMaskBits) mk[4] &= ~(1u << b);
    }
    const uint32_t lane = threadIdx.x & 63u;
    ForceAcc f = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    // Each row's union of windows is copied into one of two slots by direct
    // loads into LDS (no registers), row r + 1's copy issued before row r is
    // walked, so the copy runs behind the walk.  Lane i of a copy writes
    // entry i (and i + 64) of the slot.
    uint32_t nlo, nhi;
    bool nfit;
    auto fetch = [&](int row) {
        const bool has = hi[row] > lo[row];
        nlo = wave_umin(has ? lo[row] : 0xffffffffu);
        nhi = wave_umax(has ? hi[row] : 0u);
        nfit = nlo < nhi && nhi - nlo <= kStage;
        if (!nfit) return;
        StageSlot& sl = (row & 1) ? slot_b : slot_a;
        const uint32_t span = nhi - nlo;
        if (lane < span) {
            lds_copy16(rpos, sl.pos, (nlo + lane) << 4);
            lds_copy16(rvel, sl.vel, (nlo + lane) << 4);
        }
        if (span > 64u && lane + 64u < span) {
            lds_copy16(rpos, sl.pos + 64, (nlo + lane + 64u) << 4);
            lds_copy16(rvel, sl.vel + 64, (nlo + lane + 64u) << 4);
        }
    };
    fetch(0);
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        const uint32_t ulo = nlo, uhi = nhi;
        const bool fit = nfit;
        StageSlot& slot = (row & 1) ? slot_b : slot_a;
        if (fit) lds_copy_wait();
        if (row + 1 < kRows) fetch(row + 1);
        if (ulo >= uhi) continue;
        if (fit) {
            // Entry j of the row sits at slot.pos[j - ulo], its velocity
            // kStage records further on; the mask's entries from row_first.
            const float4* base = slot.pos - ulo;
            const float4* bm = slot.pos + (row_first(lo[row], hi[row]) - ulo);
            row_neighbors_from(
                lo[row], hi[row], mk[row],
                [&](uint32_t j) {
                    const float4 pj = base[j];
                    return j != k && dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < sc.h2;
                },
                [&](uint32_t j) { force_pair(sc, pi, vi, base[j], base[j + kStage], f); },
                [&](uint32_t i) { force_pair(sc, pi, vi, bm[i], bm[i + kStage], f); });
        } else {
            row_neighbors(
                lo[row], hi[row], mk[row],
                [&](uint32_t j) {
                    const float4 pj = load4(rpos, j << 4);
                    return j != k && dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < sc.h2;
                },
                [&](uint32_t j) { force_pair(sc, pi, vi, load4(rpos, j << 4), load4(rvel, j << 4), f); });
        }
    }
    return active ? force_finish(a, sc, k, pi, vi, f) : 0u;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void density_kernel(StepArgs a) {
    if (a.halt && *a.halt) return;
    const __amdgpu_buffer_rsrc_t rpos = buffer_of(a.pos_s, a.n);
    const __amdgpu_buffer_rsrc_t rx = buffer_of(a.soa, a.n);
    const __amdgpu_buffer_rsrc_t ry = buffer_of(a.soa + a.n, a.n);
    const __amdgpu_buffer_rsrc_t rz = buffer_of(a.soa + 2 * (size_t)a.n, a.n);
    const __amdgpu_buffer_rsrc_t rcs = buffer_of(a.cstart, a.ncells + 1u);
    for_each_tile(a.ntiles, a.queue, [&](uint32_t t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        if (k < sc.p_end) density_one(a, sc, k, rpos, rx, ry, rz, rcs);
    });
}

// density_kernel with the rows staged in LDS (1.5 KB per wave).  Eight
// waves per SIMD: the bound fits it in 59 VGPRs without spills (65 without).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void density_staged_kernel(StepArgs a) {
    if (a.halt && *a.halt) return;
    __shared__ DensitySlot s_slot[BLOCK / 64];
    const __amdgpu_buffer_rsrc_t rpos = buffer_of(a.pos_s, a.n);
    const __amdgpu_buffer_rsrc_t rcs = buffer_of(a.cstart, a.ncells + 1u);
    DensitySlot& slot = s_slot[threadIdx.x >> 6];
    for_each_tile(a.ntiles, a.queue, [&](uint32_t t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        density_one_staged(a, sc, k, k < sc.p_end, rpos, rcs, slot);
    });
}

// Six waves per SIMD (77 VGPRs, no VGPR spills).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6, 6))) void force_kernel(StepArgs a) {
    if (a.halt && *a.halt) return;
    const __amdgpu_buffer_rsrc_t rpos = buffer_of(a.pos_s, a.n);
    const __amdgpu_buffer_rsrc_t rvel = buffer_of(a.vel_s, a.n);
    const __amdgpu_buffer_rsrc_t rcs = buffer_of(a.cstart, a.ncells + 1u);
    uint32_t* queue = a.queue ? a.queue + 8 : nullptr;
    for_each_tile(a.ntiles, queue, [&](uint32_t t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        const bool active = k < sc.p_end;
        uint32_t key = 0;
        if (active) key = force_one(a, sc, k, rpos, rvel, rcs);
        // Next step's key and its slot in the per-cell count (whole wave).
        const uint32_t slot = count_slot(a.count, a.count + a.ncells, key, active);
        if (a.vmax2) {
            float4 vn = make_float4(0.f, 0.f, 0.f, 0.f);
            if (active) vn = a.vel_o[k];
            wave_max_v2(a.vmax2, ti.x, vn.x, vn.y, vn.z, active);
        }
        if (active) {
            a.key0[k] = key;
            a.rank0[k] = slot;
        }
    });
}

// force_kernel with the rows staged in LDS (two 3 KB slots per wave).  Six
// waves per SIMD: the bound fits it in 80 VGPRs (one spilled; 82 and five
// waves without it).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6, 6))) void force_staged_kernel(StepArgs a) {
    if (a.halt && *a.halt) return;
    __shared__ StageSlot s_slot_a[BLOCK / 64];
    __shared__ StageSlot s_slot_b[BLOCK / 64];
    const __amdgpu_buffer_rsrc_t rpos = buffer_of(a.pos_s, a.n);
    const __amdgpu_buffer_rsrc_t rvel = buffer_of(a.vel_s, a.n);
    const __amdgpu_buffer_rsrc_t rcs = buffer_of(a.cstart, a.ncells + 1u);
    uint32_t* queue = a.queue ? a.queue + 8 : nullptr;
    StageSlot& slot_a = s_slot_a[threadIdx.x >> 6];
    StageSlot& slot_b = s_slot_b[threadIdx.x >> 6];
    for_each_tile(a.ntiles, queue, [&](uint32_t t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        const bool active = k < sc.p_end;
        const uint32_t key = force_one_staged(a, sc, k, active, rpos, rvel, rcs, slot_a, slot_b);
        const uint32_t cslot = count_slot(a.count, a.count + a.ncells, key, active);
        if (a.vmax2) {
            float4 vn = make_float4(0.f, 0.f, 0.f, 0.f);
            if (active) vn = a.vel_o[k];
            wave_max_v2(a.vmax2, ti.x, vn.x, vn.y, vn.z, active);
        }
        if (active) {
            a.key0[k] = key;
            a.rank0[k] = cslot;
        }
    });
}

__global__ void hash_kernel(const float4* __restrict__ pos, const float4* __restrict__ vel,
                            uint32_t* __restrict__ key0, uint32_t* __restrict__ rank0, uint32_t* __restrict__ count,
                            uint32_t ncells, uint32_t* __restrict__ vmax2, const SceneConst* __restrict__ scenes,
                            const uint2* __restrict__ tiles, uint32_t ntiles) {
    uint32_t t0, t1;
    tile_span(ntiles, t0, t1);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 ti = tile_at(tiles, t);
        const SceneConst sc = scene_at(scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        const bool active = k < sc.p_end;
        uint32_t key = 0;
        if (active) {
            const float4 p = pos[k];
            key = cell_key(sc, p.x, p.y, p.z);
        }
        const uint32_t slot = count_slot(count, count + ncells, key, active);
        if (active) {
            key0[k] = key;
            rank0[k] = slot;
        }
        if (vmax2) {
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (active) v = vel[k];
            wave_max_v2(vmax2, ti.x, v.x, v.y, v.z, active);
        }
    }
}

// Neighbour sets exactly as force_one iterates them (self included).
template <int FILL>
__global__ void neighbors_kernel(StepArgs a, const uint32_t* __restrict__ offsets,
                                 uint32_t* __restrict__ out) {
    const __amdgpu_buffer_rsrc_t rcs = buffer_of(a.cstart, a.ncells + 1u);
    uint32_t t0, t1;
    tile_span(a.ntiles, t0, t1);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        if (k >= sc.p_end) continue;
        const float4 pi = a.pos_s[k];
        const RowGeom g = row_geom(sc, pi.x, pi.y, pi.z);
        uint32_t c = 0;
        const uint32_t o = FILL ? offsets[k] : 0u;
        for (int row = 0; row < kRows; ++row) {
            uint32_t lo, hi;
            row_window(rcs, sc, g, row % 3 - 1, row / 3 - 1, lo, hi);
            row_neighbors(
                lo, hi, a.mask[(size_t)row * a.n + k],
                [&](uint32_t j) {
                    const float4 pj = a.pos_s[j];
                    return dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < sc.h2;
                },
                [&](uint32_t j) {
                    if (FILL) out[o + c] = j;
                    ++c;
                });
        }
        if (!FILL) out[k] = c;
    }
}

// dt = min(dt_max, (cfl h) / (c0 + |v|max)) per scene, in this op order.
// With a frame state: nothing once halted; else scene 0's step is added to
// the frame time (float, as the host frame loop adds it) and a frame of at
// least 1/30 s marks pending.
__global__ void cfl_kernel(uint32_t nscenes, SceneConst* __restrict__ scenes, uint32_t* __restrict__ vmax2,
                           float* __restrict__ dt_out, uint32_t* __restrict__ frame) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nscenes) return;
    if (frame && frame[0]) return;
    SceneConst& sc = scenes[s];
    if (sc.cfl > 0.0f) {
        const float vmax = sqrtf(__uint_as_float(vmax2[s]));
        sc.dt = fminf(sc.dt_max, (sc.cfl * sc.h) / (sc.c0 + vmax));
    }
    vmax2[s] = 0u;
    if (s == 0) {
        *dt_out = sc.dt;
        if (frame) {
            const float ft = __uint_as_float(frame[3]) + sc.dt;
            frame[3] = __float_as_uint(ft);
            frame[2] += 1u;
            if (!((double)ft < 0.0333333333)) frame[1] = 1u;
        }
    }
}

__global__ void frame_reset_kernel(uint32_t* __restrict__ frame) {
    if (threadIdx.x < (uint32_t)kFrameWords) frame[threadIdx.x] = 0u;
}

__global__ void frame_gate_kernel(uint32_t* __restrict__ frame) {
    if (threadIdx.x == 0 && frame[1]) frame[0] = 1u;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void unpermute3_kernel(uint32_t n, const float4* __restrict__ src,
                                  const uint32_t* __restrict__ id, float* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 v = src[k];
    const uint32_t i = id[k];
    out[3 * (size_t)i] = v.x;
    out[3 * (size_t)i + 1] = v.y;
    out[3 * (size_t)i + 2] = v.z;
}

__global__ void unpermute1_kernel(uint32_t n, const float* __restrict__ src,
                                  const uint32_t* __restrict__ id, float* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[id[k]] = src[k];
}

__global__ void pack3_kernel(uint32_t n, const float* __restrict__ src, float4* __restrict__ dst) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    dst[k] = src ? make_float4(src[3 * (size_t)k], src[3 * (size_t)k + 1], src[3 * (size_t)k + 2], 0.0f)
                 : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

inline uint32_t blocks_for(size_t n, int b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

#define SPH_DISPATCH_BLOCK(KERNEL, block, grid, st, ...)                                       \
    switch (block) {                                                                          \
        case 64: hipLaunchKernelGGL(KERNEL<64>, dim3(grid), dim3(64), 0, st, __VA_ARGS__); break;   \
        case 128: hipLaunchKernelGGL(KERNEL<128>, dim3(grid), dim3(128), 0, st, __VA_ARGS__); break; \
        case 512: hipLaunchKernelGGL(KERNEL<512>, dim3(grid), dim3(512), 0, st, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL(KERNEL<256>, dim3(grid), dim3(256), 0, st, __VA_ARGS__); break;  \
    }

hipError_t launch_density(hipStream_t st, int block, uint32_t grid, const StepArgs& a, int staged) {
    if (grid == 0) return hipSuccess;
    if (staged) {
        SPH_DISPATCH_BLOCK(density_staged_kernel, block, grid, st, a);
    } else {
        SPH_DISPATCH_BLOCK(density_kernel, block, grid, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_force(hipStream_t st, int block, uint32_t grid, const StepArgs& a, int staged) {
    if (grid == 0) return hipSuccess;
    if (staged) {
        SPH_DISPATCH_BLOCK(force_staged_kernel, block, grid, st, a);
    } else {
        SPH_DISPATCH_BLOCK(force_kernel, block, grid, st, a);
    }
    return hipGetLastError();
}

template <class K>
static int occupancy(K kernel, int block) {
    int nb = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) == hipSuccess ? nb : 0;
}

int force_blocks_per_cu(int block, int staged) {
    // Looked up once per form and block size; 0 = not yet (stores nb + 1).
    static std::atomic<int> cache[2][4];
    const int b = block == 64 ? 0 : block == 128 ? 1 : block == 512 ? 3 : 2;
    std::atomic<int>& c = cache[staged ? 1 : 0][b];
    const int known = c.load(std::memory_order_relaxed);
    if (known > 0) return known - 1;
    int nb = 0;
    switch (b) {
        case 0: nb = staged ? occupancy(force_staged_kernel<64>, 64) : occupancy(force_kernel<64>, 64); break;
        case 1: nb = staged ? occupancy(force_staged_kernel<128>, 128) : occupancy(force_kernel<128>, 128); break;
        case 3: nb = staged ? occupancy(force_staged_kernel<512>, 512) : occupancy(force_kernel<512>, 512); break;
        default: nb = staged ? occupancy(force_staged_kernel<256>, 256) : occupancy(force_kernel<256>, 256); break;
    }
    c.store(nb + 1, std::memory_order_relaxed);
    return nb;
}

hipError_t launch_hash(hipStream_t st, int block, uint32_t grid, const float4* pos, const float4* vel,
                       uint32_t* key0, uint32_t* rank0, uint32_t* count, uint32_t ncells, uint32_t* vmax2,
                       const SceneConst* scenes, const uint2* tiles, uint32_t ntiles) {
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(hash_kernel, dim3(grid), dim3(block), 0, st, pos, vel, key0, rank0, count, ncells, vmax2,
                       scenes, tiles, ntiles);
    return hipGetLastError();
}

hipError_t launch_neighbors(hipStream_t st, int block, uint32_t grid, const StepArgs& a,
                            const uint32_t* offsets, uint32_t* out, int fill) {
    if (grid == 0) return hipSuccess;
    if (fill)
        hipLaunchKernelGGL(neighbors_kernel<1>, dim3(grid), dim3(block), 0, st, a, offsets, out);
    else
        hipLaunchKernelGGL(neighbors_kernel<0>, dim3(grid), dim3(block), 0, st, a, offsets, out);
    return hipGetLastError();
}

hipError_t launch_cfl(hipStream_t st, uint32_t nscenes, SceneConst* scenes, uint32_t* vmax2, float* dt_out,
                      uint32_t* frame) {
    if (nscenes == 0) return hipSuccess;
    hipLaunchKernelGGL(cfl_kernel, dim3((nscenes + 255) / 256), dim3(256), 0, st, nscenes, scenes, vmax2, dt_out,
                       frame);
    return hipGetLastError();
}

hipError_t launch_frame_reset(hipStream_t st, uint32_t* frame) {
    hipLaunchKernelGGL(frame_reset_kernel, dim3(1), dim3(64), 0, st, frame);
    return hipGetLastError();
}

hipError_t launch_frame_gate(hipStream_t st, uint32_t* frame) {
    hipLaunchKernelGGL(frame_gate_kernel, dim3(1), dim3(64), 0, st, frame);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t value, size_t count) {
    if (count == 0) return hipSuccess;
    const uint32_t g = blocks_for(count, 256) < 4096 ? blocks_for(count, 256) : 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(g), dim3(256), 0, st, p, value, count);
    return hipGetLastError();
}

hipError_t launch_unpermute3(hipStream_t st, uint32_t n, const float4* src, const uint32_t* id, float* out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(unpermute3_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, src, id, out);
    return hipGetLastError();
}

hipError_t launch_unpermute1(hipStream_t st, uint32_t n, const float* src, const uint32_t* id, float* out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(unpermute1_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, src, id, out);
    return hipGetLastError();
}

hipError_t launch_pack3(hipStream_t st, uint32_t n, const float* src, float4* dst) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pack3_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, src, dst);
    return hipGetLastError();
}

}  // namespace sph
