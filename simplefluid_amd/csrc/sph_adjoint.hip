// Hand-written reverse pass of one SPH step (gather form, no float atomics).
//
// Forward, per sorted particle i with neighbours j != i, r = x_i - x_j,
// d = |r|, hr = h - d, t = h^2 - d^2, P = p/rho^2, q = 1/rho:
//   rho_i = cD sum_j t^3                          (self included)
//   a_i   = cP sum (P_i+P_j) hr^2/d r + cV q_i sum q_j hr (v_j - v_i) + g
//   v' = v + dt a ; x~ = x + dt v' ; clamp(x~, v') -> (x', v'')
// Reverse, given xbar', vbar'' (output order = this step's sorted order):
//   adj_integrate: clamp masks, xbar~ , vbar' = vbar'' (or -e vbar'' where the
//                  velocity was reflected) + dt xbar~, abar = dt vbar'
//   adj_force:     xbar_i += sum cP(P_i+P_j)[phi da + (phi'/d)(da.r) r]
//                           + cV q_i q_j (psi'/d)(da.dv) r
//                  vbar_i += sum cV q_i q_j hr (abar_j - abar_i)
//                  Pbar_i = sum cP phi da.r ; qbar_i = sum cV q_j hr da.dv
//                  rhobar_i = Pbar_i (p' q^2 - 2 p q^3) - qbar_i q^2
//                  with da = abar_i - abar_j, dv = v_j - v_i, phi = hr^2/d,
//                  phi'/d = -t/d^3, psi' = -1; per-particle dL/dB and dL/dmu terms
//   adj_density:   xbar_i += sum -6 cD t^2 (rhobar_i + rhobar_j) r, then
//                  scatter both adjoints back to the step's input order.
// Walls, per axis d (u_lo = (x_d - box_min_d)/h, u_hi = (box_max_d - x_d)/h;
// forward: rho_i += rho0 wall_rho(u), a_i,d += cW P_i (wall_j(u_lo) - wall_j(u_hi))):
//   adj_integrate: xbar_i,d += cW P_i abar_i,d (wall_j'(u_lo) + wall_j'(u_hi)) / h,
//                  and cW abar_i,d (wall_j(u_lo) - wall_j(u_hi)) in abar.w, where
//   adj_force      starts Pbar_i from it (the term involves particle i alone)
//   adj_density:   xbar_i,d += rhobar_i rho0 (wall_rho'(u_lo) - wall_rho'(u_hi)) / h
// Each kernel reads neighbour values written by the previous kernel only.
// adj_force and adj_density come in two forms with identical results: the
// neighbours' records gathered from global memory, or each row's candidates
// staged in LDS by the wave first (the forward passes' staged form).
#include "sph_internal.hpp"

namespace sph {

namespace {

__device__ __forceinline__ void tile_span(uint32_t ntiles, uint32_t& t0, uint32_t& t1) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    t0 = lb * per;
    t1 = t0 + per < ntiles ? t0 + per : ntiles;
}

// Visit the neighbours j != k the forward force pass visited, in its order.
template <class F>
__device__ __forceinline__ void for_each_neighbor(const AdjArgs& a, const SceneConst& sc, uint32_t k,
                                                  const float4& pi, F&& f) {
    const __amdgpu_buffer_rsrc_t rcs = buffer_of(a.cstart, a.ncells + 1u);
    const RowGeom g = row_geom(sc, pi.x, pi.y, pi.z);
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        uint32_t lo, hi;
        row_window(rcs, sc, g, row % 3 - 1, row / 3 - 1, lo, hi);
        uint32_t m = a.mask[(size_t)row * a.n + k];
        if (row == 4) {
            const uint32_t b = mask_bit(lo, hi, k);
            if (b < (uint32_t)kMaskBits) m &= ~(1u << b);
        }
        row_neighbors(
            lo, hi, m,
            [&](uint32_t j) {
                const float4 pj = a.pos_s[j];
                return j != k && dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < sc.h2;
            },
            f);
    }
}

__global__ void adj_integrate_kernel(AdjArgs a) {
    uint32_t t0, t1;
    tile_span(a.ntiles, t0, t1);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        if (k >= sc.p_end) continue;
        const float4 x0 = a.pos_s[k], v0 = a.vel_s[k], ac = a.acc[k];
        const float4 xb = a.xbar[k], vb = a.vbar[k];
        const float xa[3] = {x0.x, x0.y, x0.z}, va[3] = {v0.x, v0.y, v0.z}, aa[3] = {ac.x, ac.y, ac.z};
        const float xba[3] = {xb.x, xb.y, xb.z}, vba[3] = {vb.x, vb.y, vb.z};
        float xo[3], vo[3], ab[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float v1 = va[d] + sc.dt * aa[d];
            const float xt = xa[d] + sc.dt * v1;
            float xbt = xba[d], vb1 = vba[d];
            if (xt < sc.lo[d]) {
                xbt = 0.0f;
                if (v1 < 0.0f) vb1 = -sc.e * vb1;
            } else if (xt > sc.hi[d]) {
                xbt = 0.0f;
                if (v1 > 0.0f) vb1 = -sc.e * vb1;
            }
            vb1 += sc.dt * xbt;
            xo[d] = xbt;
            vo[d] = vb1;
            ab[d] = sc.dt * vb1;
        }
        // Walls: their pressure term depends on this particle alone, so its
        // reverse (Pbar share in abar.w, xbar here) is elementwise.
        float pw = 0.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float ul = (xa[d] - sc.bmin[d]) * sc.inv_h, uh = (sc.wmax[d] - xa[d]) * sc.inv_h;
            pw += sc.cW * ab[d] * (wall_j(ul) - wall_j(uh));
            xo[d] += ((sc.cW * x0.w) * ab[d]) * (sc.inv_h * (wall_j_du(ul) + wall_j_du(uh)));
        }
        a.xs[k] = make_float4(xo[0], xo[1], xo[2], 0.0f);
        a.vs[k] = make_float4(vo[0], vo[1], vo[2], 0.0f);
        a.abar[k] = make_float4(ab[0], ab[1], ab[2], pw);
    }
}

// Per-particle sums of the adjoint force pass.
struct AdjForceAcc {
    float gx, gy, gz, hx, hy, hz, Pb, qb, Ax, Ay, Az;
};

// One neighbour's terms of the adjoint force pass (pos.w = p/rho^2, vel.w = 1/rho).
__device__ __forceinline__ void adj_force_pair(const SceneConst& sc, const float4& pi, const float4& vi,
                                               const float4& abi, const float4& pj, const float4& vj,
                                               const float4& abj, AdjForceAcc& c) {
    const float Pi = pi.w, qi = vi.w;
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    const float d2 = dist2(rx, ry, rz);
    const float inv_r = d2 > 0.0f ? __builtin_amdgcn_rsqf(d2) : 0.0f;
    const float dd = d2 * inv_r;
    const float hr = sc.h - dd;
    const float tt = sc.h2 - d2;
    const float phi = hr * hr * inv_r;
    const float dphi_r = -tt * inv_r * inv_r * inv_r;
    const float dax = abi.x - abj.x, day = abi.y - abj.y, daz = abi.z - abj.z;
    const float dvx = vj.x - vi.x, dvy = vj.y - vi.y, dvz = vj.z - vi.z;
    const float PP = sc.cP * (Pi + pj.w);
    const float dar = dax * rx + day * ry + daz * rz;
    const float dav = dax * dvx + day * dvy + daz * dvz;
    const float cq = sc.cV * qi * vj.w;
    const float coefr = PP * dphi_r * dar - cq * inv_r * dav;
    c.gx += PP * phi * dax + coefr * rx;
    c.gy += PP * phi * day + coefr * ry;
    c.gz += PP * phi * daz + coefr * rz;
    c.hx -= cq * hr * dax;
    c.hy -= cq * hr * day;
    c.hz -= cq * hr * daz;
    c.Pb += sc.cP * phi * dar;
    c.qb += sc.cV * vj.w * hr * dav;
    c.Ax += hr * vj.w * dvx;
    c.Ay += hr * vj.w * dvy;
    c.Az += hr * vj.w * dvz;
}

// The EOS reverse and the stores of particle k; returns its (dL/dB, dL/dmu) terms.
__device__ __forceinline__ double2 adj_force_finish(const AdjArgs& a, const SceneConst& sc, uint32_t k,
                                                    const float4& vi, const float4& abi, const AdjForceAcc& c) {
    const float qi = vi.w;
    // EOS derivatives (clamped Tait, exponent 7).
    const float rho = a.rho[k];
    const float x = rho * sc.inv_rho0;
    const float x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x7 = x6 * x;
    const float praw = sc.B * (x7 - 1.0f);
    const float sgn = praw < 0.0f ? (sc.attractive ? sc.attr : 0.0f) : 1.0f;
    const float p = praw * sgn;
    const float dp_drho = sgn * sc.B * 7.0f * x6 * sc.inv_rho0;
    const float dp_dB = sgn * (x7 - 1.0f);
    a.rhobar[k] = c.Pb * (dp_drho * qi * qi - 2.0f * p * qi * qi * qi) - c.qb * qi * qi;
    const float bpart = c.Pb * qi * qi * dp_dB;
    const float mpart = sc.cP * qi * (abi.x * c.Ax + abi.y * c.Ay + abi.z * c.Az);
    float4 xs = a.xs[k], vs = a.vs[k];
    xs.x += c.gx; xs.y += c.gy; xs.z += c.gz;
    vs.x += c.hx; vs.y += c.hy; vs.z += c.hz;
    a.xs[k] = xs;
    a.vs[k] = vs;
    return make_double2((double)bpart, (double)mpart);
}

// One particle of the adjoint force pass; returns its (dL/dB, dL/dmu) terms.
__device__ __forceinline__ double2 adj_force_one(const AdjArgs& a, const SceneConst& sc, uint32_t k) {
    const float4 pi = a.pos_s[k], vi = a.vel_s[k], abi = a.abar[k];
    AdjForceAcc c = {0, 0, 0, 0, 0, 0, abi.w, 0, 0, 0, 0};
    for_each_neighbor(a, sc, k, pi, [&](uint32_t j) {
        adj_force_pair(sc, pi, vi, abi, a.pos_s[j], a.vel_s[j], a.abar[j], c);
    });
    return adj_force_finish(a, sc, k, vi, abi, c);
}

// Staged forms of the adjoint force and density passes: the forward passes'
// staging applied to the reverse walk.  All nine windows and masks are formed
// first; per row the wave copies the union of its lanes' windows into its LDS
// slot (coalesced) and every lane walks its own mask over the slot; a wider
// union is read from global memory.  Same neighbours, same order, same
// arithmetic: the gathered kernels' results bit for bit.  Every lane of the
// wave takes part; inactive lanes have empty windows and store nothing.
constexpr uint32_t kStageA = 96;
static_assert(kStageA <= 128, "the copy into a slot takes two records per lane");
struct AdjForceSlot {
    float4 pos[kStageA];
    float4 vel[kStageA];
    float4 abar[kStageA];
};
struct AdjDensitySlot {
    float4 pos[kStageA];
    float rb[kStageA];
};

struct AdjRows {
    uint32_t lo[kRows], hi[kRows], mk[kRows];
};

__device__ __forceinline__ void adj_rows(const AdjArgs& a, const SceneConst& sc, uint32_t k, bool active,
                                         const float4& pi, AdjRows& r) {
    const __amdgpu_buffer_rsrc_t rcs = buffer_of(a.cstart, a.ncells + 1u);
    const RowGeom g = row_geom(sc, pi.x, pi.y, pi.z);
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        row_window(rcs, sc, g, row % 3 - 1, row / 3 - 1, r.lo[row], r.hi[row]);
        r.mk[row] = active ? a.mask[(size_t)row * a.n + k] : 0u;
        if (!active) r.lo[row] = r.hi[row] = 0u;
    }
    const uint32_t b = mask_bit(r.lo[4], r.hi[4], k);
    if (active && b < (uint32_t)kMaskBits) r.mk[4] &= ~(1u << b);
}

// kv: a particle of the tile (the lane's own when active) whose values the
// inactive lanes load in place of theirs.
__device__ __forceinline__ double2 adj_force_one_staged(const AdjArgs& a, const SceneConst& sc, uint32_t k,
                                                        bool active, uint32_t kv, AdjForceSlot& slot) {
    const float4 pi = a.pos_s[kv], vi = a.vel_s[kv], abi = a.abar[kv];
    AdjRows r;
    adj_rows(a, sc, k, active, pi, r);
    AdjForceAcc c = {0, 0, 0, 0, 0, 0, abi.w, 0, 0, 0, 0};
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int row = 0; row < kRows; ++row) {
        const bool has = r.hi[row] > r.lo[row];
        const uint32_t ulo = wave_umin(has ? r.lo[row] : 0xffffffffu);
        const uint32_t uhi = wave_umax(has ? r.hi[row] : 0u);
        if (ulo >= uhi) continue;
        if (uhi - ulo <= kStageA) {
            // Both halves of the copy (a union holds at most 2 x 64 records)
            // are loaded before either is stored, so their latencies overlap.
            const uint32_t span = uhi - ulo;
            const bool c0 = lane < span, c1 = lane + 64u < span;
            float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), v0 = p0, b0 = p0, p1 = p0, v1 = p0, b1 = p0;
            if (c0) {
                p0 = a.pos_s[ulo + lane];
                v0 = a.vel_s[ulo + lane];
                b0 = a.abar[ulo + lane];
            }
            if (c1) {
                p1 = a.pos_s[ulo + lane + 64u];
                v1 = a.vel_s[ulo + lane + 64u];
                b1 = a.abar[ulo + lane + 64u];
            }
            if (c0) {
                slot.pos[lane] = p0;
                slot.vel[lane] = v0;
                slot.abar[lane] = b0;
            }
            if (c1) {
                slot.pos[lane + 64u] = p1;
                slot.vel[lane + 64u] = v1;
                slot.abar[lane + 64u] = b1;
            }
            wave_lds_sync();
            const uint32_t e0 = row_first(r.lo[row], r.hi[row]) - ulo;
            row_neighbors_from(
                r.lo[row], r.hi[row], r.mk[row],
                [&](uint32_t j) {
                    const float4 pj = slot.pos[j - ulo];
                    return j != k && dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < sc.h2;
                },
                [&](uint32_t j) {
                    const uint32_t e = j - ulo;
                    adj_force_pair(sc, pi, vi, abi, slot.pos[e], slot.vel[e], slot.abar[e], c);
                },
                [&](uint32_t i) {
                    adj_force_pair(sc, pi, vi, abi, slot.pos[e0 + i], slot.vel[e0 + i], slot.abar[e0 + i], c);
                });
            wave_lds_sync();
        } else {
            row_neighbors(
                r.lo[row], r.hi[row], r.mk[row],
                [&](uint32_t j) {
                    const float4 pj = a.pos_s[j];
                    return j != k && dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < sc.h2;
                },
                [&](uint32_t j) { adj_force_pair(sc, pi, vi, abi, a.pos_s[j], a.vel_s[j], a.abar[j], c); });
        }
    }
    return active ? adj_force_finish(a, sc, k, vi, abi, c) : make_double2(0.0, 0.0);
}

// Eight waves per SIMD (64 VGPRs, a 36-byte spill): the backward pass was
// 3 % faster than without the bound (47.0 against 48.3 ms, 16 x 32k
// particles x 200 steps, with the walls' reverse still in this kernel).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void adj_force_kernel(AdjArgs a) {
    uint32_t t0, t1;
    tile_span(a.ntiles, t0, t1);
    // Per-tile (dL/dB, dL/dmu) of this step, reduced in double in a fixed
    // tree and added to the tile's running sum: deterministic, no atomics.
    __shared__ double2 red[256];
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        double2 part = make_double2(0.0, 0.0);
        if (k < sc.p_end) part = adj_force_one(a, sc, k);
        red[threadIdx.x] = part;
        __syncthreads();
        for (uint32_t o = blockDim.x >> 1; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                red[threadIdx.x].x += red[threadIdx.x + o].x;
                red[threadIdx.x].y += red[threadIdx.x + o].y;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double2 acc = a.tile_acc[t];
            acc.x += red[0].x;
            acc.y += red[0].y;
            a.tile_acc[t] = acc;
        }
        __syncthreads();
    }
}

// adj_force_kernel with the rows staged in LDS (4.5 KB per wave).
__global__ __launch_bounds__(256) void adj_force_staged_kernel(AdjArgs a) {
    uint32_t t0, t1;
    tile_span(a.ntiles, t0, t1);
    __shared__ double2 red[256];
    __shared__ AdjForceSlot s_slot[256 / 64];
    AdjForceSlot& slot = s_slot[threadIdx.x >> 6];
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        const bool active = k < sc.p_end;
        const double2 part = adj_force_one_staged(a, sc, k, active, active ? k : ti.y, slot);
        red[threadIdx.x] = part;
        __syncthreads();
        for (uint32_t o = blockDim.x >> 1; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                red[threadIdx.x].x += red[threadIdx.x + o].x;
                red[threadIdx.x].y += red[threadIdx.x + o].y;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double2 acc = a.tile_acc[t];
            acc.x += red[0].x;
            acc.y += red[0].y;
            a.tile_acc[t] = acc;
        }
        __syncthreads();
    }
}

// One neighbour's term of the adjoint density pass.
__device__ __forceinline__ void adj_density_pair(const SceneConst& sc, const float4& pi, float rbi, const float4& pj,
                                                 float rbj, float& gx, float& gy, float& gz) {
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    const float tt = sc.h2 - dist2(rx, ry, rz);
    const float s = -6.0f * sc.cD * tt * tt * (rbi + rbj);
    gx += s * rx;
    gy += s * ry;
    gz += s * rz;
}

// The walls' density reverse, and the adjoints of particle k scattered back
// to the step's input order.
__device__ __forceinline__ void adj_density_finish(const AdjArgs& a, const SceneConst& sc, uint32_t k,
                                                   const float4& pi, float rbi, float gx, float gy, float gz) {
    const float pk[3] = {pi.x, pi.y, pi.z};
    float gw[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float ul = (pk[d] - sc.bmin[d]) * sc.inv_h, uh = (sc.wmax[d] - pk[d]) * sc.inv_h;
        gw[d] = (rbi * sc.rho0) * (sc.inv_h * (wall_rho_du(ul) - wall_rho_du(uh)));
    }
    gx += gw[0];
    gy += gw[1];
    gz += gw[2];
    float4 xs = a.xs[k];
    xs.x += gx; xs.y += gy; xs.z += gz;
    const uint32_t src = a.perm[k];
    a.xbar[src] = xs;
    a.vbar[src] = a.vs[k];
}

// adj_density_kernel with the rows staged in LDS (1.9 KB per wave).
// Eight waves per SIMD: the bound fits it in 64 VGPRs.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void adj_density_staged_kernel(AdjArgs a) {
    uint32_t t0, t1;
    tile_span(a.ntiles, t0, t1);
    __shared__ AdjDensitySlot s_slot[256 / 64];
    AdjDensitySlot& slot = s_slot[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        const bool active = k < sc.p_end;
        const uint32_t kv = active ? k : ti.y;
        const float4 pi = a.pos_s[kv];
        const float rbi = a.rhobar[kv];
        AdjRows r;
        adj_rows(a, sc, k, active, pi, r);
        float gx = 0, gy = 0, gz = 0;
#pragma unroll
        for (int row = 0; row < kRows; ++row) {
            const bool has = r.hi[row] > r.lo[row];
            const uint32_t ulo = wave_umin(has ? r.lo[row] : 0xffffffffu);
            const uint32_t uhi = wave_umax(has ? r.hi[row] : 0u);
            if (ulo >= uhi) continue;
            if (uhi - ulo <= kStageA) {
                // Both halves of the copy are loaded before either is stored.
                const uint32_t span = uhi - ulo;
                const bool c0 = lane < span, c1 = lane + 64u < span;
                float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0;
                float q0 = 0.f, q1 = 0.f;
                if (c0) {
                    p0 = a.pos_s[ulo + lane];
                    q0 = a.rhobar[ulo + lane];
                }
                if (c1) {
                    p1 = a.pos_s[ulo + lane + 64u];
                    q1 = a.rhobar[ulo + lane + 64u];
                }
                if (c0) {
                    slot.pos[lane] = p0;
                    slot.rb[lane] = q0;
                }
                if (c1) {
                    slot.pos[lane + 64u] = p1;
                    slot.rb[lane + 64u] = q1;
                }
                wave_lds_sync();
                const uint32_t e0 = row_first(r.lo[row], r.hi[row]) - ulo;
                row_neighbors_from(
                    r.lo[row], r.hi[row], r.mk[row],
                    [&](uint32_t j) {
                        const float4 pj = slot.pos[j - ulo];
                        return j != k && dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < sc.h2;
                    },
                    [&](uint32_t j) {
                        const uint32_t e = j - ulo;
                        adj_density_pair(sc, pi, rbi, slot.pos[e], slot.rb[e], gx, gy, gz);
                    },
                    [&](uint32_t i) { adj_density_pair(sc, pi, rbi, slot.pos[e0 + i], slot.rb[e0 + i], gx, gy, gz); });
                wave_lds_sync();
            } else {
                row_neighbors(
                    r.lo[row], r.hi[row], r.mk[row],
                    [&](uint32_t j) {
                        const float4 pj = a.pos_s[j];
                        return j != k && dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < sc.h2;
                    },
                    [&](uint32_t j) { adj_density_pair(sc, pi, rbi, a.pos_s[j], a.rhobar[j], gx, gy, gz); });
            }
        }
        if (active) adj_density_finish(a, sc, k, pi, rbi, gx, gy, gz);
    }
}

__global__ void adj_density_kernel(AdjArgs a) {
    uint32_t t0, t1;
    tile_span(a.ntiles, t0, t1);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint2 ti = tile_at(a.tiles, t);
        const SceneConst sc = scene_at(a.scenes, ti.x);
        const uint32_t k = ti.y + threadIdx.x;
        if (k >= sc.p_end) continue;
        const float4 pi = a.pos_s[k];
        const float rbi = a.rhobar[k];
        float gx = 0, gy = 0, gz = 0;
        for_each_neighbor(a, sc, k, pi, [&](uint32_t j) {
            adj_density_pair(sc, pi, rbi, a.pos_s[j], a.rhobar[j], gx, gy, gz);
        });
        adj_density_finish(a, sc, k, pi, rbi, gx, gy, gz);
    }
}

// Per-scene totals of the per-tile accumulators: block s sums tiles
// [tile_first[s], tile_first[s + 1]) in a fixed order (threads stride, then
// a tree), so the result is the same run to run.
__global__ __launch_bounds__(256) void scene_tiles_kernel(const uint32_t* __restrict__ tile_first,
                                                          const double2* __restrict__ tile_acc,
                                                          double* __restrict__ acc_a, double* __restrict__ acc_b) {
    __shared__ double sa[256], sb[256];
    const uint32_t s = blockIdx.x;
    const uint32_t beg = tile_first[s], end = tile_first[s + 1];
    double va = 0.0, vb = 0.0;
    for (uint32_t i = beg + threadIdx.x; i < end; i += 256) {
        va += tile_acc[i].x;
        vb += tile_acc[i].y;
    }
    sa[threadIdx.x] = va;
    sb[threadIdx.x] = vb;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sa[threadIdx.x] += sa[threadIdx.x + o];
            sb[threadIdx.x] += sb[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        acc_a[s] = sa[0];
        acc_b[s] = sb[0];
    }
}

__global__ void gather3_kernel(uint32_t n, const float* __restrict__ g, const uint32_t* __restrict__ id,
                               float4* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t i = id[k];
    out[k] = make_float4(g[3 * i], g[3 * i + 1], g[3 * i + 2], 0.0f);
}

}  // namespace

hipError_t launch_adj_integrate(hipStream_t st, int block, uint32_t grid, const AdjArgs& a) {
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(adj_integrate_kernel, dim3(grid), dim3(block), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_adj_force(hipStream_t st, int block, uint32_t grid, const AdjArgs& a, int staged) {
    if (grid == 0) return hipSuccess;
    // __launch_bounds__(256), and the per-tile tree wants a power of two;
    // the staged form also wants whole waves.
    if (block > 256 || (block & (block - 1)) || (staged && block < 64)) return hipErrorInvalidValue;
    if (staged)
        hipLaunchKernelGGL(adj_force_staged_kernel, dim3(grid), dim3(block), 0, st, a);
    else
        hipLaunchKernelGGL(adj_force_kernel, dim3(grid), dim3(block), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_adj_density(hipStream_t st, int block, uint32_t grid, const AdjArgs& a, int staged) {
    if (grid == 0) return hipSuccess;
    if (staged && (block > 256 || block % 64)) return hipErrorInvalidValue;
    if (staged)
        hipLaunchKernelGGL(adj_density_staged_kernel, dim3(grid), dim3(block), 0, st, a);
    else
        hipLaunchKernelGGL(adj_density_kernel, dim3(grid), dim3(block), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_scene_tile_sums(hipStream_t st, uint32_t nscenes, const uint32_t* tile_first,
                                  const double2* tile_acc, double* acc_a, double* acc_b) {
    if (nscenes == 0) return hipSuccess;
    hipLaunchKernelGGL(scene_tiles_kernel, dim3(nscenes), dim3(256), 0, st, tile_first, tile_acc, acc_a, acc_b);
    return hipGetLastError();
}

hipError_t launch_gather3(hipStream_t st, uint32_t n, const float* g, const uint32_t* id, float4* out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather3_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, g, id, out);
    return hipGetLastError();
}

}  // namespace sph
