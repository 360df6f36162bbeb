// TEST INFRASTRUCTURE ONLY. Not part of the product and never linked into it.
//
// CPU restatement of one weakly-compressible SPH step, used as the parity
// checker for the HIP engine and as the timed CPU baseline.
//
// The reference (ttnghia/SimpleFluid) drives its solver through
//   Source/Simulator.cpp:42   m_SPHSolver->makeReady();
//   Source/Simulator.cpp:49   frameTime += m_SPHSolver->advanceFrame();
// but the solver itself (Banana ParticleSolvers/SPH/SPHSolver.h, included at
// Include/QtSPHSolver.h:22) is not vendored, so its arithmetic is unknown.
// This file restates the formulation the engine defines instead:
//   * support h = kernelRadius (Source/Controller.cpp:55: kernelRadius = 2/res),
//     particle radius r = h/4, lattice spacing 2r (Source/SceneManager.cpp:45);
//   * mass m = rho0 (2r)^3;
//   * rho_i = m sum_j W_poly6(|x_i - x_j|), self term included (Mueller 2003);
//   * Tait EOS p = B((rho/rho0)^7 - 1), p < 0 clamped to 0 unless the
//     attractive-pressure option (Source/Controller.cpp:61) scales it;
//   * symmetric pressure force with the spiky gradient, Mueller viscosity
//     (mu/rho_i) sum m (v_j - v_i)/rho_j lap W_visc, gravity;
//   * walls: each box face is the boundary of a half-space of fluid at rest
//     density; a particle at distance d < h from a face gets that half-space's
//     poly6 density, rho0 w_rho(d/h), and the pressure force of its particles
//     taken at the particle's own p/rho^2 (mirror), n 2 (p/rho^2) (45 rho0/h)
//     w_j(d/h), n pointing into the box, with
//       w_rho(u) = (1-u)^5 (128 + 325u + 345u^2 + 175u^3 + 35u^4) / 256
//                = (315/256) int_u^1 (1 - t^2)^4 dt,
//       w_j(u)   = (1-u)^4 (1 + 4u) / 30 = int_u^1 (1-s)^2 (s^2 - u^2) ds
//     (the half-space integrals of W_poly6 and of the spiky pressure term);
//   * symplectic Euler and a box clamp with restitution on the normal velocity.
// Every op sequence that decides a discrete outcome (cell index, neighbour
// predicate) is written exactly as the engine writes it, so cells and
// neighbour sets are comparable bit for bit.  Float sums run in the canonical
// neighbour order: rows dz = -1..1 (outer), dy = -1..1 (inner) of y/z cells,
// each row a contiguous sorted range of x cells, ascending sorted index.
// Cells are cs = h(1 + 2^-10) in y and z and cs/xsub in x; a row here spans
// the x cells within xsub of the particle's own (a superset of its
// neighbours, filtered by the predicate), so the set and the order of the
// neighbours match the engine's narrower per-row windows exactly.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// Same field meaning as the engine's sph_params.
struct Params {
    float particle_radius;
    float kernel_radius;
    float rest_density;
    float pressure_stiffness;
    float viscosity;
    float boundary_restitution;
    int   use_attractive_pressure;
    float attractive_pressure_ratio;
    float gravity[3];
    float dt;
    float box_min[3];
    float box_max[3];
    int   x_subdivision;   // x cells per y/z cell (1, 2 or 4)
    float cfl_number;      // 0: fixed dt; > 0: CFL step capped at dt
};

template <class T>
struct Consts {
    T h, h2, inv_cs, inv_csx, mass, cD, cP, cV, inv_rho0, B, e, attr, dt, r;
    T cfl, c0, dt_max;  // CFL number, EOS sound speed sqrt(7 B / rho0), dt cap
    T rho0, inv_h, cW;  // wall terms: rest density, 1/h, 2 * 45 rho0 / h
    T wmin[3], wmax[3]; // box faces
    int xsub;
    T g[3], bmin[3], lo[3], hi[3];
    int G[3];
    bool attractive;
    uint32_t ncells;
};

// Derived constants. Kernel normalisations are formed in double from the
// float inputs and rounded once to T.  Cell size is h(1 + 2^-10): a pair
// closer than h can then never land two cells apart through rounding of
// (x - box_min) * inv_cs for grids up to 1024 cells per axis.
template <class T>
inline Consts<T> derive(const Params& p) {
    Consts<T> c{};
    const T r   = (T)p.particle_radius;
    const T sp  = T(2) * r;
    c.r         = r;
    c.mass      = (T)p.rest_density * (sp * sp * sp);
    c.h         = (T)p.kernel_radius;
    c.h2        = c.h * c.h;
    const double hd = (double)(T)p.kernel_radius;
    const double h3 = hd * hd * hd, h6 = h3 * h3, h9 = h6 * h3;
    const double poly6 = 315.0 / (64.0 * M_PI * h9);
    const double spiky = 45.0 / (M_PI * h6);
    c.cD        = (T)((double)c.mass * poly6);
    c.cP        = (T)((double)c.mass * spiky);
    c.cV        = (T)((double)(T)p.viscosity * (double)c.mass * spiky);
    c.inv_rho0  = (T)(1.0 / (double)(T)p.rest_density);
    c.inv_cs    = (T)(1.0 / (hd * (1.0 + 1.0 / 1024.0)));
    c.xsub      = p.x_subdivision > 0 ? p.x_subdivision : 1;
    c.inv_csx   = c.inv_cs * (T)c.xsub;
    c.B         = (T)p.pressure_stiffness;
    c.e         = (T)p.boundary_restitution;
    c.attr      = (T)p.attractive_pressure_ratio;
    c.attractive = p.use_attractive_pressure != 0;
    c.dt        = (T)p.dt;
    c.dt_max    = (T)p.dt;
    c.cfl       = (T)p.cfl_number;
    c.c0        = (T)std::sqrt(7.0 * (double)p.pressure_stiffness / (double)p.rest_density);
    c.ncells    = 1;
    c.rho0      = (T)p.rest_density;
    c.inv_h     = (T)(1.0 / hd);
    c.cW        = (T)(90.0 * (double)(T)p.rest_density / hd);
    for (int k = 0; k < 3; ++k) {
        c.g[k]    = (T)p.gravity[k];
        c.bmin[k] = (T)p.box_min[k];
        c.wmin[k] = (T)p.box_min[k];
        c.wmax[k] = (T)p.box_max[k];
        const T ext = (T)p.box_max[k] - (T)p.box_min[k];
        int G = (int)std::ceil(ext * (k == 0 ? c.inv_csx : c.inv_cs));
        c.G[k] = G < 1 ? 1 : G;
        c.lo[k] = (T)p.box_min[k] + r;
        c.hi[k] = (T)p.box_max[k] - r;
        c.ncells *= (uint32_t)c.G[k];
    }
    return c;
}

template <class T>
inline int cell_coord(T x, T bmin, T inv_cs, int G) {
    T f = (x - bmin) * inv_cs;
    f = std::fmin(std::fmax(f, T(0)), (T)(G - 1));
    return (int)f;
}

template <class T>
inline uint32_t cell_key(const Consts<T>& c, T x, T y, T z) {
    const int cx = cell_coord(x, c.bmin[0], c.inv_csx, c.G[0]);
    const int cy = cell_coord(y, c.bmin[1], c.inv_cs, c.G[1]);
    const int cz = cell_coord(z, c.bmin[2], c.inv_cs, c.G[2]);
    return ((uint32_t)cz * (uint32_t)c.G[1] + (uint32_t)cy) * (uint32_t)c.G[0] + (uint32_t)cx;
}

template <class T>
inline T dist2(T dx, T dy, T dz) {
    return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
}

// Tait EOS with exponent 7 by a fixed multiply chain: x^2, x^4, (x^4 x^2) x.
template <class T>
inline T eos(const Consts<T>& c, T rho) {
    const T x  = rho * c.inv_rho0;
    const T x2 = x * x;
    const T x4 = x2 * x2;
    const T x7 = (x4 * x2) * x;
    T p = c.B * (x7 - T(1));
    if (p < T(0)) p = c.attractive ? p * c.attr : T(0);
    return p;
}

// Wall profiles (u = distance to the face / h); see the file comment.  The
// engine evaluates them with the same operations.
template <class T>
inline T wall_rho(T u) {
    if (!(u < T(1))) return T(0);
    u = std::fmax(u, T(0));
    const T w = T(1) - u;
    const T w2 = w * w;
    const T w5 = (w2 * w2) * w;
    const T poly = T(128) + u * (T(325) + u * (T(345) + u * (T(175) + u * T(35))));
    return (w5 * poly) * T(0.00390625);
}
template <class T>
inline T wall_j(T u) {
    if (!(u < T(1))) return T(0);
    u = std::fmax(u, T(0));
    const T w = T(1) - u;
    const T w2 = w * w;
    return ((w2 * w2) * (T(1) + T(4) * u)) * T(1.0 / 30.0);
}
// Their derivatives in u (the reverse pass).
template <class T>
inline T wall_rho_du(T u) {
    if (!(u < T(1))) return T(0);
    u = std::fmax(u, T(0));
    const T s = T(1) - u * u;
    const T s2 = s * s;
    return -T(1.23046875) * (s2 * s2);
}
template <class T>
inline T wall_j_du(T u) {
    if (!(u < T(1))) return T(0);
    u = std::fmax(u, T(0));
    const T w = T(1) - u;
    return -T(2.0 / 3.0) * u * ((w * w) * w);
}

template <class T>
struct Sim {
    Params   prm{};
    Consts<T> c{};
    uint32_t n = 0;
    // State, in the engine's state order (the sorted order of the last step).
    std::vector<T> x, y, z, vx, vy, vz;
    std::vector<uint32_t> id;
    // Intermediates of the last step.
    std::vector<uint32_t> key;         // state order, key of the input state
    std::vector<uint32_t> perm;        // sorted index -> state index
    std::vector<uint32_t> skey;        // sorted keys
    std::vector<uint32_t> cstart;      // C + 1: cstart[c] = number of particles with key < c
    std::vector<T> sx, sy, sz, svx, svy, svz;  // reordered input state
    std::vector<T> rho, pterm, irho, ax, ay, az; // sorted order
    std::vector<uint32_t> noff, nbr;   // neighbour CSR in sorted index space
    bool record_nbrs = false;
    int nthreads = 1;

    void set_params(const Params& p) {
        prm = p;
        c = derive<T>(p);
    }

    void set_state(uint32_t count, const T* pos3, const T* vel3, const uint32_t* ids) {
        n = count;
        x.resize(n); y.resize(n); z.resize(n); vx.resize(n); vy.resize(n); vz.resize(n); id.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            x[i] = pos3[3 * i]; y[i] = pos3[3 * i + 1]; z[i] = pos3[3 * i + 2];
            vx[i] = vel3 ? vel3[3 * i] : T(0);
            vy[i] = vel3 ? vel3[3 * i + 1] : T(0);
            vz[i] = vel3 ? vel3[3 * i + 2] : T(0);
            id[i] = ids ? ids[i] : i;
        }
    }

    // Stable counting sort of the state indices by key.  Chunks are counted
    // per thread and offsets laid out cell-major, thread-minor, so the result
    // equals a serial stable sort.
    void sort_keys() {
        const uint32_t C = c.ncells;
        perm.assign(n, 0);
        skey.assign(n, 0);
        int T_ = 1;
#ifdef _OPENMP
        T_ = std::min(16, std::max(1, nthreads));  // T_ x C counters
#endif
        if (T_ == 1 || n < 65536) {
            std::vector<uint32_t> cnt((size_t)C + 1, 0);
            for (uint32_t i = 0; i < n; ++i) cnt[key[i] + 1]++;
            for (uint32_t k = 0; k < C; ++k) cnt[k + 1] += cnt[k];
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t d = cnt[key[i]]++;
                perm[d] = i;
                skey[d] = key[i];
            }
            return;
        }
        std::vector<uint32_t> cnt((size_t)T_ * C, 0);
        const uint32_t chunk = (n + T_ - 1) / T_;
#pragma omp parallel num_threads(T_)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            const uint32_t b = std::min(n, chunk * t), e = std::min(n, b + chunk);
            uint32_t* my = &cnt[(size_t)t * C];
            for (uint32_t i = b; i < e; ++i) my[key[i]]++;
#pragma omp barrier
#pragma omp single
            {
                uint32_t run = 0;
                for (uint32_t k = 0; k < C; ++k)
                    for (int tt = 0; tt < T_; ++tt) {
                        uint32_t& v = cnt[(size_t)tt * C + k];
                        const uint32_t cnt_v = v;
                        v = run;
                        run += cnt_v;
                    }
            }
            for (uint32_t i = b; i < e; ++i) {
                const uint32_t d = my[key[i]]++;
                perm[d] = i;
                skey[d] = key[i];
            }
        }
    }

    void row_range(int cx, int y_, int z_, uint32_t& lo, uint32_t& hi) const {
        lo = 0; hi = 0;
        if (y_ < 0 || y_ >= c.G[1] || z_ < 0 || z_ >= c.G[2]) return;
        const int x0 = std::max(cx - c.xsub, 0), x1 = std::min(cx + c.xsub, c.G[0] - 1);
        const uint32_t base = ((uint32_t)z_ * (uint32_t)c.G[1] + (uint32_t)y_) * (uint32_t)c.G[0];
        lo = cstart[base + x0];
        hi = cstart[base + x1 + 1];
    }

    void coords(T x_, T y_, T z_, int& cx, int& cy, int& cz) const {
        cx = cell_coord(x_, c.bmin[0], c.inv_csx, c.G[0]);
        cy = cell_coord(y_, c.bmin[1], c.inv_cs, c.G[1]);
        cz = cell_coord(z_, c.bmin[2], c.inv_cs, c.G[2]);
    }

    // Adaptive step (cfl > 0): dt = min(dt_max, (cfl h) / (c0 + max|v|)) from
    // the velocities the step starts from, in this op order.  Engine-defined
    // (the reference's advanceFrame returns a variable step,
    // Source/Simulator.cpp:46-51, by a rule not in the repository).
    void cfl_step() {
        if (!(c.cfl > T(0))) return;
        T vmax2 = T(0);
        for (uint32_t i = 0; i < n; ++i) {
            const T v2 = dist2(vx[i], vy[i], vz[i]);
            if (v2 > vmax2) vmax2 = v2;
        }
        c.dt = std::fmin(c.dt_max, (c.cfl * c.h) / (c.c0 + std::sqrt(vmax2)));
    }

    void step() {
        const int nt = std::max(1, nthreads);
        (void)nt;
        cfl_step();
        key.resize(n);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (long long i = 0; i < (long long)n; ++i) key[i] = cell_key(c, x[i], y[i], z[i]);
        sort_keys();
        // Reorder.
        sx.resize(n); sy.resize(n); sz.resize(n); svx.resize(n); svy.resize(n); svz.resize(n);
        std::vector<uint32_t> sid(n);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (long long k = 0; k < (long long)n; ++k) {
            const uint32_t s = perm[k];
            sx[k] = x[s]; sy[k] = y[s]; sz[k] = z[s];
            svx[k] = vx[s]; svy[k] = vy[s]; svz[k] = vz[s];
            sid[k] = id[s];
        }
        // Cell-start table: counts, exclusive prefix sum.
        cstart.assign((size_t)c.ncells + 1, 0u);
        for (uint32_t k = 0; k < n; ++k) cstart[skey[k] + 1]++;
        for (uint32_t q = 0; q < c.ncells; ++q) cstart[q + 1] += cstart[q];
        // Density + EOS.
        rho.resize(n); pterm.resize(n); irho.resize(n);
        const T h2 = c.h2;
#pragma omp parallel for num_threads(nt) schedule(static)
        for (long long k = 0; k < (long long)n; ++k) {
            const T xi = sx[k], yi = sy[k], zi = sz[k];
            int cx, cy, cz;
            coords(xi, yi, zi, cx, cy, cz);
            T sum = 0;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy) {
                    uint32_t lo, hi;
                    row_range(cx, cy + dy, cz + dz, lo, hi);
                    for (uint32_t j = lo; j < hi; ++j) {
                        const T d2 = dist2(xi - sx[j], yi - sy[j], zi - sz[j]);
                        if (d2 < h2) {
                            const T t = h2 - d2;
                            sum = std::fma(t * t, t, sum);
                        }
                    }
                }
            T wr = T(0);
            const T pk[3] = {xi, yi, zi};
            for (int d = 0; d < 3; ++d) {
                wr += wall_rho((pk[d] - c.wmin[d]) * c.inv_h);
                wr += wall_rho((c.wmax[d] - pk[d]) * c.inv_h);
            }
            const T r = c.cD * sum + c.rho0 * wr;
            rho[k] = r;
            const T p = eos(c, r);
            const T ir = T(1) / r;
            irho[k] = ir;
            pterm[k] = (p * ir) * ir;
        }
        // Neighbour CSR (test only).
        if (record_nbrs) {
            noff.assign((size_t)n + 1, 0);
            std::vector<std::vector<uint32_t>> tmp(n);
#pragma omp parallel for num_threads(nt) schedule(static)
            for (long long k = 0; k < (long long)n; ++k) {
                const T xi = sx[k], yi = sy[k], zi = sz[k];
                int cx, cy, cz;
                coords(xi, yi, zi, cx, cy, cz);
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy) {
                        uint32_t lo, hi;
                        row_range(cx, cy + dy, cz + dz, lo, hi);
                        for (uint32_t j = lo; j < hi; ++j)
                            if (dist2(xi - sx[j], yi - sy[j], zi - sz[j]) < h2) tmp[k].push_back(j);
                    }
            }
            for (uint32_t k = 0; k < n; ++k) noff[k + 1] = noff[k] + (uint32_t)tmp[k].size();
            nbr.resize(noff[n]);
            for (uint32_t k = 0; k < n; ++k) std::copy(tmp[k].begin(), tmp[k].end(), nbr.begin() + noff[k]);
        }
        // Forces + integration + clamp.
        ax.resize(n); ay.resize(n); az.resize(n);
        std::vector<T> nx(n), ny(n), nz(n), nvx(n), nvy(n), nvz(n);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (long long k = 0; k < (long long)n; ++k) {
            const T xi = sx[k], yi = sy[k], zi = sz[k];
            const T vxi = svx[k], vyi = svy[k], vzi = svz[k];
            const T pti = pterm[k];
            int cx, cy, cz;
            coords(xi, yi, zi, cx, cy, cz);
            T apx = 0, apy = 0, apz = 0, avx = 0, avy = 0, avz = 0;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy) {
                    uint32_t lo, hi;
                    row_range(cx, cy + dy, cz + dz, lo, hi);
                    for (uint32_t j = lo; j < hi; ++j) {
                        if (j == (uint32_t)k) continue;
                        const T dx = xi - sx[j], dy_ = yi - sy[j], dz_ = zi - sz[j];
                        const T d2 = dist2(dx, dy_, dz_);
                        if (!(d2 < h2)) continue;
                        const T inv_r = d2 > T(0) ? T(1) / std::sqrt(d2) : T(0);
                        const T rr = d2 * inv_r;
                        const T hr = c.h - rr;
                        const T w = hr * hr;
                        const T s = ((pti + pterm[j]) * w) * inv_r;
                        apx = std::fma(s, dx, apx);
                        apy = std::fma(s, dy_, apy);
                        apz = std::fma(s, dz_, apz);
                        const T vv = hr * irho[j];
                        avx = std::fma(vv, svx[j] - vxi, avx);
                        avy = std::fma(vv, svy[j] - vyi, avy);
                        avz = std::fma(vv, svz[j] - vzi, avz);
                    }
                }
            const T cvi = c.cV * irho[k];
            const T cwi = c.cW * pti;
            const T pk[3] = {xi, yi, zi};
            T aw[3];
            for (int d = 0; d < 3; ++d)
                aw[d] = cwi * (wall_j((pk[d] - c.wmin[d]) * c.inv_h) - wall_j((c.wmax[d] - pk[d]) * c.inv_h));
            const T a0 = ((c.cP * apx + cvi * avx) + c.g[0]) + aw[0];
            const T a1 = ((c.cP * apy + cvi * avy) + c.g[1]) + aw[1];
            const T a2 = ((c.cP * apz + cvi * avz) + c.g[2]) + aw[2];
            ax[k] = a0; ay[k] = a1; az[k] = a2;
            T v[3] = {vxi + c.dt * a0, vyi + c.dt * a1, vzi + c.dt * a2};
            T p[3] = {xi + c.dt * v[0], yi + c.dt * v[1], zi + c.dt * v[2]};
            for (int d = 0; d < 3; ++d) {
                if (p[d] < c.lo[d]) {
                    p[d] = c.lo[d];
                    if (v[d] < T(0)) v[d] = -c.e * v[d];
                } else if (p[d] > c.hi[d]) {
                    p[d] = c.hi[d];
                    if (v[d] > T(0)) v[d] = -c.e * v[d];
                }
            }
            nx[k] = p[0]; ny[k] = p[1]; nz[k] = p[2];
            nvx[k] = v[0]; nvy[k] = v[1]; nvz[k] = v[2];
        }
        x.swap(nx); y.swap(ny); z.swap(nz); vx.swap(nvx); vy.swap(nvy); vz.swap(nvz);
        id.swap(sid);
    }
};

}  // namespace orc
