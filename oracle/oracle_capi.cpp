// TEST INFRASTRUCTURE ONLY: ctypes surface of the CPU restatement in
// sph_oracle.hpp.  Loaded by tests/ and by bench.py's CPU-baseline leg; the
// engine never links it.
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "sph_oracle.hpp"
#include "sph_adjoint_oracle.hpp"

using namespace orc;

namespace {
struct Handle {
    int fp64 = 0;
    Sim<float>  f;
    Sim<double> d;
};
}  // namespace

template <class T>
static void get_state_t(Sim<T>& s, double* pos3, double* vel3, uint32_t* ids) {
    for (uint32_t i = 0; i < s.n; ++i) {
        if (pos3) { pos3[3 * i] = s.x[i]; pos3[3 * i + 1] = s.y[i]; pos3[3 * i + 2] = s.z[i]; }
        if (vel3) { vel3[3 * i] = s.vx[i]; vel3[3 * i + 1] = s.vy[i]; vel3[3 * i + 2] = s.vz[i]; }
        if (ids) ids[i] = s.id[i];
    }
}

template <class T>
static void get_dbg_t(Sim<T>& s, uint32_t* key, uint32_t* perm, uint32_t* skey, double* rho,
                      double* acc3, uint32_t* sid) {
    for (uint32_t i = 0; i < s.n; ++i) {
        if (key) key[i] = s.key[i];
        if (perm) perm[i] = s.perm[i];
        if (skey) skey[i] = s.skey[i];
        if (rho) rho[i] = s.rho[i];
        if (acc3) { acc3[3 * i] = s.ax[i]; acc3[3 * i + 1] = s.ay[i]; acc3[3 * i + 2] = s.az[i]; }
        if (sid) sid[i] = s.id[i];
    }
}

extern "C" {

void* orc_new(int fp64, const Params* p) {
    Handle* h = new Handle();
    h->fp64 = fp64;
    if (fp64) h->d.set_params(*p); else h->f.set_params(*p);
    return h;
}

void orc_free(void* hp) { delete (Handle*)hp; }

void orc_set_threads(void* hp, int nt) {
    Handle* h = (Handle*)hp;
    h->f.nthreads = nt;
    h->d.nthreads = nt;
}

void orc_record_neighbors(void* hp, int on) {
    Handle* h = (Handle*)hp;
    h->f.record_nbrs = on != 0;
    h->d.record_nbrs = on != 0;
}

// State in the caller's order; ids may be null (0..n-1).
void orc_set_state(void* hp, uint32_t n, const double* pos3, const double* vel3, const uint32_t* ids) {
    Handle* h = (Handle*)hp;
    if (h->fp64) {
        h->d.set_state(n, pos3, vel3, ids);
    } else {
        std::vector<float> p(3 * (size_t)n), v(3 * (size_t)n);
        for (size_t i = 0; i < 3 * (size_t)n; ++i) {
            p[i] = (float)pos3[i];
            v[i] = vel3 ? (float)vel3[i] : 0.f;
        }
        h->f.set_state(n, p.data(), v.data(), ids);
    }
}

// Steps; returns wall seconds spent.
double orc_step(void* hp, int nsteps) {
    Handle* h = (Handle*)hp;
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nsteps; ++s) {
        if (h->fp64) h->d.step(); else h->f.step();
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

uint32_t orc_num_particles(void* hp) {
    Handle* h = (Handle*)hp;
    return h->fp64 ? h->d.n : h->f.n;
}

uint32_t orc_num_cells(void* hp) {
    Handle* h = (Handle*)hp;
    return h->fp64 ? h->d.c.ncells : h->f.c.ncells;
}

// Current state in state order (the sorted order of the last step).
// Step length of the last step (dt, or the CFL step).
double orc_last_dt(void* hp) {
    Handle* h = (Handle*)hp;
    return h->fp64 ? (double)h->d.c.dt : (double)h->f.c.dt;
}

void orc_get_state(void* hp, double* pos3, double* vel3, uint32_t* ids) {
    Handle* h = (Handle*)hp;
    if (h->fp64) get_state_t(h->d, pos3, vel3, ids); else get_state_t(h->f, pos3, vel3, ids);
}

// Derived constants: h, h2, inv_cs, mass, cD, cP, cV, inv_rho0, lo[3], hi[3] (14 doubles)
// and grid G[3] plus the x subdivision (4 ints).
void orc_get_constants(void* hp, double* out14, int* G3) {
    Handle* h = (Handle*)hp;
    auto put = [&](auto& c) {
        double v[14] = {(double)c.h, (double)c.h2, (double)c.inv_cs, (double)c.mass, (double)c.cD,
                        (double)c.cP, (double)c.cV, (double)c.inv_rho0,
                        (double)c.lo[0], (double)c.lo[1], (double)c.lo[2],
                        (double)c.hi[0], (double)c.hi[1], (double)c.hi[2]};
        for (int i = 0; i < 14; ++i) out14[i] = v[i];
        for (int i = 0; i < 3; ++i) G3[i] = c.G[i];
        G3[3] = c.xsub;
    };
    if (h->fp64) put(h->d.c); else put(h->f.c);
}

// Intermediates of the last step: input-state keys (state order), the sort
// permutation, sorted keys, density and acceleration (sorted order), and the
// ids in sorted order (= ids of the new state).
void orc_get_step_debug(void* hp, uint32_t* key, uint32_t* perm, uint32_t* skey, double* rho,
                        double* acc3, uint32_t* sid) {
    Handle* h = (Handle*)hp;
    if (h->fp64) get_dbg_t(h->d, key, perm, skey, rho, acc3, sid);
    else get_dbg_t(h->f, key, perm, skey, rho, acc3, sid);
}

// Cell table of the last step as (start, end) per cell, empty = (~0, 0).
void orc_get_cells(void* hp, uint32_t* start, uint32_t* end) {
    Handle* h = (Handle*)hp;
    auto& cs = h->fp64 ? h->d.cstart : h->f.cstart;
    for (size_t i = 0; i + 1 < cs.size(); ++i) {
        const bool empty = cs[i] == cs[i + 1];
        start[i] = empty ? ~0u : cs[i];
        end[i] = empty ? 0u : cs[i + 1];
    }
}

uint64_t orc_num_neighbor_entries(void* hp) {
    Handle* h = (Handle*)hp;
    return h->fp64 ? h->d.nbr.size() : h->f.nbr.size();
}

// Neighbour CSR of the last step in sorted index space (self included).
void orc_get_neighbors(void* hp, uint32_t* offsets, uint32_t* list) {
    Handle* h = (Handle*)hp;
    auto& o = h->fp64 ? h->d.noff : h->f.noff;
    auto& l = h->fp64 ? h->d.nbr : h->f.nbr;
    std::copy(o.begin(), o.end(), offsets);
    std::copy(l.begin(), l.end(), list);
}

// O(n^2) neighbour enumeration with the same predicate (float), self included.
// Returns the number of entries; when list is non-null fills offsets (n+1)
// and list (entries sorted ascending by index within each row).
uint64_t orc_bruteforce_neighbors(uint32_t n, const float* pos3, float h2, uint32_t* offsets,
                                  uint32_t* list) {
    uint64_t cnt = 0;
    if (offsets) offsets[0] = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float xi = pos3[3 * i], yi = pos3[3 * i + 1], zi = pos3[3 * i + 2];
        for (uint32_t j = 0; j < n; ++j) {
            const float d2 = dist2(xi - pos3[3 * j], yi - pos3[3 * j + 1], zi - pos3[3 * j + 2]);
            if (d2 < h2) {
                if (list) list[cnt] = j;
                ++cnt;
            }
        }
        if (offsets) offsets[i + 1] = (uint32_t)cnt;
    }
    return cnt;
}

// Float cell key of each position under the handle's constants.
void orc_cell_keys(void* hp, uint32_t n, const double* pos3, uint32_t* keys) {
    Handle* h = (Handle*)hp;
    for (uint32_t i = 0; i < n; ++i) {
        if (h->fp64)
            keys[i] = cell_key(h->d.c, pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]);
        else
            keys[i] = cell_key(h->f.c, (float)pos3[3 * i], (float)pos3[3 * i + 1], (float)pos3[3 * i + 2]);
    }
}

// ---------------------------------------------------------------------------
// Differentiable rollout (see sph_adjoint_oracle.hpp).
//
// Runs `steps` steps from (pos0, vel0) in the caller's order, returns the
// loss L = sum_i |x_i(T) - target_i|^2 * loss_scale over particles in the
// caller's (original-id) order.  When grads are non-null, also runs the
// hand-written reverse pass and returns dL/dB, dL/dmu and dL/dv0 (n x 3).
double orc_rollout_loss_grad(int fp64, const Params* p, uint32_t n, const double* pos0,
                             const double* vel0, const double* target, int steps,
                             double loss_scale, double* dB, double* dmu, double* dv0,
                             int nthreads, double B_override, double mu_override) {
    if (fp64)
        return rollout_loss_grad<double>(*p, n, pos0, vel0, target, steps, loss_scale, dB, dmu, dv0,
                                         nthreads, B_override, mu_override);
    return rollout_loss_grad<float>(*p, n, pos0, vel0, target, steps, loss_scale, dB, dmu, dv0,
                                    nthreads, B_override, mu_override);
}

}  // extern "C"
