// TEST INFRASTRUCTURE ONLY: reverse-mode restatement of the step in
// sph_oracle.hpp, used to pin the engine's hand-written adjoint.  The
// reference has no differentiable path (the adjoint is new here); this oracle is
// itself pinned by FP64 central differences and by torch float64 autograd of a
// dense pairwise restatement (tests/).
//
// Per pair (i, j), j != i, with r = x_i - x_j, d = |r|, hr = h - d,
// t = h^2 - d^2, phi = hr^2/d, phi'/d = -t/d^3, psi = hr, psi' = -1,
// P = p/rho^2, q = 1/rho, abar = dL/da:
//   xbar_i += cP (P_i+P_j) [phi (abar_i-abar_j) + (phi'/d) ((abar_i-abar_j).r) r]
//           + cV q_i q_j (psi'/d) ((abar_i-abar_j).(v_j-v_i)) r
//           - 6 cD t^2 (rhobar_i + rhobar_j) r
//   vbar_i += cV q_i q_j psi (abar_j - abar_i)
//   Pbar_i  = cP phi (abar_i-abar_j).r
//   qbar_i  = cV q_j psi (abar_i-abar_j).(v_j-v_i)
//   rhobar_i = Pbar_i (p' q^2 - 2 p q^3) - qbar_i q^2
//   Bbar += Pbar_i q_i^2 dp/dB ;  mubar += cP abar_i . q_i sum_j psi q_j (v_j - v_i)
// Walls (per axis d, u_lo = (x_d - box_min_d)/h, u_hi = (box_max_d - x_d)/h):
//   Pbar_i += cW abar_i,d (w_j(u_lo) - w_j(u_hi))
//   xbar_i,d += cW P_i abar_i,d (w_j'(u_lo) + w_j'(u_hi)) / h
//             + rhobar_i rho0 (w_rho'(u_lo) - w_rho'(u_hi)) / h
// The clamp passes xbar through only on unclamped axes and scales vbar by
// -restitution where the velocity was reflected.
#pragma once

#include "sph_oracle.hpp"

namespace orc {

template <class T>
struct Tape {
    std::vector<uint32_t> perm, noff, nbr;
    std::vector<T> sx, sy, sz, svx, svy, svz, rho, pterm, irho, ax, ay, az;
};

// dp/drho and dp/dB of the clamped Tait EOS.
template <class T>
inline void eos_grad(const Consts<T>& c, T rho, T& p, T& dp_drho, T& dp_dB) {
    const T x  = rho * c.inv_rho0;
    const T x2 = x * x;
    const T x4 = x2 * x2;
    const T x6 = x4 * x2;
    const T x7 = x6 * x;
    const T praw = c.B * (x7 - T(1));
    T s = T(1);
    if (praw < T(0)) s = c.attractive ? c.attr : T(0);
    p = praw * s;
    dp_drho = s * c.B * T(7) * x6 * c.inv_rho0;
    dp_dB = s * (x7 - T(1));
}

template <class T>
void reverse_step(const Consts<T>& c, const Tape<T>& tp, uint32_t n, std::vector<T>& xb,
                  std::vector<T>& vb, double& Bbar, double& mubar, int nt) {
    // xb, vb: 3n, sorted order of this step (= output state order) on entry;
    // input state order of this step on exit.
    std::vector<T> abar(3 * (size_t)n), xs(3 * (size_t)n), vs(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        const T a[3] = {tp.ax[i], tp.ay[i], tp.az[i]};
        const T x0[3] = {tp.sx[i], tp.sy[i], tp.sz[i]};
        const T v0[3] = {tp.svx[i], tp.svy[i], tp.svz[i]};
        for (int d = 0; d < 3; ++d) {
            const T v1 = v0[d] + c.dt * a[d];
            const T xt = x0[d] + c.dt * v1;
            T xbt = xb[3 * i + d], vb1 = vb[3 * i + d];
            if (xt < c.lo[d]) {
                xbt = T(0);
                if (v1 < T(0)) vb1 = -c.e * vb1;
            } else if (xt > c.hi[d]) {
                xbt = T(0);
                if (v1 > T(0)) vb1 = -c.e * vb1;
            }
            vb1 += c.dt * xbt;
            xs[3 * i + d] = xbt;
            vs[3 * i + d] = vb1;
            abar[3 * i + d] = c.dt * vb1;
        }
    }
    std::vector<T> rhobar(n);
    std::vector<double> bpart(n), mpart(n);
    const T cP = c.cP, cV = c.cV, cD = c.cD, h = c.h, h2 = c.h2;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const T xi = tp.sx[i], yi = tp.sy[i], zi = tp.sz[i];
        const T vxi = tp.svx[i], vyi = tp.svy[i], vzi = tp.svz[i];
        const T abx = abar[3 * i], aby = abar[3 * i + 1], abz = abar[3 * i + 2];
        const T Pi = tp.pterm[i], qi = tp.irho[i];
        T gx = 0, gy = 0, gz = 0, hx = 0, hy = 0, hz = 0, Pb = 0, qb = 0;
        T Avx = 0, Avy = 0, Avz = 0;
        for (uint32_t e = tp.noff[i]; e < tp.noff[i + 1]; ++e) {
            const uint32_t j = tp.nbr[e];
            if (j == i) continue;
            const T rx = xi - tp.sx[j], ry = yi - tp.sy[j], rz = zi - tp.sz[j];
            const T d2 = dist2(rx, ry, rz);
            const T dax = abx - abar[3 * j], day = aby - abar[3 * j + 1], daz = abz - abar[3 * j + 2];
            const T dvx = tp.svx[j] - vxi, dvy = tp.svy[j] - vyi, dvz = tp.svz[j] - vzi;
            const T qj = tp.irho[j];
            const T inv_r = d2 > T(0) ? T(1) / std::sqrt(d2) : T(0);
            const T dd = d2 * inv_r;
            const T hr = h - dd;
            const T t = h2 - d2;
            const T phi = hr * hr * inv_r;
            const T dphi_r = -t * inv_r * inv_r * inv_r;
            const T PP = cP * (Pi + tp.pterm[j]);
            const T dar = dax * rx + day * ry + daz * rz;
            const T dav = dax * dvx + day * dvy + daz * dvz;
            const T cq = cV * qi * qj;
            const T coefr = PP * dphi_r * dar - cq * inv_r * dav;  // psi' = -1
            gx += PP * phi * dax + coefr * rx;
            gy += PP * phi * day + coefr * ry;
            gz += PP * phi * daz + coefr * rz;
            hx -= cq * hr * dax;
            hy -= cq * hr * day;
            hz -= cq * hr * daz;
            Pb += cP * phi * dar;
            qb += cV * qj * hr * dav;
            Avx += hr * qj * dvx; Avy += hr * qj * dvy; Avz += hr * qj * dvz;
        }
        {
            const T pk[3] = {xi, yi, zi}, ab3[3] = {abx, aby, abz};
            T gw[3];
            for (int d = 0; d < 3; ++d) {
                const T ul = (pk[d] - c.wmin[d]) * c.inv_h, uh = (c.wmax[d] - pk[d]) * c.inv_h;
                Pb += c.cW * ab3[d] * (wall_j(ul) - wall_j(uh));
                gw[d] = ((c.cW * Pi) * ab3[d]) * (c.inv_h * (wall_j_du(ul) + wall_j_du(uh)));
            }
            gx += gw[0]; gy += gw[1]; gz += gw[2];
        }
        T p, dp_drho, dp_dB;
        eos_grad(c, tp.rho[i], p, dp_drho, dp_dB);
        rhobar[i] = Pb * (dp_drho * qi * qi - T(2) * p * qi * qi * qi) - qb * qi * qi;
        bpart[i] = (double)Pb * (double)qi * (double)qi * (double)dp_dB;
        mpart[i] = (double)cP * (double)qi * ((double)abx * Avx + (double)aby * Avy + (double)abz * Avz);
        xs[3 * i] += gx; xs[3 * i + 1] += gy; xs[3 * i + 2] += gz;
        vs[3 * i] += hx; vs[3 * i + 1] += hy; vs[3 * i + 2] += hz;
    }
    for (uint32_t i = 0; i < n; ++i) { Bbar += bpart[i]; mubar += mpart[i]; }
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const T xi = tp.sx[i], yi = tp.sy[i], zi = tp.sz[i];
        T gx = 0, gy = 0, gz = 0;
        for (uint32_t e = tp.noff[i]; e < tp.noff[i + 1]; ++e) {
            const uint32_t j = tp.nbr[e];
            if (j == i) continue;
            const T rx = xi - tp.sx[j], ry = yi - tp.sy[j], rz = zi - tp.sz[j];
            const T t = h2 - dist2(rx, ry, rz);
            const T s = T(-6) * cD * t * t * (rhobar[i] + rhobar[j]);
            gx += s * rx; gy += s * ry; gz += s * rz;
        }
        {
            const T pk[3] = {xi, yi, zi};
            T gw[3];
            for (int d = 0; d < 3; ++d) {
                const T ul = (pk[d] - c.wmin[d]) * c.inv_h, uh = (c.wmax[d] - pk[d]) * c.inv_h;
                gw[d] = (rhobar[i] * c.rho0) * (c.inv_h * (wall_rho_du(ul) - wall_rho_du(uh)));
            }
            gx += gw[0]; gy += gw[1]; gz += gw[2];
        }
        xs[3 * i] += gx; xs[3 * i + 1] += gy; xs[3 * i + 2] += gz;
    }
    // Back to the input state order of this step.
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t s = tp.perm[k];
        for (int d = 0; d < 3; ++d) {
            xb[3 * s + d] = xs[3 * k + d];
            vb[3 * s + d] = vs[3 * k + d];
        }
    }
}

// B_override / mu_override (when not NaN) replace the float parameters with
// values held in T, so FP64 central differences can use steps below float
// resolution.
template <class T>
double rollout_loss_grad(const Params& p, uint32_t n, const double* pos0, const double* vel0,
                         const double* target, int steps, double loss_scale, double* dB,
                         double* dmu, double* dv0, int nt, double B_override = NAN,
                         double mu_override = NAN) {
    Sim<T> s;
    s.set_params(p);
    if (!std::isnan(B_override)) s.c.B = (T)B_override;
    if (!std::isnan(mu_override)) s.c.cV = (T)(mu_override * (double)s.c.cP);
    s.nthreads = nt;
    std::vector<T> ps(3 * (size_t)n), vs(3 * (size_t)n);
    for (size_t i = 0; i < 3 * (size_t)n; ++i) { ps[i] = (T)pos0[i]; vs[i] = vel0 ? (T)vel0[i] : T(0); }
    s.set_state(n, ps.data(), vs.data(), nullptr);
    const bool grad = dv0 || dB || dmu;
    s.record_nbrs = grad;
    std::vector<Tape<T>> tape(grad ? steps : 0);
    for (int t = 0; t < steps; ++t) {
        s.step();
        if (grad) {
            Tape<T>& tp = tape[t];
            tp.perm = s.perm; tp.noff = s.noff; tp.nbr = s.nbr;
            tp.sx = s.sx; tp.sy = s.sy; tp.sz = s.sz;
            tp.svx = s.svx; tp.svy = s.svy; tp.svz = s.svz;
            tp.rho = s.rho; tp.pterm = s.pterm; tp.irho = s.irho;
            tp.ax = s.ax; tp.ay = s.ay; tp.az = s.az;
        }
    }
    double L = 0;
    std::vector<T> xb(3 * (size_t)n, T(0)), vb(3 * (size_t)n, T(0));
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t id = s.id[k];
        const double q[3] = {(double)s.x[k], (double)s.y[k], (double)s.z[k]};
        for (int d = 0; d < 3; ++d) {
            const double df = q[d] - target[3 * id + d];
            L += df * df;
            xb[3 * k + d] = (T)(2.0 * loss_scale * df);
        }
    }
    L *= loss_scale;
    if (!grad) return L;
    double Bbar = 0, mubar = 0;
    for (int t = steps - 1; t >= 0; --t) reverse_step(s.c, tape[t], n, xb, vb, Bbar, mubar, nt);
    if (dB) *dB = Bbar;
    if (dmu) *dmu = mubar;
    if (dv0)
        for (size_t i = 0; i < 3 * (size_t)n; ++i) dv0[i] = (double)vb[i];
    return L;
}

}  // namespace orc
