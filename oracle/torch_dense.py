"""TEST INFRASTRUCTURE ONLY: dense all-pairs restatement of the SPH step in
torch float64, differentiated by autograd.  It pins the hand-written
reverse pass (oracle/sph_adjoint_oracle.hpp and the engine's adjoint) at
small particle counts; it shares no code with either.

Same formulation as oracle/sph_oracle.hpp: poly6 density (self included),
clamped Tait EOS (exponent 7), symmetric spiky pressure force, Mueller
viscosity, gravity, the walls' half-space density and pressure terms,
symplectic Euler, box clamp with restitution.  The
neighbour predicate d^2 < h^2 is a constant mask (it is not differentiated;
every term vanishes at d = h).
"""
from __future__ import annotations

import torch


def make_consts(c: dict, params) -> dict:
    """Constants in float64 from the CPU oracle's FP64 constants."""
    return dict(h=float(c["h"]), h2=float(c["h2"]), cD=float(c["cD"]), cP=float(c["cP"]),
                inv_rho0=float(c["inv_rho0"]), lo=[float(x) for x in c["lo"]],
                hi=[float(x) for x in c["hi"]], dt=float(params.dt),
                e=float(params.boundary_restitution),
                attr=float(params.attractive_pressure_ratio) if params.use_attractive_pressure else 0.0,
                g=[float(x) for x in params.gravity], rho0=float(params.rest_density),
                wmin=[float(x) for x in params.box_min], wmax=[float(x) for x in params.box_max])


def _wall_rho(u):
    # (315/256) int_u^1 (1 - t^2)^4 dt, zero beyond the support.
    w = (1.0 - u).clamp(min=0.0)
    return w ** 5 * (128.0 + 325.0 * u + 345.0 * u ** 2 + 175.0 * u ** 3 + 35.0 * u ** 4) / 256.0


def _wall_j(u):
    # int_u^1 (1 - s)^2 (s^2 - u^2) ds, zero beyond the support.
    w = (1.0 - u).clamp(min=0.0)
    return w ** 4 * (1.0 + 4.0 * u) / 30.0


def step(x, v, B, mu, k):
    n = x.shape[0]
    r = x[:, None, :] - x[None, :, :]
    d2 = (r * r).sum(-1)
    mask = (d2 < k["h2"]).detach()
    eye = torch.eye(n, dtype=torch.bool)
    t = torch.where(mask, k["h2"] - d2, torch.zeros_like(d2))
    wmin = torch.tensor(k["wmin"], dtype=x.dtype)
    wmax = torch.tensor(k["wmax"], dtype=x.dtype)
    ul = (x - wmin) / k["h"]
    uh = (wmax - x) / k["h"]
    rho = k["cD"] * (t * t * t).sum(1) + k["rho0"] * (_wall_rho(ul) + _wall_rho(uh)).sum(1)
    xr = rho * k["inv_rho0"]
    praw = B * (xr ** 7 - 1.0)
    p = torch.where(praw < 0, praw * k["attr"], praw)
    P = p / (rho * rho)
    q = 1.0 / rho
    off = mask & ~eye
    dsafe = torch.sqrt(torch.where(off, d2, torch.ones_like(d2)))
    hr = k["h"] - dsafe
    phi = torch.where(off, hr * hr / dsafe, torch.zeros_like(d2))
    ap = k["cP"] * (((P[:, None] + P[None, :]) * phi)[..., None] * r).sum(1)
    psi = torch.where(off, hr, torch.zeros_like(d2))
    cV = mu * k["cP"]
    av = cV * q[:, None] * ((psi * q[None, :])[..., None] * (v[None, :, :] - v[:, None, :])).sum(1)
    g = torch.tensor(k["g"], dtype=x.dtype)
    aw = (90.0 * k["rho0"] / k["h"]) * P[:, None] * (_wall_j(ul) - _wall_j(uh))
    a = ap + av + g + aw
    v1 = v + k["dt"] * a
    x1 = x + k["dt"] * v1
    lo = torch.tensor(k["lo"], dtype=x.dtype)
    hi = torch.tensor(k["hi"], dtype=x.dtype)
    below = x1 < lo
    above = (x1 > hi) & ~below
    x1 = torch.where(below, lo.expand_as(x1), x1)
    x1 = torch.where(above, hi.expand_as(x1), x1)
    v1 = torch.where((below & (v1 < 0)) | (above & (v1 > 0)), -k["e"] * v1, v1)
    return x1, v1


def rollout_loss(x0, v0, B, mu, k, steps, target, scale):
    x, v = x0, v0
    for _ in range(steps):
        x, v = step(x, v, B, mu, k)
    return scale * ((x - target) ** 2).sum()


def loss_and_grads(pos0, vel0, target, params, consts, steps, scale=None):
    """Loss and autograd dL/dB, dL/dmu, dL/dv0 (float64)."""
    k = make_consts(consts, params)
    x0 = torch.tensor(pos0, dtype=torch.float64)
    v0 = torch.tensor(vel0, dtype=torch.float64, requires_grad=True)
    B = torch.tensor(float(params.pressure_stiffness), dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(float(params.viscosity), dtype=torch.float64, requires_grad=True)
    tg = torch.tensor(target, dtype=torch.float64)
    sc = 1.0 / (3 * x0.shape[0]) if scale is None else scale
    L = rollout_loss(x0, v0, B, mu, k, steps, tg, sc)
    L.backward()
    return L.item(), B.grad.item(), mu.grad.item(), v0.grad.numpy().copy()
