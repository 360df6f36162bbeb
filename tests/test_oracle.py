"""Pins of the CPU oracle (oracle/), each chosen so a plausible mistake in the
restatement fails: reference known answers, brute force, closed forms,
conservation laws, finite differences and an independent autograd path."""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def lattice(n3, r=0.005, origin=(0.0, 0.0, 0.0), jitter=0.0, seed=0):
    nx, ny, nz = n3
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.float64)
    p = np.asarray(origin) + r + 2 * r * ijk
    if jitter:
        p += np.random.default_rng(seed).uniform(-jitter, jitter, p.shape)
    return p.astype(np.float32).astype(np.float64)


# ---------------------------------------------------------------- scenes ----

def test_reference_scene_counts_match_screenshots(oracle_lib):
    from oracle import scenes
    g = json.load(open(os.path.join(HERE, "golden", "scene_counts.json")))
    for name, n in g["res24"].items():
        assert len(scenes.reference_scene(name, 24)) == n, name


def test_radius_rule_is_the_only_one_matching_counts():
    """particleRadius = kernelRadius/4 reproduces all three screenshot counts;
    neighbouring ratios do not (guards the inferred radius rule)."""
    from oracle import scenes
    g = json.load(open(os.path.join(HERE, "golden", "scene_counts.json")))["res24"]
    f32 = np.float32
    kr = f32(2.0) / f32(24)
    for div, ok in ((4.0, True), (3.0, False), (3.5, False), (5.0, False), (2.0, False)):
        r = kr / f32(div)
        counts = dict(CubeDrop=len(scenes.cube_drop(r)), SphereDrop=len(scenes.sphere_drop(r)),
                      DoubleDambreak=len(scenes.double_dambreak(r)))
        match = all(counts[k] == g[k] for k in counts)
        assert match == ok, (div, counts)


def test_cube_drop_res256_count():
    from oracle import scenes
    _, r = scenes.radii(256)
    g = int((np.float32(1.0)) / (np.float32(2.0) * r))
    assert g ** 3 == 16777216


def test_scene_positions_on_lattice():
    from oracle import scenes
    p = scenes.reference_scene("Dambreak", 24)
    _, r = scenes.radii(24)
    assert np.isclose(p[:, 0].min(), -1 + r)
    assert -0.5 - 4 * r < p[:, 2].max() <= -0.5   # last lattice point within 2 spacings of bMax
    d = np.diff(np.unique(p[:, 2]))
    assert np.allclose(d, 2 * r, rtol=1e-4)


# ---------------------------------------------------------- neighbours ----

@pytest.mark.parametrize("jitter", [0.0, 0.2 * 0.005])
def test_neighbor_sets_equal_bruteforce(oracle_lib, jitter):
    """Cell-list neighbour sets == all-pairs sets with the same predicate."""
    import oracle
    pos = lattice((10, 12, 9), jitter=jitter, seed=5)
    p = oracle.make_params()
    o = oracle.Oracle(p)
    o.record_neighbors(True)
    o.set_state(pos)
    o.step()
    off, lst = o.neighbors()
    dbg = o.step_debug()
    perm = dbg["perm"]
    sorted_pos = pos[perm].astype(np.float32)
    boff, blst = oracle.bruteforce_neighbors(sorted_pos, p.kernel_radius)
    n = len(pos)
    for k in range(n):
        a = np.sort(lst[off[k]:off[k + 1]])
        b = blst[boff[k]:boff[k + 1]]
        assert np.array_equal(a, b), k
    # symmetric and self-inclusive
    assert all(k in set(lst[off[k]:off[k + 1]]) for k in range(0, n, 97))


def test_nan_positions_have_no_neighbours(oracle_lib):
    """NaN coordinates of both signs inside a lattice (one axis or all three):
    the cell-list sets still equal the all-pairs sets, the NaN particles have
    no neighbours (not even themselves) and are nobody's neighbour."""
    import oracle
    pos = lattice((10, 12, 9), jitter=0.0005, seed=8).astype(np.float32)
    nan = {True: np.uint32(0xFFC00000).view(np.float32), False: np.uint32(0x7FC00000).view(np.float32)}
    bad = [17, 240, 333, 512, 777]
    for i, (axes, neg) in zip(bad, [((1,), True), ((1,), False), ((0,), True), ((0, 1, 2), True), ((2,), False)]):
        for ax in axes:
            pos[i, ax] = nan[neg]
    assert np.signbit(pos[17, 1]) and np.isnan(pos[17, 1])
    p = oracle.make_params(x_subdivision=4)
    o = oracle.Oracle(p)
    o.record_neighbors(True)
    o.set_state(pos.astype(np.float64))
    o.step()
    off, lst = o.neighbors()
    d = o.step_debug()
    sp = pos[d["perm"]]
    boff, blst = oracle.bruteforce_neighbors(sp, p.kernel_radius)
    assert np.array_equal(off, boff)
    for k in range(len(pos)):
        assert np.array_equal(np.sort(lst[off[k]:off[k + 1]]), blst[boff[k]:boff[k + 1]]), k
    nan_sorted = np.flatnonzero(np.isin(d["perm"], bad))
    assert (np.diff(off)[nan_sorted] == 0).all()
    assert not np.isin(lst, nan_sorted).any()
    ok = ~np.isin(d["sorted_id"], bad)
    assert np.isfinite(d["rho"]).all() and np.isfinite(d["acc"][ok]).all()


@pytest.mark.parametrize("xsub", [1, 4])
def test_clustered_cloud_neighbors_equal_bruteforce(oracle_lib, xsub):
    """Random clouds: uniform background, dense Gaussian blobs (cells with tens
    of particles, rows far longer than a neighbour mask) and particles at the
    box faces.  Cell-list sets == all-pairs sets."""
    import oracle
    rng = np.random.default_rng(17 + xsub)
    parts = [rng.uniform(0.0, 1.0, size=(1500, 3))]
    for c in rng.uniform(0.2, 0.8, size=(4, 3)):
        parts.append(c + rng.normal(0.0, 0.012, size=(500, 3)))
    face = rng.uniform(0.0, 1.0, size=(200, 3))
    face[:100, 0] = 0.0
    face[100:, 1] = 1.0
    parts.append(face)
    pos = np.clip(np.concatenate(parts), 0.0, 1.0)
    p = oracle.make_params(x_subdivision=xsub)
    o = oracle.Oracle(p)
    o.record_neighbors(True)
    o.set_state(pos)
    o.step()
    off, lst = o.neighbors()
    sp = pos[o.step_debug()["perm"]].astype(np.float32)
    boff, blst = oracle.bruteforce_neighbors(sp, p.kernel_radius)
    assert np.array_equal(off, boff)
    counts = np.diff(off)
    assert counts.max() > 100    # the blobs are dense
    for k in range(len(pos)):
        a = lst[off[k]:off[k + 1]]
        assert np.array_equal(np.sort(a), blst[boff[k]:boff[k + 1]]), k


@pytest.mark.parametrize("xsub", [2, 4])
def test_subdivided_grid_neighbors_and_density(oracle_lib, xsub):
    """x cells of cs/xsub: same neighbour sets as brute force, and density
    equal to the xsub = 1 grid up to summation order."""
    import oracle
    pos = lattice((9, 11, 8), jitter=0.001, seed=6, origin=(0.21, 0.05, 0.4))
    res = {}
    for xs in (1, xsub):
        p = oracle.make_params(x_subdivision=xs)
        o = oracle.Oracle(p)
        o.record_neighbors(True)
        o.set_state(pos)
        o.step()
        d = o.step_debug()
        off, lst = o.neighbors()
        sp = pos[d["perm"]].astype(np.float32)
        boff, blst = oracle.bruteforce_neighbors(sp, p.kernel_radius)
        assert np.array_equal(off, boff)
        for k in range(0, len(pos), 7):
            assert np.array_equal(np.sort(lst[off[k]:off[k + 1]]), blst[boff[k]:boff[k + 1]])
        assert o.constants()["G"][0] == 50 * xs
        rho = np.empty(len(pos)); rho[d["perm"]] = d["rho"]
        res[xs] = rho
    assert np.allclose(res[1], res[xsub], rtol=1e-6)


def test_lattice_interior_neighbour_count(oracle_lib):
    """Spacing h/2: 27 lattice points strictly inside radius h (33 with <=);
    float ties at exactly h land on either side, so interior counts lie in [27, 33]."""
    import oracle
    pos = lattice((12, 12, 12))
    o = oracle.Oracle(oracle.make_params())
    o.record_neighbors(True)
    o.set_state(pos)
    o.step()
    off, _ = o.neighbors()
    dbg = o.step_debug()
    sp = pos[dbg["perm"]]
    interior = np.all((sp > 0.005 + 2 * 0.01 - 1e-9) & (sp < 0.115 - 2 * 0.01 + 1e-9), axis=1)
    cnt = np.diff(off)[interior]
    assert cnt.min() >= 27 and cnt.max() <= 33
    exact = sum(1 for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3) if a * a + b * b + c * c < 4)
    assert exact == 27


def test_cell_keys_closed_form(oracle_lib):
    import oracle
    p = oracle.make_params()
    o = oracle.Oracle(p)
    c = o.constants()
    rng = np.random.default_rng(0)
    pos = rng.uniform(-0.1, 1.1, (2000, 3)).astype(np.float32).astype(np.float64)
    keys = o.cell_keys(pos)
    G = c["G"]
    f = ((pos.astype(np.float32) - np.float32(0)) * np.float32(c["inv_cs"])).astype(np.float32)
    cc = np.clip(np.floor(f), 0, G - 1).astype(np.int64)
    ref = (cc[:, 2] * G[1] + cc[:, 1]) * G[0] + cc[:, 0]
    assert np.array_equal(keys, ref)
    assert tuple(G) == (50, 50, 50)


def test_stable_sort_permutation(oracle_lib):
    import oracle
    pos = lattice((8, 8, 8), jitter=0.002, seed=3)
    o = oracle.Oracle(oracle.make_params())
    o.set_state(pos)
    o.step()
    d = o.step_debug()
    ref = np.argsort(d["key"], kind="stable")
    assert np.array_equal(d["perm"], ref)
    assert np.all(np.diff(d["sorted_key"].astype(np.int64)) >= 0)
    s, e = o.cells()
    occ = np.nonzero(e)[0]
    assert np.array_equal(occ, np.unique(d["sorted_key"]))
    assert np.all(e[occ] - s[occ] == np.bincount(d["sorted_key"])[occ])


# ------------------------------------------------------------- physics ----

def test_isolated_particle_density_and_free_fall(oracle_lib):
    """A lone particle: rho = m W(0) = cD h^6; it free-falls on the closed-form
    symplectic-Euler trajectory (no pressure or viscosity partner)."""
    import oracle
    p = oracle.make_params(pressure_stiffness=50.0)
    o = oracle.Oracle(p, fp64=True)
    o.set_state(np.array([[0.5, 0.8, 0.5]]), np.array([[0.1, 0.0, -0.2]]))
    c = o.constants()
    dt, g = float(np.float32(1e-4)), float(np.float32(-9.81))   # the params are float
    for s in range(1, 11):
        o.step()
        d = o.step_debug()
        assert d["rho"][0] == pytest.approx(c["cD"] * c["h2"] ** 3, rel=1e-12)
        x, v, _ = o.state()
        vy = s * dt * g
        y = 0.8 + dt * dt * g * s * (s + 1) / 2
        assert v[0, 1] == pytest.approx(vy, rel=1e-9)
        assert x[0, 1] == pytest.approx(y, rel=1e-9)
        assert x[0, 0] == pytest.approx(0.5 + 0.1 * dt * s, rel=1e-9)


def _half_space_sum(h, d, f):
    """sum f(dz, r2) s^3 over the midpoints of s-cubes of the half-space z < 0
    within h of (0, 0, d) (dz = d - z), at s = h/60 and h/120, combined by
    Richardson extrapolation (the midpoint sum's error is O(s^2))."""
    def grid_sum(n):
        s = h / n
        ax = (np.arange(-n, n) + 0.5) * s
        X, Y = np.meshgrid(ax, ax, indexing="ij")
        tot = 0.0
        for z in -(np.arange(0, n) + 0.5) * s:
            dz = d - z
            r2 = X ** 2 + Y ** 2 + dz * dz
            m = r2 < h * h
            tot += f(dz, r2[m]).sum()
        return tot * s ** 3
    return (4.0 * grid_sum(120) - grid_sum(60)) / 3.0


@pytest.mark.parametrize("u", [0.0, 0.25, 0.5, 0.8])
def test_wall_density_is_the_half_space_integral(oracle_lib, u):
    """A lone particle at u h above the floor: its density beyond the self term
    is rest fluid filling the half-space below the floor, summed with poly6
    over a fine grid of cubes (independent of the closed form)."""
    import oracle
    h, rho0 = 0.02, 1000.0
    o = oracle.Oracle(oracle.make_params(kernel_radius=h, particle_radius=h / 4, rest_density=rho0), fp64=True)
    o.set_state(np.array([[0.5, 0.5, u * h]]))
    o.step()
    c = o.constants()
    wall = o.step_debug()["rho"][0] - c["cD"] * c["h2"] ** 3
    brute = rho0 * _half_space_sum(h, u * h, lambda dz, r2: 315 / (64 * math.pi * h ** 9) * (h * h - r2) ** 3)
    assert wall == pytest.approx(brute, rel=1e-4)
    if u == 0.0:
        assert wall == pytest.approx(rho0 / 2, rel=1e-12)


@pytest.mark.parametrize("u", [0.25, 0.5, 0.8])
def test_wall_pressure_force_is_the_half_space_sum(oracle_lib, u):
    """A lone particle with negative pressure (attractive option at ratio 1,
    no gravity or viscosity) near the floor: its acceleration is the spiky
    pressure term of the rest fluid below the floor taken at its own p/rho^2,
    summed over a fine grid of cubes, and it points at the floor."""
    import oracle
    h, rho0, B = 0.02, 1000.0, 1000.0
    prm = oracle.make_params(kernel_radius=h, particle_radius=h / 4, rest_density=rho0, pressure_stiffness=B,
                             viscosity=0.0, gravity=(0.0, 0.0, 0.0), use_attractive_pressure=1,
                             attractive_pressure_ratio=1.0)
    o = oracle.Oracle(prm, fp64=True)
    o.set_state(np.array([[0.5, 0.5, u * h]]))
    o.step()
    d = o.step_debug()
    rho = d["rho"][0]
    q = B * ((rho / rho0) ** 7 - 1.0) / rho ** 2
    spiky = _half_space_sum(h, u * h, lambda dz, r2: (h - np.sqrt(r2)) ** 2 * dz / np.sqrt(r2))
    brute = 2 * q * rho0 * 45 / (math.pi * h ** 6) * spiky
    assert q < 0 and d["acc"][0, 2] < 0
    assert d["acc"][0, 2] == pytest.approx(brute, rel=1e-4)
    assert abs(d["acc"][0, 0]) < 1e-12 and abs(d["acc"][0, 1]) < 1e-12


def test_poly6_normalisation():
    """m sum W over a fine lattice -> rho0 * (integral of W = 1): checks 315/(64 pi h^9)."""
    h = 0.02
    s = h / 20
    ax = np.arange(-20, 21) * s
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    d2 = X ** 2 + Y ** 2 + Z ** 2
    W = np.where(d2 < h * h, (h * h - d2) ** 3, 0).sum() * 315 / (64 * math.pi * h ** 9)
    assert W * s ** 3 == pytest.approx(1.0, rel=2e-3)


def test_spiky_normalisation():
    """45/(pi h^6) is the normalisation of the spiky kernel's gradient magnitude
    (15/(pi h^6)(h-r)^3 integrates to 1)."""
    h = 0.02
    r = np.linspace(0, h, 200001)
    W = 15 / (math.pi * h ** 6) * (h - r) ** 3
    assert np.trapezoid(W * 4 * math.pi * r * r, r) == pytest.approx(1.0, rel=1e-6)
    dW = -45 / (math.pi * h ** 6) * (h - r) ** 2
    assert np.allclose(np.gradient(W, r)[1:-1], dW[1:-1], rtol=1e-3, atol=1e-3 * abs(dW).max())


# ------------------------------ pair terms against the published formulas ----
#
# These pins take nothing from the oracle but its outputs: the constants are
# written here from the papers (Mueller, Charypar and Gross 2003, "Particle-
# based fluid simulation for interactive applications": W_poly6 = 315/(64 pi
# h^9)(h^2 - r^2)^3, grad W_spiky = -45/(pi h^6)(h - r)^2 r_hat, lap W_visc =
# 45/(pi h^6)(h - r); Becker and Teschner 2007: p = B((rho/rho0)^7 - 1)) and
# from the lattice-cell mass m = rho0 (2r)^3.  A wrong normalisation (15 for
# 45, h^5 for h^6), a dropped 1/rho_i or 1/rho_j, or another mass rule moves
# the oracle's density or acceleration far outside the tolerances.

def _f32(v):
    return float(np.float32(v))   # the oracle's parameters are float


def _textbook_step(pos, vel, h, r, rho0, B, mu, g, dt, attractive=False, ratio=0.0):
    """One step of the formulation, FP64, all pairs, for particles farther
    than h from every box face (no wall terms): density with the self term,
    Tait EOS with the clamp (or the attractive ratio), symmetric pressure
    force, Laplacian viscosity, gravity, symplectic Euler."""
    m = rho0 * (2.0 * r) ** 3
    n = len(pos)
    rij = pos[:, None, :] - pos[None, :, :]
    d = np.linalg.norm(rij, axis=2)
    near = d < h
    W = np.where(near, 315.0 / (64.0 * math.pi * h ** 9) * (h * h - d * d) ** 3, 0.0)
    rho = m * W.sum(1)
    p_raw = B * ((rho / rho0) ** 7 - 1.0)
    p = np.where(p_raw < 0.0, p_raw * ratio if attractive else 0.0, p_raw)
    pair = near & ~np.eye(n, dtype=bool)
    ds = np.where(pair, d, 1.0)
    gradW = np.where(pair[..., None], (-45.0 / (math.pi * h ** 6) * (h - ds) ** 2 / ds)[..., None] * rij, 0.0)
    a_p = -(m * (p[:, None] / rho[:, None] ** 2 + p[None, :] / rho[None, :] ** 2))[..., None] * gradW
    lapW = np.where(pair, 45.0 / (math.pi * h ** 6) * (h - ds), 0.0)
    a_v = (mu / rho)[:, None, None] * (m * lapW / rho[None, :])[..., None] * (vel[None, :, :] - vel[:, None, :])
    a = a_p.sum(1) + a_v.sum(1) + np.asarray(g)
    v1 = vel + dt * a
    x1 = pos + dt * v1
    return dict(rho=rho, p=p, p_raw=p_raw, a=a, a_p=a_p.sum(1), x1=x1, v1=v1, m=m)


def _oracle_step(pos, vel, **kw):
    """One FP64 oracle step; density and acceleration back in input order."""
    import oracle
    o = oracle.Oracle(oracle.make_params(**kw), fp64=True)
    o.set_state(pos, vel)
    o.step()
    d = o.step_debug()
    rho = np.empty(len(pos)); acc = np.empty((len(pos), 3))
    rho[d["perm"]] = d["rho"]
    acc[d["perm"]] = d["acc"]
    x1, v1 = o.state_by_id()
    return rho, acc, x1, v1


def _pair_case(q, direction, h=0.02, r_over_h=0.4, seed=0):
    """Two particles q h apart along `direction`, around the box centre."""
    u = np.asarray(direction, float)
    u /= np.linalg.norm(u)
    h = _f32(h)
    c = np.array([0.5, 0.5, 0.5])
    pos = np.stack([c - 0.5 * q * h * u, c + 0.5 * q * h * u])
    vel = np.random.default_rng(seed).normal(0.0, 0.3, (2, 3))
    return pos, vel, h, _f32(r_over_h * h)


def _check_against_textbook(pos, vel, h, r, B, mu, rho0=1000.0, g=(0.0, -9.81, 0.0), dt=1e-4,
                            attractive=False, ratio=0.0):
    want = _textbook_step(pos, vel, h, r, _f32(rho0), _f32(B), _f32(mu), [_f32(x) for x in g], _f32(dt),
                          attractive, _f32(ratio))
    rho, acc, x1, v1 = _oracle_step(pos, vel, kernel_radius=h, particle_radius=r, rest_density=rho0,
                                    pressure_stiffness=B, viscosity=mu, gravity=g, dt=dt,
                                    use_attractive_pressure=int(attractive), attractive_pressure_ratio=ratio)
    np.testing.assert_allclose(rho, want["rho"], rtol=1e-12)
    scale = np.abs(want["a"]).max()
    np.testing.assert_allclose(acc, want["a"], rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(v1, want["v1"], rtol=0, atol=1e-11 * max(1.0, np.abs(want["v1"]).max()))
    np.testing.assert_allclose(x1, want["x1"], rtol=0, atol=1e-13)
    return want


def test_two_particle_step_compressed_branch(oracle_lib):
    """r = 0.4 h and 0.38 h apart: rho/rho0 ~ 1.3, p > 0; pressure and
    viscosity terms, gravity and the step all as published."""
    pos, vel, h, r = _pair_case(0.38, (1.0, 2.0, 2.0))
    w = _check_against_textbook(pos, vel, h, r, B=1000.0, mu=0.1)
    assert (w["rho"] > 1.2 * 1000.0).all() and (w["p"] > 0).all()
    # The pair repels: particle 0's pressure acceleration points away from 1.
    assert np.dot(w["a_p"][0], pos[0] - pos[1]) > 0


def test_two_particle_step_clamped_branch(oracle_lib):
    """0.9 h apart: rarefied, p < 0 is clamped to 0, so only viscosity and
    gravity act."""
    pos, vel, h, r = _pair_case(0.9, (0.0, 1.0, 0.0), seed=1)
    w = _check_against_textbook(pos, vel, h, r, B=1000.0, mu=0.2)
    assert (w["p_raw"] < 0).all() and (w["p"] == 0).all()
    assert np.abs(w["a_p"]).max() == 0.0


def test_two_particle_step_attractive_branch(oracle_lib):
    """Attractive pressure on: p < 0 is scaled by the ratio and the pair
    attracts (Source/Controller.cpp:61 exposes the option)."""
    pos, vel, h, r = _pair_case(0.7, (3.0, -1.0, 0.5), seed=2)
    w = _check_against_textbook(pos, vel, h, r, B=2000.0, mu=0.05, attractive=True, ratio=0.3)
    assert (w["p"] < 0).all()
    assert np.dot(w["a_p"][0], pos[0] - pos[1]) < 0


def test_three_particle_step_unequal_densities(oracle_lib):
    """Three particles at unequal spacings, so rho_i, p_i differ per particle:
    the p_i/rho_i^2 + p_j/rho_j^2 pairing and the 1/rho_i, 1/rho_j factors of
    the viscosity are told apart."""
    h = _f32(0.02)
    c = np.array([0.5, 0.5, 0.5])
    pos = c + h * np.array([[0.0, 0.0, 0.0], [0.33, 0.1, 0.0], [-0.1, 0.55, 0.3]])
    vel = np.random.default_rng(3).normal(0.0, 0.5, (3, 3))
    w = _check_against_textbook(pos, vel, h, _f32(0.42 * h), B=500.0, mu=0.3)
    rho = w["rho"]
    assert len(np.unique(np.round(rho, 6))) == 3 and (w["p"] > 0).any()


def test_eos_closed_form(oracle_lib):
    """The oracle's pressure, read back from its acceleration (two particles,
    no viscosity or gravity: |a| = 2 m (p/rho^2) 45/(pi h^6)(h - d)^2), is
    B((rho/rho0)^7 - 1) over compressed, near-rest and rarefied densities,
    clamped to 0 or scaled by the attractive ratio below rest."""
    h = _f32(0.02)
    r = _f32(0.4 * h)
    rho0, B = 1000.0, 3000.0
    m = rho0 * (2 * r) ** 3
    seen = set()
    for attractive in (False, True):
        for q in (0.2, 0.3, 0.38, 0.45, 0.5, 0.6, 0.8, 0.95):
            pos, _, _, _ = _pair_case(q, (2.0, 1.0, -2.0), h=0.02)
            vel = np.zeros((2, 3))
            rho, acc, _, _ = _oracle_step(pos, vel, kernel_radius=h, particle_radius=r, rest_density=rho0,
                                          pressure_stiffness=B, viscosity=0.0, gravity=(0, 0, 0),
                                          use_attractive_pressure=int(attractive),
                                          attractive_pressure_ratio=0.25)
            d = np.linalg.norm(pos[0] - pos[1])
            u = (pos[0] - pos[1]) / d
            a_along = float(np.dot(acc[0], u))
            p_oracle = a_along * rho[0] ** 2 / (2.0 * m * 45.0 / (math.pi * h ** 6) * (h - d) ** 2)
            x = rho[0] / rho0
            p = B * (x ** 7 - 1.0)
            if p < 0:
                p = 0.25 * p if attractive else 0.0
            assert p_oracle == pytest.approx(p, rel=1e-9, abs=1e-9 * B), (q, attractive)
            seen.add((x > 1.0, attractive))
    assert seen == {(True, False), (False, False), (True, True), (False, True)}


def test_interior_lattice_density_is_the_closed_lattice_sum(oracle_lib):
    """A particle inside a perfect lattice of spacing 2r = h/2: its density is
    m 315/(64 pi h^9) sum over lattice offsets o with |o| < h of
    (h^2 - |o|^2)^3 = rho0 (315 / 512 pi) (1 + 6 (3/4)^3 + 12 (1/2)^3 + 8 (1/4)^3)
    = 1.00978 rho0 (the 6 offsets at exactly h add 0).  Checked for two
    rest densities, so the mass rule m = rho0 (2r)^3 is pinned too."""
    h = _f32(0.02)
    r = _f32(h / 4)
    s = 2.0 * r
    ijk = np.stack(np.meshgrid(*[np.arange(-3, 4)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = 0.5 + s * ijk.astype(np.float64)
    centre = int(np.flatnonzero((ijk == 0).all(1))[0])
    closed = 315.0 / (512.0 * math.pi) * (1 + 6 * 0.75 ** 3 + 12 * 0.5 ** 3 + 8 * 0.25 ** 3)
    assert closed == pytest.approx(1.00978, abs=1e-5)
    for rho0 in (1000.0, 750.0):
        rho, acc, _, _ = _oracle_step(pos, np.zeros_like(pos), kernel_radius=h, particle_radius=r,
                                      rest_density=rho0, gravity=(0, 0, 0), pressure_stiffness=100.0)
        assert rho[centre] == pytest.approx(_f32(rho0) * closed, rel=1e-9)
        # The lattice is symmetric about the centre: no net force there.
        assert np.abs(acc[centre]).max() < 1e-9 * np.abs(acc).max()


def test_momentum_conservation_without_gravity(oracle_lib):
    """Pairwise antisymmetric pressure and viscosity forces: sum_i m a_i = 0."""
    import oracle
    pos = lattice((8, 8, 8), jitter=0.001, seed=7, origin=(0.3, 0.3, 0.3))
    rng = np.random.default_rng(1)
    vel = rng.normal(0, 0.5, pos.shape)
    o = oracle.Oracle(oracle.make_params(gravity=(0, 0, 0), pressure_stiffness=500.0, viscosity=0.5), fp64=True)
    o.set_state(pos, vel)
    o.step()
    a = o.step_debug()["acc"]
    assert np.abs(a.sum(0)).max() < 1e-9 * np.abs(a).sum(0).max()
    assert np.abs(a).max() > 1.0


def test_uniform_velocity_has_no_viscous_force(oracle_lib):
    """Galilean invariance: adding a uniform velocity changes no acceleration."""
    import oracle
    pos = lattice((6, 6, 6), jitter=0.001, seed=2, origin=(0.3, 0.3, 0.3))
    p = oracle.make_params(gravity=(0, 0, 0), viscosity=1.0)
    rng = np.random.default_rng(3)
    vel = rng.normal(0, 0.2, pos.shape)
    a = []
    for shift in (0.0, 0.7):
        o = oracle.Oracle(p, fp64=True)
        o.set_state(pos, vel + shift)
        o.step()
        a.append(o.step_debug()["acc"])
    assert np.allclose(a[0], a[1], rtol=1e-9, atol=1e-9 * np.abs(a[0]).max())


def test_wall_clamp_and_restitution(oracle_lib):
    """A particle driven through the floor is clamped to box_min + r and its
    normal velocity is reflected by -restitution."""
    import oracle
    p = oracle.make_params(boundary_restitution=0.25)
    o = oracle.Oracle(p, fp64=True)
    o.set_state(np.array([[0.5, 0.0051, 0.5]]), np.array([[0.0, -2.0, 0.0]]))
    o.step()
    x, v, _ = o.state()
    f = lambda v: float(np.float32(v))     # the params are float
    vy = -2.0 + f(1e-4) * f(-9.81)
    assert x[0, 1] == f(0.005)             # box_min + r
    assert v[0, 1] == pytest.approx(-f(0.25) * vy, rel=1e-12)


def test_cfl_step_closed_form(oracle_lib):
    """CFL step = min(dt, cfl h / (c0 + max|v|)), c0 = sqrt(7 B / rho0), from
    the velocities a step starts from; float32 arithmetic in that order.  A
    planted error (c0 without the factor 7) would change the step."""
    import oracle
    f32 = np.float32
    rng = np.random.default_rng(5)
    pos = lattice((6, 6, 6), jitter=1e-4, seed=5)
    vel = rng.normal(0.0, 0.7, size=pos.shape)
    for cfl, dt_max, B in ((0.3, 5e-3, 5000.0), (0.3, 1e-4, 5000.0), (0.0, 1e-4, 50.0)):
        p = oracle.make_params(cfl_number=cfl, dt=dt_max, pressure_stiffness=B)
        o = oracle.Oracle(p, fp64=False)
        o.set_state(pos, vel)
        o.step()
        v32 = vel.astype(np.float32)
        v2 = (v32[:, 2] * v32[:, 2] + (v32[:, 1] * v32[:, 1] + v32[:, 0] * v32[:, 0])).max()
        # fma chain vs mul+add differ by one rounding at most: compare closed
        # form in float64 to float32 resolution.
        c0 = f32(np.sqrt(7.0 * f32(B) / 1000.0))
        want = min(f32(dt_max), (f32(cfl) * f32(0.02)) / (c0 + f32(np.sqrt(v2)))) if cfl > 0 else f32(dt_max)
        assert o.last_dt() == pytest.approx(float(want), rel=3e-7), (cfl, dt_max, B)
    # Zero velocity: exactly cfl h / c0 when below the cap.
    p = oracle.make_params(cfl_number=0.25, dt=1.0, pressure_stiffness=7000.0)
    o = oracle.Oracle(p, fp64=True)
    o.set_state(pos)
    o.step()
    assert o.last_dt() == pytest.approx(0.25 * float(f32(0.02)) / np.sqrt(7.0 * 7000.0 / 1000.0), rel=1e-7)


def test_float_vs_double_step_agree(oracle_lib):
    import oracle
    pos = lattice((10, 20, 10), jitter=1e-4, seed=9)
    res = []
    for fp64 in (False, True):
        o = oracle.Oracle(oracle.make_params(), fp64=fp64)
        o.set_state(pos)
        o.step(3)
        res.append(o.state_by_id()[0])
    assert np.abs(res[0] - res[1]).max() < 1e-6


def test_golden_fixture_tiny(oracle_lib):
    """Committed FP64 vectors (tests/golden/tiny_step.npz, written by
    tests/golden/make_golden.py) still match the oracle."""
    import oracle
    g = np.load(os.path.join(HERE, "golden", "tiny_step.npz"))
    o = oracle.Oracle(oracle.make_params(pressure_stiffness=float(g["B"]), viscosity=float(g["mu"])), fp64=True)
    o.record_neighbors(True)
    o.set_state(g["pos"], g["vel"])
    o.step()
    d = o.step_debug()
    off, lst = o.neighbors()
    assert np.array_equal(d["perm"], g["perm"])
    assert np.array_equal(off, g["nbr_off"]) and np.array_equal(lst, g["nbr_list"])
    assert np.allclose(d["rho"], g["rho"], rtol=1e-12)
    assert np.allclose(d["acc"], g["acc"], rtol=1e-10, atol=1e-12)


def test_golden_fixture_wall(oracle_lib):
    """Committed FP64 vectors of a step in the box corner
    (tests/golden/wall_step.npz, written by tests/golden/make_golden.py once
    the wall terms matched their half-space sums) still match the oracle."""
    import oracle
    g = np.load(os.path.join(HERE, "golden", "wall_step.npz"))
    o = oracle.Oracle(oracle.make_params(pressure_stiffness=float(g["B"]), viscosity=float(g["mu"])), fp64=True)
    o.set_state(g["pos"], g["vel"])
    o.step()
    d = o.step_debug()
    assert np.array_equal(d["perm"], g["perm"])
    assert np.allclose(d["rho"], g["rho"], rtol=1e-12)
    assert np.allclose(d["acc"], g["acc"], rtol=1e-10, atol=1e-12)
    # The walls act: the floor's half-space makes up the floor layer's missing
    # neighbours (its mean density is the rest density to 1 %), and the layer
    # is held up against gravity.
    floor = g["pos"][d["perm"], 1] < 0.0075
    assert abs(g["rho"][floor].mean() - 1000.0) < 10.0
    assert g["acc"][floor, 1].mean() > -9.81


# ------------------------------------------------------------- adjoint ----

def _rollout_case(n3=(6, 6, 6), steps=8, seed=0):
    rng = np.random.default_rng(seed)
    pos = lattice(n3, jitter=1e-6, seed=seed)
    vel = rng.normal(0, 0.05, pos.shape)
    target = pos + rng.normal(0, 0.002, pos.shape)
    return pos, vel, target


@pytest.mark.parametrize("B,mu", [(50.0, 0.01), (500.0, 0.2)])
def test_adjoint_oracle_matches_fp64_central_differences(oracle_lib, B, mu):
    import oracle
    pos, vel, target = _rollout_case()
    p = oracle.make_params(pressure_stiffness=B, viscosity=mu)
    steps = 10
    L, dB, dmu, dv0 = oracle.rollout_loss_grad(p, pos, vel, target, steps)

    def loss(**kw):
        q = oracle.make_params(pressure_stiffness=kw.get("B", B), viscosity=kw.get("mu", mu))
        return oracle.rollout_loss_grad(q, pos, kw.get("vel", vel), target, steps, grad=False)

    # The float parameters are perturbed in float32, so use the exact float deltas.
    def fd_param(name, val):
        e = 1e-6 * max(1.0, abs(val))
        hi_v, lo_v = float(np.float32(val + e)), float(np.float32(val - e))
        return (loss(**{name: hi_v}) - loss(**{name: lo_v})) / (hi_v - lo_v)

    assert dB == pytest.approx(fd_param("B", B), rel=2e-4)
    assert dmu == pytest.approx(fd_param("mu", mu), rel=2e-4)
    rng = np.random.default_rng(1)
    u = rng.normal(size=vel.shape)
    e = 1e-6
    fd = (loss(vel=vel + e * u) - loss(vel=vel - e * u)) / (2 * e)
    assert float((dv0 * u).sum()) == pytest.approx(fd, rel=2e-5)


def test_adjoint_oracle_matches_torch_autograd(oracle_lib):
    import oracle
    from oracle import torch_dense
    pos, vel, target = _rollout_case(n3=(5, 6, 5), seed=4)
    p = oracle.make_params(pressure_stiffness=300.0, viscosity=0.1)
    steps = 12
    L, dB, dmu, dv0 = oracle.rollout_loss_grad(p, pos, vel, target, steps)
    c = oracle.Oracle(p, fp64=True).constants()
    Lt, dBt, dmut, dv0t = torch_dense.loss_and_grads(pos, vel, target, p, c, steps)
    assert L == pytest.approx(Lt, rel=1e-10)
    assert dB == pytest.approx(dBt, rel=1e-7)
    assert dmu == pytest.approx(dmut, rel=1e-7)
    assert np.abs(dv0 - dv0t).max() <= 1e-7 * np.abs(dv0t).max()


def test_adjoint_oracle_detects_planted_error(oracle_lib):
    """The FD check above has teeth: dropping the viscosity term of the force
    makes the oracle's own gradient disagree with its own differences."""
    import oracle
    pos, vel, target = _rollout_case(seed=2)
    p = oracle.make_params(pressure_stiffness=50.0, viscosity=0.3)
    L, dB, dmu, dv0 = oracle.rollout_loss_grad(p, pos, vel, target, 6)
    q = oracle.make_params(pressure_stiffness=50.0, viscosity=0.0)
    L0 = oracle.rollout_loss_grad(q, pos, vel, target, 6, grad=False)
    # a planted sign error on dmu would flip this finite-difference secant
    assert dmu * (L - L0) > 0
