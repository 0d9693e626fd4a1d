"""Pins the oracle's ring forces and update (src/rings/integration.jl:32-351)
with polygon geometry, finite differences of the area energy and the
overdamped closed form."""
import math

import numpy as np
import pytest


def ring_sys(oracle_mod, nmax, nrings, L, par, inter, types=None, np_type=None, cols=0, nc=0, dt=0.01):
    return oracle_mod.Sys(n=nmax * nrings, bl=(0.0, 0.0), size=(L, L), periodic=True, kind=oracle_mod.KIND_RINGS,
                          p=[0, 0, 0, 0], dt=dt, cols=cols, rows=cols, nc=nc, nmax=nmax, nrings=nrings,
                          types=types, np_type=np_type if np_type is not None else [nmax],
                          ring_par=np.asarray(par, dtype=float), inter=np.asarray(inter, dtype=float))


def ngon(center, n, side, phase=0.0):
    R = side / (2 * math.sin(math.pi / n))
    a = phase + np.arange(n) * 2 * math.pi / n
    return np.stack([center[0] + R * np.cos(a), center[1] + R * np.sin(a)], axis=1)


def shoelace(p):
    x, y = p[:, 0], p[:, 1]
    return 0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))


# p0, relax, vo, mu, rot, k_area, k_spring, l_spring
PAR = [3.5, 1.0, 1.0, 1.0, 0.0, 1.0, 20.0, 1.0]
NO_PAIR = [10.0, 0.0, 1.0, 1e-9]  # pair range ~0: only springs and area act


@pytest.mark.parametrize("n", [3, 5, 10, 20])
def test_ngon_area(oracle_mod, n):
    s = ring_sys(oracle_mod, n, 1, 100.0, PAR, NO_PAIR)
    pos = ngon((50.0, 50.0), n, 1.0, 0.3)
    cpos = s.rings_continuous_pos(pos)
    assert s.rings_area(cpos, 1) == pytest.approx(n / (4 * math.tan(math.pi / n)), rel=1e-13)


def test_relaxed_ring_has_zero_force(oracle_mod):
    n, l = 10, 1.0
    area = n * l * l / (4 * math.tan(math.pi / n))
    p0 = n * l / math.sqrt(area)  # A0 = (n l / p0)^2 = area
    par = list(PAR)
    par[0] = p0
    s = ring_sys(oracle_mod, n, 1, 100.0, par, NO_PAIR)
    f, areas = s.rings_forces(ngon((50.0, 50.0), n, l))
    assert areas[0] == pytest.approx(area, rel=1e-13)
    assert np.abs(f).max() < 1e-12


def test_area_force_is_minus_gradient(oracle_mod):
    n = 8
    rng = np.random.default_rng(0)
    pos = ngon((50.0, 50.0), n, 1.0) + rng.normal(scale=0.05, size=(n, 2))
    par = list(PAR)
    par[6] = 0.0  # no springs
    k_area, p0, l0 = par[5], par[0], par[7]
    A0 = (n * l0 / p0) ** 2
    s = ring_sys(oracle_mod, n, 1, 100.0, par, NO_PAIR)
    f, _ = s.rings_forces(pos)
    E = lambda p: 0.5 * k_area * (shoelace(p) - A0) ** 2
    h = 1e-3  # E is quadratic in each coordinate: central differences are exact
    for i in range(n):
        for d in range(2):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += h
            pm[i, d] -= h
            g = (E(pp) - E(pm)) / (2 * h)
            assert f[i, d] == pytest.approx(-g, rel=1e-6, abs=1e-9)


def test_spring_force_is_minus_gradient(oracle_mod):
    n = 6
    rng = np.random.default_rng(1)
    pos = ngon((50.0, 50.0), n, 1.0) + rng.normal(scale=0.1, size=(n, 2))
    par = list(PAR)
    par[5] = 0.0  # no area term
    k, l = par[6], par[7]
    s = ring_sys(oracle_mod, n, 1, 100.0, par, NO_PAIR)
    f, _ = s.rings_forces(pos)
    E = lambda p: sum(0.5 * k * (math.hypot(*(p[i] - p[(i + 1) % n])) - l) ** 2 for i in range(n))
    h = 1e-6
    for i in range(n):
        for d in range(2):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += h
            pm[i, d] -= h
            assert f[i, d] == pytest.approx(-(E(pp) - E(pm)) / (2 * h), rel=1e-6, abs=1e-9)


def test_unwrap_across_boundary(oracle_mod):
    n, L = 10, 20.0
    pos = ngon((0.2, 19.9), n, 1.0)  # straddles both periodic edges
    pos = np.mod(pos, L)
    s = ring_sys(oracle_mod, n, 1, L, PAR, NO_PAIR)
    c = s.rings_continuous_pos(pos)
    steps = np.hypot(*(np.diff(c, axis=0).T))
    assert np.all(steps < 1.5)  # continuous
    assert np.allclose(np.mod(c, L), pos, atol=1e-12)
    assert s.rings_area(c, 1) == pytest.approx(n / (4 * math.tan(math.pi / n)), rel=1e-12)


def test_same_ring_pair_rules(oracle_mod):
    """Neighbours along the ring never interact; same-ring pairs repel but do
    not attract; different rings attract with k_atr."""
    n = 6
    inter = [10.0, 3.0, 1.0, 1.5]
    s = ring_sys(oracle_mod, n, 2, 100.0, PAR, inter)
    pos = np.zeros((12, 2))
    pos[:6] = ngon((20.0, 20.0), 6, 1.0)
    pos[6:] = ngon((60.0, 60.0), 6, 1.0)
    pos[2] = pos[0] + [0.5, 0.0]  # ids 1 and 3 (non adjacent, same ring) at 0.5: repel
    f = s.pair_force(pos, 1, 3)
    assert f[0] == pytest.approx(-10.0 * (0.5 - 1.0) * -1.0, rel=1e-12)
    assert np.all(s.pair_force(pos, 1, 2) == 0)  # adjacent
    assert np.all(s.pair_force(pos, 1, 6) == 0)  # adjacent across the wrap (diff = n-1)
    pos[2] = pos[0] + [1.2, 0.0]  # same ring, attractive zone -> 0
    assert np.all(s.pair_force(pos, 1, 3) == 0)
    pos[7] = pos[0] + [1.2, 0.0]  # other ring, attractive zone -> k_atr
    f = s.pair_force(pos, 1, 8)
    assert f[0] == pytest.approx(-3.0 * (1.2 - 1.0) * -1.0, rel=1e-12)


def test_overdamped_drift_closed_form(oracle_mod):
    """Relaxed ring, no pair forces, rot_diff = 0: every particle moves by
    vo*(cos t, sin t)*dt and the polarisation stays put (vcm parallel to pol)."""
    n, l, dt = 10, 1.0, 0.01
    area = n * l * l / (4 * math.tan(math.pi / n))
    par = list(PAR)
    par[0] = n * l / math.sqrt(area)
    s = ring_sys(oracle_mod, n, 1, 100.0, par, NO_PAIR, dt=dt)
    pos = ngon((50.0, 50.0), n, l)
    theta = 0.7
    p, pol, f, _, _ = s.rings_steps(pos, [theta], 1)
    assert np.allclose(p - pos, dt * np.array([math.cos(theta), math.sin(theta)]), atol=1e-13)
    assert pol[0] == pytest.approx(theta, abs=1e-12)


def test_polarisation_relaxes_toward_velocity(oracle_mod):
    """With an external drift from the area force absent and vo = 0, a ring
    pushed sideways turns its polarisation by (1/tau) asin(cross) dt."""
    n = 10
    par = list(PAR)
    s = ring_sys(oracle_mod, n, 1, 100.0, par, NO_PAIR)
    pos = ngon((50.0, 50.0), n, 1.0)
    noise = np.array([0.5])
    par_rot = list(par)
    par_rot[4] = 0.2
    s2 = ring_sys(oracle_mod, n, 1, 100.0, par_rot, NO_PAIR)
    _, pol0, _, _, _ = s.rings_steps(pos, [0.0], 1)
    _, pol1, _, _, _ = s2.rings_steps(pos, [0.0], 1, noise=noise)
    assert pol1[0] - pol0[0] == pytest.approx(math.sqrt(2 * 0.2 * 0.01) * 0.5, rel=1e-12)


@pytest.mark.parametrize("only_rings", [False, True])
@pytest.mark.parametrize("chunks", [True, False])
def test_contact_neighbors_match_brute_force(oracle_mod, only_rings, chunks):
    """ParticleNeighbors (src/rings/neighbors.jl:117-137): the oracle's lists,
    filled along its chunk traversal, hold exactly the pairs closer than
    tol * dist_eq found by brute force, symmetrically."""
    import sys

    sys.path.insert(0, "mavi.jl_amd")
    from mavi_hip import workloads as wl

    st, sp, cfg, ic = wl.rings_mixture(nrings=36, seed=3)
    rng = np.random.default_rng(1)
    pos = np.mod(st.pos + rng.normal(scale=0.08, size=st.pos.shape), sp.geometry_cfg.length)
    if not chunks:
        ic.chunks_cfg = None
    o = oracle_mod.from_rings(st, sp, cfg, ic)
    tol = 1.1
    counts, lists = o.rings_neighbors(pos, tol=tol, only_rings=only_rings)
    L = sp.geometry_cfg.length
    n, nmax = len(pos), st.nmax
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    dist = np.sqrt((d ** 2).sum(-1))
    ring = np.arange(n) // nmax
    close = (dist < 1.0 * tol) & ~np.eye(n, dtype=bool)
    if only_rings:
        close &= ring[:, None] != ring[None, :]
    assert counts.sum() > (20 if only_rings else n)  # contacts exist
    for i in range(n):
        got = set(int(v) - 1 for v in lists[i, :counts[i]])
        assert got == set(np.nonzero(close[i])[0].tolist())
        assert np.all(lists[i, counts[i]:] == -1)
