"""Pins the oracle's self-propelled dynamics (Szabo and run-and-tumble,
src/integration.jl:68-109, 433-489, 517-535) with properties that do not
restate its formulas: the Szabo force is minus the gradient of a potential
written here independently, free particles move on straight lines, the
polarisation turns towards the force, the rotational noise has variance
2 Dr dt, and tumbles happen exactly when the uniform draw is below rate dt.
"""
import math

import numpy as np
import pytest

SZ = dict(k_rep=10.0, k_adh=0.75, r_eq=1.0, r_max=1.1)  # examples/szabo.jl


def szabo_sys(oracle_mod, n, L, *, vo=1.0, mu=1.0, relax=1.0, rot=0.0, dt=0.01, cols=0, nc=0, par=SZ):
    return oracle_mod.Sys(n=n, bl=(0.0, 0.0), size=(L, L), periodic=True, kind=oracle_mod.KIND_SZABO,
                          p=[par["k_rep"], par["k_adh"], par["r_eq"], par["r_max"]], dt=dt, cols=cols, rows=cols,
                          nc=nc, radius=par["r_eq"] / 2, ap=[vo, mu, relax, rot])


def rtp_sys(oracle_mod, n, L, *, vo=1.0, sigma=1.0, eps=1.0, rate=1.0, dt=1e-3, cols=0, nc=0):
    cut = 2 ** (1 / 6) * sigma
    return oracle_mod.Sys(n=n, bl=(0.0, 0.0), size=(L, L), periodic=True, kind=oracle_mod.KIND_RTP,
                          p=[sigma, eps, cut, 0.0], dt=dt, cols=cols, rows=cols, nc=nc, radius=cut / 2,
                          ap=[vo, rate, 0.0, 0.0])


def szabo_energy(d, par=SZ):
    """A potential whose gradient is the Szabo pair force on each branch:
    dU/dd = f_mod d (d - r_eq), so U = f_mod (d^3/3 - r_eq d^2/2)."""
    k_rep, k_adh, r_eq, r_max = par["k_rep"], par["k_adh"], par["r_eq"], par["r_max"]
    f_mod = k_adh / r_eq if d > r_eq else k_rep / (r_max - r_eq)
    return f_mod * (d ** 3 / 3.0 - r_eq * d * d / 2.0)


@pytest.mark.parametrize("d", [0.55, 0.8, 0.97, 1.03, 1.07, 1.099])
def test_szabo_force_is_minus_gradient(oracle_mod, d):
    s = szabo_sys(oracle_mod, 2, 50.0)
    ang = 0.37
    p1 = np.array([20.0, 20.0])
    p2 = p1 + d * np.array([math.cos(ang), math.sin(ang)])
    f = s.pair_force(np.stack([p1, p2]), 1, 2)  # force on particle 1
    h = 1e-6
    g = np.zeros(2)
    for k in range(2):
        e = np.zeros(2)
        e[k] = h
        up = szabo_energy(float(np.linalg.norm(p1 + e - p2)))
        dn = szabo_energy(float(np.linalg.norm(p1 - e - p2)))
        g[k] = (up - dn) / (2 * h)
    assert np.allclose(f, -g, rtol=1e-7, atol=1e-9)


def test_szabo_force_zero_at_rest_and_beyond_range(oracle_mod):
    s = szabo_sys(oracle_mod, 2, 50.0)
    for d in (SZ["r_eq"], SZ["r_max"] + 1e-9, 2.0):
        pos = np.array([[10.0, 10.0], [10.0 + d, 10.0]])
        assert np.all(s.pair_force(pos, 1, 2) == 0.0)
    # repulsive inside r_eq, attractive between r_eq and r_max
    pos = np.array([[10.0, 10.0], [10.9, 10.0]])
    assert s.pair_force(pos, 1, 2)[0] < 0.0
    pos = np.array([[10.0, 10.0], [11.05, 10.0]])
    assert s.pair_force(pos, 1, 2)[0] > 0.0


def test_szabo_cell_list_equals_all_pairs(oracle_mod):
    rng = np.random.default_rng(3)
    n, L, cols = 400, 24.0, 20  # chunk 1.2 >= r_max
    pos = rng.uniform(0, L, size=(n, 2))
    a = szabo_sys(oracle_mod, n, L)
    b = szabo_sys(oracle_mod, n, L, cols=cols, nc=oracle_mod.chunk_capacity(L / cols, L / cols, 0.5) * 4)
    fa = a.forces(pos)
    fb = b.forces(pos)
    assert np.allclose(fa, fb, rtol=1e-12, atol=1e-12)
    assert np.abs(fb.sum(axis=0)).max() < 1e-10  # Newton's third law


def test_free_self_propulsion_is_a_straight_line(oracle_mod):
    n, L, k, dt = 3, 1000.0, 50, 0.01
    s = szabo_sys(oracle_mod, n, L, vo=1.0, rot=0.0, dt=dt)
    pos0 = np.array([[100.0, 100.0], [300.0, 500.0], [700.0, 200.0]])
    ang0 = np.array([0.3, 2.0, -1.2])
    pos, ang, _ = s.sp_steps(pos0, ang0, k)
    exp = pos0 + k * dt * np.stack([np.cos(ang0), np.sin(ang0)], axis=1)
    assert np.allclose(pos, exp, rtol=0, atol=1e-11)
    assert np.allclose(ang, ang0, rtol=0, atol=1e-14)  # vel parallel to pol: no torque


def test_polarisation_turns_towards_the_force(oracle_mod):
    # particle 1 heads along +x; particle 2 sits just below it inside r_eq, so
    # the repulsion pushes particle 1 along +y: its polarisation must turn
    # counter-clockwise (theta grows), particle 2's clockwise
    s = szabo_sys(oracle_mod, 2, 100.0, vo=1.0, mu=1.0, relax=1.0, rot=0.0, dt=0.01)
    pos = np.array([[50.0, 50.0], [50.0, 49.2]])
    ang = np.array([0.0, math.pi])
    _, a, _ = s.sp_steps(pos, ang, 1)
    assert a[0] > 0.0
    assert a[1] > math.pi


def test_rotational_noise_variance(oracle_mod):
    n, L, dt, rot = 4000, 4000.0, 0.01, 0.3
    rng = np.random.default_rng(5)
    s = szabo_sys(oracle_mod, n, L, vo=1.0, rot=rot, dt=dt)
    g = np.arange(n)
    pos = np.stack([(g % 60) * 60.0 + 5.0, (g // 60) * 50.0 + 5.0], axis=1)  # far apart: no forces
    ang = rng.uniform(0, 2 * math.pi, n)
    noise = rng.standard_normal(n)
    _, a, _ = s.sp_steps(pos, ang, 1, noise=noise)
    d = a - ang
    assert np.allclose(d, math.sqrt(2 * rot * dt) * noise, rtol=1e-12, atol=1e-15)
    assert np.var(d) == pytest.approx(2 * rot * dt, rel=0.08)


def test_rtp_force_is_truncated_lj(oracle_mod):
    s = rtp_sys(oracle_mod, 2, 50.0)
    cut = 2 ** (1 / 6)
    for d, sign in ((0.95, -1.0), (1.05, -1.0)):
        pos = np.array([[10.0, 10.0], [10.0 + d, 10.0]])
        f = s.pair_force(pos, 1, 2)
        # minus the gradient of 4 (d^-12 - d^-6): repulsive below the minimum
        fe = -(4 * (-12 * d ** -13 + 6 * d ** -7))
        assert f[0] == pytest.approx(-fe, rel=1e-13)
        assert sign * f[0] > 0.0
    pos = np.array([[10.0, 10.0], [10.0 + cut * (1 + 1e-12), 10.0]])
    assert np.all(s.pair_force(pos, 1, 2) == 0.0)


def test_rtp_tumbles_exactly_below_rate_dt(oracle_mod):
    n, L, dt, rate = 6, 1000.0, 1e-2, 2.0
    s = rtp_sys(oracle_mod, n, L, rate=rate, dt=dt)
    pos = np.stack([np.arange(n) * 100.0 + 10.0, np.full(n, 500.0)], axis=1)
    ang0 = np.linspace(0.1, 0.6, n)
    thr = rate * dt
    u = np.array([0.0, thr * 0.999, thr, thr * 1.001, 0.5, 0.99])
    v = np.array([0.25, 0.5, 0.75, 0.1, 0.2, 0.3])
    noise = np.stack([u, v], axis=1).reshape(-1)
    p, a, _ = s.sp_steps(pos, ang0, 1, noise=noise)
    tumbled = u < thr
    assert np.array_equal(a[tumbled], 2 * math.pi * v[tumbled])
    assert np.array_equal(a[~tumbled], ang0[~tumbled])
    # the move of this step used the angle before the tumble
    exp = pos + dt * np.stack([np.cos(ang0), np.sin(ang0)], axis=1)
    assert np.allclose(p, exp, rtol=0, atol=1e-12)


def test_sp_periodic_wrap(oracle_mod):
    s = szabo_sys(oracle_mod, 1, 10.0, vo=1.0, dt=0.5)
    p, _, _ = s.sp_steps(np.array([[9.9, 5.0]]), np.array([0.0]), 1)
    assert p[0, 0] == pytest.approx(0.4, abs=1e-12)
