"""Pins the oracle's newton_step!/walls!/observables (src/integration.jl:271-515,
src/quantities.jl) with free flight, an analytic two-body oscillator, energy
conservation, time reversal and the wall closed forms."""
import math

import numpy as np
import pytest

from test_oracle_forces import ht_sys, lj_sys, random_dense


def test_free_flight_exact(oracle_mod):
    # particles far apart, HT range tiny: F = 0 and x_{n+1} = x_n + (v dt + 0)
    rng = np.random.default_rng(0)
    n = 50
    pos = np.stack([np.arange(n) * 3.0 + 5.0, np.full(n, 50.0)], axis=1)
    vel = rng.normal(size=(n, 2))
    s = oracle_mod.Sys(n=n, bl=(0, 0), size=(200.0, 100.0), periodic=False, kind=oracle_mod.KIND_HT,
                       p=[1.0, 1.0, 0.1, 0.1], dt=1e-3, cols=20, rows=10, nc=16, radius=0.05)
    p2, v2, f, _, _ = s.newton_steps(pos, vel, 1)
    assert np.array_equal(p2, pos + vel * 1e-3)
    assert np.array_equal(v2, vel) and np.all(f == 0)


def test_two_body_harmonic(oracle_mod):
    """HT with k_rep = k_atr = k is a linear spring of constant k/deq; the
    separation obeys mu d'' = -(k/deq)(d - deq), mu = 1/2."""
    k, deq, A = 4.0, 1.0, 0.05
    dt, nsteps = 1e-3, 2000
    s = oracle_mod.Sys(n=2, bl=(0, 0), size=(20.0, 20.0), periodic=False, kind=oracle_mod.KIND_HT,
                       p=[k, k, deq, 5.0], dt=dt, cols=0, rows=0, radius=0.5)
    pos = np.array([[10.0 - (deq + A) / 2, 10.0], [10.0 + (deq + A) / 2, 10.0]])
    vel = np.zeros((2, 2))
    p, v, _, _, _ = s.newton_steps(pos, vel, nsteps)
    omega = math.sqrt(2 * k / deq)
    d_exact = deq + A * math.cos(omega * nsteps * dt)
    d = p[1, 0] - p[0, 0]
    assert abs(d - d_exact) < 10 * A * (omega * dt) ** 2
    # centre of mass at rest, exactly symmetric motion
    assert p[0, 0] + p[1, 0] == pytest.approx(20.0, abs=1e-13)


def test_energy_conservation_lj(oracle_mod):
    rng = np.random.default_rng(1)
    n, L = 200, 20.0
    pos = random_dense(rng, n, L, dmin=1.05)
    vel = rng.normal(size=(n, 2)) * 0.5
    vel -= vel.mean(axis=0)
    s = lj_sys(oracle_mod, n=n, L=L, periodic=True, cols=0)
    s.s.dt = 1e-3
    _, _, _, ke, pe = s.newton_steps(pos, vel, 400, record=True)
    e = ke + pe
    assert np.abs(e - e[0]).max() < 2e-3 * abs(ke[0])
    # halving dt reduces the fluctuation ~4x (second order integrator)
    s.s.dt = 5e-4
    _, _, _, ke2, pe2 = s.newton_steps(pos, vel, 800, record=True)
    e2 = ke2 + pe2
    r = np.abs(e - e[0]).max() / np.abs(e2 - e2[0]).max()
    assert 2.5 < r < 6.0


def test_time_reversal(oracle_mod):
    rng = np.random.default_rng(2)
    n, L = 60, 15.0
    pos = random_dense(rng, n, L, dmin=1.1)
    vel = rng.normal(size=(n, 2)) * 0.3
    s = lj_sys(oracle_mod, n=n, L=L, periodic=False, cols=0)
    s.s.dt = 1e-3
    s.s.radius = -1.0  # keep walls out of the way (no reflection)
    p1, v1, _, _, _ = s.newton_steps(pos, vel, 300)
    p2, v2, _, _, _ = s.newton_steps(p1, -v1, 300)
    assert np.abs(p2 - pos).max() < 1e-9
    assert np.abs(v2 + vel).max() < 1e-9


def test_rigid_wall_reflection(oracle_mod):
    # one particle moving right, radius 0.5, wall at x = 10: velocity flips
    # the first step after x + r > 10 and keeps flipping while still out
    dt = 0.01
    s = oracle_mod.Sys(n=1, bl=(0, 0), size=(10.0, 10.0), periodic=False, kind=oracle_mod.KIND_HT,
                       p=[1, 1, 0.1, 0.1], dt=dt, cols=0, radius=0.5)
    pos = np.array([[9.45, 5.0]])
    vel = np.array([[1.0, 0.0]])
    p, v, _, _, _ = s.newton_steps(pos, vel, 1)  # x = 9.46, inside
    assert v[0, 0] == 1.0
    p, v, _, _, _ = s.newton_steps(pos, vel, 6)  # 9.51 > 9.5 at step 6 -> flipped
    assert v[0, 0] == -1.0 and v[0, 1] == 0.0


def test_periodic_wrap(oracle_mod):
    s = oracle_mod.Sys(n=1, bl=(-5.0, -5.0), size=(10.0, 10.0), periodic=True, kind=oracle_mod.KIND_HT,
                       p=[1, 1, 0.1, 0.1], dt=0.1, cols=0, radius=0.5)
    p, v, _, _, _ = s.newton_steps(np.array([[4.95, -4.99]]), np.array([[1.0, -0.5]]), 1)
    assert p[0, 0] == pytest.approx(5.05 - 10.0, abs=1e-14)
    assert p[0, 1] == pytest.approx(-5.04 + 10.0, abs=1e-14)


def test_observables_brute_force(oracle_mod):
    rng = np.random.default_rng(7)
    n, L = 40, 12.0
    pos = random_dense(rng, n, L, dmin=0.9)
    vel = rng.normal(size=(n, 2))
    s = lj_sys(oracle_mod, n=n, L=L, periodic=True)
    assert s.kinetic_energy(vel) == pytest.approx(0.5 * sum(vx * vx + vy * vy for vx, vy in vel), rel=1e-14)
    pe = 0.0
    for i in range(n):
        for j in range(i + 1, n):
            d = pos[i] - pos[j]
            d = d - L * np.round(d / L)
            r = math.hypot(*d)
            pe += 4 * ((1 / r) ** 12 - (1 / r) ** 6)
    assert s.potential_energy_lj(pos) == pytest.approx(pe, rel=1e-12)


def _lj_pair(ri, rj):
    """Force on i from j for LJ(sigma = 1, eps = 1) with no cutoff, written from
    U(d) = 4 (d^-12 - d^-6): F_i = -U'(d) (ri - rj) / d."""
    dr = np.asarray(ri, dtype=np.float64) - np.asarray(rj, dtype=np.float64)
    d = math.hypot(*dr)
    return (48.0 / d ** 13 - 24.0 / d ** 7) / d * dr


def _stale_setup(oracle_mod):
    """Rigid 8 x 6 box, 4 x 3 chunks of 2 x 2, LJ with no cutoff (the
    reference's LenJonesCfg): the pair set is "same or adjacent chunk".
    A (chunk column 1) and C (column 4) are at rest; B moves right from column
    2 (x = 3.95) into column 3 (x' = 4.05) during the drift."""
    s = oracle_mod.Sys(n=3, bl=(0.0, 0.0), size=(8.0, 6.0), periodic=False, kind=oracle_mod.KIND_LJ,
                       p=[1.0, 1.0, 0.0, 0.0], dt=0.01, cols=4, rows=3, nc=8, radius=0.2)
    pos = np.array([[1.9, 3.0], [3.95, 3.0], [6.1, 3.0]])
    vel = np.array([[0.0, 0.0], [10.0, 0.0], [0.0, 0.0]])
    return s, pos, vel


def test_second_evaluation_keeps_stale_chunks(oracle_mod):
    """update_verlet! (src/integration.jl:415-431) evaluates the second force
    on the chunks binned before the drift.  B leaves A's neighbourhood and
    enters C's during the drift, so the stale pair set {A-B} and the rebinned
    one {B-C} give different F' and different velocities."""
    s, pos, vel = _stale_setup(oracle_mod)
    dt = 0.01
    # the chunk columns before and after the drift (1-based, as the reference)
    assert [s.cell(*p)[2] for p in pos] == [1, 2, 4]
    f0 = np.zeros((3, 2))
    f0[0] = _lj_pair(pos[0], pos[1])  # A-B: chunks 1 and 2 are adjacent
    f0[1] = -f0[0]  # B-C: chunks 2 and 4 are not; C has no partner
    x1 = pos + vel * dt + f0 * (dt * dt / 2)
    assert s.cell(*x1[1])[2] == 3
    # stale chunks: the second evaluation keeps A-B at the new positions
    f1 = np.zeros((3, 2))
    f1[0] = _lj_pair(x1[0], x1[1])
    f1[1] = -f1[0]
    v_stale = vel + dt / 2 * (f0 + f1)
    # rebinned before the second evaluation: B-C instead of A-B
    g1 = np.zeros((3, 2))
    g1[1] = _lj_pair(x1[1], x1[2])
    g1[2] = -g1[1]
    v_rebin = vel + dt / 2 * (f0 + g1)
    p, v, f, _, _ = s.newton_steps(pos, vel, 1)
    assert np.allclose(p, x1, rtol=0, atol=1e-14)
    assert np.allclose(v, v_stale, rtol=1e-12, atol=1e-15)
    # the two semantics differ by far more than the tolerance, on all three
    # particles (the motion is along x)
    assert np.abs(v_stale - v_rebin)[:, 0].min() > 1e-4
    assert np.abs(v - v_rebin)[:, 0].min() > 1e-4
    # the forces left in the system are the second (stale) evaluation
    assert np.allclose(f, f1, rtol=1e-12, atol=1e-15)


def test_stale_pair_set_across_two_steps(oracle_mod):
    """The next step bins afresh: its first evaluation uses B-C (B now in
    chunk 3), not A-B."""
    s, pos, vel = _stale_setup(oracle_mod)
    p1, v1, _, _, _ = s.newton_steps(pos, vel, 1)
    p2, v2, f2, _, _ = s.newton_steps(pos, vel, 2)
    dt = 0.01
    g0 = np.zeros((3, 2))
    g0[1] = _lj_pair(p1[1], p1[2])
    g0[2] = -g0[1]
    x2 = p1 + v1 * dt + g0 * (dt * dt / 2)
    assert np.allclose(p2, x2, rtol=0, atol=1e-13)
    assert abs(g0[0]).max() == 0.0 and abs(g0[1]).max() > 1e-3
