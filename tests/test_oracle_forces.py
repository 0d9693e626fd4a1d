"""Pins the oracle's pair forces and force loop (src/configs.jl:354-397,
src/integration.jl:37-226) with closed forms, finite differences of the
potential energy, brute force and conservation laws."""
import math

import numpy as np
import pytest


def lj_sys(oracle_mod, n=2, L=100.0, periodic=False, rcut=0.0, cols=0, nc=0, sigma=1.0, eps=1.0):
    return oracle_mod.Sys(n=n, bl=(0.0, 0.0), size=(L, L), periodic=periodic, kind=oracle_mod.KIND_LJ,
                          p=[sigma, eps, rcut, 0], cols=cols, rows=cols, nc=nc)


def ht_sys(oracle_mod, n=2, L=100.0, periodic=False, k_rep=10.0, k_atr=2.0, deq=1.0, dmax=1.2, cols=0, nc=0):
    return oracle_mod.Sys(n=n, bl=(0.0, 0.0), size=(L, L), periodic=periodic, kind=oracle_mod.KIND_HT,
                          p=[k_rep, k_atr, deq, dmax], cols=cols, rows=cols, nc=nc)


@pytest.mark.parametrize("sigma,eps", [(1.0, 1.0), (2.0, 4.0), (0.7, 0.3)])
def test_lj_closed_forms(oracle_mod, sigma, eps):
    s = lj_sys(oracle_mod, sigma=sigma, eps=eps)
    c = np.array([50.0, 50.0])
    # |F| = 24 eps / sigma at d = sigma, repulsive (pushes particle 1 away from 2)
    pos = np.array([c + [sigma, 0.0], c])
    f = s.pair_force(pos, 1, 2)
    assert f[0] == pytest.approx(24 * eps / sigma, rel=1e-14) and f[1] == 0.0
    # zero at the minimum 2^(1/6) sigma
    rmin = 2 ** (1 / 6) * sigma
    pos = np.array([c + [0.0, rmin], c])
    assert abs(s.pair_force(pos, 1, 2)[1]) < 1e-13 * 24 * eps / sigma
    # minus the derivative of U = 4 eps ((s/r)^12 - (s/r)^6), central differences
    U = lambda r: 4 * eps * ((sigma / r) ** 12 - (sigma / r) ** 6)
    for r in (0.95 * sigma, 1.3 * sigma, 2.0 * sigma, 2.9 * sigma):
        h = 1e-6 * r
        dU = (U(r + h) - U(r - h)) / (2 * h)
        direction = np.array([0.6, 0.8])
        pos = np.array([c + r * direction, c])
        f = s.pair_force(pos, 1, 2)
        assert np.allclose(f, -dU * direction, rtol=1e-7, atol=1e-12)
    # antisymmetry f(j,i) = -f(i,j), bitwise
    pos = np.array([c + [1.1, 0.3], c])
    assert np.array_equal(s.pair_force(pos, 2, 1), -s.pair_force(pos, 1, 2))


def test_lj_cutoff(oracle_mod):
    s = lj_sys(oracle_mod, rcut=2.5)
    c = np.array([50.0, 50.0])
    assert np.all(s.pair_force(np.array([c + [2.5000001, 0], c]), 1, 2) == 0)
    assert s.pair_force(np.array([c + [2.4999999, 0], c]), 1, 2)[0] != 0


def test_harmtrunc_piecewise(oracle_mod):
    k_rep, k_atr, deq, dmax = 10.0, 2.0, 1.0, 1.2
    s = ht_sys(oracle_mod, k_rep=k_rep, k_atr=k_atr, deq=deq, dmax=dmax)
    c = np.array([50.0, 50.0])
    for d, k in ((0.75, k_rep), (0.998046875, k_rep), (1.125, k_atr), (1.1875, k_atr)):
        f = s.pair_force(np.array([c + [d, 0], c]), 1, 2)
        # force on 1: -k (d/deq - 1) along +x (positive = repulsive)
        assert f[0] == pytest.approx(-k * (d / deq - 1), rel=1e-13)
    assert np.all(s.pair_force(np.array([c + [1.2000001, 0], c]), 1, 2) == 0)
    assert abs(s.pair_force(np.array([c + [1.0, 0], c]), 1, 2)[0]) == 0.0


def test_min_image(oracle_mod):
    L = 10.0
    s = lj_sys(oracle_mod, L=L, periodic=True)
    # strict >: exactly L/2 is not wrapped
    assert s.calc_diff([5.0, 0.0], [0.0, 0.0])[0] == 5.0
    assert s.calc_diff([5.0 + 1e-12, 0.0], [0.0, 0.0])[0] == pytest.approx(-5.0 + 1e-12, abs=1e-15)
    assert s.calc_diff([0.1, 9.9], [9.9, 0.1])[0] == pytest.approx(0.2, abs=1e-14)
    assert s.calc_diff([0.1, 9.9], [9.9, 0.1])[1] == pytest.approx(-0.2, abs=1e-14)
    s2 = lj_sys(oracle_mod, L=L, periodic=False)
    assert s2.calc_diff([0.1, 9.9], [9.9, 0.1])[0] == pytest.approx(-9.8)


def random_dense(rng, n, L, dmin=0.8):
    """Random positions in [0, L)^2 with a minimum separation, measured with
    the minimum image so the set is valid for periodic boxes too."""
    pts = np.empty((0, 2))
    while len(pts) < n:
        p = rng.uniform(0, L, 2)
        if len(pts):
            d = pts - p
            d -= L * np.round(d / L)
            if np.hypot(*d.T).min() < dmin:
                continue
        pts = np.vstack([pts, p])
    return pts


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("kind", ["ht", "lj_cut"])
def test_cells_equal_brute_force(oracle_mod, periodic, kind):
    """The reference's check_chunks technique: for finite-range potentials the
    chunk loop and the all-pairs loop see the same interactions."""
    rng = np.random.default_rng(3)
    L, cols, n = 30.0, 10, 500
    pos = random_dense(rng, n, L)
    if kind == "ht":
        mk = lambda c, nc: ht_sys(oracle_mod, n=n, L=L, periodic=periodic, dmax=2.9, cols=c, nc=nc)
    else:
        mk = lambda c, nc: lj_sys(oracle_mod, n=n, L=L, periodic=periodic, rcut=2.9, cols=c, nc=nc)
    f_cells = mk(cols, 64).forces(pos)
    f_all = mk(0, 0).forces(pos)
    assert np.allclose(f_cells, f_all, rtol=1e-12, atol=1e-12)
    if periodic:
        assert np.abs(f_cells.sum(axis=0)).max() < 1e-13 * np.abs(f_cells).sum()  # Newton's third law


def test_threaded_equals_sequencial(oracle_mod):
    rng = np.random.default_rng(4)
    L, cols, n = 40.0, 16, 900
    pos = random_dense(rng, n, L)
    s = lj_sys(oracle_mod, n=n, L=L, periodic=True, cols=cols, nc=64)
    f1 = s.forces(pos)
    for nt in (2, 3, 8):
        assert np.allclose(s.forces_threaded(pos, nt), f1, rtol=1e-13, atol=1e-13)


def test_absum_bounds_force(oracle_mod):
    rng = np.random.default_rng(6)
    pos = random_dense(rng, 300, 25.0)
    s = lj_sys(oracle_mod, n=300, L=25.0, periodic=True, cols=10, nc=64)
    f, ab = s.forces(pos, with_absum=True)
    assert np.all(np.abs(f).sum(axis=1) <= ab * (1 + 1e-12))
