"""Pins the oracle's invasion detection (check_invasions!, find_invasions!,
polygons_intersect, point_line_intersect; src/rings/integration.jl:379-509)
independently of its formulas: the parametric crossing of the +x ray with
the open segment, hand polygons with known insides, a winding-number
point-in-polygon over random overlapping rings, and the chunk restriction
(pairs of rings whose centres of mass sit in the same or adjacent r_chunks)."""
import math

import numpy as np
import pytest


def rings_sys(oracle_mod, nrings, nmax, L=24.0, periodic=True, np_type=None):
    np_type = [nmax] if np_type is None else np_type
    return oracle_mod.Sys(n=nrings * nmax, bl=(0.0, 0.0), size=(L, L), periodic=periodic,
                          kind=oracle_mod.KIND_RINGS, p=[0.0] * 4, nmax=nmax, nrings=nrings, np_type=np_type,
                          ring_par=np.zeros(8), inter=np.array([1.0, 1.0, 1.0, 1.2]))


def crosses(p, a, b):
    """Does the ray {p + t (1, 0), t > 0} meet the open segment (a, b)?"""
    (px, py), (ax, ay), (bx, by) = p, a, b
    if ay == by:
        return False  # parallel to the ray: the reference never counts it
    t = (py - ay) / (by - ay)
    if not (0.0 < t < 1.0):
        return False
    return ax + t * (bx - ax) > px


def winding(p, poly):
    tot = 0.0
    for k in range(len(poly)):
        a = poly[k] - p
        b = poly[(k + 1) % len(poly)] - p
        tot += math.atan2(a[0] * b[1] - a[1] * b[0], a @ b)
    return round(tot / (2 * math.pi))


def test_point_line_intersect_matches_parametric_crossing(oracle_mod):
    rng = np.random.default_rng(0)
    agree = 0
    for _ in range(20000):
        p, a, b = rng.uniform(-1, 1, size=(3, 2))
        if rng.random() < 0.1:
            b[0] = a[0]  # vertical
        assert oracle_mod.point_line_intersect(p, a, b) == crosses(p, a, b)
        agree += crosses(p, a, b)
    assert agree > 1000
    # horizontal segments and end heights never count
    assert not oracle_mod.point_line_intersect((0.0, 0.5), (1.0, 0.5), (2.0, 0.5))
    assert not oracle_mod.point_line_intersect((0.0, 0.0), (1.0, 0.0), (2.0, 1.0))
    assert not oracle_mod.point_line_intersect((0.0, 1.0), (1.0, 0.0), (2.0, 1.0))
    assert oracle_mod.point_line_intersect((0.0, 0.5), (1.0, 0.0), (1.0, 1.0))
    assert oracle_mod.point_line_intersect((0.0, 0.5), (1.0, 1.0), (1.0, 0.0))
    assert not oracle_mod.point_line_intersect((1.0, 0.5), (1.0, 0.0), (1.0, 1.0))  # ray starts on it


def place(nmax, polys):
    pts = np.zeros((len(polys) * nmax, 2))
    for r, poly in enumerate(polys):
        pts[r * nmax:r * nmax + len(poly)] = poly
    return pts


def test_hand_polygons(oracle_mod):
    nmax = 6
    square = np.array([[0.0, 0.0], [4.0, 0.0], [4.0, 4.0], [0.0, 4.0]])
    # an L-shaped (concave) ring and a ring whose points fall inside the
    # square, in the L's notch, and inside the L's arm
    ell = np.array([[5.0, 0.0], [9.0, 0.0], [9.0, 1.0], [6.0, 1.0], [6.0, 4.0], [5.0, 4.0]])
    probe = np.array([[1.0, 1.3], [3.0, 2.7], [7.0, 2.5], [8.0, 0.5], [5.5, 3.5], [10.0, 0.5]])
    # the square has 4 points: typed rings give each ring its own count
    s = oracle_mod.Sys(n=3 * nmax, bl=(0.0, 0.0), size=(12.0, 12.0), periodic=False, kind=oracle_mod.KIND_RINGS,
                       p=[0.0] * 4, nmax=nmax, nrings=3, np_type=[4, 6], types=[1, 2, 2],
                       ring_par=np.zeros(16), inter=np.tile([1.0, 1.0, 1.0, 1.2], 4))
    pts = place(nmax, [square, ell, probe])
    cms = np.array([square.mean(0), ell.mean(0), probe.mean(0)])
    inv = s.rings_invasions(pts, cms)
    # pair (1, 2): none (neither holds a vertex of the other); probe points
    # 1, 2 inside the square; 4 (arm) and 5 (vertical bar) inside the L,
    # 3 (notch) and 6 outside.  (The probe's own self-crossing polygon may
    # hold vertices of the others: those records have invasor 1 or 2.)
    assert not [v for v in inv if {v[0], v[1]} == {1, 2}]
    assert [v for v in inv if v[0] == 3] == [(3, 1, 2 * nmax + 1), (3, 1, 2 * nmax + 2), (3, 2, 2 * nmax + 4),
                                             (3, 2, 2 * nmax + 5)]


def test_point_level_with_a_vertex_is_missed(oracle_mod):
    """A point at the height of the diamond's left and right vertices: the
    strict height test drops all four edges there, so the reference counts
    no crossing (a property of its test, kept)."""
    nmax = 4
    diamond = np.array([[2.0, 0.0], [4.0, 2.0], [2.0, 4.0], [0.0, 2.0]])
    tri = np.array([[1.5, 2.0], [2.5, 1.0], [2.5, 3.0], [2.5, 3.0]])
    s = oracle_mod.Sys(n=2 * nmax, bl=(0.0, 0.0), size=(8.0, 8.0), periodic=False, kind=oracle_mod.KIND_RINGS,
                       p=[0.0] * 4, nmax=nmax, nrings=2, np_type=[4, 3], types=[1, 2],
                       ring_par=np.zeros(16), inter=np.tile([1.0, 1.0, 1.0, 1.2], 4))
    pts = place(nmax, [diamond, tri[:3]])
    inv = s.rings_invasions(pts, np.array([diamond.mean(0), tri[:3].mean(0)]))
    assert (2, 1, nmax + 1) not in inv  # (1.5, 2) is inside but level with two vertices
    assert (2, 1, nmax + 2) in inv and (2, 1, nmax + 3) in inv


def random_rings(seed, nrings, nmax, L, radius=1.0):
    rng = np.random.default_rng(seed)
    th = 2 * math.pi * np.arange(nmax) / nmax
    pts = np.zeros((nrings * nmax, 2))
    for r in range(nrings):
        c = rng.uniform(0, L, 2)
        rad = radius * (1 + 0.15 * rng.uniform(-1, 1, nmax))
        pts[r * nmax:(r + 1) * nmax] = c + np.stack([rad * np.cos(th), rad * np.sin(th)], 1)
    cms = pts.reshape(nrings, nmax, 2).mean(1)
    return pts, cms


def brute(pts, nrings, nmax, pair_ok=lambda a, b: True):
    out = set()
    polys = pts.reshape(nrings, nmax, 2)
    for a in range(nrings):
        for b in range(nrings):
            if a == b or not pair_ok(a, b):
                continue
            for i in range(nmax):
                if winding(polys[a, i], polys[b]) != 0:
                    out.add((a + 1, b + 1, a * nmax + i + 1))
    return out


def test_all_pairs_equal_winding_number(oracle_mod):
    nrings, nmax, L = 40, 12, 14.0
    pts, cms = random_rings(1, nrings, nmax, L)
    s = rings_sys(oracle_mod, nrings, nmax, L=L)
    inv = s.rings_invasions(pts, cms)
    assert len(inv) == len(set(inv))
    want = brute(pts, nrings, nmax)
    assert len(want) > 30
    assert set(inv) == want
    # push order: each pair's invasions come together, the first ring's points first
    for (a, b, p), (c, d, q) in zip(inv, inv[1:]):
        if (a, b) == (c, d):
            assert p < q


@pytest.mark.parametrize("periodic", [True, False])
def test_chunks_restrict_to_adjacent_centres(oracle_mod, periodic):
    nrings, nmax, L = 60, 10, 16.0
    pts, cms = random_rings(2, nrings, nmax, L - 2.0)
    pts += 1.0
    cms += 1.0
    s = rings_sys(oracle_mod, nrings, nmax, L=L, periodic=periodic)
    # coarse chunks (4 > ring diameter ~2.3): every overlapping pair is adjacent
    coarse = s.rings_invasions(pts, cms, 4, 4, 64)
    assert set(coarse) == set(s.rings_invasions(pts, cms))
    # fine chunks (0.8): only pairs whose centres sit in adjacent chunks
    n = 20
    cell = np.floor(cms / (L / n)).astype(int)
    row = (n - 1) - cell[:, 1]

    def adj(a, b):
        dr, dc = abs(row[a] - row[b]), abs(cell[a, 0] - cell[b, 0])
        if periodic:
            dr, dc = min(dr, n - dr), min(dc, n - dc)
        return dr <= 1 and dc <= 1

    fine = s.rings_invasions(pts, cms, n, n, 64)
    want = brute(pts, nrings, nmax, adj)
    assert set(fine) == want
    assert len(fine) < len(coarse)  # the restriction dropped pairs


def test_binning_errors(oracle_mod):
    nrings, nmax, L = 6, 8, 12.0
    pts, cms = random_rings(3, nrings, nmax, L)
    s = rings_sys(oracle_mod, nrings, nmax, L=L)
    far = cms.copy()
    far[2] = (L + 5.0, 3.0)  # beyond the one-chunk margin the row/col clamps allow
    with pytest.raises(RuntimeError, match="error 2"):
        s.rings_invasions(pts, far, 4, 4, 8)
    near = cms.copy()
    near[2] = (L + 1.0, 3.0)  # within one chunk of the edge: clamped into the last column
    s.rings_invasions(pts, near, 4, 4, 8)
    same = np.tile([[6.5, 6.5]], (nrings, 1))
    with pytest.raises(RuntimeError, match="error 1"):
        s.rings_invasions(pts, same, 4, 4, 5)
    s.rings_invasions(pts, same, 4, 4, 6)
