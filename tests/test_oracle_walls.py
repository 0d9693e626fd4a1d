"""Pins the oracle's walls list (ManyWalls of rigid/periodic rectangles,
rigid and slippery circles, slippery lines, potential walls;
src/integration.jl:228-403, src/configs.jl:99-162, 255-261) with geometry
and energies written here independently of its formulas."""
import math

import numpy as np
import pytest

R_P = 0.5  # particle radius used by the slippery walls


def sys_with(oracle_mod, walls, *, periodic=False, L=40.0, bl=(-20.0, -20.0), kind=None, p=(1.0, 1.0, 0.0, 0.0)):
    kind = oracle_mod.KIND_LJ if kind is None else kind
    return oracle_mod.Sys(n=4, bl=bl, size=(L, L), periodic=periodic, kind=kind, p=list(p), dt=1e-3, radius=R_P,
                          walls=walls)


def circle(oracle_mod, kind, c, R, **kw):
    return dict(kind=kind, geom=(c[0], c[1], R, 0.0), **kw)


def test_rigid_circle_reflects(oracle_mod):
    s = sys_with(oracle_mod, [circle(oracle_mod, oracle_mod.WK_RIGID_CIRCLE, (0.0, 0.0), 10.0)])
    rng = np.random.default_rng(0)
    for _ in range(20):
        ang = rng.uniform(0, 2 * math.pi)
        r = rng.uniform(9.6, 10.2)  # beyond R - r_p = 9.5
        pos = np.zeros((4, 2))
        vel = np.zeros((4, 2))
        pos[1] = (r * math.cos(ang), r * math.sin(ang))
        vel[1] = rng.normal(size=2)
        _, v = s.walls_apply(pos, vel, 1, R_P)
        n = pos[1] / np.linalg.norm(pos[1])
        assert np.linalg.norm(v[1]) == pytest.approx(np.linalg.norm(vel[1]), rel=1e-13)
        assert v[1] @ n == pytest.approx(-(vel[1] @ n), abs=1e-13)
        t = np.array([-n[1], n[0]])
        assert v[1] @ t == pytest.approx(vel[1] @ t, abs=1e-13)
    pos = np.zeros((4, 2))
    pos[2] = (3.0, 4.0)  # inside: untouched
    vel = np.ones((4, 2))
    _, v = s.walls_apply(pos, vel, 2, R_P)
    assert np.array_equal(v, vel)


def test_slippery_circle_pushes_to_contact(oracle_mod):
    c = np.array([2.0, -1.0])
    s = sys_with(oracle_mod, [circle(oracle_mod, oracle_mod.WK_SLIPPERY_CIRCLE, c, 3.0)])
    for ang, d in ((0.3, 3.2), (2.0, 1.0), (-1.4, 3.49)):
        pos = np.zeros((4, 2))
        pos[0] = c + d * np.array([math.cos(ang), math.sin(ang)])
        p, _ = s.walls_apply(pos, None, 0, R_P)
        q = p[0] - c
        assert np.linalg.norm(q) == pytest.approx(3.0 + R_P, rel=1e-14)
        assert math.atan2(q[1], q[0]) == pytest.approx(ang, abs=1e-12)
    pos = np.zeros((4, 2))
    pos[0] = c + (4.0, 0.0)  # beyond contact: untouched
    p, _ = s.walls_apply(pos, None, 0, R_P)
    assert np.array_equal(p, pos)


def test_slippery_line_interior_and_corner(oracle_mod):
    seg = np.array([[-5.0, 1.0, 5.0, 1.0]])  # horizontal, normal (0, 1)
    s = sys_with(oracle_mod, [dict(kind=oracle_mod.WK_SLIPPERY_LINES, lines=seg)])
    for x, y, want in ((0.0, 1.3, 1.0 + R_P), (2.0, 0.8, 1.0 - R_P), (-4.9, 1.49, 1.0 + R_P)):
        pos = np.zeros((4, 2))
        pos[3] = (x, y)
        p, _ = s.walls_apply(pos, None, 3, R_P)
        assert p[3, 0] == x
        assert p[3, 1] == pytest.approx(want, abs=1e-14)
    # corner beyond the segment end: pushed to distance r from the endpoint
    pos = np.zeros((4, 2))
    pos[0] = (5.2, 1.1)
    p, _ = s.walls_apply(pos, None, 0, R_P)
    assert np.linalg.norm(p[0] - (5.0, 1.0)) == pytest.approx(R_P, rel=1e-14)
    # far from the segment: untouched
    pos[0] = (5.6, 1.0)
    p, _ = s.walls_apply(pos, None, 0, R_P)
    assert np.array_equal(p, pos)


def test_slippery_lines_add_corrections_from_the_start_position(oracle_mod):
    a = np.array([[0.0, 0.0, 10.0, 0.0]])
    b = np.array([[0.0, 0.0, 0.0, 10.0]])
    pos = np.zeros((4, 2))
    pos[0] = (0.3, 0.2)
    both = sys_with(oracle_mod, [dict(kind=oracle_mod.WK_SLIPPERY_LINES, lines=np.vstack([a, b]))])
    pa, _ = sys_with(oracle_mod, [dict(kind=oracle_mod.WK_SLIPPERY_LINES, lines=a)]).walls_apply(pos, None, 0, R_P)
    pb, _ = sys_with(oracle_mod, [dict(kind=oracle_mod.WK_SLIPPERY_LINES, lines=b)]).walls_apply(pos, None, 0, R_P)
    p, _ = both.walls_apply(pos, None, 0, R_P)
    assert np.allclose(p[0], pos[0] + (pa[0] - pos[0]) + (pb[0] - pos[0]), rtol=0, atol=1e-15)
    assert np.allclose(p[0], (R_P, R_P), atol=1e-15)


def ht_energy(d, k_rep, k_atr, deq, dmax):
    if d > dmax:
        return 0.0
    k = k_rep if d < deq else k_atr
    return 0.5 * k * deq * (d / deq - 1.0) ** 2


def lj_energy(d, sigma=1.0, eps=1.0):
    return 4 * eps * ((sigma / d) ** 12 - (sigma / d) ** 6)


def fd_grad(fun, p, h=1e-6):
    g = np.zeros(2)
    for k in range(2):
        e = np.zeros(2)
        e[k] = h
        g[k] = (fun(p + e) - fun(p - e)) / (2 * h)
    return g


@pytest.mark.parametrize("mode,sign", [(2, 1.0), (0, 1.0), (1, -1.0)])
def test_potential_circle_is_minus_gradient(oracle_mod, mode, sign):
    # HarmTrunc wall potential (k_rep 10, k_atr 2, dist_eq 1, dist_max 1.5);
    # particle outside the circle: inside flag +1, so Outside keeps the
    # distance and Inside negates it
    ht = (10.0, 2.0, 1.0, 1.5)
    c, R = np.array([1.0, 2.0]), 4.0
    w = circle(oracle_mod, oracle_mod.WK_POTENTIAL_CIRCLE, c, R, mode=mode, pot_kind=oracle_mod.KIND_HT, pot=ht)
    s = sys_with(oracle_mod, [w])
    for ang, d in ((0.4, 0.7), (2.5, 1.2)):
        p0 = c + (R + d) * np.array([math.cos(ang), math.sin(ang)])
        pos = np.zeros((4, 2))
        pos[1] = p0
        f = s.wall_forces(pos, 1)

        def energy(p):
            dist = sign * abs(np.linalg.norm(p - c) - R)
            return ht_energy(dist, *ht)

        if sign > 0:
            assert np.allclose(f, -fd_grad(energy, p0), rtol=1e-6, atol=1e-8)
        else:
            # a negative processed distance: the wall force of the formula,
            # fmod/dist * dr, with fmod from the distance's sign branch
            dist = -d
            fmod = -ht[0] * (dist / ht[2] - 1.0)
            dr = (p0 - c) - (p0 - c) / np.linalg.norm(p0 - c) * R
            assert np.allclose(f, fmod / dist * dr, rtol=1e-13)


def test_potential_lines_lj_is_minus_gradient(oracle_mod):
    seg = np.array([[0.0, 0.0, 6.0, 0.0], [6.0, 0.0, 6.0, 6.0]])
    w = dict(kind=oracle_mod.WK_POTENTIAL_LINES, mode=2, pot_kind=oracle_mod.KIND_LJ, pot=(1.0, 1.0, 0.0, 0.0),
             lines=seg)
    s = sys_with(oracle_mod, [w])

    def dist_seg(p, a, b):
        t = np.clip((p - a) @ (b - a) / ((b - a) @ (b - a)), 0, 1)
        return np.linalg.norm(p - (a + t * (b - a)))

    def energy(p):
        return sum(lj_energy(dist_seg(p, q[:2], q[2:])) for q in seg)

    for p0 in (np.array([3.0, 1.1]), np.array([5.0, 1.05]), np.array([-0.8, 0.5])):
        pos = np.zeros((4, 2))
        pos[0] = p0
        assert np.allclose(s.wall_forces(pos, 0), -fd_grad(energy, p0), rtol=1e-6, atol=1e-6)


def test_walls_apply_in_list_order(oracle_mod):
    # a periodic box first, then a slippery circle at the box edge: the wrap
    # happens, then the push uses the space's minimum image
    box = dict(kind=oracle_mod.WK_PERIODIC_RECT, geom=(0.0, 0.0, 20.0, 20.0))
    circ = circle(oracle_mod, oracle_mod.WK_SLIPPERY_CIRCLE, (0.5, 10.0), 2.0)
    s = oracle_mod.Sys(n=2, bl=(0.0, 0.0), size=(20.0, 20.0), periodic=True, kind=oracle_mod.KIND_LJ,
                       p=[1.0, 1.0, 0.0, 0.0], dt=1e-3, radius=R_P, walls=[box, circ])
    pos = np.array([[20.3, 10.0], [5.0, 5.0]])  # wraps to 0.3, then sits inside the circle
    p, _ = s.walls_apply(pos, None, 0, R_P)
    q = p[0] - (0.5, 10.0)
    q -= 20.0 * np.round(q / 20.0)
    assert np.linalg.norm(q) == pytest.approx(2.0 + R_P, rel=1e-14)


def test_verlet_in_rigid_circle_keeps_speed(oracle_mod):
    # one free particle bouncing in a circle centred at the origin: the
    # reflections keep its speed
    w = circle(oracle_mod, oracle_mod.WK_RIGID_CIRCLE, (0.0, 0.0), 5.0)
    s = oracle_mod.Sys(n=1, bl=(-5.0, -5.0), size=(10.0, 10.0), periodic=False, kind=oracle_mod.KIND_LJ,
                       p=[1.0, 1.0, 2.5, 0.0], dt=1e-2, radius=R_P, walls=[w])
    pos, vel, _, _, _ = s.newton_steps(np.array([[0.0, 0.0]]), np.array([[3.0, 1.7]]), 2000)
    assert np.linalg.norm(vel[0]) == pytest.approx(math.hypot(3.0, 1.7), rel=1e-12)
    assert np.linalg.norm(pos[0]) < 5.0
