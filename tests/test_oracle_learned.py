"""Pins the learned-propagator oracle (oracle/learned_oracle.py): the
shortlisted neighbour search against brute force, the rotation against its
inverse, and the feature layout on a hand-built configuration."""
import numpy as np

import learned_oracle as lo
import mlp_oracle as mo


def test_neighbour_lists_match_brute_force():
    rng = np.random.default_rng(0)
    L, H = 12.0, 9.0
    pos = np.stack([rng.uniform(-1, L + 1, 400), rng.uniform(0, H, 400)], axis=1)  # some outside [0, L)
    a, da = lo.neighbour_lists(pos, L, H, 1.25)
    b, db = lo.neighbour_lists(pos, L, H, 1.25, brute=True)
    for x, y, dx, dy in zip(a, b, da, db):
        assert np.array_equal(x, y)
        assert np.array_equal(dx, dy)


def test_rotation_inverse():
    rng = np.random.default_rng(1)
    th = rng.uniform(-10, 10, 50)
    v = rng.normal(size=(50, 2))
    r = lo.rot(np.cos(th), np.sin(th), v)
    back = lo.rot(np.cos(-th), np.sin(-th), r)
    assert np.allclose(back, v, atol=1e-14)
    assert np.allclose(np.linalg.norm(r, axis=1), np.linalg.norm(v, axis=1), atol=1e-14)


def test_feature_layout_by_hand():
    """Two 4-particle rings; ring 0's polarisation points along +y, so a
    neighbour at +x in the lab frame appears at -y in the ring frame."""
    nmax = 4
    sq = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    pos = np.concatenate([sq + 2.0, sq + np.array([4.1, 2.0])])  # rings 0 and 1, gap 1.1 in x
    pol = np.array([np.pi / 2, 0.0])
    forces = np.zeros_like(pos)
    areas = np.array([1.0, 1.0])
    one = np.array([1.0])
    X, e, over = lo.features(pos, pol, forces, areas, 20.0, 20.0, nmax, np.array([0, 0]), np.array([4]), one * 4,
                             one * 1.0, one * 1.0, one, False, 1.25, 0.01)
    assert not over
    Xf = mo.bf16_bits_to_f64(X)
    # particle 1 (ring 0, at (3, 2)): ring neighbours 0 at (2,2) and 2 at (3,3), other ring: 4 at (4.1, 2)
    s = Xf[1]
    assert np.allclose(s[2:4], lo.rot(0.0, 1.0, np.array([-1.0, 0.0])), atol=1e-2)   # prev
    assert np.allclose(s[4:6], lo.rot(0.0, 1.0, np.array([0.0, 1.0])), atol=1e-2)    # next
    assert np.allclose(s[6:8], [1.0, 0.0], atol=1e-6)                                 # v0 p in its own frame
    assert abs(s[1] - (1.0 - 1.0) / 1.0) < 1e-12                                      # A0 = (4*1/4)^2 = 1
    assert np.allclose(s[8:14], 0.0, atol=1e-12)  # same-ring neighbours move alike; no area change or turn
    nb = s[32:].reshape(32, 4)
    present = np.abs(nb[:, 2])
    assert present.sum() == 3
    # sorted by distance: the two ring neighbours (d = 1) by id, then particle 4 (d = 1.1)
    assert np.allclose(nb[0], [*lo.rot(0.0, 1.0, np.array([-1.0, 0.0])), 1, 0], atol=1e-2)
    assert np.allclose(nb[1], [*lo.rot(0.0, 1.0, np.array([0.0, 1.0])), 1, 0], atol=1e-2)
    # ring 1 moves along +x, ring 0 along +y: the gap to particle 4 opens at rate 1
    assert np.allclose(nb[2], [*lo.rot(0.0, 1.0, np.array([1.1, 0.0])), -1, 1.0], atol=1e-2)
    assert np.allclose(e[1], [0.0, 0.01], atol=1e-15)
