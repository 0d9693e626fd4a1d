"""Pins the oracle's chunk binning (src/chunks.jl) without reusing its formulas:
exact rational arithmetic for Julia's float div, enumeration for the stencil,
invariants for the chunk arrays."""
import math
from fractions import Fraction

import numpy as np
import pytest


def exact_trunc_quotient(a: float, b: float) -> int:
    """Julia's div(a, b) for Float64 is the truncated quotient of the real
    numbers a/b; computed here with exact rationals."""
    q = Fraction(a) / Fraction(b)
    return math.trunc(q)


def adversarial_values(cell, ncell, rng):
    vals = []
    for k in range(0, ncell + 1):
        x = k * cell
        vals += [x, np.nextafter(x, -np.inf), np.nextafter(x, np.inf)]
    vals += list(rng.uniform(0, cell * ncell, 200))
    return [v for v in vals if v >= 0.0]


def test_jl_div_is_exact_truncated_quotient(oracle_mod):
    rng = np.random.default_rng(5)
    for cell in (2.5486, 43.327 / 17, 1223.914 / 489, 0.1, 1.0, 3.0):
        for a in adversarial_values(cell, 20, rng):
            got = oracle_mod.jl_div(a, cell)
            assert got == exact_trunc_quotient(a, cell), (a, cell)
    # negative numerators truncate toward zero
    for a in (-1e-12, -0.3, -2.7, -5.0):
        assert oracle_mod.jl_div(a, 2.5) == exact_trunc_quotient(a, 2.5)


def test_float_division_differs_where_div_does_not():
    # the case the exact quotient exists for: a/b rounds up onto an integer
    found = False
    for k in range(1, 2000):
        b = 0.1
        a = np.nextafter(k * b, -np.inf)
        if math.trunc(a / b) != exact_trunc_quotient(a, b):
            found = True
            break
    assert found


def test_chunk_capacity_hand_values(oracle_mod):
    r_lj = 2 ** (1 / 6) / 2
    # 17x17 chunks over the 32x32 LJ lattice: chunk 2.5486, ceil(2.27)=3 -> (4*4)*2
    assert oracle_mod.chunk_capacity(43.32704 / 17, 43.32704 / 17, r_lj) == 32
    # chunk equal to 2 radii: ceil(1)=1 -> (2*2)*2
    assert oracle_mod.chunk_capacity(1.0, 1.0, 0.5) == 8
    # chunk 1.2, radius 0.5 -> ceil(1.2)=2 -> 3*3*2
    assert oracle_mod.chunk_capacity(1.2, 1.2, 0.5) == 18
    from mavi_hip.configs import chunk_capacity

    for cl, r in ((2.5486, r_lj), (1.0, 0.5), (1.2, 0.5), (7.3, 0.4)):
        assert chunk_capacity(cl, cl * 1.1, r) == oracle_mod.chunk_capacity(cl, cl * 1.1, r)


def exact_cell(x, y, bl, size, cols, rows):
    L, H = size
    cl, ch = L / cols, H / rows
    row = exact_trunc_quotient((-y + bl[1]) + H, ch) + 1
    col = exact_trunc_quotient(x - bl[0], cl) + 1
    row -= 1 if row == rows + 1 else 0
    col -= 1 if col == cols + 1 else 0
    return row, col


@pytest.mark.parametrize("periodic", [False, True])
def test_bin_invariants_and_exact_cells(oracle_mod, periodic):
    rng = np.random.default_rng(11)
    L, H, cols, rows = 43.32704, 30.5, 17, 11
    n = 900
    pos = np.stack([rng.uniform(0, L, n), rng.uniform(0, H, n)], axis=1)
    # adversarial: exact chunk boundaries, +-1 ulp, the top/right edge (clamp)
    cl, ch = L / cols, H / rows
    edge = []
    for k in range(cols + 1):
        for x in (k * cl, np.nextafter(k * cl, -np.inf), np.nextafter(k * cl, np.inf)):
            if 0 <= x <= L:
                edge.append((x, H / 3))
    for k in range(rows + 1):
        for y in (k * ch, np.nextafter(k * ch, -np.inf), np.nextafter(k * ch, np.inf)):
            if 0 <= y <= H:
                edge.append((L / 2, y))
    edge += [(L, H), (0.0, 0.0), (L, 0.0), (0.0, H)]
    pos = np.concatenate([pos, np.asarray(edge)])
    n = len(pos)
    s = oracle_mod.Sys(n=n, bl=(0.0, 0.0), size=(L, H), periodic=periodic, kind=oracle_mod.KIND_LJ,
                       p=[1, 1, 0, 0], cols=cols, rows=rows, nc=64)
    err, cp, counts = s.bin(pos)
    assert err == 0
    assert counts.sum() == n
    seen = []
    for r in range(rows):
        for c in range(cols):
            ids = cp[: counts[r, c], r, c]
            assert list(ids) == sorted(ids)  # ascending inside a chunk
            for i in ids:
                assert exact_cell(pos[i - 1, 0], pos[i - 1, 1], (0, 0), (L, H), cols, rows) == (r + 1, c + 1)
            seen += list(ids)
    assert sorted(seen) == list(range(1, n + 1))


def test_bin_overflow_and_out_of_box(oracle_mod):
    s = oracle_mod.Sys(n=3, bl=(0.0, 0.0), size=(10.0, 10.0), periodic=False, kind=oracle_mod.KIND_LJ,
                       p=[1, 1, 0, 0], cols=3, rows=3, nc=2)
    err, _, _ = s.bin(np.array([[1.0, 1.0], [1.1, 1.0], [1.2, 1.0]]))
    assert err == 1  # capacity exceeded
    # up to one chunk past the right edge maps to the last column (the N+1
    # clamp); two chunks past is a BoundsError
    err, _, _ = s.bin(np.array([[1.0, 1.0], [11.0, 1.0], [1.2, 5.0]]))
    assert err == 0
    err, _, _ = s.bin(np.array([[1.0, 1.0], [14.0, 1.0], [1.2, 5.0]]))
    assert err == 2


def _adjacent_pairs(rows, cols, periodic):
    pairs = set()
    for r in range(1, rows + 1):
        for c in range(1, cols + 1):
            for dr in (-1, 0, 1):
                for dc in (-1, 0, 1):
                    if dr == dc == 0:
                        continue
                    rr, cc = r + dr, c + dc
                    if periodic:
                        rr = (rr - 1) % rows + 1
                        cc = (cc - 1) % cols + 1
                    elif not (1 <= rr <= rows and 1 <= cc <= cols):
                        continue
                    pairs.add(frozenset([(r, c), (rr, cc)]))
    return pairs


@pytest.mark.parametrize("rows,cols", [(3, 3), (4, 7), (9, 5), (17, 17)])
@pytest.mark.parametrize("periodic", [False, True])
def test_stencil_covers_each_adjacency_once(oracle_mod, rows, cols, periodic):
    nb = oracle_mod.neighbors(rows, cols, periodic)
    listed = []
    for cell, lst in nb.items():
        for other in lst:
            assert other != cell
            listed.append(frozenset([cell, other]))
    assert len(listed) == len(set(listed))  # no repeats
    assert set(listed) == _adjacent_pairs(rows, cols, periodic)


def test_particle_ids_typed_rings(oracle_mod):
    # typed rings list the first num_particles[type] slots of every ring
    s = oracle_mod.Sys(n=3 * 10, bl=(0, 0), size=(50, 50), periodic=True, kind=oracle_mod.KIND_RINGS,
                       p=[0, 0, 0, 0], nmax=10, nrings=3, types=[1, 2, 1], np_type=[10, 5],
                       ring_par=np.ones((2, 8)), inter=np.ones(16))
    ids = s.particle_ids()
    assert list(ids) == list(range(1, 11)) + list(range(11, 16)) + list(range(21, 31))
