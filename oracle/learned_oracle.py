"""CPU oracle of the learned propagator's feature builder and training
targets (numpy, FP64, then float32 -> BF16 like the device).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product
never imports it.

The features are this project's own definition (the reference has no learned
path); what they are built from follows the reference's rings model:
ring ids from the scalar index (src/rings/states.jl:137-140), calc_diff with
the minimum image (src/integration.jl:43-48), the target area
A0 = (n l / p0)^2 (src/rings/integration.jl:170) and the overdamped velocity
v = v0 (cos th, sin th) + mu F of update! (src/rings/integration.jl:300-351).

Per particle i of ring r with polarisation th (R = rotation by -th), with
u_i = v0 (cos th, sin th) + mu F_i:
  self  [0] type index, [1] (A_r - A0)/A0, [2:4] R d(prev), [4:6] R d(next),
        [6:8] R u_i, [8:10] R (u_prev - u_i), [10:12] R (u_next - u_i),
        [12] dA_r/dt / A0 with dA/dt = 1/2 sum_i cross(u_i, x_{i+1} - x_{i-1}),
        [13] asin(p x v_cm / |v_cm|), v_cm = mean u over the ring, [14:32] zero
  neighbours: the 32 nearest j != i with |d_ij| <= r_feat, ascending by
        (|d|^2, id): [32+4q : 32+4q+2] R d_ij, [+2] +1 same ring / -1 other,
        [+3] (u_j - u_i) . d_ij / |d_ij|
  explicit displacement e_i = k dt (v0 p + mu F_i)
  target = R (x_i(t + k dt) - x_i(t) - e_i) / s_t   (minimum image)

Parity pinned by: the neighbour sets against a brute-force all-pairs search
(scipy cKDTree is used only to shortlist), the rotation against its inverse
and the same-ring flags against the id arithmetic (tests/test_oracle_learned.py).
"""
from __future__ import annotations

import numpy as np

from mlp_oracle import f32_to_bf16_bits

N_NB, N_SELF, N_IN = 32, 32, 160


def min_image(d, size):
    half = size / 2.0
    return d - (np.abs(d) > half) * np.copysign(size, d)


def rot(cs, sn, v):
    return np.stack([cs * v[..., 0] + sn * v[..., 1], -sn * v[..., 0] + cs * v[..., 1]], axis=-1)


def neighbour_lists(pos, L, H, r_feat, brute=False):
    """For each particle, the ids j != i within r_feat (minimum image), sorted
    by (d^2, id).  Returns (lists, d2 lists)."""
    n = len(pos)
    out, outd = [], []
    if brute:
        for i in range(n):
            dx = min_image(pos[:, 0] - pos[i, 0], L)
            dy = min_image(pos[:, 1] - pos[i, 1], H)
            d2 = dx * dx + dy * dy
            ok = np.sqrt(d2) <= r_feat
            ok[i] = False
            js = np.nonzero(ok)[0]
            order = np.lexsort((js, d2[js]))
            out.append(js[order])
            outd.append(d2[js][order])
        return out, outd
    from scipy.spatial import cKDTree

    wrapped = np.stack([np.mod(pos[:, 0], L), np.mod(pos[:, 1], H)], axis=1)
    wrapped = np.where(wrapped >= np.array([L, H]), 0.0, wrapped)
    tree = cKDTree(wrapped, boxsize=(L, H))
    cand = tree.query_ball_point(wrapped, r_feat * (1 + 1e-9) + 1e-12)
    for i in range(n):
        js = np.asarray([j for j in cand[i] if j != i], dtype=np.int64)
        dx = min_image(pos[js, 0] - pos[i, 0], L)
        dy = min_image(pos[js, 1] - pos[i, 1], H)
        d2 = dx * dx + dy * dy
        ok = np.sqrt(d2) <= r_feat
        js, d2 = js[ok], d2[ok]
        order = np.lexsort((js, d2))
        out.append(js[order])
        outd.append(d2[order])
    return out, outd


def features(pos, pol, forces, areas, L, H, nmax, ring_type, np_t, p0_t, vo_t, mu_t, l_t, typed, r_feat, kdt):
    """Features (BF16 bits, [n][160]), explicit displacements ([n][2]) and
    whether a particle had more than 32 neighbours."""
    n = len(pos)
    ids = np.arange(n)
    ring = ids // nmax
    i = ids - ring * nmax
    t = ring_type[ring]
    npr = np_t[t]
    sn, cs = np.sin(pol[ring]), np.cos(pol[ring])
    area0 = npr * l_t[t] / p0_t[t]
    area0 = area0 * area0
    prev = ring * nmax + np.where(i == 0, npr - 1, i - 1)
    nxt = ring * nmax + np.where(i == npr - 1, 0, i + 1)
    d = lambda a, b: np.stack([min_image(pos[a, 0] - pos[b, 0], L), min_image(pos[a, 1] - pos[b, 1], H)], axis=1)
    dp = rot(cs, sn, d(prev, ids))
    dn = rot(cs, sn, d(nxt, ids))
    u = np.stack([vo_t[t] * cs + mu_t[t] * forces[:, 0], vo_t[t] * sn + mu_t[t] * forces[:, 1]], axis=1)
    euler = kdt * u
    ur = rot(cs, sn, u)
    dup = rot(cs, sn, u[prev] - u)
    dun = rot(cs, sn, u[nxt] - u)
    # per-ring rates, summed in ring order
    nr = len(pol)
    rt = ring_type
    da = np.zeros(nr)
    vx = np.zeros(nr)
    vy = np.zeros(nr)
    rings = np.arange(nr)
    for q in range(int(np_t.max())):
        ok = np_t[rt] > q
        rr = rings[ok]
        npq = np_t[rt[rr]]
        me = rr * nmax + q
        pv = rr * nmax + np.where(q == 0, npq - 1, q - 1)
        nx = rr * nmax + np.where(q == npq - 1, 0, q + 1)
        dd = d(nx, pv)
        da[rr] += u[me, 0] * dd[:, 1] - u[me, 1] * dd[:, 0]
        vx[rr] += u[me, 0]
        vy[rr] += u[me, 1]
    da = 0.5 * da
    a0r = np_t[rt] * l_t[rt] / p0_t[rt]
    a0r = a0r * a0r
    rate = da / a0r
    vx = vx / np_t[rt]
    vy = vy / np_t[rt]
    speed = np.sqrt(vx * vx + vy * vy)
    rs, rc = np.sin(pol), np.cos(pol)
    with np.errstate(invalid="ignore", divide="ignore"):
        cr = np.where(speed != 0.0, (rc * vy - rs * vx) / speed, 0.0)
    omega = np.arcsin(np.clip(cr, -1.0, 1.0))
    X = np.zeros((n, N_IN), dtype=np.float64)
    X[:, 0] = t if typed else 0.0
    X[:, 1] = (areas[ring] - area0) / area0
    X[:, 2:4], X[:, 4:6], X[:, 6:8] = dp, dn, ur
    X[:, 8:10], X[:, 10:12] = dup, dun
    X[:, 12], X[:, 13] = rate[ring], omega[ring]
    lists, d2s = neighbour_lists(pos, L, H, r_feat)
    for a in range(n):
        js = lists[a][:N_NB]
        if len(js) == 0:
            continue
        dl = np.stack([min_image(pos[js, 0] - pos[a, 0], L), min_image(pos[js, 1] - pos[a, 1], H)], axis=1)
        dd = rot(cs[a], sn[a], dl)
        same = (js // nmax) == ring[a]
        rdot = ((u[js, 0] - u[a, 0]) * dl[:, 0] + (u[js, 1] - u[a, 1]) * dl[:, 1]) / np.sqrt(d2s[a][:N_NB])
        blk = np.zeros((len(js), 4))
        blk[:, 0:2] = dd
        blk[:, 2] = np.where(same, 1.0, -1.0)
        blk[:, 3] = rdot
        X[a, N_SELF:N_SELF + 4 * len(js)] = blk.ravel()
    over = max((len(v) for v in lists), default=0) > N_NB
    return f32_to_bf16_bits(X.astype(np.float32)), euler, over


def targets(pos0, pos1, pol0, euler, nmax, L, H, s_t):
    ring = np.arange(len(pos0)) // nmax
    sn, cs = np.sin(pol0[ring]), np.cos(pol0[ring])
    dx = min_image(pos1[:, 0] - pos0[:, 0], L)
    dy = min_image(pos1[:, 1] - pos0[:, 1], H)
    r = rot(cs, sn, np.stack([dx - euler[:, 0], dy - euler[:, 1]], axis=1))
    return (r / s_t).astype(np.float32)
