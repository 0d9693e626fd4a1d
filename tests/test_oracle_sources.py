"""Pins the variable ring population (VarRingsIds, sources and sinks;
src/rings/states.jl:20-39, 173-223, src/rings/sources.jl,
src/rings/integration.jl:353-372) in the oracle and the mirror's Source
construction: box geometry by hand, ChunksChecker lists by brute-force
rectangle overlap, spawns exactly where no particle of the step's start
lies in a (padded) box, removals exactly where the centre of mass the step
computed lies in the sink, the lowest free slot taken, uids counting up,
and the reference's own check that chunks and no chunks give the same run
(test/tests_rings/runtests.jl:6, check_chunks(:source))."""
import math

import numpy as np
import pytest


def source_case(oracle_mod, *, use_chunks=True, nslots=24, pol=0.0, length_d=8, seed=1):
    """test/tests_rings/rings_utils.jl:101-190 (create_system_source) in a
    shorter box, with a fixed spawn polarisation so that rings travel to
    the sink."""
    import mavi_hip as mv
    from mavi_hip import rings
    from mavi_hip.systems import Chunks

    n = 10
    inter = rings.HarmTruncCfg(k_rep=30, k_atr=3, dist_eq=1, dist_max=1.5)
    cfg = rings.RingsCfg(p0=3.545, relax_time=100, vo=1, mobility=1, rot_diff=0, k_area=3, k_spring=30,
                         l_spring=0.8, num_particles=n, interaction_finder=inter)
    ring_d = rings.equilibrium_ring_radius(cfg, n) * 2
    geom = mv.RectangleCfg(length=length_d * ring_d, height=6 * ring_d)
    r = (rings.equilibrium_area(cfg, n) / math.pi) ** 0.5
    spawn = rings.create_circle((0, 0), r, n)
    rng = np.random.default_rng(seed)
    rp = np.full((nslots, n, 2), 0.0)
    state = rings.RingsState(rings_pos=rp, pol=rings.random_pol(nslots, rng),
                             active_state=rings.ActiveState(np.zeros(nslots, bool)))
    space = mv.SpaceCfg(geometry_cfg=geom, wall_type=mv.PeriodicWalls())
    max_size = 1.5 * 1.2
    pch = mv.ChunksCfg(int(geom.length // max_size), int(geom.height // max_size)) if use_chunks else None
    ic = rings.RingsIntCfg(dt=0.01, p_chunks_cfg=pch)
    chunks = Chunks(pch.num_cols, pch.num_rows, geom, cfg.particle_radius()) if use_chunks else None
    src = rings.Source(rings.SourceCfg(bottom_left=(ring_d, geom.height / 2 - ring_d), spawn_pos=spawn, size=(2, 2),
                                       spawn_pol=pol, pad=cfg.particle_radius()), chunks)
    sink = rings.SinkCfg(mv.RectangleCfg(length=ring_d * 2, height=ring_d * 2,
                                         bottom_left=(geom.length - 3 * ring_d, geom.height / 2 - ring_d)))
    return state, space, cfg, ic, [src, sink], ring_d


def make_oracle(oracle_mod, state, space, cfg, ic, sources, pol_draws=None):
    o = oracle_mod.from_rings(state, space, cfg, ic)
    o.set_var_rings(state.rings_ids.mask.astype(np.int32), oracle_mod.source_dicts(sources, state.nmax),
                    pol_draws=pol_draws)
    o.rings_var_init(state.pos)
    return o


def test_source_boxes_and_chunk_lists(oracle_mod):
    state, space, cfg, ic, (src, sink), ring_d = source_case(oracle_mod)
    b = np.array(src.bbox)
    L, H = b[0, 2], b[0, 3]
    pad = src.cfg.pad
    # size (2, 2), offset 0: boxes tile from bottom_left, j (y) fastest
    bl = np.array(src.cfg.bottom_left)
    want = [bl + (i * L, j * H) for i in range(2) for j in range(2)]
    assert np.allclose(b[:, :2], want, rtol=0, atol=1e-12)
    sp = np.asarray(src.cfg.spawn_pos)
    assert L == pytest.approx(sp[:, 0].max() - sp[:, 0].min() + 2 * pad, rel=1e-15)
    for k, pts in enumerate(src.bbox_spawn_pos):
        assert np.allclose(pts.min(0), b[k, :2] + pad, rtol=0, atol=1e-12)  # the template sits pad inside its box
        assert np.allclose(pts - pts[0], sp - sp[0], rtol=0, atol=1e-12)
    # ChunksChecker: every chunk overlapping the padded box (chunks are
    # smaller than the boxes, where the reference's edge test is an overlap test)
    g = space.geometry_cfg
    cols, rows = ic.chunks_cfg.num_cols, ic.chunks_cfg.num_rows
    cl, chh = g.length / cols, g.height / rows
    assert cl < L and chh < H
    for k in range(4):
        x0, y0 = b[k, 0] - pad, b[k, 1] - pad
        x1, y1 = x0 + L + 2 * pad, y0 + H + 2 * pad
        brute = []
        for row in range(rows):
            for col in range(cols):
                cx0, cy1 = col * cl, g.height - row * chh
                if cx0 <= x1 and cx0 + cl >= x0 and cy1 - chh <= y1 and cy1 >= y0:
                    brute.append(row * cols + col)
        assert sorted(src.chunk_ids[k]) == brute


def test_first_step_fills_every_box(oracle_mod):
    state, space, cfg, ic, sources, ring_d = source_case(oracle_mod, pol="random")
    draws = np.random.default_rng(4).random(64)
    o = make_oracle(oracle_mod, state, space, cfg, ic, sources, pol_draws=draws)
    pos, pol, *_ = o.rings_steps(state.pos, state.pol, 1)
    mask, ids, uids, na, cms, used = o.var_state()
    n = len(mask)
    assert na == 4 and list(np.nonzero(mask)[0]) == [0, 1, 2, 3]  # the lowest free slots
    assert list(uids[:4]) == [n + 1, n + 2, n + 3, n + 4]
    assert used == 4
    assert np.array_equal(ids[:4], [1, 2, 3, 4]) and np.array_equal(ids[4:], np.arange(5, n + 1))
    # the spawned polarisations are 2 pi u in box order; the step then turned them
    assert np.allclose(pol[:4], 2 * math.pi * draws[:4], atol=1e-3)
    # the empty slots keep their values
    assert np.array_equal(pos[40:], state.pos[40:])


def brute_step_check(o, state, sources, pos_before, mask_before, cms_now, mask_after, spawned_boxes):
    """Re-derives one step's sink and source decisions from the state at the
    step's start."""
    src, sink = sources
    act = np.nonzero(mask_before)[0]
    p_act = np.concatenate([pos_before[r * state.nmax:(r + 1) * state.nmax] for r in act] +
                           [np.zeros((0, 2))])
    pad = src.cfg.pad
    empty = []
    for bx, by, L, H in src.bbox:
        inside = ((p_act[:, 0] >= bx - pad) & (p_act[:, 0] <= bx + L + pad) &
                  (p_act[:, 1] >= by - pad) & (p_act[:, 1] <= by + H + pad))
        empty.append(not inside.any())
    g = sink.geometry_cfg
    removed = [r for r in act if (g.bottom_left[0] <= cms_now[r, 0] <= g.bottom_left[0] + g.length and
                                  g.bottom_left[1] <= cms_now[r, 1] <= g.bottom_left[1] + g.height)]
    return empty, removed


@pytest.mark.parametrize("use_chunks", [True, False])
def test_spawns_and_removals_follow_the_rules(oracle_mod, use_chunks):
    state, space, cfg, ic, sources, ring_d = source_case(oracle_mod, use_chunks=use_chunks)
    o = make_oracle(oracle_mod, state, space, cfg, ic, sources)
    pos, pol = state.pos.copy(), state.pol.copy()
    nm = state.nmax
    spawned = removed = 0
    mask_prev, _, uids_prev, _, _, _ = o.var_state()
    cpos_prev = o.var["cpos"].copy()
    for step in range(1400):
        pos_before = pos.copy()
        pos, pol, *_ = o.rings_steps(pos, pol, 1)
        mask, ids, uids, na, cms, _ = o.var_state()
        # update_cms! of this step: the mean of the previous continuous positions
        act0 = np.nonzero(mask_prev)[0]
        cms_now = np.zeros((len(mask), 2))
        for r in act0:
            cms_now[r] = cpos_prev[r * nm:(r + 1) * nm].sum(0) / nm
        empty, rem = brute_step_check(o, state, sources, pos_before, mask_prev, cms_now, mask, None)
        new = np.nonzero(mask & ~mask_prev.astype(bool))[0] if mask.dtype == bool else \
            np.nonzero((mask == 1) & (mask_prev == 0))[0]
        gone = np.nonzero((mask == 0) & (mask_prev == 1))[0]
        assert sorted(gone) == sorted(rem)
        assert len(new) == sum(empty)  # one ring per empty box (free slots remain)
        # lowest free slots, after this step's removals
        free = [r for r in range(len(mask)) if mask_prev[r] == 0 or r in rem]
        assert list(new) == free[:len(new)]
        assert all(uids[r] > uids_prev.max() for r in new)
        spawned += len(new)
        removed += len(gone)
        assert na == mask.sum() == int(mask_prev.sum()) + len(new) - len(gone)
        mask_prev, uids_prev = mask, uids
        cpos_prev = o.var["cpos"].copy()
    assert spawned > 4 and removed > 0  # the run spawned and removed rings


def test_chunks_equal_no_chunks(oracle_mod):
    """check_chunks(:source): the same run with and without particle chunks."""
    runs = []
    for use_chunks in (True, False):
        state, space, cfg, ic, sources, ring_d = source_case(oracle_mod, use_chunks=use_chunks, pol="random")
        draws = np.random.default_rng(9).random(400)
        o = make_oracle(oracle_mod, state, space, cfg, ic, sources, pol_draws=draws)
        pos, pol, *_ = o.rings_steps(state.pos, state.pol, 1500)
        runs.append((pos, pol, o.var_state()))
    (p1, q1, v1), (p2, q2, v2) = runs
    assert np.array_equal(v1[0], v2[0]) and np.array_equal(v1[2], v2[2]) and v1[5] == v2[5]
    act = np.nonzero(v1[0])[0]
    sl = np.concatenate([np.arange(r * 10, r * 10 + 10) for r in act])
    assert len(act) > 4
    assert np.abs(p1[sl] - p2[sl]).max() < 1e-9
