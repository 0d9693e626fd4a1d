"""Invasions of rings on the GPU (check_invasions!, update_invasions!,
src/rings/integration.jl:379-520, with r_chunks and InvasionsCfg,
src/rings/rings.jl:188-202, src/rings/configs.jl:334-350) against the
oracle through the C ABI.  Bar: the same set of (invasor, invaded, p_id)
records, exactly, at the ring points and centres of mass the check used."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def overlap_case(*, n=10, cols=8, rows=7, radius_k=1.5, r_div=4.5, chunks=True, r_chunks=True, every=0,
                 types=False, seed=3):
    """Rings of the reference's create_system_normal lattice
    (test/tests_rings/rings_utils.jl:27-99) blown up by radius_k so that
    neighbouring rings overlap."""
    import mavi_hip as mv
    from mavi_hip import rings

    rng = np.random.default_rng(seed)
    if types:
        i1 = rings.HarmTruncCfg(k_rep=13, k_atr=0.1, dist_eq=1, dist_max=1.1)
        i2 = rings.HarmTruncCfg(k_rep=13, k_atr=0.1, dist_eq=0.8, dist_max=0.88)
        pd = 0.9
        ip = rings.HarmTruncCfg(k_rep=13, k_atr=10, dist_eq=pd, dist_max=pd * 1.2)
        finder = rings.InteractionMatrix([[i1, ip], [ip, i2]])
        nps = [10, 7]
        cfg = rings.RingsCfg(p0=3.5, relax_time=1, vo=1, mobility=1, rot_diff=0.5, k_area=1, k_spring=20,
                             l_spring=[0.8, 0.64], num_particles=nps, interaction_finder=finder)
        tl = rng.integers(1, 3, size=cols * rows)
        rp, geom = rings.rectangular_grid(num_cols=cols, num_rows=rows, num_particles=nps, p_radius=[0.5, 0.4],
                                          types=list(tl), pad_x=0.1, pad_y=0.1, radius_k=radius_k)
        state = rings.RingsState(rings_pos=rp, pol=rings.random_pol(cols * rows, rng), types=tl, num_particles=nps)
        dmax = 1.1 * 1.1
    else:
        inter = rings.HarmTruncCfg(k_rep=20, k_atr=4, dist_eq=1, dist_max=1.2)
        cfg = rings.RingsCfg(p0=3.5, relax_time=1, vo=1.0, mobility=1, rot_diff=0.05, k_area=1, k_spring=20,
                             l_spring=1, num_particles=n, interaction_finder=inter)
        rp, geom = rings.rectangular_grid(num_cols=cols, num_rows=rows, num_particles=n,
                                          p_radius=cfg.particle_radius(), pad_x=0.1, pad_y=0.1, radius_k=radius_k)
        state = rings.RingsState(rings_pos=rp, pol=rings.random_pol(cols * rows, rng))
        dmax = 1.2 * 1.1
    # shift by a random offset so that rings straddle the periodic edges
    shift = rng.uniform(0, 1, 2) * (geom.length, geom.height)
    state.pos[:] = np.mod(state.pos + shift, (geom.length, geom.height))
    space = mv.SpaceCfg(geometry_cfg=geom, wall_type=mv.PeriodicWalls())
    pch = mv.ChunksCfg(int(geom.length // dmax), int(geom.height // dmax)) if chunks else None
    rch = mv.ChunksCfg(max(3, int(geom.length // r_div)), max(3, int(geom.height // r_div))) if r_chunks else None
    ic = rings.RingsIntCfg(dt=0.01, p_chunks_cfg=pch, r_chunks_cfg=rch,
                           invasions_cfg=None if every is None else rings.InvasionsCfg(every))
    return state, space, cfg, ic


def oracle_list(oracle_mod, state, space, cfg, ic, sys_):
    """The oracle's check at the device's ring points and centres of mass."""
    o = oracle_mod.from_rings(state, space, cfg, ic)
    _, cms, cpos = sys_.dev.rings_info()
    cols, rows, nc = sys_.r_chunks_grid or (0, 0, 0)
    return o, o.rings_invasions(cpos, cms, cols, rows, nc), cpos, cms


def as_set(recs):
    return {(int(a), int(b), int(c)) for a, b, c in recs}


@pytest.mark.parametrize("name", ["chunks", "all_pairs", "types"])
def test_check_at_creation(gpu_ctx, oracle_mod, name):
    from mavi_hip import rings

    say(f"[invasions creation] {name}")
    state, space, cfg, ic = overlap_case(r_chunks=name != "all_pairs", types=name == "types")
    pos0 = state.pos.copy()
    sys_ = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic)
    dev = sys_.dev.rings_invasions()
    o, ref, cpos, cms = oracle_list(oracle_mod, state, space, cfg, ic, sys_)
    say(f"  {len(dev)} records, oracle {len(ref)}")
    assert len(ref) > state.num_rings  # the case has invasions
    assert len(ref) == len(set(ref))
    assert as_set(dev) == set(ref)
    assert [tuple(r) for r in dev] == sorted(ref)  # ordered by (invasor, invaded, p_id)
    # the device's ring points and centres of mass are the oracle's
    cref = o.rings_continuous_pos(pos0)
    ids = state.particle_ids()
    assert np.array_equal(cpos[ids], cref[ids])
    nr, nm = state.num_rings, state.nmax
    cm_ref = np.array([cref[r * nm:r * nm + state.ring_num_particles(r + 1)].sum(0) /
                       state.ring_num_particles(r + 1) for r in range(nr)])
    assert np.allclose(cms, cm_ref, rtol=0, atol=1e-12 * space.geometry_cfg.length)
    assert sys_.invasions[0] == rings.Invasion(*map(int, dev[0]))


def test_scheduled_checks_follow_the_oracle(gpu_ctx, oracle_mod):
    """InvasionsCfg(5): after 11 steps the last check ran at num_steps = 10,
    at the ring points and centres of mass that step's forces! saw, which
    the device still holds."""
    from mavi_hip import rings

    say("[invasions schedule] host noise")
    state, space, cfg, ic = overlap_case(every=5)
    sys_ = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic)
    assert len(sys_.invasions) == 0  # steps_to_update > 0: no check at creation
    rng = np.random.default_rng(5)
    sys_.advance(11, rng.standard_normal((11, state.num_rings)))
    dev = sys_.dev.rings_invasions()
    _, ref, _, _ = oracle_list(oracle_mod, state, space, cfg, ic, sys_)
    say(f"  {len(dev)} records")
    assert len(ref) > 0
    assert as_set(dev) == set(ref)


def test_schedule_with_graphs_and_split_calls(gpu_ctx):
    """Device noise (graph replays between checks): the check of step 10 is
    the same whether the 12 steps run in one call or in several, and equals
    a manual check_invasions after 11 steps."""
    from mavi_hip import rings

    say("[invasions schedule] graphs")
    out = []
    for split in ((12,), (3, 4, 5), None):
        state, space, cfg, ic = overlap_case(every=5 if split else None, seed=11)
        sys_ = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic, seed=17)
        if split:
            for k in split:
                sys_.advance(k)
            out.append(as_set(sys_.dev.rings_invasions()))
        else:
            sys_.advance(11)
            out.append(as_set((i.invasor, i.invaded, i.p_id) for i in sys_.check_invasions()))
    say(f"  {[len(o) for o in out]} records")
    assert len(out[0]) > 0
    assert out[0] == out[1] == out[2]


def test_grid_errors(gpu_ctx):
    import mavi_hip as mv
    from mavi_hip import rings

    say("[invasions errors]")
    state, space, cfg, ic = overlap_case(every=None, r_chunks=False)
    sys_ = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic)
    # capacity 1 per ring chunk over a 3x3 grid: a chunk holds several rings
    sys_.dev.rings_invasions_cfg(False, 0, 3, 3, 1)
    with pytest.raises(mv.MaviError) as e:
        sys_.dev.rings_invasions_check()
    assert e.value.status == mv._lib.MAVI_E_CELL_OVERFLOW
    # periodic grids below 3 chunks per side are refused, as for particles
    with pytest.raises(mv.MaviError):
        sys_.dev.rings_invasions_cfg(False, 0, 2, 5, 8)
    # a centre of mass more than one chunk off the grid: one big ring whose
    # first particle wrapped across the left edge
    g = space.geometry_cfg
    nmax = state.nmax
    th = 2 * np.pi * np.arange(nmax) / nmax
    big = np.stack([g.length - 2.5 + 3.0 * np.cos(th), g.height / 2 + 3.0 * np.sin(th)], 1)
    state.rings_pos[0] = np.mod(big, (g.length, g.height))
    sys_.upload()  # ring points and centres of mass of this state
    _, cms, _ = sys_.dev.rings_info()
    cl = g.length / 20
    assert cms[0, 0] == pytest.approx(-2.5, abs=1e-9) and 2.5 > cl  # more than a chunk off the left edge
    sys_.dev.rings_invasions_cfg(False, 0, 20, 20, 64)
    with pytest.raises(mv.MaviError) as e:
        sys_.dev.rings_invasions_check()
    assert e.value.status == mv._lib.MAVI_E_OUT_OF_BOX
