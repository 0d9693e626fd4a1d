"""GPU parity for the Rings path (src/rings/integration.jl) against the oracle,
with host-supplied rotational noise so both sides consume the same draws."""
import math
import sys

import numpy as np
import pytest

from helpers import FORCE_RTOL, TRAJ_RTOL, force_rel_err, traj_err

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def normal_case(use_chunks=True, n=10, cols=13, rows=13, seed=24042001):
    """The reference's create_system_normal (test/tests_rings/rings_utils.jl:27-99)
    with numpy's RNG in place of MersenneTwister."""
    import mavi_hip as mv
    from mavi_hip import rings

    inter = rings.HarmTruncCfg(k_rep=20, k_atr=4, dist_eq=1, dist_max=1 + 0.2)
    cfg = rings.RingsCfg(p0=3.5, relax_time=1, vo=1.0, mobility=1, rot_diff=0.05, k_area=1, k_spring=20,
                         l_spring=1, num_particles=n, interaction_finder=inter)
    rp, geom = rings.rectangular_grid(num_cols=cols, num_rows=rows, num_particles=n,
                                      p_radius=cfg.particle_radius(), pad_x=0.1, pad_y=0.1)
    rng = np.random.default_rng(seed)
    state = rings.RingsState(rings_pos=rp, pol=rings.random_pol(cols * rows, rng))
    space = mv.SpaceCfg(geometry_cfg=geom, wall_type=mv.PeriodicWalls())
    max_size = inter.dist_max * 1.1
    chunks = mv.ChunksCfg(int(geom.length // max_size), int(geom.height // max_size)) if use_chunks else None
    ic = mv.IntCfg(dt=0.01, chunks_cfg=chunks, device=mv.GPU())
    return state, space, cfg, ic


def types_case(use_chunks=True, seed=7):
    """The reference's create_system_types (test/tests_rings/rings_utils.jl:193-297)."""
    import mavi_hip as mv
    from mavi_hip import rings

    i1 = rings.HarmTruncCfg(k_rep=13, k_atr=0.1, dist_eq=1, dist_max=1 * 1.1)
    i2 = rings.HarmTruncCfg(k_rep=13, k_atr=0.1, dist_eq=0.8, dist_max=0.8 * 1.1)
    pd = i1.dist_eq / 2 + i2.dist_eq / 2
    ip = rings.HarmTruncCfg(k_rep=13, k_atr=10, dist_eq=pd, dist_max=pd * 1.2)
    finder = rings.InteractionMatrix([[i1, ip], [ip, i2]])
    nps = [10, 5]
    cfg = rings.RingsCfg(p0=3.5, relax_time=1, vo=1, mobility=1, rot_diff=0.5, k_area=1, k_spring=20,
                         l_spring=[i1.dist_eq * 0.8, i2.dist_eq * 0.8], num_particles=nps, interaction_finder=finder)
    rng = np.random.default_rng(seed)
    types = rng.integers(1, 3, size=25)
    rp, geom = rings.rectangular_grid(num_cols=5, num_rows=5, num_particles=nps, p_radius=[0.5, 0.4],
                                      types=list(types), pad_x=0.1, pad_y=0.1)
    state = rings.RingsState(rings_pos=rp, pol=rings.random_pol(25, rng), types=types, num_particles=nps)
    space = mv.SpaceCfg(geometry_cfg=geom, wall_type=mv.PeriodicWalls())
    max_size = max(i.dist_max for i in (i1, i2, ip)) * 1.1
    chunks = mv.ChunksCfg(int(geom.length // max_size), int(geom.height // max_size)) if use_chunks else None
    ic = mv.IntCfg(dt=0.01, chunks_cfg=chunks, device=mv.GPU())
    return state, space, cfg, ic


CASES = {
    "big_rings": lambda: normal_case(True, n=40, cols=5, rows=4),  # > 32 particles: one-thread-per-ring kernel
    "normal": lambda: normal_case(True),
    "normal_nochunks": lambda: normal_case(False),
    "types": lambda: types_case(True),
    "types_nochunks": lambda: types_case(False),
}


def build(name):
    from mavi_hip import rings

    state, space, cfg, ic = CASES[name]()
    pos0 = state.pos.copy()
    pol0 = state.pol.copy()
    sys_ = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic)
    return sys_, pos0, pol0, (state, space, cfg, ic)


@pytest.mark.parametrize("name", list(CASES))
def test_rings_forces(gpu_ctx, oracle_mod, name):
    say(f"[rings forces] {name}")
    sys_, pos0, pol0, (state, space, cfg, ic) = build(name)
    o = oracle_mod.from_rings(state, space, cfg, ic)
    f_ref, areas_ref = o.rings_forces(pos0)
    # sum_j |f_ij| over pair terms, plus the springs/area magnitudes bound the error scale
    _, ab = o.forces(pos0, with_absum=True)
    ab = ab + np.abs(f_ref).sum(axis=1)
    e = force_rel_err(sys_.forces, f_ref, ab)
    say(f"  rel err {e:.3e}")
    assert e <= FORCE_RTOL
    ids = state.particle_ids()
    assert np.allclose(sys_.areas, areas_ref, rtol=1e-13, atol=1e-13)
    # pads (typed rings) never receive a force
    pads = np.setdiff1d(np.arange(len(pos0)), ids)
    assert np.all(sys_.forces[pads] == 0)


@pytest.mark.parametrize("name", list(CASES))
def test_rings_trajectory(gpu_ctx, oracle_mod, name):
    say(f"[rings trajectory] {name}")
    sys_, pos0, pol0, (state, space, cfg, ic) = build(name)
    nsteps = 100
    rng = np.random.default_rng(99)
    noise = rng.standard_normal((nsteps, state.num_rings))
    sys_.upload()
    sys_.advance(nsteps, noise)
    sys_.download()
    o = oracle_mod.from_rings(state, space, cfg, ic)
    p_ref, pol_ref, f_ref, areas_ref, cms_ref = o.rings_steps(pos0, pol0, nsteps, noise)
    ids = state.particle_ids()
    ep, _ = traj_err(sys_.state.pos[ids], None, p_ref[ids], None, space.geometry_cfg.length)
    epol = float(np.abs(sys_.state.pol - pol_ref).max())
    raw = float(np.abs(sys_.state.pos[ids] - p_ref[ids]).max())
    say(f"  pos err {ep:.3e} (raw {raw:.3e}) pol err {epol:.3e}")
    assert ep <= TRAJ_RTOL and epol <= 1e-9
    assert np.allclose(sys_.cms, cms_ref, rtol=0, atol=1e-9 * space.geometry_cfg.length)


def test_rings_device_noise_statistics(gpu_ctx):
    """Device rotational noise: with vo = 0, no forces and the alignment term
    switched off (relax_time -> infinity) the polarisation performs a random
    walk of variance 2 D_r t per ring (checked statistically).  A finite
    relax_time would let round-off forces steer every ring toward its own
    fixed direction."""
    import mavi_hip as mv
    from mavi_hip import rings

    say("[rings device noise]")
    n, nrings = 6, 4096
    inter = rings.HarmTruncCfg(k_rep=1, k_atr=0, dist_eq=0.5, dist_max=0.5)
    side = 1.0
    area = n * side * side / (4 * math.tan(math.pi / n))
    cfg = rings.RingsCfg(p0=n * side / math.sqrt(area), relax_time=1e30, vo=0.0, mobility=1, rot_diff=0.5, k_area=1,
                         k_spring=1, l_spring=side, num_particles=n, interaction_finder=inter)
    m = 64
    R = side / (2 * math.sin(math.pi / n))
    ang = np.arange(n) * 2 * math.pi / n
    k = np.arange(nrings)
    centers = np.stack([(k % m + 0.5) * 4.0, (k // m + 0.5) * 4.0], axis=1)
    rp = centers[:, None, :] + R * np.stack([np.cos(ang), np.sin(ang)], axis=1)[None]
    state = rings.RingsState(rings_pos=rp, pol=np.zeros(nrings))
    space = mv.SpaceCfg(geometry_cfg=mv.RectangleCfg(length=4.0 * m, height=4.0 * m), wall_type=mv.PeriodicWalls())
    ic = mv.IntCfg(dt=0.01, chunks_cfg=mv.ChunksCfg(64, 64), device=mv.GPU())
    pol0 = state.pol.copy()
    sys_ = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic, seed=1234)
    sys_.advance(200)
    sys_.download()
    var = float(np.var(sys_.state.pol))
    expect = 2 * 0.5 * 0.01 * 200
    assert abs(var - expect) < 0.1 * expect
    assert abs(float(np.mean(sys_.state.pol))) < 4 * math.sqrt(expect / nrings)
    # same seed reproduces, another seed does not
    st2 = rings.RingsState(rings_pos=rp, pol=pol0)
    again = rings.RingsSystem(state=st2, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic, seed=1234)
    again.advance(200)
    again.download()
    assert np.array_equal(again.state.pol, sys_.state.pol)
    st3 = rings.RingsState(rings_pos=rp, pol=pol0)
    other = rings.RingsSystem(state=st3, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic, seed=99)
    other.advance(200)
    other.download()
    assert np.corrcoef(other.state.pol, sys_.state.pol)[0, 1] < 0.1


def test_rings_chunks_equal_nochunks(gpu_ctx):
    """The reference's check_chunks on the GPU: 100 steps with and without
    chunks agree (test/tests_rings/tests_general.jl:16-29, bar 1e-4)."""
    from mavi_hip import rings

    say("[rings chunks vs none]")
    a, _, _, _ = build("normal")
    b, _, _, _ = build("normal_nochunks")
    rng = np.random.default_rng(5)
    noise = rng.standard_normal((100, a.state.num_rings))
    a.advance(100, noise)
    b.advance(100, noise)
    a.download()
    b.download()
    d = float(((a.state.pos - b.state.pos) ** 2).sum())
    assert d < 1e-4


@pytest.mark.parametrize("shape", [(256, 1, 0), (256, 1, 1), (128, 0, 1)])
@pytest.mark.parametrize("name", ["normal", "types"])
def test_rings_kernel_shapes(gpu_ctx, oracle_mod, name, shape):
    from mavi_hip import _lib

    wg, lds, tp = shape
    say(f"[rings shape] {name} {shape}")
    sys_, pos0, pol0, (state, space, cfg, ic) = build(name)
    sys_.dev.set_tile(wg)
    sys_.dev.set_option(_lib.MAVI_OPT_LDS_TILES, lds)
    sys_.dev.set_option(_lib.MAVI_OPT_TWO_PHASE, tp)
    rng = np.random.default_rng(99)
    noise = rng.standard_normal((50, state.num_rings))
    sys_.advance(50, noise)
    sys_.download()
    o = oracle_mod.from_rings(state, space, cfg, ic)
    p_ref, pol_ref, _, _, _ = o.rings_steps(pos0, pol0, 50, noise)
    ids = state.particle_ids()
    ep, _ = traj_err(sys_.state.pos[ids], None, p_ref[ids], None, space.geometry_cfg.length)
    assert ep <= TRAJ_RTOL and float(np.abs(sys_.state.pol - pol_ref).max()) <= 1e-9


@pytest.mark.parametrize("ptype", ["all", "rings"])
@pytest.mark.parametrize("chunks", [True, False])
def test_contact_neighbors_vs_oracle(gpu_ctx, oracle_mod, ptype, chunks):
    """ParticleNeighbors (src/rings/neighbors.jl:117-137): after k steps the
    lists hold the pairs of the last forces! (positions after k-1 steps),
    equal as sets to the oracle's (the device sorts them by id)."""
    from mavi_hip import rings as R
    from mavi_hip import workloads as wl

    st, sp, cfg, ic = wl.rings_mixture(nrings=80, seed=7)
    if not chunks:
        ic.chunks_cfg = None
    k = 30
    noise = np.random.default_rng(3).standard_normal((k, st.num_rings)) * 0.3
    o = oracle_mod.from_rings(st, sp, cfg, ic)
    pos_km1, _, _, _, _ = o.rings_steps(st.pos.copy(), st.pol.copy(), k - 1, noise=noise[:k - 1])
    s = R.RingsSystem(state=st, space_cfg=sp, dynamic_cfg=cfg, int_cfg=ic,
                      p_neighbors_cfg=R.NeighborsCfg(type=ptype, tol=1.1))
    s.advance(k, noise)
    counts, lists = s.dev.rings_neighbors()
    rc, rl = o.rings_neighbors(pos_km1, tol=1.1, only_rings=(ptype == "rings"))
    assert counts.sum() > 0
    assert np.array_equal(counts, rc)
    for i in range(len(counts)):
        assert np.array_equal(lists[i, :counts[i]], np.sort(rl[i, :rc[i]]))
        assert np.all(lists[i, counts[i]:] == -1)
    assert np.array_equal(s.get_neigh_list(1), lists[0, :counts[0]])


def test_rings_fused_keys_graphs_equal_single_steps(gpu_ctx):
    """The ring update does the next binning's first pass.  16 steps in one
    call (two-step graphs that start from those keys) equal 16 one-step calls
    (direct launches) bit for bit, also after the state is uploaded again in
    the middle of the run (the keys are dropped and the binning starts from
    scratch)."""
    say("[rings fused keys] graphs vs single steps")
    a, pos0, pol0, _ = build("normal")
    b, _, _, _ = build("normal")
    for s_ in (a, b):
        s_.upload()
    a.advance(16)
    for _ in range(16):
        b.advance(1)
    for s_ in (a, b):
        s_.download()
    assert np.array_equal(a.state.pos, b.state.pos) and np.array_equal(a.state.pol, b.state.pol)
    mid_pos, mid_pol = a.state.pos.copy(), a.state.pol.copy()
    for s_ in (a, b):  # a new state over the kept keys of the last step
        s_.state.pos[...] = pos0
        s_.state.pol[...] = pol0
        s_.upload()
    a.advance(8)
    for _ in range(8):
        b.advance(1)
    for s_ in (a, b):
        s_.download()
    assert np.array_equal(a.state.pos, b.state.pos) and np.array_equal(a.state.pol, b.state.pol)
    assert not np.array_equal(a.state.pos, mid_pos) or not np.array_equal(a.state.pol, mid_pol)
