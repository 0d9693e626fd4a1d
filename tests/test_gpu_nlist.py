"""Neighbour lists (MAVI_OPT_NLIST) of finite-range plain systems against the
oracle: the same pairs as the reference's chunk loop (every pair within
range at the binned positions; the step's chunks kept for the second
evaluation), so forces and trajectories meet the parity bars; rollbacks,
the fallback to the chunk gather and the independence from rebuild times."""
import sys

import numpy as np
import pytest

from helpers import FORCE_RTOL, TRAJ_RTOL, traj_err

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make(case):
    import mavi_hip as mv
    from test_gpu_parity import plain_case

    pos, vel, space, dyn, ic = plain_case(case)
    s = mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)
    return s, (pos, vel, space, dyn, ic)


@pytest.mark.parametrize("case,eligible", [("lj_periodic_small_cut", True), ("lj_gas_1024_cut", True),
                                           ("ht_rigid", True), ("ht_periodic", True),
                                           ("lj_gas_1024_ref", False), ("lj_periodic_small_ref", False)])
def test_eligibility(gpu_ctx, case, eligible):
    s, _ = make(case)
    info = s.dev.nlist_info()
    assert info["eligible"] == eligible and info["active"] == eligible


@pytest.mark.parametrize("case", ["lj_periodic_small_cut", "lj_gas_1024_cut", "ht_rigid", "ht_periodic"])
@pytest.mark.parametrize("speed", [1.0, 8.0])
def test_trajectory_vs_oracle(gpu_ctx, oracle_mod, case, speed):
    """100 steps against the oracle; at 8x the speed many particles change
    chunk every step, so the second evaluation's stale-chunk test is busy."""
    say(f"[nlist] {case} x{speed}")
    s, (pos, vel, space, dyn, ic) = make(case)
    vel = vel * speed
    s.state.vel[...] = vel
    s.upload()
    s.advance(100)
    s.download()
    info = s.dev.nlist_info()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, f_ref, _, _ = o.newton_steps(pos, vel, 100)
    ep, ev = traj_err(s.state.pos, s.state.vel, p_ref, v_ref, space.geometry_cfg.length,
                      periodic=ic is not None and space.wall_type.__class__.__name__ == "PeriodicWalls")
    say(f"  pos {ep:.2e} vel {ev:.2e} {info}")
    assert info["active"] and info["rebuilds"] >= 1
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL
    assert np.abs(s.forces - f_ref).max() <= 1e-9 * max(np.abs(f_ref).max(), 1e-300)


def test_rollbacks_keep_parity(gpu_ctx, oracle_mod):
    """A skin of 5% of the range and 64-step blocks: blocks overrun the
    displacement bound, roll back and run again after a rebuild."""
    from mavi_hip import _lib

    s, (pos, vel, space, dyn, ic) = make("lj_periodic_small_cut")
    vel = vel * 4.0
    s.state.vel[...] = vel
    s.upload()
    s.dev.set_option(_lib.MAVI_OPT_NLIST_SKIN, 50)
    s.dev.set_option(_lib.MAVI_OPT_NLIST_BLOCK, 64)
    s.advance(200)
    s.download()
    info = s.dev.nlist_info()
    say(f"[nlist rollbacks] {info}")
    assert info["rollbacks"] >= 1 and info["active"]
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, _, _, _ = o.newton_steps(pos, vel, 200)
    ep, ev = traj_err(s.state.pos, s.state.vel, p_ref, v_ref, space.geometry_cfg.length)
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL


def test_fallback_when_lists_overflow(gpu_ctx, oracle_mod):
    """A skin of twice the range gives lists of more than 64 partners: the
    lists switch off and the chunk gather runs the steps, with parity."""
    from mavi_hip import _lib

    s, (pos, vel, space, dyn, ic) = make("lj_periodic_small_cut")
    s.dev.set_option(_lib.MAVI_OPT_NLIST_SKIN, 2000)
    s.advance(30)
    s.download()
    info = s.dev.nlist_info()
    say(f"[nlist overflow] {info}")
    assert info["eligible"] and not info["active"]
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, _, _, _ = o.newton_steps(pos, vel, 30)
    ep, ev = traj_err(s.state.pos, s.state.vel, p_ref, v_ref, space.geometry_cfg.length)
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL


def test_independent_of_rebuild_times(gpu_ctx):
    """Lists are sorted by partner id: runs with other skins, blocks and call
    patterns (so other rebuild times and slot orders) are bit-identical, and
    so is a run with force reuse off."""
    from mavi_hip import _lib

    runs = []
    for skin, block, calls, reuse in ((0, 0, [60], 1), (300, 8, [60], 1), (80, 32, [7] * 8 + [4], 1),
                                      (0, 0, [1] * 60, 1), (0, 0, [60], 0)):
        s, _ = make("ht_periodic")
        s.state.vel[...] *= 6.0
        s.upload()
        s.dev.set_option(_lib.MAVI_OPT_NLIST_SKIN, skin)
        s.dev.set_option(_lib.MAVI_OPT_NLIST_BLOCK, block)
        s.dev.set_option(_lib.MAVI_OPT_REUSE_FORCES, reuse)
        for c in calls:
            s.advance(c)
        s.download()
        runs.append((s.state.pos.copy(), s.state.vel.copy(), s.forces.copy(), s.dev.nlist_info()))
    say(f"[nlist rebuild times] {[r[3]['rebuilds'] for r in runs]}")
    assert len({r[3]["rebuilds"] for r in runs}) > 1
    for r in runs[1:]:
        assert np.array_equal(r[0], runs[0][0]) and np.array_equal(r[1], runs[0][1])
        assert np.array_equal(r[2], runs[0][2])


def test_stale_pair_within_cutoff_left_out(gpu_ctx, oracle_mod):
    """LJ cut at 1.9 on 2 x 2 chunks: B starts two chunk columns from A and
    drifts to within the cutoff of A; the second evaluation keeps the
    step-start chunks, which are not adjacent, so the pair is left out there
    (B's velocity is unchanged) and counts from the next step on."""
    import mavi_hip as mv

    pos = np.array([[1.95, 3.0], [4.01, 3.0], [7.5, 5.5]])
    vel = np.array([[0.0, 0.0], [-20.0, 0.0], [0.0, 0.0]])
    dyn = mv.LenJonesCfg(sigma=1.0, epsilon=1.0, cutoff=1.9)
    space = mv.SpaceCfg(wall_type=mv.RigidWalls(), geometry_cfg=mv.RectangleCfg(length=8.0, height=6.0))
    ic = mv.IntCfg(dt=0.01, chunks_cfg=mv.ChunksCfg(num_cols=4, num_rows=3), device=mv.GPU())
    o = oracle_mod.from_system(3, space, dyn, ic)
    for nsteps in (1, 2, 3):
        s = mv.System(state=mv.SecondLawState(pos=pos.copy(), vel=vel.copy()), space_cfg=space, dynamic_cfg=dyn,
                      int_cfg=ic)
        assert s.dev.nlist_info()["active"]
        s.advance(nsteps)
        s.download()
        p_ref, v_ref, f_ref, _, _ = o.newton_steps(pos, vel, nsteps)
        assert np.allclose(s.state.pos, p_ref, rtol=0, atol=1e-13)
        assert np.allclose(s.state.vel, v_ref, rtol=1e-12, atol=1e-15)
        if nsteps == 1:
            # B drifted to 3.81: 1.86 from A, inside the cutoff, chunks 0 and 2 at the start
            assert abs(s.state.pos[1, 0] - 3.81) < 1e-12
            assert s.state.vel[1, 0] == -20.0 and np.all(s.forces == 0.0)
        else:
            assert s.state.vel[1, 0] != -20.0  # the next steps bin afresh: the pair acts


@pytest.mark.parametrize("skin", [150, 1000])
def test_small_periodic_list_grid(gpu_ctx, oracle_mod, skin):
    """A periodic box 3.2 wide on 3 x 3 chunks with LJ cut at 1: the list
    grid has 2 cells a side at skin 0.15 and 1 at skin 1.0, so the 3 x 3 cell
    stencil must visit each cell once (the wrapped neighbours coincide)."""
    import mavi_hip as mv
    from mavi_hip import _lib

    rng = np.random.default_rng(7)
    a = 3.2 / 3
    g = (np.arange(3) + 0.5) * a
    pos = np.stack(np.meshgrid(g, g, indexing="xy"), axis=-1).reshape(-1, 2)
    pos = pos + rng.uniform(-0.04, 0.04, size=pos.shape)  # separations > 0.98
    vel = rng.standard_normal(size=pos.shape) * 0.5
    dyn = mv.LenJonesCfg(sigma=0.9, epsilon=1.0, cutoff=1.0)
    space = mv.SpaceCfg(wall_type=mv.PeriodicWalls(), geometry_cfg=mv.RectangleCfg(length=3.2, height=3.2))
    ic = mv.IntCfg(dt=0.002, chunks_cfg=mv.ChunksCfg(num_cols=3, num_rows=3), device=mv.GPU())
    s = mv.System(state=mv.SecondLawState(pos=pos.copy(), vel=vel.copy()), space_cfg=space, dynamic_cfg=dyn,
                  int_cfg=ic)
    s.dev.set_option(_lib.MAVI_OPT_NLIST_SKIN, skin)
    s.advance(200)
    s.download()
    info = s.dev.nlist_info()
    say(f"[nlist small grid skin {skin}] {info}")
    assert info["active"]
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, f_ref, _, _ = o.newton_steps(pos, vel, 200)
    ep, ev = traj_err(s.state.pos, s.state.vel, p_ref, v_ref, 3.2, periodic=True)
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL
    assert np.abs(s.forces - f_ref).max() <= 1e-9 * max(np.abs(f_ref).max(), 1e-300)


@pytest.mark.parametrize("nside", [12, 15])
def test_list_grid_without_nearest_image_staging(gpu_ctx, oracle_mod, nside):
    """Periodic LJ boxes whose list grid is 4 x 4 (nside 12) or 6 x 6 (nside
    15) cells: under 7 cells a side the build tests every candidate with the
    minimum image (the nearest image of a candidate to the cell's first
    particle need not be the nearest to the others), with interior cells on
    the three-run stencil path; 100 steps at 4x the speed against the oracle."""
    import mavi_hip as mv
    from mavi_hip import workloads as wl

    pos, vel, space, dyn, ic = wl.lj_periodic(nside=nside, density=0.7, seed=5, cutoff=True)
    vel = vel * 4.0
    s = mv.System(state=mv.SecondLawState(pos=pos.copy(), vel=vel.copy()), space_cfg=space, dynamic_cfg=dyn,
                  int_cfg=ic)
    s.advance(100)
    s.download()
    info = s.dev.nlist_info()
    say(f"[nlist grid nside {nside}] {info}")
    assert info["active"] and info["rebuilds"] >= 1
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, f_ref, _, _ = o.newton_steps(pos, vel, 100)
    ep, ev = traj_err(s.state.pos, s.state.vel, p_ref, v_ref, space.geometry_cfg.length, periodic=True)
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL
    assert np.abs(s.forces - f_ref).max() <= 1e-9 * max(np.abs(f_ref).max(), 1e-300)


def test_block_sized_by_stale_step_rolls_back(gpu_ctx, oracle_mod):
    """Blocks sized by the displacement bound (block option 0): after 60
    steps the velocities are raised 6x, so the next block, sized at the old
    step length, overruns the bound; it rolls back, the step estimate takes
    the new length, and the run keeps parity with the oracle."""
    s, (pos, vel, space, dyn, ic) = make("lj_periodic_small_cut")
    s.advance(60)
    s.download()
    info0 = s.dev.nlist_info()
    pos1, vel1 = s.state.pos.copy(), s.state.vel.copy() * 6.0
    s.state.vel[...] = vel1
    s.upload()
    s.advance(100)
    s.download()
    info = s.dev.nlist_info()
    say(f"[nlist stale step] {info0} -> {info}")
    assert info["active"] and info["rollbacks"] > info0["rollbacks"]
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, _, _, _ = o.newton_steps(pos1, vel1, 100)
    ep, ev = traj_err(s.state.pos, s.state.vel, p_ref, v_ref, space.geometry_cfg.length)
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL
