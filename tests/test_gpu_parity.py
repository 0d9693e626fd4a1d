"""GPU parity: libmavi_hip (through its C ABI) against the CPU oracle.

Bars: chunk arrays bit-exact; forces within 1e-12 of the oracle relative to
sum_j |f_ij|; 100-step trajectories within 1e-9 of the box size (positions)
and of the largest speed (velocities).  Full-size (1M particle) cases are
checked through size-independent properties."""
import math
import sys

import numpy as np
import pytest

from helpers import FORCE_RTOL, TRAJ_RTOL, force_rel_err, random_gas, traj_err

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_plain(pos, vel, space, dyn, int_cfg):
    import mavi_hip as mv

    st = mv.SecondLawState(pos=pos, vel=vel)
    return mv.System(state=st, space_cfg=space, dynamic_cfg=dyn, int_cfg=int_cfg)


def plain_case(name):
    import mavi_hip as mv
    from mavi_hip import workloads as wl

    rng = np.random.default_rng(hash(name) % 2**32)
    if name == "lj_gas_1024_ref":
        return wl.lj_gas_1024(cutoff=None)
    if name == "lj_gas_1024_cut":
        return wl.lj_gas_1024(cutoff=2.5)
    if name.startswith("lj_periodic_small"):
        return wl.lj_periodic(nside=48, density=0.7, seed=3, cutoff=name.endswith("cut"))
    if name == "ht_rigid":
        n, L = 600, 30.0
        pos = random_gas(rng, n, L, dmin=0.7)
        vel = rng.normal(size=(n, 2)) * 0.5
        dyn = mv.HarmTruncCfg(k_rep=10.0, k_atr=1.0, dist_eq=1.0, dist_max=1.2)
        space = mv.SpaceCfg(wall_type=mv.RigidWalls(), geometry_cfg=mv.RectangleCfg(length=L, height=L))
        ic = mv.IntCfg(dt=1e-3, chunks_cfg=mv.ChunksCfg(num_cols=20, num_rows=20), device=mv.GPU())
        return pos, vel, space, dyn, ic
    if name == "ht_periodic":
        n, L = 900, 32.0
        pos = random_gas(rng, n, L, dmin=0.7)
        vel = rng.normal(size=(n, 2)) * 0.5
        dyn = mv.HarmTruncCfg(k_rep=10.0, k_atr=1.0, dist_eq=1.0, dist_max=1.2)
        space = mv.SpaceCfg(wall_type=mv.PeriodicWalls(),
                            geometry_cfg=mv.RectangleCfg(length=L, height=L, bottom_left=(-3.0, 5.0)))
        pos = pos + np.array([-3.0, 5.0])
        ic = mv.IntCfg(dt=1e-3, chunks_cfg=mv.ChunksCfg(num_cols=26, num_rows=25), device=mv.GPU())
        return pos, vel, space, dyn, ic
    if name == "lj_allpairs":
        n, L = 300, 25.0
        pos = random_gas(rng, n, L, dmin=0.95)
        vel = rng.normal(size=(n, 2)) * 0.3
        dyn = mv.LenJonesCfg(sigma=1.0, epsilon=1.0)
        space = mv.SpaceCfg(wall_type=mv.PeriodicWalls(), geometry_cfg=mv.RectangleCfg(length=L, height=L))
        ic = mv.IntCfg(dt=1e-3, chunks_cfg=None, device=mv.GPU())
        return pos, vel, space, dyn, ic
    raise KeyError(name)


PLAIN = ["lj_gas_1024_ref", "lj_gas_1024_cut", "lj_periodic_small_ref", "lj_periodic_small_cut", "ht_rigid",
         "ht_periodic", "lj_allpairs"]


@pytest.mark.parametrize("name", [n for n in PLAIN if n != "lj_allpairs"])
def test_binning_bitexact(gpu_ctx, oracle_mod, name):
    say(f"[binning] {name}")
    pos, vel, space, dyn, ic = plain_case(name)
    sys_ = make_plain(pos, vel, space, dyn, ic)
    cp, counts = sys_.update_chunks()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    err, cp_ref, counts_ref = o.bin(pos)
    assert err == 0
    assert np.array_equal(counts, counts_ref)
    for r in range(counts.shape[0]):
        for c in range(counts.shape[1]):
            k = counts[r, c]
            assert np.array_equal(cp[:k, r, c], cp_ref[:k, r, c]), (r, c)


def test_binning_adversarial_boundaries(gpu_ctx, oracle_mod):
    """Particles exactly on chunk edges, +-1 ulp, and on the top/right edge."""
    import mavi_hip as mv

    say("[binning] adversarial")
    L, H, cols, rows = 43.32704, 30.5, 17, 11
    cl, ch = L / cols, H / rows
    pts = []
    for k in range(cols + 1):
        for x in (k * cl, np.nextafter(k * cl, -1), np.nextafter(k * cl, 1e9)):
            if 0 <= x <= L:
                pts.append((x, H * 0.37))
    for k in range(rows + 1):
        for y in (k * ch, np.nextafter(k * ch, -1), np.nextafter(k * ch, 1e9)):
            if 0 <= y <= H:
                pts.append((L * 0.61, y))
    pts += [(L, H), (0.0, 0.0), (L, 0.0), (0.0, H)]
    pos = np.asarray(pts, dtype=np.float64)
    vel = np.zeros_like(pos)
    dyn = mv.HarmTruncCfg(k_rep=1.0, k_atr=1.0, dist_eq=0.01, dist_max=0.02)
    space = mv.SpaceCfg(wall_type=mv.RigidWalls(), geometry_cfg=mv.RectangleCfg(length=L, height=H))
    ic = mv.IntCfg(dt=1e-3, chunks_cfg=mv.ChunksCfg(num_cols=cols, num_rows=rows), device=mv.GPU())
    sys_ = make_plain(pos, vel, space, dyn, ic)
    cp, counts = sys_.update_chunks()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    err, cp_ref, counts_ref = o.bin(pos)
    assert err == 0 and np.array_equal(counts, counts_ref)
    for r in range(rows):
        for c in range(cols):
            k = counts[r, c]
            assert np.array_equal(cp[:k, r, c], cp_ref[:k, r, c])


@pytest.mark.parametrize("name", PLAIN)
def test_forces_parity(gpu_ctx, oracle_mod, name):
    say(f"[forces] {name}")
    pos, vel, space, dyn, ic = plain_case(name)
    sys_ = make_plain(pos, vel, space, dyn, ic)
    f = sys_.calc_forces().copy()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    f_ref, ab = o.forces(pos, with_absum=True)
    e = force_rel_err(f, f_ref, ab)
    say(f"  rel err {e:.3e}")
    assert e <= FORCE_RTOL


@pytest.mark.parametrize("name", PLAIN)
def test_trajectory_100_steps(gpu_ctx, oracle_mod, name):
    import mavi_hip as mv

    say(f"[trajectory] {name}")
    pos, vel, space, dyn, ic = plain_case(name)
    sys_ = make_plain(pos, vel, space, dyn, ic)
    mv.run_system(sys_, num_steps=100)
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, f_ref, _, _ = o.newton_steps(pos, vel, 100)
    ep, ev = traj_err(sys_.state.pos, sys_.state.vel, p_ref, v_ref, space.geometry_cfg.length)
    say(f"  pos err {ep:.3e}  vel err {ev:.3e}")
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL
    assert sys_.time_info.num_steps == 100


def test_observables(gpu_ctx, oracle_mod):
    import mavi_hip as mv

    say("[observables]")
    pos, vel, space, dyn, ic = plain_case("lj_gas_1024_ref")
    sys_ = make_plain(pos, vel, space, dyn, ic)
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    assert mv.kinetic_energy(sys_) == pytest.approx(o.kinetic_energy(vel), rel=1e-13)
    assert mv.potential_energy(sys_) == pytest.approx(o.potential_energy_lj(pos), rel=1e-12)
    ke, pe = sys_.dev.run_record(20)
    _, _, _, ke_ref, pe_ref = o.newton_steps(pos, vel, 20, record=True)
    assert np.allclose(ke, ke_ref, rtol=1e-12) and np.allclose(pe, pe_ref, rtol=1e-12)


def test_resident_equals_per_step(gpu_ctx):
    import mavi_hip as mv

    say("[resident vs per-step]")
    pos, vel, space, dyn, ic = plain_case("lj_periodic_small_cut")
    a = make_plain(pos, vel, space, dyn, ic)
    b = make_plain(pos, vel, space, dyn, ic)
    mv.run_system(a, num_steps=20)
    for _ in range(20):
        mv.newton_step(b)
    assert np.array_equal(a.state.pos, b.state.pos) and np.array_equal(a.state.vel, b.state.vel)
    assert a.time_info.time == b.time_info.time


def test_tf_matches_reference_time_loop(gpu_ctx):
    import mavi_hip as mv

    pos, vel, space, dyn, ic = plain_case("lj_gas_1024_cut")
    a = make_plain(pos, vel, space, dyn, ic)
    mv.run_system(a, tf=0.0105)
    t, n = 0.0, 0
    while t < 0.0105:
        t += 1e-3
        n += 1
    assert a.time_info.num_steps == n


def test_deterministic(gpu_ctx):
    say("[determinism]")
    pos, vel, space, dyn, ic = plain_case("lj_periodic_small_ref")
    a = make_plain(pos, vel, space, dyn, ic)
    b = make_plain(pos, vel, space, dyn, ic)
    a.advance(30)
    b.advance(30)
    a.download()
    b.download()
    assert np.array_equal(a.state.pos, b.state.pos) and np.array_equal(a.forces, b.forces)


def test_error_paths(gpu_ctx):
    """Overflow and out-of-box are reported as status codes (the reference's
    BoundsError), forced from the host with a small capacity / bad upload."""
    import mavi_hip as mv
    from mavi_hip import _lib

    say("[errors]")
    d = _lib.SysDesc()
    d.n = 3
    d.size[0] = d.size[1] = 10.0
    d.wall = _lib.MAVI_WALL_RIGID
    d.potential = _lib.MAVI_POT_HARMTRUNC
    d.pot[0], d.pot[1], d.pot[2], d.pot[3] = 1.0, 1.0, 1.0, 1.2
    d.dt = 1e-3
    d.cols = d.rows = 3
    d.nc = 2
    d.radius = 0.5
    dev = _lib.DeviceSystem(gpu_ctx, d)
    dev.upload(np.array([[1.0, 1.0], [1.1, 1.0], [1.2, 1.0]]), np.zeros((3, 2)))
    with pytest.raises(mv.CellOverflowError):
        dev.bin(export=False)
    dev.upload(np.array([[1.0, 1.0], [25.0, 1.0], [1.2, 5.0]]), np.zeros((3, 2)))
    with pytest.raises(mv.OutOfBoxError):
        dev.step(1)
    dev.upload(np.array([[1.0, 1.0], [np.nan, 1.0], [1.2, 5.0]]), np.zeros((3, 2)))
    with pytest.raises(mv.MaviError):
        dev.bin(export=False)
    # the system still works after an error
    dev.upload(np.array([[1.0, 1.0], [5.0, 5.0], [8.2, 5.0]]), np.zeros((3, 2)))
    dev.step(2)
    d.cols = d.rows = 2
    d.wall = _lib.MAVI_WALL_PERIODIC
    with pytest.raises(mv.UnsupportedOnGPU):
        _lib.DeviceSystem(gpu_ctx, d)


def test_custom_force_not_on_gpu(gpu_ctx):
    import mavi_hip as mv

    class RadialForce:  # examples/custom_force.jl:17-37
        force, min_dist = 1.0, 1.0

        def particle_radius(self):
            return 0.5

    pos = np.array([[1.0, 1.0], [2.0, 2.0]])
    space = mv.SpaceCfg(wall_type=mv.RigidWalls(), geometry_cfg=mv.RectangleCfg(length=5, height=5))
    with pytest.raises(mv.UnsupportedOnGPU):
        make_plain(pos, np.zeros_like(pos), space, RadialForce(), mv.IntCfg(dt=1e-3))


def test_full_size_properties(gpu_ctx, oracle_mod):
    """1,048,576 LJ particles (density 0.7, periodic, cutoff 2.5): bit-exact
    binning, Newton's third law, momentum conservation over 20 steps."""
    from mavi_hip import workloads as wl

    say("[full size] 1M binning")
    pos, vel, space, dyn, ic = wl.lj_periodic()
    sys_ = make_plain(pos, vel, space, dyn, ic)
    cp, counts = sys_.update_chunks()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    err, cp_ref, counts_ref = o.bin(pos)
    assert err == 0 and np.array_equal(counts, counts_ref)
    mask = np.arange(cp.shape[0])[:, None, None] < counts[None, :, :]
    assert np.array_equal(np.where(mask, cp, 0), np.where(mask, cp_ref, 0))
    say("[full size] forces")
    f = sys_.calc_forces()
    scale = np.abs(f).sum()
    assert np.abs(f.sum(axis=0)).max() <= 1e-13 * scale
    say("[full size] 20 steps")
    p0 = vel.sum(axis=0)
    sys_.advance(20)
    sys_.download()
    p1 = sys_.state.vel.sum(axis=0)
    assert np.abs(p1 - p0).max() <= 1e-9 * np.abs(vel).sum()
    assert np.all(np.isfinite(sys_.state.pos))
    g = space.geometry_cfg
    assert np.all(sys_.state.pos >= 0) and np.all(sys_.state.pos <= g.length)


def test_full_size_force_sample_vs_oracle(gpu_ctx, oracle_mod):
    """1M-particle forces against the oracle on the same input (full oracle
    force loop, a few seconds on one core)."""
    from mavi_hip import workloads as wl

    say("[full size] forces vs oracle")
    pos, vel, space, dyn, ic = wl.lj_periodic()
    sys_ = make_plain(pos, vel, space, dyn, ic)
    f = sys_.calc_forces()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    f_ref, ab = o.forces(pos, with_absum=True)
    e = force_rel_err(f, f_ref, ab)
    say(f"  rel err {e:.3e}")
    assert e <= FORCE_RTOL


def test_large_grid_binning_and_steps(gpu_ctx, oracle_mod):
    """A grid of 4800 x 1000 = 4.8M chunks: the one-pass scan runs 293 tiles
    of 16384 chunks, so its look-back crosses two windows of 256 tiles.
    Binning bit-exact against the oracle, then forces and 10 steps."""
    import mavi_hip as mv

    say("[large grid] 4.8M chunks, 120k particles")
    rng = np.random.default_rng(11)
    cols, rows = 4800, 1000
    L, H = float(cols), float(rows)
    n = 120_000
    pos = np.stack([rng.uniform(0.0, L, n), rng.uniform(0.0, H, n)], axis=1)
    vel = rng.normal(size=(n, 2)) * 0.3
    dyn = mv.HarmTruncCfg(k_rep=10.0, k_atr=1.0, dist_eq=1.0, dist_max=1.0)
    space = mv.SpaceCfg(wall_type=mv.PeriodicWalls(), geometry_cfg=mv.RectangleCfg(length=L, height=H))
    ic = mv.IntCfg(dt=1e-3, chunks_cfg=mv.ChunksCfg(num_cols=cols, num_rows=rows), device=mv.GPU())
    assert cols * rows > 256 * 16384
    sys_ = make_plain(pos, vel, space, dyn, ic)
    cp, counts = sys_.update_chunks()
    o = oracle_mod.from_system(n, space, dyn, ic)
    err, cp_ref, counts_ref = o.bin(pos)
    assert err == 0 and np.array_equal(counts, counts_ref)
    mask = np.arange(cp.shape[0])[:, None, None] < counts[None, :, :]
    assert np.array_equal(np.where(mask, cp, 0), np.where(mask, cp_ref, 0))
    del cp, cp_ref, mask
    say("[large grid] forces")
    f = sys_.calc_forces().copy()
    f_ref, ab = o.forces(pos, with_absum=True)
    assert force_rel_err(f, f_ref, ab) <= FORCE_RTOL
    say("[large grid] 10 steps")
    mv.run_system(sys_, num_steps=10)
    p_ref, v_ref, _, _, _ = o.newton_steps(pos, vel, 10)
    ep, ev = traj_err(sys_.state.pos, sys_.state.vel, p_ref, v_ref, L)
    say(f"  pos err {ep:.3e}  vel err {ev:.3e}")
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL


SHAPES = [(256, 0, 0), (128, 0, 1), (256, 1, 0), (256, 1, 1), (64, 0, 0), (512, 0, 1)]


@pytest.mark.parametrize("name", ["lj_gas_1024_ref", "lj_periodic_small_cut", "ht_rigid", "ht_periodic"])
@pytest.mark.parametrize("shape", SHAPES)
def test_kernel_shapes_parity(gpu_ctx, oracle_mod, name, shape):
    """Every pair-kernel shape the autotuner can pick (workgroup, LDS tiles,
    two-phase list) meets the force and 100-step trajectory bars."""
    from mavi_hip import _lib

    wg, lds, tp = shape
    say(f"[shape] {name} wg={wg} lds={lds} two_phase={tp}")
    pos, vel, space, dyn, ic = plain_case(name)
    sys_ = make_plain(pos, vel, space, dyn, ic)
    sys_.dev.set_tile(wg)
    sys_.dev.set_option(_lib.MAVI_OPT_LDS_TILES, lds)
    sys_.dev.set_option(_lib.MAVI_OPT_TWO_PHASE, tp)
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    f_ref, ab = o.forces(pos, with_absum=True)
    assert force_rel_err(sys_.calc_forces(), f_ref, ab) <= FORCE_RTOL
    sys_.advance(100)
    sys_.download()
    p_ref, v_ref, _, _, _ = o.newton_steps(pos, vel, 100)
    ep, ev = traj_err(sys_.state.pos, sys_.state.vel, p_ref, v_ref, space.geometry_cfg.length)
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL


def test_autotune_restores_state(gpu_ctx, oracle_mod):
    say("[autotune]")
    pos, vel, space, dyn, ic = plain_case("lj_periodic_small_cut")
    sys_ = make_plain(pos, vel, space, dyn, ic)
    best, shape = sys_.dev.autotune(10)
    assert best > 0 and shape["workgroup"] in (64, 128, 256, 512)
    sys_.download()
    assert np.array_equal(sys_.state.pos, pos) and np.array_equal(sys_.state.vel, vel)
    sys_.advance(50)
    sys_.download()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, _, _, _ = o.newton_steps(pos, vel, 50)
    ep, ev = traj_err(sys_.state.pos, sys_.state.vel, p_ref, v_ref, space.geometry_cfg.length)
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL


@pytest.mark.parametrize("name", [n for n in PLAIN if n != "lj_allpairs"])
def test_force_reuse_bitexact(gpu_ctx, name):
    """Reusing the previous step's second force evaluation (MAVI_OPT_REUSE_FORCES,
    the default) gives bit-identical trajectories to evaluating every time,
    also across calls that invalidate it (forces, binning, graph and direct
    paths, odd step counts)."""
    from mavi_hip import _lib

    say(f"[reuse] {name}")
    pos, vel, space, dyn, ic = plain_case(name)
    vel = vel * 8.0  # many chunk crossings and periodic wraps
    a = make_plain(pos.copy(), vel.copy(), space, dyn, ic)
    b = make_plain(pos.copy(), vel.copy(), space, dyn, ic)
    b.dev.set_option(_lib.MAVI_OPT_REUSE_FORCES, 0)
    for s_ in (a, b):
        s_.dev.step(37)
        s_.dev.forces()
        s_.dev.step(5)
        s_.dev.bin(export=False)
        s_.dev.step(3)
        s_.dev.step(12)
        s_.download()
    assert np.array_equal(a.state.pos, b.state.pos)
    assert np.array_equal(a.state.vel, b.state.vel)
    assert np.array_equal(a.forces, b.forces)
