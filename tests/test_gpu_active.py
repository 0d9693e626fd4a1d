"""Self-propelled particles on the GPU (SzaboCfg and RunTumbleCfg with a
SelfPropelledState, szabo_step! / rtp_step!, src/integration.jl:433-535)
against the CPU oracle, through the C ABI.

Bars: forces within 1e-12 of the oracle relative to sum_j |f_ij|; with the
same host noise, one step within 1e-13 of the box size (positions) and
1e-12 rad (angles), and 100 steps within 10x the spread of two oracle runs
whose initial angles differ by 1e-15 (the device's sin/cos/asin and the
host libm's may differ by an ulp, which the dynamics amplifies); with the
device Philox stream, the
statistics the dynamics fixes (variance 2 Dr dt of the rotational noise,
tumble probability rate dt, uniform new angles), bit-identical replays
and graph == direct launches."""
import math
import sys

import numpy as np
import pytest

from helpers import FORCE_RTOL, TRAJ_RTOL, force_rel_err

pytestmark = pytest.mark.gpu

ANG_TOL = 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sp_case(name, nside=24, seed=1):
    from mavi_hip import workloads as wl

    if name == "szabo":
        pos, ang, sp, dyn, ic = wl.szabo_lattice(nside=nside, seed=seed)
        rng = np.random.default_rng(seed)
        pos = np.mod(pos + rng.uniform(-0.08, 0.08, size=pos.shape), sp.geometry_cfg.length)
    else:
        pos, ang, sp, dyn, ic = wl.rtp_gas(nside=nside, seed=seed, density=0.85, tumble_rate=5.0)
    return pos, ang, sp, dyn, ic


def make_sp(pos, ang, sp, dyn, ic, seed=0, chunks=True):
    import mavi_hip as mv

    if not chunks:
        ic = mv.IntCfg(dt=ic.dt, chunks_cfg=None, device=mv.GPU())
    st = mv.SelfPropelledState(pos=pos, pol_angle=ang)
    return mv.System(state=st, space_cfg=sp, dynamic_cfg=dyn, int_cfg=ic, seed=seed), ic


def host_noise(name, nsteps, n, seed):
    rng = np.random.default_rng(seed)
    if name == "szabo":
        return rng.standard_normal((nsteps, n))
    return rng.random((nsteps, n, 2))


@pytest.mark.parametrize("name", ["szabo", "rtp"])
@pytest.mark.parametrize("chunks", [True, False])
def test_forces_vs_oracle(gpu_ctx, oracle_mod, name, chunks):
    say(f"[sp forces] {name} chunks={chunks}")
    pos, ang, sp, dyn, ic = sp_case(name)
    system, ic2 = make_sp(pos, ang, sp, dyn, ic, chunks=chunks)
    f = system.calc_forces()
    o = oracle_mod.from_system(len(pos), sp, dyn, ic2)
    fr, ab = o.forces(pos, with_absum=True)
    assert np.count_nonzero(ab) > len(pos) // 4  # the case has interactions
    assert force_rel_err(f, fr, ab) <= FORCE_RTOL


def wrapped_max(a, b, L):
    d = np.asarray(a) - np.asarray(b)
    d -= L * np.round(d / L)
    return float(np.abs(d).max())


@pytest.mark.parametrize("name", ["szabo", "rtp"])
def test_one_step_with_host_noise(gpu_ctx, oracle_mod, name):
    say(f"[sp one step] {name}")
    pos, ang, sp, dyn, ic = sp_case(name)
    n = len(pos)
    noise = host_noise(name, 1, n, 10)
    system, _ = make_sp(pos, ang, sp, dyn, ic)
    system.advance(1, noise=noise)
    system.download()
    o = oracle_mod.from_system(n, sp, dyn, ic)
    pr, ar, fr = o.sp_steps(pos, ang, 1, noise=noise)
    L = sp.geometry_cfg.length
    assert wrapped_max(system.state.pos, pr, L) <= 1e-13 * L
    assert np.abs(system.state.pol_angle - ar).max() <= 1e-12
    assert np.allclose(system.forces, fr, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", ["szabo", "rtp"])
def test_trajectory_with_host_noise(gpu_ctx, oracle_mod, name):
    """100 steps: the device and the oracle stay as close as the dynamics
    lets two runs stay whose initial angles differ by 1e-15 (device and host
    sin/cos/asin differ by an ulp; the Szabo step at dt 0.01 amplifies such
    differences by orders of magnitude)."""
    say(f"[sp trajectory] {name}")
    pos, ang, sp, dyn, ic = sp_case(name)
    n, k = len(pos), 100
    noise = host_noise(name, k, n, 11)
    system, _ = make_sp(pos, ang, sp, dyn, ic)
    system.advance(k, noise=noise)
    system.download()
    o = oracle_mod.from_system(n, sp, dyn, ic)
    pr, ar, fr = o.sp_steps(pos, ang, k, noise=noise)
    pq, aq, _ = o.sp_steps(pos, ang + 1e-15 * np.random.default_rng(1).standard_normal(n), k, noise=noise)
    L = sp.geometry_cfg.length
    ep, ea = wrapped_max(system.state.pos, pr, L), float(np.abs(system.state.pol_angle - ar).max())
    sp_, sa = wrapped_max(pq, pr, L), float(np.abs(aq - ar).max())
    say(f"[sp trajectory] {name}: device vs oracle |dx| {ep:.2e} |dtheta| {ea:.2e}; "
        f"oracle sensitivity |dx| {sp_:.2e} |dtheta| {sa:.2e}")
    assert ep <= max(10 * sp_, TRAJ_RTOL * L)
    assert ea <= max(10 * sa, ANG_TOL)
    assert ep <= 1e-4 * L and ea <= 1e-3
    if name == "rtp":
        assert np.count_nonzero(np.abs(ar - ang) > 1e-6) > n // 4  # tumbles happened


def test_szabo_device_noise_statistics(gpu_ctx):
    import mavi_hip as mv

    # isolated particles (no pair forces): one step turns each polarisation
    # by sqrt(2 Dr dt) xi, xi standard normal
    nside, rot, dt = 256, 0.2, 0.01
    a = 3.0
    L = nside * a
    g = np.arange(nside * nside)
    pos = np.stack([(g % nside + 0.5) * a, (g // nside + 0.5) * a], axis=1)
    ang = np.random.default_rng(0).random(len(pos)) * 2 * math.pi
    dyn = mv.SzaboCfg(vo=1.0, mobility=1.0, relax_time=1.0, k_rep=10.0, k_adh=0.75, r_eq=1.0, r_max=1.1,
                      rot_diff=rot)
    sp = mv.SpaceCfg(wall_type=mv.PeriodicWalls(), geometry_cfg=mv.RectangleCfg(length=L, height=L))
    ic = mv.IntCfg(dt=dt, chunks_cfg=mv.ChunksCfg(num_cols=nside, num_rows=nside), device=mv.GPU())
    system, _ = make_sp(pos, ang, sp, dyn, ic, seed=7)
    system.advance(1)
    system.download()
    d = (system.state.pol_angle - ang) / math.sqrt(2 * rot * dt)
    say(f"[szabo noise] mean {d.mean():.4f} var {d.var():.4f}")
    assert abs(d.mean()) < 5 / math.sqrt(len(d))
    assert abs(d.var() - 1.0) < 5 * math.sqrt(2 / len(d))
    # positions moved by vo pol dt exactly as the free motion
    exp = pos + dt * np.stack([np.cos(ang), np.sin(ang)], axis=1)
    assert np.allclose(system.state.pos, exp, rtol=0, atol=1e-12)


def test_rtp_device_tumbles(gpu_ctx):
    import mavi_hip as mv

    nside, rate, dt = 512, 2.0, 5e-3
    a = 3.0
    L = nside * a
    g = np.arange(nside * nside)
    pos = np.stack([(g % nside + 0.5) * a, (g // nside + 0.5) * a], axis=1)
    ang = np.full(len(pos), 1.0)
    dyn = mv.RunTumbleCfg(vo=1.0, sigma=1.0, epsilon=1.0, tumble_rate=rate)
    sp = mv.SpaceCfg(wall_type=mv.PeriodicWalls(), geometry_cfg=mv.RectangleCfg(length=L, height=L))
    ic = mv.IntCfg(dt=dt, chunks_cfg=mv.ChunksCfg(num_cols=nside, num_rows=nside), device=mv.GPU())
    system, _ = make_sp(pos, ang, sp, dyn, ic, seed=9)
    system.advance(1)
    system.download()
    t = system.state.pol_angle != 1.0
    p = rate * dt
    n = len(pos)
    say(f"[rtp tumbles] {t.sum()} of {n}, expected {p * n:.0f}")
    assert abs(t.sum() - p * n) < 5 * math.sqrt(n * p * (1 - p))
    new = system.state.pol_angle[t] / (2 * math.pi)
    assert new.min() >= 0.0 and new.max() < 1.0
    hist, _ = np.histogram(new, bins=8, range=(0, 1))
    assert hist.min() > 0.6 * len(new) / 8


@pytest.mark.parametrize("name", ["szabo", "rtp"])
def test_graph_equals_direct_and_replays(gpu_ctx, name):
    from mavi_hip import _lib

    say(f"[sp graphs] {name}")
    pos, ang, sp, dyn, ic = sp_case(name, nside=40)
    a, _ = make_sp(pos, ang, sp, dyn, ic, seed=3)
    b, _ = make_sp(pos, ang, sp, dyn, ic, seed=3)
    c, _ = make_sp(pos, ang, sp, dyn, ic, seed=4)
    b.dev.set_option(_lib.MAVI_OPT_GRAPHS, 0)
    for s_ in (a, b, c):
        s_.advance(7)
        s_.advance(12)
        s_.download()
    assert np.array_equal(a.state.pos, b.state.pos)
    assert np.array_equal(a.state.pol_angle, b.state.pol_angle)
    assert np.array_equal(a.forces, b.forces)
    assert not np.array_equal(a.state.pol_angle, c.state.pol_angle)  # another noise stream
    # a state uploaded over the chunk keys the last update left: the keys
    # are dropped and the binning starts from scratch, graphs or not
    for s_ in (a, b):
        s_.state.pos[...] = pos
        s_.state.pol_angle[...] = ang
        s_.upload()
        s_.advance(9)
        s_.download()
    assert np.array_equal(a.state.pos, b.state.pos)
    assert np.array_equal(a.state.pol_angle, b.state.pol_angle)


def test_unsupported_configurations(gpu_ctx):
    import mavi_hip as mv
    from mavi_hip.batch import Batch

    pos, ang, sp, dyn, ic = sp_case("szabo", nside=8)
    rigid = mv.SpaceCfg(wall_type=mv.RigidWalls(), geometry_cfg=sp.geometry_cfg)
    with pytest.raises(mv.UnsupportedOnGPU):
        make_sp(pos, ang, rigid, dyn, ic)
    system, _ = make_sp(pos, ang, sp, dyn, ic)
    with pytest.raises(mv.MaviError):
        system.dev.observables()
    b = Batch()
    with pytest.raises(mv.UnsupportedOnGPU):
        b.add(system.state, sp, dyn, ic)
