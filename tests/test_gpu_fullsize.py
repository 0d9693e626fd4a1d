"""Parity at the sizes of the benchmark workloads, through the C ABI, against
the CPU oracle:

  1. 1024 LJ particles, rigid walls, 1000 steps with kinetic and potential
     energy recorded after every step (cut at 2.5 sigma, and the
     reference's LJ without cutoff);
  2. 1,048,576 LJ particles, periodic, density 0.7: 10 steps;
  3. 40,000 active rings x 20 (800k particles): binning, forces, 5 steps
     with host noise, and 20 steps with the device noise stream fed to the
     oracle (oracle/philox.py restates it);
  4. the 512-replica ensemble (log-uniform 1k-64k particles, half LJ gases,
     half ring mixtures): 100 steps in one batch, every replica bit-identical
     to the same system stepped alone, three replicas (the largest of each
     kind and the smallest rings replica) against the oracle;
  5. 200k-particle ring system (10,000 rings): learned-propagator features
     and training targets against the feature oracle.

Bars as in the other parity tests: forces 1e-12 relative to sum_j |f_ij|,
trajectories 1e-9 of the box (positions) and of the largest speed
(velocities), energies 1e-9 relative, chunk arrays and batch-vs-alone
bit-exact."""
import os
import sys

import numpy as np
import pytest

import philox
from helpers import FORCE_RTOL, TRAJ_RTOL, force_rel_err, traj_err

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def plain_system(pos, vel, space, dyn, ic):
    import mavi_hip as mv

    return mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)


@pytest.mark.parametrize("cutoff", [2.5, None])
def test_lj_gas_1024_1000_steps_energies(gpu_ctx, oracle_mod, cutoff):
    from mavi_hip import workloads as wl

    say(f"[lj 1024] 1024 LJ, 1000 steps, energies every step (cutoff {cutoff})")
    pos, vel, space, dyn, ic = wl.lj_gas_1024(cutoff=cutoff)
    s = plain_system(pos, vel, space, dyn, ic)
    ke, pe = s.dev.run_record(1000)
    s.download()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, _, ke_ref, pe_ref = o.newton_steps(pos, vel, 1000, record=True)
    ep, ev = traj_err(s.state.pos, s.state.vel, p_ref, v_ref, space.geometry_cfg.length, periodic=False)
    eke = float(np.abs(ke - ke_ref).max() / np.abs(ke_ref).max())
    epe = float(np.abs(pe - pe_ref).max() / np.abs(pe_ref).max())
    say(f"  pos {ep:.2e} vel {ev:.2e} KE {eke:.2e} PE {epe:.2e}")
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL
    assert eke <= TRAJ_RTOL and epe <= TRAJ_RTOL


def test_lj_1m_10_steps_vs_oracle(gpu_ctx, oracle_mod):
    from mavi_hip import workloads as wl

    say("[lj 1M] 1M LJ, 10 steps vs oracle")
    pos, vel, space, dyn, ic = wl.lj_periodic()
    s = plain_system(pos, vel, space, dyn, ic)
    s.advance(10)
    s.download()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    p_ref, v_ref, f_ref, _, _ = o.newton_steps(pos, vel, 10, nthreads=NTHREADS)
    ep, ev = traj_err(s.state.pos, s.state.vel, p_ref, v_ref, space.geometry_cfg.length)
    say(f"  pos {ep:.2e} vel {ev:.2e}")
    assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL
    assert np.abs(s.forces - f_ref).max() <= 1e-9 * np.abs(f_ref).max()


def rings_800k():
    from mavi_hip import rings
    from mavi_hip import workloads as wl

    st, sp, cfg, ic = wl.rings_mixture(nrings=40000)
    pos0, pol0 = st.pos.copy(), st.pol.copy()
    s = rings.RingsSystem(state=st, space_cfg=sp, dynamic_cfg=cfg, int_cfg=ic, seed=2024)
    return s, (st, sp, cfg, ic), pos0, pol0


def test_rings_800k_binning_forces_steps(gpu_ctx, oracle_mod):
    say("[rings 800k] 40,000 rings x 20: binning, forces")
    s, (st, sp, cfg, ic), pos0, pol0 = rings_800k()
    o = oracle_mod.from_rings(st, sp, cfg, ic)
    cp, counts = s.dev.bin(export=True)
    err, cp_ref, counts_ref = o.bin(pos0)
    assert err == 0 and np.array_equal(counts, counts_ref)
    mask = np.arange(cp.shape[0])[:, None, None] < counts[None, :, :]
    assert np.array_equal(np.where(mask, cp, 0), np.where(mask, cp_ref, 0))
    del cp, cp_ref, mask
    f = s.dev.forces()
    f_ref, areas_ref = o.rings_forces(pos0)
    _, ab = o.forces(pos0, with_absum=True)
    e = force_rel_err(f, f_ref, ab + np.abs(f_ref).sum(axis=1))
    say(f"  forces rel err {e:.2e}")
    assert e <= FORCE_RTOL
    say("[rings 800k] 5 steps, host noise")
    noise = np.random.default_rng(33).standard_normal((5, st.num_rings))
    s.upload()
    s.advance(5, noise)
    s.download()
    p_ref, pol_ref, _, _, cms_ref = o.rings_steps(pos0, pol0, 5, noise)
    L = sp.geometry_cfg.length
    ep, _ = traj_err(s.state.pos, None, p_ref, None, L)
    epol = float(np.abs(s.state.pol - pol_ref).max())
    say(f"  pos {ep:.2e} pol {epol:.2e}")
    assert ep <= TRAJ_RTOL and epol <= 1e-9
    assert np.allclose(s.cms, cms_ref, rtol=0, atol=1e-9 * L)


def test_rings_800k_device_noise_vs_oracle(gpu_ctx, oracle_mod):
    """20 steps with the device's Philox4x32-10 rotational noise; the oracle
    is fed the same draws (oracle/philox.py, pinned by the Random123
    known-answer vectors)."""
    say("[rings 800k] 20 steps, device noise")
    s, (st, sp, cfg, ic), pos0, pol0 = rings_800k()
    seed, counter = s.dev.noise_state()
    assert counter == 0
    s.advance(20)
    s.download()
    # the step's binning advances the counter before the ring update draws:
    # step j of a run that starts at counter c draws at c + j
    noise = philox.ring_noise(seed, counter + 1, 20, st.num_rings)
    o = oracle_mod.from_rings(st, sp, cfg, ic)
    p_ref, pol_ref, _, _, _ = o.rings_steps(pos0, pol0, 20, noise)
    ep, _ = traj_err(s.state.pos, None, p_ref, None, sp.geometry_cfg.length)
    epol = float(np.abs(s.state.pol - pol_ref).max())
    say(f"  pos {ep:.2e} pol {epol:.2e}")
    assert ep <= TRAJ_RTOL and epol <= 1e-9


def test_batch_512_replicas(gpu_ctx, oracle_mod):
    from mavi_hip import batch as mb
    from mavi_hip import rings
    from mavi_hip import workloads as wl
    from test_gpu_batch import single_run

    say("[batch 512] 512 replicas: building")
    ens = wl.batch_ensemble()
    assert len(ens) == 512
    b = mb.Batch(ctx=gpu_ctx)
    inits = []
    for kind, case, seed in ens:
        if kind == "plain":
            pos, vel, space, dyn, ic = case
            from mavi_hip import SecondLawState

            b.add(SecondLawState(pos=pos, vel=vel), space, dyn, ic)
            inits.append((pos.copy(), vel.copy(), None))
        else:
            st, space, cfg, ic = case
            b.add(st, space, cfg, ic, seed=seed)
            inits.append((st.pos.copy(), None, st.pol.copy()))
    b.finalize()
    for i, (p, v, pol) in enumerate(inits):
        b.upload(i, p, v, pol)
    info = b.info()
    say(f"  {info['particles'] if 'particles' in info else info}")
    say("[batch 512] 100 steps in one batch")
    b.run(100)
    out = [b.download(i) for i in range(len(ens))]
    forces = [b.get_forces(i) for i in range(len(ens))]
    sizes = [len(p) for p, _, _ in inits]
    for i, (kind, case, _) in enumerate(ens):
        pos, vel, pol = out[i]
        assert np.all(np.isfinite(pos)) and np.all(np.isfinite(forces[i]))
        if kind == "plain":
            p0 = inits[i][1].sum(axis=0)
            assert np.abs(vel.sum(axis=0) - p0).max() <= 1e-9 * np.abs(inits[i][1]).sum()
    say("[batch 512] every replica against the same system stepped alone")
    for i, (kind, case, seed) in enumerate(ens):
        if kind == "plain":
            pos, vel, space, dyn, ic = case
            case = (inits[i][0].copy(), inits[i][1].copy(), space, dyn, ic)
        else:
            st, space, cfg, ic = case
            # (copies: the system writes its downloads into the state's arrays)
            st = rings.RingsState(rings_pos=inits[i][0].reshape(st.num_rings, st.nmax, 2).copy(), pol=inits[i][2].copy(),
                                  num_particles=st.num_particles, types=st.types)
            case = (st, space, cfg, ic)
        want = single_run(kind, case, seed, [100])[0]
        pos, vel, pol = out[i]
        assert np.array_equal(pos, want[0]), f"replica {i} ({kind}, n={sizes[i]}) positions"
        if vel is not None:
            assert np.array_equal(vel, want[1]), f"replica {i} velocities"
        if pol is not None:
            assert np.array_equal(pol, want[2]), f"replica {i} polarisations"
        assert np.array_equal(forces[i], want[3]), f"replica {i} forces"
    say("[batch 512] three replicas against the oracle")
    plain_ids = [i for i, e in enumerate(ens) if e[0] == "plain"]
    ring_ids = [i for i, e in enumerate(ens) if e[0] == "rings"]
    picks = [max(plain_ids, key=lambda i: sizes[i]), max(ring_ids, key=lambda i: sizes[i]),
             min(ring_ids, key=lambda i: sizes[i])]
    for i in picks:
        kind, case, seed = ens[i]
        pos, vel, pol = out[i]
        if kind == "plain":
            _, _, space, dyn, ic = case
            o = oracle_mod.from_system(sizes[i], space, dyn, ic)
            p_ref, v_ref, _, _, _ = o.newton_steps(inits[i][0], inits[i][1], 100, nthreads=NTHREADS)
            ep, ev = traj_err(pos, vel, p_ref, v_ref, space.geometry_cfg.length)
        else:
            st, space, cfg, ic = case
            o = oracle_mod.from_rings(st, space, cfg, ic)
            noise = philox.ring_noise(seed, 1, 100, st.num_rings)
            p_ref, pol_ref, _, _, _ = o.rings_steps(inits[i][0], inits[i][2], 100, noise)
            ep, _ = traj_err(pos, None, p_ref, None, space.geometry_cfg.length)
            ev = float(np.abs(pol - pol_ref).max())
        say(f"  replica {i} ({kind}, n={sizes[i]}): pos {ep:.2e} vel/pol {ev:.2e}")
        assert ep <= TRAJ_RTOL and ev <= TRAJ_RTOL


def test_learned_features_targets_200k(gpu_ctx):
    """10,000 rings x 20: features of every particle and the k-step targets of
    one snapshot against oracle/learned_oracle.py."""
    from mavi_hip import learned, rings
    from mavi_hip import workloads as wl
    from test_gpu_learned import K, RF, bits_close, host_state, oracle_features
    import learned_oracle as lo

    say("[learned 200k] 200k features and targets")
    st, sp, cfg, ic = wl.rings_mixture(nrings=10000, seed=100, chunk=RF)
    system = rings.RingsSystem(state=st, space_cfg=sp, dynamic_cfg=cfg, int_cfg=ic, seed=100)
    system.advance(40)
    net = learned.MLP(ctx=gpu_ctx, seed=1)
    lp = learned.Learned(system.dev, net, k=K, r_feat=RF, target_scale=1e-3)
    x, euler = lp.features()
    (xr, er, over), _ = oracle_features(system, cfg)
    assert not over and x.shape == (200000, 160)
    dmax, frac = bits_close(x, xr)
    say(f"  features: max bf16 ulps {dmax}, fraction differing {frac:.2e}")
    if dmax > 1:  # name the rows and feature columns that differ
        bad = np.abs(x.astype(np.int32) - xr.astype(np.int32)) > 1
        rows = np.nonzero(bad.any(axis=1))[0]
        cols = np.bincount(np.nonzero(bad)[1], minlength=160)
        say(f"  {len(rows)} rows differ, e.g. {rows[:8].tolist()}; columns {np.nonzero(cols)[0][:40].tolist()}")
        r0 = rows[0]
        say(f"  row {r0}: device {x[r0, :48].tolist()}")
        say(f"  row {r0}: oracle {xr[r0, :48].tolist()}")
    assert dmax <= 1 and frac <= 1e-3
    assert np.abs(euler - er).max() <= 1e-12 * max(np.abs(er).max(), 1e-300)
    pos0, pol0, _, _ = host_state(system)
    lp.collect(nsnap=1, stride=0)
    system.download()
    n = len(pos0)
    assert net.dataset_size() == n
    xd, td = net.dataset_get(0, n)
    assert np.array_equal(xd, x)
    g = sp.geometry_cfg
    tr = lo.targets(pos0, system.state.pos.copy(), pol0, euler, system.state.nmax, g.length, g.height, 1e-3)
    scale = np.abs(tr).max()
    say(f"  targets: scale {scale:.3g}, max err {np.abs(td - tr).max():.2e}")
    assert np.abs(td - tr).max() <= 1e-6 * scale + 1e-6


def test_stale_chunks_second_evaluation(gpu_ctx, oracle_mod):
    """The hand-built crossing of the oracle's stale-chunk test on the GPU: B
    leaves A's chunk neighbourhood during the drift; the second evaluation
    keeps A-B (stale chunks) and the device agrees with the oracle and the
    closed form, with and without force reuse, over one and three steps."""
    import mavi_hip as mv
    from mavi_hip import _lib
    from test_oracle_integrator import _lj_pair, _stale_setup

    say("[stale chunks] three-particle crossing")
    o, pos, vel = _stale_setup(oracle_mod)
    dyn = mv.LenJonesCfg(sigma=1.0, epsilon=1.0)
    space = mv.SpaceCfg(wall_type=mv.RigidWalls(), geometry_cfg=mv.RectangleCfg(length=8.0, height=6.0))
    ic = mv.IntCfg(dt=0.01, chunks_cfg=mv.ChunksCfg(num_cols=4, num_rows=3), device=mv.GPU())
    for reuse in (1, 0):
        for nsteps in (1, 3):
            s = plain_system(pos.copy(), vel.copy(), space, dyn, ic)
            s.dev.set_option(_lib.MAVI_OPT_REUSE_FORCES, reuse)
            s.advance(nsteps)
            s.download()
            o2 = oracle_mod.from_system(3, space, dyn, ic)
            p_ref, v_ref, f_ref, _, _ = o2.newton_steps(pos, vel, nsteps)
            assert np.allclose(s.state.pos, p_ref, rtol=0, atol=1e-13)
            assert np.allclose(s.state.vel, v_ref, rtol=1e-12, atol=1e-15)
            assert np.allclose(s.forces, f_ref, rtol=1e-12, atol=1e-15)
            if nsteps == 1:
                dt = 0.01
                f0 = np.zeros((3, 2))
                f0[0] = _lj_pair(pos[0], pos[1])
                f0[1] = -f0[0]
                x1 = pos + vel * dt + f0 * (dt * dt / 2)
                f1 = np.zeros((3, 2))
                f1[0] = _lj_pair(x1[0], x1[1])
                f1[1] = -f1[0]
                assert np.allclose(s.state.vel, vel + dt / 2 * (f0 + f1), rtol=1e-12, atol=1e-15)
