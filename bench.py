#!/usr/bin/env python3
"""Benchmark of the step hot path on one MI355X.

Workload (value): 1,048,576 LJ particles, periodic square box at number
density 0.7, jittered lattice (seed 1), Gaussian velocities T = 1, cutoff
2.5 sigma, 489x489 chunks, dt = 1e-3, FP64, newton_step! semantics (binning,
two force evaluations with stale chunks, velocity Verlet, periodic walls).
The state is resident in HBM when the timed region starts.

The system runs on Verlet neighbour lists (the library's default for a
finite-range plain system): one kernel per step, k_nl_second (the second
force evaluation, the kick and walls, the next step's first evaluation and
drift), plus a list rebuild every ~30 steps and a position write-back per
block of 30 steps.

Prints ONE JSON line.  `roofline` is for k_nl_second: achieved =
algorithmic bytes per launch / average launch duration measured with HIP
events on the library's stream.  Algorithmic bytes per particle: the state
read and written (drifted position 16 + 16, velocity 16 + 16, forces 16 +
16), the packed list (2 B per partner, mean rho pi (rc (1 + skin))^2
partners) and its count (4 B).  `traffic` is the measured HBM bytes per
launch from the committed rocprofv3 counter summary.  `cpu_baseline` is the
CPU oracle (a restatement of the reference's loops, OpenMP over chunk
columns like the reference's Threaded mode) timed on this host on a bounded
sample of the same workload.  `forms.lj_gather` times the same workload on
the chunk gather (lists off) and `forms.lj_reference` the reference's
LenJonesCfg as it is (no cutoff: chunk-membership pair set, lists do not
apply).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "mavi.jl_amd", ROOT / "oracle"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Algorithmic work per particle-step (FP64, SoA, one force evaluation):
#   bytes: state x,y,vx,vy read+write 64, F_old read + F_new write 32,
#          binning 40 (read x,y 16, key 4, permutation 4, reordered x,y 16),
#          force-phase read of reordered x,y 16                 -> 152 B
#   flops: 6.87 unique pairs within rc (0.5 rho pi rc^2) x 21 + 14 -> 158
B_STEP = 152.0
FLOP_STEP = 158.0
NL_SKIN = 0.15  # the library's default list skin (fraction of the range)


def nl_bytes_per_particle(density: float, rc: float) -> float:
    """Algorithmic bytes per particle of one k_nl_second launch."""
    import math

    partners = density * math.pi * (rc * (1.0 + NL_SKIN)) ** 2
    return 96.0 + 2.0 * partners + 4.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(name: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(pos, vel, space, dyn, ic, steps: int, threads: int):
    import oracle

    o = oracle.from_system(len(pos), space, dyn, ic)
    t0 = time.perf_counter()
    o.newton_steps(pos, vel, steps, nthreads=threads)
    dt = time.perf_counter() - t0
    return len(pos) * steps / dt, dt


def bench_rings(args):
    """Active-ring mixture (40,000 rings x 20 particles): Rings step! with
    device rotational noise, timed like the plain path."""
    import mavi_hip as mv
    from mavi_hip import rings, workloads as wl

    st, sp, cfg, ic = wl.rings_mixture(nrings=args.rings)
    pos0, pol0 = st.pos.copy(), st.pol.copy()
    n = int(len(st.pos))
    log(f"[bench] rings: {args.rings} rings x 20 = {n} particles, chunks {ic.chunks_cfg.num_cols}^2")
    system = rings.RingsSystem(state=st, space_cfg=sp, dynamic_cfg=cfg, int_cfg=ic, seed=2)
    dev = system.dev
    dev.step(args.warmup)
    t0 = time.perf_counter()
    dev.step(args.rings_steps)
    wall = time.perf_counter() - t0
    dev.prof_enable(True)
    dev.step(20)
    prof = {name: dev.prof_get(k) for k, name in ((0, "bin"), (1, "pair"), (2, "ring"))}
    dev.prof_enable(False)
    out = {"workload": f"active rings binary mixture, {args.rings} rings x 20 particles, periodic, ring packing "
                       "fraction 0.8, springs k=20 l=1, HarmTrunc repulsion k=10 radius 1, overdamped v0=1 "
                       "D_r=0.1, dt 1e-3, device Philox noise",
           "particles": n, "steps": args.rings_steps, "value": n * args.rings_steps / wall,
           "unit": "particle-steps/s", "ms_per_step": wall / args.rings_steps * 1e3,
           "kernel_avg_us": {k: v[0] / max(v[1], 1) * 1e3 for k, v in prof.items()}}
    # check_invasions! at the last step's ring points, ring chunks of about
    # one ring diameter (r_chunks, capacity from the equilibrium ring radius)
    g = sp.geometry_cfg
    r_out = max(rings.equilibrium_ring_radius(cfg, 20, t) for t in (1, 2))
    rcols = int(g.length // (2 * r_out))
    grid = rings.ring_chunks_grid(st, sp, cfg, mv.ChunksCfg(rcols, rcols))
    dev.rings_invasions_cfg(False, 0, *grid)
    dev.rings_invasions_check()
    t0 = time.perf_counter()
    for _ in range(args.inv_checks):
        dev.rings_invasions_check()
    inv_ms = (time.perf_counter() - t0) / args.inv_checks * 1e3
    nrec = len(dev.rings_invasions())
    out["invasions"] = {"grid": list(grid), "checks": args.inv_checks, "check_ms": inv_ms, "records": nrec,
                        "note": "host wall clock per mavi_rings_invasions_check (ring boxes + ring grid + "
                                "crossing tests + one record-count read back)"}
    if not args.no_cpu:
        import oracle

        o = oracle.from_rings(st, sp, cfg, ic)
        _, cms_d, cpos_d = dev.rings_info()
        t0 = time.perf_counter()
        o.rings_invasions(cpos_d, cms_d, *grid)
        out["invasions"]["cpu_baseline_ms"] = (time.perf_counter() - t0) * 1e3
        out["invasions"]["cpu_baseline_kind"] = "port, one core, the same check"
        threads = host_cores()
        t0 = time.perf_counter()
        o.rings_steps(pos0, pol0, args.rings_cpu_steps, nthreads=threads)
        dt = time.perf_counter() - t0
        n1 = max(2, args.rings_cpu_steps // 8)
        t0 = time.perf_counter()
        o.rings_steps(pos0, pol0, n1)
        dt1 = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * args.rings_cpu_steps / dt, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{n} particles x {args.rings_cpu_steps} Rings step! on {threads} threads "
                                         f"(Threaded-mode pair forces, per-ring threads; {dt:.1f} s)",
                               "one_core": {"value": n * n1 / dt1, "sample": f"{n1} steps on one core ({dt1:.1f} s)"}}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def bench_rings_sources(args, fixed_ms):
    """The rings mixture with a variable population, a source of 8 x 8 spawn
    boxes and a sink: every step adds update_cms!, the sink and spawn-box
    tests, a read back of a few words and the host's source/sink logic; a
    step that changes the population also rebuilds the state."""
    import mavi_hip as mv
    from mavi_hip import rings, workloads as wl

    st, sp, cfg, ic, sources = wl.rings_sources(nrings=args.rings)
    pos0, pol0 = st.pos.copy(), st.pol.copy()
    mask0 = st.rings_ids.mask.copy()
    system = rings.RingsSystem(state=st, space_cfg=sp, dynamic_cfg=cfg, int_cfg=ic, seed=2, source_cfg=sources,
                               rng=np.random.default_rng(11))
    dev = system.dev
    na0 = dev.rings_var_state()[3]
    log(f"[bench] rings with sources: {args.rings} ring slots, {na0} active")
    system.advance(args.warmup)
    uid0 = dev.rings_var_state()[2].max()
    nrec, t0 = [], time.perf_counter()
    k = 0
    while k < args.sources_steps:
        blk = min(100, args.sources_steps - k)
        system.advance(blk)
        nrec.append(dev.rings_var_state()[3])
        k += blk
    wall = time.perf_counter() - t0
    _, _, uids, na, used, _ = dev.rings_var_state()
    nact = float(np.mean(nrec)) * 20
    out = {"workload": f"rings_mixture ({args.rings} ring slots x 20) with a variable population: {int(na0)} rings "
                       "active at the start; a source of 8 x 8 spawn boxes (random polarisation) and a rectangular "
                       "sink; device Philox rotational noise",
           "steps": args.sources_steps, "ms_per_step": wall / args.sources_steps * 1e3,
           "value": nact * args.sources_steps / wall, "unit": "active particle-steps/s",
           "active_rings_end": int(na), "spawned": int(uids.max() - uid0),
           "ratio_ms_vs_fixed_rings": (wall / args.sources_steps * 1e3) / fixed_ms if fixed_ms else None}
    if not args.no_cpu:
        import oracle

        o = oracle.from_rings(st, sp, cfg, ic)
        o.set_var_rings(mask0.astype(np.int32), oracle.source_dicts(system.sources, st.nmax),
                        pol_draws=np.random.default_rng(11).random(4096))
        o.rings_var_init(pos0)
        t0 = time.perf_counter()
        o.rings_steps(pos0, pol0, args.sources_cpu_steps)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": na0 * 20 * args.sources_cpu_steps / dt, "unit": "active particle-steps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{args.sources_cpu_steps} steps of the same system on one core ({dt:.1f} s)"}
    return out


def bench_experiments(args):
    """run_experiment_batch over LJ gases (nside^2 particles each, velocities
    scaled by the value) with a DelayedCol collector and A/B checkpoints,
    routed to one device batch; compared with the bare batch stepping the
    same replicas the same number of steps without the runner's files."""
    import shutil
    import tempfile

    import mavi_hip as mv
    from mavi_hip import batch as mb
    from mavi_hip import experiments as ex
    from mavi_hip import workloads as wl

    nside, nrep = args.exp_nside, args.exp_replicas
    pos, vel, space, dyn, ic = wl.lj_periodic(nside=nside)
    ic = mv.IntCfg(dt=ic.dt, chunks_cfg=ic.chunks_cfg, device=mv.GPU())
    init = mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)

    def get_system(init_system, value, idx):
        return mv.System(state=mv.SecondLawState(pos=pos, vel=vel * value), space_cfg=space, dynamic_cfg=dyn,
                         int_cfg=ic)

    steps = args.exp_steps
    tf = steps * ic.dt * (1 - 1e-9)
    root = Path(tempfile.mkdtemp(prefix="mavi_exp_"))
    try:
        batch = ex.ExperimentBatch(exp_cfg=ex.ExperimentCfg(tf=tf, root=str(root),
                                                            checkpoint_cfg=ex.CheckpointCfg(delta_time=tf / 4)),
                                   col_cfg=ex.DelayedCfg(delay_time=tf / 8), init_system=init,
                                   values=list(np.linspace(0.5, 1.5, nrep)))
        log(f"[bench] experiments: {nrep} x {nside * nside} particles, {steps} steps")
        t0 = time.perf_counter()
        res = ex.run_experiment_batch(batch, get_system, verbose=False)
        wall = time.perf_counter() - t0
        done = sum((root / "data" / str(i + 1) / ".done").is_file() for i in range(nrep))
    finally:
        shutil.rmtree(root, ignore_errors=True)
    b = mb.Batch()
    for i in range(nrep):
        b.add(mv.SecondLawState(pos=pos, vel=vel), space, dyn, ic)
    b.finalize()
    for i in range(nrep):
        b.upload(i, pos, vel * (0.5 + i / max(nrep - 1, 1)))
    b.run(steps)
    t0 = time.perf_counter()
    b.run(steps)
    bare = time.perf_counter() - t0
    n = nrep * nside * nside
    return {"workload": f"run_experiment_batch: {nrep} LJ gases of {nside * nside} particles (lj_periodic), "
                        f"{steps} steps each, DelayedCol every {steps // 8} steps, checkpoints every {steps // 4}, "
                        "final states saved",
            "value": n * steps / wall, "unit": "particle-steps/s (files included)", "wall_s": wall,
            "experiments_done": done, "errors": sum(r is not None for r in res),
            "bare_batch_value": n * steps / bare, "runner_overhead": wall / bare}


def bench_lj_variant(args, plain_value, kind):
    """The headline workload on the chunk gather (lists off: one binning and
    two pair-force launches per step) or as the reference's LenJonesCfg
    expresses it (no cutoff: every pair in adjacent chunks interacts)."""
    import mavi_hip as mv
    from mavi_hip import workloads as wl

    pos, vel, space, dyn, ic = wl.lj_periodic(nside=args.nside, cutoff=(kind == "gather"))
    n = len(pos)
    system = mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)
    dev = system.dev
    if kind == "gather":
        dev.set_option(mv._lib.MAVI_OPT_NLIST, 0)
    dev.step(args.warmup)
    t0 = time.perf_counter()
    dev.step(args.variant_steps)
    wall = time.perf_counter() - t0
    value = n * args.variant_steps / wall
    info = dev.nlist_info()
    wk = ("plain LJ as the headline (cutoff 2.5 sigma) on the chunk gather: binning every step, two pair-force "
          "launches" if kind == "gather" else
          "plain LJ, reference LenJonesCfg (no cutoff): the pair set is every pair in the same or adjacent "
          "chunks of the 489x489 grid, as calc_forces! visits them")
    return {"workload": wk, "particles": n, "steps": args.variant_steps, "value": value,
            "unit": "particle-steps/s", "ms_per_step": wall / args.variant_steps * 1e3,
            "ratio_vs_headline": value / plain_value, "neighbour_lists": info["active"]}


def bench_walls(args, plain_value):
    """The plain 1M LJ workload inside a walls list (ManyWalls: periodic box,
    16 slippery circles, 8 slippery lines, one potential circle): the cost
    of the walls on the step (the walls-carrying instantiations of the pair
    kernels, with force reuse)."""
    import mavi_hip as mv
    from mavi_hip import workloads as wl

    pos, vel, space, dyn, ic = wl.lj_obstacles(nside=args.nside)
    n = len(pos)
    log(f"[bench] walls: {n} particles, {len(space.pairs())} walls")
    system = mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)
    dev = system.dev
    dev.step(args.warmup)
    t0 = time.perf_counter()
    dev.step(args.walls_steps)
    wall = time.perf_counter() - t0
    value = n * args.walls_steps / wall
    return {"workload": f"plain LJ as the headline, {n} particles (those within 1.5 of an obstacle left out), "
                        "walls list: PeriodicWalls box, 16 SlipperyWalls circles (radius 8), SlipperyWalls "
                        "LinesCfg of 8 segments (length 40), PotentialWalls circle (HarmTrunc k_rep 20, range 1)",
            "particles": n, "steps": args.walls_steps, "value": value, "unit": "particle-steps/s",
            "ms_per_step": wall / args.walls_steps * 1e3, "ratio_vs_plain": value / plain_value}


def bench_active(args, kind):
    """Self-propelled particles, nside^2 = 1M: szabo_step! (examples/szabo.jl
    parameters, dt 0.01) or rtp_step! (examples/rtp.jl parameters, density
    0.8, dt 1e-3), device Philox noise.  One step = binning + one fused
    pair-force/update launch."""
    import mavi_hip as mv
    from mavi_hip import workloads as wl

    if kind == "szabo":
        pos, ang, sp, dyn, ic = wl.szabo_lattice(nside=args.active_nside)
        what = ("Szabo self-propelled particles, lattice spacing r_eq in a periodic box, vo 1, mobility 1, "
                "relax_time 1, k_rep 10, k_adh 0.75, r_eq 1, r_max 1.1, rot_diff 0.01, dt 0.01")
    else:
        pos, ang, sp, dyn, ic = wl.rtp_gas(nside=args.active_nside)
        what = ("run-and-tumble particles, density 0.8 in a periodic box, vo 1, LJ sigma 1 epsilon 1 cut at "
                "2^(1/6) sigma, tumble_rate 1, dt 1e-3")
    n = len(pos)
    log(f"[bench] {kind}: {n} particles, chunks {ic.chunks_cfg.num_cols}^2")
    system = mv.System(state=mv.SelfPropelledState(pos=pos, pol_angle=ang), space_cfg=sp, dynamic_cfg=dyn,
                       int_cfg=ic, seed=6)
    dev = system.dev
    dev.step(args.warmup)
    t0 = time.perf_counter()
    dev.step(args.active_steps)
    wall = time.perf_counter() - t0
    dev.prof_enable(True)
    dev.step(20)
    prof = {name: dev.prof_get(k) for k, name in ((0, "bin"), (1, "pair_update"))}
    dev.prof_enable(False)
    out = {"workload": f"{what}, {n} particles, {ic.chunks_cfg.num_cols}^2 chunks, device Philox noise",
           "particles": n, "steps": args.active_steps, "value": n * args.active_steps / wall,
           "unit": "particle-steps/s", "ms_per_step": wall / args.active_steps * 1e3,
           "kernel_avg_us": {k: v[0] / max(v[1], 1) * 1e3 for k, v in prof.items()}}
    if not args.no_cpu:
        import oracle

        o = oracle.from_system(n, sp, dyn, ic)
        threads = host_cores()
        t0 = time.perf_counter()
        o.sp_steps(pos, ang, args.active_cpu_steps, nthreads=threads)
        dt = time.perf_counter() - t0
        n1 = max(2, args.active_cpu_steps // 4)
        t0 = time.perf_counter()
        o.sp_steps(pos, ang, n1)
        dt1 = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * args.active_cpu_steps / dt, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{n} particles x {args.active_cpu_steps} steps on {threads} threads "
                                         f"(Threaded-mode pair forces), zero noise ({dt:.1f} s)",
                               "one_core": {"value": n * n1 / dt1, "sample": f"{n1} steps on one core ({dt1:.1f} s)"}}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def bench_batched(args, plain_value):
    """Replica batch: 512 replicas (log-uniform 1k-64k particles, half LJ
    gases, half active-ring mixtures), every phase one launch over all
    replicas, 100-step graphs.  Compared with the plain path stepping the
    same replicas one after another (mavi_step per replica, its own 2-step
    graphs), and with the 1M plain figure."""
    import mavi_hip as mv
    from mavi_hip import _lib, rings, systems
    from mavi_hip import batch as mb
    from mavi_hip import workloads as wl

    ens = wl.batch_ensemble(nrep=args.batch)
    ctx = _lib.Context.for_device(0)
    b = mb.Batch(ctx=ctx)
    inits = []
    for kind, case, seed in ens:
        if kind == "plain":
            pos, vel, space, dyn, ic = case
            b.add(mv.SecondLawState(pos=pos, vel=vel), space, dyn, ic)
            inits.append((pos, vel, None))
        else:
            st, space, cfg, ic = case
            b.add(st, space, cfg, ic, seed=seed)
            inits.append((st.pos, None, st.pol))
    b.finalize()
    b.set_option(mb.MAVI_BOPT_GRAPH_STEPS, args.batch_graph)
    for i, (p, v, pol) in enumerate(inits):
        b.upload(i, p, v, pol)
    info = b.info()
    n = info["particles"]
    log(f"[bench] batched: {info}")
    best_us, shape = b.autotune(trial_steps=4)
    log(f"[bench] batched autotune: {best_us:.1f} us/step with {shape}")
    b.run(args.batch_graph)  # warm-up + graph capture
    t0 = time.perf_counter()
    b.run(args.batch_steps)
    wall = time.perf_counter() - t0
    value = n * args.batch_steps / wall
    b.prof_enable(True)
    b.run(10)
    prof = {name: b.prof_get(k) for k, name in ((0, "bin"), (1, "pair"), (2, "ring"))}
    b.prof_enable(False)
    del b

    # the same replicas through the plain single-system path, one after another
    seq_steps = args.batch_seq_steps
    if seq_steps <= 0:
        return {"particles": n, "replicas": info["replicas"], "value": value, "unit": "particle-steps/s",
                "ms_per_step": wall / args.batch_steps * 1e3, "shape": shape,
                "kernel_avg_ms_per_step": {k: v[0] / 10 for k, v in prof.items()}}
    devs = []
    for (kind, case, seed), (p, v, pol) in zip(ens, inits):
        if kind == "plain":
            pos, vel, space, dyn, ic = case
            d, _ = systems.describe(mv.SecondLawState(pos=pos, vel=vel), space, dyn, ic)
            dev = _lib.DeviceSystem(ctx, d)
        else:
            st, space, cfg, ic = case
            d, _, tables = rings.describe(st, space, cfg, ic, seed)
            dev = _lib.DeviceSystem(ctx, d, keep=tables)
        dev.upload(p, v, pol)
        dev.step(4)  # warm-up + graph capture
        devs.append(dev)
    t0 = time.perf_counter()
    for dev in devs:
        dev.step(seq_steps)
    wall_seq = time.perf_counter() - t0
    seq_value = n * seq_steps / wall_seq
    del devs
    return {
        "workload": f"{info['replicas']} replicas, particle counts log-uniform in [1000, 64000] (seed 3), even "
                    "replicas LJ gases (density 0.7, cutoff 2.5, periodic), odd replicas active-ring mixtures "
                    "(20-particle rings, packing fraction 0.8, device noise); dt 1e-3",
        "particles": n, "replicas": info["replicas"], "chunks": info["chunks"], "steps": args.batch_steps,
        "value": value, "unit": "particle-steps/s", "ms_per_step": wall / args.batch_steps * 1e3,
        "graph_steps": args.batch_graph, "shape": shape,
        "kernel_avg_ms_per_step": {k: v[0] / 10 for k, v in prof.items()},
        "sequential_plain": {"value": seq_value, "unit": "particle-steps/s", "steps": seq_steps,
                             "how": "mavi_step(replica, steps) for each replica in turn"},
        "speedup_vs_sequential_plain": value / seq_value,
        "ratio_vs_plain_1M": value / plain_value if plain_value else None,
        "error": "none: every replica is bit-identical to its single-system run (tests/test_gpu_batch.py), so "
                 "g(r) and segregation index are unchanged",
    }


def bench_learned(args):
    """Learned propagator on the active-ring mixture (nrings x 20 particles,
    chunks >= the 1.25 feature radius).  Training: snapshots of plain GPU
    trajectories from seeds 100.. (equilibrated first), network trained on
    the GPU by the project's kernels.  Evaluation on a held-out seed: plain
    FP64 steps vs learned steps (k = 10) from the same equilibrated state;
    g(r) and the segregation index averaged over snapshots, with a second
    plain run (other noise seed) as the run-to-run reference."""
    import mavi_hip as mv
    from mavi_hip import learned, rings, observables as obs, workloads as wl

    k = 10
    ctx = mv._lib.Context.for_device(0)
    net = learned.MLP(ctx=ctx, seed=7)
    t_collect = 0.0
    for q in range(args.learned_seeds):
        seed = 100 + q
        st, sp, cfg, ic = wl.rings_mixture(nrings=args.learned_rings, seed=seed, chunk=1.25)
        s = rings.RingsSystem(state=st, space_cfg=sp, dynamic_cfg=cfg, int_cfg=ic, seed=seed, ctx=ctx)
        s.dev.step(args.learned_equil)
        lp = learned.Learned(s.dev, net, k=k, r_feat=1.25, target_scale=1e-3, seed=seed)
        t0 = time.perf_counter()
        lp.collect(nsnap=args.learned_snaps, stride=20, rows_per_snap=args.learned_rows, seed=seed)
        t_collect += time.perf_counter() - t0
        del lp, s
    rows = net.dataset_size()
    log(f"[bench] learned: {rows} training rows from {args.learned_seeds} seeds")
    t0 = time.perf_counter()
    losses = net.train(steps=args.learned_train, batch=8192, lr=1e-3, seed=3)
    t_train = time.perf_counter() - t0

    # held-out seed
    seed = 199
    st, sp, cfg, ic = wl.rings_mixture(nrings=args.learned_rings, seed=seed, chunk=1.25)
    ref = rings.RingsSystem(state=st, space_cfg=sp, dynamic_cfg=cfg, int_cfg=ic, seed=seed, ctx=ctx)
    ref.dev.step(args.learned_equil)
    ref.download()
    pos0, pol0 = ref.state.pos.copy(), ref.state.pol.copy()
    g = sp.geometry_cfg
    L, H = g.length, g.height
    types = np.asarray(st.types)
    n = len(pos0)
    nmac = args.learned_eval // k
    nsnap = 4

    def run(kind, noise_seed):
        st2, sp2, cfg2, ic2 = wl.rings_mixture(nrings=args.learned_rings, seed=seed, chunk=1.25)
        s = rings.RingsSystem(state=st2, space_cfg=sp2, dynamic_cfg=cfg2, int_cfg=ic2, seed=noise_seed, ctx=ctx)
        s.dev.upload(pos0, None, pol0)
        lp = learned.Learned(s.dev, net, k=k, r_feat=1.25, target_scale=1e-3, seed=noise_seed) if kind else None
        if lp:
            lp.run(2)  # graph capture
            s.dev.upload(pos0, None, pol0)
        else:
            s.dev.step(2 * k)
            s.dev.upload(pos0, None, pol0)
        gs, segs, wall, acc = [], [], 0.0, 0
        per = nmac // nsnap
        for _ in range(nsnap):
            t0 = time.perf_counter()
            if lp:
                acc += lp.run(per)
            else:
                s.dev.step(per * k)
            wall += time.perf_counter() - t0
            pos = np.zeros((n, 2))
            s.dev.download(pos, None, None)
            r, gr = obs.radial_distribution(pos, L, H, r_max=4.0, dr=0.05)
            gs.append(gr)
            segs.append(obs.segregation_index(pos, types, 20, L, H))
        if lp:
            stats.update(lp.stats())
        return r, np.mean(gs, axis=0), float(np.mean(segs)), n * per * nsnap * k / wall, acc

    stats = {}
    r, g_plain, seg_plain, v_plain, _ = run(False, 1001)
    _, g_plain2, seg_plain2, _, _ = run(False, 1002)
    _, g_learn, seg_learn, v_learn, acc = run(True, 1003)
    gate_stats = dict(stats)

    # the gate where it fires: the trained network with its output layer
    # (the last 2 x 128 weights and 2 biases) scaled up, from the same state
    fires = {}
    p0 = net.params
    for scale in (30.0, 300.0):
        p = p0.copy()
        p[-(2 * 128 + 2):] *= scale
        net.params = p
        st3, sp3, cfg3, ic3 = wl.rings_mixture(nrings=args.learned_rings, seed=seed, chunk=1.25)
        s3 = rings.RingsSystem(state=st3, space_cfg=sp3, dynamic_cfg=cfg3, int_cfg=ic3, seed=1004, ctx=ctx)
        s3.dev.upload(pos0, None, pol0)
        lp3 = learned.Learned(s3.dev, net, k=k, r_feat=1.25, target_scale=1e-3, seed=1004)
        nl3 = 40
        t0 = time.perf_counter()
        a3 = lp3.run(nl3)
        w3 = time.perf_counter() - t0
        fires[f"output_x{scale:g}"] = {"learned_steps": nl3, "accepted": a3, "gate": lp3.stats(),
                                       "value": n * nl3 * k / w3, "unit": "particle-steps/s (incl. graph capture)"}
        del lp3, s3
    net.params = p0
    us_mlp = net.time_forward(n, 20)
    flop = 2.0 * n * (160 * 128 + 128 * 128 * 2 + 128 * 2)
    return {
        "gate": gate_stats,
        "gate_fires": fires,
        "workload": f"active rings binary mixture, {args.learned_rings} rings x 20 = {n} particles (ring physics of "
                    "the 800k mixture, chunks >= 1.25), k = 10 plain steps per learned step, held-out seed 199",
        "particles": n, "k": k, "value": v_learn, "unit": "particle-steps/s",
        "plain_value": v_plain, "speedup_vs_plain": v_learn / v_plain,
        "accepted": acc, "learned_steps": nmac,
        "error": {"g_r_rel_l2": obs.relative_l2(g_learn, g_plain),
                  "g_r_rel_l2_plain_vs_plain": obs.relative_l2(g_plain2, g_plain),
                  "segregation_index": {"plain": seg_plain, "plain_other_noise": seg_plain2, "learned": seg_learn,
                                        "abs_error": abs(seg_learn - seg_plain)},
                  "how": f"{nsnap} snapshots over {args.learned_eval} steps from one equilibrated state; g(r) to r = 4 "
                         "in bins of 0.05"},
        "training": {"seeds": f"100..{99 + args.learned_seeds}", "snapshots": args.learned_seeds * args.learned_snaps,
                     "rows": rows, "rows_per_snapshot": args.learned_rows, "steps": args.learned_train,
                     "batch": 8192, "loss_first": float(losses[:20].mean()), "loss_last": float(losses[-50:].mean()),
                     "collect_s": t_collect, "train_s": t_train},
        "mlp_kernel": {"rows": n, "avg_us": us_mlp, "bf16_TFLOPs": flop / (us_mlp * 1e-6) / 1e12,
                       "frac_of_bf16_peak": flop / (us_mlp * 1e-6) / 2.5e15},
    }


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=48, help="steps of the CPU sample (all host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prof-steps", type=int, default=100)
    ap.add_argument("--variant-steps", type=int, default=300,
                    help="steps of the lj_gather and lj_reference forms (0 = skip)")
    ap.add_argument("--rings", type=int, default=40000, help="rings in the rings form (0 = skip)")
    ap.add_argument("--rings-steps", type=int, default=500)
    ap.add_argument("--rings-cpu-steps", type=int, default=40)
    ap.add_argument("--inv-checks", type=int, default=20, help="timed invasion checks in the rings form")
    ap.add_argument("--sources-steps", type=int, default=500, help="steps of the rings-with-sources form (0 = skip)")
    ap.add_argument("--sources-cpu-steps", type=int, default=10)
    ap.add_argument("--exp-replicas", type=int, default=32, help="experiments in the runner form (0 = skip)")
    ap.add_argument("--exp-nside", type=int, default=100)
    ap.add_argument("--exp-steps", type=int, default=400)
    ap.add_argument("--walls-steps", type=int, default=500, help="timed steps of the walls form (0 = skip)")
    ap.add_argument("--active-nside", type=int, default=1024,
                    help="side of the Szabo and run-and-tumble lattices (0 = skip those forms)")
    ap.add_argument("--active-steps", type=int, default=500)
    ap.add_argument("--active-cpu-steps", type=int, default=20)
    ap.add_argument("--batch", type=int, default=512, help="replicas in the batched form (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--batch-graph", type=int, default=100, help="steps per captured graph of the batch")
    ap.add_argument("--batch-seq-steps", type=int, default=100,
                    help="steps per replica of the sequential run (0 = skip it)")
    ap.add_argument("--learned-rings", type=int, default=10000, help="rings of the learned form (0 = skip)")
    ap.add_argument("--learned-seeds", type=int, default=20)
    ap.add_argument("--learned-snaps", type=int, default=30, help="snapshots per training seed")
    ap.add_argument("--learned-rows", type=int, default=40000, help="sampled particles per snapshot")
    ap.add_argument("--learned-equil", type=int, default=1000)
    ap.add_argument("--learned-train", type=int, default=4000)
    ap.add_argument("--learned-eval", type=int, default=10000, help="plain-step equivalents of the evaluation")
    ap.add_argument("--no-graphs", action="store_true",
                    help="direct kernel launches instead of graph replay (for kernel traces)")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="MiB copied device-to-device 4 times before the run (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU decomposition is a design note, not code).", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy > 0:
        import torch

        nbytes = args.calib_copy << 20
        a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {nbytes} bytes read + {nbytes} bytes written, 4 launches")
        del a, b

    import mavi_hip as mv
    from mavi_hip import workloads as wl

    if args.no_graphs:
        mv._lib.USE_GRAPHS = False
    pos, vel, space, dyn, ic = wl.lj_periodic(nside=args.nside)
    n = len(pos)
    log(f"[bench] building system: {n} particles, chunks {ic.chunks_cfg.num_cols}x{ic.chunks_cfg.num_rows}")
    system = mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)
    dev = system.dev
    log(f"[bench] warmup {args.warmup} steps")
    dev.step(args.warmup)
    log(f"[bench] timed {args.steps} steps")
    t0 = time.perf_counter()
    dev.step(args.steps)  # returns after the library's stream is synchronised
    t1 = time.perf_counter()
    wall = t1 - t0
    value = n * args.steps / wall
    ms_per_step = wall / args.steps * 1e3

    nl0 = dev.nlist_info()
    log(f"[bench] event-timed pass, {args.prof_steps} steps")
    dev.prof_enable(True)
    dev.step(args.prof_steps)
    f_ms, f_n = dev.prof_get(mv._lib.MAVI_PROF_FORCE)
    r_ms, r_n = dev.prof_get(mv._lib.MAVI_PROF_FORCE_REUSE)
    b_ms, b_n = dev.prof_get(mv._lib.MAVI_PROF_BIN)
    dev.prof_enable(False)
    nl1 = dev.nlist_info()
    f_avg_s = f_ms / max(f_n, 1) / 1e3
    r_avg_s = r_ms / max(r_n, 1) / 1e3
    b_avg_s = b_ms / max(b_n, 1) / 1e3
    b_launch = nl_bytes_per_particle(0.7, 2.5)
    achieved = b_launch * n / f_avg_s / 1e9
    traffic = read_traffic("nl_second_lj_1M_traffic.txt")
    step_gbs = B_STEP * n / (wall / args.steps) / 1e9
    fp64_tflops = FLOP_STEP * n / (wall / args.steps) / 1e12

    cpu = None
    if not args.no_cpu:
        threads = host_cores()
        log(f"[bench] CPU oracle sample: {args.cpu_steps} steps of the same system, {threads} threads")
        v, dt = cpu_baseline(pos, vel, space, dyn, ic, args.cpu_steps, threads)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{n} particles x {args.cpu_steps} newton_step! ({dt:.1f} s), OpenMP over chunk "
                         "columns like Threaded mode; restatement, not reference (no Julia on the host)"}

    out = {
        "metric": "particle-steps/sec on 1 GPU: plain FP64 (1M LJ, periodic, rho=0.7, rc=2.5)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain LJ, {n} particles, periodic square box, density 0.7, cutoff 2.5 sigma "
                               "(LenJonesCfg cutoff: an extension of this library; the reference's LenJonesCfg has "
                               "none, see forms.lj_reference), "
                               f"{ic.chunks_cfg.num_cols}x{ic.chunks_cfg.num_rows} chunks, dt 1e-3, "
                               "newton_step! (velocity Verlet, 2 force evaluations per step), Verlet neighbour lists",
                   "particles": n, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_nl_second (second evaluation, kick, walls, next first evaluation and drift)",
                     "avg_launch_us": f_avg_s * 1e6, "launches": f_n,
                     "bytes_per_particle_alg": b_launch, "bytes_per_launch_alg": b_launch * n},
        "step_rates": {"alg_GBs": step_gbs, "fp64_alg_TFLOPs": fp64_tflops,
                       "fp64_frac": fp64_tflops / FP64_PEAK_TFLOPS,
                       "rebuild_launch_avg_us": b_avg_s * 1e6, "rebuild_launches": b_n,
                       "start_eval_avg_us": r_avg_s * 1e6, "start_eval_launches": r_n,
                       "list_rebuilds_per_1000_steps": 1000.0 * (nl1["rebuilds"] - nl0["rebuilds"]) / args.prof_steps,
                       "list_rollbacks": nl1["rollbacks"], "neighbour_lists": nl1["active"]},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = value / cpu["value"]
    forms = {}
    del system, dev
    if args.variant_steps > 0:
        log("[bench] chunk-gather and reference (no cutoff) forms")
        forms["lj_gather"] = bench_lj_variant(args, value, "gather")
        forms["lj_reference"] = bench_lj_variant(args, value, "reference")
    if args.rings > 0:
        forms["rings"] = bench_rings(args)
        if args.sources_steps > 0:
            forms["rings_sources"] = bench_rings_sources(args, forms["rings"]["ms_per_step"])
    if args.exp_replicas > 0:
        forms["experiments"] = bench_experiments(args)
    if args.walls_steps > 0:
        forms["walls"] = bench_walls(args, value)
    if args.active_nside > 0:
        forms["szabo"] = bench_active(args, "szabo")
        forms["run_and_tumble"] = bench_active(args, "rtp")
    if args.batch > 0:
        log(f"[bench] batched form, {args.batch} replicas")
        forms["batched"] = bench_batched(args, value)
    if args.learned_rings > 0:
        log(f"[bench] learned form, {args.learned_rings} rings")
        forms["learned"] = bench_learned(args)
    if forms:
        out["forms"] = forms
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
