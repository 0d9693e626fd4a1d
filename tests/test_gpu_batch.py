"""Replica batches (mavi_batch_*): every replica of a mixed batch must be
bit-identical to the same system stepped alone through mavi_step, whose own
parity with the oracle the other GPU tests establish.  The batch is the
device side of run_experiment_batch (src/experiments.jl:246-300)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def replica_cases():
    """(name, kind, builder) for a mixed batch of uneven sizes: LJ with and
    without cutoff, rigid and periodic walls, HarmTrunc, rings untyped and
    typed, particle counts that are and are not multiples of 64."""
    from mavi_hip import workloads as wl
    from test_gpu_parity import plain_case
    from test_gpu_rings import normal_case, types_case

    return [
        ("lj_gas_1024_ref", "plain", lambda: plain_case("lj_gas_1024_ref")),
        ("rings_normal", "rings", lambda: normal_case(True)),
        ("ht_periodic", "plain", lambda: plain_case("ht_periodic")),
        ("lj_periodic_400", "plain", lambda: wl.lj_periodic(nside=20, density=0.7, seed=5, cutoff=True)),
        ("rings_types", "rings", lambda: types_case(True)),
        ("ht_rigid", "plain", lambda: plain_case("ht_rigid")),
        ("rings_small", "rings", lambda: normal_case(True, n=12, cols=7, rows=6, seed=11)),
        ("lj_periodic_small_cut", "plain", lambda: plain_case("lj_periodic_small_cut")),
    ]


def single_run(kind, case, seed, steps_list):
    """The replica alone: System / RingsSystem, mavi_step per entry of
    steps_list; returns the state after each entry."""
    import mavi_hip as mv
    from mavi_hip import rings

    from mavi_hip import _lib

    out = []
    if kind == "plain":
        pos, vel, space, dyn, ic = case
        s = mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)
        # replicas run the chunk gather; a system alone would take neighbour
        # lists (same pairs, another summation order)
        s.dev.set_option(_lib.MAVI_OPT_NLIST, 0)
        for k in steps_list:
            s.advance(k)
            s.download()
            out.append((s.state.pos.copy(), s.state.vel.copy(), None, s.forces.copy()))
    else:
        state, space, cfg, ic = case
        s = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic, seed=seed)
        for k in steps_list:
            s.advance(k)
            s.download()
            out.append((s.state.pos.copy(), None, s.state.pol.copy(), s.forces.copy()))
    return out


def build_batch(cases, ctx):
    from mavi_hip import batch as mb
    from mavi_hip import rings

    b = mb.Batch(ctx=ctx)
    inits = []
    for (name, kind, _), case, seed in cases:
        if kind == "plain":
            pos, vel, space, dyn, ic = case
            from mavi_hip import SecondLawState

            st = SecondLawState(pos=pos, vel=vel)
            b.add(st, space, dyn, ic)
            inits.append((st.pos.copy(), st.vel.copy(), None))
        else:
            state, space, cfg, ic = case
            b.add(state, space, cfg, ic, seed=seed)
            inits.append((state.pos.copy(), None, state.pol.copy()))
    b.finalize()
    for i, (p, v, pol) in enumerate(inits):
        b.upload(i, p, v, pol)
    return b


def assert_same(tag, got, want):
    for what, g, w in zip(("pos", "vel", "pol", "forces"), got, want):
        if w is None:
            continue
        assert g is not None, (tag, what)
        if not np.array_equal(g, w):
            d = np.nanmax(np.abs(np.asarray(g) - np.asarray(w)))
            raise AssertionError(f"{tag}: {what} differs from the single system (max |diff| {d:.3e})")


def test_batch_bitexact_vs_single(gpu_ctx):
    """Mixed batch of 8 replicas, graph of 4 steps: 7 steps (graph + 3 direct)
    then 10 more (one direct step to return to the even copy, 2 graphs, 1
    direct); every replica equals its single-system run bit for bit."""
    from mavi_hip import batch as mb

    specs = replica_cases()
    steps_list = [7, 10]
    singles = []
    for i, (name, kind, mk) in enumerate(specs):
        say(f"[batch] single run of replica {i} {name}")
        singles.append(single_run(kind, mk(), 100 + i, steps_list))
    cases = [(spec, spec[2](), 100 + i) for i, spec in enumerate(specs)]
    b = build_batch(cases, gpu_ctx)
    info = b.info()
    say(f"[batch] info {info}")
    assert info["replicas"] == len(specs)
    assert info["padded_slots"] % 64 == 0 and info["padded_slots"] >= info["particles"]
    b.set_option(mb.MAVI_BOPT_GRAPH_STEPS, 4)
    for j, k in enumerate(steps_list):
        say(f"[batch] run {k} steps")
        b.run(k)
        for i, (name, kind, _) in enumerate(specs):
            pos, vel, pol = b.download(i)
            f = b.get_forces(i)
            assert_same(f"{name} after {sum(steps_list[:j + 1])} steps", (pos, vel, pol, f), singles[i][j])


@pytest.mark.parametrize("wg,listed,graphs", [(64, 0, 0), (512, 1, 1), (128, 1, 0)])
def test_batch_shapes_bitexact(gpu_ctx, wg, listed, graphs):
    """Workgroup size, the two-phase candidate list and graph replay change
    nothing in the results."""
    from mavi_hip import batch as mb

    specs = [s for s in replica_cases() if s[0] in ("lj_periodic_400", "rings_types", "ht_periodic")]
    say(f"[batch shapes] wg={wg} listed={listed} graphs={graphs}")
    want = [single_run(kind, mk(), 7 + i, [6])[0] for i, (_, kind, mk) in enumerate(specs)]
    b = build_batch([(spec, spec[2](), 7 + i) for i, spec in enumerate(specs)], gpu_ctx)
    b.set_option(mb.MAVI_BOPT_WG_PLAIN, wg)
    b.set_option(mb.MAVI_BOPT_WG_RINGS, wg)
    b.set_option(mb.MAVI_BOPT_LIST_PLAIN, listed)
    b.set_option(mb.MAVI_BOPT_LIST_RINGS, listed)
    b.set_option(mb.MAVI_BOPT_GRAPH_STEPS, 2)
    b.set_option(mb.MAVI_BOPT_GRAPHS, graphs)
    b.run(6)
    for i, (name, _, _) in enumerate(specs):
        pos, vel, pol = b.download(i)
        assert_same(name, (pos, vel, pol, b.get_forces(i)), want[i])


def test_batch_autotune_restores_state(gpu_ctx):
    specs = [s for s in replica_cases() if s[0] in ("lj_gas_1024_ref", "rings_normal")]
    say("[batch autotune]")
    want = [single_run(kind, mk(), 3 + i, [5])[0] for i, (_, kind, mk) in enumerate(specs)]
    b = build_batch([(spec, spec[2](), 3 + i) for i, spec in enumerate(specs)], gpu_ctx)
    best, shape = b.autotune(trial_steps=2)
    say(f"[batch autotune] {best:.1f} us/step, {shape}")
    assert best > 0 and shape["wg_plain"] in (64, 128, 256, 512) and shape["wg_rings"] in (64, 128, 256, 512)
    # best_us is a whole step of the mixed batch with both chosen shapes: at
    # least the pair and ring kernels of both kinds that one step launches
    from mavi_hip import _lib

    b.prof_enable(True)
    b.run(4)
    kern_ms = sum(b.prof_get(w)[0] for w in (_lib.MAVI_PROF_BIN, _lib.MAVI_PROF_FORCE, _lib.MAVI_PROF_RINGS))
    b.prof_enable(False)
    say(f"[batch autotune] profiled kernels {1e3 * kern_ms / 4:.1f} us/step")
    assert best >= 0.5 * 1e3 * kern_ms / 4
    b = build_batch([(spec, spec[2](), 3 + i) for i, spec in enumerate(specs)], gpu_ctx)
    b.autotune(trial_steps=2)
    b.run(5)
    for i, (name, _, _) in enumerate(specs):
        pos, vel, pol = b.download(i)
        assert_same(name, (pos, vel, pol, b.get_forces(i)), want[i])


def test_batch_errors(gpu_ctx):
    """A chunk overflow in one replica is reported with its index; unsupported
    replicas and runs without state are refused before any launch."""
    import mavi_hip as mv
    from mavi_hip import _lib
    from mavi_hip import batch as mb
    from test_gpu_rings import normal_case

    say("[batch errors]")
    good = [s for s in replica_cases() if s[0] == "lj_periodic_400"][0]
    pos, vel, space, dyn, ic = good[2]()
    b = mb.Batch(ctx=gpu_ctx)
    b.add(mv.SecondLawState(pos=pos, vel=vel), space, dyn, ic)
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
    b.add_desc(d)
    with pytest.raises(mv.MaviError):
        b.run(1)  # not finalized
    b.finalize()
    with pytest.raises(mv.MaviError):
        b.run(1)  # no state
    b.upload(0, pos, vel)
    b.upload(1, np.array([[1.0, 1.0], [1.1, 1.0], [1.2, 1.0]]), np.zeros((3, 2)))
    with pytest.raises(mv.CellOverflowError, match="replica 1"):
        b.run(1)
    # unsupported replicas
    b2 = mb.Batch(ctx=gpu_ctx)
    state, space, cfg, ic = normal_case(False)
    with pytest.raises(mv.UnsupportedOnGPU):
        b2.add(state, space, cfg, ic)
    state, space, cfg, ic = normal_case(True, n=40, cols=5, rows=4)
    with pytest.raises(mv.UnsupportedOnGPU):
        b2.add(state, space, cfg, ic)
