"""Rings with a variable population on the GPU (VarRingsIds, sources and
sinks; src/rings/states.jl:20-39, 173-223, src/rings/sources.jl,
src/rings/integration.jl:353-372) against the oracle through the C ABI.
Bars: the population (mask, ids, uids, active count, polarisation draws
used) equal to the oracle's after every block of steps; positions and
polarisations of the active rings within 1e-9 of the box size; with every
ring active and no source, bit-identical to the fixed population."""
import sys

import numpy as np
import pytest

from test_oracle_sources import make_oracle, source_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build_gpu(oracle_mod, use_chunks, pol, seed=7):
    from mavi_hip import rings

    state, space, cfg, ic, sources, ring_d = source_case(oracle_mod, use_chunks=use_chunks, pol=pol)
    pos0, pol0 = state.pos.copy(), state.pol.copy()
    src, sink = sources
    sys_ = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic,
                             source_cfg=[src.cfg, sink], rng=np.random.default_rng(seed))
    o = make_oracle(oracle_mod, state, space, cfg, ic, sources,
                    pol_draws=np.random.default_rng(seed).random(4096) if pol == "random" else None)
    return sys_, o, pos0, pol0, space


@pytest.mark.parametrize("use_chunks", [True, False])
@pytest.mark.parametrize("pol", [0.0, "random"])
def test_population_follows_the_oracle(gpu_ctx, oracle_mod, use_chunks, pol):
    say(f"[sources] chunks={use_chunks} pol={pol}")
    sys_, o, pos, pol_, space = build_gpu(oracle_mod, use_chunks, pol)
    L = space.geometry_cfg.length
    nm = sys_.state.nmax
    events = 0
    na_prev = 0
    for block in range(14):
        sys_.advance(100)
        sys_.download()
        pos, pol_, *_ = o.rings_steps(pos, pol_, 100)
        mask, ids, uids, na, cms, used = o.var_state()
        dmask, dids, duids, dna, dused, _ = sys_.dev.rings_var_state()
        assert np.array_equal(dmask, mask), f"block {block}"
        assert np.array_equal(dids, ids) and np.array_equal(duids, uids)
        assert dna == na and dused == used
        act = np.nonzero(mask)[0]
        sl = np.concatenate([np.arange(r * nm, (r + 1) * nm) for r in act] + [np.zeros(0, int)])
        d = sys_.state.pos[sl] - pos[sl]
        d -= L * np.round(d / L)
        assert np.abs(d).max(initial=0.0) <= 1e-9 * L
        assert np.abs(sys_.state.pol[act] - pol_[act]).max(initial=0.0) <= 1e-9
        assert np.array_equal(sys_.state.rings_ids.mask, mask != 0)
        events += abs(na - na_prev)
        na_prev = na
    say(f"  final active {na}, max uid {uids.max()}, draws used {used}")
    assert uids.max() > len(mask) + 6  # spawns happened
    assert events > 6


def test_all_active_equals_fixed_population(gpu_ctx):
    """ActiveState(true) and no source: the variable path steps exactly as
    the fixed one (the same kernels on the same state)."""
    from mavi_hip import rings
    from test_gpu_rings import normal_case

    out = []
    for var in (False, True):
        state, space, cfg, ic = normal_case(True, n=10, cols=8, rows=8)
        if var:
            state = rings.RingsState(rings_pos=state.rings_pos.copy(), pol=state.pol.copy(),
                                     active_state=rings.ActiveState(True))
        sys_ = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic, seed=5)
        sys_.advance(60)
        sys_.download()
        out.append((state.pos.copy(), state.pol.copy(), sys_.forces.copy(), sys_.cms.copy()))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_errors(gpu_ctx, oracle_mod):
    import mavi_hip as mv
    from mavi_hip import rings

    state, space, cfg, ic, sources, ring_d = source_case(oracle_mod, pol="random")
    src, sink = sources
    sys_ = rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic, source_cfg=[src.cfg, sink])
    # without a pool the spawns draw from the built-in stream
    sys_.advance(1)
    assert sys_.dev.rings_var_state()[5] == 4 and sys_.dev.rings_var_state()[3] == 4
    big = rings.SourceCfg(bottom_left=src.cfg.bottom_left, spawn_pos=src.cfg.spawn_pos, size=(9, 8), spawn_pol=0.0)
    state2, space2, cfg2, ic2, _, _ = source_case(oracle_mod)
    with pytest.raises(mv.UnsupportedOnGPU):
        rings.RingsSystem(state=state2, space_cfg=space2, dynamic_cfg=cfg2, int_cfg=ic2, source_cfg=[big])
