"""Host logic of the experiment runner and the config serialisation
(src/experiments.jl, src/serder.jl): configs survive the JSON round trip;
the event-driven loop (the device runs the steps between collections and
checkpoints in one call) leaves the same files as the reference's
per-step loop; checkpoints alternate between folders 1 and 2."""
import json
import math
from pathlib import Path

import numpy as np
import pytest


def roundtrip(obj):
    from mavi_hip import serder

    return serder.load_component(json.loads(json.dumps(serder.get_obj_save_data_json(obj))))


def same(a, b):
    if isinstance(a, np.ndarray):
        return np.array_equal(a, b)
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
    if isinstance(a, dict):
        return a.keys() == b.keys() and all(same(a[k], b[k]) for k in a)
    if hasattr(a, "__dict__"):
        return type(a) is type(b) and same(vars(a), vars(b))
    return a == b


def test_configs_roundtrip():
    import mavi_hip as mv
    from mavi_hip import rings

    space = mv.SpaceCfg.from_list([(mv.PeriodicWalls(), mv.RectangleCfg(length=40, height=30)),
                                   (mv.SlipperyWalls(), mv.CircleCfg(radius=3, center=(5, 6))),
                                   (mv.PotentialWalls(mv.HarmTruncCfg(k_rep=2, k_atr=1, dist_eq=1, dist_max=1.5),
                                                      "inside"), mv.CircleCfg(radius=4, center=(20, 10)))])
    i1 = rings.HarmTruncCfg(k_rep=13, k_atr=0.1, dist_eq=1, dist_max=1.1)
    i2 = rings.HarmTruncCfg(k_rep=13, k_atr=0.1, dist_eq=0.8, dist_max=0.88)
    rcfg = rings.RingsCfg(p0=[3.5, 3.6], relax_time=1, vo=1, mobility=1, rot_diff=0.5, k_area=1, k_spring=20,
                          l_spring=[0.8, 0.64], num_particles=[10, 7],
                          interaction_finder=rings.InteractionMatrix([[i1, i2], [i2, i1]]))
    ic = rings.RingsIntCfg(dt=0.01, p_chunks_cfg=mv.ChunksCfg(10, 8), r_chunks_cfg=mv.ChunksCfg(4, 3),
                           invasions_cfg=rings.InvasionsCfg(5))
    objs = [space, mv.LenJonesCfg(sigma=1, epsilon=1, cutoff=2.5), mv.SzaboCfg(vo=1, mobility=1, relax_time=1,
            k_rep=10, k_adh=0.75, r_eq=1, r_max=1.1, rot_diff=0.01), rcfg, ic, mv.TimeInfo(num_steps=7, time=0.07),
            rings.SourceCfg(bottom_left=(1.0, 2.0), spawn_pos=np.zeros((3, 2)), size=(2, 2), pad=0.5),
            rings.SinkCfg(mv.RectangleCfg(length=2, height=3, bottom_left=(4, 5)))]
    for o in objs:
        assert same(o, roundtrip(o)), type(o).__name__


class FakeDev:
    def __init__(self):
        self.seed, self.counter = 3, 0

    def noise_state(self):
        return self.seed, self.counter

    def set_noise_state(self, seed, counter):
        self.seed, self.counter = seed, counter


class FakeSystem:
    """A host stand-in with the System interface: free flight, x += v dt."""

    def __init__(self, n=16, dt=0.01):
        import mavi_hip as mv

        rng = np.random.default_rng(0)
        self.state = mv.SecondLawState(pos=rng.uniform(0, 10, (n, 2)), vel=rng.normal(size=(n, 2)))
        self.space_cfg = mv.SpaceCfg(wall_type=mv.PeriodicWalls(), geometry_cfg=mv.RectangleCfg(length=10, height=10))
        self.dynamic_cfg = mv.LenJonesCfg(sigma=1, epsilon=1)
        self.int_cfg = mv.IntCfg(dt=dt)
        self.time_info = mv.TimeInfo()
        self.dev = FakeDev()
        self.calls = 0
        self._dev_pos = self.state.pos.copy()

    def advance(self, k):
        self.calls += 1
        for _ in range(k):
            self._dev_pos = self._dev_pos + self.state.vel * self.int_cfg.dt
        self.dev.counter += k
        self.time_info.num_steps += k
        for _ in range(k):
            self.time_info.time += self.int_cfg.dt

    def download(self):
        self.state.pos[:] = self._dev_pos

    def upload(self):
        self._dev_pos = self.state.pos.copy()


def run(tmp, per_step):
    from mavi_hip import experiments as ex

    s = FakeSystem()
    cfg = ex.ExperimentCfg(tf=2.0, root=str(tmp), checkpoint_cfg=ex.CheckpointCfg(delta_time=0.5))
    col = ex.ManyColsCfg({"a": ex.DelayedCfg(delay_time=0.37), "b": ex.DelayedCfg(delay_time=0.8)})
    e = ex.Experiment(cfg, col, s)
    ex.run_experiment(e, (lambda sy: False) if per_step else None)
    return s


def files(root):
    out = {}
    for p in sorted(Path(root).rglob("*")):
        if p.is_file() and p.name != ".done":
            if p.suffix == ".npz":
                with np.load(p) as z:
                    out[str(p.relative_to(root))] = {k: z[k].tolist() for k in z.files}
            else:
                out[str(p.relative_to(root))] = p.read_text()
    return out


def test_event_loop_equals_per_step_loop(tmp_path):
    a = run(tmp_path / "a", per_step=False)
    b = run(tmp_path / "b", per_step=True)
    assert a.time_info.num_steps == b.time_info.num_steps == 200 and a.time_info.time == b.time_info.time
    assert b.calls == 200 and a.calls < 20  # the device ran the stretches between events in one call
    fa, fb = files(tmp_path / "a"), files(tmp_path / "b")
    assert fa.keys() == fb.keys()
    for k in fa:
        if k.endswith("experiment_configs.json") or k.endswith("configs.json"):
            continue  # they name their root
        assert fa[k] == fb[k], k
    info = json.loads((tmp_path / "a" / "checkpoint" / "checkpoint_info.json").read_text())
    assert info["is_valid"] == {"1": True, "2": True} and info["last_used"] in ("1", "2")
    # checkpoints every 0.5 of time: 4 in 2.0, the last in folder 2
    assert info["last_used"] == "2" and info["last_time"] == pytest.approx(2.0, abs=0.011)


def test_delayed_collector_keeps_the_older_snapshot(tmp_path):
    from mavi_hip import experiments as ex

    s = FakeSystem()
    col = ex.get_collector(ex.DelayedCfg(delay_time=0.25), None, s)
    times = []
    for k in range(100):
        s.advance(1)
        s.download()
        col.collect(s)
        times.append(col.state.last_save_time)
    # a snapshot each time more than 0.25 passed: at t > 0.25, > 0.51, ...
    assert len(col.state.sys_states) == 2
    t_saved = sorted(x[0].time for x in col.state.sys_states)
    assert t_saved[1] - t_saved[0] > 0.25 and t_saved[1] <= 1.0 + 1e-12
