"""Mirror of Mavi.Experiments (src/experiments.jl) for the GPU path.

    exp = Experiment(ExperimentCfg(tf=..., root=..., checkpoint_cfg=CheckpointCfg(...)), col_cfg, system)
    run_experiment(exp)
    batch = ExperimentBatch(exp_cfg=..., col_cfg=..., init_system=..., values=[...])
    run_experiment_batch(batch, get_system)

The loop of run_experiment is the reference's: step, collect, checkpoint,
stop test, until time_info.time >= tf.  The state stays on the device
between the steps at which something reads it: the step counts up to the
next collection, checkpoint or the end come from the same time arithmetic
(time += dt per step), the device runs them in one call, and the collectors
and checkpoints see the downloaded state exactly where the per-step loop
would.  A stop_func (any user test of the state) makes the loop download
after every step.

run_experiment_batch runs the experiments of a batch that share dt, time
and step count, are batchable kinds (LJ / HarmTrunc without a walls list,
fixed rings) and have no stop_func as one device batch (mavi_batch_*),
every replica bit-identical to its experiment run alone; the others run
one after the other.  Folders and files follow the reference: root/
experiment_configs.json, init_system/, data/<i>/ (configs.json,
experiment_configs.json, data/, checkpoint/{1,2}, checkpoint_info.json,
final_state/, .done).
"""
from __future__ import annotations

import copy
import json
import math
import time as _time
from dataclasses import dataclass, field
from datetime import datetime
from pathlib import Path

import numpy as np

from . import serder
from .serder import (get_obj_save_data_json, load_component, load_component_json, load_system, save_component_json,
                     save_state, save_system, save_system_configs)

COL_DIRNAME = "data"
FINAL_STATE_NAME = "final_state"
CHECKPOINT_DIRNAME = "checkpoint"
CHECKPOINT_INFO_NAME = "checkpoint_info.json"
EXP_COL_CONFIGS_NAME = "experiment_configs.json"


# --- configs ----------------------------------------------------------------
@dataclass
class CheckpointCfg:  # src/experiments.jl:33-36
    delta_time: float
    save_data: bool = False


@dataclass
class Checkpoint:  # src/experiments.jl:38-43
    last_used: str = "2"
    last_time: float = 0.0
    is_valid: dict = field(default_factory=lambda: {"1": False, "2": False})


@dataclass
class ExperimentCfg:  # src/experiments.jl:45-50
    tf: float
    root: str
    save_final_state: bool = True
    checkpoint_cfg: CheckpointCfg | None = None


# --- collectors (src/experiments.jl:23-31, 360-459) -------------------------
class ColCfg:
    pass


class Collector:
    """collect(system) runs after every step; due(t) says whether collect
    acts at time t (the default: always, which makes the runner download
    every step)."""

    def due(self, t: float) -> bool:
        return True

    def collect(self, system):
        pass

    def save_data(self, path):
        pass


def get_collector(col_cfg, exp_cfg, system, state=None) -> Collector:
    return col_cfg.get_collector(exp_cfg, system, state)


@dataclass
class DelayedCfg(ColCfg):  # src/experiments.jl:411-414
    delay_time: float

    def get_collector(self, exp_cfg, system, state=None):
        return DelayedCol(self, system, state)


@dataclass
class DelayedState:  # src/experiments.jl:416-419
    last_save_time: float
    sys_states: list  # at most 2 (time_info, state arrays) snapshots, oldest first


class DelayedCol(Collector):
    """Keeps the system as it was about delay_time ago (src/experiments.jl:
    421-452): a snapshot each time more than delay_time passed since the
    last one, the older of the two kept ones saved."""

    def __init__(self, cfg: DelayedCfg, system, state: DelayedState | None = None):
        self.cfg = cfg
        if state is None:
            state = DelayedState(system.time_info.time, [_snapshot(system)])
        self.state = state

    def due(self, t):
        return t - self.state.last_save_time > self.cfg.delay_time

    def collect(self, system):
        t = system.time_info.time
        if self.due(t):
            self.state.last_save_time = t
            self.state.sys_states.append(_snapshot(system))
            self.state.sys_states = self.state.sys_states[-2:]  # CircularBuffer(2)

    def save_data(self, path):
        times = [s[0].time for s in self.state.sys_states]
        snap = self.state.sys_states[int(np.argmin(times))]
        save_component_json(snap[0], path, "time_info")
        _save_snapshot_state(snap, Path(path) / "state")


@dataclass
class ManyColsCfg(ColCfg):  # src/experiments.jl:364-366
    cfgs: dict

    def get_collector(self, exp_cfg, system, state=None):
        return ManyCols(self, exp_cfg, system, state)


class ManyCols(Collector):
    def __init__(self, cfg: ManyColsCfg, exp_cfg, system, state=None):
        self.cfg = cfg
        self.cols = {name: get_collector(c, exp_cfg, system, None if state is None else state[name])
                     for name, c in cfg.cfgs.items()}
        self.state = {name: c.state for name, c in self.cols.items()}

    def due(self, t):
        return any(c.due(t) for c in self.cols.values())

    def collect(self, system):
        for c in self.cols.values():
            c.collect(system)

    def save_data(self, path):
        for name, c in self.cols.items():
            p = Path(path) / str(name)
            p.mkdir(parents=True, exist_ok=True)
            c.save_data(p)


def _snapshot(system):
    """(time_info, state arrays, device noise position, state kind) now:
    the deepcopy of (time_info, state, rng) the reference keeps."""
    st = system.state
    if hasattr(st, "rings_pos"):
        arrays = {"rings_pos": st.rings_pos.copy(), "pol": st.pol.copy()}
        if st.rings_ids is not None:
            arrays["active_state"] = st.rings_ids.mask.copy()
    elif hasattr(st, "pol_angle"):
        arrays = {"pos": st.pos.copy(), "pol_angle": st.pol_angle.copy()}
    else:
        arrays = {"pos": st.pos.copy(), "vel": st.vel.copy()}
    return (copy.deepcopy(system.time_info), arrays, system.dev.noise_state(), type(st).__name__)


def _save_snapshot_state(snap, path):
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    np.savez(path / "data.npz", **snap[1])
    with open(path / "type.json", "w") as f:
        json.dump({"type": snap[3]}, f)
    with open(path / "rng.json", "w") as f:
        json.dump({"device_seed": snap[2][0], "device_counter": snap[2][1]}, f)


serder.register(CheckpointCfg, ExperimentCfg, DelayedCfg, ManyColsCfg, DelayedState)


# --- experiment -------------------------------------------------------------
class Experiment:  # src/experiments.jl:52-73
    def __init__(self, cfg: ExperimentCfg, col_cfg, system, custom_step=None, *, col=None, checkpoint=None):
        self.cfg = cfg
        self.col_cfg = col_cfg
        self.system = system
        self.custom_step = custom_step
        self.col = col if col is not None else get_collector(col_cfg, cfg, system)
        if checkpoint is None and cfg.checkpoint_cfg is not None:
            checkpoint = Checkpoint(last_time=system.time_info.time)
        self.checkpoint = checkpoint


def save_experiment_configs(col_cfg, exp_cfg):
    Path(exp_cfg.root).mkdir(parents=True, exist_ok=True)
    with open(Path(exp_cfg.root) / EXP_COL_CONFIGS_NAME, "w") as f:
        json.dump({"experiment": get_obj_save_data_json(exp_cfg), "collector": get_obj_save_data_json(col_cfg)}, f,
                  indent=1)


def load_configs(path):
    with open(path) as f:
        d = json.load(f)
    return {k: load_component(v) for k, v in d.items()}


def _cp_due(cp_cfg, cp, t):
    return cp_cfg is not None and not (t - cp.last_time < cp_cfg.delta_time)


def check_checkpoint(experiment):
    """src/experiments.jl:458-488: alternate checkpoint folders 1 and 2."""
    cfg = experiment.cfg.checkpoint_cfg
    if cfg is None:
        return
    cp = experiment.checkpoint
    t = experiment.system.time_info.time
    if t - cp.last_time < cfg.delta_time:
        return
    cp.last_time = t
    name = "2" if cp.last_used == "1" else "1"
    cp_dir = Path(experiment.cfg.root) / CHECKPOINT_DIRNAME
    path = cp_dir / name
    cp.is_valid[name] = False
    _write_cp_info(cp_dir, cp)
    save_state(experiment.system, path / "sys_state")
    save_component_json(experiment.system.time_info, path / "time_info", "time_info")
    _save_col_state(experiment.col, path / "col_state")
    if cfg.save_data:
        experiment.col.save_data(Path(experiment.cfg.root) / COL_DIRNAME)
    cp.is_valid[name] = True
    cp.last_used = name
    _write_cp_info(cp_dir, cp)


def _write_cp_info(cp_dir, cp):
    cp_dir.mkdir(parents=True, exist_ok=True)
    with open(cp_dir / CHECKPOINT_INFO_NAME, "w") as f:
        json.dump({"last_used": cp.last_used, "last_time": cp.last_time, "is_valid": cp.is_valid}, f)


def _save_col_state(col, path):
    """Collector state: JSON for the small fields, npz for snapshot arrays."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    arrays = {}

    def enc(c, prefix):
        if isinstance(c, ManyCols):
            return {"type": "ManyCols", "data": {k: enc(v, f"{prefix}{k}.") for k, v in c.cols.items()}}
        if isinstance(c, DelayedCol):
            snaps = []
            for j, sn in enumerate(c.state.sys_states):
                names = {}
                for k, v in sn[1].items():
                    key = f"{prefix}{j}.{k}"
                    arrays[key] = v
                    names[k] = key
                snaps.append({"time_info": get_obj_save_data_json(sn[0]), "arrays": names, "noise": list(sn[2]),
                              "kind": sn[3]})
            return {"type": "DelayedCol", "data": {"last_save_time": c.state.last_save_time, "snaps": snaps}}
        return {"type": "custom", "data": get_obj_save_data_json(getattr(c, "state", None))}

    meta = enc(col, "")
    np.savez(path / "arrays.npz", **arrays)
    with open(path / "col_state.json", "w") as f:
        json.dump(meta, f)


def _load_col_state(path):
    path = Path(path)
    with open(path / "col_state.json") as f:
        meta = json.load(f)
    with np.load(path / "arrays.npz", allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}

    def dec(d):
        if d["type"] == "ManyCols":
            return {k: dec(v) for k, v in d["data"].items()}
        if d["type"] == "DelayedCol":
            snaps = [(load_component(sn["time_info"]), {k: arrays[v] for k, v in sn["arrays"].items()},
                      tuple(sn["noise"]), sn["kind"]) for sn in d["data"]["snaps"]]
            return DelayedState(d["data"]["last_save_time"], snaps)
        return load_component(d["data"])

    return dec(meta)


def _steps_to_event(system, experiment, stop_func):
    """Steps until the next one after which the loop reads the state: a
    collection, a checkpoint, the end, or every step with a stop_func."""
    if stop_func is not None or experiment.custom_step is not None:
        return 1
    t, dt, tf = system.time_info.time, system.int_cfg.dt, experiment.cfg.tf
    cp_cfg, cp, col = experiment.cfg.checkpoint_cfg, experiment.checkpoint, experiment.col
    k = 0
    while True:
        t += dt
        k += 1
        if t >= tf or col.due(t) or _cp_due(cp_cfg, cp, t) or k >= 1_000_000:
            return k


def run_experiment(experiment: Experiment, stop_func=None, prog_kwargs=None):
    """src/experiments.jl:94-144."""
    system, col, cfg = experiment.system, experiment.col, experiment.cfg
    save_system_configs(system, cfg.root)
    save_experiment_configs(experiment.col_cfg, cfg)
    col_path = Path(cfg.root) / COL_DIRNAME
    col_path.mkdir(parents=True, exist_ok=True)
    if experiment.checkpoint is not None:
        (Path(cfg.root) / CHECKPOINT_DIRNAME / "1").mkdir(parents=True, exist_ok=True)
        (Path(cfg.root) / CHECKPOINT_DIRNAME / "2").mkdir(parents=True, exist_ok=True)
    while system.time_info.time < cfg.tf:
        k = _steps_to_event(system, experiment, stop_func)
        if experiment.custom_step is not None:
            experiment.custom_step(system)
        else:
            system.advance(k)
        system.download()
        col.collect(system)
        check_checkpoint(experiment)
        if stop_func is not None and stop_func(system):
            break
    col.save_data(col_path)
    if cfg.save_final_state:
        save_system(system, Path(cfg.root) / FINAL_STATE_NAME)
    with open(Path(cfg.root) / ".done", "w") as f:
        f.write(f"Experiment concluded on {datetime.now().isoformat()}\n")


def load_experiment(root, custom_step=None) -> Experiment:
    """src/experiments.jl:490-520: the last valid of checkpoints 1 and 2."""
    root = Path(root)
    cp_path = root / CHECKPOINT_DIRNAME / CHECKPOINT_INFO_NAME
    if not cp_path.is_file():
        raise FileNotFoundError(f"Checkpoint info file does not exist at {cp_path}")
    info = json.load(open(cp_path))
    cp = Checkpoint(last_used=info["last_used"], last_time=info["last_time"], is_valid=info["is_valid"])
    name = cp.last_used
    if not cp.is_valid[name]:
        name = "2" if name == "1" else "1"
        if not cp.is_valid[name]:
            raise RuntimeError("Every checkpoint is corrupted!")
    path = root / CHECKPOINT_DIRNAME / name
    configs = load_configs(root / EXP_COL_CONFIGS_NAME)
    exp_cfg, col_cfg = configs["experiment"], configs["collector"]
    exp_cfg.root = str(root)
    system = load_system(root / "configs.json", path / "sys_state", path / "time_info")
    col_state = _load_col_state(path / "col_state")
    col = get_collector(col_cfg, exp_cfg, system, col_state)
    return Experiment(exp_cfg, col_cfg, system, custom_step, col=col, checkpoint=cp)


# --- batches ----------------------------------------------------------------
@dataclass
class ExperimentBatch:  # src/experiments.jl:146-159
    exp_cfg: ExperimentCfg
    col_cfg: object
    init_system: object
    values: list
    custom_step: object = None


def add_experiments(batch: ExperimentBatch, extra_values) -> ExperimentBatch:
    return ExperimentBatch(batch.exp_cfg, batch.col_cfg, batch.init_system, list(batch.values) + list(extra_values),
                           batch.custom_step)


def set_final_time(batch: ExperimentBatch, tf) -> ExperimentBatch:
    """src/experiments.jl:173-199."""
    batch = copy.copy(batch)
    batch.exp_cfg = copy.copy(batch.exp_cfg)
    batch.exp_cfg.tf = tf
    root = Path(batch.exp_cfg.root)
    for entry in sorted((root / "data").iterdir()) if (root / "data").is_dir() else ():
        done = entry / ".done"
        if done.is_file():
            done.unlink()
        cfg_path = entry / EXP_COL_CONFIGS_NAME
        if cfg_path.is_file():
            c = load_configs(cfg_path)
            e = c["experiment"]
            e.tf = tf
            save_experiment_configs(c["collector"], e)
    save_experiment_configs(batch.col_cfg, batch.exp_cfg)
    return batch


def _batchable(system) -> bool:
    from . import rings as _rings
    from .configs import HarmTruncCfg, LenJonesCfg

    if getattr(system, "sources", None) is not None:
        return False
    if isinstance(system, _rings.RingsSystem):
        return system.space_cfg.wall_type.__class__.__name__ == "PeriodicWalls"
    return (isinstance(system.dynamic_cfg, (LenJonesCfg, HarmTruncCfg)) and
            system.space_cfg.wall_type.__class__.__name__ in ("PeriodicWalls", "RigidWalls") and
            not getattr(system, "self_propelled", False))


def run_experiment_batch(batch: ExperimentBatch, get_system, stop_func=None, verbose=True, use_threads=True,
                         use_device_batch=True):
    """src/experiments.jl:246-300: every value's experiment in data/<i>/,
    skipping finished ones and resuming from checkpoints."""
    exp_cfg, col_cfg = batch.exp_cfg, batch.col_cfg
    root = Path(exp_cfg.root)
    root.mkdir(parents=True, exist_ok=True)
    save_experiment_configs(col_cfg, exp_cfg)
    with open(root / "values.json", "w") as f:
        json.dump(get_obj_save_data_json(list(batch.values)), f)
    save_system(batch.init_system, root / "init_system")
    t1 = _time.time()
    results: list = [None] * len(batch.values)
    pending = []
    for i, value in enumerate(batch.values):
        idx = i + 1
        exp_root = root / "data" / str(idx)
        try:
            if (exp_root / ".done").is_file():
                _log(verbose, idx, "Experiment already completed")
                continue
            try:
                experiment = load_experiment(exp_root, batch.custom_step)
                _log(verbose, idx, "Experiment Loaded!")
            except (FileNotFoundError, RuntimeError, KeyError, OSError):
                cfg_i = ExperimentCfg(tf=exp_cfg.tf, root=str(exp_root), save_final_state=exp_cfg.save_final_state,
                                      checkpoint_cfg=exp_cfg.checkpoint_cfg)
                experiment = Experiment(cfg_i, col_cfg, get_system(batch.init_system, value, idx),
                                        batch.custom_step)
            pending.append((idx, experiment))
        except Exception as e:  # noqa: BLE001 - reported like the reference
            results[i] = e
            _log(verbose, idx, f"Error setting up experiment: {e!r}")
    groups: dict = {}
    solo = []
    for idx, ex in pending:
        s = ex.system
        if use_device_batch and stop_func is None and ex.custom_step is None and _batchable(s):
            key = (s.int_cfg.dt, s.time_info.time, s.time_info.num_steps, type(s).__name__)
            groups.setdefault(key, []).append((idx, ex))
        else:
            solo.append((idx, ex))
    for key, members in groups.items():
        if len(members) == 1:
            solo.extend(members)
            continue
        try:
            _run_group(members)
        except Exception as e:  # noqa: BLE001
            for idx, _ in members:
                results[idx - 1] = e
                _log(verbose, idx, f"Error during experiment: {e!r}")
    for idx, ex in solo:
        try:
            run_experiment(ex, stop_func)
        except Exception as e:  # noqa: BLE001
            results[idx - 1] = e
            _log(verbose, idx, f"Error during experiment: {e!r}")
            save_system(ex.system, Path(ex.cfg.root) / "error_system")
    if verbose:
        failed = [i + 1 for i, r in enumerate(results) if isinstance(r, Exception)]
        if failed:
            print("\nFailed experiments: ", failed)
        print(f"\nTotal time: {_time.time() - t1:.1f} s")
    return results


def _log(verbose, idx, msg):
    if verbose:
        print(f"{idx}: {msg}")


class _BatchView:
    """A replica of a device batch seen through the System interface the
    collectors and checkpoints use (state, time_info, dev.noise_state)."""

    def __init__(self, system, batch, i):
        self._system, self._batch, self._i = system, batch, i

    def __getattr__(self, name):
        return getattr(self._system, name)

    def download(self):
        pos, vel, pol = self._batch.download(self._i)
        st = self._system.state
        st.pos[:] = pos
        if vel is not None:
            st.vel[:] = vel
        if pol is not None:
            st.pol[:] = pol


def _run_group(members):
    """Experiments that share dt, time and steps, in one device batch: the
    same event schedule for all (the exp_cfg and col_cfg are the batch's)."""
    from .batch import Batch

    b = Batch()
    ids = []
    for idx, ex in members:
        s = ex.system
        seed = s.dev.noise_state()[0]
        ids.append(b.add(s.state, s.space_cfg, s.dynamic_cfg, s.int_cfg, seed=seed))
    b.finalize()
    for (idx, ex), i in zip(members, ids):
        st = ex.system.state
        if hasattr(st, "rings_pos"):
            b.upload(i, st.pos, None, st.pol)
        else:
            b.upload(i, st.pos, st.vel)
    b.set_noise_counter(members[0][1].system.dev.noise_state()[1])
    lead = members[0][1]
    views = []
    for (idx, ex), i in zip(members, ids):
        v = _BatchView(ex.system, b, i)
        views.append(v)
        save_system_configs(ex.system, ex.cfg.root)
        save_experiment_configs(ex.col_cfg, ex.cfg)
        (Path(ex.cfg.root) / COL_DIRNAME).mkdir(parents=True, exist_ok=True)
    counter = b.noise_counter()
    solo = [ex.system for _, ex in members]
    for (idx, ex), v in zip(members, views):
        ex.system = v  # collectors and checkpoints read the replica in the batch
    try:
        while lead.system.time_info.time < lead.cfg.tf:
            k = _steps_to_event(lead.system, lead, None)
            b.run(k)
            counter += k
            for (idx, ex), v in zip(members, views):
                ex.system.time_info.num_steps += k
                for _ in range(k):
                    ex.system.time_info.time += ex.system.int_cfg.dt
                ex.system.dev.set_noise_state(ex.system.dev.noise_state()[0], counter)
                v.download()
                ex.col.collect(ex.system)
                check_checkpoint(ex)
        for idx, ex in members:
            ex.col.save_data(Path(ex.cfg.root) / COL_DIRNAME)
            if ex.cfg.save_final_state:
                save_system(ex.system, Path(ex.cfg.root) / FINAL_STATE_NAME)
            with open(Path(ex.cfg.root) / ".done", "w") as f:
                f.write(f"Experiment concluded on {datetime.now().isoformat()}\n")
    finally:
        for (idx, ex), s0 in zip(members, solo):
            s0.upload()  # the experiment's own system holds the final state
            ex.system = s0


def load_experiment_batch(root, custom_step=None) -> ExperimentBatch:
    """src/experiments.jl:522-548."""
    root = Path(root)
    configs = load_configs(root / EXP_COL_CONFIGS_NAME)
    exp_cfg, col_cfg = configs["experiment"], configs["collector"]
    values = load_experiment_batch_values(root)
    init_system = load_system(root / "init_system")
    if Path(exp_cfg.root) != root:
        exp_cfg.root = str(root)
        for entry in sorted((root / "data").iterdir()) if (root / "data").is_dir() else ():
            cfg_path = entry / EXP_COL_CONFIGS_NAME
            if cfg_path.is_file():
                c = load_configs(cfg_path)
                e = c["experiment"]
                e.root = str(entry)
                save_experiment_configs(c["collector"], e)
    return ExperimentBatch(exp_cfg, col_cfg, init_system, values, custom_step)


def load_experiment_batch_values(root):
    with open(Path(root) / "values.json") as f:
        return load_component(json.load(f))
