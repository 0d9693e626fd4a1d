"""Mirror of Mavi.MaviSerder and Mavi.Rings.RingsSerder (src/serder.jl,
src/rings/serder.jl) for the GPU path.

    save_system(system, root)   root/configs.json and root/state/
    load_system(root)           the system, continuing exactly where it was

configs.json keeps the reference's layout: "sys_type" and, for each of
space_cfg, dynamic_cfg, int_cfg, info, debug_info and time_info, a
{"type": ..., "data": ...} record (get_obj_save_data_json, src/serder.jl:
28-33).  The state folder holds type.json, data.npz (the arrays the
reference serializes: pos/vel, pos/pol_angle, or rings_pos, pol, types,
num_particles, active mask) and rng.json: the device noise stream's seed
and step counter (the serialized system.rng's role), the host generator of
random spawn polarisations and the draws it already handed to the device.
Saving downloads the device state first (the device -> host sync).
"""
from __future__ import annotations

import dataclasses
import json
from pathlib import Path

import numpy as np

from . import configs as _configs
from . import rings as _rings
from . import systems as _systems

_REGISTRY: dict[str, type] = {}


def _key(cls) -> str:
    return f"{cls.__module__.rsplit('.', 1)[-1]}.{cls.__name__}"


def register(*classes):
    """Make classes loadable by name (user collectors and configs too)."""
    for c in classes:
        _REGISTRY[_key(c)] = c


register(_configs.RectangleCfg, _configs.CircleCfg, _configs.Line2D, _configs.LinesCfg, _configs.ManyGeometries,
         _configs.RigidWalls, _configs.PeriodicWalls, _configs.SlipperyWalls, _configs.PotentialWalls,
         _configs.ManyWalls, _configs.SpaceCfg, _configs.HarmTruncCfg, _configs.LenJonesCfg, _configs.SzaboCfg,
         _configs.RunTumbleCfg, _configs.Sequencial, _configs.Threaded, _configs.GPU, _configs.ChunksCfg,
         _configs.IntCfg, _systems.TimeInfo, _rings.HarmTruncCfg, _rings.InteractionMatrix, _rings.RingsCfg,
         _rings.NeighborsCfg, _rings.InvasionsCfg, _rings.IntCfgExtra, _rings.SourceCfg, _rings.SinkCfg)


def get_obj_save_data_json(obj):
    """{"type": ..., "data": ...} record of obj (src/serder.jl:28-33)."""
    if obj is None or isinstance(obj, (bool, int, float, str)):
        return obj
    if isinstance(obj, np.generic):
        return obj.item()
    if isinstance(obj, np.ndarray):
        return {"type": "ndarray", "data": {"dtype": str(obj.dtype), "shape": list(obj.shape),
                                            "values": obj.ravel().tolist()}}
    if isinstance(obj, (list, tuple)):
        return {"type": type(obj).__name__, "data": [get_obj_save_data_json(x) for x in obj]}
    if isinstance(obj, dict):
        return {"type": "dict", "data": {str(k): get_obj_save_data_json(v) for k, v in obj.items()}}
    key = _key(type(obj))
    if key not in _REGISTRY:
        raise TypeError(f"{key} is not registered for saving (serder.register)")
    if dataclasses.is_dataclass(obj):
        data = {f.name: get_obj_save_data_json(getattr(obj, f.name)) for f in dataclasses.fields(obj)}
    else:
        data = {k: get_obj_save_data_json(v) for k, v in vars(obj).items() if not k.startswith("_")}
    return {"type": key, "data": data}


def load_component(record):
    """Inverse of get_obj_save_data_json."""
    if not isinstance(record, dict):
        return record
    t, d = record["type"], record["data"]
    if t == "ndarray":
        return np.asarray(d["values"], dtype=d["dtype"]).reshape(d["shape"])
    if t in ("list", "tuple"):
        out = [load_component(x) for x in d]
        return out if t == "list" else tuple(out)
    if t == "dict":
        return {k: load_component(v) for k, v in d.items()}
    cls = _REGISTRY[t]
    vals = {k: load_component(v) for k, v in d.items()}
    if dataclasses.is_dataclass(cls):
        return cls(**vals)
    obj = cls.__new__(cls)
    obj.__dict__.update(vals)
    return obj


def save_component_json(obj, root, name):
    Path(root).mkdir(parents=True, exist_ok=True)
    with open(Path(root) / f"{name}.json", "w") as f:
        json.dump(get_obj_save_data_json(obj), f, indent=1)


def load_component_json(path):
    with open(path) as f:
        return load_component(json.load(f))


def _is_rings(system) -> bool:
    # (a replica seen through the experiment runner's batch view wraps its system)
    return isinstance(getattr(system, "_system", system), _rings.RingsSystem)


def _info_data(system):
    if not _is_rings(system):
        return None
    src = None
    if system.sources is not None:
        src = [s.cfg if isinstance(s, _rings.Source) else s for s in system.sources]
    return {"p_neighbors_cfg": system.p_neighbors_cfg, "r_neighbors_cfg": None, "source_cfg": src,
            "user_data": None}


def save_system_configs(system, root, name="configs"):
    """src/serder.jl:41-55."""
    Path(root).mkdir(parents=True, exist_ok=True)
    data = {
        "sys_type": "RingsSys" if _is_rings(system) else "StandardSys",
        "space_cfg": get_obj_save_data_json(system.space_cfg),
        "dynamic_cfg": get_obj_save_data_json(system.dynamic_cfg),
        "int_cfg": get_obj_save_data_json(system.int_cfg),
        "info": get_obj_save_data_json(_info_data(system)),
        "debug_info": None,
        "time_info": get_obj_save_data_json(system.time_info),
    }
    with open(Path(root) / f"{name}.json", "w") as f:
        json.dump(data, f, indent=1)


def save_state(system, path):
    """The state (downloaded from the device) and the noise stream's position."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    system.download()
    st = system.state
    if _is_rings(system):
        arrays = {"rings_pos": st.rings_pos, "pol": st.pol,
                  "num_particles": np.atleast_1d(np.asarray(st.num_particles, dtype=np.int64))}
        if st.types is not None:
            arrays["types"] = st.types
        if st.rings_ids is not None:
            arrays["active_state"] = st.rings_ids.mask
        kind = "RingsState"
    elif isinstance(st, _systems.SelfPropelledState):
        arrays, kind = {"pos": st.pos, "pol_angle": st.pol_angle}, "SelfPropelledState"
    else:
        arrays, kind = {"pos": st.pos, "vel": st.vel}, "SecondLawState"
    np.savez(path / "data.npz", **arrays)
    with open(path / "type.json", "w") as f:
        json.dump({"type": kind, "num_particles_scalar": bool(np.ndim(getattr(st, "num_particles", 0)) == 0)}, f)
    seed, counter = system.dev.noise_state()
    rng = {"device_seed": seed, "device_counter": counter}
    if _is_rings(system) and system.sources is not None:
        if system._rng is not None:
            rng["host"] = system._rng.bit_generator.state
        rng["pending_pol_draws"] = system.pending_pol_draws().tolist()
        rng["pol_stream"] = system.dev.rings_var_state()[5]
    with open(path / "rng.json", "w") as f:
        json.dump(rng, f)


def save_system(system, root):
    """Save a system in the folder root (src/serder.jl:58-63)."""
    save_system_configs(system, root)
    save_state(system, Path(root) / "state")


def load_state(path):
    path = Path(path)
    with open(path / "type.json") as f:
        meta = json.load(f)
    with np.load(path / "data.npz", allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    kind = meta["type"]
    if kind == "RingsState":
        nump = arrays["num_particles"]
        nump = int(nump[0]) if meta.get("num_particles_scalar", True) else [int(v) for v in nump]
        active = arrays.get("active_state")
        return _rings.RingsState(rings_pos=arrays["rings_pos"], pol=arrays["pol"], num_particles=nump,
                                 types=arrays.get("types"),
                                 active_state=None if active is None else _rings.ActiveState(active))
    if kind == "SelfPropelledState":
        return _systems.SelfPropelledState(pos=arrays["pos"], pol_angle=arrays["pol_angle"])
    return _systems.SecondLawState(pos=arrays["pos"], vel=arrays["vel"])


def load_system(root, state_path=None, time_info_path=None):
    """load_system(root) or load_system(configs_path, state_path,
    time_info_path) (src/serder.jl:119-165, src/rings/serder.jl:97-116)."""
    if state_path is None:
        configs_path, state_path = Path(root) / "configs.json", Path(root) / "state"
    else:
        configs_path = Path(root)
    with open(configs_path) as f:
        cfg = json.load(f)
    space_cfg = load_component(cfg["space_cfg"])
    dynamic_cfg = load_component(cfg["dynamic_cfg"])
    int_cfg = load_component(cfg["int_cfg"])
    time_info = load_component(cfg["time_info"])
    if time_info_path is not None:
        time_info = load_component_json(Path(time_info_path) / "time_info.json")
    state = load_state(state_path)
    rng_path = Path(state_path) / "rng.json"
    rng = json.load(open(rng_path)) if rng_path.is_file() else None
    seed = rng["device_seed"] if rng else 0
    if cfg["sys_type"] == "RingsSys":
        info = load_component(cfg["info"]) or {}
        host = None
        if rng and "host" in rng:
            host = np.random.default_rng()
            host.bit_generator.state = rng["host"]
        system = _rings.RingsSystem(state=state, space_cfg=space_cfg, dynamic_cfg=dynamic_cfg, int_cfg=int_cfg,
                                    time_info=time_info, seed=seed, p_neighbors_cfg=info.get("p_neighbors_cfg"),
                                    source_cfg=info.get("source_cfg"), rng=host,
                                    pol_draws=(rng or {}).get("pending_pol_draws"),
                                    pol_stream=(rng or {}).get("pol_stream", 0))
    else:
        system = _systems.System(state=state, space_cfg=space_cfg, dynamic_cfg=dynamic_cfg, int_cfg=int_cfg,
                                 time_info=time_info, seed=seed)
    if rng:
        system.dev.set_noise_state(rng["device_seed"], rng["device_counter"])
    return system
