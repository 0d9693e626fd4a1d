"""Mirror of Mavi.Integration / Mavi.RunSystem / Mavi.Quantities on the GPU.

  newton_step(system)       newton_step!, src/integration.jl:507-515
  szabo_step / rtp_step     szabo_step! / rtp_step!, src/integration.jl:517-535
  get_step_function(system) src/integration.jl:537-548, src/rings/integration.jl:545
  calc_forces(system)       calc_forces!, src/integration.jl:226
  update_chunks(system)     update_chunks!, src/chunks.jl:150-161
  run_system(system; ...)   src/run_system.jl:7-23 (resident fast path)
  kinetic_energy / potential_energy   src/quantities.jl:12-18, 46-66

The per-step functions keep the reference's sync model: the host state is
uploaded, one step runs on the GPU and the state is downloaded, so user code
may read or edit ``system.state`` between steps.  ``run_system`` without a
custom step function uploads once, runs every step on the device and
downloads at the end.
"""
from __future__ import annotations


def _synced_step(system):
    # the library runs the system's own dynamics for its kind
    system.upload()
    system.advance(1)
    system.download()


def newton_step(system):
    """newton_step!, src/integration.jl:507-515."""
    _synced_step(system)


def szabo_step(system):
    """szabo_step!, src/integration.jl:517-525 (device noise stream)."""
    _synced_step(system)


def rtp_step(system):
    """rtp_step!, src/integration.jl:527-535 (device noise stream)."""
    _synced_step(system)


def calc_forces(system):
    return system.calc_forces()


def update_chunks(system):
    return system.update_chunks()


def get_step_function(system):
    from .rings import RingsSystem, step

    if isinstance(system, RingsSystem):
        return step
    from .configs import RunTumbleCfg, SzaboCfg

    if isinstance(system.dynamic_cfg, SzaboCfg):
        return szabo_step
    if isinstance(system.dynamic_cfg, RunTumbleCfg):
        return rtp_step
    return newton_step


def _steps_until(system, tf):
    # while system.time_info.time < tf: step  (time accumulates dt per step)
    t = system.time_info.time
    n = 0
    while t < tf:
        t += system.int_cfg.dt
        n += 1
    return n


def run_system(system, tf=None, num_steps=None, step_func=None):
    """src/run_system.jl:7-23."""
    if step_func is not None:
        if tf is not None:
            while system.time_info.time < tf:
                step_func(system)
        else:
            for _ in range(int(num_steps)):
                step_func(system)
        return
    n = _steps_until(system, tf) if tf is not None else int(num_steps)
    system.upload()
    system.advance(n)
    system.download()


def kinetic_energy(system):
    """kinetic_energy(system.state) computed on the device."""
    system.upload()
    ke, _ = system.dev.observables()
    return ke


def potential_energy(system):
    """potential_energy(system, ::LenJonesCfg): all pairs, as the reference."""
    system.upload()
    _, pe = system.dev.observables()
    return pe
