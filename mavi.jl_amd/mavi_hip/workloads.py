"""Deterministic synthetic inputs for the benchmark workloads.

Each generator states its RNG (numpy PCG64 via default_rng(seed)).  They
return plain arrays plus the configs, so the same input feeds the GPU path
and the CPU oracle.

  lj_gas_1024        1024 LJ particles (32x32 lattice, offset 0.4), rigid
                     walls, speed 0.2 at a uniform angle (seed 0), 17x17
                     chunks, LJ cut at 2.5 sigma (cutoff=None gives the
                     reference's stencil-ranged LJ), dt = 1e-3
  lj_periodic        1024x1024 LJ particles, periodic square box at number
                     density 0.7, jittered lattice (+-0.1 sigma, seed 1),
                     Gaussian velocities T = 1 with zero net momentum,
                     cutoff 2.5 sigma, chunks of side >= 2.5, dt = 1e-3
  rings_mixture      active rings: nrings x 20 particles, periodic box at
                     ring packing fraction 0.8 (nrings * pi * R_out^2 / L^2,
                     R_out = regular 20-gon circumradius + particle radius),
                     springs k = 20, l = 1, HarmTrunc repulsion k = 10 with
                     interaction radius 1 (dist_eq = dist_max = 1), two types
                     50/50 (seed 2) whose target areas are 1.0 and 0.9 of the
                     regular n-gon area, v0 = 1, D_r = 0.1, dt = 1e-3
  batch_ensemble     512 independent replicas, particle counts log-uniform
                     in [1000, 64000] (seed 3); even replicas are lj_periodic
                     gases (nside = round(sqrt(N)), seed 1000 + i), odd ones
                     rings_mixture systems (round(N / 20) rings of 20,
                     seed 2000 + i, device noise seed 3000 + i)
  lj_obstacles       lj_periodic with a walls list: the periodic box first,
                     then 16 slippery circles (radius 8, on a 4x4 grid),
                     8 slippery line segments (length 40) and one potential
                     circle (HarmTrunc k_rep 20, range 1, repulsion);
                     particles within 1.5 of an obstacle are left out
  szabo_lattice      nside x nside Szabo particles (examples/szabo.jl
                     parameters: vo 1, mobility 1, relax_time 1, k_rep 10,
                     k_adh 0.75, r_eq 1, r_max 1.1, rot_diff 0.01), square
                     lattice of spacing r_eq in a periodic box, polarisations
                     uniform (seed), chunks of side >= r_max, dt = 0.01
  rtp_gas            nside x nside run-and-tumble particles (examples/rtp.jl
                     parameters: vo 1, sigma 1, epsilon 1, tumble_rate 1) on a
                     lattice at number density 0.8 (spacing 1.118, inside the
                     2^(1/6) sigma range) jittered by +-0.02 (seed), periodic,
                     chunks of side >= 2^(1/6) sigma, dt = 1e-3
  rings_sources      rings_mixture with a variable population: the rings
                     whose centres lie within one ring radius of a source
                     region start inactive, and 20% of the others (seed 6);
                     one source of 8 x 8 spawn boxes (a regular 20-gon of the
                     mixture's placement radius each, pad = particle radius,
                     random spawn polarisation) and one rectangular sink of
                     width 4 ring radii and height L / 2 at the far side
"""
from __future__ import annotations

import math

import numpy as np

from .configs import (ChunksCfg, GPU, IntCfg, LenJonesCfg, PeriodicWalls, RectangleCfg, RigidWalls, RunTumbleCfg,
                      SpaceCfg, SzaboCfg, particle_radius)
from .init_states import rectangular_grid


def lj_gas_1024(nx=32, ny=32, cutoff=2.5):
    dyn = LenJonesCfg(sigma=1.0, epsilon=1.0, cutoff=cutoff)
    r = particle_radius(dyn)
    pos, geom = rectangular_grid(nx, ny, 0.4, r)
    rng = np.random.default_rng(0)
    ang = rng.random(len(pos)) * 2 * math.pi
    vel = np.stack([0.2 * np.cos(ang), 0.2 * np.sin(ang)], axis=1)
    cols = int(geom.length // 2.5)  # chunk side >= 2.5 sigma (17 for 32x32)
    rows = int(geom.height // 2.5)
    space = SpaceCfg(wall_type=RigidWalls(), geometry_cfg=geom)
    int_cfg = IntCfg(dt=1e-3, chunks_cfg=ChunksCfg(num_cols=cols, num_rows=rows), device=GPU())
    return pos, vel, space, dyn, int_cfg


def lj_periodic(nside=1024, density=0.7, rc=2.5, seed=1, cutoff=True):
    n = nside * nside
    L = math.sqrt(n / density)
    a = L / nside
    rng = np.random.default_rng(seed)
    ix, iy = np.meshgrid(np.arange(nside), np.arange(nside), indexing="xy")
    pos = np.stack([(ix.ravel() + 0.5) * a, (iy.ravel() + 0.5) * a], axis=1)
    pos += rng.uniform(-0.1, 0.1, size=pos.shape)
    vel = rng.standard_normal(size=pos.shape)
    vel -= vel.mean(axis=0)
    cols = int(L // rc)
    dyn = LenJonesCfg(sigma=1.0, epsilon=1.0, cutoff=rc if cutoff else None)
    space = SpaceCfg(wall_type=PeriodicWalls(), geometry_cfg=RectangleCfg(length=L, height=L))
    int_cfg = IntCfg(dt=1e-3, chunks_cfg=ChunksCfg(num_cols=cols, num_rows=cols), device=GPU())
    return pos, vel, space, dyn, int_cfg


def rings_mixture(nrings=40000, nper=20, phi=0.8, seed=2, rot_diff=0.1, k_area=1.0, chunk=1.0):
    """Binary active-ring mixture (see the module docstring).  The target area
    of a ring is A0 = (n l / p0)^2 (src/rings/integration.jl:170); the
    contraction c of a type maps to p0 = n l / sqrt(c * A_ngon(n, l)).
    Chunks are the largest whose side is >= ``chunk`` (>= the interaction
    range 1; the learned propagator's features need >= their radius)."""
    from .rings import HarmTruncCfg, RingsCfg, RingsState

    rng = np.random.default_rng(seed)
    types = rng.integers(1, 3, size=nrings).astype(np.int32)
    l0 = 1.0
    p_r = 0.5
    area_ngon = nper * l0 * l0 / (4 * math.tan(math.pi / nper))
    p0_a = nper * l0 / math.sqrt(area_ngon)
    p0_b = nper * l0 / math.sqrt(0.9 * area_ngon)
    inter = HarmTruncCfg(k_rep=10.0, k_atr=0.0, dist_eq=2 * p_r, dist_max=2 * p_r)
    cfg = RingsCfg(p0=[p0_a, p0_b], relax_time=[1.0, 1.0], vo=[1.0, 1.0], mobility=[1.0, 1.0],
                   rot_diff=[rot_diff, rot_diff], k_area=[k_area, k_area], k_spring=[20.0, 20.0],
                   l_spring=[l0, l0], interaction_finder=inter, num_particles=[nper, nper])
    ring_r = l0 / (2 * math.sin(math.pi / nper))  # circumradius of the regular n-gon of side l0
    r_out = ring_r + p_r
    L = math.sqrt(nrings * math.pi * r_out * r_out / phi)
    m = int(math.ceil(math.sqrt(nrings)))
    spacing = L / m
    ang = np.arange(nper) * 2 * math.pi / nper
    circle = np.stack([np.cos(ang), np.sin(ang)], axis=1) * (ring_r * 0.97)
    k = np.arange(nrings)
    centers = np.stack([(k % m + 0.5) * spacing, (k // m + 0.5) * spacing], axis=1)
    rings_pos = np.mod(centers[:, None, :] + circle[None, :, :], L)
    pol = rng.random(nrings) * 2 * math.pi
    state = RingsState(rings_pos=rings_pos, pol=pol, num_particles=[nper, nper], types=types)
    cols = int(L // max(chunk, 2 * p_r))
    space = SpaceCfg(wall_type=PeriodicWalls(), geometry_cfg=RectangleCfg(length=L, height=L))
    int_cfg = IntCfg(dt=1e-3, chunks_cfg=ChunksCfg(num_cols=cols, num_rows=cols), device=GPU())
    return state, space, cfg, int_cfg


def batch_ensemble(nrep=512, nmin=1000, nmax=64000, seed=3):
    """Replica list [(kind, case, noise_seed)], kind "plain" with case =
    lj_periodic(...) or "rings" with case = rings_mixture(...)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=nrep))
    out = []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            nside = max(8, int(round(math.sqrt(n))))
            out.append(("plain", lj_periodic(nside=nside, density=0.7, seed=1000 + i, cutoff=True), 0))
        else:
            nr = max(2, int(round(n / 20)))
            out.append(("rings", rings_mixture(nrings=nr, seed=2000 + i), 3000 + i))
    return out


def lj_obstacles(nside=1024, seed=1):
    """(pos, vel, space, dyn, int_cfg) of lj_periodic inside a walls list (see
    the module docstring)."""
    from .configs import CircleCfg, LinesCfg, ManyGeometries, PotentialWalls, SlipperyWalls
    from .configs import HarmTruncCfg as HT

    pos, vel, space, dyn, ic = lj_periodic(nside=nside, seed=seed)
    box = space.geometry_cfg
    L = box.length
    circles = [CircleCfg(radius=8.0, center=((i + 0.5) * L / 4, (j + 0.5) * L / 4)) for i in range(4) for j in range(4)]
    segs = []
    for k in range(8):
        x0, y0 = (k + 0.25) * L / 8, (0.1 + 0.1 * k) * L
        segs.append(((x0, y0), (x0 + 40.0, y0)))
    lines = LinesCfg(segs)
    pc = CircleCfg(radius=5.0, center=(0.13 * L, 0.87 * L))
    keep = np.ones(len(pos), dtype=bool)
    for c in circles + [pc]:
        keep &= np.linalg.norm(pos - np.array(c.center), axis=1) > c.radius + 1.5
    for (a, b) in segs:
        inside = (pos[:, 0] > a[0] - 1.5) & (pos[:, 0] < b[0] + 1.5)
        keep &= ~(inside & (np.abs(pos[:, 1] - a[1]) < 1.5))
    pos, vel = pos[keep], vel[keep]
    vel = vel - vel.mean(axis=0)
    pairs = [(PeriodicWalls(), box)] + [(SlipperyWalls(), c) for c in circles] + [(SlipperyWalls(), lines),
             (PotentialWalls(HT(k_rep=20.0, k_atr=0.0, dist_eq=1.0, dist_max=1.0)), pc)]
    space = SpaceCfg.from_list(pairs)
    return pos, vel, space, dyn, ic


def szabo_lattice(nside=1024, seed=4, rot_diff=0.01, dt=0.01):
    """Szabo particles on a square lattice (see the module docstring).
    Returns (pos, pol_angle, space, dyn, int_cfg)."""
    dyn = SzaboCfg(vo=1.0, mobility=1.0, relax_time=1.0, k_rep=10.0, k_adh=0.75, r_eq=1.0, r_max=1.1,
                   rot_diff=rot_diff)
    a = dyn.r_eq
    L = nside * a
    ix, iy = np.meshgrid(np.arange(nside), np.arange(nside), indexing="xy")
    pos = np.stack([(ix.ravel() + 0.5) * a, (iy.ravel() + 0.5) * a], axis=1)
    rng = np.random.default_rng(seed)
    ang = rng.random(len(pos)) * 2 * math.pi
    cols = int(L // dyn.r_max)
    space = SpaceCfg(wall_type=PeriodicWalls(), geometry_cfg=RectangleCfg(length=L, height=L))
    int_cfg = IntCfg(dt=dt, chunks_cfg=ChunksCfg(num_cols=cols, num_rows=cols), device=GPU())
    return pos, ang, space, dyn, int_cfg


def rtp_gas(nside=1024, density=0.8, seed=5, tumble_rate=1.0, dt=1e-3, jitter=0.02):
    """Run-and-tumble particles on a jittered lattice (see the module
    docstring).  Returns (pos, pol_angle, space, dyn, int_cfg)."""
    dyn = RunTumbleCfg(vo=1.0, sigma=1.0, epsilon=1.0, tumble_rate=tumble_rate)
    n = nside * nside
    L = math.sqrt(n / density)
    a = L / nside
    rng = np.random.default_rng(seed)
    ix, iy = np.meshgrid(np.arange(nside), np.arange(nside), indexing="xy")
    pos = np.stack([(ix.ravel() + 0.5) * a, (iy.ravel() + 0.5) * a], axis=1)
    pos += rng.uniform(-jitter, jitter, size=pos.shape)
    ang = rng.random(n) * 2 * math.pi
    cols = int(L // dyn.cutoff())
    space = SpaceCfg(wall_type=PeriodicWalls(), geometry_cfg=RectangleCfg(length=L, height=L))
    int_cfg = IntCfg(dt=dt, chunks_cfg=ChunksCfg(num_cols=cols, num_rows=cols), device=GPU())
    return pos, ang, space, dyn, int_cfg


def rings_sources(nrings=40000, seed=6):
    """Returns (state, space_cfg, dynamic_cfg, int_cfg, [SourceCfg, SinkCfg])."""
    from .rings import ActiveState, RingsState, SinkCfg, SourceCfg

    st, sp, cfg, ic = rings_mixture(nrings=nrings)
    L = sp.geometry_cfg.length
    nper, l0, p_r = 20, 1.0, 0.5
    ring_r = l0 / (2 * math.sin(math.pi / nper))
    r_out = ring_r + p_r
    ang = np.arange(nper) * 2 * math.pi / nper
    spawn = np.stack([np.cos(ang), np.sin(ang)], axis=1) * (ring_r * 0.97)
    box = 2 * ring_r * 0.97 + 2 * p_r  # spawn box side (template extent + 2 pad)
    x0, y0 = 0.1 * L, 0.5 * L - 4 * box
    m = int(math.ceil(math.sqrt(nrings)))
    k = np.arange(nrings)
    centres = np.stack([(k % m + 0.5) * (L / m), (k // m + 0.5) * (L / m)], axis=1)
    inside = ((centres[:, 0] > x0 - r_out) & (centres[:, 0] < x0 + 8 * box + r_out) &
              (centres[:, 1] > y0 - r_out) & (centres[:, 1] < y0 + 8 * box + r_out))
    rng = np.random.default_rng(seed)
    mask = ~inside & (rng.random(nrings) >= 0.2)
    state = RingsState(rings_pos=st.rings_pos, pol=st.pol, num_particles=st.num_particles, types=st.types,
                       active_state=ActiveState(mask))
    sources = [SourceCfg(bottom_left=(x0, y0), spawn_pos=spawn, spawn_pol="random", pad=p_r, size=(8, 8)),
               SinkCfg(RectangleCfg(length=4 * r_out, height=0.5 * L, bottom_left=(0.8 * L, 0.25 * L)))]
    return state, sp, cfg, ic, sources
