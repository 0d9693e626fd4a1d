"""Walls lists on the GPU (ManyWalls / ManyGeometries with rigid and
periodic rectangles, rigid and slippery circles, slippery lines and
potential walls; src/integration.jl:228-403) against the CPU oracle,
through the C ABI, for the three integrators that apply walls: newton_step!,
the rings step! and the self-propelled steps.

Bars: one step within 1e-13 of the box size (positions), 1e-12 relative
(velocities, forces); longer runs within 1e-9 of the box size, or, where the
dynamics amplifies ulp-level differences (rings, Szabo), within 10x the
spread of two oracle runs whose initial state differs by 1e-15."""
import math
import sys

import numpy as np
import pytest

from helpers import TRAJ_RTOL

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def wrapped_max(a, b, L=None):
    d = np.asarray(a) - np.asarray(b)
    if L is not None:
        d -= L * np.round(d / L)
    return float(np.abs(d).max())


def print_energy_case():
    """examples/print_energy.jl: five LJ particles in a rigid circle, no chunks."""
    import mavi_hip as mv

    pos = np.array([[1.0, -1.7], [-2.5, 2.1], [3.3, -3.8], [-4.0, 4.4], [5.0, -5.4]])
    vel = np.array([[0.3, 1.0], [2.0, 0.0], [5.7, 7.8], [9.8, 0.12], [3.0, 2.2]])
    space = mv.SpaceCfg(wall_type=mv.RigidWalls(), geometry_cfg=mv.CircleCfg(radius=10.0, center=(0.0, 0.0)))
    dyn = mv.LenJonesCfg(sigma=2.0, epsilon=4.0)
    ic = mv.IntCfg(dt=0.01, chunks_cfg=None, device=mv.GPU())
    return pos, vel, space, dyn, ic


def wall_force_case():
    """examples/wall_force.jl: 10x10 HarmTrunc particles in a rigid box with a
    potential circle (outside mode) and potential lines (repulsion)."""
    import mavi_hip as mv
    from mavi_hip.init_states import rectangular_grid

    dyn = mv.HarmTruncCfg(k_rep=10.0, k_atr=1.0, dist_eq=1.0, dist_max=1.2)
    r = mv.particle_radius(dyn)
    pos, g = rectangular_grid(10, 10, 0.4, r)
    rng = np.random.default_rng(3)
    ang = rng.uniform(0, 2 * math.pi, len(pos))
    vel = 0.2 * np.stack([np.cos(ang), np.sin(ang)], axis=1)  # random_vel(num_p, 1/5)
    box = mv.RectangleCfg(length=2 * g.length, height=g.height, bottom_left=g.bottom_left)
    l, h = box.length, box.height
    circle = mv.CircleCfg(radius=3 * r, center=(l / 2 + l / 4, h / 2))
    lines = mv.LinesCfg([((3 / 4 * l, 1 / 4 * h), (3 / 4 * l, 3 / 4 * h)), ((1 / 2 * l, 1 / 2 * h), (3 / 4 * l, 1 / 2 * h))])
    wpot = mv.HarmTruncCfg(k_rep=20.0, k_atr=0.0, dist_eq=r, dist_max=r * 1.1)
    space = mv.SpaceCfg.from_list([(mv.RigidWalls(), box), (mv.PotentialWalls(wpot, mode="outside"), circle),
                                   (mv.PotentialWalls(wpot), lines)])
    # extra particles within dist_eq of one wall each (the circle's centre
    # is where the two lines meet; they sit clear of the other walls), moving
    # towards it, so the potential walls act from the first step
    cx, cy = circle.center
    d = circle.radius + 0.45
    extra = np.array([[cx + d, cy], [cx + d / math.sqrt(2), cy - d / math.sqrt(2)],
                      [3 / 4 * l - 0.45, 1 / 3 * h], [0.6 * l, 1 / 2 * h + 0.45]])
    xv = np.array([[-0.3, 0.0], [-0.2, 0.2], [0.3, 0.05], [0.0, -0.3]])
    pos, vel = np.vstack([pos, extra]), np.vstack([vel, xv])
    ic = mv.IntCfg(dt=1e-3, chunks_cfg=mv.ChunksCfg(num_cols=18, num_rows=18), device=mv.GPU())
    return pos, vel, space, dyn, ic


def slippery_lines_case():
    """An LJ gas in a periodic box with slippery line segments and a slippery
    circle as obstacles (lines placed away from the particles)."""
    import mavi_hip as mv
    from mavi_hip import workloads as wl

    pos, vel, space, dyn, ic = wl.lj_periodic(nside=24, density=0.5, seed=8, cutoff=True)
    L = space.geometry_cfg.length
    a = L / 24
    y0 = 6 * a  # between lattice rows 5 and 6
    lines = mv.LinesCfg([((2 * a, y0), (12 * a, y0)), ((16 * a, 3 * a), (16 * a, 10 * a))])
    circle = mv.CircleCfg(radius=1.5, center=(18.0 * a, 18.0 * a))
    keep = np.linalg.norm(pos - np.array(circle.center), axis=1) > circle.radius + 1.0
    keep &= np.abs(pos[:, 1] - y0) > 0.7
    keep &= np.abs(pos[:, 0] - 16 * a) > 0.7
    pos, vel = pos[keep], vel[keep]
    space = mv.SpaceCfg.from_list([(mv.PeriodicWalls(), space.geometry_cfg), (mv.SlipperyWalls(), lines),
                                   (mv.SlipperyWalls(), circle)])
    return pos, vel, space, dyn, ic


@pytest.mark.parametrize("case", ["print_energy", "wall_force", "slippery_lines"])
def test_verlet_with_walls_vs_oracle(gpu_ctx, oracle_mod, case):
    import mavi_hip as mv

    say(f"[walls verlet] {case}")
    pos, vel, space, dyn, ic = {"print_energy": print_energy_case, "wall_force": wall_force_case,
                                "slippery_lines": slippery_lines_case}[case]()
    o = oracle_mod.from_system(len(pos), space, dyn, ic)
    L = mv.get_bounding_box(space.geometry_cfg).length
    per = isinstance(space.main()[0], mv.PeriodicWalls)
    for k in (1, 300, 2000):
        s = mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)
        s.advance(k)
        s.download()
        pr, vr, fr, _, _ = o.newton_steps(pos, vel, k)
        ep = wrapped_max(s.state.pos, pr, L if per else None) / L
        ev = float(np.abs(s.state.vel - vr).max()) / max(float(np.abs(vr).max()), 1e-300)
        say(f"[walls verlet] {case} {k} steps: |dx|/L {ep:.2e} |dv|/vmax {ev:.2e}")
        if k < 2000:
            assert ep <= (1e-13 if k == 1 else TRAJ_RTOL)
            assert ev <= (1e-12 if k == 1 else TRAJ_RTOL)
        else:
            # long enough for collisions and wall contacts to amplify the
            # per-step rounding differences (device and host LJ round
            # differently at every step): bounded by the oracle stepped with
            # ulp-sized perturbations of the velocities at every step
            pq, vq = pos.copy(), vel.copy()
            rng = np.random.default_rng(7)
            for _ in range(k):
                vq = vq * (1 + 2e-16 * rng.choice([-1.0, 1.0], size=vq.shape))
                pq, vq, _, _, _ = o.newton_steps(pq, vq, 1)
            sp_ = wrapped_max(pq, pr, L if per else None) / L
            say(f"[walls verlet] {case}: oracle under per-step ulp noise |dx|/L {sp_:.2e}")
            assert ep <= max(10 * sp_, TRAJ_RTOL)
    if case == "print_energy":
        assert np.all(np.linalg.norm(s.state.pos, axis=1) < 10.0)


@pytest.mark.parametrize("case", ["wall_force", "slippery_lines"])
def test_walls_force_reuse_bitexact(gpu_ctx, case):
    """Reusing the previous step's second evaluation (the default) with a
    walls list gives bit-identical trajectories to evaluating every time."""
    import mavi_hip as mv
    from mavi_hip import _lib

    say(f"[walls reuse] {case}")
    pos, vel, space, dyn, ic = {"wall_force": wall_force_case, "slippery_lines": slippery_lines_case}[case]()
    a = mv.System(state=mv.SecondLawState(pos=pos, vel=vel * 5.0), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)
    b = mv.System(state=mv.SecondLawState(pos=pos, vel=vel * 5.0), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)
    b.dev.set_option(_lib.MAVI_OPT_REUSE_FORCES, 0)
    for s_ in (a, b):
        s_.advance(41)
        s_.advance(6)
        s_.download()
    assert np.array_equal(a.state.pos, b.state.pos)
    assert np.array_equal(a.state.vel, b.state.vel)


def rings_obstacles_case():
    """examples/rings_circle_obs.jl: rings in a periodic box with two slippery
    circles; rings overlapping the circles are left out."""
    import mavi_hip as mv
    from mavi_hip import rings as R
    from mavi_hip import workloads as wl

    st, space, cfg, ic = wl.rings_mixture(nrings=64, seed=5)
    g = space.geometry_cfg
    L = g.length
    ring_r = 1.0 / (2 * math.sin(math.pi / 20))
    c1 = mv.CircleCfg(radius=2 * ring_r, center=(L / 2, L / 4))
    c2 = mv.CircleCfg(radius=2 * ring_r, center=(L / 2, 3 * L / 4))
    cen = st.rings_pos.mean(axis=1)
    keep = np.ones(st.num_rings, dtype=bool)
    for c in (c1, c2):
        keep &= np.linalg.norm(cen - np.array(c.center), axis=1) > c.radius + ring_r + 1.0
    st2 = R.RingsState(rings_pos=st.rings_pos[keep], pol=st.pol[keep], num_particles=st.num_particles,
                       types=st.types[keep])
    space2 = mv.SpaceCfg.from_list([(mv.PeriodicWalls(), g), (mv.SlipperyWalls(), c1), (mv.SlipperyWalls(), c2)])
    return st2, space2, cfg, ic


def test_rings_with_slippery_obstacles_vs_oracle(gpu_ctx, oracle_mod):
    from mavi_hip import rings as R

    st, space, cfg, ic = rings_obstacles_case()
    say(f"[walls rings] {st.num_rings} rings")
    o = oracle_mod.from_rings(st, space, cfg, ic)
    L = space.geometry_cfg.list[0].length
    k = 200
    noise = np.random.default_rng(2).standard_normal((k, st.num_rings)) * 0.3
    pos0, pol0 = st.pos.copy(), st.pol.copy()
    s = R.RingsSystem(state=st, space_cfg=space, dynamic_cfg=cfg, int_cfg=ic)
    s.dev.rings_step(1, noise[:1])
    s.download()
    p1, q1, _, _, _ = o.rings_steps(pos0, pol0, 1, noise=noise[:1])
    assert wrapped_max(s.state.pos, p1, L) <= 1e-13 * L
    s.dev.rings_step(k - 1, noise[1:])
    s.download()
    pr, qr, _, _, _ = o.rings_steps(pos0, pol0, k, noise=noise)
    pq, qq, _, _, _ = o.rings_steps(pos0, pol0 + 1e-15, k, noise=noise)
    ep, sp_ = wrapped_max(s.state.pos, pr, L), wrapped_max(pq, pr, L)
    say(f"[walls rings] {k} steps: device vs oracle {ep:.2e}, oracle sensitivity {sp_:.2e}")
    assert ep <= max(10 * sp_, TRAJ_RTOL * L)
    # the obstacles are respected: no particle centre inside a circle
    for c in space.geometry_cfg.list[1:]:
        d = s.state.pos - np.array(c.center)
        d -= L * np.round(d / L)
        assert np.linalg.norm(d, axis=1).min() > c.radius


def test_szabo_with_obstacles_vs_oracle(gpu_ctx, oracle_mod):
    import mavi_hip as mv
    from mavi_hip import workloads as wl

    pos, ang, space, dyn, ic = wl.szabo_lattice(nside=24, seed=3)
    L = space.geometry_cfg.length
    c1 = mv.CircleCfg(radius=2.0, center=(6.0, 6.0))
    c2 = mv.CircleCfg(radius=2.5, center=(17.0, 15.0))
    keep = (np.linalg.norm(pos - c1.center, axis=1) > 2.6) & (np.linalg.norm(pos - c2.center, axis=1) > 3.1)
    pos, ang = pos[keep], ang[keep]
    wpot = mv.HarmTruncCfg(k_rep=20.0, k_atr=0.0, dist_eq=0.5, dist_max=0.55)
    space = mv.SpaceCfg.from_list([(mv.PeriodicWalls(), space.geometry_cfg), (mv.SlipperyWalls(), c1),
                                   (mv.PotentialWalls(wpot, mode="repulsion"), c2)])
    n, k = len(pos), 100
    noise = np.random.default_rng(4).standard_normal((k, n))
    o = oracle_mod.from_system(n, space, dyn, ic)
    s = mv.System(state=mv.SelfPropelledState(pos=pos, pol_angle=ang), space_cfg=space, dynamic_cfg=dyn, int_cfg=ic)
    s.advance(1, noise=noise[:1])
    s.download()
    p1, a1, _ = o.sp_steps(pos, ang, 1, noise=noise[:1])
    assert wrapped_max(s.state.pos, p1, L) <= 1e-13 * L
    assert np.abs(s.state.pol_angle - a1).max() <= 1e-12
    s.advance(k - 1, noise=noise[1:])
    s.download()
    pr, ar, _ = o.sp_steps(pos, ang, k, noise=noise)
    pq, aq, _ = o.sp_steps(pos, ang + 1e-15, k, noise=noise)
    ep, sp_ = wrapped_max(s.state.pos, pr, L), wrapped_max(pq, pr, L)
    say(f"[walls szabo] {k} steps: device vs oracle {ep:.2e}, oracle sensitivity {sp_:.2e}")
    assert ep <= max(10 * sp_, TRAJ_RTOL * L)


def test_walls_unsupported(gpu_ctx):
    import mavi_hip as mv
    from mavi_hip import rings as R
    from mavi_hip import workloads as wl

    pos, vel, space, dyn, ic = wl.lj_periodic(nside=8, density=0.5, seed=1)
    circ = mv.CircleCfg(radius=1.0, center=(2.0, 2.0))
    bad = mv.SpaceCfg.from_list([(mv.SlipperyWalls(), circ), (mv.PeriodicWalls(), space.geometry_cfg)])
    with pytest.raises(mv.UnsupportedOnGPU):  # periodic walls must be the main wall
        mv.System(state=mv.SecondLawState(pos=pos, vel=vel), space_cfg=bad, dynamic_cfg=dyn, int_cfg=ic)
    p2, a2, sp2, dyn2, ic2 = wl.szabo_lattice(nside=8)
    rigid = mv.SpaceCfg.from_list([(mv.PeriodicWalls(), sp2.geometry_cfg), (mv.RigidWalls(), circ)])
    with pytest.raises(mv.UnsupportedOnGPU):  # a self-propelled state has no velocity to reflect
        mv.System(state=mv.SelfPropelledState(pos=p2, pol_angle=a2), space_cfg=rigid, dynamic_cfg=dyn2, int_cfg=ic2)
    st, sp3, cfg3, ic3 = wl.rings_mixture(nrings=16, seed=1)
    rig3 = mv.SpaceCfg.from_list([(mv.PeriodicWalls(), sp3.geometry_cfg), (mv.RigidWalls(), circ)])
    with pytest.raises(mv.UnsupportedOnGPU):
        R.RingsSystem(state=st, space_cfg=rig3, dynamic_cfg=cfg3, int_cfg=ic3)
