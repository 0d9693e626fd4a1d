"""Shared helpers for the parity tests (GPU path vs the CPU oracle)."""
from __future__ import annotations

import numpy as np

# Tolerances (FP64):
#   single-step forces   max_i |F_gpu,i - F_ref,i|_inf / max(sum_j |f_ij|, TINY) <= 1e-12
#   100-step trajectory  max |x_gpu - x_ref| <= 1e-9 * L, max |v_gpu - v_ref| <= 1e-9 * max|v_ref|
FORCE_RTOL = 1e-12
TRAJ_RTOL = 1e-9
TINY = 1e-300


def force_rel_err(f_gpu, f_ref, absum):
    d = np.abs(np.asarray(f_gpu) - np.asarray(f_ref)).max(axis=1)
    den = np.maximum(np.asarray(absum), TINY)
    r = np.where(d == 0.0, 0.0, d / den)
    return float(r.max()) if len(r) else 0.0


def traj_err(pos_g, vel_g, pos_r, vel_r, L, periodic=True):
    """Position error with the minimum image (a particle wrapped on one side
    and not on the other is the same point), relative to the box size."""
    d = np.asarray(pos_g) - np.asarray(pos_r)
    if periodic:
        d = d - L * np.round(d / L)
    ep = float(np.abs(d).max()) / L
    ev = 0.0
    if vel_r is not None:
        ev = float(np.abs(vel_g - vel_r).max()) / max(float(np.abs(vel_r).max()), 1e-300)
    return ep, ev


def random_gas(rng, n, L, dmin=0.9, periodic=True):
    """n points in [0, L)^2 with a minimum (minimum-image) separation, drawn on
    a jittered grid so it scales to large n."""
    m = int(np.ceil(np.sqrt(n)))
    a = L / m
    assert a > dmin + 0.05, "too dense for the jittered grid"
    jit = (a - dmin) / 2 * 0.95
    k = np.arange(n)
    base = np.stack([(k % m + 0.5) * a, (k // m + 0.5) * a], axis=1)
    pos = base + rng.uniform(-jit, jit, size=base.shape)
    return pos
