"""Structural observables of rings configurations, used to state the error of
the learned propagator against plain FP64 trajectories.

  radial_distribution   g(r) of all particle pairs (minimum image), ideal-gas
                        normalised, on fixed bins
  segregation_index     mean over rings of the fraction of their contact
                        rings (any particle pair within r_contact) that have
                        the same type: 0.5 for a random 50/50 mixture, 1 for
                        full demixing
"""
from __future__ import annotations

import numpy as np


def _wrap(pos, L, H, bl=(0.0, 0.0)):
    p = np.stack([np.mod(pos[:, 0] - bl[0], L), np.mod(pos[:, 1] - bl[1], H)], axis=1)
    return np.where(p >= np.array([L, H]), 0.0, p)


def radial_distribution(pos, L, H, r_max=6.0, dr=0.05, bl=(0.0, 0.0)):
    """(bin centres, g(r)) of the periodic configuration ``pos`` ([n][2])."""
    from scipy.spatial import cKDTree

    p = _wrap(np.asarray(pos, dtype=np.float64), L, H, bl)
    tree = cKDTree(p, boxsize=(L, H))
    pairs = tree.query_pairs(r_max, output_type="ndarray")
    d = p[pairs[:, 0]] - p[pairs[:, 1]]
    d[:, 0] -= L * np.round(d[:, 0] / L)
    d[:, 1] -= H * np.round(d[:, 1] / H)
    r = np.sqrt((d * d).sum(axis=1))
    edges = np.arange(0.0, r_max + dr / 2, dr)
    hist, _ = np.histogram(r, bins=edges)
    n = len(p)
    rho = n / (L * H)
    shell = np.pi * (edges[1:] ** 2 - edges[:-1] ** 2)
    g = 2.0 * hist / (n * rho * shell)
    return 0.5 * (edges[1:] + edges[:-1]), g


def segregation_index(pos, ring_types, nmax, L, H, r_contact=1.25, bl=(0.0, 0.0)):
    """Mean like-type fraction of each ring's contact rings."""
    from scipy.spatial import cKDTree

    p = _wrap(np.asarray(pos, dtype=np.float64), L, H, bl)
    tree = cKDTree(p, boxsize=(L, H))
    pairs = tree.query_pairs(r_contact, output_type="ndarray")
    ra, rb = pairs[:, 0] // nmax, pairs[:, 1] // nmax
    keep = ra != rb
    ra, rb = ra[keep], rb[keep]
    rp = np.unique(np.stack([np.minimum(ra, rb), np.maximum(ra, rb)], axis=1), axis=0)
    types = np.asarray(ring_types)
    nr = len(types)
    like = np.zeros(nr)
    tot = np.zeros(nr)
    same = types[rp[:, 0]] == types[rp[:, 1]]
    for a, b in ((0, 1), (1, 0)):
        np.add.at(tot, rp[:, a], 1.0)
        np.add.at(like, rp[:, a], same.astype(np.float64))
    ok = tot > 0
    return float(np.mean(like[ok] / tot[ok])) if ok.any() else float("nan")


def relative_l2(a, b):
    """||a - b|| / ||b||."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
