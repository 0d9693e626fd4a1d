"""Mirror of Mavi.InitStates.rectangular_grid (src/init_states.jl:34-57)."""
from __future__ import annotations

import numpy as np

from .configs import RectangleCfg


def rectangular_grid(num_p_x, num_p_y, offset, radius):
    """Particles on a grid with spacing offset*radius between surfaces.

    Returns (pos (N,2), geometry_cfg), N = num_p_x*num_p_y, x fastest."""
    xs, ys = [], []
    current_y = radius * (offset + 1)
    for _ in range(num_p_y):
        current_x = -radius
        for _ in range(num_p_x):
            xs.append(current_x + radius * (2 + offset))
            ys.append(current_y)
            current_x = xs[-1]
        current_y += radius * (2 + offset)
    geometry_cfg = RectangleCfg(
        length=num_p_x * 2 * radius + radius * offset * (num_p_x + 1),
        height=num_p_y * 2 * radius + radius * offset * (num_p_y + 1),
    )
    return np.stack([np.asarray(xs), np.asarray(ys)], axis=1), geometry_cfg
