"""D3Q27 lattice velocities in the reference's order (NREL/marbles
Source/Stencil.H:49-79) and host-side moments of population arrays."""
from __future__ import annotations

import numpy as np

EV = np.array([
    (0, 0, 0),
    (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1),
    (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1),
    (1, 1, 1), (-1, -1, -1), (1, -1, 1), (-1, 1, -1), (1, -1, -1), (-1, 1, 1),
    (1, 1, -1), (-1, -1, 1),
], dtype=np.int64)


def velocity_density(f: np.ndarray):
    """([u, v, w], rho) from populations f of shape (27, ...): rho = sum f,
    u = sum c f / rho (0 where rho = 0)."""
    f = np.asarray(f, dtype=np.float64)
    rho = f.sum(axis=0)
    safe = np.where(rho != 0, rho, 1.0)
    vel = [np.tensordot(EV[:, d].astype(np.float64), f, axes=(0, 0)) / safe for d in range(3)]
    return vel, rho
