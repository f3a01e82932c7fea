"""Mesh description helpers shared by the Python side.

The authoritative mesh (face radii by bisection, N_phi per annulus and the
random phi offsets) is computed by the C library (``sim_set_rz``,
``sim_set_N_p``, ``cell_create``); this module only restates the explicit
N_phi policy so that callers can build meshes for other codes.
"""
from __future__ import annotations

import math

import numpy as np


def n_p_round(r_faces, r_min: float, npcap: int) -> np.ndarray:
    """N_phi = floor(2 pi r_c / dr + 0.5) per annulus, capped at NPCAP > 0.

    Like the reference's sim_set_N_p (src/Sim/sim_set.c:17-29), an annulus
    whose centre lies inside R_Min uses the width of the next annulus."""
    rf = np.asarray(r_faces, dtype=np.float64)
    NR = len(rf) - 1
    out = np.empty(NR, dtype=np.int32)
    for i in range(NR):
        r_cell = 0.5 * (rf[i] + rf[i + 1])
        dr = rf[i + 1] - rf[i]
        if r_cell < r_min:
            dr = rf[i + 2] - rf[i + 1]
        n = int(math.floor(2.0 * math.pi * r_cell / dr + 0.5))
        if npcap > 0 and n > npcap:
            n = npcap
        out[i] = n
    return out
