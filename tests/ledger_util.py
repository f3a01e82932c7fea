"""Helpers of the conservation-ledger tests (tests/test_ledger.py,
tests/test_gpu_ledger.py): interior sums and the per-substep identity.

For one substep with RK weight RK, every interior cell's conserved state
becomes (1 - RK) U + RK U_RK + its faces' increments + its sources, then
goes through cons2prim and prim2cons.  Summed over the interior cells the
increments of faces between two interior cells cancel, so
  sum U' = (1 - RK) sum U + RK sum U_RK + B_in - B_out + S
with B the increments of the boundary r faces (inner: the interior cell is
the face's R, outer: its L) and S the sources: none for the mass (no sink),
for the angular momentum the torque term sum dt dV rho F_phi r
(src/Cell/cell_source.c:171 with a non-rotating frame) plus the viscous
flux's LLL part, which riemann_AddFlux adds conservatively
(src/Riemann/Riemann_routines.c:390-391)."""
import numpy as np

DDD, LLL = 0, 3


def interior_mask(N_p, NZ=1, ng_in=1, ng_out=2):
    NR = len(N_p)
    mask = np.zeros(int(np.sum(N_p)) * NZ, dtype=bool)
    off = np.concatenate([[0], np.cumsum(N_p)])
    for i in range(ng_in, NR - ng_out):
        mask[off[i]:off[i + 1]] = True
    return mask


def predicted(before, rk, RK, led, nq, q):
    """the interior sum of component q after the substep"""
    nv = nq + 3
    b_in, b_out = led[q], led[nv + q]
    pred = (1.0 - RK) * before + RK * rk + b_in - b_out
    if q == LLL:
        pred += led[nq + 1] - led[nv + nq + 1] + led[2 * nv]
    return pred
