"""The reference's checkpoint rows and restart (src/IO, src/main.c:77-119),
restated for the tests on the oracle.

``restart_oracle`` is pinned by tests/test_oracle_golden.py against the
reference's own restart (the ``io/`` records of the golden fixtures, written
by io_setbuf / io_readbuf compiled from the reference's sources)."""
import numpy as np


def centres(N_p, r_faces, z_faces):
    """columns 1 and 2 of the checkpoint rows: the annulus and layer centres
    0.5 (FacePos(i-1) + FacePos(i)) of every cell, in cell order
    (io_buffer.c:35-36)"""
    rc = 0.5 * (r_faces[:-1] + r_faces[1:])
    zc = 0.5 * (z_faces[:-1] + z_faces[1:])
    r = np.concatenate([np.full(int(n), rc[i]) for i, n in enumerate(N_p)])
    return np.tile(r, len(zc)), np.repeat(zc, int(np.sum(N_p)))


def restart_oracle(params, N_p, r_faces, z_faces, dphi0, data, T, gm):
    """the reference's restart on the oracle: tiph and primitives from the
    rows (v_z = 0 in 2D, io_buffer.c:70-72), dphi as cell_create left it,
    the time, the masses' r, phi, M and omega, then gravMass_clean_pi and the
    set-up of src/main.c:106-119"""
    from oracle.oracle import Oracle
    NZ = len(z_faces) - 1
    prim = data[:, 3:].T.copy()
    if NZ == 1:
        prim[4] = 0.0
    o = Oracle(params, N_p, r_faces, z_faces)
    o.set_state(tiph=data[:, 0].copy(), prim=prim, dphi=dphi0)
    o.init()
    sc = o.scalars()
    sc[0] = T
    for m in range(2):
        sc[3 + 10 * m: 3 + 10 * m + 4] = gm[m]
    for k in (4, 14):  # gravMass_clean_pi, GravMass_misc.c:9-17
        while sc[k] > 2. * np.pi:
            sc[k] -= 2. * np.pi
        while sc[k] < 0.0:
            sc[k] += 2. * np.pi
    o.set_scalars(sc)
    return o
