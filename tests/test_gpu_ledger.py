"""Conservation ledger of the HIP path (the class API's timestep_substep,
which runs the same kernels as timestep_rk2): over 5 RK2 steps of outflow
disks, per substep, the interior mass and angular momentum change by exactly
what the boundary r faces carry plus the gravitational torque term
(tests/ledger_util.py), to 1e-13 of the total.  The GPU's boundary fluxes
and torque terms also agree with the oracle's per substep, and the adjoint's
loss (the time average of the same torque terms) with their sum."""
import ctypes

import numpy as np
import pytest

from ledger_util import DDD, LLL, interior_mask, predicted
from test_ledger import CASES, STEPS, setup

pytestmark = pytest.mark.gpu


def gpu_ledger(case):
    from disco2_amd._lib import check, dptr, lib
    L = lib()
    p, s, prim0 = setup(**CASES[case])
    sim, mpi = s.sim, s.mpi
    gm = L.gravMass_create(2)
    cells = L.cell_create(sim, mpi)
    check()
    ts = L.timestep_create(sim)
    ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)(L.gravMass_init_ptr(sim))(gm, sim)
    L.cell_upload(cells, 0, dptr(np.ascontiguousarray(prim0)))
    L.gravMass_clean_pi(gm, sim)
    L.cell_calc_cons(cells, sim)
    assert L.disco_cell_ledger_enable(cells, 1) == 0
    nq = s.NQ
    mask = interior_mask(s.N_p)

    def field(f):
        out = np.empty(nq * s.nzones)
        L.cell_download(cells, f, dptr(out))
        return out.reshape(nq, -1)[:, mask]

    recs = []
    t0 = L.timestep_get_t(ts)
    for _ in range(STEPS):  # timestep_rk2 through the class API, src/TimeStep/timestep_algorithm.c:11-41
        L.cell_clear_w(cells, sim)
        L.cell_set_w(cells, sim)
        L.timestep_set_dt(ts, cells, sim, gm)
        L.cell_copy(cells, sim)
        L.gravMass_copy(gm, sim)
        L.timestep_set_RK(ts, 0.0)
        for fac, RK in ((1.0, 0.0), (0.5, 0.5)):
            u0, rk = field(1), field(2)
            L.timestep_substep(ts, cells, sim, gm, mpi, fac)
            led = np.empty(2 * (nq + 3) + 1)
            assert L.disco_cell_ledger(cells, dptr(led)) == 0
            check()
            recs.append((RK, u0, rk, led, field(1)))
            L.gravMass_move(sim, gm, L.timestep_get_t(ts), fac * L.timestep_dt(ts))
            L.timestep_set_RK(ts, 0.5)
        L.timestep_update_t(ts)
    T = L.timestep_get_t(ts) - t0
    L.cell_destroy(cells, sim)
    L.timestep_destroy(ts)
    L.gravMass_destroy(gm)
    return p, s, prim0, recs, T


def oracle_ledgers(p, s, prim0):
    from oracle.oracle import Oracle
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=prim0)
    o.init()
    out = []
    for _ in range(STEPS):
        o.step_begin()
        for fac in (1.0, 0.5):
            o.substep(fac)
            out.append(o.ledger())
        o.step_end()
    return out


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_conserves_mass_and_angular_momentum(case):
    print(f"case {case}", flush=True)
    p, s, prim0, recs, T = gpu_ledger(case)
    nq = s.NQ
    for n, (RK, u0, rk, led, u1) in enumerate(recs):
        for q in (DDD, LLL):
            pred = predicted(u0[q].sum(), rk[q].sum(), RK, led, nq, q)
            scale = np.abs(u0[q]).sum()
            assert abs(u1[q].sum() - pred) <= 1e-13 * scale, (case, n, q, u1[q].sum(), pred)
    # the same boundary fluxes and torque terms as the oracle's
    ref = oracle_ledgers(p, s, prim0)
    for n, ((_, _, _, led, _), r) in enumerate(zip(recs, ref)):
        scale = np.max(np.abs(r))
        assert np.max(np.abs(led - r)) <= 1e-11 * scale, (case, n, led, r)
    # the adjoint's loss is the time average of these torque terms
    from disco2_amd.api import Batch, Setup
    b = Batch([Setup(p)])
    b.upload("prim", prim0)
    b.init()
    loss = b.torque_loss(STEPS)
    b.close()
    tq = sum(led[2 * (nq + 3)] for (_, _, _, led, _) in recs)
    assert abs(loss - tq / T) <= 1e-12 * abs(tq / T), (loss, tq / T)
