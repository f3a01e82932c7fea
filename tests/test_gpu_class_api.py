"""The reference's class API driven one call at a time on the GPU, against
the reference's own outputs (tests/golden/*.npz, unit/* records written by
the reference build through oracle/ref_harness.c):

* timestep_substep(..., 1.0) after cell_clear_w, cell_set_w, timestep_set_dt,
  cell_copy, gravMass_copy and timestep_set_RK(0) -- the reference's state
  after its first substep (unit/sub1);
* the same substep replayed from its pieces exactly as
  src/TimeStep/timestep_routines.c:16-113 calls them: face_create,
  cell_clean_pi, cell_adjust_RK_cons, cell_plm_p / _rz, one riemann_create /
  riemann_setup_p|rz / riemann_AddFlux / riemann_destroy per face,
  cell_add_src, cell_add_visc_src, cell_add_split_fictitious,
  cell_update_phi, cell_update_dphi, gravMass_update_RK, cell_calc_prim, the
  r and z boundaries (outflow or fixed), cell_bc_damp, the syncs and
  cell_calc_cons -- against unit/sub1 as well;
* timestep_forward_euler, two steps (unit/euler1, unit/euler2);
* cell_mindt against the reference's first dt, cell_prim2cons against the
  reference's initial conserved state, and the edges of the unit-level API.

Bitwise where no libm transcendental enters (the vortex cases), otherwise
within 1e-11 relative per zone (device sin/cos/pow differ from glibc by ulps).
"""
import ctypes

import numpy as np
import pytest

from compare import bitwise_equal, describe, max_rel, zone_rel
from golden_util import Golden, golden_names

pytestmark = pytest.mark.gpu

BITWISE = {"vortex_r16", "vortex_scalar_hll_noinner", "plugin_vortex_fixed", "vortex_three_scalars",
           "vortex_two_scalars"}
RTOL = 1e-11
UNIT = [n for n in golden_names() if Golden(n).has("unit/sub1/prim")]
EULER = [n for n in golden_names() if Golden(n).has("unit/euler1/prim")]
FIELDS = {"prim": 0, "cons": 1, "RKcons": 2, "tiph": 5, "RKtiph": 6, "dphi": 7, "wiph": 8}
VP = ctypes.c_void_p


class Driver:
    """main.c's set-up through the class API (src/main.c:60-119) on the
    fixture's initial primitives."""

    def __init__(self, g):
        from disco2_amd._lib import check, dptr, lib
        from disco2_amd.api import Setup
        self.L, self.check, self.dptr = lib(), check, dptr
        L = self.L
        self.g = g
        self.s = Setup(g.params)
        assert np.array_equal(self.s.N_p, g.N_p)
        self.sim, self.mpi = self.s.sim, self.s.mpi
        self.gm = L.gravMass_create(2)
        self.cells = L.cell_create(self.sim, self.mpi)
        check()
        self.ts = L.timestep_create(self.sim)
        L.cell_clean_pi(self.cells, self.sim)
        ctypes.CFUNCTYPE(None, VP, VP)(L.gravMass_init_ptr(self.sim))(self.gm, self.sim)
        prim = np.ascontiguousarray(g.prim0)
        L.cell_upload(self.cells, 0, dptr(prim))
        L.gravMass_clean_pi(self.gm, self.sim)
        L.cell_syncproc_r(self.cells, self.sim, self.mpi)
        L.cell_syncproc_z(self.cells, self.sim, self.mpi)
        L.cell_calc_cons(self.cells, self.sim)
        check()

    def first_substep_setup(self):
        """timestep_rk2's lines before its first substep
        (src/TimeStep/timestep_algorithm.c:13-23)"""
        L = self.L
        L.cell_clear_w(self.cells, self.sim)
        L.cell_set_w(self.cells, self.sim)
        L.timestep_set_dt(self.ts, self.cells, self.sim, self.gm)
        L.cell_copy(self.cells, self.sim)
        L.gravMass_copy(self.gm, self.sim)
        L.timestep_set_RK(self.ts, 0.0)

    def field(self, f):
        n = self.s.nzones * (self.s.NQ if f in ("prim", "cons", "RKcons") else 1)
        out = np.empty(n)
        self.L.cell_download(self.cells, FIELDS[f], self.dptr(out))
        return out.reshape(self.s.NQ, -1) if f in ("prim", "cons", "RKcons") else out

    def scalars(self):
        L = self.L
        gm = [[L.gravMass_r(self.gm, p), L.gravMass_phi(self.gm, p), L.gravMass_M(self.gm, p),
               L.gravMass_omega(self.gm, p), L.gravMass_Mdot(self.gm, p), L.gravMass_Macc(self.gm, p)]
              for p in range(2)]
        return np.array([L.timestep_get_t(self.ts), L.timestep_dt(self.ts)]), np.array(gm)

    def close(self):
        L = self.L
        L.cell_destroy(self.cells, self.sim)
        L.timestep_destroy(self.ts)
        L.gravMass_destroy(self.gm)

    def compare(self, name, key, fields=("prim", "cons", "RKcons", "tiph", "RKtiph", "dphi", "wiph")):
        g = self.g
        if g.params.get("BoundTypeR") == 0 or g.params.get("BoundTypeZ") == 0:
            # a fixed boundary's single-cell function also writes its ghost
            # cells' wiph, which cell_clear_w resets before any use; the
            # device keeps one wiph per annulus and skips that dead write
            fields = tuple(f for f in fields if f != "wiph")
        for f in fields:
            k = f"{key}/{f}"
            if not g.has(k):
                continue
            got, ref = self.field(f), g.get(k)
            if name in BITWISE:
                assert bitwise_equal(got, ref), f"{key} {f}: {describe(got, ref)}"
            elif f in ("prim", "cons", "RKcons"):
                assert zone_rel(got, ref) < RTOL, f"{key} {f}: {zone_rel(got, ref)}"
            else:
                assert max_rel(got, ref) < RTOL, f"{key} {f}: {max_rel(got, ref)}"
        t_dt, gm = self.scalars()
        ref_t = g.get(f"{key}/t_dt")
        ref_gm = g.get(f"{key}/gravmass").reshape(2, 12)[:, [0, 1, 2, 3, 8, 9]]
        if name in BITWISE:
            assert bitwise_equal(t_dt, ref_t), f"{key} t, dt {t_dt} {ref_t}"
            assert bitwise_equal(gm, ref_gm), f"{key} masses {gm} {ref_gm}"
        else:
            assert max_rel(t_dt, ref_t) < RTOL
            assert np.max(np.abs(gm - ref_gm)) < 1e-12, f"{key} masses {gm} {ref_gm}"


def replay_substep(d, fac):
    """timestep_substep from its pieces, src/TimeStep/timestep_routines.c:16-113"""
    L, sim, cells, gm, ts, mpi = d.L, d.sim, d.cells, d.gm, d.ts, d.mpi
    dt = fac * L.timestep_dt(ts)
    RK = 0.0 if fac == 1.0 else 0.5
    NR, NZ = L.sim_N(sim, 0), L.sim_N(sim, 1)
    Fr = L.face_create(cells, sim, ts, 0)
    Fz = L.face_create(cells, sim, ts, 1) if L.sim_N_global(sim, 1) > 1 else None
    L.cell_clean_pi(cells, sim)
    L.gravMass_clean_pi(gm, sim)
    L.cell_clear_w(cells, sim)
    L.cell_set_w(cells, sim)
    L.cell_adjust_RK_cons(cells, sim, RK)
    L.cell_plm_p(cells, sim)
    L.cell_plm_rz(cells, sim, Fr, ts, mpi, 0)
    R = L.riemann_create(sim)
    for k in range(NZ):
        for i in range(NR):
            for j in range(L.sim_N_p(sim, i)):
                L.riemann_setup_p(R, cells, sim, i, j, k, 1)
                L.riemann_AddFlux(R, sim, gm, dt)
    for n in range(L.timestep_n(ts, NR - 1, 0)):
        L.riemann_setup_rz(R, Fr, sim, n, 0)
        L.riemann_AddFlux(R, sim, gm, dt)
    if Fz:
        L.cell_plm_rz(cells, sim, Fz, ts, mpi, 1)
        for n in range(L.timestep_n(ts, NZ - 1, 1)):
            L.riemann_setup_rz(R, Fz, sim, n, 2)
            L.riemann_AddFlux(R, sim, gm, dt)
    L.riemann_destroy(R)
    L.cell_add_src(cells, sim, gm, dt)
    if L.sim_EXPLICIT_VISCOSITY(sim) > 0.0:
        L.cell_add_visc_src(cells, sim, gm, dt)
    L.cell_add_split_fictitious(cells, sim, gm, dt)
    L.cell_update_phi(cells, sim, gm, RK, dt)
    L.cell_update_dphi(cells, sim)
    L.gravMass_update_RK(gm, sim, RK)
    L.cell_calc_prim(cells, sim)
    single = L.cell_single_init_ptr(sim) if L.sim_BoundTypeR(sim) == 0 or L.sim_DAMP_TIME(sim) > 0.0 else None
    if L.sim_BoundTypeR(sim) == 1:
        L.cell_boundary_outflow_r(cells, Fr, sim, mpi, ts)
    elif L.sim_BoundTypeR(sim) == 0:
        L.cell_boundary_fixed_r(cells, sim, gm, mpi, single)
    if Fz:
        if L.sim_BoundTypeZ(sim) == 1:
            L.cell_boundary_outflow_z(cells, Fz, sim, mpi, ts)
        elif L.sim_BoundTypeZ(sim) == 0:
            L.cell_boundary_fixed_z(cells, sim, gm, mpi, L.cell_single_init_ptr(sim))
    L.face_destroy(Fr)
    if Fz:
        L.face_destroy(Fz)
    if L.sim_DAMP_TIME(sim) > 0.0:
        L.cell_bc_damp(cells, sim, gm, dt, single)
    L.cell_syncproc_r(cells, sim, mpi)
    L.cell_syncproc_z(cells, sim, mpi)
    L.cell_calc_cons(cells, sim)
    d.check()


@pytest.mark.parametrize("name", UNIT)
def test_timestep_substep_matches_reference(name):
    print(f"case {name}", flush=True)
    d = Driver(Golden(name))
    d.first_substep_setup()
    d.L.timestep_substep(d.ts, d.cells, d.sim, d.gm, d.mpi, 1.0)
    d.check()
    d.compare(name, "unit/sub1")
    d.close()


@pytest.mark.parametrize("name", UNIT)
def test_substep_pieces_match_reference(name):
    print(f"case {name}", flush=True)
    d = Driver(Golden(name))
    d.first_substep_setup()
    replay_substep(d, 1.0)
    d.compare(name, "unit/sub1")
    d.close()


@pytest.mark.parametrize("name", EULER)
def test_forward_euler_matches_reference(name):
    print(f"case {name}", flush=True)
    d = Driver(Golden(name))
    for st in (1, 2):
        d.L.timestep_forward_euler(d.ts, d.sim, d.cells, d.gm, d.mpi)
        d.check()
        d.compare(name, f"unit/euler{st}")
    d.close()


@pytest.mark.parametrize("name", ["vortex_r16", "binary_r16", "disk3d_r8z4"])
def test_mindt_and_conversions(name):
    """cell_mindt on the initial state is the reference's first dt (Time_Max
    is far away); cell_prim2cons of the initial primitives with the
    cell_calc_cons volume is the reference's initial conserved state; after
    cell_calc_prim the unit-level conserved state converts back to within
    round-off."""
    g = Golden(name)
    d = Driver(g)
    L = d.L
    L.cell_set_w(d.cells, d.sim)
    dt = L.cell_mindt(d.cells, d.sim, d.gm)
    ref_dt = g.get("unit/sub1/t_dt")[1]
    if name in BITWISE:
        assert dt == ref_dt
    else:
        assert abs(dt - ref_dt) <= 1e-12 * ref_dt
    cons = d.field("cons")
    assert bitwise_equal(cons, g.get("unit/init/cons")), describe(cons, g.get("unit/init/cons"))
    # the host conversion routines on one cell (cell_conversion.c:11-103)
    NQ = d.s.NQ
    prim = d.field("prim")  # after the z sync of set-up
    pr = np.ascontiguousarray(prim[:, 5])
    u = np.empty(NQ)
    r = 0.5 * (d.s.r_faces[0] + d.s.r_faces[1])
    dV = 0.5 * (d.s.r_faces[1] ** 2 - d.s.r_faces[0] ** 2) * g.dphi0[5] * (d.s.z_faces[1] - d.s.z_faces[0])
    L.cell_prim2cons(d.dptr(pr), d.dptr(u), r, dV, d.sim)
    assert bitwise_equal(u, cons[:, 5])
    back = np.empty(NQ)
    L.cell_cons2prim(d.dptr(u), d.dptr(back), r, dV, d.sim)
    assert np.max(np.abs(back - pr) / (np.abs(pr) + 1e-300)) < 1e-14
    L.cell_calc_prim(d.cells, d.sim)
    d.check()
    assert zone_rel(d.field("prim"), prim) < 1e-14
    d.close()


def test_unit_api_edges():
    """cells of their own, cell_add_cons on the device conserved state, and
    the errors where the reference would crash or exit"""
    from disco2_amd._lib import DiscoError
    g = Golden("binary_r16")
    d = Driver(g)
    L = d.L
    # cell_single_create + the disk's own single-cell plugin on it
    c = L.cell_single_create(d.sim)
    assert L.cell_prim(c, 0) == 0.0 and L.cell_tiph(c) == 0.0
    L.cell_single_destroy(c)
    # cell_add_cons: one zone's conserved density moves by exactly the amount
    before = d.field("cons")
    cell = L.cell_single(d.cells, 5, 3, 0)
    z = L.cell_zone_index(cell)
    L.cell_add_cons(cell, 0, 0.25)
    after = d.field("cons")
    assert after[0, z] == before[0, z] + 0.25
    after[0, z] = before[0, z]
    assert bitwise_equal(after, before)
    L.cell_calc_cons(d.cells, d.sim)  # back to prim2cons of the primitives
    assert bitwise_equal(d.field("cons"), before)
    assert L.cell_Cool(d.cells, 5, 3, 0) == 0.0
    assert L.sim_ZeroPsiBndry(d.sim) == 0
    # riemann_AddFlux before a setup, a foreign plugin for the fixed boundary
    R = L.riemann_create(d.sim)
    L.riemann_AddFlux(R, d.sim, d.gm, 1e-3)
    with pytest.raises(DiscoError):
        d.check()
    L.riemann_destroy(R)
    L.cell_boundary_fixed_r(d.cells, d.sim, d.gm, d.mpi, ctypes.cast(L.cell_init_ptr(d.sim), VP))
    with pytest.raises(DiscoError):
        d.check()
    d.close()
