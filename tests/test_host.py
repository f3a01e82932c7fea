"""CPU tests of the product's host side: the C ABI loads and exports every
declared entry point, and the mesh set-up (parameter reading, bisection face
radii, N_phi, phi offsets) is bitwise equal to the reference's."""
import ctypes
import ctypes.util
import os
import re

import numpy as np
import pytest

from compare import bitwise_equal
from golden_util import Golden, golden_names

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "disco.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"\b([a-zA-Z_][a-zA-Z0-9_]*)\s*\(", text):
        n = m.group(1)
        if n in ("if", "void", "int", "double", "struct", "return"):
            continue
        names.add(n)
    # function-pointer return type: void (*gravMass_init_ptr(struct Sim *))(...)
    names.discard("Sim")
    return sorted(n for n in names if not n.startswith("__"))


def test_library_exports_every_declared_symbol():
    from disco2_amd._lib import lib
    L = lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, f"missing exports: {missing}"
    assert len(declared_symbols()) > 150


def test_every_reference_class_function_is_declared_and_exported():
    """Every function the reference's class headers declare (src/Headers/*.h,
    listed in tests/golden/reference_api.json) is declared in include/disco.h
    and exported by the library: the drop-in boundary is the whole class API."""
    import json
    from disco2_amd._lib import lib
    api = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_api.json")))
    assert len(api) == 199
    declared = set(declared_symbols())
    L = lib()
    missing = [f"{h}:{n} {name}" for h, n, name in api if name not in declared or not hasattr(L, name)]
    assert not missing, missing


@pytest.mark.parametrize("name", golden_names())
def test_mesh_matches_reference(name):
    from disco2_amd.api import Setup
    g = Golden(name)
    s = Setup(g.params)
    assert bitwise_equal(s.r_faces, g.r_faces)
    assert bitwise_equal(s.z_faces, g.z_faces)
    assert np.array_equal(s.N_p, g.N_p)
    assert bitwise_equal(s.tiph0, g.tiph0)
    assert bitwise_equal(s.dphi0, g.dphi0)
    p2 = dict(g.params)
    p2["N_p_round"] = 0
    assert np.array_equal(Setup(p2).N_p, g.get("N_p_ref_formula"))


def test_phi_offsets_follow_the_c_library_generator():
    """The first offsets equal 2 pi rand()/RAND_MAX after srand(666) of this
    host's C library (the reference's cell_create draws)."""
    from disco2_amd.api import Setup
    from disco2_amd import configs
    p, _, _ = configs.vortex(8, r_min=0.5, r_max=1.5)
    s = Setup(p)
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(666)
    off = s.tiph0[np.concatenate([[0], np.cumsum(s.N_p)[:-1]])]
    for i in range(s.NR):
        expect = 2.0 * np.pi * libc.rand() / 2147483647.0
        while expect > 2.0 * np.pi:
            expect -= 2.0 * np.pi
        assert off[i] == expect


def test_parfile_reader_matches_parameters(tmp_path):
    from disco2_amd import configs
    from disco2_amd._lib import check, lib
    from disco2_amd.params import write_parfile
    p, _, _ = configs.binary(32, r_min=0.3, r_max=3.0, q=0.37, alpha=0.02)
    path = tmp_path / "b.par"
    write_parfile(path, p)
    L = lib()
    mpi = L.mpisetup_create(0, None)
    sim = L.sim_create(mpi)
    assert L.sim_read_par_file(sim, mpi, str(path).encode()) == 0
    check()
    L.sim_alloc_arr(sim, mpi)
    L.sim_set_rz(sim, mpi)
    L.sim_set_N_p(sim)
    L.sim_set_misc(sim, mpi)
    assert L.sim_N(sim, 0) == 32 + 3
    assert L.sim_N_global(sim, 1) == 1
    assert L.sim_NUM_Q(sim) == 5
    L.sim_destroy(sim)
    # a missing key is an error, as in the reference reader
    text = path.read_text().replace("CFL ", "XCFL ")
    path.write_text(text)
    sim = L.sim_create(mpi)
    assert L.sim_read_par_file(sim, mpi, str(path).encode()) == 1
    with pytest.raises(Exception, match="CFL"):
        check()
    L.sim_destroy(sim)
    L.mpisetup_destroy(mpi)


def test_frame_rotation_closed_forms():
    from disco2_amd.api import Setup
    from disco2_amd import configs
    from disco2_amd._lib import lib
    p, _, _ = configs.binary(8, r_min=0.3, r_max=1.0, w_analytic_type=1)
    s = Setup(p)
    L = lib()
    for r in (0.3, 1.0, 4.0):
        assert abs(L.sim_rOm_a(s.sim, r, 1.0) - r ** -0.5) < 1e-15 * r ** -0.5
        assert abs(L.sim_rdrOm_a(s.sim, r, 1.0) + 1.5 * r ** -1.5) < 1e-15 * r ** -1.5


def test_gravmass_host_orbit():
    """gravMass_move keeps the circular orbit closed form phi(t) = phi0 + t
    for a = 1, M = 1 (src/GravMass/GravMass_misc.c:31-64)."""
    from disco2_amd.api import Setup
    from disco2_amd import configs
    from disco2_amd._lib import lib
    p, _, _ = configs.binary(8, r_min=0.3, r_max=1.0, q=0.25)
    s = Setup(p)
    L = lib()
    gm = L.gravMass_create(2)
    fn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)(L.gravMass_init_ptr(s.sim))
    fn(gm, s.sim)
    assert abs(L.gravMass_M(gm, 0) - 0.8) < 1e-15 and abs(L.gravMass_r(gm, 0) - 0.2) < 1e-15
    for _ in range(10):
        L.gravMass_move(s.sim, gm, 0.0, 0.1)
    assert abs(L.gravMass_phi(gm, 0) - 1.0) < 1e-14
    assert abs(L.gravMass_phi(gm, 1) - (np.pi + 1.0)) < 1e-14
    L.gravMass_destroy(gm)


def test_negative_ghost_radius_is_rejected():
    """NumR 32 on r in [0.3, 10]: the inner ghost face would sit at r < 0 and
    the reference's N_p formula gives -3 for that annulus; set-up must fail
    with an error instead of sizing arrays with it."""
    from disco2_amd import configs
    from disco2_amd._lib import DiscoError
    from disco2_amd.api import Setup
    p, _, _ = configs.binary(32)
    with pytest.raises(DiscoError, match="N_p"):
        Setup(p)
    p, _, _ = configs.binary(48)  # dr = 0.2 < R_Min: fine
    s = Setup(p)
    assert s.N_p.min() >= 2
