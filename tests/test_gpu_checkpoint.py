"""Checkpoints and restart (src/IO) on the GPU.

The reference's io_hdf5.c needs parallel HDF5 (io_hdf5.c:37-38
H5Pset_fapl_mpio), which this image lacks, so no file written by the
reference is at hand.  Its io_buffer.c, io_setup.c, io_create_destroy.c and
io_access.c are built into the reference harness, and the golden fixtures
hold io_setbuf's rows, io_setup's schedule and the reference's restart from
those rows (the io/ records): the file's Data, GravMass and T and the
restarted run are checked against them, through the batch API and through
the class API (io_setbuf, io_hdf5_out, io_hdf5_in, io_readbuf, io_setup).
Larger cases are checked against the oracle, whose restart rules
(tests/ckpt_util.py) the CPU tests pin to the same records:

* layout: the file of disco_batch_checkpoint_write after some RK2 steps,
  read back with the HDF5 library, holds T = t, GravMass = (r, phi, M,
  omega) of each mass and Data with one row per cell in the reference's cell
  order (ghosts included): tiph, the annulus centre, the layer centre, the
  primitives, all equal to the oracle's state after the same steps (bitwise
  for the vortex, 1e-11 for the binary disk);
* restart: a fresh disk restarted from the file (src/main.c:95-105: tiph and
  primitives from the file, v_z = 0 in 2D, dphi as cell_create left it, the
  masses' r, phi, M, omega) and stepped on, against the oracle restarted the
  same way: bitwise for the vortex, 1e-11 for the binary;
* the class API in the C driver replaying src/main.c:92-139: it writes
  checkpoint_NNNN.h5 on the reference's schedule, and with Restart 1 it
  continues from input.h5.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import ckpt_util
from compare import bitwise_equal, describe, max_rel, zone_rel
from golden_util import Golden, golden_names

pytestmark = pytest.mark.gpu


def h5lib():
    for name in ("libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            pass
    return None


def read_h5(path, name):
    L = h5lib()
    if L is None:
        pytest.skip("no HDF5 library")
    hid = ctypes.c_int64
    L.H5open()
    L.H5Fopen.restype = hid
    L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
    L.H5Dopen2.restype = hid
    L.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
    L.H5Dget_space.restype = hid
    L.H5Dget_space.argtypes = [hid]
    L.H5Sget_simple_extent_ndims.argtypes = [hid]
    L.H5Sget_simple_extent_dims.argtypes = [hid, ctypes.c_void_p, ctypes.c_void_p]
    L.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
    for f in ("H5Dclose", "H5Sclose", "H5Fclose"):
        getattr(L, f).argtypes = [hid]
    native = hid.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    f = L.H5Fopen(str(path).encode(), 0, 0)
    assert f >= 0
    d = L.H5Dopen2(f, name.encode(), 0)
    assert d >= 0, name
    sp = L.H5Dget_space(d)
    nd = L.H5Sget_simple_extent_ndims(sp)
    dims = (ctypes.c_ulonglong * 2)()
    L.H5Sget_simple_extent_dims(sp, ctypes.byref(dims), None)
    shape = tuple(dims[i] for i in range(nd))
    buf = np.empty(int(np.prod(shape)))
    assert L.H5Dread(d, native, 0, 0, 0, buf.ctypes.data) >= 0
    L.H5Sclose(sp)
    L.H5Dclose(d)
    L.H5Fclose(f)
    return buf.reshape(shape)


def case(kind):
    from disco2_amd import configs
    if kind == "vortex":
        return configs.vortex(24)
    if kind == "disk3d":
        return configs.disk3d(8, 4, 24, R_Min=0.5, R_Max=2.5)
    return configs.binary(32, r_min=0.5, r_max=4.0, q=0.5)


def oracle_after(p, s, prim0, steps):
    from oracle.oracle import Oracle
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=prim0)
    o.init()
    o.step(steps)
    return o


def check(a, b, exact, what):
    """bitwise, or (transcendental physics: device and host libm differ by
    ulps) within 1e-11 per zone (compare.zone_rel) / of 2 pi for angles"""
    from compare import zone_rel
    if exact:
        assert np.array_equal(a, b), what
    elif np.ndim(b) == 2:
        assert zone_rel(a, b) < 1e-11, what
    else:
        assert np.max(np.abs(np.asarray(a) - np.asarray(b))) < 1e-11 * 2 * np.pi * max(1.0, np.max(np.abs(b))), what


def centres(s):
    return ckpt_util.centres(s.N_p, s.r_faces, s.z_faces)


def restart_oracle(p, s, data, T, gm):
    return ckpt_util.restart_oracle(p, s.N_p, s.r_faces, s.z_faces, s.dphi0, data, T, gm)


@pytest.mark.parametrize("kind", ["vortex", "binary", "disk3d"])
def test_batch_checkpoint_and_restart(kind, tmp_path):
    from disco2_amd import ics
    from disco2_amd.api import Batch, Setup
    p, ic, kw = case(kind)
    s = Setup(p)
    prim0 = ics.make_ic(ic, s.geometry(), p, **kw)
    steps, more = 8, 8
    print(f"checkpoint: {kind}, {steps} steps, restart, {more} more", flush=True)
    b = Batch([s])
    b.upload("prim", prim0)
    b.init()
    b.run(steps)
    path = tmp_path / "ck.h5"
    b.checkpoint(path)
    sc = b.scalars()
    b.close()
    exact = kind == "vortex"
    o = oracle_after(p, s, prim0, steps)
    T = read_h5(path, "T")
    gm = read_h5(path, "GravMass")
    data = read_h5(path, "Data")
    assert T.shape == (1,) and gm.shape == (2, 4) and data.shape == (s.nzones, s.NQ + 3)
    assert T[0] == sc[0]
    assert np.array_equal(gm, np.array([sc[3:7], sc[13:17]]))
    r, z = centres(s)
    assert np.array_equal(data[:, 1], r) and np.array_equal(data[:, 2], z)
    check(data[:, 0], o.get("tiph"), exact, "tiph")
    check(data[:, 3:].T, o.get("prim"), exact, "prim")
    check(T, o.scalars()[:1], exact, "t")
    # restart into a fresh disk and step on
    b2 = Batch([Setup(p)])
    b2.upload("prim", prim0)
    b2.init()
    b2.restart(path)
    b2.run(more)
    o2 = restart_oracle(p, s, data, T[0], gm)
    o2.step(more)
    check(b2.download("prim"), o2.get("prim"), exact, "prim after restart")
    check(b2.download("tiph"), o2.get("tiph"), exact, "tiph after restart")
    check(b2.scalars()[:1], o2.scalars()[:1], exact, "t after restart")
    b2.close()


def test_c_driver_checkpoints_and_restarts(tmp_path):
    from disco2_amd import ics
    from disco2_amd.api import Setup
    from disco2_amd.params import write_parfile
    from test_c_driver import build_driver
    p, ic, kw = case("vortex")
    p = dict(p, Time_Max=10.0, Num_Checkpoints=100)  # a checkpoint whenever t passes a multiple of 0.1
    s = Setup(p)
    prim0 = ics.make_ic(ic, s.geometry(), p, **kw)
    exe = build_driver(str(tmp_path))
    run1 = tmp_path / "run1"
    run1.mkdir()
    write_parfile(str(run1 / "run.par"), p)
    np.ascontiguousarray(prim0).tofile(str(run1 / "prim.bin"))
    steps = 10
    print("C driver: vortex, checkpoints, then restart", flush=True)
    r = subprocess.run([exe, "run.par", "prim.bin", str(steps), "out.bin"], capture_output=True, text=True,
                       timeout=120, cwd=str(run1))
    assert r.returncode == 0, r.stderr
    cks = sorted(f for f in os.listdir(run1) if f.startswith("checkpoint_"))
    assert cks and cks == ["checkpoint_%04d.h5" % k for k in range(len(cks))], cks
    last = run1 / cks[-1]
    T = read_h5(last, "T")[0]
    # the checkpoint is the state at the first step past its time: find it on the oracle
    o = oracle_after(p, s, prim0, 0)
    n = 0
    while o.scalars()[0] < T:
        o.step(1)
        n += 1
    assert o.scalars()[0] == T
    data = read_h5(last, "Data")
    assert np.array_equal(data[:, 0], o.get("tiph")) and np.array_equal(data[:, 3:].T, o.get("prim"))
    # restart from it (Restart 1, input.h5) for the remaining steps
    run2 = tmp_path / "run2"
    run2.mkdir()
    write_parfile(str(run2 / "run.par"), dict(p, Restart=1))
    shutil.copy(last, run2 / "input.h5")
    r = subprocess.run([exe, "run.par", "-", str(steps - n), "out.bin"], capture_output=True, text=True,
                       timeout=120, cwd=str(run2))
    assert r.returncode == 0, r.stderr
    out = np.fromfile(str(run2 / "out.bin"))
    nz = s.nzones
    prim, tiph, t = out[:s.NQ * nz].reshape(s.NQ, nz), out[s.NQ * nz:(s.NQ + 1) * nz], out[-2]
    o2 = restart_oracle(p, s, data, T, read_h5(last, "GravMass"))
    o2.step(steps - n)
    assert np.array_equal(prim, o2.get("prim"))
    assert np.array_equal(tiph, o2.get("tiph"))
    assert t == o2.scalars()[0]


# fixtures with the reference's checkpoint rows and restart; vortex cases
# have no libm transcendental and are compared bit for bit
IO_CASES = [n for n in golden_names() if Golden(n).has("io/buffer")]
IO_BITWISE = {"vortex_r16", "vortex_io_schedule"}


def same(name, got, ref, what):
    if name in IO_BITWISE:
        assert bitwise_equal(got, ref), f"{what}: {describe(got, ref)}"
    elif np.ndim(ref) == 2:
        assert zone_rel(got, ref) < 1e-11, f"{what}: {zone_rel(got, ref)}"
    else:
        assert max_rel(got, ref) < 1e-11, f"{what}: {max_rel(got, ref)}"


def check_file(name, g, path):
    """the file's T, GravMass and Data against io_setbuf's rows"""
    T, gm, data = read_h5(path, "T"), read_h5(path, "GravMass"), read_h5(path, "Data")
    ref = g.get("io/buffer").reshape(g.Nz, g.NQ + 3)
    assert data.shape == ref.shape and gm.shape == (2, 4) and T.shape == (1,)
    assert bitwise_equal(data[:, 1:3], ref[:, 1:3])  # the centres: no arithmetic beyond the mesh
    same(name, data[:, 0], ref[:, 0], "tiph")
    same(name, data[:, 3:].T.copy(), ref[:, 3:].T.copy(), "prim")
    same(name, T, g.get("io/state/t_dt")[:1], "T")
    same(name, gm.ravel(), g.get("io/gravmass_buffer"), "GravMass")


def check_state(name, g, key, prim, tiph, dphi, t):
    same(name, prim, g.get(f"{key}/prim"), f"{key} prim")
    same(name, tiph, g.get(f"{key}/tiph"), f"{key} tiph")
    same(name, dphi, g.get(f"{key}/dphi"), f"{key} dphi")
    same(name, np.array([t]), g.get(f"{key}/t_dt")[:1], f"{key} t")


@pytest.mark.parametrize("name", IO_CASES)
def test_batch_checkpoint_matches_reference_rows(name, tmp_path):
    """disco_batch_checkpoint_write after the fixture's steps holds the
    reference's io_setbuf rows; disco_batch_checkpoint_read into a fresh disk
    gives the reference's restarted state, and the steps after it its state"""
    from disco2_amd.api import Batch, Setup
    g = Golden(name)
    n, more = (int(x) for x in g.get("io/steps"))
    print(f"checkpoint rows: {name}, {n} steps, restart, {more} more", flush=True)
    b = Batch([Setup(g.params)])
    b.upload("prim", g.prim0)
    b.init()
    b.run(n)
    path = tmp_path / "ck.h5"
    b.checkpoint(path)
    b.close()
    check_file(name, g, path)
    b2 = Batch([Setup(g.params)])
    b2.upload("prim", g.prim0)
    b2.init()
    b2.restart(path)
    check_state(name, g, "io/restart0", b2.download("prim"), b2.download("tiph"), b2.download("dphi"),
                b2.scalars()[0])
    b2.run(more)
    check_state(name, g, "io/restart", b2.download("prim"), b2.download("tiph"), b2.download("dphi"),
                b2.scalars()[0])
    b2.close()


@pytest.mark.parametrize("name", IO_CASES)
def test_class_api_checkpoint_matches_reference_rows(name, tmp_path, monkeypatch):
    """src/main.c's checkpoint and restart through the class API:
    io_setbuf + io_setup + io_hdf5_out after the fixture's steps, then
    cell_create, io_hdf5_in (input.h5), io_readbuf, gravMass_clean_pi,
    io_setup, the syncs and cell_calc_cons (src/main.c:74-119) and more steps"""
    from test_gpu_class_api import Driver, VP
    g = Golden(name)
    n, more = (int(x) for x in g.get("io/steps"))
    print(f"class API checkpoint: {name}, {n} steps, restart, {more} more", flush=True)
    monkeypatch.chdir(tmp_path)
    d = Driver(g)
    L, sim, mpi = d.L, d.sim, d.mpi
    for _ in range(n):
        L.timestep_rk2(d.ts, sim, d.cells, d.gm, mpi)
    io = L.io_create(sim)
    L.io_allocbuf(io, sim)
    L.io_setbuf(io, d.cells, sim, d.gm)
    L.io_setup(io, sim, d.ts)
    assert L.io_tcheck(io) == g.get("io/tcheck")[0] and L.io_nfile(io) == g.get("io/nfile")[0]
    fname = "checkpoint_%04d.h5" % L.io_nfile(io)
    L.io_hdf5_out(io, sim, d.ts)
    L.io_deallocbuf(io)
    L.io_destroy(io)
    d.check()
    d.close()
    check_file(name, g, tmp_path / fname)
    shutil.copy(tmp_path / fname, tmp_path / "input.h5")
    # the restart: src/main.c:74-119 with Restart 1
    r = Driver.__new__(Driver)
    r.L, r.check, r.dptr, r.g, r.s, r.sim, r.mpi = L, d.check, d.dptr, g, d.s, sim, mpi
    r.gm = L.gravMass_create(2)
    r.cells = L.cell_create(sim, mpi)
    L.cell_clean_pi(r.cells, sim)
    L.cell_syncproc_r(r.cells, sim, mpi)
    L.cell_syncproc_z(r.cells, sim, mpi)
    r.ts = L.timestep_create(sim)
    io = L.io_create(sim)
    L.io_allocbuf(io, sim)
    L.io_hdf5_in(io, sim, r.ts)
    L.io_readbuf(io, r.cells, sim, r.gm)
    L.io_deallocbuf(io)
    L.gravMass_clean_pi(r.gm, sim)
    L.io_setup(io, sim, r.ts)
    assert L.io_tcheck(io) == g.get("io/restart_tcheck")[0] and L.io_nfile(io) == g.get("io/restart_nfile")[0]
    L.io_destroy(io)
    L.cell_syncproc_r(r.cells, sim, mpi)
    L.cell_syncproc_z(r.cells, sim, mpi)
    L.cell_calc_cons(r.cells, sim)
    r.check()
    check_state(name, g, "io/restart0", r.field("prim"), r.field("tiph"), r.field("dphi"), L.timestep_get_t(r.ts))
    for _ in range(more):
        L.timestep_rk2(r.ts, sim, r.cells, r.gm, mpi)
    r.check()
    check_state(name, g, "io/restart", r.field("prim"), r.field("tiph"), r.field("dphi"), L.timestep_get_t(r.ts))
    gm = np.array([[L.gravMass_r(r.gm, p), L.gravMass_phi(r.gm, p), L.gravMass_M(r.gm, p),
                    L.gravMass_omega(r.gm, p)] for p in range(2)])
    same(name, gm.ravel(), g.get("io/restart/gravmass").reshape(2, 12)[:, :4].ravel(), "masses after restart")
    r.close()
