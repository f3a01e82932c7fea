"""Checkpoints and restart (src/IO) on the GPU.

The reference's IO needs parallel HDF5 (io_hdf5.c:37-38 H5Pset_fapl_mpio),
which this image lacks, so no file written by the reference is at hand: the
file layout is the one io_setbuf / io_readbuf / io_hdf5_out / io_hdf5_in
define, and the tests pin it by content, against the oracle (the CPU
restatement that is bitwise the reference):

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
    rc = 0.5 * (s.r_faces[:-1] + s.r_faces[1:])
    zc = 0.5 * (s.z_faces[:-1] + s.z_faces[1:])
    r = np.concatenate([np.full(n, rc[i]) for i, n in enumerate(s.N_p)])
    return np.tile(r, s.NZ), np.repeat(zc, int(s.N_p.sum()))


def restart_oracle(p, s, data, T, gm):
    """the reference's restart on the oracle: tiph and primitives from the
    file (v_z = 0 in 2D), dphi from cell_create, t and the masses"""
    from oracle.oracle import Oracle
    prim = data[:, 3:].T.copy()
    if s.NZ == 1:
        prim[4] = 0.0
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=data[:, 0].copy(), prim=prim, dphi=s.dphi0)
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


@pytest.mark.parametrize("kind", ["vortex", "binary"])
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
