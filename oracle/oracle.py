"""ctypes driver for the CPU restatement (liboracle.so) and for the reference
build (oracle/_ref/disco_ref) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "disco_ref")

# slot order of enum OI_* / OD_* in disco_oracle.h
I_SLOTS = ["InitialDataType", "GravMassType", "BoundTypeR", "BoundTypeZ", "NoInnerBC",
           "w_numeric_type", "w_analytic_type", "RiemannSolver", "ViscConst", "Grav_2D",
           "RhoSinkOn", "PHI_ORDER", "Cooling", "NumZ", None, None, None, None, "NUM_C"]
D_SLOTS = ["Adiabatic_Index", "CFL", "PLM", "Explicit_Viscosity", "G_EPS", "Rho_Floor",
           "Cs_Floor", "Cs_Cap", "Vel_Cap", "MassRatio", "OrbShrinkTscale", "OrbShrinkT0",
           "Time_Max", "DAMP_TIME", "PoRho_r1", "RDAMP_INNER", "RDAMP_OUTER", "R_Min", "R_Max"]

_lib = None


def build(quiet=True):
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, dp, ip, dp]
        for name in ["oracle_destroy", "oracle_init", "oracle_plm_p", "oracle_set_w"]:
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = None
        for name in ["oracle_set_tiph", "oracle_set_dphi", "oracle_set_prim", "oracle_scalars",
                     "oracle_set_scalars"]:
            getattr(L, name).argtypes = [vp, dp]
            getattr(L, name).restype = None
        L.oracle_nzones.argtypes = [vp]
        L.oracle_nzones.restype = ctypes.c_int
        L.oracle_initial_data.argtypes = [vp]
        L.oracle_initial_data.restype = ctypes.c_int
        L.oracle_rk2.argtypes = [vp]
        L.oracle_rk2.restype = ctypes.c_int
        L.oracle_run.argtypes = [vp, ctypes.c_int]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_forward_euler.argtypes = [vp]
        L.oracle_forward_euler.restype = ctypes.c_int
        L.oracle_step_begin.argtypes = [vp]
        L.oracle_step_begin.restype = ctypes.c_int
        L.oracle_step_substep.argtypes = [vp, ctypes.c_double]
        L.oracle_step_substep.restype = ctypes.c_int
        L.oracle_step_end.argtypes = [vp]
        L.oracle_step_end.restype = None
        L.oracle_ledger.argtypes = [vp, dp]
        L.oracle_ledger.restype = None
        L.oracle_get.argtypes = [vp, ctypes.c_int, dp]
        L.oracle_put.argtypes = [vp, ctypes.c_int, dp]
        L.oracle_face_create.argtypes = [vp, ctypes.c_int]
        L.oracle_face_create.restype = ctypes.c_int
        L.oracle_face_export.argtypes = [vp, ctypes.c_int, ip, dp, ip]
        L.oracle_plm_rz.argtypes = [vp, ctypes.c_int]
        L.oracle_torque.argtypes = [vp]
        L.oracle_torque.restype = ctypes.c_double
        L.oracle_mindt.argtypes = [vp]
        L.oracle_mindt.restype = ctypes.c_double
        L.oracle_last_error.argtypes = [vp]
        L.oracle_last_error.restype = ctypes.c_int
        L.oracle_diag_neq.argtypes = [vp]
        L.oracle_diag_neq.restype = ctypes.c_int
        L.oracle_diag_measure.argtypes = [vp, ctypes.c_double, dp, dp, dp, dp, dp]
        L.oracle_diag_measure.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


FIELDS = {"prim": 0, "cons": 1, "RKcons": 2, "grad": 3, "gradp": 4,
          "tiph": 5, "RKtiph": 6, "dphi": 7, "wiph": 8}


class Oracle:
    """One simulation on the CPU restatement, on a given mesh."""

    def __init__(self, params: dict, N_p, r_faces, z_faces):
        self.params = params
        self.N_p = np.ascontiguousarray(N_p, dtype=np.int32)
        self.r_faces = np.ascontiguousarray(r_faces, dtype=np.float64)
        self.z_faces = np.ascontiguousarray(z_faces, dtype=np.float64)
        self.NR, self.NZ = len(self.N_p), len(self.z_faces) - 1
        self.NQ = params["NUM_C"] + params["NUM_N"]
        ng = params["ng"]
        ivals = []
        for k in I_SLOTS:
            ivals.append(0 if k is None else int(params[k]))
        ivals[14] = 1                       # Nghost_min[R] on the inner-boundary rank
        ivals[15] = ng                      # Nghost_max[R]
        ivals[16] = 0 if params["NumZ"] == 1 else ng
        ivals[17] = 0 if params["NumZ"] == 1 else ng
        self.iparams = np.array(ivals, dtype=np.int32)
        self.dparams = np.array([float(params[k]) for k in D_SLOTS], dtype=np.float64)
        L = lib()
        self.h = L.oracle_create(self.NR, self.NZ, self.NQ, _ip(self.N_p), _dp(self.r_faces),
                                 _dp(self.z_faces), _ip(self.iparams), _dp(self.dparams))
        self.Nz = L.oracle_nzones(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def set_state(self, tiph=None, prim=None, dphi=None):
        L = lib()
        if tiph is not None:
            L.oracle_set_tiph(self.h, _dp(np.ascontiguousarray(tiph, dtype=np.float64)))
        if dphi is not None:
            L.oracle_set_dphi(self.h, _dp(np.ascontiguousarray(dphi, dtype=np.float64)))
        if prim is not None:
            L.oracle_set_prim(self.h, _dp(np.ascontiguousarray(prim, dtype=np.float64)))

    def initial_data(self):
        """the InitialDataType plugin over the current cells (before init())"""
        e = lib().oracle_initial_data(self.h)
        if e:
            raise RuntimeError(f"oracle error {e}")

    def init(self):
        lib().oracle_init(self.h)

    def step(self, n=1):
        e = lib().oracle_run(self.h, n)
        if e:
            raise RuntimeError(f"oracle error {e}")

    def step_begin(self):
        """timestep_rk2's set-up before its first substep"""
        e = lib().oracle_step_begin(self.h)
        if e:
            raise RuntimeError(f"oracle error {e}")

    def substep(self, fac):
        """one substep (fac 1.0, then 0.5) with its gravMass_move"""
        e = lib().oracle_step_substep(self.h, fac)
        if e:
            raise RuntimeError(f"oracle error {e}")

    def step_end(self):
        lib().oracle_step_end(self.h)

    def ledger(self):
        """the last substep's boundary r-face increments (inner, outer: NQ + 3
        each) and torque term"""
        out = np.empty(2 * (self.NQ + 3) + 1)
        lib().oracle_ledger(self.h, _dp(out))
        return out

    def euler(self, n=1):
        """n timestep_forward_euler steps"""
        for _ in range(n):
            e = lib().oracle_forward_euler(self.h)
            if e:
                raise RuntimeError(f"oracle error {e}")

    def get(self, field):
        n = self.Nz * (self.NQ if FIELDS[field] <= 4 else 1)
        out = np.empty(n)
        lib().oracle_get(self.h, FIELDS[field], _dp(out))
        return out.reshape(self.NQ, self.Nz) if FIELDS[field] <= 4 else out

    def put(self, field, arr):
        lib().oracle_put(self.h, FIELDS[field], _dp(np.ascontiguousarray(arr, dtype=np.float64)))

    def torque(self):
        """sum of dt*dV*rho*F_phi*r over interior cells and substeps so far"""
        return lib().oracle_torque(self.h)

    def scalars(self):
        out = np.empty(23)
        lib().oracle_scalars(self.h, _dp(out))
        return out

    def set_scalars(self, v):
        """t, dt, RK, then 2 x {r, phi, M, omega, RK_r, RK_phi, RK_M, RK_omega, Mdot, Macc}"""
        lib().oracle_set_scalars(self.h, _dp(np.ascontiguousarray(v, dtype=np.float64)))

    def faces(self, direction):
        L = lib()
        nf = L.oracle_face_create(self.h, direction)
        LR = np.empty(2 * max(nf, 1), dtype=np.int32)
        geo = np.empty(6 * max(nf, 1))
        offs = np.empty(self.NR if direction == 0 else self.NZ, dtype=np.int32)
        L.oracle_face_export(self.h, direction, _ip(LR), _dp(geo), _ip(offs))
        return LR[:2 * nf].reshape(-1, 2), geo[:6 * nf].reshape(-1, 6), offs

    def plm_p(self):
        lib().oracle_plm_p(self.h)

    def plm_rz(self, direction):
        lib().oracle_plm_rz(self.h, direction)

    def mindt(self):
        return lib().oracle_mindt(self.h)

    def diag_measure(self, dtout, rz):
        """one measurement of diagnostics_set on the current state:
        (vector (NRg, 17), scalar (16,), equat (Neq, 18), mdot (12,))"""
        L = lib()
        nrg = int(self.params["NumR"])  # sim_N_global(R_DIR) on one rank
        neq = L.oracle_diag_neq(self.h)
        vec, sc, eq, md = np.zeros(nrg * 17), np.zeros(16), np.zeros(max(neq, 1) * 18), np.zeros(12)
        L.oracle_diag_measure(self.h, float(dtout), _dp(np.ascontiguousarray(rz, dtype=np.float64)), _dp(vec),
                              _dp(sc), _dp(eq), _dp(md))
        return vec.reshape(nrg, 17), sc, eq[:neq * 18].reshape(neq, 18), md


class Diagnostics:
    """The diagnostics' schedule, running sums and text files
    (src/Diagnostics/diagnostics_create_destroy.c:9-118, diagnostics_set.c,
    diagnostics_print.c:11-92), one rank, around anything with a
    diag_measure(dtout, rz) method: the oracle here, the GPU batch in tests.
    The HDF5 equatorial dump is not written by this restatement."""

    NUM_DIAG = 16

    def __init__(self, source, params, t):
        self.src = source
        T = float(params["Time_Max"])
        self.dtm = T / int(params["Num_Diag_Measure"])
        self.dtd = T / int(params["Num_Diag_Dump"])
        self.tm, self.td = self.dtm, self.dtd
        while self.tm < t:
            self.tm += self.dtm
        while self.td < t:
            self.td += self.dtd
        self.toutprev = self.toutprev_dump = t
        self.rz = np.array([params["R_Min"], params["R_Max"], params["Z_Min"], params["Z_Max"]], dtype=np.float64)
        self.vec = None
        self.scal = np.zeros(self.NUM_DIAG)
        self.equat = None

    def set(self, t, outdir=None):
        """diagnostics_set: True when it measured"""
        if not t > self.tm:
            return False
        v, sc, eq, md = self.src.diag_measure(t - self.toutprev, self.rz)
        if outdir is not None:
            with open(os.path.join(outdir, "DiagMdot.dat"), "a") as f:
                f.write(" ".join("%e" % x for x in (t, *md)) + "\n")
        self.vec = v.copy() if self.vec is None else self.vec + v
        self.scal = self.scal + sc
        self.equat = eq
        self.toutprev = t
        self.tm += self.dtm
        return True

    def print(self, t, outdir=None):
        """diagnostics_print: True when it dumped"""
        if not t > self.td:
            return False
        dt_dump = t - self.toutprev_dump
        if outdir is not None:
            vec = self.vec if self.vec is not None else np.zeros((0, 17))
            with open(os.path.join(outdir, "DiagVector_%010.2f.dat" % t), "w") as f:
                for row in vec:
                    f.write("".join("%e " % (x / dt_dump) for x in row) + "\n")
            with open(os.path.join(outdir, "DiagScalar.dat"), "a") as f:
                f.write("%e " % t + "".join("%e " % (x / dt_dump) for x in self.scal) + "\n")
        self.toutprev_dump = t
        if self.vec is not None:
            self.vec = np.zeros_like(self.vec)
        self.scal = np.zeros(self.NUM_DIAG)
        self.td += self.dtd
        return True


# ---------------------------------------------------------------- reference build
def read_records(path) -> dict:
    b = open(path, "rb").read()
    o, d = 0, {}
    while o < len(b):
        nl = struct.unpack_from("<i", b, o)[0]
        o += 4
        nm = b[o:o + nl].decode()
        o += nl
        ty, cnt = struct.unpack_from("<iq", b, o)
        o += 12
        dt, sz = (np.float64, 8) if ty == 0 else (np.int32, 4)
        d[nm] = np.frombuffer(b, dt, cnt, o).copy()
        o += cnt * sz
    return d


def write_icfile(path, NR, NZ, NQ, N_p=None, tiph=None, prim=None, plugin=False):
    """plugin: the reference then runs its InitialDataType plugin over the
    cells (after any tiph / prim given here)"""
    flags = (1 if N_p is not None else 0) | (2 if tiph is not None else 0) | (4 if prim is not None else 0)
    flags |= 8 if plugin else 0
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", 0x44534331, NR, NZ, NQ, flags))
        if N_p is not None:
            f.write(np.ascontiguousarray(N_p, dtype="<i4").tobytes())
        if tiph is not None:
            f.write(np.ascontiguousarray(tiph, dtype="<f8").tobytes())
        if prim is not None:
            f.write(np.ascontiguousarray(prim, dtype="<f8").tobytes())


def have_reference() -> bool:
    return os.path.exists(REF_BIN)


def run_reference(args, cwd, timeout=600):
    r = subprocess.run([REF_BIN] + list(args), cwd=cwd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"reference run failed ({r.returncode}): {r.stderr[-2000:]}")
    return r
