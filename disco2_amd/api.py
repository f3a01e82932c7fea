"""Python front end of the C ABI: simulation set-up and batched runs.

``Setup`` drives the reference's set-up sequence (src/main.c:60-71) through
the class API; ``Batch`` holds one or many disks on the GPU and advances them
with RK2 steps.  Field arrays use the reference's zone order (layer k,
annulus i, cell j) and are q-major for the NQ-vectors.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, dptr, iptr, lib

FIELDS = {"prim": 0, "cons": 1, "RKcons": 2, "grad": 3, "gradp": 4,
          "tiph": 5, "RKtiph": 6, "dphi": 7, "wiph": 8}
VECTOR_FIELDS = {"prim", "cons", "RKcons", "grad", "gradp"}


class Setup:
    """A Sim built from a parameter dictionary (parfile keys)."""

    def __init__(self, params: dict, N_p=None):
        L = lib()
        self.params = dict(params)
        self.mpi = L.mpisetup_create(0, None)
        self.sim = L.sim_create(self.mpi)
        for k, v in params.items():
            if L.sim_set_param(self.sim, k.encode(), float(v)) != 0:
                raise KeyError(f"unknown parameter {k}")
        L.sim_alloc_arr(self.sim, self.mpi)
        L.sim_set_rz(self.sim, self.mpi)
        L.sim_set_N_p(self.sim)
        if N_p is not None:
            arr = np.ascontiguousarray(N_p, dtype=np.int32)
            L.sim_set_N_p_array(self.sim, iptr(arr))
        L.sim_set_misc(self.sim, self.mpi)
        check()
        self.NR = L.sim_N(self.sim, 0)
        self.NZ = L.sim_N(self.sim, 1)
        self.NQ = L.sim_NUM_Q(self.sim)
        self.N_p = np.array([L.sim_N_p(self.sim, i) for i in range(self.NR)], dtype=np.int32)
        self.r_faces = np.array([L.sim_FacePos(self.sim, i - 1, 0) for i in range(self.NR + 1)])
        self.z_faces = np.array([L.sim_FacePos(self.sim, k - 1, 1) for k in range(self.NZ + 1)])
        self.nzones = int(self.N_p.sum()) * self.NZ
        self.tiph0 = np.empty(self.nzones)
        self.dphi0 = np.empty(self.nzones)
        L.disco_initial_edges(self.sim, dptr(self.tiph0), dptr(self.dphi0))

    def geometry(self) -> dict:
        return {"N_p": self.N_p, "r_faces": self.r_faces, "z_faces": self.z_faces,
                "tiph": self.tiph0, "dphi": self.dphi0}

    def initial_prim(self, prim=None):
        """The primitives the disk's InitialDataType plugin sets (on the
        host), over prim (default: zeros, as cell_create leaves them)."""
        out = np.zeros((self.NQ, self.nzones)) if prim is None else np.array(prim, dtype=np.float64, order="C")
        rc = lib().disco_initial_prim(self.sim, dptr(self.tiph0), dptr(self.dphi0), dptr(out))
        check()
        if rc != 0:
            raise RuntimeError("initial data failed")
        return out

    def interior_zones(self) -> int:
        L = lib()
        i0, i1 = L.sim_Nghost_min(self.sim, 0), self.NR - L.sim_Nghost_max(self.sim, 0)
        nz = L.sim_N_global(self.sim, 1)
        return int(self.N_p[i0:i1].sum()) * nz

    def close(self):
        L = lib()
        if getattr(self, "sim", None):
            L.sim_destroy(self.sim)
            L.mpisetup_destroy(self.mpi)
            self.sim = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """One or many independent disks resident on the GPU."""

    def __init__(self, setups):
        self.setups = list(setups)
        L = lib()
        arr = (ctypes.c_void_p * len(self.setups))(*[s.sim for s in self.setups])
        self.h = L.disco_batch_create(len(self.setups), arr)
        check()
        if not self.h:
            raise RuntimeError("batch creation failed")
        self.NQ = self.setups[0].NQ

    def nzones(self, disk=0):
        return self.setups[disk].nzones

    def upload(self, field, arr, disk=0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        lib().disco_batch_upload(self.h, disk, FIELDS[field], dptr(a))
        check()

    def download(self, field, disk=0):
        n = self.setups[disk].nzones
        out = np.empty(n * (self.NQ if field in VECTOR_FIELDS else 1))
        lib().disco_batch_download(self.h, disk, FIELDS[field], dptr(out))
        check()
        return out.reshape(self.NQ, n) if field in VECTOR_FIELDS else out

    def init(self):
        lib().disco_batch_init(self.h)
        check()

    def run(self, nsteps, graph=True, sync=True):
        lib().disco_batch_run(self.h, int(nsteps), 1 if graph else 0)
        if sync:
            self.synchronize()

    def synchronize(self):
        lib().disco_batch_synchronize(self.h)
        check()

    def last_ms(self):
        return lib().disco_batch_last_ms(self.h)

    def profile(self, nsteps):
        """Per-kernel (total ms, launches) over nsteps eager steps."""
        maxk = 64
        names = ctypes.create_string_buffer(32 * maxk)
        ms = np.zeros(maxk)
        cnt = np.zeros(maxk, dtype=np.int32)
        nk = lib().disco_batch_profile(self.h, int(nsteps), names, dptr(ms), iptr(cnt), maxk)
        check()
        raw = names.raw
        return {raw[32 * k:32 * k + 32].split(b"\0")[0].decode(): (float(ms[k]), int(cnt[k])) for k in range(nk)}

    def scalars(self, disk=0):
        out = np.empty(23)
        lib().disco_batch_scalars(self.h, disk, dptr(out))
        check()
        return out

    def set_scalars(self, values, disk=0):
        """t, dt, RK, then 2 x {r, phi, M, omega, RK_r, RK_phi, RK_M, RK_omega,
        Mdot, Macc} (the layout of scalars())"""
        lib().disco_batch_set_scalars(self.h, disk, dptr(np.ascontiguousarray(values, dtype=np.float64)))
        check()

    def diag_measure(self, dtout, rz, disk=0, equat=True):
        """One measurement of the reference's diagnostics_set on the GPU:
        (vector (NumR, 17), scalar (16,), equat (Neq, 18), mdot (12,)) as
        the reference adds them to its time averages (dtout = t - previous
        measurement; rz = R_Min, R_Max, Z_Min, Z_Max).  equat=False leaves
        the equatorial cells on the GPU (returns None for them)."""
        L = lib()
        nrg, neq = ctypes.c_int(), ctypes.c_int()
        L.disco_batch_diag_sizes(self.h, disk, ctypes.byref(nrg), ctypes.byref(neq))
        check()
        vec, sc = np.zeros(nrg.value * 17), np.zeros(16)
        eq, md = np.zeros(max(neq.value, 1) * 18), np.zeros(12)
        rc = L.disco_batch_diag_measure(self.h, disk, float(dtout), dptr(np.ascontiguousarray(rz, dtype=np.float64)),
                                        dptr(vec), dptr(sc), dptr(eq) if equat else None, dptr(md))
        check()
        if rc != 0:
            raise RuntimeError("diagnostics failed")
        return vec.reshape(nrg.value, 17), sc, eq[:neq.value * 18].reshape(neq.value, 18) if equat else None, md

    def autotune(self, nsteps=10):
        """Time every launch schedule on a scratch copy of the batch and keep
        the fastest; returns (chosen schedule, {schedule: ms/step})."""
        ms = np.zeros(1024)  # DISCO_NSCHEDULES
        k = lib().disco_batch_autotune(self.h, int(nsteps), dptr(ms))
        check()
        if k < 0:
            raise RuntimeError("autotune failed")
        return k, {i: float(v) for i, v in enumerate(ms) if v >= 0}

    def set_schedule(self, k):
        lib().disco_batch_set_schedule(self.h, int(k))
        check()

    def schedule(self):
        return lib().disco_batch_schedule(self.h)

    def initial_data(self, disk=0):
        """Set the disk's primitives with its InitialDataType plugin (the
        reference's src/main.c:103-104); run before init()."""
        rc = lib().disco_batch_initial_data(self.h, disk)
        check()
        if rc != 0:
            raise RuntimeError("initial data failed")

    def checkpoint(self, path, disk=0):
        """Write the disk's time, masses and cells in the reference's
        checkpoint layout (HDF5: T, GravMass, Data)."""
        rc = lib().disco_batch_checkpoint_write(self.h, disk, str(path).encode())
        check()
        if rc != 0:
            raise RuntimeError("checkpoint failed")

    def restart(self, path, disk=0):
        """Read a checkpoint into the disk as the reference restarts (after
        init(); dphi is left as it is)."""
        rc = lib().disco_batch_checkpoint_read(self.h, disk, str(path).encode())
        check()
        if rc != 0:
            raise RuntimeError("restart failed")

    def torque_loss(self, nsteps):
        """Run nsteps RK2 steps; the time-averaged gravitational torque on the
        gas (the loss of the adjoint)."""
        out = np.zeros(1)
        rc = lib().disco_batch_torque_loss(self.h, int(nsteps), dptr(out))
        check()
        if rc != 0:
            raise RuntimeError("torque_loss failed")
        return float(out[0])

    def adjoint(self, nsteps, checkpoint_every=20):
        """Run nsteps RK2 steps forward and the adjoint back: returns
        (loss, dL/d alpha, dL/d q, dL/d prim0 as (NQ, Nzones))."""
        n = self.nzones(0)
        loss, ga, gq = np.zeros(1), np.zeros(1), np.zeros(1)
        gp = np.zeros(self.setups[0].NQ * n)
        rc = lib().disco_batch_adjoint(self.h, int(nsteps), int(checkpoint_every), dptr(loss), dptr(ga), dptr(gq),
                                       dptr(gp))
        check()
        if rc != 0:
            raise RuntimeError("adjoint failed")
        return float(loss[0]), float(ga[0]), float(gq[0]), gp.reshape(self.setups[0].NQ, n)

    def close(self):
        if getattr(self, "h", None):
            lib().disco_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_disk(params: dict, ic_kind: str, ic_kwargs: dict, N_p=None):
    """Setup + initial state for one disk: (setup, prim0)."""
    from . import ics
    s = Setup(params, N_p=N_p)
    prim0 = ics.make_ic(ic_kind, s.geometry(), params, **ic_kwargs)
    return s, prim0
