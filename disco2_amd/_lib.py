"""ctypes binding of libdisco2.so (the C ABI declared in include/disco.h).

The library is built in-tree by ``python -m disco2_amd.build`` (or
``__graft_entry__.build()``).  There is no fallback: when the library is
missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdisco2.so")

_lib = None

vp = ctypes.c_void_p
ci = ctypes.c_int
cd = ctypes.c_double
cll = ctypes.c_longlong
ccp = ctypes.c_char_p
pd = ctypes.POINTER(ctypes.c_double)
pi = ctypes.POINTER(ctypes.c_int)

# name: (restype, argtypes)
SIGNATURES = {
    "mpisetup_create": (vp, [ci, vp]),
    "mpisetup_destroy": (None, [vp]),
    "sim_create": (vp, [vp]),
    "sim_destroy": (None, [vp]),
    "sim_read_par_file": (ci, [vp, vp, ccp]),
    "sim_set_param": (ci, [vp, ccp, cd]),
    "sim_alloc_arr": (None, [vp, vp]),
    "sim_set_rz": (None, [vp, vp]),
    "sim_set_N_p": (None, [vp]),
    "sim_set_N_p_array": (None, [vp, pi]),
    "sim_set_misc": (None, [vp, vp]),
    "sim_N": (ci, [vp, ci]),
    "sim_N_p": (ci, [vp, ci]),
    "sim_N_global": (ci, [vp, ci]),
    "sim_FacePos": (cd, [vp, ci, ci]),
    "sim_NUM_Q": (ci, [vp]),
    "sim_get_T_MAX": (cd, [vp]),
    "sim_NumGravMass": (ci, [vp]),
    "sim_Nghost_min": (ci, [vp, ci]),
    "sim_Nghost_max": (ci, [vp, ci]),
    "sim_rOm_a": (cd, [vp, cd, cd]),
    "sim_rdrOm_a": (cd, [vp, cd, cd]),
    "disco_initial_edges": (None, [vp, pd, pd]),
    "gravMass_create": (vp, [ci]),
    "gravMass_destroy": (None, [vp]),
    "gravMass_init_ptr": (vp, [vp]),
    "gravMass_clean_pi": (None, [vp, vp]),
    "gravMass_r": (cd, [vp, ci]),
    "gravMass_phi": (cd, [vp, ci]),
    "gravMass_M": (cd, [vp, ci]),
    "gravMass_omega": (cd, [vp, ci]),
    "gravMass_Mdot": (cd, [vp, ci]),
    "gravMass_Macc": (cd, [vp, ci]),
    "gravMass_move": (None, [vp, vp, cd, cd]),
    "gravMass_copy": (None, [vp, vp]),
    "gravMass_update_RK": (None, [vp, vp, cd]),
    "cell_create": (vp, [vp, vp]),
    "cell_destroy": (None, [vp, vp]),
    "cell_nzones": (cll, [vp]),
    "cell_upload": (None, [vp, ci, pd]),
    "cell_download": (None, [vp, ci, pd]),
    "cell_single": (vp, [vp, ci, ci, ci]),
    "cell_prim": (cd, [vp, ci]),
    "cell_tiph": (cd, [vp]),
    "cell_dphi": (cd, [vp]),
    "cell_wiph": (cd, [vp]),
    "cell_grad": (cd, [vp, ci]),
    "cell_gradp": (cd, [vp, ci]),
    "cell_set_prim": (None, [vp, ci, ci, ci, ci, cd]),
    "cell_set_tiph": (None, [vp, ci, ci, ci, cd]),
    "cell_clean_pi": (None, [vp, vp]),
    "cell_syncproc_r": (None, [vp, vp, vp]),
    "cell_syncproc_z": (None, [vp, vp, vp]),
    "cell_calc_cons": (None, [vp, vp]),
    "cell_plm_p": (None, [vp, vp]),
    "cell_plm_rz": (None, [vp, vp, vp, vp, vp, ci]),
    "cell_set_w": (None, [vp, vp]),
    "cell_mindt": (cd, [vp, vp, vp]),
    "cell_zone_index": (cll, [vp]),
    "face_create": (vp, [vp, vp, vp, ci]),
    "face_destroy": (None, [vp]),
    "face_count": (ci, [vp]),
    "face_L_pointer": (vp, [vp, ci]),
    "face_R_pointer": (vp, [vp, ci]),
    "face_dA": (cd, [vp, ci]),
    "face_cm": (cd, [vp, ci]),
    "face_r": (cd, [vp, ci]),
    "face_deltaL": (cd, [vp, ci]),
    "face_deltaR": (cd, [vp, ci]),
    "timestep_create": (vp, [vp]),
    "timestep_destroy": (None, [vp]),
    "timestep_rk2": (None, [vp, vp, vp, vp, vp]),
    "timestep_forward_euler": (None, [vp, vp, vp, vp, vp]),
    "timestep_substep": (None, [vp, vp, vp, vp, vp, cd]),
    "timestep_get_t": (cd, [vp]),
    "timestep_dt": (cd, [vp]),
    "timestep_set_t": (None, [vp, cd]),
    "timestep_n": (ci, [vp, ci, ci]),
    "disco_batch_create": (vp, [ci, ctypes.POINTER(vp)]),
    "disco_batch_destroy": (None, [vp]),
    "disco_batch_nzones": (cll, [vp, ci]),
    "disco_batch_upload": (None, [vp, ci, ci, pd]),
    "disco_batch_download": (None, [vp, ci, ci, pd]),
    "disco_batch_init": (None, [vp]),
    "disco_batch_run": (None, [vp, ci, ci]),
    "disco_batch_scalars": (None, [vp, ci, pd]),
    "disco_batch_synchronize": (None, [vp]),
    "disco_batch_last_ms": (cd, [vp]),
    "disco_batch_profile": (ci, [vp, ci, ctypes.c_char_p, pd, pi, ci]),
    "disco_calibrate_copy": (cd, [cll]),
    "disco_batch_torque_loss": (ci, [vp, ci, pd]),
    "disco_batch_torque_loss_noalloc": (ci, [vp, ci, pd]),
    "disco_batch_adjoint": (ci, [vp, ci, ci, pd, pd, pd, pd]),
    "disco_batch_diag_sizes": (ci, [vp, ci, pi, pi]),
    "disco_batch_diag_measure": (ci, [vp, ci, cd, pd, pd, pd, pd, pd]),
    "disco_batch_set_scalars": (None, [vp, ci, pd]),
    "disco_batch_diag_equat": (ci, [vp, pd]),
    "disco_batch_initial_data": (ci, [vp, ci]),
    "disco_batch_autotune": (ci, [vp, ci, pd]),
    "disco_batch_set_schedule": (ci, [vp, ci]),
    "disco_batch_schedule": (ci, [vp]),
    "disco_initial_prim": (ci, [vp, pd, pd, pd]),
    "disco_batch_checkpoint_write": (ci, [vp, ci, ccp]),
    "disco_batch_checkpoint_read": (ci, [vp, ci, ccp]),
    "io_create": (vp, [vp]),
    "io_destroy": (None, [vp]),
    "io_allocbuf": (None, [vp, vp]),
    "io_deallocbuf": (None, [vp]),
    "io_setbuf": (None, [vp, vp, vp, vp]),
    "io_readbuf": (None, [vp, vp, vp, vp]),
    "io_hdf5_out": (None, [vp, vp, vp]),
    "io_hdf5_in": (None, [vp, vp, vp]),
    "io_tcheck": (cd, [vp]),
    "io_nfile": (ci, [vp]),
    "io_setup": (None, [vp, vp, vp]),
    "diagnostics_create": (vp, [vp, vp, vp]),
    "diagnostics_destroy": (None, [vp, vp]),
    "diagnostics_set": (None, [vp, vp, vp, vp, vp, vp]),
    "diagnostics_print": (None, [vp, vp, vp, vp]),
    "diagnostics_tdiag_measure": (cd, [vp]),
    "diagnostics_tdiag_dump": (cd, [vp]),
    "cell_single_init_ptr": (vp, [vp]),
    "cell_init_ptr": (vp, [vp]),
    "cell_clear_w": (None, [vp, vp]),
    "cell_calc_prim": (None, [vp, vp]),
    "cell_prim2cons": (None, [pd, pd, cd, cd, vp]),
    "cell_cons2prim": (None, [pd, pd, cd, cd, vp]),
    "cell_Cool": (cd, [vp, ci, ci, ci]),
    "cell_single_create": (vp, [vp]),
    "cell_single_destroy": (None, [vp]),
    "cell_add_cons": (None, [vp, ci, cd]),
    "cell_copy": (None, [vp, vp]),
    "cell_adjust_RK_cons": (None, [vp, vp, cd]),
    "cell_add_src": (None, [vp, vp, vp, cd]),
    "cell_add_visc_src": (None, [vp, vp, vp, cd]),
    "cell_add_visc_src_old": (None, [vp, vp, cd]),
    "cell_add_split_fictitious": (None, [vp, vp, vp, cd]),
    "cell_update_phi": (None, [vp, vp, vp, cd, cd]),
    "cell_update_dphi": (None, [vp, vp]),
    "cell_boundary_outflow_r": (None, [vp, vp, vp, vp, vp]),
    "cell_boundary_outflow_z": (None, [vp, vp, vp, vp, vp]),
    "cell_boundary_fixed_r": (None, [vp, vp, vp, vp, vp]),
    "cell_boundary_fixed_z": (None, [vp, vp, vp, vp, vp]),
    "cell_bc_damp": (None, [vp, vp, vp, cd, vp]),
    "cell_print": (None, [vp, ci, ci, ci]),
    "riemann_create": (vp, [vp]),
    "riemann_destroy": (None, [vp]),
    "riemann_setup_rz": (None, [vp, vp, vp, ci, ci]),
    "riemann_setup_p": (None, [vp, vp, vp, ci, ci, ci, ci]),
    "riemann_AddFlux": (None, [vp, vp, vp, cd]),
    "diagnostics_reset": (None, [vp, vp, vp, vp]),
    "sim_ZeroPsiBndry": (ci, [vp]),
    "sim_NoInnerBC": (ci, [vp]),
    "sim_BoundTypeR": (ci, [vp]),
    "sim_BoundTypeZ": (ci, [vp]),
    "sim_DAMP_TIME": (cd, [vp]),
    "sim_EXPLICIT_VISCOSITY": (cd, [vp]),
    "timestep_set_dt": (None, [vp, vp, vp, vp]),
    "timestep_set_RK": (None, [vp, cd]),
    "timestep_update_t": (None, [vp]),
    "disco_cell_ledger_enable": (ci, [vp, ci]),
    "disco_cell_ledger": (ci, [vp, pd]),
    "disco_last_error": (ci, []),
    "disco_last_error_msg": (ccp, []),
    "disco_clear_error": (None, []),
    "disco_version": (ccp, []),
}


class DiscoError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DiscoError(f"{LIB_PATH} is not built: run python -m disco2_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check():
    """Raise DiscoError if the library recorded an error (and clear it)."""
    L = lib()
    code = L.disco_last_error()
    if code:
        msg = L.disco_last_error_msg().decode()
        L.disco_clear_error()
        raise DiscoError(f"disco error {code}: {msg}")


def dptr(a):
    return a.ctypes.data_as(pd)


def iptr(a):
    return a.ctypes.data_as(pi)
