"""Named problem set-ups: the benchmark workloads and small test variants.

Each entry returns ``(params, ic_kind, ic_kwargs)``.  ``params`` is a full
parameter dictionary (see ``params.py``); ``N_p_round=1`` selects
N_phi = round(2 pi r_c / dr) per annulus instead of the reference's
truncating aspect-ratio formula (src/Sim/sim_set.c:17-29).
"""
from __future__ import annotations

from .params import (make_params, C_RIGID, C_KEPLER, C_FIXED, A_KEPLER, A_MILOS, GRAV_NONE,
                     GRAV_SINGLE, GRAV_BINARY, HLLC, HLL, BOUND_FIXED, BOUND_OUTFLOW,
                     SHEAR, VORTEX, SSTEST, TORUS, MILOS_MACFADYEN, RAD_DOM, MIDDLE)


def vortex(numr=64, r_min=0.5, r_max=5.0, **over):
    """2D isentropic vortex, annuli at the mass-weighted mean phi velocity."""
    p = make_params(NumR=numr, R_Min=r_min, R_Max=r_max, NumZ=1, Adiabatic_Index=1.4,
                    w_numeric_type=C_RIGID, GravMassType=GRAV_NONE, RiemannSolver=HLLC,
                    PLM=1.0, CFL=0.4, N_p_round=1)
    p.update(over)
    return p, "vortex", {"seed": 1}


def binary(numr=512, r_min=0.3, r_max=10.0, q=1.0, alpha=0.01, mach=20.0, **over):
    """2D disk around an equal-mass binary with alpha viscosity, Keplerian annuli."""
    p = make_params(NumR=numr, R_Min=r_min, R_Max=r_max, NumZ=1, Adiabatic_Index=5.0 / 3.0,
                    GravMassType=GRAV_BINARY, MassRatio=q, G_EPS=0.05, Grav_2D=1,
                    Explicit_Viscosity=alpha, ViscConst=0, w_numeric_type=C_KEPLER,
                    RiemannSolver=HLLC, PLM=1.0, CFL=0.4, N_p_round=1)
    p.update(over)
    return p, "binary_disk", {"seed": 1, "mach": mach}


def disk3d(numr=256, numz=64, npcap=1024, **over):
    """3D vertically isothermal Keplerian disk around a single mass, HLL."""
    p = make_params(NumR=numr, R_Min=0.5, R_Max=8.0, NumZ=numz, Z_Min=-0.5, Z_Max=0.5,
                    NPCAP=npcap, Adiabatic_Index=5.0 / 3.0, GravMassType=GRAV_SINGLE,
                    Grav_2D=0, G_EPS=0.0, Explicit_Viscosity=-1.0, w_numeric_type=C_KEPLER,
                    RiemannSolver=HLL, PLM=1.0, CFL=0.4, N_p_round=1)
    p.update(over)
    return p, "disk3d", {"seed": 1}


# Set-ups whose initial state comes from the reference's own initial-data
# plugins (InitialDataType, src/Cell/initial_data/cell_init_*.c), with fixed r
# boundaries and boundary damping where the plugin allows them.  ic_kind
# "plugin": the disk's primitives are set by Batch.initial_data().
def plugin(kind, numr=16, **over):
    base = dict(NumR=numr, NumZ=1, N_p_round=1, CFL=0.4, PLM=1.0)
    if kind == "shear":  # viscous spreading of a Gaussian shear layer at x = 3
        base.update(InitialDataType=SHEAR, R_Min=2.0, R_Max=4.0, Adiabatic_Index=1.4, w_numeric_type=C_FIXED,
                    GravMassType=GRAV_NONE, Explicit_Viscosity=0.05, ViscConst=1, BoundTypeR=BOUND_FIXED)
    elif kind == "vortex":  # the Gresho-like vortex of radius 1 and its passive marker
        base.update(InitialDataType=VORTEX, R_Min=0.2, R_Max=0.95, Adiabatic_Index=1.4, w_numeric_type=C_RIGID,
                    GravMassType=GRAV_NONE, NUM_N=1, BoundTypeR=BOUND_FIXED)
    elif kind == "sstest":
        base.update(InitialDataType=SSTEST, R_Min=0.5, R_Max=8.0, Adiabatic_Index=5.0 / 3.0,
                    w_numeric_type=C_KEPLER, GravMassType=GRAV_BINARY, MassRatio=1.0, G_EPS=0.05,
                    Explicit_Viscosity=0.01, PoRho_r1=0.01, BoundTypeR=BOUND_FIXED, DAMP_TIME=1.0,
                    RDAMP_INNER=1.0, RDAMP_OUTER=6.0)
    elif kind == "middle":
        base.update(InitialDataType=MIDDLE, R_Min=0.5, R_Max=6.0, Adiabatic_Index=5.0 / 3.0,
                    w_numeric_type=C_KEPLER, w_analytic_type=A_MILOS, GravMassType=GRAV_BINARY, MassRatio=0.5,
                    G_EPS=0.05, Explicit_Viscosity=0.01, PoRho_r1=0.01, BoundTypeR=BOUND_OUTFLOW,
                    DAMP_TIME=0.5, RDAMP_INNER=1.0, RDAMP_OUTER=4.5)
    elif kind == "milos":
        base.update(InitialDataType=MILOS_MACFADYEN, R_Min=1.0, R_Max=12.0, Adiabatic_Index=5.0 / 3.0,
                    w_numeric_type=C_KEPLER, GravMassType=GRAV_BINARY, MassRatio=0.5, G_EPS=0.05,
                    Explicit_Viscosity=0.01, PoRho_r1=0.01, Rho_Floor=1e-5, BoundTypeR=BOUND_FIXED,
                    DAMP_TIME=2.0, RDAMP_INNER=2.0, RDAMP_OUTER=10.0)
    elif kind == "rad_dom":
        base.update(InitialDataType=RAD_DOM, R_Min=1.0, R_Max=5.0, Adiabatic_Index=5.0 / 3.0,
                    w_numeric_type=C_KEPLER, GravMassType=GRAV_SINGLE, G_EPS=0.0, Explicit_Viscosity=0.1,
                    BoundTypeR=BOUND_FIXED)
    elif kind == "torus":  # 3D torus around a single mass, periodic in z
        base.update(InitialDataType=TORUS, R_Min=0.5, R_Max=1.5, NumZ=4, Z_Min=-0.3, Z_Max=0.3,
                    Adiabatic_Index=5.0 / 3.0, w_numeric_type=C_KEPLER, GravMassType=GRAV_SINGLE, Grav_2D=0,
                    G_EPS=0.0, RiemannSolver=HLL, BoundTypeR=BOUND_OUTFLOW)
    else:
        raise ValueError(f"unknown plugin set-up {kind!r}")
    base.update(over)
    return make_params(**base), "plugin", {}


# The benchmark workloads (sizes of the measurement definition)
def c1_vortex():
    return vortex(64)


def c2_binary():
    return binary(512)


def c3_disk3d():
    return disk3d(256, 64, 1024)


def c5_adjoint():
    return binary(256, r_min=0.3, r_max=8.0, q=0.5, alpha=0.01)


def c4_batch(ndisk=64, seed=2024, numr_range=(64, 384)):
    """Independent binary disks for one batched launch: NumR uniform in
    numr_range on r in [0.3, 10], mass ratio log-uniform in [0.01, 1], alpha
    log-uniform in [1e-3, 1e-1], Mach uniform in [10, 30]; the density noise
    of disk d is seeded with d."""
    import numpy as np
    rs = np.random.RandomState(seed)
    out = []
    for d in range(ndisk):
        numr = int(rs.randint(numr_range[0], numr_range[1] + 1))
        q = float(10 ** rs.uniform(-2.0, 0.0))
        alpha = float(10 ** rs.uniform(-3.0, -1.0))
        mach = float(rs.uniform(10.0, 30.0))
        p, kind, kw = binary(numr, r_min=0.3, r_max=10.0, q=q, alpha=alpha, mach=mach)
        kw = dict(kw, seed=1, disk=d)
        out.append((p, kind, kw))
    return out


WORKLOADS = {
    "vortex2d_r64": c1_vortex,
    "binary2d_r512": c2_binary,
    "disk3d_r256z64": c3_disk3d,
    "binary2d_r256_adjoint": c5_adjoint,
}
