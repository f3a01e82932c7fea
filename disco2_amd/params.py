"""Run parameters of a Disco2 simulation and their parfile text form.

The key set, spelling and types are those read by the reference's
``sim_read_par_file`` (src/Sim/sim_readpar.c:55-106): a whitespace
separated ``key value`` file where the first match wins, trailing text is
ignored and unknown keys are ignored.  All 52 keys are required by the
reference reader.  Two extra keys understood only by this library
(``N_p_round`` and ``Seed``) are ignored by the reference reader.
"""
from __future__ import annotations

import re

# (key, type) in the reference's read order, src/Sim/sim_readpar.c:55-106
PARAM_KEYS = [
    ("Restart", int), ("InitialDataType", int), ("GravMassType", int),
    ("MassRatio", float), ("OrbShrinkTscale", float), ("OrbShrinkT0", float),
    ("BoundTypeR", int), ("BoundTypeZ", int), ("NoInnerBC", int),
    ("NumR", int), ("NumZ", int), ("ng", int),
    ("R_Min", float), ("R_Max", float), ("Z_Min", float), ("Z_Max", float),
    ("NP_CONST", int), ("NPCAP", int), ("aspect", float),
    ("NUM_C", int), ("NUM_N", int), ("Time_Max", float),
    ("Num_Checkpoints", int), ("Num_Diag_Dump", int), ("Num_Diag_Measure", int),
    ("w_numeric_type", int), ("RiemannSolver", int), ("Adiabatic_Index", float),
    ("Cooling", int), ("Explicit_Viscosity", float), ("ViscConst", int),
    ("PoRho_r1", float), ("CFL", float), ("PLM", float), ("Grav_2D", int),
    ("G_EPS", float), ("RhoSinkOn", int), ("PHI_ORDER", int),
    ("Rho_Floor", float), ("Cs_Floor", float), ("Cs_Cap", float), ("Vel_Cap", float),
    ("DAMP_TIME", float), ("RDAMP_INNER", float), ("RDAMP_OUTER", float),
    ("RLogScale", float), ("ZLogScale", float), ("HiResSigma", float),
    ("HiResR0a", float), ("HiResR0b", float), ("HiResFac", float),
    ("w_analytic_type", int),
]
KEY_TYPES = dict(PARAM_KEYS)
EXTRA_KEYS = {"N_p_round": int}

# enums of src/Headers/header.h:1-13
C_FIXED, C_WCELL, C_RIGID, C_KEPLER, C_OMEGA20, C_MILOS, C_POWERLAW = range(7)
A_FIXED, A_KEPLER, A_OMEGA20, A_MILOS, A_POWERLAW = range(5)
BOUND_FIXED, BOUND_OUTFLOW, BOUND_PERIODIC = range(3)
# InitialDataType (src/Headers/header.h:8)
SHEAR, VORTEX, SSTEST, TORUS, MILOS_MACFADYEN, RAD_DOM, MIDDLE = range(7)
GRAV_NONE, GRAV_SINGLE, GRAV_BINARY = range(3)
HLLC, HLL = 0, 1

DEFAULTS = {
    "Restart": 0, "InitialDataType": VORTEX, "GravMassType": GRAV_NONE,
    "MassRatio": 1.0, "OrbShrinkTscale": 1e19, "OrbShrinkT0": 0.0,
    "BoundTypeR": BOUND_OUTFLOW, "BoundTypeZ": BOUND_PERIODIC, "NoInnerBC": 0,
    "NumR": 64, "NumZ": 1, "ng": 2,
    "R_Min": 0.5, "R_Max": 5.0, "Z_Min": -0.5, "Z_Max": 0.5,
    "NP_CONST": -1, "NPCAP": 100000, "aspect": 1.0,
    "NUM_C": 5, "NUM_N": 0, "Time_Max": 1e30,
    "Num_Checkpoints": 1, "Num_Diag_Dump": 1, "Num_Diag_Measure": 1,
    "w_numeric_type": C_RIGID, "RiemannSolver": HLLC, "Adiabatic_Index": 1.4,
    "Cooling": 0, "Explicit_Viscosity": -1.0, "ViscConst": 0,
    "PoRho_r1": 1.0, "CFL": 0.4, "PLM": 1.0, "Grav_2D": 1,
    "G_EPS": 0.0, "RhoSinkOn": 0, "PHI_ORDER": 2,
    "Rho_Floor": 1e-5, "Cs_Floor": 1e-4, "Cs_Cap": 100.0, "Vel_Cap": 100.0,
    "DAMP_TIME": -1.0, "RDAMP_INNER": -0.1, "RDAMP_OUTER": 10.8,
    "RLogScale": 1e6, "ZLogScale": 1e6, "HiResSigma": 1e6,
    "HiResR0a": 1e6, "HiResR0b": 1e6, "HiResFac": 0.0,
    "w_analytic_type": A_FIXED,
    "N_p_round": 0,
}


def make_params(**overrides) -> dict:
    p = dict(DEFAULTS)
    for k, v in overrides.items():
        if k not in KEY_TYPES and k not in EXTRA_KEYS:
            raise KeyError(f"unknown parameter {k!r}")
        p[k] = v
    return p


def _fmt(v) -> str:
    if isinstance(v, float):
        return repr(float(v))
    return str(int(v))


def parfile_text(params: dict) -> str:
    lines = ["// Disco parameter file"]
    for k, _ in PARAM_KEYS:
        lines.append(f"{k} {_fmt(params[k])}")
    for k in EXTRA_KEYS:
        if k in params:
            lines.append(f"{k} {_fmt(params[k])}")
    return "\n".join(lines) + "\n"


def write_parfile(path, params: dict) -> None:
    with open(path, "w") as f:
        f.write(parfile_text(params))


_NUMBER = re.compile(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)")


def read_parfile(path) -> dict:
    """The parameters of a parfile, read with the reference's rule (readvar,
    src/Sim/sim_readpar.c:10-46): the first line whose first word is the key;
    the value follows after any of the characters "\\t :=>_"; integers are
    the truncated double.  A missing key raises KeyError (the reference's
    reader reports it and fails)."""
    with open(path) as f:
        lines = f.readlines()
    out = {}
    for key, typ in list(PARAM_KEYS) + list(EXTRA_KEYS.items()):
        line = next((s for s in lines if s.split() and s.split()[0] == key), None)
        if line is None:
            if key in EXTRA_KEYS:
                continue
            raise KeyError(f"cant find {key}")
        rest = line[len(key):].lstrip("\t :=>_")
        m = _NUMBER.match(rest)
        if not m:
            raise ValueError(f"{key}: no number in {line.strip()!r}")
        v = float(m.group(1))
        out[key] = int(v) if typ is int else v
    return out
