"""Writes the golden fixtures of tests/golden/ from the reference build.

Run in a container where the reference sources exist and oracle/_ref has
been built (``make -C oracle ref``):

    python tests/golden/make_golden.py [case ...]

For every case it writes the parfile, asks the reference for its mesh,
builds the seeded initial state (disco2_amd.ics) or lets the reference's
own initial-data plugin set it (the plugin cases), runs the reference's
unit pieces (faces, PLM slopes, the first substep), a few RK2 steps and,
for some cases, the checkpoint buffer and a restart from it, and
stores inputs and outputs as a compressed .npz.  The fixtures are data only.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from disco2_amd import configs, ics  # noqa: E402
from disco2_amd.geometry import n_p_round  # noqa: E402
from disco2_amd.params import read_parfile, write_parfile  # noqa: E402
from oracle import oracle as orc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF_PARFILES = "/root/reference/parfiles"


def cases():
    c = {}
    c["vortex_r16"] = (configs.vortex(16), dict(steps=20, dumps=[1, 2, 20], unit=True, euler=2, io=(5, 5)))
    p, k, kw = configs.vortex(12, r_min=0.5, r_max=3.5, NUM_N=1, PLM=2.0, RiemannSolver=1, NoInnerBC=1)
    c["vortex_scalar_hll_noinner"] = ((p, k, dict(kw, x0=1.8)), dict(steps=12, dumps=[1, 12], unit=True))
    p, k, kw = configs.vortex(12, r_min=0.5, r_max=3.5, NUM_N=3, PLM=2.0)
    c["vortex_three_scalars"] = ((p, k, dict(kw, x0=1.8)), dict(steps=8, dumps=[1, 8], unit=True))
    p, k, kw = configs.vortex(12, r_min=0.5, r_max=3.5, NUM_N=2, PLM=1.5)
    c["vortex_two_scalars"] = ((p, k, dict(kw, x0=1.6)), dict(steps=8, dumps=[1, 8], unit=True))
    c["binary_r16"] = (configs.binary(16, r_min=0.3, r_max=3.0), dict(steps=20, dumps=[1, 2, 20], unit=True,
                                                                          euler=2, io=(5, 5)))
    p, k, kw = configs.binary(16, r_min=0.3, r_max=3.0, q=0.5, w_numeric_type=0, w_analytic_type=1,
                              RLogScale=1.0, RhoSinkOn=1, ViscConst=0)
    c["binary_akepler_log_sink"] = ((p, k, kw), dict(steps=12, dumps=[1, 12], unit=True))
    p, k, kw = configs.binary(16, r_min=0.3, r_max=3.0, alpha=0.02, ViscConst=1, w_numeric_type=2)
    c["binary_viscconst_rigid"] = ((p, k, kw), dict(steps=8, dumps=[1, 8], unit=True))
    p, k, kw = configs.disk3d(8, 4, 24, R_Min=0.5, R_Max=2.5)
    c["disk3d_r8z4"] = ((p, k, kw), dict(steps=10, dumps=[1, 2, 10], unit=True, euler=2, io=(3, 3)))
    p, k, kw = configs.disk3d(8, 4, 20, R_Min=0.5, R_Max=2.5, N_p_round=0, HiResFac=1.0,
                              HiResSigma=0.3, HiResR0a=1.0, ZLogScale=0.5)
    c["disk3d_refNp_hires"] = ((p, k, kw), dict(steps=6, dumps=[1, 6], unit=True))
    # the checkpoint schedule of io_setup (a checkpoint every 0.025 in time)
    c["vortex_io_schedule"] = (configs.vortex(12, Time_Max=1.0, Num_Checkpoints=40),
                               dict(steps=1, dumps=[1], unit=False, light=True, io=(6, 3)))
    c["vortex_r64_100"] = (configs.vortex(64), dict(steps=100, dumps=[100], unit=False, light=True))
    # diagnostics (src/Diagnostics): measured every step; the vortex and the
    # 3D disk dump every few measurements (time-averaged profiles), the binary
    # (whose first step is long) every step
    p, k, kw = configs.vortex(12, r_min=0.5, r_max=3.5, NUM_N=1, PLM=2.0, RiemannSolver=1, NoInnerBC=1,
                              Time_Max=10.0, Num_Diag_Measure=400, Num_Diag_Dump=100)
    c["diag_vortex_scalar"] = ((p, k, dict(kw, x0=1.8)), dict(diag=10))
    p, k, kw = configs.binary(16, r_min=0.3, r_max=3.0, Time_Max=1.0, Num_Diag_Measure=2000, Num_Diag_Dump=1200)
    c["diag_binary_r16"] = ((p, k, kw), dict(diag=8))
    p, k, kw = configs.disk3d(8, 4, 24, R_Min=0.5, R_Max=2.5, Time_Max=10.0, Num_Diag_Measure=100,
                              Num_Diag_Dump=40)
    c["diag_disk3d_r8z4"] = ((p, k, kw), dict(diag=8))
    # the reference's own initial-data plugins (src/Cell/initial_data), with
    # fixed r boundaries and boundary damping; prim0 is the plugin's output
    c["plugin_shear_fixed"] = (configs.plugin("shear", 16, NPCAP=64), dict(steps=8, dumps=[1, 8], unit=True, plugin=True))
    c["plugin_vortex_fixed"] = (configs.plugin("vortex", 16, R_Max=0.9),
                                dict(steps=8, dumps=[1, 8], unit=True, plugin=True))
    c["plugin_sstest_fixed_damp"] = (configs.plugin("sstest", 16), dict(steps=8, dumps=[1, 8], unit=True, plugin=True))
    c["plugin_middle_outflow_damp"] = (configs.plugin("middle", 16),
                                       dict(steps=8, dumps=[1, 8], unit=True, plugin=True))
    c["plugin_milos_fixed_damp"] = (configs.plugin("milos", 16), dict(steps=8, dumps=[1, 8], unit=True, plugin=True))
    c["plugin_rad_dom_fixed"] = (configs.plugin("rad_dom", 16), dict(steps=8, dumps=[1, 8], unit=True, plugin=True))
    c["plugin_torus3d"] = (configs.plugin("torus", 10, NPCAP=32), dict(steps=6, dumps=[1, 6], unit=True, plugin=True))
    # (no viscosity: the reference exits once v_z != 0 meets viscosity)
    c["plugin_middle3d_fixed_rz"] = (configs.plugin("middle", 8, R_Min=1.0, R_Max=4.0, NumZ=4, Z_Min=-0.2, Z_Max=0.2,
                                                    NPCAP=48, Grav_2D=0, BoundTypeR=0, BoundTypeZ=0, DAMP_TIME=-1.0,
                                                    Explicit_Viscosity=-1.0),
                                     dict(steps=6, dumps=[1, 6], unit=True, plugin=True))
    # the reference's shipped parfiles (parfiles/), run as they are with their
    # own plugins; large meshes, so only digests of the states are kept
    for nm, pf, steps in (("code_shear", "code_tests/shear.par", 4), ("code_vortex", "code_tests/vortex.par", 10),
                          ("code_SStest", "code_tests/SStest.par", 40), ("code_torus", "torus.par", 3)):
        c[nm] = ((os.path.join(REF_PARFILES, pf), "plugin", {}), dict(steps=steps, digest=True))
    return c


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def make_digest_case(name, parfile, opts, tmp):
    """the reference on one of its own parfiles: SHA-256 of the state arrays
    at the start and after the steps, t, dt and the masses"""
    params = read_parfile(parfile)
    par = os.path.join(tmp, name + ".par")
    write_parfile(par, params)
    g = orc.read_records(_run(["geom", par, "-", os.path.join(tmp, "g.bin")], tmp, "g.bin"))
    NR, NZ, NQ = (int(x) for x in g["dims"][:3])
    ic = os.path.join(tmp, "ic.bin")
    orc.write_icfile(ic, NR, NZ, NQ, plugin=True)
    steps = opts["steps"]
    r = orc.read_records(_run(["run", par, ic, os.path.join(tmp, "r.bin"), str(steps), str(steps)], tmp, "r.bin"))
    if f"step{steps}/prim" not in r:
        raise RuntimeError(f"{name}: the reference stopped early")
    out = {"params": np.array(json.dumps(params)), "N_p": g["N_p"], "r_faces": g["r_faces"], "z_faces": g["z_faces"],
           "dims": g["dims"], "parfile": np.array(os.path.relpath(parfile, os.path.dirname(REF_PARFILES))),
           "digest/tiph0": np.array(sha(g["tiph0"])), "digest/dphi0": np.array(sha(g["dphi0"]))}
    for st in (0, steps):
        for f in ("prim", "cons", "tiph", "dphi", "RKtiph"):
            out[f"digest/step{st}/{f}"] = np.array(sha(r[f"step{st}/{f}"]))
        out[f"digest/step{st}/t_dt"] = r[f"step{st}/t_dt"]
        out[f"digest/step{st}/gravmass"] = r[f"step{st}/gravmass"]
    out["digest/steps"] = np.array([steps], dtype=np.int32)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return path


def make_case(name, spec, opts, tmp):
    if opts.get("digest"):
        return make_digest_case(name, spec[0], opts, tmp)
    params, ic_kind, ic_kw = spec
    par = os.path.join(tmp, name + ".par")
    write_parfile(par, params)
    g0 = orc.read_records(_run(["geom", par, "-", os.path.join(tmp, "g0.bin")], tmp, "g0.bin"))
    NR, NZ, NQ = (int(x) for x in g0["dims"][:3])
    if params.get("N_p_round", 0):
        N_p = n_p_round(g0["r_faces"], params["R_Min"], params["NPCAP"])
    else:
        N_p = g0["N_p"]
    icg = os.path.join(tmp, "icg.bin")
    orc.write_icfile(icg, NR, NZ, NQ, N_p=N_p)
    g = orc.read_records(_run(["geom", par, icg, os.path.join(tmp, "g.bin")], tmp, "g.bin"))
    geom = {"N_p": N_p, "r_faces": g["r_faces"], "z_faces": g["z_faces"],
            "tiph": g["tiph0"], "dphi": g["dphi0"]}
    ic = os.path.join(tmp, "ic.bin")
    if opts.get("plugin"):
        # the reference's InitialDataType plugin sets the cells; its output
        # (the state before the first step) is the fixture's prim0
        orc.write_icfile(ic, NR, NZ, NQ, N_p=N_p, plugin=True)
        r0 = orc.read_records(_run(["run", par, ic, os.path.join(tmp, "r0.bin"), "0"], tmp, "r0.bin"))
        prim0 = r0["step0/prim"].reshape(NQ, -1)
    else:
        prim0 = ics.make_ic(ic_kind, geom, params, **ic_kw)
        orc.write_icfile(ic, NR, NZ, NQ, N_p=N_p, prim=prim0)
    out = {"params": np.array(json.dumps(params)), "N_p": N_p, "N_p_ref_formula": g0["N_p"],
           "r_faces": g["r_faces"], "z_faces": g["z_faces"], "tiph0": g["tiph0"], "dphi0": g["dphi0"],
           "prim0": prim0.ravel(), "dims": g["dims"]}
    if opts.get("diag"):
        # the text files the reference appends to land in a directory of their own
        dd = os.path.join(tmp, name + "_diag")
        os.makedirs(dd, exist_ok=True)
        orc.run_reference(["diag", par, ic, os.path.join(dd, "d.bin"), str(opts["diag"])], dd)
        r = orc.read_records(os.path.join(dd, "d.bin"))
        for k, v in r.items():
            if "/state/" in k and not k.endswith(("/prim", "/tiph", "/dphi", "/gravmass", "/t_dt")):
                continue
            if k.startswith(("diag/", "meas", "dump")):
                out["diag/" + k if not k.startswith("diag/") else k] = v
        for fn in sorted(os.listdir(dd)):
            if fn.endswith(".dat"):
                with open(os.path.join(dd, fn)) as f:
                    out["file/" + fn] = np.array(f.read())
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        return path
    if opts.get("unit"):
        u = orc.read_records(_run(["unit", par, ic, os.path.join(tmp, "u.bin")], tmp, "u.bin"))
        for k, v in u.items():
            if k.startswith(("faces_", "gradp", "grad_", "init/cons", "sub1/")):
                out["unit/" + k] = v
    if opts.get("euler"):  # timestep_forward_euler steps from the same initial state
        u = orc.read_records(_run(["euler", par, ic, os.path.join(tmp, "e.bin"), str(opts["euler"])], tmp, "e.bin"))
        for k, v in u.items():
            if k.startswith("euler"):
                out["unit/" + k] = v
    if opts.get("io"):  # checkpoint buffer (io_setbuf), io_setup, restart (io_readbuf) and steps after it
        n, more = opts["io"]
        u = orc.read_records(_run(["io", par, ic, os.path.join(tmp, "i.bin"), str(n), str(more)], tmp, "i.bin"))
        for k, v in u.items():
            if k.startswith("io/") and not k.endswith(("/cons", "/RKcons", "/RKtiph", "/wiph")):
                out[k] = v
        out["io/steps"] = np.array([n, more], dtype=np.int32)
    args = ["run", par, ic, os.path.join(tmp, "r.bin"), str(opts["steps"])] + [str(d) for d in opts["dumps"]]
    r = orc.read_records(_run(args, tmp, "r.bin"))
    missing = [d for d in opts["dumps"] if f"step{d}/prim" not in r]
    if missing or "dts" not in r:  # the reference exits with status 0 on some errors
        raise RuntimeError(f"{name}: the reference stopped before steps {missing}")
    for k, v in r.items():
        if k.startswith("step0/") or k in ("r_faces", "z_faces", "N_p", "dims"):
            continue
        if opts.get("light") and not k.endswith(("/prim", "/tiph", "/t_dt", "/gravmass")) and k != "dts":
            continue
        out["run/" + k] = v
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return path


def _run(args, tmp, outname):
    orc.run_reference(args, tmp)
    return os.path.join(tmp, outname)


def main(argv):
    if not orc.have_reference():
        raise SystemExit("build the reference harness first: make -C oracle ref")
    want = set(argv)
    with tempfile.TemporaryDirectory() as tmp:
        for name, (spec, opts) in cases().items():
            if want and name not in want:
                continue
            path = make_case(name, spec, opts, tmp)
            print(f"{name}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
