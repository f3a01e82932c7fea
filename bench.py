#!/usr/bin/env python3
"""Benchmark: FP64 zone-updates/s of the hydro RK2 step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]

One "step" is one timestep_rk2 (two substeps) over the whole mesh; a
zone-update is one interior zone advanced by one step.  The headline
workload is the largest single-GPU configuration: the 3D disk with NumR 256
on r in [0.5, 8], NumZ 64 on z in [-0.5, 0.5], N_phi = round(2 pi r / dr)
capped at 1024 (12,496,704 interior zones), HLL, FP64, synthetic seeded
state.  The binary disk (NumR 512, 874,488 interior zones, HLLC, alpha
viscosity) is measured beside it.  The timed region replays a captured RK2
step graph K times with the state resident in HBM.

Printed: one JSON line with the metric; the HBM roofline of the dominant
kernel (the update, which carries the step's algorithmic state traffic:
160 B per interior zone per substep, over its HIP-event duration) and of the
whole step (320 B per zone-update); the HBM bytes and FP64 operations the
kernels actually moved, read from rocprofv3 counter passes of the library as
built (FETCH_SIZE, WRITE_SIZE and SQ_INSTS_VALU_FLOPS_FP64, the byte counters
scaled by what they read for a plain copy of known size); and the reference
CPU code timed on this host, on all cores under MPI and on one core.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, published
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector, published
B_ALG_PER_ZONE_UPDATE = 320  # 8 * (6*NUM_Q + 10) bytes at NUM_Q = 5
# The substep's algorithmic state traffic (state {prim, tiph, dphi} read and
# written, the RK snapshot written in the first substep and read in the
# second: 160 B per interior zone per substep) is all the update launch's: it
# reads the state and writes the new state and the snapshot; every other
# launch reads the state again or touches intermediates only.
B_ALG_UPDATE_PER_ZONE = B_ALG_PER_ZONE_UPDATE // 2
CALIB_DOUBLES = 1 << 27      # the plain copy the byte counters are scaled by (1 GiB in, 1 GiB out)
HOST_CORES = 16              # this GPU's share of the host's cores


def face_counts(n_p, nz_layers):
    """Overlap faces between neighbouring annuli (N_in + N_out per r pair) and
    layers (2 N per z pair), plus coincident edges, which have measure zero
    for the seeded offsets."""
    nr = int(nz_layers * (2 * int(n_p.sum()) - int(n_p[0]) - int(n_p[-1])))
    nzf = int(2 * int(n_p.sum()) * (nz_layers - 1)) if nz_layers > 1 else 0
    return nr, nzf


def kernel_touched_bytes(name, nz, nf_r, nf_z, nq, visc):
    """Bytes one launch must move if it touched every array it reads or
    writes once (intermediates included, re-reads of its gathers not): the
    per-kernel model reported beside the algorithmic figure.  nz: zones with
    ghosts; nf_r, nf_z: r and z overlap faces."""
    d = 8
    fx = nq + (3 if visc else 0)  # flux increments stored per face
    nf_dir = nf_z if name.endswith("_z") else nf_r
    if name == "k_update":
        # per zone read: prim, gradp, tiph, RKtiph, dphi, grad[vr, omega] (viscous),
        # face-list cursors (2 ints), RKcons (second substep only: half a vector
        # per launch); per zone written: prim, tiph, RKtiph, dphi, RKcons (first
        # substep: half a vector); per r and z face: the stored increments
        rd = nz * (d * (2 * nq + 3 + (2 if visc else 0) + nq / 2) + 8) + (nf_r + nf_z) * d * fx
        wr = nz * d * (nq + 3 + nq / 2)
        return int(rd + wr)
    if name.startswith("k_flux"):
        # per face: cell indices, centre phi, overlap length in, increments out;
        # per zone: prim, grad, gradp, tiph, dphi
        return nf_dir * (24 + 8 + d * fx) + nz * d * (3 * nq + 2)
    if name in ("k_plm_r", "k_plm_z"):
        # face slopes evaluated in place from the faces' two cells: per face cell
        # indices, centre phi, overlap length; per zone prim, gradp, tiph, dphi,
        # cursors in, grad out
        return nf_dir * (16 + 2 * d) + nz * (d * (2 * nq + 2) + 8) + nz * d * nq
    if name == "k_plm_rz":
        # both directions in one pass: the r and z faces, the zone inputs once,
        # two cursor pairs, the r and z slopes out
        return (nf_r + nf_z) * (16 + 2 * d) + nz * (d * (2 * nq + 2) + 16) + 2 * nz * d * nq
    if name == "k_plm_p":
        return nz * d * (nq + 1) + nz * d * nq
    if name == "k_plm_p_dt":
        return nz * d * (nq + 2) + nz * d * nq
    if name.startswith("k_faces"):
        # tiph, dphi in; per inner cell its first-face offset, per outer cell its
        # run start out; per face cell indices, width and centre out
        return nz * (2 * d + 4 + 4) + nf_dir * 24
    return None


def _reference_time(params, setup, prim0, steps, ranks, timeout):
    """The reference build's timing mode on this host: wall clock of the step
    loop of its slowest rank (stdout to /dev/null)."""
    from oracle import oracle as orc
    from disco2_amd.params import write_parfile
    with tempfile.TemporaryDirectory() as tmp:
        par = os.path.join(tmp, "w.par")
        ic = os.path.join(tmp, "w.ic")
        write_parfile(par, params)
        orc.write_icfile(ic, setup.NR, setup.NZ, setup.NQ, N_p=setup.N_p, prim=prim0)
        cmd = [orc.REF_BIN, "time", par, ic, str(steps)]
        if ranks > 1:
            cmd = [MPIRUN, "-np", str(ranks)] + cmd
        r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({r.returncode}): {r.stderr[-500:]}")
        return json.loads(r.stderr.strip().splitlines()[-1])


MPIRUN = "/opt/conda/bin/mpirun"


def cpu_baseline(workload, params, setup, prim0, steps, one_core_layers=8):
    """The reference C code (the oracle/_ref build of its own sources, gcc -O3)
    on this host's cores: under MPI on HOST_CORES ranks over the whole
    workload (the reference's own domain split, src/MPIsetup), and on one
    rank over a bounded sample (for the 3D disk: the same disk with
    one_core_layers layers instead of 64).  Without the reference build, the
    CPU restatement (oracle/) on one core."""
    from oracle import oracle as orc
    n_int = setup.interior_zones()
    if orc.have_reference():
        out = None
        try:
            ranks = min(HOST_CORES, os.cpu_count() or 1)
            if ranks > 1 and os.path.exists(MPIRUN):
                rec = _reference_time(params, setup, prim0, steps, ranks, timeout=600)
                out = {"value": rec["zone_updates_per_s"], "unit": "zone-updates/s", "cores": rec["ranks"],
                       "kind": "reference",
                       "sample": f"{steps} RK2 step(s) of the whole {workload} workload ({rec['interior_zones']} "
                                 f"interior zones), reference Disco2 C code built with gcc -O3, {rec['ranks']} MPI "
                                 f"ranks (MPICH, its own domain split), stdout to /dev/null, step loop of the "
                                 f"slowest rank"}
        except Exception as e:
            print(f"[bench] reference under MPI unavailable: {e}", file=sys.stderr, flush=True)
        try:
            p1, s1, pr1, note = params, setup, prim0, "the whole workload"
            if params["NumZ"] > one_core_layers:
                from disco2_amd import ics
                from disco2_amd.api import Setup
                p1 = dict(params, NumZ=one_core_layers,
                          Z_Min=params["Z_Min"] * one_core_layers / params["NumZ"],
                          Z_Max=params["Z_Max"] * one_core_layers / params["NumZ"])
                s1 = Setup(p1)
                from disco2_amd import configs
                _, kind, kw = configs.WORKLOADS[workload]()
                pr1 = ics.make_ic(kind, s1.geometry(), p1, **kw)
                note = f"{one_core_layers} of its {params['NumZ']} layers (same annuli and layer height)"
            rec = _reference_time(p1, s1, pr1, steps, 1, timeout=600)
            one = {"value": rec["zone_updates_per_s"], "unit": "zone-updates/s", "cores": 1, "kind": "reference",
                   "sample": f"{steps} RK2 step(s) of {note}: {rec['interior_zones']} interior zones, 1 rank"}
            if out is None:
                out = one
            else:
                out["one_core"] = one
        except Exception as e:
            print(f"[bench] reference on one core unavailable: {e}", file=sys.stderr, flush=True)
        if out is not None:
            return out
    o = orc.Oracle(params, setup.N_p, setup.r_faces, setup.z_faces)
    o.set_state(tiph=setup.tiph0, prim=prim0)
    o.init()
    t0 = time.perf_counter()
    o.step(steps)
    el = time.perf_counter() - t0
    return {"value": n_int * steps / el, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} RK2 steps of the same workload, CPU restatement (oracle/), 1 thread"}


def _kernel_key(name):
    """bench's short kernel names from the demangled ones rocprofv3 prints"""
    base = name.split("(")[0].replace("void ", "").split("<")[0].strip()
    return base


def counter_passes(workload, timeout=240):
    """HBM bytes and FP64 operations of the library as built, from three rocprofv3
    counter passes over a child run (bench.py --counters-child): FETCH_SIZE,
    WRITE_SIZE (KiB per dispatch) and SQ_INSTS_VALU_FLOPS_FP64 (per wave
    instruction: x 64 lanes).  The child copies CALIB_DOUBLES doubles (the
    reference point of the byte counters), runs one step, copies a few doubles
    (a marker) and runs the measured step eagerly.  Returns per kernel name
    (measured step) the bytes and FLOPs per launch, the byte-counter scales,
    and the step totals; None when rocprofv3 is not there or a pass fails."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE", "SQ_INSTS_VALU_FLOPS_FP64"):
            out = os.path.join(tmp, counter)
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "p", "--", sys.executable,
                   os.path.join(ROOT, "bench.py"), "--counters-child", "--workload", workload]
            env = dict(os.environ, TMPDIR=tmp)
            pr = subprocess.Popen(cmd, cwd=tmp, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                                  text=True, start_new_session=True)
            try:
                _, err = pr.communicate(timeout=timeout)
            except subprocess.TimeoutExpired:
                os.killpg(pr.pid, signal.SIGKILL)
                pr.communicate()
                print(f"[bench] counter pass {counter} timed out", file=sys.stderr, flush=True)
                return None
            files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
            if not files:
                print(f"[bench] counter pass {counter} wrote nothing ({pr.returncode}): {err[-400:]}",
                      file=sys.stderr, flush=True)
                return None
            per = {}
            for r in csv.DictReader(open(files[0])):
                key = (int(r["Dispatch_Id"]), r["Kernel_Name"])
                per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
            seq = [(d, k, v) for (d, k), v in sorted(per.items())]
            copies = [i for i, (_, k, _) in enumerate(seq) if _kernel_key(k) == "k_copy"]
            if len(copies) < 2:
                return None
            calib = seq[copies[0]][2]
            step = seq[copies[-1] + 1:]
            kern = {}
            for _, k, v in step:
                kk = _kernel_key(k)
                kern.setdefault(kk, []).append(v)
            res[counter] = {"calib": calib, "kernels": kern, "step_total": sum(v for _, _, v in step)}
    kib = CALIB_DOUBLES * 8 / 1024.0
    fs = kib / res["FETCH_SIZE"]["calib"]
    ws = kib / res["WRITE_SIZE"]["calib"]
    out = {"fetch_scale": fs, "write_scale": ws, "kernels": {}}
    for k, v in res["FETCH_SIZE"]["kernels"].items():
        w = res["WRITE_SIZE"]["kernels"].get(k, [0.0])
        f = res["SQ_INSTS_VALU_FLOPS_FP64"]["kernels"].get(k, [0.0])
        out["kernels"][k] = {"launches": len(v),
                             "bytes_per_launch": (sum(v) * fs + sum(w) * ws) * 1024.0 / len(v),
                             "flop_per_launch": 64.0 * sum(f) / max(1, len(f))}
    out["step_bytes"] = (res["FETCH_SIZE"]["step_total"] * fs + res["WRITE_SIZE"]["step_total"] * ws) * 1024.0
    out["step_flop"] = 64.0 * res["SQ_INSTS_VALU_FLOPS_FP64"]["step_total"]
    return out


def counters_child(workload, schedule=None):
    """the run the counter passes profile (see counter_passes); schedule: a
    launch schedule to set instead of the default"""
    from disco2_amd import ics
    from disco2_amd._lib import lib
    from disco2_amd.api import Batch, Setup
    from disco2_amd import configs
    params, ic_kind, ic_kw = configs.WORKLOADS[workload]()
    s = Setup(params)
    b = Batch([s])
    b.upload("prim", ics.make_ic(ic_kind, s.geometry(), params, **ic_kw))
    b.init()
    if schedule is not None:
        b.set_schedule(schedule)
    lib().disco_calibrate_copy(CALIB_DOUBLES)
    b.run(1, graph=False)
    lib().disco_calibrate_copy(64)
    b.run(1, graph=False)
    b.close()
    return 0


def batched_leg(steps, warmup, ndisk=64):
    """Many independent disks: one batched launch set against one launch set
    per disk run disk after disk.  Both run the same disks for the same number
    of RK2 steps from the same initial states."""
    from disco2_amd import configs, ics
    from disco2_amd.api import Batch, Setup
    specs = configs.c4_batch(ndisk)
    setups, prims = [], []
    for p, kind, kw in specs:
        s = Setup(p)
        setups.append(s)
        prims.append(ics.make_ic(kind, s.geometry(), p, **kw))
    n_int = sum(s.interior_zones() for s in setups)
    big = Batch(setups)
    for d, pr in enumerate(prims):
        big.upload("prim", pr, disk=d)
    big.init()
    k, cand = big.autotune(3)
    big.run(warmup)
    big.synchronize()
    t0 = time.perf_counter()
    big.run(steps, sync=False)
    big.synchronize()
    t_batched = time.perf_counter() - t0
    big.close()
    # per disk: its own captured step graph (the strong baseline), and its
    # kernels launched one by one as the class API's timestep_rk2 does
    t_plain = t_eager = 0.0
    for s, pr in zip(setups, prims):
        b = Batch([s])
        b.upload("prim", pr)
        b.init()
        b.run(warmup)
        b.synchronize()
        t0 = time.perf_counter()
        b.run(steps, sync=False)
        b.synchronize()
        t_plain += time.perf_counter() - t0
        b.run(1, graph=False)
        t0 = time.perf_counter()
        b.run(steps, graph=False, sync=False)
        b.synchronize()
        t_eager += time.perf_counter() - t0
        b.close()
    return {"workload": f"{ndisk} binary disks, NumR uniform in [64, 384], q, alpha, Mach drawn per disk",
            "interior_zones": n_int, "steps": steps,
            "batched": {"value": n_int * steps / t_batched, "ms_per_step": 1e3 * t_batched / steps},
            "per_disk_launches": {"value": n_int * steps / t_plain, "ms_per_step": 1e3 * t_plain / steps,
                                  "graph": True},
            "per_disk_eager_launches": {"value": n_int * steps / t_eager, "ms_per_step": 1e3 * t_eager / steps,
                                        "graph": False},
            "speedup": t_plain / t_batched, "speedup_vs_eager": t_eager / t_batched, "unit": "zone-updates/s",
            "schedule": {"chosen": k, "ms_per_step": {str(c): round(v, 4) for c, v in cand.items()}}}


def adjoint_leg(steps=200, every=20):
    """The adjoint at the binary2d_r256_adjoint size: 200 RK2 steps forward and
    the hand-written adjoint back, against the same 200 steps forward alone
    (graph replay).  Two ways: the whole run recorded (one segment, no
    recompute) and checkpoints every 20 steps (each earlier segment recomputed
    from its checkpoint).  Each adjoint time includes its own forward pass."""
    from disco2_amd import configs, ics
    from disco2_amd.api import Batch, Setup
    p, kind, kw = configs.c5_adjoint()
    s = Setup(p)
    prim0 = ics.make_ic(kind, s.geometry(), p, **kw)
    b = Batch([s])
    b.upload("prim", prim0)
    b.init()
    b.run(2)  # warm the graph and the kernels
    b.upload("prim", prim0)
    b.init()
    b.synchronize()
    t0 = time.perf_counter()
    b.run(steps, sync=False)
    b.synchronize()
    t_fwd = time.perf_counter() - t0
    b.close()
    res = {"workload": "binary2d_r256_adjoint", "interior_zones": s.interior_zones(), "steps": steps,
           "forward_ms_per_step": 1e3 * t_fwd / steps}
    # the batch keeps the adjoint's working memory (the segment record) from
    # its first call, as it keeps its state buffers: the first call, which
    # allocates it, is timed apart
    b = Batch([s])
    for name, ev in (("first_call", steps), ("record", steps), ("checkpointed", every)):
        b.upload("prim", prim0)
        b.upload("tiph", s.tiph0)
        b.upload("dphi", s.dphi0)
        b.init()
        b.synchronize()
        t0 = time.perf_counter()
        loss, ga, gq, gp = b.adjoint(steps, checkpoint_every=ev)
        t_adj = time.perf_counter() - t0
        res[name] = {"checkpoint_every": ev, "adjoint_ms_per_step": 1e3 * t_adj / steps,
                     "ratio": t_adj / t_fwd, "loss": loss, "grad_alpha": ga, "grad_q": gq}
    b.close()
    return res


def diagnostics_leg(batch, params, reps=10):
    """The reference's diagnostics measurement (diagnostics_set) on the bench
    disk: wall clock of disco_batch_diag_measure, the two device kernels and
    the copies of the radial profiles (NumR x 17) and scalars; and once more
    with the equatorial cells (one per zone, 18 doubles) copied to the host,
    which the reference does at every dump."""
    import numpy as np
    rz = np.array([params["R_Min"], params["R_Max"], params["Z_Min"], params["Z_Max"]], dtype=np.float64)
    batch.diag_measure(1.0, rz, equat=False)  # allocates the scratch
    batch.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.diag_measure(1.0, rz, equat=False)
    wall = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    _, _, eq, _ = batch.diag_measure(1.0, rz, equat=True)
    wall_eq = time.perf_counter() - t0
    return {"ms_per_measurement": round(1e3 * wall, 4), "ms_with_equatorial_copy": round(1e3 * wall_eq, 4),
            "equatorial_cells": int(eq.shape[0]), "measurements": reps}


def measure_workload(name, args, batch_out=None):
    """Time one single-disk workload: autotuned schedule, K graph-replayed
    steps (synchronised on both sides), per-kernel HIP-event durations of a
    few eager steps, and the update's roofline."""
    from disco2_amd import configs, ics
    from disco2_amd.api import Batch, Setup
    params, ic_kind, ic_kw = configs.WORKLOADS[name]()
    setup = Setup(params)
    prim0 = ics.make_ic(ic_kind, setup.geometry(), params, **ic_kw)
    n_int = setup.interior_zones()
    batch = Batch([setup])
    batch.upload("prim", prim0)
    batch.init()
    graph = not args.no_graph
    schedule = None
    if args.schedule is not None:
        batch.set_schedule(args.schedule)
        schedule = {"chosen": args.schedule, "ms_per_step": None}
    elif not args.no_autotune:
        k, cand = batch.autotune(args.autotune_steps)
        schedule = {"chosen": k, "ms_per_step": {str(c): round(v, 4) for c, v in cand.items()}}
        print(f"[bench] {name}: autotuned schedule {k}: {schedule['ms_per_step']}", file=sys.stderr, flush=True)
    print(f"[bench] {name}: {setup.nzones} zones ({n_int} interior), warmup {args.warmup}", file=sys.stderr,
          flush=True)
    batch.run(args.warmup, graph=graph)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.run(args.steps, graph=graph, sync=False)
    batch.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = batch.last_ms()
    ms_per_step = 1e3 * wall / args.steps
    value = n_int * args.steps / wall
    print(f"[bench] {name}: timed {args.steps} steps: {ms_per_step:.3f} ms/step (events {ev_ms / args.steps:.3f})",
          file=sys.stderr, flush=True)
    prof = batch.profile(args.profile_steps)
    nf_r, nf_z = face_counts(setup.N_p, setup.NZ)
    nq = setup.NQ
    visc = params["Explicit_Viscosity"] > 0
    upd_ms, upd_n = prof["k_update"]
    upd_avg = upd_ms / upd_n
    alg = B_ALG_UPDATE_PER_ZONE * n_int
    ach = alg / (upd_avg * 1e-3) / 1e9
    roof = {"kernel": "k_update", "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "avg_launch_ms": round(upd_avg, 5),
            "alg_bytes_per_launch": int(alg),
            "alg_basis": f"{B_ALG_UPDATE_PER_ZONE} B per interior zone per substep (the substep's state and RK "
                         f"snapshot traffic) x {n_int} interior zones",
            "touched_bytes_per_launch": kernel_touched_bytes("k_update", setup.nzones, nf_r, nf_z, nq, visc)}
    kernels = {k: {"avg_ms": round(v[0] / v[1], 5), "launches_per_step": v[1] / args.profile_steps,
                   "share": round(v[0] / sum(x[0] for x in prof.values()), 4),
                   "touched_bytes_per_launch": kernel_touched_bytes(k, setup.nzones, nf_r, nf_z, nq, visc)}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}
    step_ach = value * B_ALG_PER_ZONE_UPDATE / 1e9
    res = {"workload": name, "params": params, "setup": setup, "prim0": prim0, "n_int": n_int,
           "value": value, "ms_per_step": ms_per_step, "graph": graph, "schedule": schedule,
           "roofline": roof, "kernels": kernels,
           "step_roofline": {"alg_bytes_per_zone_update": B_ALG_PER_ZONE_UPDATE, "achieved": round(step_ach, 1),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_ach / HBM_PEAK_GBS, 4)},
           "faces": {"r": nf_r, "z": nf_z}}
    if batch_out is not None:
        batch_out.append(batch)
    else:
        batch.close()
    return res


def add_counters(res, cnt):
    """the measured traffic and FLOPs of the counter passes into a workload's
    record (None: the counters could not be read)"""
    if not cnt:
        res["counters"] = None
        return
    upd = cnt["kernels"].get("k_update")
    if upd:
        res["roofline"]["traffic"] = int(upd["bytes_per_launch"])
        res["roofline"]["traffic_over_alg"] = round(upd["bytes_per_launch"] / res["roofline"]["alg_bytes_per_launch"], 2)
        fl = upd["flop_per_launch"] / (res["roofline"]["avg_launch_ms"] * 1e-3) / 1e12
        res["flop_roofline"] = {"kernel": "k_update", "bound": "fp64", "achieved": round(fl, 2),
                                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(fl / FP64_PEAK_TFLOPS, 4),
                                "flop_per_launch": upd["flop_per_launch"]}
    n = res["n_int"]
    st = cnt["step_flop"] / (res["ms_per_step"] * 1e-3) / 1e12
    res["step_roofline"]["measured_bytes_per_step"] = int(cnt["step_bytes"])
    res["step_roofline"]["measured_bytes_per_zone_update"] = round(cnt["step_bytes"] / n, 1)
    res["step_flop_roofline"] = {"flop_per_zone_update": round(cnt["step_flop"] / n, 1), "achieved": round(st, 2),
                                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(st / FP64_PEAK_TFLOPS, 4)}
    res["counters"] = {"fetch_scale": round(cnt["fetch_scale"], 4), "write_scale": round(cnt["write_scale"], 4),
                       "per_kernel": {k: {"launches_per_step": v["launches"],
                                          "bytes_per_launch": int(v["bytes_per_launch"]),
                                          "fp64_flop_per_launch": v["flop_per_launch"]}
                                      for k, v in sorted(cnt["kernels"].items())},
                       "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE / SQ_INSTS_VALU_FLOPS_FP64, one pass "
                                 "each, over one eager step of the library as built (bench.py --counters-child)"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="disk3d_r256z64")
    ap.add_argument("--secondary", default="binary2d_r512", help="also measured and reported beside ('' for none)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch schedule")
    ap.add_argument("--autotune-steps", type=int, default=5)
    ap.add_argument("--schedule", type=int, default=None,
                    help="use this launch schedule (0 .. DISCO_NSCHEDULES - 1) instead of autotuning")
    ap.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--counters-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg and print its JSON (for profiling it)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched leg and print its JSON (for profiling it)")
    ap.add_argument("--diag-only", action="store_true",
                    help="run only the diagnostics leg on the secondary workload and print its JSON")
    args = ap.parse_args(argv)
    if args.counters_child:
        return counters_child(args.workload, args.schedule)
    if args.gpus != 1:
        print("bench.py measures one GPU: disco2_amd has no multi-GPU path (one disk or batch per "
              "GPU); run with --gpus 1", file=sys.stderr)
        return 2
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(args.adjoint_steps, args.adjoint_every)))
        return 0
    if args.batched_only:
        print(json.dumps(batched_leg(args.batched_steps, args.warmup)))
        return 0
    if args.diag_only:
        from disco2_amd import configs, ics
        from disco2_amd.api import Batch, Setup
        params, kind, kw = configs.WORKLOADS[args.secondary or args.workload]()
        s = Setup(params)
        b = Batch([s])
        b.upload("prim", ics.make_ic(kind, s.geometry(), params, **kw))
        b.init()
        print(json.dumps({"workload": args.secondary, "diagnostics": diagnostics_leg(b, params)}))
        b.close()
        return 0

    head = measure_workload(args.workload, args)
    second = None
    keep = []
    if args.secondary:
        second = measure_workload(args.secondary, args, batch_out=keep)
    if not args.no_counters:
        print("[bench] counter passes", file=sys.stderr, flush=True)
        add_counters(head, counter_passes(args.workload))
        if second:
            add_counters(second, counter_passes(args.secondary))

    batched = None
    if not args.no_batched:
        print("[bench] batched disks", file=sys.stderr, flush=True)
        batched = batched_leg(args.batched_steps, 2)

    diagnostics = None
    if keep:
        print("[bench] diagnostics", file=sys.stderr, flush=True)
        diagnostics = dict(workload=args.secondary, **diagnostics_leg(keep[0], second["params"]))
        keep[0].close()

    adjoint = None
    if not args.no_adjoint:
        print("[bench] adjoint", file=sys.stderr, flush=True)
        adjoint = adjoint_leg(args.adjoint_steps, args.adjoint_every)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.workload, head["params"], head["setup"], head["prim0"], args.cpu_steps)

    def config(r):
        p = r["params"]
        return {"workload": r["workload"], "schedule": r["schedule"], "NumR": p["NumR"], "NumZ": p["NumZ"],
                "zones": r["setup"].nzones, "interior_zones": r["n_int"], "faces": r["faces"],
                "riemann": "HLLC" if p["RiemannSolver"] == 0 else "HLL",
                "viscosity": p["Explicit_Viscosity"] > 0, "graph": r["graph"], "parallelism": "single disk, 1 GPU"}

    out = {
        "metric": "FP64 zone-updates/s on 1 GPU (plain single disk)",
        "value": head["value"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": config(head),
        "roofline": head["roofline"],
        "step_roofline": head["step_roofline"],
        "flop_roofline": head.get("flop_roofline"),
        "step_flop_roofline": head.get("step_flop_roofline"),
        "counters": head.get("counters"),
        "kernels": head["kernels"],
    }
    if second:
        out["secondary"] = {"value": second["value"], "unit": "zone-updates/s", "ms_per_step": second["ms_per_step"],
                            "config": config(second), "roofline": second["roofline"],
                            "step_roofline": second["step_roofline"],
                            "flop_roofline": second.get("flop_roofline"),
                            "step_flop_roofline": second.get("step_flop_roofline"), "counters": second.get("counters"),
                            "kernels": second["kernels"]}
    out.update({"batched": batched, "adjoint": adjoint, "diagnostics": diagnostics, "cpu_baseline": cpu})
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
