#!/usr/bin/env python3
"""Benchmark: FP64 zone-updates/s of the hydro RK2 step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]

One "step" is one timestep_rk2 (two substeps) over the whole mesh; a
zone-update is one interior zone advanced by one step.  The default workload
is the 2D binary disk with NumR = 512 on r in [0.3, 10], N_phi =
round(2 pi r / dr) (874,488 interior zones), FP64, HLLC, alpha viscosity,
synthetic seeded initial state.  The timed region replays a captured RK2
step graph K times with the state resident in HBM.

Printed: one JSON line with the metric, the HBM roofline of the dominant
kernel (algorithmic bytes per launch over its HIP-event duration), the
whole-step roofline, and the reference CPU path timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, published
B_ALG_PER_ZONE_UPDATE = 320  # 8 * (6*NUM_Q + 10) bytes at NUM_Q = 5

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE, then
# WRITE_SIZE (KiB per dispatch, one counter per run, commands and per-kernel
# table in profiles/binary2d_r512_hbm_traffic.md), scaled by what the same
# counters read for a plain 1 GiB copy: FETCH_SIZE x 2.0, WRITE_SIZE x 1.0.
# Counters cannot be read from inside the timed run, so the figure is
# reported only for the workload and kernel it was measured on.
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
MEASURED_TRAFFIC_KIB = {("binary2d_r512", "k_update"): (135872, 72707)}

# FP64 arithmetic, measured the same way (rocprofv3 --pmc
# SQ_INSTS_VALU_FLOPS_FP64 ..., profiles/binary2d_r512_fp64_flops.md): the
# counter gives FLOPs per wave instruction (an FMA counts 2; it equals
# ADD + MUL + 2 FMA + TRANS of the per-type instruction counters), times 64
# lanes; masked lanes at tile ends are counted, so these are upper bounds.
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector, published
MEASURED_FLOP_PER_LAUNCH = {("binary2d_r512", "k_update"): 2.258e7 * 64}
MEASURED_FLOP_PER_STEP = {"binary2d_r512": 1.1626e8 * 64}  # every kernel of one RK2 step


def r_face_count(n_p, nz_layers):
    """Overlap faces between neighbouring annuli: N_in + N_out per pair (plus
    coincident edges, which have measure zero for the seeded offsets)."""
    return int(nz_layers * (2 * int(n_p.sum()) - int(n_p[0]) - int(n_p[-1])))


def kernel_alg_bytes(name, nz, nfaces_r, nq, visc):
    """Algorithmic (compulsory) bytes of one launch of a kernel: every array
    the kernel must touch, read or written once, whatever the re-reads of
    its gathers.  nz: zones incl. ghosts; nfaces_r: r overlap faces."""
    d = 8
    fx = nq + (3 if visc else 0)  # flux increments stored per r face
    if name == "k_update":
        # per zone read: prim, gradp (phi-face states), tiph, RKtiph, dphi,
        # grad[vr, omega] (viscous flux and source), face-list cursors (2 ints),
        # RKcons (read in the second substep only: half a vector per launch);
        # per zone written: prim, tiph, RKtiph, dphi, RKcons (first substep:
        # half a vector per launch); per r face: the stored increments
        rd = nz * (d * (2 * nq + 3 + (2 if visc else 0) + nq / 2) + 8) + nfaces_r * d * fx
        wr = nz * d * (nq + 3 + nq / 2)
        return int(rd + wr)
    if name == "k_flux_r":
        # per face: cell indices, centre phi, overlap length in; increments out;
        # per zone: prim, grad, gradp, tiph, dphi
        return nfaces_r * (24 + 8 + d * fx) + nz * d * (3 * nq + 2)
    if name == "k_plm_r":
        # the face slopes are evaluated in place from the faces' two cells:
        # per face: cell indices, centre phi, overlap length; per zone: prim,
        # gradp, tiph, dphi and the cursors in, grad out
        return nfaces_r * (16 + 2 * d) + nz * (d * (2 * nq + 2) + 8) + nz * d * nq
    if name == "k_plm_p":
        return nz * d * (nq + 1) + nz * d * nq
    if name == "k_plm_p_dt":
        # the phi slopes plus the CFL minimum of the same cells (tiph for the
        # alpha viscosity's position)
        return nz * d * (nq + 2) + nz * d * nq
    if name == "k_mindt":
        return nz * d * (nq + 2)
    if name == "k_faces_r":
        # tiph, dphi in; per inner cell its first-face offset, per outer cell
        # its run start out; per face cell indices, width and centre out
        return nz * (2 * d + 4 + 4) + nfaces_r * 24
    return None


def cpu_baseline(params, setup, prim0, steps):
    """The reference C code (oracle/_ref build) on this host, single thread;
    falls back to the CPU restatement of the oracle when the reference build
    is missing."""
    from oracle import oracle as orc
    from disco2_amd.params import write_parfile
    n_int = setup.interior_zones()
    if orc.have_reference():
        with tempfile.TemporaryDirectory() as tmp:
            par = os.path.join(tmp, "w.par")
            ic = os.path.join(tmp, "w.ic")
            write_parfile(par, params)
            orc.write_icfile(ic, setup.NR, setup.NZ, setup.NQ, N_p=setup.N_p, prim=prim0)
            try:
                r = orc.run_reference(["time", par, ic, str(steps)], tmp, timeout=900)
                rec = json.loads(r.stderr.strip().splitlines()[-1])
                return {"value": rec["zone_updates_per_s"], "unit": "zone-updates/s", "cores": 1,
                        "kind": "reference",
                        "sample": f"{steps} RK2 steps of the same workload ({rec['interior_zones']} interior "
                                  f"zones), reference Disco2 C code built with gcc -O3, 1 MPI rank, stdout to "
                                  f"/dev/null, step loop only"}
            except Exception as e:  # fall through to the restatement
                print(f"reference baseline unavailable: {e}", file=sys.stderr)
    o = orc.Oracle(params, setup.N_p, setup.r_faces, setup.z_faces)
    o.set_state(tiph=setup.tiph0, prim=prim0)
    o.init()
    t0 = time.perf_counter()
    o.step(steps)
    el = time.perf_counter() - t0
    return {"value": n_int * steps / el, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} RK2 steps of the same workload, CPU restatement (oracle/), 1 thread"}


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


def adjoint_leg(steps=200, every=50):
    """The adjoint at the binary2d_r256_adjoint size: 200 RK2 steps forward and
    the hand-written adjoint back, against the same 200 steps forward alone
    (graph replay).  Two ways: the whole run recorded (one segment, no
    recompute) and checkpoints every 50 steps (each earlier segment recomputed
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="binary2d_r512")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch schedule")
    ap.add_argument("--autotune-steps", type=int, default=5)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=50)
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg and print its JSON (for profiling it)")
    ap.add_argument("--diag-only", action="store_true",
                    help="run only the diagnostics leg on the workload and print its JSON (for profiling it)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: disco2_amd has no multi-GPU path (one disk or batch per "
              "GPU); run with --gpus 1", file=sys.stderr)
        return 2
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(args.adjoint_steps, args.adjoint_every)))
        return 0

    from disco2_amd import configs, ics
    from disco2_amd.api import Batch, Setup

    params, ic_kind, ic_kw = configs.WORKLOADS[args.workload]()
    setup = Setup(params)
    prim0 = ics.make_ic(ic_kind, setup.geometry(), params, **ic_kw)
    n_int = setup.interior_zones()
    batch = Batch([setup])
    batch.upload("prim", prim0)
    batch.init()
    if args.diag_only:
        print(json.dumps({"workload": args.workload, "diagnostics": diagnostics_leg(batch, params)}))
        batch.close()
        return 0
    graph = not args.no_graph
    schedule = None
    if not args.no_autotune:
        k, cand = batch.autotune(args.autotune_steps)
        schedule = {"chosen": k, "ms_per_step": {str(c): round(v, 4) for c, v in cand.items()}}
        print(f"[bench] autotuned schedule {k}: {schedule['ms_per_step']}", file=sys.stderr, flush=True)
    print(f"[bench] {args.workload}: {setup.nzones} zones ({n_int} interior), warmup {args.warmup}",
          file=sys.stderr, flush=True)
    batch.run(args.warmup, graph=graph)

    # timed region: K steps, synchronised on both sides
    batch.synchronize()
    t0 = time.perf_counter()
    batch.run(args.steps, graph=graph, sync=False)
    batch.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = batch.last_ms()
    ms_per_step = 1e3 * wall / args.steps
    value = n_int * args.steps / wall
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.3f} ms/step (events {ev_ms / args.steps:.3f})",
          file=sys.stderr, flush=True)

    # per-kernel HIP-event durations on the batch's stream
    prof = batch.profile(args.profile_steps)
    nfr = r_face_count(setup.N_p, setup.NZ)
    nq = setup.NQ
    visc = params["Explicit_Viscosity"] > 0
    dom = max(prof.items(), key=lambda kv: kv[1][0])
    dom_name, (dom_ms, dom_n) = dom
    dom_avg_ms = dom_ms / dom_n
    alg = kernel_alg_bytes(dom_name, setup.nzones, nfr, nq, visc)
    roof = {"kernel": dom_name, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": None, "traffic": None, "avg_launch_ms": round(dom_avg_ms, 5)}
    if alg:
        ach = alg / (dom_avg_ms * 1e-3) / 1e9
        roof.update(achieved=round(ach, 1), frac=round(ach / HBM_PEAK_GBS, 4), alg_bytes_per_launch=int(alg))
    meas = MEASURED_TRAFFIC_KIB.get((args.workload, dom_name))
    if meas:
        roof["traffic"] = int((meas[0] * FETCH_SCALE + meas[1] * WRITE_SCALE) * 1024)
    flop_roof = None
    fl = MEASURED_FLOP_PER_LAUNCH.get((args.workload, dom_name))
    if fl:
        fl_ach = fl / (dom_avg_ms * 1e-3) / 1e12
        flop_roof = {"kernel": dom_name, "bound": "fp64", "achieved": round(fl_ach, 2), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(fl_ach / FP64_PEAK_TFLOPS, 4), "flop_per_launch": fl}
    fs = MEASURED_FLOP_PER_STEP.get(args.workload)
    if fs:
        st_ach = fs / (ms_per_step * 1e-3) / 1e12
        flop_step = {"flop_per_zone_update": round(fs / n_int, 1), "achieved": round(st_ach, 2),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(st_ach / FP64_PEAK_TFLOPS, 4)}
    else:
        flop_step = None
    step_ach = value * B_ALG_PER_ZONE_UPDATE / 1e9
    kernels = {k: {"avg_ms": round(v[0] / v[1], 5), "share": round(v[0] / sum(x[0] for x in prof.values()), 4)}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}

    batched = None
    if not args.no_batched:
        print("[bench] batched disks", file=sys.stderr, flush=True)
        batched = batched_leg(args.batched_steps, 2)

    print("[bench] diagnostics", file=sys.stderr, flush=True)
    diagnostics = diagnostics_leg(batch, params)

    adjoint = None
    if not args.no_adjoint:
        print("[bench] adjoint", file=sys.stderr, flush=True)
        adjoint = adjoint_leg(args.adjoint_steps, args.adjoint_every)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(params, setup, prim0, args.cpu_steps)

    out = {
        "metric": "FP64 zone-updates/s on 1 GPU (plain single disk)",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "schedule": schedule, "NumR": params["NumR"], "NumZ": params["NumZ"],
                   "zones": setup.nzones, "interior_zones": n_int, "riemann": "HLLC" if params["RiemannSolver"] == 0 else "HLL",
                   "graph": graph, "parallelism": "single disk, 1 GPU"},
        "roofline": roof,
        "step_roofline": {"alg_bytes_per_zone_update": B_ALG_PER_ZONE_UPDATE, "achieved": round(step_ach, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_ach / HBM_PEAK_GBS, 4)},
        "flop_roofline": flop_roof,
        "step_flop_roofline": flop_step,
        "kernels": kernels,
        "batched": batched,
        "adjoint": adjoint,
        "diagnostics": diagnostics,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    batch.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
