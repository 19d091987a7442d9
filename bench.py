"""Throughput of the explicit residual-and-update path on one MI355X.

A step is one Explicit_Step (all RK stages of every family) over the whole grid,
with the state resident in HBM.  Metric: Mcell-stage updates/s = interior
cells x families x RK stages x steps / time.  Prints one JSON line.

Default workload: the 2-D crossing jets (1024 x 1024 cells, AG, HLLE, MUSCL van
Leer with the Jameson sensor, drag + heat, SSP-RK3).  The CPU baseline is the
OpenMP restatement of the same loops (oracle/liboracle.so) on a bounded sample
of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from ice_amd import cases, native  # noqa: E402

HBM_PEAK_GBS = 8000.0


def stage_kernel_bytes(case, batched: bool) -> float:
    """Algorithmic HBM bytes per cell-stage of one family in the fused stage kernel:
    read the stage state and the step's start state, write the new state (3 nc
    per family); the geometry the faces and the cell read (1-D 6, 2-D 11, 3-D 16
    doubles) and the frozen gas (9 doubles) when coupled, once per launch.  A plain
    launch covers one family; a batched launch all families of the blocks, which
    share the geometry and gas reads."""
    geo = {1: 6, 2: 11, 3: 16}[case.cfg.ndir] + (9 if case.cfg.coupled else 0)
    state = sum(3 * native.NCOND[f.closure] for f in case.fams)
    nf = len(case.fams)
    if batched:
        return 8.0 * (state + geo) / nf
    return 8.0 * (state / nf + geo)


WORKLOADS = {
    "jets2d": lambda mode: cases.case_crossing_jets(n=1024, steps=500, mode=mode),
    "riemann1d": lambda mode: cases.case_riemann_1d(n=2000, steps=1000, mode=mode),
    "multiblock3d": lambda mode: cases.case_multiblock(mode=mode),
    "evap3d": lambda mode: cases.case_evaporation(n=256, steps=50, mode=mode),
    "combined": lambda mode: cases.case_combined(ntree=88, npairs=4, emin=24, emax=96, pair_n=24, mode=mode),
    "combined_full": lambda mode: cases.case_combined(ntree=88, npairs=4, emin=32, emax=192, pair_n=24, mode=mode),
}
DESCR = {
    "jets2d": "2-D crossing jets, 1024x1024 cells, AG closure, HLLE, MUSCL van Leer + Jameson sensor, "
              "drag Schiller-Naumann + Ranz-Marshall heat, SSP-RK3, FP64",
    "riemann1d": "1-D IG Riemann problem, 2000 cells, MUSCL minmod, Rusanov, drag, SSP-RK3, FP64",
    "multiblock3d": "3-D multi-block tree, 64 blocks (edges log-uniform 16-128, 6.26M cells), 3 IG families, "
                    "Rusanov, MUSCL van Leer, drag, SSP-RK3, FP64",
    "evap3d": "3-D evaporation-heavy block, 256^3 cells, MK closure, Saurel, MUSCL van Leer, drag Schiller-Naumann + "
              "Ranz-Marshall heat + grey-body radiation + CEM evaporation with Langmuir-Knudsen interface and Stefan "
              "blowing, per-cell Yinf 0-0.3, gas 300-2000 K / 1-5 bar, SSP-RK3, FP64",
    "combined": "combined configuration scaled to 10.2M cells: 88-block connection tree (edges log-uniform 24-96) + "
                "4 chimera pairs, AG x 2 families, HLLE, full interphase physics with CEM/Langmuir-Knudsen evaporation "
                "and Stefan blowing, gas as the evaporation case, steady SSP-RK3 with IRS (beta 0.5), FP64",
    "combined_full": "combined configuration at full size, 48.6M cells in 96 blocks: 88-block connection tree (edges "
                     "log-uniform 32-192) + 4 chimera pairs, AG x 2 families, HLLE, full interphase physics with "
                     "CEM/Langmuir-Knudsen evaporation and Stefan blowing, steady SSP-RK3 with IRS (beta 0.5), FP64",
}


def cpu_baseline(case, max_steps: int, budget_s: float = 12.0):
    """The oracle (OpenMP C++ restatement) on the same case, one step at a time,
    until `max_steps` steps or `budget_s` seconds of CPU work: (Mcell-stage/s,
    seconds, steps)."""
    lib = native.bind(C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so")), "orc_")
    r = cases.Runner(lib, case, "orc_")
    t0 = time.perf_counter()
    n = 0
    while n < max_steps:
        r.step(1)
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    r.close()
    return case.cell_stages(n) / dt / 1e6, dt, n


def traffic_of(workload: str, mode: str, cells_per_launch: float):
    """HBM bytes per launch of the stage kernel from the committed rocprofv3 PMC
    summary of this workload (profiles/traffic_<workload>_<mode>.json: FETCH_SIZE
    and WRITE_SIZE per launch, scaled by the factors a 1 GiB copy gives), scaled to
    this run's cells per launch; None when no summary exists."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{mode}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    return t["bytes_per_cell"] * cells_per_launch


def surrogate_compare(gpu, args):
    """The same workload on the plain path (the reference's Picard solve) and with
    the converged solve without the network (secant from the equilibrium rate),
    timed like the main run; the map evaluations of each."""
    res = {}
    for label, mode in (("plain", 0), ("no_network", 2)):
        case = WORKLOADS[args.workload](native.MODE_PLAIN)
        for f in case.fams:
            f.surrogate = mode
        r = cases.Runner(gpu, case, "ice_gpu_")
        if args.warmup:
            r.step(args.warmup)
        t0 = time.perf_counter()
        r.step(args.steps)
        res[f"{label}_value"] = case.cell_stages(args.steps) / (time.perf_counter() - t0) / 1e6
        r.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="jets2d", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="plain", choices=["plain", "batched", "surrogate"],
                    help="surrogate: evaporation cases, the network-seeded interface solve (trained first, untimed)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default stage tile (otherwise the library's autotuner picks it, untimed)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB three times on the device (the PMC traffic calibration kernel) and exit")
    ap.add_argument("--cpu-steps", type=int, default=50,
                    help="at most this many steps of the CPU baseline sample, cut at about 12 s (0: skip)")
    ap.add_argument("--kernel-steps", type=int, default=2, help="steps of the per-kernel event pass")
    ap.add_argument("--compare-steps", type=int, default=10,
                    help="plain mode: steps of a batched run of the same workload, reported beside (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copied 1 GiB three times (FETCH_SIZE / WRITE_SIZE reference)"}))
        return

    mode = native.MODE_BATCHED if args.mode == "batched" else native.MODE_PLAIN
    t_start = time.perf_counter()

    def progress(what):
        print(f"[bench] {time.perf_counter() - t_start:7.1f} s  {what}", file=sys.stderr, flush=True)

    case = WORKLOADS[args.workload](mode)
    progress(f"case built: {case.cells} cells, {len(case.blocks)} blocks")
    gpu = native.load()
    surrogate = None
    if args.mode == "surrogate":
        if not any(f.evap_select and f.intf_select == 1 for f in case.fams):
            print(json.dumps({"error": "--mode surrogate needs a workload with a Langmuir-Knudsen interface solve"}))
            sys.exit(2)
        surrogate = surrogate_compare(gpu, args)
        for f in case.fams:
            f.surrogate = 1
    run = cases.Runner(gpu, case, "ice_gpu_")
    if surrogate is not None:
        rep = np.zeros(6, np.float32)
        for p in range(1, case.cfg.ngroups + 1):
            if gpu.ice_gpu_surrogate_train(run.ctx, p, 200, 512, 3e-3, 7, rep) != 0:
                raise RuntimeError(run.error())
        if gpu.ice_gpu_surrogate_retrain_every(run.ctx, 200, 20) != 0:
            raise RuntimeError(run.error())
        surrogate["retrain_every_steps"] = 200
        surrogate["training"] = {"samples": int(rep[0]), "loss_initial": float(rep[1]), "loss_final": float(rep[2]),
                                 "rate_ratio_abs_err_mean": float(rep[3]), "rate_ratio_abs_err_max": float(rep[4]),
                                 "train_ms": float(rep[5])}
    progress("context set up")
    tuned = None
    if not args.no_autotune:
        best = np.zeros(4, np.int32)
        trial_ms = np.zeros(16)
        if gpu.ice_gpu_autotune(run.ctx, 2, best, trial_ms) != 0:
            raise RuntimeError(run.error())
        tuned = {"tile": best[:3].tolist(), "dispatch": "items" if best[3] else "queues",
                 "trial_ms_per_step": [round(float(x), 4) for x in trial_ms if np.isfinite(x) and x > 0]}
    progress("autotuned" if tuned else "no autotune")
    if args.warmup:
        run.step(args.warmup)
    t0 = time.perf_counter()
    st = run.step(args.steps)
    wall = time.perf_counter() - t0
    step_ms = C.c_double()
    gpu.ice_gpu_timing(run.ctx, C.byref(step_ms), None, None)
    step_event_ms = step_ms.value
    # the dominant kernel, timed live with HIP events on the context's stream
    # (bracketing each launch; a separate pass so the brackets do not slow the timed region)
    progress("timed region done")
    gpu.ice_gpu_set_kernel_timing(run.ctx, 1)
    run.step(args.kernel_steps)
    gpu.ice_gpu_set_kernel_timing(run.ctx, 0)
    kern_ms, launches = C.c_double(), C.c_int64()
    gpu.ice_gpu_timing(run.ctx, None, C.byref(kern_ms), C.byref(launches))
    cell_stages = case.cell_stages(args.steps)
    value = cell_stages / wall / 1e6
    avg_kernel_s = kern_ms.value / max(launches.value, 1) / 1e3
    cells_per_launch = case.cell_stages(args.kernel_steps) / max(launches.value, 1)
    bytes_per_launch = cells_per_launch * stage_kernel_bytes(case, args.mode == "batched")
    achieved = bytes_per_launch / avg_kernel_s / 1e9 if avg_kernel_s > 0 else None
    run.close()

    # the other path of the metric's comparison on the same workload, after the timed
    # region: the batched path beside a plain run (a short run of its own)
    other = None
    if args.mode == "plain" and args.compare_steps > 0:
        bcase = WORKLOADS[args.workload](native.MODE_BATCHED)
        brun = cases.Runner(gpu, bcase, "ice_gpu_")
        if not args.no_autotune:
            best = np.zeros(4, np.int32)
            if gpu.ice_gpu_autotune(brun.ctx, 2, best, np.zeros(16)) != 0:
                raise RuntimeError(brun.error())
        brun.step(2)
        t1 = time.perf_counter()
        brun.step(args.compare_steps)
        bwall = time.perf_counter() - t1
        brun.close()
        other = {"mode": "batched", "steps": args.compare_steps,
                 "value": bcase.cell_stages(args.compare_steps) / bwall / 1e6}

    cpu = None
    if args.cpu_steps > 0:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        sample = WORKLOADS[args.workload](native.MODE_PLAIN)
        v, secs, nsteps = cpu_baseline(sample, args.cpu_steps)
        cpu = {"value": v, "unit": "Mcell-stage/s", "cores": threads, "kind": "port",
               "sample": f"{nsteps} steps of the same workload from the same initial state "
                         f"({secs:.1f} s on {threads} OpenMP threads)"}

    out = {
        "metric": "Mcell-stage updates/s (plain vs batched+surrogate, 1 MI355X) at fixed L2 err",
        "value": value,
        "unit": "Mcell-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": DESCR[args.workload], "mode": args.mode, "cells": case.cells,
                   "families": case.cfg.ngroups, "rk_stages": case.cfg.nrk,
                   "event_ms_per_step": step_event_ms / args.steps, "iter": st.iter, "autotune": tuned,
                   "compared": other},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "kernel": "k_stage_tiled / k_stage_persist (fused interior flux + source + residual + RK update)",
                     "kernel_avg_ms": avg_kernel_s * 1e3, "bytes_per_cell": stage_kernel_bytes(case, args.mode == "batched"),
                     "launches": launches.value},
        "cpu_baseline": cpu,
    }
    if surrogate is not None:
        surrogate["value"] = value
        surrogate["speedup_vs_plain"] = value / surrogate["plain_value"]
        out["surrogate"] = surrogate
    traffic = traffic_of(args.workload, args.mode, cells_per_launch)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic
    print(json.dumps(out))


if __name__ == "__main__":
    main()
