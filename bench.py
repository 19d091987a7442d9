"""Throughput of the explicit residual-and-update path on one MI355X.

A step is one Explicit_Step (all RK stages of every family) over the whole grid,
with the state resident in HBM.  Metric: Mcell-stage updates/s = interior
cells x families x RK stages x steps / time.  Prints one JSON line.

Default workload: the full combined configuration, the largest that fits one
GPU -- 96 blocks (an 88-block connection tree with edges log-uniform on
[32, 192] and 4 chimera overset pairs), 146.8 M cells, AG closure x 2 families,
HLLE, MUSCL van Leer, drag + heat + grey-body radiation + CEM evaporation with
the Langmuir-Knudsen interface and Stefan blowing, steady SSP-RK3 with implicit
residual smoothing -- on the batched path with the learned seed of the
interface solve in the loop.  The same context then runs the batched path with
the plain interface solve and the plain path, so the line carries both sides
of the metric's comparison.  The CPU baseline is the OpenMP restatement of the
same loops (oracle/liboracle.so) on a bounded, reduced sample of the workload.
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
    """Algorithmic HBM bytes per family-cell-stage of the fused stage kernel.  Per
    family: read the stage state (nc), read the step's start state (nc, not with
    IRS, where the update is a kernel of its own), write the new state or the
    residual (nc), read the cell's dt (1, steady runs); once per launch: the
    geometry the faces and the cell read (1-D 6, 2-D 11, 3-D 16 doubles) and, when
    coupled, the frozen gas (9 doubles, 10 with a per-cell far-field vapour
    fraction).  A plain launch covers one family; a batched launch all families of
    the blocks, which share the geometry and gas reads."""
    cfg = case.cfg
    geo = {1: 6, 2: 11, 3: 16}[cfg.ndir] + ((9 + (1 if cfg.var_yinf else 0)) if cfg.coupled else 0)
    per = 0
    for f in case.fams:
        nc = native.NCOND[f.closure]
        per += nc + (0 if cfg.irs else nc) + nc + (0 if cfg.time_accurate else 1)
    nf = len(case.fams)
    if batched:
        return 8.0 * (per + geo) / nf
    return 8.0 * (per / nf + geo)


def path_bytes(case) -> float:
    """HBM bytes the whole step must move per family-cell-stage on this path: each
    family's U_stage, U_0 and U_next (3 nc), the geometry (3-D 16 doubles), the gas
    (9) when coupled, the per-family dt read each stage (steady), the
    once-per-step dt pass spread over the stages, and with IRS the residual
    buffers written and read once each (the stage's residual, and one more with
    three resolved directions or with one: 2 nc per buffer) plus U_stage read
    again by the update (nc)."""
    cfg = case.cfg
    nd = cfg.ndir
    g = {1: 6, 2: 11, 3: 16}[nd] + (9 if cfg.coupled else 0)
    b = g
    dtp = 0
    nbuf = 1 if nd == 2 else 2
    for f in case.fams:
        nc = native.NCOND[f.closure]
        b += 3 * nc + ((2 * nbuf + 1) * nc if cfg.irs else 0) + (0 if cfg.time_accurate else 1)
        dtp += nc + (1 if cfg.coupled else 0) + (0 if cfg.time_accurate else 1)
    b += (dtp + 3 * nd) / cfg.nrk
    return 8.0 * b / len(case.fams)


WORKLOADS = {
    "combined_full": lambda mode: cases.case_combined_full(mode=mode),
    "combined": lambda mode: cases.case_combined(ntree=88, npairs=4, emin=24, emax=96, pair_n=24, mode=mode),
    "jets2d": lambda mode: cases.case_crossing_jets(n=1024, steps=500, mode=mode),
    "riemann1d": lambda mode: cases.case_riemann_1d(n=2000, steps=1000, mode=mode),
    "multiblock3d": lambda mode: cases.case_multiblock(mode=mode),
    "evap3d": lambda mode: cases.case_evaporation(n=256, steps=50, mode=mode),
}
# workloads too large for the autotuner's device copy of the state: their tile is
# tuned on a smaller workload of the same closure, physics and block shapes
TUNE_PROXY = {"combined_full": "combined"}
# the bounded sample the CPU baseline steps (the workload itself where it fits)
CPU_SAMPLE = {
    "combined_full": ("the combined configuration reduced to 16 tree blocks (edges 24-64) and 2 chimera pairs",
                      lambda: cases.case_combined(ntree=16, npairs=2, emin=24, emax=64, pair_n=16)),
}
DESCR = {
    "combined_full": "combined configuration at full size: 96 blocks, 146.8M cells -- an 88-block connection tree "
                     "(edges log-uniform 32-192) + 4 chimera overset pairs of 128x64x64 blocks -- AG x 2 families, "
                     "HLLE, MUSCL van Leer, drag Schiller-Naumann + Ranz-Marshall heat + grey-body radiation + CEM "
                     "evaporation with Langmuir-Knudsen interface and Stefan blowing, gas T 300-2000 K / p 1-5 bar / "
                     "Yinf 0-0.3, steady SSP-RK3 with IRS (beta 0.5), FP64",
    "combined": "combined configuration scaled to 10.2M cells: 88-block connection tree (edges log-uniform 24-96) + "
                "4 chimera pairs, AG x 2 families, HLLE, full interphase physics with CEM/Langmuir-Knudsen evaporation "
                "and Stefan blowing, gas as the evaporation case, steady SSP-RK3 with IRS (beta 0.5), FP64",
    "jets2d": "2-D crossing jets, 1024x1024 cells, AG closure, HLLE, MUSCL van Leer + Jameson sensor, "
              "drag Schiller-Naumann + Ranz-Marshall heat, SSP-RK3, FP64",
    "riemann1d": "1-D IG Riemann problem, 2000 cells, MUSCL minmod, Rusanov, drag, SSP-RK3, FP64",
    "multiblock3d": "3-D multi-block tree, 64 blocks (edges log-uniform 16-128, 6.26M cells), 3 IG families, "
                    "Rusanov, MUSCL van Leer, drag, SSP-RK3, FP64",
    "evap3d": "3-D evaporation-heavy block, 256^3 cells, MK closure, Saurel, MUSCL van Leer, drag Schiller-Naumann + "
              "Ranz-Marshall heat + grey-body radiation + CEM evaporation with Langmuir-Knudsen interface and Stefan "
              "blowing, per-cell Yinf 0-0.3, gas 300-2000 K / 1-5 bar, SSP-RK3, FP64",
}


def has_lk(case) -> bool:
    return any(f.evap_select and f.intf_select == 1 for f in case.fams)


def cpu_threads() -> int:
    """Host threads the CPU baseline uses: the CPUs this process may run on,
    capped by OMP_NUM_THREADS (the host share a job on the GPU box is given)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(avail, env) if env > 0 else avail


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


def counters_of(workload: str, mode: str):
    """The committed rocprofv3 PMC summary of this workload's stage kernel
    (profiles/traffic_<workload>_<mode>.json: HBM bytes per cell-stage from
    FETCH_SIZE and WRITE_SIZE, scaled by the factors a 1 GiB copy gives, and FP64
    flops per cell-stage from the SQ_INSTS_VALU_*_F64 counters), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{mode}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def autotune(gpu, run, steps: int = 2):
    best = np.zeros(4, np.int32)
    trial_ms = np.zeros(24)
    if gpu.ice_gpu_autotune(run.ctx, steps, best, trial_ms) != 0:
        raise RuntimeError(run.error())
    return {"tile": best[:3].tolist(), "dispatch": int(best[3]),
            "trial_ms_per_step": [round(float(x), 4) for x in trial_ms if np.isfinite(x) and x > 0]}


def device_used_gb():
    """Device memory in use (hipMemGetInfo: total - free), GB; None if the runtime
    is not found."""
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return None
    free, total = C.c_size_t(), C.c_size_t()
    if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        return None
    return (total.value - free.value) / 1e9


def kernel_pass(gpu, run, steps: int = 1):
    """`steps` steps with HIP events around every launch of the stage kernel and of
    the deferred interface-solve kernel: (stage ms, stage launches, solve ms, solve
    launches) summed over the steps."""
    gpu.ice_gpu_set_kernel_timing(run.ctx, 1)
    run.step(steps)
    gpu.ice_gpu_set_kernel_timing(run.ctx, 0)
    km, nl = C.c_double(), C.c_int64()
    gpu.ice_gpu_timing(run.ctx, None, C.byref(km), C.byref(nl))
    lm, ll = C.c_double(), C.c_int64()
    gpu.ice_gpu_lk_timing(run.ctx, C.byref(lm), C.byref(ll))
    return km.value, nl.value, lm.value, ll.value


def timed_steps(run, steps: int) -> float:
    t0 = time.perf_counter()
    run.step(steps)
    return time.perf_counter() - t0


def lk_stats(gpu, run, case, steps: int = 1):
    """The interface solves of `steps` steps: per family solves, map evaluations,
    fallbacks to the plain solve and the largest evaluation count; the fraction of
    the family's cell-stages that ran the solve."""
    if gpu.ice_gpu_set_lk_stats(run.ctx, 1) != 0:
        raise RuntimeError(run.error())
    run.step(steps)
    out = []
    for p in range(1, case.cfg.ngroups + 1):
        cnt = np.zeros(4, np.int64)
        if gpu.ice_gpu_lk_stats(run.ctx, p, cnt) != 0:
            raise RuntimeError(run.error())
        solves = int(cnt[0])
        out.append({"family": p, "solves": solves, "map_evals_per_solve": float(cnt[1]) / max(solves, 1),
                    "fallbacks": int(cnt[2]), "max_evals": int(cnt[3]),
                    "fraction_of_cell_stages": solves / float(case.cells * case.cfg.nrk * steps)})
    gpu.ice_gpu_set_lk_stats(run.ctx, 0)
    return out


def extra_runs(gpu, steps: int, progress):
    """The other headline targets, each on its own workload after the main one:
    the 3-D multi-block IG tree on the batched path, and the 256^3 evaporation
    block with the learned seed against the plain interface solve."""
    out = {}
    case = WORKLOADS["multiblock3d"](native.MODE_BATCHED)
    run = cases.Runner(gpu, case, "ice_gpu_")
    tuned = autotune(gpu, run)
    run.step(1)
    wall = timed_steps(run, steps)
    km, nl, _, _ = kernel_pass(gpu, run)
    avg_s = km / max(nl, 1) / 1e3
    bpc = stage_kernel_bytes(case, True)
    cpl = case.cell_stages(1) / max(nl, 1)
    ach = cpl * bpc / avg_s / 1e9 if avg_s > 0 else None
    ctr = counters_of("multiblock3d", "batched")
    fp64 = None
    if ctr and ctr.get("fp64_flop_per_cell") and avg_s > 0:
        fp64 = ctr["fp64_flop_per_cell"] * cpl / avg_s / 1e12
    out["multiblock3d_batched"] = {"workload": DESCR["multiblock3d"], "steps": steps, "fp64_tflops": fp64,
                                   "traffic_bytes_per_cell": ctr["bytes_per_cell"] if ctr else None,
                                   "value": case.cell_stages(steps) / wall / 1e6, "autotune": tuned,
                                   "kernel_avg_ms": avg_s * 1e3, "bytes_per_cell": bpc, "achieved_gbs": ach,
                                   "frac": ach / HBM_PEAK_GBS if ach else None}
    run.close()
    progress("multi-block tree done")
    ev = {}
    case = WORKLOADS["evap3d"](native.MODE_PLAIN)
    for label, smode in (("plain", 0), ("no_network", 2), ("network", 1)):
        for f in case.fams:
            f.surrogate = smode
        run = cases.Runner(gpu, case, "ice_gpu_")
        if smode == 1:
            rep = np.zeros(6, np.float32)
            if gpu.ice_gpu_surrogate_train(run.ctx, 1, 200, 512, 3e-3, 7, rep) != 0:
                raise RuntimeError(run.error())
            ev["training"] = {"samples": int(rep[0]), "loss_final": float(rep[2]),
                              "rate_ratio_abs_err_mean": float(rep[3]), "train_ms": float(rep[5])}
        # the first step from the shared initial state: the same cells need the solve
        # in every mode, so the deferred-solve kernel's time compares the solves
        _, _, lm, ll = kernel_pass(gpu, run)
        ev[f"{label}_solve_kernel_ms_first_step"] = lm
        ev[f"{label}_value"] = case.cell_stages(steps) / timed_steps(run, steps) / 1e6
        ev[f"{label}_lk"] = lk_stats(gpu, run, case)[0]
        run.close()
    ev["speedup_network_vs_plain"] = ev["network_value"] / ev["plain_value"]
    ev["solve_kernel_speedup_network_vs_plain"] = (ev["plain_solve_kernel_ms_first_step"] /
                                                   max(ev["network_solve_kernel_ms_first_step"], 1e-9))
    ev["workload"] = DESCR["evap3d"]
    out["evap3d_surrogate"] = ev
    progress("evaporation block done")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="combined_full", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="batched", choices=["plain", "batched"])
    ap.add_argument("--surrogate", default="auto", choices=["auto", "on", "off"],
                    help="the learned seed of the interface solve (auto: on where the workload has one)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default stage tile (otherwise the library's autotuner picks it, untimed)")
    ap.add_argument("--tile", default=None,
                    help="stage tile tx,ty,tz (with --dispatch): set instead of autotuned, as for a profile")
    ap.add_argument("--dispatch", type=int, default=None, help="batched dispatch with --tile: 0 queues, 1 items, 2 items family by family")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB three times on the device (the PMC traffic calibration kernel) and exit")
    ap.add_argument("--cpu-steps", type=int, default=50,
                    help="at most this many steps of the CPU baseline sample, cut at about 12 s (0: skip)")
    ap.add_argument("--kernel-steps", type=int, default=1, help="steps of the per-kernel event pass")
    ap.add_argument("--compare-steps", type=int, default=4,
                    help="steps of each comparison run on the same context (0: skip)")
    ap.add_argument("--extras", type=int, default=-1,
                    help="steps of the extra workloads (3-D multi-block batched, evaporation seeded); "
                         "-1: 10 with the default workload, 0: skip")
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

    batched = args.mode == "batched"
    mode = native.MODE_BATCHED if batched else native.MODE_PLAIN
    t_start = time.perf_counter()

    def progress(what):
        print(f"[bench] {time.perf_counter() - t_start:7.1f} s  {what}", file=sys.stderr, flush=True)

    gpu = native.load()
    tuned = None
    if args.tile:
        tuned = {"tile": [int(x) for x in args.tile.split(",")], "dispatch": args.dispatch or 0, "on": "given"}
    elif not args.no_autotune and args.workload in TUNE_PROXY:
        pcase = WORKLOADS[TUNE_PROXY[args.workload]](mode)
        prun = cases.Runner(gpu, pcase, "ice_gpu_")
        tuned = autotune(gpu, prun)
        tuned["on"] = TUNE_PROXY[args.workload]
        prun.close()
        del prun, pcase
        progress(f"autotuned on {TUNE_PROXY[args.workload]}: {tuned['tile']} dispatch {tuned['dispatch']}")
    case = WORKLOADS[args.workload](mode)
    progress(f"case built: {case.cells} cells, {len(case.blocks)} blocks")
    seeded = args.surrogate == "on" or (args.surrogate == "auto" and has_lk(case))
    if seeded and not has_lk(case):
        print(json.dumps({"error": "--surrogate on needs a workload with a Langmuir-Knudsen interface solve"}))
        sys.exit(2)
    for f in case.fams:
        f.surrogate = 1 if seeded else 0
    run = cases.Runner(gpu, case, "ice_gpu_", workers=8)
    used_gb = device_used_gb()
    progress("context set up" + (f" ({used_gb:.1f} GB of device memory in use)" if used_gb else ""))
    surrogate = None
    if seeded:
        surrogate = {"retrain_every_steps": 200, "training": []}
        for p in range(1, case.cfg.ngroups + 1):
            rep = np.zeros(6, np.float32)
            if gpu.ice_gpu_surrogate_train(run.ctx, p, 200, 512, 3e-3, 7, rep) != 0:
                raise RuntimeError(run.error())
            surrogate["training"].append(
                {"family": p, "samples": int(rep[0]), "loss_initial": float(rep[1]), "loss_final": float(rep[2]),
                 "rate_ratio_abs_err_mean": float(rep[3]), "rate_ratio_abs_err_max": float(rep[4]),
                 "train_ms": float(rep[5])})
        if gpu.ice_gpu_surrogate_retrain_every(run.ctx, 200, 20) != 0:
            raise RuntimeError(run.error())
        progress("networks trained")
    if tuned is not None:
        if gpu.ice_gpu_set_tiles(run.ctx, *tuned["tile"]) != 0 or gpu.ice_gpu_set_dispatch(run.ctx, tuned["dispatch"]):
            raise RuntimeError(run.error())
    elif not args.no_autotune:
        tuned = autotune(gpu, run)
        progress("autotuned")
    it0 = C.c_int32()
    gpu.ice_gpu_get_counters(run.ctx, C.byref(it0), None, None)
    if args.warmup:
        run.step(args.warmup)
    progress("warmed up")
    t0 = time.perf_counter()
    st = run.step(args.steps)
    wall = time.perf_counter() - t0
    step_ms = C.c_double()
    gpu.ice_gpu_timing(run.ctx, C.byref(step_ms), None, None)
    step_event_ms = step_ms.value
    progress("timed region done")
    # the dominant kernel, timed live with HIP events on the context's stream
    # (bracketing each launch; a separate pass so the brackets do not slow the timed region)
    kern_ms, launches, lk_ms, lk_launches = kernel_pass(gpu, run, args.kernel_steps)
    cell_stages = case.cell_stages(args.steps)
    value = cell_stages / wall / 1e6
    avg_kernel_s = kern_ms / max(launches, 1) / 1e3
    cells_per_launch = case.cell_stages(args.kernel_steps) / max(launches, 1)
    bpc = stage_kernel_bytes(case, batched)
    bytes_per_launch = cells_per_launch * bpc
    achieved = bytes_per_launch / avg_kernel_s / 1e9 if avg_kernel_s > 0 else None
    lk = lk_stats(gpu, run, case) if has_lk(case) else None
    progress("kernel and interface-solve passes done")

    # the other sides of the metric's comparison, on the same context after the
    # timed region: the batched path with the plain interface solve, then the plain path
    compared = []
    if args.compare_steps > 0:
        legs = []
        if seeded:
            legs.append(("batched, plain interface solve" if batched else "plain, plain interface solve", mode, 0))
        if batched:
            legs.append(("plain", native.MODE_PLAIN, 0))
        for label, m, smode in legs:
            for p in range(1, case.cfg.ngroups + 1):
                gpu.ice_gpu_set_surrogate(run.ctx, p, smode)
            gpu.ice_gpu_set_mode(run.ctx, m)
            run.step(1)   # the graph of the new arguments, or the plain path's first launches
            w = timed_steps(run, args.compare_steps)
            _, _, lm, ll = kernel_pass(gpu, run)
            compared.append({"mode": label, "steps": args.compare_steps,
                             "value": case.cell_stages(args.compare_steps) / w / 1e6,
                             "solve_kernel_ms_per_step": lm if ll else None})
        progress("comparison runs done")
    it1 = C.c_int32()
    gpu.ice_gpu_get_counters(run.ctx, C.byref(it1), None, None)
    run.close()
    del run

    extras = None
    n_extra = args.extras if args.extras >= 0 else (10 if args.workload == "combined_full" else 0)
    if n_extra > 0:
        extras = extra_runs(gpu, n_extra, progress)

    cpu = None
    if args.cpu_steps > 0:
        threads = cpu_threads()
        os.environ["OMP_NUM_THREADS"] = str(threads)
        what, make = CPU_SAMPLE.get(args.workload, ("the same workload", lambda: WORKLOADS[args.workload](native.MODE_PLAIN)))
        sample = make()
        v, secs, nsteps = cpu_baseline(sample, args.cpu_steps)
        cpu = {"value": v, "unit": "Mcell-stage/s", "cores": threads, "kind": "port",
               "sample": f"{nsteps} steps of {what} ({sample.cells} cells) from its initial state, plain interface "
                         f"solve ({secs:.1f} s on {threads} OpenMP threads)"}
        progress("CPU baseline done")

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
        "config": {"workload": DESCR[args.workload] + (
                       "; learned BF16 MFMA seed of the interface solve in the loop, trained before the timed steps "
                       "and retrained every 200 steps (no retraining falls inside the timed steps)" if seeded else ""),
                   "mode": args.mode + ("+surrogate" if seeded else ""), "cells": case.cells,
                   "families": case.cfg.ngroups, "rk_stages": case.cfg.nrk, "device_gb_in_use": used_gb,
                   "event_ms_per_step": step_event_ms / args.steps, "iter_before": it0.value, "iter_after": it1.value,
                   "autotune": tuned, "compared": compared, "interface_solve": lk,
                   "solve_kernel_ms_per_step": (lk_ms / args.kernel_steps) if lk_launches else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "kernel": "k_stage_persist / k_stage_items / k_stage_tiled (fused interior flux + source + "
                               "residual, and the RK update where there is no IRS; the cells that run the interface "
                               "solve finish in k_lk_finish, timed apart)",
                     "kernel_avg_ms": avg_kernel_s * 1e3, "bytes_per_cell": bpc, "launches": launches,
                     "path_bytes_per_cell": path_bytes(case),
                     "path_frac": value * 1e6 * path_bytes(case) / (HBM_PEAK_GBS * 1e9)},
        "cpu_baseline": cpu,
    }
    if surrogate is not None:
        out["surrogate"] = surrogate
    if extras is not None:
        out["extras"] = extras
    ctr = counters_of(args.workload, args.mode)
    if ctr is not None:
        out["roofline"]["traffic"] = ctr["bytes_per_cell"] * cells_per_launch
        if ctr.get("fp64_flop_per_cell") and avg_kernel_s > 0:
            tf = ctr["fp64_flop_per_cell"] * cells_per_launch / avg_kernel_s / 1e12
            out["roofline"]["fp64"] = {"achieved": tf, "peak": 78.6, "unit": "TFLOP/s", "frac": tf / 78.6,
                                       "flop_per_cell": ctr["fp64_flop_per_cell"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
