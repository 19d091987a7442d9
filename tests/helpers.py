"""Comparison helpers for the parity tests."""
import numpy as np

from ice_amd import cases


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """largest relative L2 difference over the variables (first axis)"""
    worst = 0.0
    for v in range(a.shape[0]):
        den = np.linalg.norm(b[v])
        num = np.linalg.norm(a[v] - b[v])
        worst = max(worst, num / den if den > 0 else num)
    return worst


def run_both(gpu, oracle, case, steps):
    rg = cases.Runner(gpu, case, "ice_gpu_")
    ro = cases.Runner(oracle, case, "orc_")
    sg = rg.step(steps)
    so = ro.step(steps)
    return rg, ro, sg, so


def compare_states(rg, ro, case):
    worst = 0.0
    for b in range(1, len(case.blocks) + 1):
        for p in range(1, case.cfg.ngroups + 1):
            worst = max(worst, rel_l2(rg.state(b, p), ro.state(b, p)))
    return worst
