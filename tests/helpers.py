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


GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)), "golden")


def doisneau(name: str, mode: int = 0):
    """(Case, reference fields per block (rho, u, v, n), iterations) of one of the
    reference's Doisneau validation cases, from its committed fixture."""
    import json
    import os
    z = np.load(os.path.join(GOLDEN, f"doisneau_{name}.npz"))
    nb = int(z["nblocks"])
    from ice_amd import geometry
    nodes = [geometry.cartesian_nodes(z[f"x{b}"], z[f"y{b}"], z[f"z{b}"]) for b in range(1, nb + 1)]
    ic = [z[f"ic{b}"] for b in range(1, nb + 1)]
    iters = int(z["iterations"])
    case = cases.case_from_inputs(f"doisneau_{name}", json.loads(str(z["ini"])), nodes, ic, z["reci"], z["recd"],
                                  z["donor"], z["weight"], float(z["rho_mat"]), float(z["cs"]), steps=iters,
                                  mode=mode)
    return case, [z[f"ref{b}"] for b in range(1, nb + 1)], iters


def doisneau_errors(runner, case, ref):
    """verify.py's two measures (test/Doisneau/IG/verify.py:43-61) over all blocks:
    the RMS density difference and the number density's relative L2 difference;
    plus the largest relative L2 difference of rho, u, v, n."""
    da, dn, sa, sn, n = 0.0, 0.0, 0.0, 0.0, 0
    worst = 0.0
    got_all, ref_all = [], []
    for b, r in enumerate(ref, 1):
        s = runner.state(b, 1)
        g = np.stack([s[0], s[1], s[2], s[-1]])
        got_all.append(g.reshape(4, -1))
        ref_all.append(r.reshape(4, -1))
    g = np.concatenate(got_all, axis=1)
    r = np.concatenate(ref_all, axis=1)
    rms_rho = float(np.sqrt(np.mean((g[0] - r[0]) ** 2)))
    rel_n = float(np.linalg.norm(g[3] - r[3]) / np.linalg.norm(r[3]))
    return rms_rho, rel_n, rel_l2(g, r), float(np.sqrt(np.mean(r[0] ** 2)))
