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


def conserved(q: np.ndarray, fam, tab=None) -> np.ndarray:
    """The conservative state (Lib_Model_*.f90 prim_2_cons) of primitive cells q
    (nc, ...) of a family: rho, rho u, and per closure the energy rho e + kinetic
    part (with e = cs T, or the property table's e(T)), IG's rho |u|^2 + 3P, AG's
    rho u u + P; the number density."""
    from ice_amd import native, properties
    rho, u = q[0], q[1:4]
    nc = q.shape[0]
    T = q[nc - 2]
    if tab is not None and tab.flags & properties.CS_VARIES:
        e = properties.table_e(tab, T)
    else:
        e = (tab.col(properties.CP)[0] if tab is not None else fam.cs) * T
    out = [rho, rho * u[0], rho * u[1], rho * u[2]]
    if fam.closure == native.MK:
        out += [rho * e + 0.5 * rho * np.sum(u * u, axis=0)]
    elif fam.closure == native.IG:
        c4 = rho * np.sum(u * u, axis=0) + 3.0 * q[4]
        out += [c4, rho * e + 0.5 * c4]
    else:
        P = q[4:10]
        uu = [u[0] * u[0], u[0] * u[1], u[0] * u[2], u[1] * u[1], u[1] * u[2], u[2] * u[2]]
        c = [rho * uu[m] + P[m] for m in range(6)]
        out += c + [rho * e + 0.5 * (c[0] + c[3] + c[5])]
    return np.stack(out + [q[nc - 1]])


def compare_states(rg, ro, case):
    """The largest relative L2 difference, over blocks, families and conserved
    variables, between two runs of a case (the bar of the parity tests is on the
    conservative state)."""
    worst = 0.0
    tabs = case.meta.get("tables", {})
    for b in range(1, len(case.blocks) + 1):
        for p in range(1, case.cfg.ngroups + 1):
            fam, tab = case.fams[p - 1], tabs.get(p)
            a, o = rg.state(b, p), ro.state(b, p)
            worst = max(worst, rel_l2(conserved(a, fam, tab), conserved(o, fam, tab)), rel_l2(a, o))
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
