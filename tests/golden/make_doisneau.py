"""Writes tests/golden/doisneau_<case>.npz from the reference's own Doisneau
validation cases (test/Doisneau/{MK,IG,AG,IG-chimera,IG-split4}: input.ini,
INPUT/part-ic.tec, INPUT/part-bc.txt, INPUT/part-properties.dat and the stored
reference/part-field.tec, written by the reference after 5000 steady RK2
iterations).  Run in a container where the reference is mounted:

    python tests/golden/make_doisneau.py [/root/reference]

Each fixture holds data only: the run settings of input.ini, the node axes of
each block (the meshes are tensor products), the initial primitive cells, the
parsed boundary table, the material's density and heat capacity from the
property table, and the reference's final density, u, v and number density.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ice_amd import formats  # noqa: E402

CASES = ["MK", "IG", "AG", "IG-chimera", "IG-split4"]


def KEEP(nvar):
    """the reference fields kept: density, u, v and number density"""
    return [0, 1, 2, nvar - 1]


def axes(nodes):
    x, y, z = nodes[0][:, 0, 0], nodes[1][0, :, 0], nodes[2][0, 0, :]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    if not (np.array_equal(X, nodes[0]) and np.array_equal(Y, nodes[1]) and np.array_equal(Z, nodes[2])):
        raise SystemExit("a mesh that is not a tensor product: store its nodes")
    return x, y, z


def material(path):
    """rho and cp of a property table whose Cp and Density columns are constant"""
    rows = []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            try:
                rows.append([float(v) for v in t])
            except ValueError:
                continue
    a = np.array([r for r in rows if len(r) == 4])
    if not (np.all(a[:, 1] == a[0, 1]) and np.all(a[:, 2] == a[0, 2])):
        raise SystemExit(f"{path}: a property table that is not constant")
    return float(a[0, 2]), float(a[0, 1])


def main(ref):
    for c in CASES:
        d = os.path.join(ref, "test", "Doisneau", c)
        ini = formats.read_ini(os.path.join(d, "input.ini"))
        _, ic = formats.read_tecplot(os.path.join(d, "INPUT", "part-ic.tec"))
        names, out = formats.read_tecplot(os.path.join(d, "reference", "part-field.tec"))
        dims = [z.cell_dims for z in ic]
        reci, recd, donor, weight = formats.read_bc(os.path.join(d, "INPUT", "part-bc.txt"), dims, 1)
        rho_mat, cs = material(os.path.join(d, "INPUT", "part-properties.dat"))
        arrs = {"ini": np.array(json.dumps(ini)), "nblocks": np.array(len(ic)), "reci": reci, "recd": recd,
                "donor": donor, "weight": weight, "rho_mat": np.array(rho_mat), "cs": np.array(cs),
                "ref_names": np.array(json.dumps([names[3 + v] for v in KEEP(len(names) - 3)])),
                "iterations": np.array(max(int(float(z.aux.get("ICE_ITERATION", "0"))) for z in out))}
        for b, (zi, zo) in enumerate(zip(ic, out), 1):
            if not np.array_equal(zi.nodes, zo.nodes):
                raise SystemExit(f"{c}: block {b} of the solution is not on the mesh of the initial condition")
            x, y, z = axes(zi.nodes)
            arrs[f"x{b}"], arrs[f"y{b}"], arrs[f"z{b}"] = x, y, z
            arrs[f"ic{b}"] = zi.cells
            arrs[f"ref{b}"] = zo.cells[KEEP(zo.cells.shape[0])]
        np.savez_compressed(os.path.join(HERE, f"doisneau_{c}.npz"), **arrs)
        print(c, [z.cell_dims for z in ic], "records", reci.shape[0], "donors", donor.shape[0],
              "iterations", int(arrs["iterations"]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
