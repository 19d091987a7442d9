"""Writes tests/golden/refcase_<name>.npz from the reference's own invariant-checked
cases: test/Axis/{IG,MK,AG} (a one-degree wedge about x: the curvilinear mesh,
the 200 side faces and the 300 axis row), test/Reflect/AG (a sheared AG blob
along a 200 plane), test/Thermal/{IG,AG} (a pressurised blob in a closed box),
test/Riemann/AG (an anisotropic Sod tube, checked against the exact gamma = 3
solution) and test/Berthon/{SCS,RCS,RCR,SCS-hlle} (AG Riemann problems, with
the exact profiles the reference tabulates in Results/<case>/<case>_exact.tec).
Their verify.py scripts pass on properties of the solution, not on a stored
field, so the fixtures hold the inputs (and, for Berthon, the exact profiles).
Run where the reference is mounted:

    python tests/golden/make_reference_cases.py [/root/reference]

Each fixture holds data: the run settings of input.ini, the block's nodes
(3, ni+1, nj+1, nk+1) (the wedge is not a tensor product), the initial
primitive cells, the parsed boundary table, and the material's density and
heat capacity from [ICE-Physics].
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ice_amd import formats  # noqa: E402

CASES = ["Axis/IG", "Axis/MK", "Axis/AG", "Reflect/AG", "Thermal/IG", "Thermal/AG", "Riemann/AG",
         "Berthon/SCS", "Berthon/RCS", "Berthon/RCR", "Berthon/SCS-hlle"]


def material(d, ini):
    """rho and cp of the material: [ICE-Physics] density and specific-heat, else the
    constant columns of INPUT/part-properties.dat (T, cp, rho, ...)"""
    phys = ini.get("ICE-Physics", {})
    if "density" in phys:
        return float(phys["density"]), float(phys["specific-heat"])
    rows = []
    with open(os.path.join(d, "INPUT", "part-properties.dat")) as fh:
        for line in fh:
            try:
                rows.append([float(v) for v in line.split()])
            except ValueError:
                continue
    a = np.array([r for r in rows if len(r) == 4])
    if not (np.all(a[:, 1] == a[0, 1]) and np.all(a[:, 2] == a[0, 2])):
        raise SystemExit(f"{d}: a property table that is not constant")
    return float(a[0, 2]), float(a[0, 1])


def read_exact(path):
    """The piecewise-linear exact profiles of a Berthon case, one Tecplot zone each:
    {zone title: (x, value)}"""
    lines = open(path).read().splitlines()
    zones, i = {}, 0
    while i < len(lines):
        m = re.search(r'ZONE\s+T\s*=\s*"([^"]+)"', lines[i])
        if not m:
            i += 1
            continue
        n = int(re.search(r"\bI\s*=\s*(\d+)", lines[i + 1]).group(1))
        i += 3
        vals = []
        while len(vals) < 2 * n:
            vals += [float(t) for t in lines[i].split()]
            i += 1
        zones[m.group(1)] = (np.array(vals[:n]), np.array(vals[n:]))
    return zones


def main(ref):
    for c in CASES:
        d = os.path.join(ref, "test", c)
        ini = formats.read_ini(os.path.join(d, "input.ini"))
        _, ic = formats.read_tecplot(os.path.join(d, "INPUT", "part-ic.tec"))
        if len(ic) != 1:
            raise SystemExit(f"{c}: one block expected")
        dims = [z.cell_dims for z in ic]
        reci, recd, donor, weight = formats.read_bc(os.path.join(d, "INPUT", "part-bc.txt"), dims, 1)
        rho_mat, cs = material(d, ini)
        arrs = {"ini": np.array(json.dumps(ini)), "reci": reci, "recd": recd, "donor": donor, "weight": weight,
                "rho_mat": np.array(rho_mat), "cs": np.array(cs), "nodes": ic[0].nodes, "ic": ic[0].cells}
        if c.startswith("Berthon/"):
            base = c.split("/")[1].replace("-hlle", "")
            ex = read_exact(os.path.join(ref, "test", "Berthon", "Results", base, f"{base}_exact.tec"))
            arrs["exact_zones"] = np.array(json.dumps(sorted(ex)))
            for z, (x, v) in ex.items():
                arrs[f"exact_x_{z}"], arrs[f"exact_v_{z}"] = x, v
        name = c.replace("/", "_")
        np.savez_compressed(os.path.join(HERE, f"refcase_{name}.npz"), **arrs)
        print(c, dims, "records", reci.shape[0], "codes", sorted(set(reci[:, 5].tolist())))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
