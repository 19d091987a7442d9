"""Writes tests/golden/refcase_<name>.npz from the reference's own invariant-checked
cases: test/Axis/{IG,MK,AG} (a one-degree wedge about x: the curvilinear mesh,
the 200 side faces and the 300 axis row), test/Reflect/AG (a sheared AG blob
along a 200 plane) and test/Thermal/{IG,AG} (a pressurised blob in a closed
box).  Their verify.py scripts pass on properties of the solution, not on a
stored field, so the fixtures hold the inputs only.  Run where the reference is
mounted:

    python tests/golden/make_reference_cases.py [/root/reference]

Each fixture holds data: the run settings of input.ini, the block's nodes
(3, ni+1, nj+1, nk+1) (the wedge is not a tensor product), the initial
primitive cells, the parsed boundary table, and the material's density and
heat capacity from [ICE-Physics].
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ice_amd import formats  # noqa: E402

CASES = ["Axis/IG", "Axis/MK", "Axis/AG", "Reflect/AG", "Thermal/IG", "Thermal/AG"]


def main(ref):
    for c in CASES:
        d = os.path.join(ref, "test", c)
        ini = formats.read_ini(os.path.join(d, "input.ini"))
        _, ic = formats.read_tecplot(os.path.join(d, "INPUT", "part-ic.tec"))
        if len(ic) != 1:
            raise SystemExit(f"{c}: one block expected")
        dims = [z.cell_dims for z in ic]
        reci, recd, donor, weight = formats.read_bc(os.path.join(d, "INPUT", "part-bc.txt"), dims, 1)
        phys = ini.get("ICE-Physics", {})
        arrs = {"ini": np.array(json.dumps(ini)), "reci": reci, "recd": recd, "donor": donor, "weight": weight,
                "rho_mat": np.array(float(phys["density"])), "cs": np.array(float(phys["specific-heat"])),
                "nodes": ic[0].nodes, "ic": ic[0].cells}
        name = c.replace("/", "_")
        np.savez_compressed(os.path.join(HERE, f"refcase_{name}.npz"), **arrs)
        print(c, dims, "records", reci.shape[0], "codes", sorted(set(reci[:, 5].tolist())))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
