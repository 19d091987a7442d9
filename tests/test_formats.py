"""The readers of ICE's input files (ice_amd/formats.py): the ATLAS bc.txt
parser's refusals (Setup_BC, src/lib/io/IO_BC.f90:144-326), Tecplot BLOCK
round trips, and, where the reference is mounted, that parsing its Doisneau
inputs gives the committed fixtures."""
import os

import numpy as np
import pytest

from ice_amd import formats
from ice_amd.native import (R_DONOR0, R_NCHI1, R_NCHI2, R_NORMAL_DIR, R_TYPE, RD_ALPHA, RD_MASSFLUX, RD_RADIUS,
                            RD_TEMPERATURE)

from helpers import GOLDEN

REF = "/root/reference/test/Doisneau"


def bc_file(tmp_path, body):
    p = tmp_path / "bc.txt"
    p.write_text(body)
    return str(p)


def all_cells(dims):
    ni, nj, nk = dims
    for f in range(1, 7):
        if f <= 2:
            for k in range(1, nk + 1):
                for j in range(1, nj + 1):
                    yield f, (1 if f == 1 else ni), j, k
        elif f <= 4:
            for k in range(1, nk + 1):
                for i in range(1, ni + 1):
                    yield f, i, (1 if f == 3 else nj), k
        else:
            for j in range(1, nj + 1):
                for i in range(1, ni + 1):
                    yield f, i, j, (1 if f == 5 else nk)


def write_block(dims, special=None):
    lines = []
    for f, i, j, k in all_cells(dims):
        s = (special or {}).get((f, i, j, k))
        if s is None:
            lines.append(f"1 {i} {j} {k} {f} 400")
        else:
            lines.extend(s(i, j, k, f))
    return "\n".join(lines) + "\n"


def test_bc_payloads(tmp_path):
    dims = (3, 2, 1)
    special = {
        (1, 1, 1, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 402", "0.5 10.0 normal normal 300.0 1e-6 0 Dirac 0"],
        (1, 1, 2, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 402", "0.5 10.0 0.3 0.0 300.0 1e-6 0 Dirac 0"],
        (2, 3, 1, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 102", "1 2", "1 1 1 1 1.0", "1 2 1 1 0.25",
                                          "1 2 2 1 0.75"],
    }
    reci, recd, donor, weight = formats.read_bc(bc_file(tmp_path, write_block(dims, special)), [dims], 2)
    ncell = 2 * (2 * 1 + 3 * 1 + 3 * 2)
    assert reci.shape[0] == 2 * ncell                          # fanned out over two families
    assert list(reci[0:2, R_TYPE]) == [402, 402] and list(reci[0:2, 6]) == [1, 2]
    assert reci[0, R_NORMAL_DIR] == 1 and reci[2, R_NORMAL_DIR] == 0
    assert recd[2, RD_ALPHA] == 0.3 and recd[0, RD_MASSFLUX] == 0.5
    c = [n for n in range(reci.shape[0]) if reci[n, R_TYPE] == 102]
    assert len(c) == 2 and reci[c[0], R_NCHI1] == 1 and reci[c[0], R_NCHI2] == 2 and reci[c[0], R_DONOR0] == 0
    assert donor.tolist() == [[1, 1, 1, 1], [1, 2, 1, 1], [1, 2, 2, 1]] and weight.tolist() == [1.0, 0.25, 0.75]


def test_bc_payload_commas(tmp_path):
    """The payload is read list-directed (IO_BC.f90:261-265), where a comma
    separates values as a blank does: ICE's Axis/MK case writes 'normal, normal,'."""
    dims = (3, 2, 1)
    special = {
        (1, 1, 1, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 402", "0.5 10.0 normal, normal, 300.0 1e-6"],
        (1, 1, 2, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 402", "0.5,10.0,0.3,0.0,300.0,1e-6"],
    }
    reci, recd, _, _ = formats.read_bc(bc_file(tmp_path, write_block(dims, special)), [dims], 1)
    assert reci[0, R_NORMAL_DIR] == 1 and recd[0, RD_MASSFLUX] == 0.5 and recd[0, RD_RADIUS] == 1e-6
    assert reci[1, R_NORMAL_DIR] == 0 and recd[1, RD_ALPHA] == 0.3 and recd[1, RD_TEMPERATURE] == 300.0


@pytest.mark.parametrize("bad, msg", [
    ({(1, 1, 1, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 55"]}, "bad BC code"),
    ({(1, 1, 1, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 777"]}, "unsupported"),
    ({(1, 1, 1, 1): lambda i, j, k, f: [f"1 2 {j} {k} {f} 400"]}, "not on face"),
    ({(1, 1, 1, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 402", "0.5 10 normal 0.0 300 1e-6"]}, "both normal"),
    ({(1, 1, 1, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 402", "0.5 0.0 0.1 0.0 300 1e-6"]}, "zero velocity"),
    ({(1, 1, 1, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 102", "1 1", "1 1 1 1 0.1", "1 2 1 1 1.0"]},
     "sum to ~0"),
    ({(1, 1, 1, 1): lambda i, j, k, f: [f"1 {i} {j} {k} {f} 401", "0.5 1.0 0.1 0.0 1.0 1e-6 0 Dirac 0"]}, "10th"),
])
def test_bc_refusals(tmp_path, bad, msg):
    with pytest.raises(formats.FormatError, match=msg):
        formats.read_bc(bc_file(tmp_path, write_block((3, 2, 1), bad)), [(3, 2, 1)], 1)


def test_bc_truncated(tmp_path):
    body = write_block((3, 2, 1)).split("\n")
    with pytest.raises(formats.FormatError, match="ends before"):
        formats.read_bc(bc_file(tmp_path, "\n".join(body[:-3])), [(3, 2, 1)], 1)


def test_tecplot_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    z = formats.Zone("b", (4, 3, 2), rng.random((3, 4, 3, 2)), rng.random((5, 3, 2, 1)), {})
    p = str(tmp_path / "f.tec")
    formats.write_tecplot(p, ["x", "y", "z", "a", "b", "c", "d", "e"], [z, z])
    names, zs = formats.read_tecplot(p)
    assert names == ["x", "y", "z", "a", "b", "c", "d", "e"] and len(zs) == 2
    for q in zs:
        assert np.array_equal(q.nodes, z.nodes) and np.array_equal(q.cells, z.cells)


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's Doisneau inputs are not mounted")
@pytest.mark.parametrize("name", ["IG", "IG-chimera", "IG-split4"])
def test_parsing_reference_inputs_gives_fixture(name):
    z = np.load(os.path.join(GOLDEN, f"doisneau_{name}.npz"))
    _, ic = formats.read_tecplot(os.path.join(REF, name, "INPUT", "part-ic.tec"))
    reci, recd, donor, weight = formats.read_bc(os.path.join(REF, name, "INPUT", "part-bc.txt"),
                                                [q.cell_dims for q in ic], 1)
    assert np.array_equal(reci, z["reci"]) and np.array_equal(recd, z["recd"])
    assert np.array_equal(donor, z["donor"]) and np.array_equal(weight, z["weight"])
