"""The Fortran host (ice_amd/fortran): ICE's driver loop, INI registry, input
readers and block metrics in Fortran, calling the path through the
ISO_C_BINDING module ice_gpu_m.  It is run on the reference's Doisneau cases,
written out in ICE's own file formats (input.ini, INPUT/part-ic.tec,
INPUT/part-bc.txt, INPUT/part-properties.dat), and its OUTPUT/part-field.tec
is compared with the field the reference stored.  On a host without a GPU it
runs over the oracle's C-ABI (oracle/_abi/libice_gpu.so, the same symbols)."""
import json
import os
import subprocess

import numpy as np
import pytest

from ice_amd import formats

from helpers import GOLDEN, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ice_amd", "fortran", "ice_driver")
ABI = os.path.join(ROOT, "oracle", "_abi")


def write_case(d, name, mode="plain", extra=None):
    z = np.load(os.path.join(GOLDEN, f"doisneau_{name}.npz"))
    nb = int(z["nblocks"])
    os.makedirs(os.path.join(d, "INPUT"), exist_ok=True)
    ini = json.loads(str(z["ini"]))
    ini.setdefault("ICE-Device", {})["mode"] = mode
    for sec, kv in (extra or {}).items():
        ini.setdefault(sec, {}).update(kv)
    formats.write_ini(os.path.join(d, "input.ini"), ini)
    zones = []
    for b in range(1, nb + 1):
        x, y, zz = z[f"x{b}"], z[f"y{b}"], z[f"z{b}"]
        X, Y, Z = np.meshgrid(x, y, zz, indexing="ij")
        zones.append(formats.Zone(f"B{b}", (x.size, y.size, zz.size), np.stack([X, Y, Z]), z[f"ic{b}"], {}))
    nc = zones[0].cells.shape[0]
    formats.write_tecplot(os.path.join(d, "INPUT", "part-ic.tec"), ["x", "y", "z"] + [f"v{m}" for m in range(nc)],
                          zones)
    formats.write_bc(os.path.join(d, "INPUT", "part-bc.txt"), z["reci"], z["recd"], z["donor"], z["weight"])
    rho, cp = float(z["rho_mat"]), float(z["cs"])
    with open(os.path.join(d, "INPUT", "part-properties.dat"), "w") as fh:
        fh.write('TITLE = "Mass Thermodynamic Properties"\nVARIABLES = "Temperature", "Cp", "Density", "Enthalpy"\n')
        fh.write('ZONE T="A"\nI=3, F=POINT\n')
        for T in (1.0, 2.0, 3.0):
            fh.write(f"{T} {cp} {rho} {cp * T}\n")
    return [z[f"ref{b}"] for b in range(1, nb + 1)]


def run_driver(d, env_extra, timeout=600):
    if not os.path.exists(DRIVER):
        pytest.fail("ice_amd/fortran/ice_driver is missing: run __graft_entry__.build()")
    env = dict(os.environ, **env_extra)
    return subprocess.run([DRIVER, d], env=env, capture_output=True, text=True, timeout=timeout)


def compare(d, ref):
    _, out = formats.read_tecplot(os.path.join(d, "OUTPUT", "part-field.tec"))
    g = np.concatenate([np.stack([z.cells[0], z.cells[1], z.cells[2], z.cells[-1]]).reshape(4, -1) for z in out], 1)
    r = np.concatenate([q.reshape(4, -1) for q in ref], axis=1)
    return rel_l2(g[[0, 3]], r[[0, 3]]), rel_l2(g, r)


def test_driver_on_oracle_abi(tmp_path):
    """IG-chimera through the Fortran host over the oracle: 5000 RK2 iterations."""
    d = str(tmp_path / "case")
    ref = write_case(d, "IG-chimera")
    p = run_driver(d, {"LD_LIBRARY_PATH": ABI, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS", "8")})
    print(p.stdout[-600:], p.stderr[-600:])
    assert p.returncode == 0
    assert "iter    5000" in p.stdout
    rn, worst = compare(d, ref)
    print("rho/n rel L2", rn, "worst", worst)
    assert rn <= 1e-12 and worst <= 1e-12


def test_driver_refuses_unknown_key(tmp_path):
    d = str(tmp_path / "case")
    write_case(d, "IG", extra={"ICE-Numerics": {"cfl-typo": "0.5"}})
    p = run_driver(d, {"LD_LIBRARY_PATH": ABI})
    assert p.returncode != 0
    assert "[ERROR] [ICE::Read_Inifile]" in p.stderr and "cfl-typo" in p.stderr


def test_driver_refuses_irs_time_accurate(tmp_path):
    d = str(tmp_path / "case")
    write_case(d, "IG", extra={"ICE-Numerics": {"irs": ".true.", "time-accurate": ".true."}})
    p = run_driver(d, {"LD_LIBRARY_PATH": ABI})
    assert p.returncode != 0 and "time-accurate = false" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "batched"])
@pytest.mark.parametrize("name", ["IG-split4", "IG-chimera", "AG"])
def test_driver_on_gpu(tmp_path, name, mode):
    print(f"case fortran driver {name} {mode}", flush=True)
    d = str(tmp_path / "case")
    ref = write_case(d, name, mode, extra={"ICE-Device": {"autotune": ".true."}} if name == "IG-split4" else None)
    p = run_driver(d, {}, timeout=300)
    print(p.stdout[-400:], p.stderr[-400:])
    assert p.returncode == 0
    rn, worst = compare(d, ref)
    print("rho/n rel L2", rn, "worst", worst)
    assert rn <= 1e-12 and worst <= (1e-8 if name == "AG" else 1e-12)
