"""Particles entering through extrapolation faces (400): count_400_inflow and
report_400_inflow (src/lib/diagnostic/Mod_Inflow400.f90:52-171).

The known answers are those of ICE's own test, test/fast/inflow-400/run.py:1-38,
175-202, on its 16-cell duct (cases.case_duct400): a cloud moving along -x enters
through face 2 at iteration 1 with one face in all 40 stages (|w| = 10 m/s for MK,
|w|/sqrt(P/rho) = 2 for IG, |w|/sqrt(Pnn/rho) = 1 for AG); an inlet, empty cells,
a cloud below rho_empty, an inclined empty face, a cloud at rest and IG/AG at rest
up to round-off print nothing; a cloud dragged in by the gas prints a first line at
iteration 1 and a second above 10 times it, in 39 stages.  The multigrid case of
that test has no counterpart here (one grid level).

The CPU tests pin the oracle to those answers; the GPU tests compare the device's
lines and totals with the oracle's, exactly."""
import re

import numpy as np
import pytest

from ice_amd import cases, inflow400
from ice_amd.native import AG, IG, MK

U = 10.0
INLET = (1.0e-2, U, 300.0, 1e-5)   # 402 payload: mass flux, speed, temperature, radius

QUIET = {
    "none": dict(u=U, codes={1: 402, 2: 400, 3: 400, 4: 400, 5: 0, 6: 0}, inlet=INLET),
    "empty": dict(cells=range(5, 13)),
    "sparse": dict(rho0=1e-8),
    "inclined": dict(cells=range(5, 13), tilt=0.5),
    "rest": dict(u=0.0, cells=range(5, 13)),
    "still-ig": dict(closure=IG, u=-1e-15),
    "still-ag": dict(closure=AG, u=-1e-15),
}

# the printed lines of ICE's log, as its test reads them (run.py:47-52)
FIRST = re.compile(r' 400 boundary of block (\d+), face (\d+)(?:, family \d+)?: particles entering '
                   r'at iteration (-?\d+), faces(?: above 10 times its last line)?: (\d+), largest (\S+): (\S+)')
TOTAL = re.compile(r' 400 boundary of block (\d+), face (\d+)(?:, family \d+)?: particles entered in '
                   r'(\d+) stages, first at iteration (-?\d+) \(faces: (\d+), largest (\S+): (\S+?)(?: m/s)?\), '
                   r'largest over the run: (\S+)')


def run(lib, prefix, case):
    r = cases.Runner(lib, case, prefix)
    r.step(case.steps)
    lines, dropped = r.inflow400_lines()
    totals = r.inflow400_totals()
    r.close()
    assert dropped == 0
    return lines, totals


def texts(case, lines, totals):
    cl = [f.closure for f in case.fams]
    ng = case.cfg.ngroups
    return ([inflow400.line_text(x, cl, ng) for x in lines], [inflow400.total_text(t, cl, ng) for t in totals])


def near(text, value):
    return abs(float(text) - value) <= 1e-4 * value


def one_boundary(first, total, label, x, stages):
    assert len(first) == 1 and len(total) == 1, (first, total)
    b, f, it, nf, lab, val = FIRST.match(first[0]).groups()
    assert (b, f, it, nf, lab) == ("1", "2", "1", "1", label) and near(val, x), first
    b, f, ns, it, nf, lab, val, run_max = TOTAL.match(total[0]).groups()
    assert (b, f, ns, it, nf, lab) == ("1", "2", str(stages), "1", "1", label), total
    assert near(val, x) and near(run_max, x), total


ENTERING = [(MK, "|w|", U), (IG, "|w|/sqrt(P/rho)", U / 5.0), (AG, "|w|/sqrt(Pnn/rho)", U / 10.0)]


@pytest.mark.parametrize("closure,label,x", ENTERING, ids=["mk", "ig", "ag"])
def test_oracle_entering(oracle, closure, label, x):
    case = cases.case_duct400(closure)
    first, total = texts(case, *run(oracle, "orc_", case))
    one_boundary(first, total, label, x, 40)


@pytest.mark.parametrize("name", sorted(QUIET))
def test_oracle_quiet(oracle, name):
    case = cases.case_duct400(**QUIET[name])
    lines, totals = run(oracle, "orc_", case)
    assert not lines and not totals, texts(case, lines, totals)


def test_oracle_growth(oracle):
    case = cases.case_duct400(u=0.0, gas_u=-U)
    lines, totals = run(oracle, "orc_", case)
    first, total = texts(case, lines, totals)
    assert len(first) >= 2 and len(total) == 1, first
    f0, f1 = FIRST.match(first[0]).groups(), FIRST.match(first[1]).groups()
    assert f0[:2] == ("1", "2") and f0[2] == "1" and "above 10 times" not in first[0]
    assert "above 10 times its last line" in first[1] and float(f1[5]) > 10.0 * float(f0[5])
    assert TOTAL.match(total[0]).groups()[2] == "39"


def test_text_of_three_digit_exponent():
    """ES12.4E3 where the exponent has three digits (Mod_Inflow400.f90:239-243)"""
    assert inflow400.number_text(AG, 1.234567e150) == "1.2346E+150"
    assert inflow400.number_text(MK, 10.0) == "1.0000E+01 m/s"
    assert inflow400.number_text(IG, 2.5e-100) == "2.5000E-100"


def same(gl, ol):
    assert len(gl) == len(ol)
    for g, o in zip(gl, ol):
        for name, _ in g._fields_:
            assert getattr(g, name) == getattr(o, name), (name, getattr(g, name), getattr(o, name))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1], ids=["plain", "batched"])
@pytest.mark.parametrize("name", ["mk", "ig", "ag", "growth"] + sorted(QUIET))
def test_gpu_lines_equal_oracle(gpu, oracle, name, mode):
    kw = {"mk": {}, "ig": dict(closure=IG), "ag": dict(closure=AG), "growth": dict(u=0.0, gas_u=-U)}.get(name)
    kw = dict(QUIET[name]) if kw is None else kw
    case = cases.case_duct400(mode=mode, **kw)
    gl, gt = run(gpu, "ice_gpu_", case)
    ol, ot = run(oracle, "orc_", case)
    same(gl, ol)
    same(gt, ot)


@pytest.mark.gpu
def test_gpu_log_is_drained_and_counts_persist(gpu):
    """Lines are handed over once; the totals keep counting across reads."""
    case = cases.case_duct400()
    r = cases.Runner(gpu, case, "ice_gpu_")
    r.step(5)
    lines, dropped = r.inflow400_lines()
    assert len(lines) == 1 and dropped == 0 and lines[0].iter == 1
    r.step(5)
    lines, dropped = r.inflow400_lines()
    assert lines == [] and dropped == 0
    (t,) = r.inflow400_totals()
    assert (t.block, t.face, t.stages, t.first_iter) == (1, 2, 20, 1)
    np.testing.assert_equal(t.largest, t.first_largest)
    r.close()


def write_duct_inputs(d, closure_word, mode="plain"):
    """The MK duct in ICE's own input files, for the Fortran host."""
    import os

    from ice_amd import formats
    case = cases.case_duct400()
    os.makedirs(os.path.join(d, "INPUT"), exist_ok=True)
    formats.write_ini(os.path.join(d, "input.ini"), {
        "ICE-Parameters": {"iter-threshold": "20"},
        "ICE-Numerics": {"time-scheme": "RK2", "cfl": "0.5", "time-accurate": ".true.",
                         "space-reconstruction": "MUSCL", "flux-limiter": "vanleer"},
        "ICE-Family1": {"closure": closure_word},
        "ICE-IO": {"shell-diter": "1"},
        "ICE-Device": {"mode": mode}})
    x = np.linspace(0.0, 0.16, 17)
    y = z = np.array([0.0, 0.01])
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    prim = case.blocks[0].prim[0].reshape((6, 20, 5, 5), order="F")[:, 2:18, 2:3, 2:3]
    formats.write_tecplot(os.path.join(d, "INPUT", "part-ic.tec"), ["x", "y", "z"] + [f"v{m}" for m in range(6)],
                          [formats.Zone("B1", (17, 2, 2), np.stack([X, Y, Z]), prim, {})])
    formats.write_bc(os.path.join(d, "INPUT", "part-bc.txt"), case.reci, case.recd, case.donor, case.weight)
    with open(os.path.join(d, "INPUT", "part-properties.dat"), "w") as fh:
        fh.write('TITLE = "Mass Thermodynamic Properties"\nVARIABLES = "Temperature", "Cp", "Density", "Enthalpy"\n')
        fh.write('ZONE T="A"\nI=3, F=POINT\n')
        for T in (1.0, 2.0, 3.0):
            fh.write(f"{T} 900.0 1000.0 {900.0 * T}\n")


def driver_report(stdout):
    lines = [x for x in stdout.split("\n") if x.startswith(" 400 boundary of ")]
    return [x for x in lines if FIRST.match(x)], [x for x in lines if TOTAL.match(x)]


def test_fortran_host_prints_the_report(tmp_path):
    """The Fortran host over the oracle's C-ABI prints the reference's lines."""
    import os

    from test_fortran_host import ABI, run_driver
    d = str(tmp_path / "duct")
    write_duct_inputs(d, "MK")
    p = run_driver(d, {"LD_LIBRARY_PATH": ABI, "OMP_NUM_THREADS": "2"})
    assert p.returncode == 0, p.stdout[-800:] + p.stderr[-800:]
    one_boundary(*driver_report(p.stdout), "|w|", U, 40)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_gpu_fortran_host_prints_the_report(tmp_path, mode):
    from test_fortran_host import run_driver
    d = str(tmp_path / "duct")
    write_duct_inputs(d, "MK", mode)
    p = run_driver(d, {}, timeout=300)
    assert p.returncode == 0, p.stdout[-800:] + p.stderr[-800:]
    one_boundary(*driver_report(p.stdout), "|w|", U, 40)


def two_family_duct(mode=0):
    """The duct with an MK and an IG family side by side (one table for both)"""
    mk, ig = cases.case_duct400(MK, mode=mode), cases.case_duct400(IG, mode=mode)
    mk.cfg.ngroups = 2
    tb = cases.TableBuilder(2)
    codes = {1: 400, 2: 400, 3: 400, 4: 400, 5: 0, 6: 0}

    def normal_dir(ri, fijk):
        ri[:, cases.R_NORMAL_DIR] = 1

    tb.add_block(1, mk.blocks[0].dims, lambda f, i, j, k: np.vectorize(codes.get)(f), normal_dir)
    reci, recd, donor, weight = tb.arrays()
    blk = cases.Block(mk.blocks[0].dims, mk.blocks[0].geo, [mk.blocks[0].prim[0], ig.blocks[0].prim[0]])
    return cases.Case("duct400x2", mk.cfg, [mk.fams[0], ig.fams[0]], [blk], reci, recd, donor, weight, mk.steps)


def test_oracle_two_families_name_each_family(oracle):
    case = two_family_duct()
    first, total = texts(case, *run(oracle, "orc_", case))
    assert len(first) == 2 and len(total) == 2, first
    assert first[0].startswith(" 400 boundary of block 1, face 2, family 1: particles entering at iteration 1")
    assert first[1].startswith(" 400 boundary of block 1, face 2, family 2: particles entering at iteration 1")
    assert "|w|: 1.0000E+01 m/s" in first[0] and "|w|/sqrt(P/rho): 2.0000E+00" in first[1]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1], ids=["plain", "batched"])
def test_gpu_two_families_equal_oracle(gpu, oracle, mode):
    """plain: the reference's order; batched: the same lines, interleaved by stage"""
    case = two_family_duct(mode)
    gl, gt = run(gpu, "ice_gpu_", case)
    ol, ot = run(oracle, "orc_", case)
    key = lambda x: (x.family, x.iter, x.block, x.face)  # noqa: E731
    if mode == 0:
        same(gl, ol)
    else:
        same(sorted(gl, key=key), sorted(ol, key=key))
    same(gt, ot)
