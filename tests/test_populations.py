"""Multi-population boundary files (ICE's test/fast/populations/run.py): ATLAS
writes one copy of the boundary table per (material, population) pair, block by
block, and copy c belongs to family c (IO_BC.f90:21-64, 144-200).  Restated on a
16 x 4 MK duct with a 402 inlet on face 1 and 30 RK2 steps:

* proportional: three families, copies whose mass fluxes are 1, 2 and 4 times a
  reference, and an initial loading scaled the same way; the equations are linear
  in the loading at fixed velocity and temperature, so family p comes out 2^(p-1)
  times family 1 in rho and n, and equal to it in u, v, w and T, to 1e-12;
* shared: one copy, three families: the record speaks for all of them, and the
  families stay equal bit for bit;
* mismatch: two copies for three families is refused, as the reference refuses it.
"""
import os

import numpy as np
import pytest

from ice_amd import cases, formats, geometry, native
from ice_amd.cases import Block, Case, TableBuilder, make_config, make_family, padded_state
from ice_amd.native import MK, R_NORMAL_DIR, RD_MASSFLUX, RD_RADIUS, RD_TEMPERATURE, RD_VELOCITY

NX, NY, LX, LY = 16, 4, 0.16, 0.04
GP0, VINJ, TINJ, RINJ, RHO_AL, CS = 2.0, 10.0, 300.0, 1e-5, 1000.0, 900.0
RHO0 = 1e-3
N0 = RHO0 / (RHO_AL * 4.0 / 3.0 * np.pi * RINJ ** 3)


def duct(tmp_path, copies, families, scaled, mode=0):
    """The duct with `families` MK families and its boundary file written with
    `copies` copies (one per family, or one for all), then read back."""
    dims = (NX, NY, 1)
    nodes = geometry.cartesian_nodes(np.linspace(0.0, LX, NX + 1), np.linspace(0.0, LY, NY + 1), [0.0, LX / NX])
    geo = geometry.metrics(nodes)
    scale = [2.0 ** p if scaled else 1.0 for p in range(families)]

    def fill(s):
        def f(I, J, K):
            z = np.zeros(I.shape)
            return np.stack([z + RHO0 * s, z, z, z, z + TINJ, z + N0 * s])
        return f

    prims = [padded_state(dims, 6, fill(s)) for s in scale]
    codes = {1: 402, 2: 400, 3: 400, 4: 400, 5: 0, 6: 0}
    tb = TableBuilder(1)

    def payload(rd, fijk):
        on = fijk[:, 0] == 1
        rd[on, RD_MASSFLUX], rd[on, RD_VELOCITY], rd[on, RD_TEMPERATURE], rd[on, RD_RADIUS] = GP0, VINJ, TINJ, RINJ

    def normal_dir(ri, fijk):
        ri[:, R_NORMAL_DIR] = 1

    tb.add_block(1, dims, lambda f, i, j, k: np.vectorize(codes.get)(f), normal_dir, payload)
    ri1, rd1, dn, w = tb.arrays()
    # the table of `copies` families: copy c carries the mass flux 2^(c-1) GP0 (scaled)
    ri = np.repeat(ri1, copies, axis=0)
    rd = np.repeat(rd1, copies, axis=0)
    ri[:, native.R_P] = np.tile(np.arange(1, copies + 1, dtype=np.int32), ri1.shape[0])
    for c in range(copies):
        rd[c::copies, RD_MASSFLUX] *= 2.0 ** c if scaled else 1.0
    path = os.path.join(tmp_path, "part-bc.txt")
    formats.write_bc(path, ri, rd, dn, w, ngroups=copies, copies=copies > 1)
    reci, recd, donor, weight = formats.read_bc(path, [dims], families)
    cfg = make_config(ngroups=families, nrk=2, time_accurate=True, ndir=2, limiter=3, cfl=0.8, dt_max=1e-4,
                      mode=mode)
    fams = [make_family(MK, rho_mat=RHO_AL, cs=CS) for _ in range(families)]
    return Case("populations", cfg, fams, [Block(dims, geo, prims)], reci, recd, donor, weight, 30)


def test_copies_reach_their_families(tmp_path):
    case = duct(tmp_path, copies=3, families=3, scaled=True)
    on = case.reci[:, native.R_TYPE] == 402
    for p in (1, 2, 3):
        mine = on & (case.reci[:, native.R_P] == p)
        assert mine.sum() == NY and np.all(case.recd[mine, RD_MASSFLUX] == GP0 * 2.0 ** (p - 1))


def test_copy_count_mismatch_is_refused(tmp_path):
    with pytest.raises(formats.FormatError, match="copies of the boundary table"):
        duct(tmp_path, copies=2, families=3, scaled=False)


def test_copy_out_of_order_is_refused(tmp_path):
    duct(tmp_path, copies=2, families=2, scaled=True)
    path = os.path.join(tmp_path, "part-bc.txt")
    lines = open(path).read().split("\n")
    # swap the first two records of copy 2 (each 402 record is a header and a payload line)
    recs, k = [], 0
    while k < len(lines) and lines[k].strip():
        n = 2 if lines[k].split()[5] == "402" else 1
        recs.append(lines[k:k + n])
        k += n
    ncell = 2 * (NY * 1 + NX * 1 + NX * NY)
    recs[ncell], recs[ncell + 1] = recs[ncell + 1], recs[ncell]
    with open(path, "w") as fh:
        fh.write("\n".join(sum(recs, [])) + "\n")
    with pytest.raises(formats.FormatError, match="does not repeat the cells"):
        formats.read_bc(path, [(NX, NY, 1)], 2)


def check_proportional(r):
    f1 = r.state(1, 1)
    assert f1[0].max() > RHO0 * 1.01                 # the inlet has injected
    for p in (2, 3):
        fp = r.state(1, p)
        s = 2.0 ** (p - 1)
        for v in (0, 5):
            assert np.max(np.abs(fp[v] - s * f1[v])) <= 1e-12 * np.max(np.abs(s * f1[v]))
        for v in (1, 2, 3, 4):
            assert np.max(np.abs(fp[v] - f1[v])) <= 1e-12 * max(np.max(np.abs(f1[v])), 1e-300)


def check_shared(r):
    f1 = r.state(1, 1)
    for p in (2, 3):
        assert np.array_equal(r.state(1, p), f1)


def test_oracle_proportional_and_shared(tmp_path, oracle):
    r = cases.Runner(oracle, duct(tmp_path, 3, 3, True), "orc_")
    r.step(30)
    check_proportional(r)
    r = cases.Runner(oracle, duct(tmp_path, 1, 3, False), "orc_")
    r.step(30)
    check_shared(r)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_proportional_and_shared(tmp_path, gpu, oracle, mode):
    print(f"case populations mode {mode}", flush=True)
    for copies, scaled, check in ((3, True, check_proportional), (1, False, check_shared)):
        rg = cases.Runner(gpu, duct(tmp_path, copies, 3, scaled, mode=mode), "ice_gpu_")
        rg.step(30)
        check(rg)
        ro = cases.Runner(oracle, duct(tmp_path, copies, 3, scaled), "orc_")
        ro.step(30)
        for p in (1, 2, 3):
            assert np.array_equal(rg.state(1, p), ro.state(1, p))
