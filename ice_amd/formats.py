"""ICE's on-disk inputs, read into the layouts of include/ice_gpu.h.

* Tecplot ASCII, BLOCK packing, nodal x y z then cell-centred variables, one
  zone per block: the initial condition ``INPUT/<prefix>ic.tec`` and the
  solution ``<prefix>field.tec`` (read_ic / write_solution,
  src/lib/io/IO_Solution.f90:36-38, ORION-written, 17 significant digits).
* The ATLAS boundary file ``INPUT/<prefix>bc.txt``: per boundary cell a header
  ``b i j k f code`` and a code-dependent payload (Setup_BC,
  src/lib/io/IO_BC.f90:65-344); the records are fanned out over the families,
  record idx = (n-1)*ngroups + p (:299).
* The INI run file ``input.ini`` (sections and ``key = value`` lines).

Setup_BC's refusals that concern the layout are kept: a code outside {0,
100..999}, a short payload, a chimera layer whose weights sum below 0.5, a cell
off its block, and inlet angles that are not both numbers or both ``normal``.
"""
from __future__ import annotations

import re
from typing import Dict, List, Sequence, Tuple

import numpy as np

from .native import (R_B, R_BS, R_DONOR0, R_F, R_FS, R_I, R_IS, R_J, R_JS, R_K, R_KS, R_NCHI1, R_NCHI2,
                     R_NORMAL_DIR, R_P, R_TYPE, RD_ALPHA, RD_BETA, RD_GFIXED, RD_MASSFLUX, RD_RADIUS,
                     RD_TEMPERATURE, RD_VELOCITY, REC_ND, REC_NI)


class FormatError(ValueError):
    """A file that Setup_BC / read_ic would refuse; the message names file and line."""


# ------------------------------------------------------------------ Tecplot
class Zone:
    """One block of a Tecplot BLOCK file: I, J, K node counts, nodes (3, I, J, K)
    and cell variables (nvar, I-1, J-1, max(K-1, 1)), Fortran index order."""

    def __init__(self, title: str, dims: Tuple[int, int, int], nodes: np.ndarray, cells: np.ndarray, aux: Dict):
        self.title = title
        self.I, self.J, self.K = dims
        self.nodes = nodes
        self.cells = cells
        self.aux = aux

    @property
    def cell_dims(self) -> Tuple[int, int, int]:
        return self.I - 1, self.J - 1, max(self.K - 1, 1)


def _zone_int(head: str, key: str) -> int:
    m = re.search(rf"\b{key}\s*=\s*(\d+)", head, re.IGNORECASE)
    if m is None:
        raise FormatError(f"zone header without {key}: {head.strip()[:80]}")
    return int(m.group(1))


def read_tecplot(path: str) -> Tuple[List[str], List[Zone]]:
    """Variable names and zones of a Tecplot ASCII BLOCK file with nodal x, y, z
    and cell-centred variables (the layout ICE reads and writes)."""
    with open(path) as fh:
        text = fh.read()
    m = re.search(r"(?im)^\s*VARIABLES\s*=(.*?)(?=^\s*ZONE)", text, re.S | re.M)
    if m is None:
        raise FormatError(f"{path}: no VARIABLES line before the first ZONE")
    names = re.findall(r'"([^"]*)"|(\S+)', m.group(1).replace(",", " "))
    names = [a or b for a, b in names]
    pieces = re.split(r"(?im)^\s*ZONE\b", text)[1:]
    zones = []
    for z in pieces:
        head, _, body = z.partition("\n")
        I, J, K = _zone_int(head, "I"), _zone_int(head, "J"), _zone_int(head, "K")
        t = re.search(r"\bT\s*=\s*\"?([^,\"]+)", head)
        aux = {}
        keep = []
        for line in body.split("\n"):
            s = line.strip()
            if s.upper().startswith(("DATASETAUXDATA", "AUXDATA")):
                a = re.match(r"\S+\s+(\S+)\s*=\s*\"?([^\"]*)\"?", s)
                if a:
                    aux[a.group(1)] = a.group(2)
                continue
            keep.append(line)
        vals = np.array(" ".join(keep).split(), dtype=np.float64)
        nn = I * J * K
        nc = (I - 1) * (J - 1) * max(K - 1, 1)
        nvar = (vals.size - 3 * nn) // nc
        if vals.size != 3 * nn + nvar * nc or nvar < 1:
            raise FormatError(f"{path}: zone {t.group(1) if t else '?'} holds {vals.size} values, not 3*{nn} nodal "
                              f"+ n*{nc} cell values")
        nodes = vals[:3 * nn].reshape(3, K, J, I).transpose(0, 3, 2, 1)
        cells = vals[3 * nn:].reshape(nvar, max(K - 1, 1), J - 1, I - 1).transpose(0, 3, 2, 1)
        zones.append(Zone(t.group(1).strip() if t else "", (I, J, K), np.ascontiguousarray(nodes),
                          np.ascontiguousarray(cells), aux))
    return names, zones


def write_tecplot(path: str, names: Sequence[str], zones: Sequence[Zone]) -> None:
    """A solution file in the layout read_tecplot reads (17 significant digits)."""
    with open(path, "w") as fh:
        fh.write(" VARIABLES = " + " ".join(names) + "\n")
        for b, z in enumerate(zones, 1):
            nv = z.cells.shape[0]
            fh.write(f" ZONE  T = Block{b}, I={z.I}, J={z.J}, K={z.K}, DATAPACKING=BLOCK, "
                     f"VARLOCATION=([1-3]=NODAL,[4-{3 + nv}]=CELLCENTERED)\n")
            for arr in list(z.nodes) + list(z.cells):
                for v in arr.ravel(order="F"):
                    fh.write(f" {v:.17E}\n")


# ------------------------------------------------------------------ BC file
_NEEDS_PAYLOAD = {101, 201, 102, 401, 402, 403}


def _dir_token(tok: str, where: str) -> Tuple[float, bool]:
    if tok.strip().lower() == "normal":
        return 0.0, True
    try:
        return float(tok), False
    except ValueError:
        raise FormatError(f"{where}: a direction that is neither a number nor normal: '{tok}'") from None


def _bc_records(path: str, dims: Sequence[Tuple[int, int, int]]):
    """Every record of an ATLAS bc.txt in file order (Setup_BC's reading rules,
    IO_BC.f90:144-326; Count_BC_Records walks the same rules): a list of (ri,
    rd, donors, weights, line) with ri (REC_NI,) int32 and rd (REC_ND,) float64
    of the header and payload, and the chimera record's donors and weights."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    pos = 0

    def nxt(need=True):
        nonlocal pos
        while pos < len(lines) and not lines[pos].strip():
            pos += 1
        if pos >= len(lines):
            if need:
                raise FormatError(f"{path}: the file ends before the last boundary cell")
            return None, None
        pos += 1
        # list-directed input (IO_BC.f90:261-265): commas separate values as blanks do
        return pos, [t for t in re.split(r"[,\s]+", lines[pos - 1].strip()) if t]

    out = []
    while True:
        ln, tok = nxt(need=False)
        if tok is None:
            return out
        n = len(out)
        where = f"{path}, line {ln}, record {n + 1}"
        ri = np.zeros(REC_NI, np.int32)
        rd = np.zeros(REC_ND, np.float64)
        dl: List[Tuple[int, int, int, int]] = []
        wl: List[float] = []
        if len(tok) < 6:
            raise FormatError(f"{where}: a header needs b i j k f code")
        try:
            b, i, j, k, f, code = (int(x) for x in tok[:6])
        except ValueError:
            raise FormatError(f"{where}: a header value that is not an integer") from None
        if code != 0 and not 100 <= code <= 999:
            raise FormatError(f"{where}: bad BC code {code} (the ATLAS schema: 6 columns, a 3-digit code or 0)")
        if not 1 <= b <= len(dims):
            raise FormatError(f"{where}: block {b} is not in the mesh")
        ni, nj, nk = dims[b - 1]
        cell_ok = 1 <= i <= ni and 1 <= j <= nj and 1 <= k <= nk and 1 <= f <= 6
        on_face = {1: i == 1, 2: i == ni, 3: j == 1, 4: j == nj, 5: k == 1, 6: k == nk}.get(f, False)
        if not (cell_ok and on_face):
            raise FormatError(f"{where}: cell ({b},{i},{j},{k}) is not on face {f} of its block")
        ri[R_B], ri[R_I], ri[R_J], ri[R_K], ri[R_F], ri[R_TYPE] = b, i, j, k, f, code
        if code in (101, 201):
            ln, tok = nxt()
            if len(tok) < 9:
                raise FormatError(f"{path}, line {ln}: a connection payload needs bs is js ks fs d11 d12 d21 d22")
            ri[R_BS], ri[R_IS], ri[R_JS], ri[R_KS], ri[R_FS] = (int(x) for x in tok[:5])
        elif code == 102:
            ln, tok = nxt()
            n1, n2 = int(tok[0]), int(tok[1])
            ri[R_NCHI1], ri[R_NCHI2] = n1, n2
            for _ in range(n1 + n2):
                ln, tok = nxt()
                db, di, dj, dk = (int(x) for x in tok[:4])
                if not 1 <= db <= len(dims):
                    raise FormatError(f"{path}, line {ln}: donor block {db} is not in the mesh")
                dl.append((db, di, dj, dk))
                wl.append(float(tok[4]))
            if sum(wl[:n1]) < 0.5 or sum(wl[n1:]) < 0.5:
                raise FormatError(f"{where}: a chimera ghost layer whose donor weights sum to ~0")
        elif 401 <= code <= 403:
            ln, tok = nxt()
            if len(tok) < 6:
                raise FormatError(f"{path}, line {ln}: an inlet payload needs gp v alpha beta T radius")
            a, na = _dir_token(tok[2], f"{path}, line {ln}")
            be, nb = _dir_token(tok[3], f"{path}, line {ln}")
            if na != nb:
                raise FormatError(f"{path}, line {ln}: alpha and beta are both normal or both numbers")
            rd[RD_MASSFLUX], rd[RD_VELOCITY] = float(tok[0]), float(tok[1])
            rd[RD_ALPHA], rd[RD_BETA] = a, be
            rd[RD_TEMPERATURE], rd[RD_RADIUS] = float(tok[4]), float(tok[5])
            ri[R_NORMAL_DIR] = int(na)
            if code == 401:
                if len(tok) < 10:
                    raise FormatError(f"{path}, line {ln}: a 401 line without its 10th field G")
                rd[RD_GFIXED] = float(tok[9])
            if rd[RD_MASSFLUX] != 0.0 and (rd[RD_VELOCITY] == 0.0 or not rd[RD_RADIUS] > 0.0):
                raise FormatError(f"{path}, line {ln}: a mass flux with a zero velocity or a non-positive radius")
        elif code not in (0, 200, 300, 301, 400):
            raise FormatError(f"{where}: unsupported ATLAS BC code {code}")
        out.append((ri, rd, dl, wl, ln))


def read_bc(path: str, dims: Sequence[Tuple[int, int, int]], ngroups: int):
    """Boundary table of an ATLAS bc.txt file (Setup_BC, IO_BC.f90:65-344).

    The file holds one copy of the table for all families, or one copy per
    family (one per material and population pair), block by block with every
    copy of a block together over the same cells; copy c then belongs to family
    c.  Any other number of copies is refused.  Returns (reci, recd, donor,
    weight) in the C-ABI layout: reci (nrec, REC_NI) int32 and recd (nrec,
    REC_ND) float64, nrec = ncell * ngroups in record order idx = (n-1)*ngroups
    + p; donor (nd, 4) int32 (b, i, j, k) and weight (nd,), record r's donors
    from reci[r, R_DONOR0] on (layer 1, then 2).
    """
    recs = _bc_records(path, dims)
    per_block = [2 * (nj * nk + ni * nk + ni * nj) for ni, nj, nk in dims]
    ncell = sum(per_block)
    if len(recs) < ncell:
        raise FormatError(f"{path}: the file ends before the last boundary cell")
    if len(recs) % ncell:
        raise FormatError(f"{path}: the file holds {len(recs)} records, which is not a whole number of copies "
                          f"of the {ncell} boundary cells of this mesh")
    ncopy = len(recs) // ncell
    if ncopy not in (1, ngroups):
        raise FormatError(f"{path}: the file carries {ncopy} copies of the boundary table but this run has "
                          f"{ngroups} particle families")
    reci = np.zeros((ncell * ngroups, REC_NI), np.int32)
    recd = np.zeros((ncell * ngroups, REC_ND), np.float64)
    donor: List[Tuple[int, int, int, int]] = []
    weight: List[float] = []

    def put(n, p, rec, shared_donor=None):
        ri, rd, dl, wl, _ = rec
        idx = n * ngroups + (p - 1)
        reci[idx], recd[idx] = ri, rd
        reci[idx, R_P] = p
        if ri[R_TYPE] == 102:
            if shared_donor is None:
                reci[idx, R_DONOR0] = len(donor)
                donor.extend(dl)
                weight.extend(wl)
            else:
                reci[idx, R_DONOR0] = shared_donor
        return reci[idx, R_DONOR0]

    if ncopy == 1:
        # one record for every family, each carrying its own (b, i, j, k, f): any order
        for n, rec in enumerate(recs):
            d0 = put(n, 1, rec)
            for p in range(2, ngroups + 1):
                put(n, p, rec, shared_donor=d0)
    else:
        pos, base = 0, 0
        for b, nbb in enumerate(per_block, 1):
            keys = []
            for c in range(1, ncopy + 1):
                for r in range(nbb):
                    rec = recs[pos]
                    pos += 1
                    ri, ln = rec[0], rec[4]
                    if ri[R_B] != b:
                        raise FormatError(f"{path}, line {ln}: record {pos} belongs to block {ri[R_B]} where the mesh "
                                          f"expects block {b}: a multi-population file lists the blocks in mesh "
                                          f"order, every copy of one block together")
                    key = (int(ri[R_I]), int(ri[R_J]), int(ri[R_K]), int(ri[R_F]))
                    if c == 1:
                        keys.append(key)
                    elif keys[r] != key:
                        raise FormatError(f"{path}, line {ln}: copy {c} of block {b} does not repeat the cells of "
                                          f"copy 1 in the same order")
                    put(base + r, c, rec)
            base += nbb
    dn = np.asarray(donor, np.int32).reshape(-1, 4)
    return (np.ascontiguousarray(reci), np.ascontiguousarray(recd), np.ascontiguousarray(dn),
            np.asarray(weight, np.float64))


# ------------------------------------------------------------------ INI
def read_ini(path: str) -> Dict[str, Dict[str, str]]:
    """Sections -> {key: value} of an INI file (``;`` and ``#`` start comments)."""
    out: Dict[str, Dict[str, str]] = {}
    sec = None
    with open(path) as fh:
        for ln, raw in enumerate(fh, 1):
            line = re.split(r"[;#]", raw, maxsplit=1)[0].strip()
            if not line:
                continue
            if line.startswith("["):
                sec = line.strip("[]").strip()
                out.setdefault(sec, {})
                continue
            if "=" not in line or sec is None:
                raise FormatError(f"{path}, line {ln}: neither a [section] nor key = value")
            k, v = line.split("=", 1)
            out[sec][k.strip()] = v.strip()
    return out


def write_bc(path: str, reci: np.ndarray, recd: np.ndarray, donor: np.ndarray, weight: np.ndarray,
             ngroups: int = 1, copies: bool = False) -> None:
    """An ATLAS bc.txt from a boundary table in the C-ABI layout: one copy for all
    families (family 1's records), or with copies one copy per family, block by
    block, as ATLAS writes a multi-population phase (the inverse of read_bc)."""
    order = list(range(0, reci.shape[0], ngroups))
    if copies and ngroups > 1:
        order = []
        blocks = reci[::ngroups, R_B]
        for b in range(1, int(blocks.max()) + 1):
            cells = [n for n in range(0, reci.shape[0], ngroups) if reci[n, R_B] == b]
            for p in range(ngroups):
                order.extend(n + p for n in cells)
    with open(path, "w") as fh:
        for n in order:
            r, d = reci[n], recd[n]
            code = int(r[R_TYPE])
            fh.write(f"{r[R_B]:8d}{r[R_I]:8d}{r[R_J]:8d}{r[R_K]:8d}{r[R_F]:8d}{code:8d}\n")
            if code in (101, 201):
                fh.write(f"{r[R_BS]:8d}{r[R_IS]:8d}{r[R_JS]:8d}{r[R_KS]:8d}{r[R_FS]:8d}       1       0       0       1\n")
            elif code == 102:
                n1, n2, d0 = int(r[R_NCHI1]), int(r[R_NCHI2]), int(r[R_DONOR0])
                fh.write(f"{n1:8d}{n2:8d}\n")
                for s in range(d0, d0 + n1 + n2):
                    fh.write(f"{donor[s, 0]:8d}{donor[s, 1]:8d}{donor[s, 2]:8d}{donor[s, 3]:8d} {weight[s]:.17E}\n")
            elif 401 <= code <= 403:
                ang = ("normal", "normal") if r[R_NORMAL_DIR] else (f"{d[RD_ALPHA]:.17E}", f"{d[RD_BETA]:.17E}")
                line = (f" {d[RD_MASSFLUX]:.17E} {d[RD_VELOCITY]:.17E} {ang[0]} {ang[1]} {d[RD_TEMPERATURE]:.17E} "
                        f"{d[RD_RADIUS]:.17E} 0.0 Dirac 0.0")
                if code == 401:
                    line += f" {d[RD_GFIXED]:.17E}"
                fh.write(line + "\n")


def write_ini(path: str, ini: Dict[str, Dict[str, str]]) -> None:
    with open(path, "w") as fh:
        for sec, kv in ini.items():
            fh.write(f"[{sec}]\n")
            for k, v in kv.items():
                fh.write(f"{k} = {v}\n")
            fh.write("\n")
