"""The slots not integrated between mirror faces one cell apart: setup_freeze
(src/lib/numerics/time/explicit/Mod_Freeze.f90:37-141) with the plane test of
mirror_planes / plane_fit / thickness (src/lib/numerics/space/Mod_Metrics.f90:286-355).

On a 2-D mesh whose k faces all mirror (200, 300, 301) and whose k node planes
are planes normal to z, w is frozen (with AG also P13 and P23); on a 1-D mesh
likewise v between mirror j faces normal to y (with AG also P12 and P23); an
axisymmetric wedge (delthe != 0) freezes w when its k faces are all 200 and its k
node planes are the planes at -/+ delthe/2 about z = 0, and stops otherwise.  The
frozen slots are zeroed in the initial state; the stages zero their source and
residual (ice_family_t.frozen)."""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from .native import NCOND, NFROZEN_MAX, R_F, R_TYPE

MESH_FIT = 1e-2    # Mod_Metrics.f90:12
MIRRORS = (200, 300, 301)


class FreezeError(ValueError):
    """setup_freeze stops the run (an axisymmetric wedge that is not one)."""


@dataclasses.dataclass
class Freeze:
    freeze_v: bool
    freeze_w: bool
    frozen: List[List[int]]          # per family, 1-based slots (the primitive and the conservative slot alike)
    warning: Optional[str] = None    # the reference's [WARNING] line, if it prints one


def thickness(node: np.ndarray, d: int) -> float:
    """thickness (Mod_Metrics.f90:286-303): the largest coordinate difference
    between the node planes 0 and 1 of index direction d (2: j, 3: k).
    node: (3, ni+1, nj+1, nk+1)."""
    if d == 2:
        return float(np.max(np.abs(node[:, :, 1, :] - node[:, :, 0, :])))
    return float(np.max(np.abs(node[:, :, :, 1] - node[:, :, :, 0])))


def plane_fit(node: np.ndarray, d: int, delthe: float = 0.0) -> float:
    """plane_fit (Mod_Metrics.f90:332-355): the largest distance of a node of the
    planes 0 and 1 of direction d from its mirror plane, over the block's
    thickness (huge when the block has none)."""
    thick = thickness(node, d)
    off = 0.0
    for s in (0, 1):
        if d == 2:
            y = node[1, :, s, :]
            off = max(off, float(np.max(np.abs(y - y[0, 0]))))
        elif delthe == 0.0:
            z = node[2, :, :, s]
            off = max(off, float(np.max(np.abs(z - z[0, 0]))))
        else:
            ang = (s - 0.5) * delthe
            off = max(off, float(np.max(np.abs(node[2, :, :, s] * np.cos(ang) - node[1, :, :, s] * np.sin(ang)))))
    return off / thick if thick > 0.0 else float(np.finfo(float).max)


def mirror_planes(nodes: Sequence[np.ndarray], d: int, delthe: float = 0.0):
    """mirror_planes (Mod_Metrics.f90:312-327): (mirror, fit) over all blocks."""
    fit = max(plane_fit(n, d, delthe) for n in nodes)
    return fit <= MESH_FIT, fit


def setup_freeze(ndir: int, reci: np.ndarray, nodes: Sequence[np.ndarray], closures: Sequence[int],
                 delthe: float = 0.0) -> Freeze:
    """The frozen slots of every family from the boundary table (all records, every
    family's), the blocks' nodes and the families' closures."""
    f = reci[:, R_F]
    t = reci[:, R_TYPE]
    jf = (f == 3) | (f == 4)
    kf = (f == 5) | (f == 6)
    mir = np.isin(t, MIRRORS)
    nj, nk = int(jf.sum()), int(kf.sum())
    nmj, nmk = int((jf & mir).sum()), int((kf & mir).sum())
    other = np.flatnonzero(kf & (t != 200))
    freeze_v = freeze_w = False
    warning = None
    if ndir == 2 and delthe != 0.0:
        if other.size:
            r = reci[other[0]]
            raise FreezeError(f" [ERROR] [ICE::setup_freeze] face {int(r[R_F])} of the cell (block i j) carries "
                              f"{int(r[R_TYPE])}: the k faces of an axisymmetric wedge must carry 200, the mirror that "
                              "keeps the flow axisymmetric")
        mirror, fit = mirror_planes(nodes, 3, delthe)
        if not mirror:
            raise FreezeError(f" [ERROR] [ICE::setup_freeze] the k node planes of the axisymmetric wedge are not the "
                              f"planes at -/+ delthe/2 about z = 0 (a node off its plane by {fit:.2E} of the thickness "
                              "of its block): the wedge must be symmetric about the plane z = 0")
        freeze_w = True
    elif ndir <= 2:
        if nk > 0 and nmk == nk:
            mirror, fit = mirror_planes(nodes, 3)
            freeze_w = mirror
            if not mirror:
                warning = (f" [WARNING] [ICE::setup_freeze] mirror k faces that are not planes normal to z (a node "
                           f"off by {fit:.2E} of the thickness of its block): w is integrated, and its mode across "
                           "the cells can make a thin layer fail")
        if ndir == 1 and nj > 0 and nmj == nj:
            mirror, fit = mirror_planes(nodes, 2)
            freeze_v = mirror
            if not mirror:
                warning = (f" [WARNING] [ICE::setup_freeze] mirror j faces that are not planes normal to y (a node "
                           f"off by {fit:.2E} of the thickness of its block): v is integrated, and its mode across "
                           "the cells can make thin cells fail")
    frozen = []
    for cl in closures:
        odd = {3: freeze_v, 4: freeze_w}
        if NCOND[cl] == 12:
            odd.update({6: freeze_v, 7: freeze_w, 9: freeze_v or freeze_w})
        frozen.append([s for s in sorted(odd) if odd[s]])
    if (freeze_v or freeze_w) and warning is None:
        ag = any(NCOND[c] == 12 for c in closures)
        if freeze_v and freeze_w:
            what, faces = ("v and w (with AG also P12, P13 and P23)" if ag else "v and w"), "j and k"
        elif freeze_v:
            what, faces = ("v (with AG also P12 and P23)" if ag else "v"), "j"
        else:
            what, faces = ("w (with AG also P13 and P23)" if ag else "w"), "k"
        if ndir == 2 and delthe != 0.0:
            warning = f" [WARNING] [ICE::setup_freeze] axisymmetric wedge: the swirl {what} set to zero and not integrated"
        else:
            mesh = "2-D plane" if ndir == 2 else "1-D"
            warning = (f" [WARNING] [ICE::setup_freeze] {mesh} mesh with mirror {faces} faces: {what} set to zero "
                       "and not integrated")
    return Freeze(freeze_v, freeze_w, frozen, warning)


def apply_freeze(case, nodes: Sequence[np.ndarray], delthe: float = 0.0) -> Freeze:
    """setup_freeze on a Case: each family's ice_family_t.frozen, and the frozen
    slots of its initial state (ghosts included) set to zero."""
    fz = setup_freeze(case.cfg.ndir, case.reci, nodes, [f.closure for f in case.fams], delthe)
    for p, fam in enumerate(case.fams):
        slots = fz.frozen[p]
        fam.nfrozen = len(slots)
        for m in range(NFROZEN_MAX):
            fam.frozen[m] = slots[m] if m < len(slots) else 0
        nc = NCOND[fam.closure]
        for B in case.blocks:
            a = B.prim[p].reshape(-1, nc)  # one row per padded cell (the slot index is the fastest)
            for s in slots:
                a[:, s - 1] = 0.0
    return fz


__all__ = ["Freeze", "FreezeError", "setup_freeze", "apply_freeze", "mirror_planes", "plane_fit", "thickness",
           "MESH_FIT"]
