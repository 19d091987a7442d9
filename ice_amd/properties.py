"""Tabulated material properties for ice_gpu_set_properties: the columns and
flags that ICE's Load_Table (src/lib/io/Load_Table.f90:131-147, 160-340)
derives from a property table, with its node and column checks.

The table sits on consecutive integer kelvins from Tmin.  Density and cp always
come from it (rho_varies / cs_varies say whether they vary); the energy is
e = h - hOff with hOff the enthalpy at 0 K along the first segment
(table_h_offset, :535-543).  Psat, Lv and Cpv serve a material that evaporates,
Emissivity every material (all zero: absent; above 1: cut to 1)."""
from __future__ import annotations

import dataclasses
from typing import Optional, Sequence

import numpy as np

from .formats import FormatError

RHO, CP, E, PSAT, LV, CPV, EMISS, NCOL = range(8)
RHO_VARIES, CS_VARIES, USE_PSAT, USE_LV, USE_CPV, USE_EMISS = 1, 2, 4, 8, 16, 32


@dataclasses.dataclass
class Table:
    t_min: int
    n: int
    cols: np.ndarray      # (NCOL * n,) float64, column after column
    flags: int
    h_off: float

    def col(self, c: int) -> np.ndarray:
        return self.cols[c * self.n:(c + 1) * self.n]


def table_h_offset(T, cp, h) -> float:
    """Enthalpy at 0 K along the first segment (:535-543)."""
    if np.all(cp == cp[0]):
        return float(h[0] - cp[0] * T[0])
    return float(h[0] - T[0] * (h[1] - h[0]))


def _check_nodes(T):
    """check_table_nodes (:480-496)"""
    if T.size < 2:
        raise FormatError("a property table needs at least two rows")
    if not np.all(np.isfinite(T)):
        raise FormatError("a temperature that is not finite")
    t0 = int(np.rint(T[0]))
    if t0 < 0:
        raise FormatError("a temperature below 0 K")
    if np.any(np.abs(T - (t0 + np.arange(T.size))) > 1e-6):
        raise FormatError("rows not on consecutive integer kelvins")
    return t0


def _check_columns(T, cp, rho, h, relative):
    """check_table_columns (:501-532)"""
    if not (np.all(np.isfinite(cp)) and np.all(np.isfinite(rho)) and np.all(np.isfinite(h))):
        raise FormatError("a value that is not finite")
    if np.any(rho <= 0.0):
        raise FormatError("a density that is not positive")
    if np.any(cp <= 0.0):
        raise FormatError("a cp that is not positive")
    if np.any(h[1:] <= h[:-1]):
        raise FormatError("an enthalpy that does not increase")
    if np.all(cp == cp[0]):
        off = table_h_offset(T, cp, h)
        if np.any(np.abs(h - cp[0] * T - off) > 1e-6 * np.maximum(np.abs(h), cp[0] * T)):
            raise FormatError("an enthalpy that does not follow the constant cp")
        if relative and abs(off) > 1e-6 * cp[0] * T[0]:
            raise FormatError("a relative enthalpy with an offset (write it as Enthalpy_abs)")
    else:
        dh = h[1:] - h[:-1]
        trap = 0.5 * (cp[:-1] + cp[1:]) * (T[1:] - T[:-1])
        tol = np.maximum(1e-3 * dh, 1e-6 * np.maximum(np.abs(h[:-1]), np.abs(h[1:])))
        if np.any(np.abs(dh - trap) > tol):
            raise FormatError("an enthalpy that does not follow the cp column")


def _optional(col: Optional[Sequence[float]], n: int, name: str):
    """An optional positive column (validate_positive_column, :449-461): None when absent."""
    if col is None:
        return None
    a = np.asarray(col, float)
    if a.size != n:
        raise FormatError(f"{name} column of {a.size} rows, the table has {n}")
    if not np.all(np.isfinite(a)):
        raise FormatError(f"{name} column: a value that is not finite")
    if np.all(a == 0.0):
        return None
    if np.any(a == 0.0):
        raise FormatError(f"{name} column: zero on some rows and not on others")
    if np.any(a < 0.0):
        raise FormatError(f"{name} column: a value that is not positive")
    return a


def material_table(T, cp, rho, h, psat=None, lv=None, cpv=None, emiss=None, evaporates: bool = False,
                   relative: bool = True) -> Table:
    """The columns and flags of one material's table, as Load_Table sets them."""
    T, cp, rho, h = (np.asarray(x, float) for x in (T, cp, rho, h))
    n = T.size
    if not (cp.size == rho.size == h.size == n):
        raise FormatError("columns of different lengths")
    t0 = _check_nodes(T)
    _check_columns(T, cp, rho, h, relative)
    off = table_h_offset(T, cp, h)
    cols = np.zeros((NCOL, n))
    cols[RHO], cols[CP], cols[E] = rho, cp, h - off
    flags = (RHO_VARIES if np.any(rho != rho[0]) else 0) | (CS_VARIES if np.any(cp != cp[0]) else 0)
    if evaporates:
        for c, bit, data, name in ((PSAT, USE_PSAT, psat, "Psat"), (LV, USE_LV, lv, "Lv"), (CPV, USE_CPV, cpv, "Cpv")):
            a = _optional(data, n, name)
            if a is not None:
                cols[c] = a
                flags |= bit
    a = _optional(emiss, n, "Emissivity")
    if a is not None:
        cols[EMISS] = np.minimum(a, 1.0)
        flags |= USE_EMISS
    return Table(t0, n, np.ascontiguousarray(cols.ravel()), flags, off)


def table_value(tab: Table, c: int, T: float) -> float:
    """The table's value at T (Lib_Properties.f90:233-250), for host-side scalars such as
    Lv(298.15 K) and Cpv(boiling-temperature) that Load_Table stores (:254-266)."""
    if T != T:
        return T
    lo, hi = tab.t_min, tab.t_min + tab.n - 1
    t = tab.col(c)
    Tc = min(max(T, float(lo)), float(hi))
    if Tc >= hi:
        return float(t[-1])
    i = int(Tc)
    return float(t[i - lo] + (t[i + 1 - lo] - t[i - lo]) * (Tc - i))


def table_e(tab: Table, T):
    """The table's energy e(T) (mat_e, Lib_Properties.f90:163-200) on an array of
    temperatures: linear between the nodes, the line through e(0) = 0 below Tmin
    (the first segment when Tmin = 0), the last segment's slope above Tmax."""
    T = np.asarray(T, float)
    t = tab.col(E)
    lo, hi = tab.t_min, tab.t_min + tab.n - 1
    i = np.clip(np.floor(T).astype(np.int64), lo, hi - 1)
    mid = t[i - lo] + (t[i + 1 - lo] - t[i - lo]) * (T - i)
    below = t[0] * (T / lo) if lo > 0 else (t[1] - t[0]) * T
    above = t[-1] + (t[-1] - t[-2]) * (T - hi)
    return np.where(T < lo, below, np.where(T >= hi, above, mid))
