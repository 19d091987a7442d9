"""The text of the 400-inflow report, as ICE prints it: the lines of
count_400_inflow and the totals of report_400_inflow
(src/lib/diagnostic/Mod_Inflow400.f90:118-122,162-165,204-249), from the
records that the C-ABI hands over (ice_gpu_inflow400_lines / _totals)."""
from __future__ import annotations

from .native import AG, IG

_LABEL = {IG: "|w|/sqrt(P/rho): ", AG: "|w|/sqrt(Pnn/rho): "}


def number_text(closure: int, x: float) -> str:
    """ES11.4, or ES12.4E3 for a three-digit exponent, with m/s for MK."""
    s = "%.4E" % x
    return s if closure in (IG, AG) else s + " m/s"


def measure_text(closure: int, x: float) -> str:
    return _LABEL.get(closure, "|w|: ") + number_text(closure, x)


def boundary_name(block: int, face: int, family: int, ngroups: int) -> str:
    s = "block %d, face %d" % (block, face)
    if ngroups > 1:
        s += ", family %d" % family
    return s


def line_text(line, closures, ngroups: int) -> str:
    """One line of a stage (count_400_inflow)."""
    which = ", faces above 10 times its last line: " if line.relative else ", faces: "
    return (" 400 boundary of %s: particles entering at iteration %d%s%d, largest %s"
            % (boundary_name(line.block, line.face, line.family, ngroups), line.iter, which, line.faces,
               measure_text(closures[line.family - 1], line.largest)))


def total_text(t, closures, ngroups: int) -> str:
    """One boundary's line at the end of the run (report_400_inflow)."""
    cl = closures[t.family - 1]
    return (" 400 boundary of %s: particles entered in %d stages, first at iteration %d (faces: %d, largest %s), "
            "largest over the run: %s"
            % (boundary_name(t.block, t.face, t.family, ngroups), t.stages, t.first_iter, t.first_faces,
               measure_text(cl, t.first_largest), number_text(cl, t.largest)))
