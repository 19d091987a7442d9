"""The coupler's side of births and absorptions (src/lib/coupling/Mod_Exchange.f90:
birth_packet :212-224, absorption_packet :228-245, add_birth_rate :264-280,
add_absorption_rate :284-324): the packets a coupler hands to a family, and the
import rates (conservative slots) that ice_gpu_put_imports holds in force.

A packet of mass m (an amount per unit volume, or a rate per unit volume and
time) arriving with velocity u at temperature T is [m, m u, m (e(T) + |u|^2/2)],
with e = cs T for a material of constant heat capacity; a birth adds the number
of drops of one diameter d, m / (pi/6 rho_mat d^3)."""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np

from .native import AG, IG, NCOND


def absorption_packet(m: float, u: Sequence[float], T: float, cs: float) -> np.ndarray:
    """absorption_packet (Mod_Exchange.f90:228-245) of a constant-cp material."""
    u = np.asarray(u, float)
    e = cs * T
    return np.array([m, m * u[0], m * u[1], m * u[2], m * (e + 0.5 * float(np.dot(u, u)))])


def birth_packet(m: float, u: Sequence[float], T: float, d: float, rho_mat: float, cs: float) -> np.ndarray:
    """birth_packet (Mod_Exchange.f90:212-224): the absorption packet and the number of
    drops of diameter d."""
    pk = np.zeros(6)
    pk[:5] = absorption_packet(m, u, T, cs)
    pk[5] = m / (math.pi / 6.0 * rho_mat * d ** 3)
    return pk


def birth_rate(closure: int, pk: Sequence[float]) -> np.ndarray:
    """add_birth_rate's import vector (Mod_Exchange.f90:264-280) of one cell: the
    packet in slots 1:4, nbase-1 (energy) and nbase (number)."""
    nc = NCOND[closure]
    r = np.zeros(nc)
    r[:4] = pk[:4]
    r[nc - 2] = pk[4]
    r[nc - 1] = pk[5]
    return r


def absorption_rate(closure: int, pk: Sequence[float]) -> np.ndarray:
    """add_absorption_rate's import vector (Mod_Exchange.f90:284-324) of one cell: the
    mass, momentum and energy, and with IG / AG the m u u of a mass arriving with one
    velocity; the number does not change."""
    nc = NCOND[closure]
    pk = np.asarray(pk, float)
    if not pk[0] > 0.0:
        raise ValueError("the mass absorbed must be positive")
    u = pk[1:4] / pk[0]
    r = np.zeros(nc)
    r[:4] = pk[:4]
    r[nc - 2] = pk[4]
    if closure == IG:
        r[4] = pk[0] * float(np.dot(u, u))
    elif closure == AG:
        r[4:10] = pk[0] * np.array([u[0] * u[0], u[0] * u[1], u[0] * u[2], u[1] * u[1], u[1] * u[2], u[2] * u[2]])
    return r
