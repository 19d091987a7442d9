"""ctypes binding of the C-ABI in include/ice_gpu.h.

The structures mirror the header field for field.  `load()` finds the in-tree
``ice_amd/libice_gpu.so`` built by ``__graft_entry__.build()``; there is no
other implementation to fall back to, so a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libice_gpu.so")

GC = 2
NEP = 8
MK, IG, AG = 0, 1, 2
SAUREL, RUSANOV, HLLE = 0, 1, 2
LIMITERS = ["IORD", "minmod", "vanalbada", "vanleer", "ospre", "umist", "osher", "sweby", "mc", "koren",
            "superbee"]
MODE_PLAIN, MODE_BATCHED = 0, 1

# record fields (ICE_R_* / ICE_RD_*)
R_B, R_I, R_J, R_K, R_F, R_TYPE, R_P, R_BS, R_IS, R_JS, R_KS, R_FS, R_NORMAL_DIR, R_NCHI1, R_NCHI2, \
    R_DONOR0 = range(16)
REC_NI = 16
RD_MASSFLUX, RD_VELOCITY, RD_TEMPERATURE, RD_RADIUS, RD_ALPHA, RD_BETA, RD_GFIXED = range(7)
REC_ND = 7

NCOND = {MK: 6, IG: 7, AG: 12}
NFROZEN_MAX = 6


class Family(C.Structure):
    _fields_ = [("closure", C.c_int32), ("riemann", C.c_int32), ("nfrozen", C.c_int32),
                ("frozen", C.c_int32 * NFROZEN_MAX), ("evap_select", C.c_int32), ("intf_select", C.c_int32),
                ("surrogate", C.c_int32), ("rho_mat", C.c_double), ("cs", C.c_double), ("lv", C.c_double),
                ("emiss", C.c_double), ("ep", C.c_double * NEP)]


class Config(C.Structure):
    _fields_ = [("ngroups", C.c_int32), ("nrk", C.c_int32), ("time_accurate", C.c_int32),
                ("cfl_ramp_iter", C.c_int32), ("ndir", C.c_int32), ("limiter", C.c_int32),
                ("shock_detector", C.c_int32), ("irs", C.c_int32), ("coupled", C.c_int32),
                ("drag_select", C.c_int32), ("heat_select", C.c_int32), ("blow_select", C.c_int32),
                ("mode", C.c_int32), ("var_yinf", C.c_int32), ("reserved", C.c_int32 * 2),
                ("cfl", C.c_double), ("dt_max", C.c_double), ("tau_factor", C.c_double),
                ("irs_beta", C.c_double)]


class Status(C.Structure):
    _fields_ = [("n_floored", C.c_int64), ("bad_state", C.c_int32), ("bad_block", C.c_int32),
                ("bad_i", C.c_int32), ("bad_j", C.c_int32), ("bad_k", C.c_int32), ("bad_p", C.c_int32),
                ("bad_recon", C.c_int32), ("iter", C.c_int32), ("dtglobal", C.c_double), ("time", C.c_double)]


class Inflow400Line(C.Structure):
    _fields_ = [("iter", C.c_int32), ("block", C.c_int32), ("face", C.c_int32), ("family", C.c_int32),
                ("faces", C.c_int32), ("relative", C.c_int32), ("largest", C.c_double)]


class Inflow400Total(C.Structure):
    _fields_ = [("block", C.c_int32), ("face", C.c_int32), ("family", C.c_int32), ("stages", C.c_int32),
                ("first_iter", C.c_int32), ("first_faces", C.c_int32), ("first_largest", C.c_double),
                ("largest", C.c_double)]


INFLOW400_LOG = 4096
_DP = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_IP = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_FP = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_LP = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
MLP_PARAMS = 6288

# symbols of include/ice_gpu.h with their argument types
SIGNATURES = {
    "ice_gpu_create": [C.c_int, C.POINTER(Config), C.POINTER(Family), C.c_int, _IP, C.POINTER(C.c_void_p)],
    "ice_gpu_set_geometry": [C.c_void_p, C.c_int, _DP, _DP, _DP, _DP, _DP, _DP],
    "ice_gpu_set_boundaries": [C.c_void_p, C.c_int, _IP, _DP, C.c_int, _IP, _DP],
    "ice_gpu_get_bc_maps": [C.c_void_p, _IP, _IP, _IP, C.c_int, C.POINTER(C.c_int32), _IP, _IP, C.c_int],
    "ice_gpu_put_state": [C.c_void_p, C.c_int, C.c_int, _DP],
    "ice_gpu_get_state": [C.c_void_p, C.c_int, C.c_int, _DP],
    "ice_gpu_put_gas": [C.c_void_p, C.c_int, _DP, _DP, _DP, _DP, _DP],
    "ice_gpu_put_gas_yinf": [C.c_void_p, C.c_int, _DP],
    "ice_gpu_put_tau": [C.c_void_p, C.c_int, C.c_int, _DP],
    "ice_gpu_get_tau": [C.c_void_p, C.c_int, C.c_int, _DP],
    "ice_gpu_step": [C.c_void_p, C.c_int, C.POINTER(Status)],
    "ice_gpu_set_keep_source": [C.c_void_p, C.c_int],
    "ice_gpu_get_source": [C.c_void_p, C.c_int, C.c_int, _DP, _DP],
    "ice_gpu_residual_norm": [C.c_void_p, _DP, _DP],
    "ice_gpu_residual_norm_noalloc": [C.c_void_p, _DP, _DP],
    "ice_gpu_timing": [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)],
    "ice_gpu_set_kernel_timing": [C.c_void_p, C.c_int],
    "ice_gpu_lk_timing": [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
    "ice_gpu_set_tiles": [C.c_void_p, C.c_int, C.c_int, C.c_int],
    "ice_gpu_set_dispatch": [C.c_void_p, C.c_int],
    "ice_gpu_autotune": [C.c_void_p, C.c_int, _IP, np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")],
    "ice_gpu_mlp_grad": [C.c_void_p, C.c_int, _FP, _FP, _FP, _FP, _FP],
    "ice_gpu_mlp_forward": [C.c_void_p, C.c_int, _FP, _FP, _FP],
    "ice_gpu_surrogate_train": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, _FP],
    "ice_gpu_surrogate_retrain_every": [C.c_void_p, C.c_int, C.c_int],
    "ice_gpu_surrogate_retrain": [C.c_void_p, C.c_int, C.c_int],
    "ice_gpu_surrogate_params": [C.c_void_p, C.c_int, _FP],
    "ice_gpu_surrogate_set_params": [C.c_void_p, C.c_int, _FP],
    "ice_gpu_set_lk_stats": [C.c_void_p, C.c_int],
    "ice_gpu_lk_stats": [C.c_void_p, C.c_int, _LP],
    "ice_gpu_restrict": [C.c_void_p, C.c_void_p],
    "ice_gpu_prolong": [C.c_void_p, C.c_void_p],
    "ice_gpu_set_res_threshold": [C.c_void_p, C.c_double],
    "ice_gpu_res_stop": [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)],
    "ice_gpu_set_properties": [C.c_void_p, C.c_int, C.c_int, C.c_int, _DP, C.c_int32],
    "ice_gpu_inflow400_lines": [C.c_void_p, C.POINTER(Inflow400Line), C.c_int, C.POINTER(C.c_int32),
                                C.POINTER(C.c_int64)],
    "ice_gpu_inflow400_totals": [C.c_void_p, C.POINTER(Inflow400Total), C.c_int, C.POINTER(C.c_int32)],
    "ice_gpu_fill_ghosts": [C.c_void_p],
    "ice_gpu_put_imports": [C.c_void_p, C.c_int, C.c_int, C.c_void_p],
    "ice_gpu_add_births": [C.c_void_p, C.c_int, C.c_int, _IP, _DP],
    "ice_gpu_set_mode": [C.c_void_p, C.c_int],
    "ice_gpu_set_surrogate": [C.c_void_p, C.c_int, C.c_int],
    "ice_gpu_set_counters": [C.c_void_p, C.c_int32, C.c_double, C.c_double],
    "ice_gpu_get_counters": [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double)],
    "ice_gpu_last_error": [C.c_void_p, C.c_char_p, C.c_int],
    "ice_gpu_destroy": [C.c_void_p],
}

_LIB = None


def load(path: str = LIB_PATH):
    """The HIP library; raises when it was not built (no CPU fallback exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() to compile the HIP library")
    lib = C.CDLL(path)
    bind(lib, "ice_gpu_")
    _LIB = lib
    return lib


def bind(lib, prefix: str):
    """Set argtypes/restype of every C-ABI symbol, under `prefix` (the oracle uses orc_)."""
    for name, args in SIGNATURES.items():
        sym = prefix + name[len("ice_gpu_"):]
        if not hasattr(lib, sym):
            continue
        f = getattr(lib, sym)
        f.argtypes = args
        f.restype = None if name == "ice_gpu_destroy" else C.c_int
    return lib
