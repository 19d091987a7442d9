"""The C-ABI library: every function include/ice_gpu.h declares is exported, and
the ctypes binding names the same set (no GPU needed: nothing is called)."""
import os
import re
import subprocess

import pytest

from ice_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "ice_gpu.h")).read()
    return sorted(set(re.findall(r"\b(ice_gpu_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_header():
    if not os.path.exists(native.LIB_PATH):
        pytest.fail("ice_amd/libice_gpu.so is missing: run __graft_entry__.build()")
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing


def test_binding_matches_header():
    assert sorted(native.SIGNATURES) == declared()


def test_oracle_exports_same_api():
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    # device-only: timing, tiles, the on-GPU training of the learned seed and its solve counters,
    # and the residual norm on the device's scratch array
    skip = {"ice_gpu_timing", "ice_gpu_set_kernel_timing", "ice_gpu_lk_timing", "ice_gpu_set_tiles", "ice_gpu_set_keep_source",
            "ice_gpu_autotune", "ice_gpu_set_dispatch",
            "ice_gpu_surrogate_train", "ice_gpu_surrogate_params", "ice_gpu_surrogate_retrain_every", "ice_gpu_set_lk_stats", "ice_gpu_lk_stats",
            "ice_gpu_surrogate_retrain", "ice_gpu_residual_norm_noalloc"}
    missing = ["orc_" + s[len("ice_gpu_"):] for s in declared() if s not in skip
               and "orc_" + s[len("ice_gpu_"):] not in exported]
    assert not missing, missing


def test_load_without_gpu():
    """The library loads (dlopen) on a host without a GPU."""
    lib = native.load()
    assert hasattr(lib, "ice_gpu_step")


def test_fortran_module_binds_every_symbol():
    """ice_gpu_m.f90 has a bind(C) interface for every function the header declares"""
    src = open(os.path.join(ROOT, "ice_amd", "fortran", "ice_gpu_m.f90")).read()
    bound = set(re.findall(r"name='(ice_gpu_[a-z0-9_]+)'", src))
    missing = [s for s in declared() if s not in bound]
    assert not missing, missing
