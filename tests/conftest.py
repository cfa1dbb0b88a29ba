"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built HIP
library; everything else runs on the CPU (oracle, CLI, host logic)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

ORACLE_LIB = os.path.join(ROOT, "oracle", "_bin", "liboracle.so")
ORACLE_CLI = os.path.join(ROOT, "oracle", "_bin", "rdcu_cpu")
HIP_CLI = os.path.join(ROOT, "rdcu_amd", "bin", "rdcu")
ORACLE_SWEEP = os.path.join(ROOT, "oracle", "_bin", "rdcu_cpu_sweep")
HIP_SWEEP = os.path.join(ROOT, "rdcu_amd", "bin", "rdcu_sweep")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def olib():
    import ctypes
    from rdcu_amd.abi import load
    lib = load(ORACLE_LIB)
    D = ctypes.POINTER(ctypes.c_double)
    lib.oracle_fft.argtypes = [ctypes.c_int, ctypes.c_int, D, D]
    lib.oracle_r2c.argtypes = [ctypes.c_int, D, D]
    lib.oracle_c2r.argtypes = [ctypes.c_int, D, D]
    lib.oracle_plan.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    lib.oracle_unit_root.argtypes = [ctypes.c_long, ctypes.c_long, D]
    lib.oracle_unit_root.restype = None
    lib.oracle_pow_m02.argtypes = [ctypes.c_double]
    lib.oracle_pow_m02.restype = ctypes.c_double
    lib.oracle_pairsum.argtypes = [D, ctypes.c_long]
    lib.oracle_pairsum.restype = ctypes.c_double
    lib.oracle_last_norm.argtypes = [ctypes.c_void_p]
    lib.oracle_last_norm.restype = ctypes.c_double
    lib.oracle_set_threads.argtypes = [ctypes.c_int]
    lib.oracle_set_threads.restype = None
    lib.oracle_factors.argtypes = [ctypes.c_void_p, D, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int), D]
    lib.oracle_loss_tl.argtypes = [ctypes.c_void_p, D, ctypes.c_double, ctypes.c_longlong, D,
                                   ctypes.c_int, D, D, D, D, D, D]
    return lib


def loss_tl(olib, it, y0, t0, hs, t_out, target, dcoef, dy0):
    """Oracle tangent-linear: (loss, d loss along (dcoef, dy0))."""
    import ctypes
    from rdcu_amd.abi import _ptr
    arr = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
           for a in (y0, hs, t_out, target, dcoef, dy0)]
    L, dL = ctypes.c_double(), ctypes.c_double()
    rc = olib.oracle_loss_tl(it.h, _ptr(arr[0]), t0, arr[1].size, _ptr(arr[1]), arr[2].size,
                             _ptr(arr[2]), _ptr(arr[3]), _ptr(arr[4]), _ptr(arr[5]),
                             ctypes.byref(L), ctypes.byref(dL))
    assert rc == 0, olib.rdcu_last_error()
    return L.value, dL.value


@pytest.fixture(scope="session")
def hlib():
    from rdcu_amd.abi import hip_lib
    return hip_lib()


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
