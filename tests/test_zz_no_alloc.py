"""rdcu_rhs_noalloc: rdcu_rhs with every intermediate device buffer taken from
the library's static per-GPU scratch array instead of device allocations.

Checked through the C ABI against rdcu_rhs on the same handles, bit for bit,
on two calls in a row (the second on a smaller problem, so the scratch holds
the first call's larger contents), plus the rejection of a problem whose
buffers exceed the array, and unchanged free device memory across a first
call that rdcu_rhs would have served by allocating its factor buffers.
"""
import ctypes

import numpy as np
import pytest

from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic
from rdcu_amd.abi import RdcuError, EINVAL

pytestmark = pytest.mark.gpu


def _free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_noalloc_rhs_equals_rhs_on_two_calls(hlib):
    big = Problem(2, [128, 96], [60.0, 40.0], cgle_terms(2, 2.0, 0.8))
    small = Problem(2, [64], [30.0], cgle_terms(1))
    yb = glibc_rand_ic(2, big.N, seed=3, A=1.0)
    ys = glibc_rand_ic(2, small.N, seed=4, A=1.0)
    print("case noalloc 128x96 then 64", flush=True)
    with Integrator(hlib, big, yb) as gb, Integrator(hlib, small, ys) as gs:
        before = _free_device_bytes()
        f1, Y1 = gb.rhs(yb, factors=True, noalloc=True)  # first call: factors too
        assert _free_device_bytes() == before
        f2, Y2 = gs.rhs(ys, factors=True, noalloc=True)  # smaller, over the first's contents
        f3 = gs.rhs(ys, noalloc=True)
        rf1, rY1 = gb.rhs(yb, factors=True)
        rf2, rY2 = gs.rhs(ys, factors=True)
        assert np.array_equal(f1, rf1) and np.array_equal(Y1, rY1)
        assert np.array_equal(f2, rf2) and np.array_equal(Y2, rY2)
        assert np.array_equal(f3, rf2)
        # the handle's own stepping is unaffected by the calls in between
        with Integrator(hlib, small, ys) as fresh:
            assert gs.step(0.0, 0.01) == fresh.step(0.0, 0.01)
            assert np.array_equal(gs.state(), fresh.state())


def test_noalloc_rhs_rejects_more_than_the_scratch(hlib):
    prob = Problem(2, [1024, 1024], [100.0, 100.0], cgle_terms(2, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=5, A=1.0)
    print("case noalloc 1024^2 with factors: too large", flush=True)
    with Integrator(hlib, prob, y0) as g:
        with pytest.raises(RdcuError) as e:
            g.rhs(y0, factors=True, noalloc=True)
        assert e.value.code == EINVAL and "scratch" in str(e.value)
        f = g.rhs(y0, noalloc=True)  # without the factor blocks it fits
        assert np.array_equal(f, g.rhs(y0))
