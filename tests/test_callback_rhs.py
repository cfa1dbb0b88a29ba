"""Integrators whose right-hand side is a callback (rdcu_create_fn, the dydt
of dp45_init, dp45.h:8), on the CPU oracle: the callback sees the stage
times of dp45.cu:90-106, and a callback that evaluates a dictionary gives the
dictionary's own steps where the error norm's lines coincide (1-D)."""
import ctypes

import numpy as np
import pytest

from rdcu_amd import CallbackIntegrator, Integrator, Problem, RdcuError, cgle_terms, glibc_rand_ic

D = ctypes.POINTER(ctypes.c_double)


def host_array(ptr, n):
    return np.ctypeslib.as_array(ctypes.cast(ptr, D), shape=(n,))


def test_linear_callback_is_dp5_stability_polynomial(olib):
    """y' = lam*y through a callback: one fixed step multiplies y by R(z)."""
    lam, h = -3.0, 0.07

    def fn(t, y, f, stream):
        host_array(f, 2)[:] = lam * host_array(y, 2)
        return 0

    with CallbackIntegrator(olib, 1, 2, np.array([1.0, -2.0]), fn, fixed=True) as it:
        t, hh, acc = it.step(0.0, h)
        y = it.state()
    z = lam * h
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24 + z ** 5 / 120 + z ** 6 / 600
    assert acc and t == h and abs(y[0] - R) <= 2e-15 * abs(R)


def test_callback_sees_the_stage_times(olib):
    """dydt is called at t (k1), then at t + c_s*h for s = 2..7 (dp45.cu:90-106)."""
    seen = []

    def fn(t, y, f, stream):
        seen.append(t)
        host_array(f, 3)[:] = np.cos(t) * host_array(y, 3)
        return 0

    with CallbackIntegrator(olib, 1, 3, np.ones(3), fn) as it:
        it.step(0.5, 0.2)
    c = [1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0]
    assert seen == [0.5] + [0.5 + 0.2 * ci for ci in c] + [0.5 + 0.2]


def test_dictionary_callback_equals_dictionary_problem_in_1d(olib):
    """A callback that evaluates the 1-D CGLE dictionary through rdcu_rhs_device
    (host memory on the CPU build) gives the dictionary problem's own attempts
    bit for bit: in 1-D the norm's lines are the fields in both."""
    prob = Problem(2, [128], [50.0], cgle_terms(1))
    y0 = glibc_rand_ic(2, 128, seed=3, A=1.0)
    with Integrator(olib, prob, y0) as dic, Integrator(olib, prob, y0) as ref:
        def fn(t, y, f, stream):
            return olib.rdcu_rhs_device(dic.h, y, f)

        with CallbackIntegrator(olib, 2, 128, y0, fn) as cb:
            lc, lr = [], []
            cb.run(0.0, 0.01, 0.5, callback=lambda t, h, ok, cur: lc.append((t, h, ok)))
            ref.run(0.0, 0.01, 0.5, callback=lambda t, h, ok, cur: lr.append((t, h, ok)))
            assert lc == lr and len(lc) > 10
            assert np.array_equal(cb.state(), ref.state())
            assert np.array_equal(cb.eval(lc[-1][0] - 0.3 * lc[-1][1]),
                                  ref.eval(lr[-1][0] - 0.3 * lr[-1][1]))


def test_callback_errors_and_dictionary_only_calls(olib):
    def bad(t, y, f, stream):
        return 7

    with CallbackIntegrator(olib, 1, 4, np.zeros(4), bad) as it:
        with pytest.raises(RdcuError) as e:
            it.step(0.0, 0.1)
        assert e.value.code == 6  # RDCU_ECALLBACK
        with pytest.raises(RdcuError) as e:
            it.rhs(np.zeros(4))
        assert e.value.code == 2  # RDCU_EUNSUPPORTED

    def boom(t, y, f, stream):
        raise ValueError("callback failed")

    with CallbackIntegrator(olib, 1, 4, np.zeros(4), boom) as it:
        with pytest.raises(ValueError):
            it.step(0.0, 0.1)
