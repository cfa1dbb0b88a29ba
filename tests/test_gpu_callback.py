"""GPU integrators whose right-hand side is a callback (rdcu_create_fn)
against the oracle's: a callback evaluating the 2-D CGLE dictionary on the
device (rdcu_rhs_device), and a non-autonomous host callback, bit for bit
(attempt log, state, dense output); in 1-D the dictionary callback also
equals the GPU's own dictionary problem."""
import ctypes

import numpy as np
import pytest

from rdcu_amd import CallbackIntegrator, Integrator, Problem, cgle_terms, glibc_rand_ic

pytestmark = pytest.mark.gpu

D = ctypes.POINTER(ctypes.c_double)


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _runs(a, b, t1, h0=0.01):
    la, lb = [], []
    a.run(0.0, h0, t1, callback=lambda t, h, ok, cur: la.append((t, h, ok)))
    b.run(0.0, h0, t1, callback=lambda t, h, ok, cur: lb.append((t, h, ok)))
    return la, lb


@pytest.mark.parametrize("Ns,Ls", [([64, 64], [50.0, 50.0]), ([256], [100.0])], ids=["2d", "1d"])
def test_dictionary_callback_bitwise(hlib, olib, Ns, Ls):
    hip = _hip()
    prob = Problem(2, Ns, Ls, cgle_terms(len(Ns), 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=11, A=1.0)
    print(f"case callback dictionary {Ns}", flush=True)
    with Integrator(hlib, prob, y0) as gd, Integrator(olib, prob, y0) as od:
        def gfn(t, y, f, stream):
            if hip.hipStreamSynchronize(stream):
                return 1
            return hlib.rdcu_rhs_device(gd.h, y, f)

        def ofn(t, y, f, stream):
            return olib.rdcu_rhs_device(od.h, y, f)

        with CallbackIntegrator(hlib, 2, prob.N, y0, gfn) as g, \
                CallbackIntegrator(olib, 2, prob.N, y0, ofn) as o:
            lg, lo = _runs(g, o, 0.3)
            assert lg == lo and len(lg) > 10
            assert np.array_equal(g.state(), o.state())
            te = lg[-1][0] - 0.4 * lg[-1][1]
            assert np.array_equal(g.eval(te), o.eval(te))
            if len(Ns) == 1:  # the norm's lines are the fields: the dictionary run itself
                with Integrator(hlib, prob, y0) as ref:
                    lr = []
                    ref.run(0.0, 0.01, 0.3, callback=lambda t, h, ok, cur: lr.append((t, h, ok)))
                    assert lr == lg and np.array_equal(ref.state(), g.state())


def test_time_dependent_host_callback_bitwise(hlib, olib):
    """f = cos(t) * y - 0.5 * y^3, evaluated by numpy on the host: through
    device copies on the GPU side, in place on the oracle's.  The stage times
    and the integrator's arithmetic are the same bits on both sides."""
    hip = _hip()
    n, N = 3, 5000  # two error chunks per field on the GPU
    y0 = glibc_rand_ic(n, N, seed=5, A=1.0)

    def rhs(t, y):
        return np.cos(t) * y - 0.5 * (y * y * y)

    buf = np.empty(n * N)

    def gfn(t, y, f, stream):
        assert hip.hipStreamSynchronize(stream) == 0
        assert hip.hipMemcpy(buf.ctypes.data, y, buf.nbytes, 2) == 0
        out = np.ascontiguousarray(rhs(t, buf))
        assert hip.hipMemcpy(f, out.ctypes.data, out.nbytes, 1) == 0
        return 0

    def ofn(t, y, f, stream):
        ya = np.ctypeslib.as_array(ctypes.cast(y, D), shape=(n * N,))
        np.ctypeslib.as_array(ctypes.cast(f, D), shape=(n * N,))[:] = rhs(t, ya)
        return 0

    print("case callback host numpy", flush=True)
    with CallbackIntegrator(hlib, n, N, y0, gfn, atol=1e-9, rtol=1e-9) as g, \
            CallbackIntegrator(olib, n, N, y0, ofn, atol=1e-9, rtol=1e-9) as o:
        lg, lo = _runs(g, o, 2.0, h0=0.5)  # the first attempts are rejected
        assert lg == lo and any(not a for _, _, a in lg) and len(lg) > 20
        assert np.array_equal(g.state(), o.state())
