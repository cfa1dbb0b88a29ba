"""Pins the oracle's DP45 (dp45.cu:6-169): order of accuracy, the stability
polynomial, dense output, and every branch of the step control."""
import ctypes
from decimal import Decimal, getcontext

import numpy as np
import pytest

from rdcu_amd import Integrator, Problem, Term

D = ctypes.POINTER(ctypes.c_double)


def scalar_problem(terms, n=1):
    """ODE on a 2-point 1-D grid with no derivative factors: every grid point
    evolves independently."""
    return Problem(n, [2], [1.0], terms)


def test_linear_step_is_dp5_stability_polynomial(olib):
    """One fixed step of y' = lam*y multiplies y by the 5th-order DP stability
    polynomial R(z) = sum_{k<=5} z^k/k! + z^6/600."""
    for lam, h in [(-1.0, 0.1), (2.0, 0.05), (-30.0, 0.01), (0.5, 0.7)]:
        prob = scalar_problem([Term(lam, 0, [(0, 1)])])
        with Integrator(olib, prob, np.array([1.0, -2.0]), fixed=True) as it:
            t, hh, acc = it.step(0.0, h)
            y = it.state()
        z = lam * h
        R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24 + z ** 5 / 120 + z ** 6 / 600
        assert acc and t == h and hh == h
        assert abs(y[0] - R) <= 2e-15 * abs(R) and abs(y[1] + 2 * R) <= 2e-15 * abs(2 * R)


def test_fifth_order_convergence(olib):
    """Oscillator u0' = u1, u1' = -u0 from (1, 0): exact (cos t, -sin t).
    The global error of fixed steps falls by 2^5 per halving."""
    prob = scalar_problem([Term(1.0, 0, [(1, 1)]), Term(-1.0, 1, [(0, 1)])], n=2)
    errs = []
    for nsteps in (8, 16, 32):
        h = 2.0 / nsteps
        with Integrator(olib, prob, np.array([1.0, 1.0, 0.0, 0.0]), fixed=True) as it:
            t, _ = it.run(0.0, h, 2.0 - 1e-9)
            y = it.state()
        errs.append(max(abs(y[0] - np.cos(t)), abs(y[2] + np.sin(t))))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 28 < r1 < 36 and 28 < r2 < 36, (errs, r1, r2)


def test_dense_output(olib):
    """x=0 gives y_old exactly; x=1 gives y_new to roundoff; interior points
    are 4th-order accurate (error falls ~2^5 per halving of h)."""
    prob = scalar_problem([Term(1.0, 0, [(0, 2)])])
    errs = []
    for h in (0.2, 0.1, 0.05):
        with Integrator(olib, prob, np.array([0.5, 0.5]), fixed=True) as it:
            y0 = it.state().copy()
            t, _, _ = it.step(0.0, h)
            y1 = it.state()
            assert np.array_equal(it.eval(0.0), y0)
            assert np.abs(it.eval(t) - y1).max() <= 4e-16 * np.abs(y1).max()
            te = 0.37 * h
            errs.append(abs(it.eval(te)[0] - 0.5 / (1 - 0.5 * te)))
    assert errs[0] / errs[1] > 20 and errs[1] / errs[2] > 20, errs


def test_error_estimate_order(olib):
    """The embedded error estimate of y' = -y scales as h^5."""
    prob = scalar_problem([Term(-1.0, 0, [(0, 1)])])
    norms = []
    for h in (0.04, 0.02, 0.01):
        with Integrator(olib, prob, np.array([1.0, 1.0]), atol=1e-8, rtol=1e-8) as it:
            it.step(0.0, h)
            norms.append(olib.oracle_last_norm(it.h))
    assert 28 < norms[0] / norms[1] < 36 and 28 < norms[1] / norms[2] < 36, norms


def pow_m02_exact(x):
    getcontext().prec = 50
    return float(Decimal(x) ** Decimal(-0.2))


def test_accept_growth_formula(olib):
    """Accepted: t += h; h *= min(10, 0.9*norm^-0.2); h = min(h, hmax)."""
    prob = scalar_problem([Term(-1.0, 0, [(0, 1)])])
    with Integrator(olib, prob, np.array([1.0, 1.0])) as it:
        t, h, acc = it.step(0.0, 0.05)
        norm = olib.oracle_last_norm(it.h)
    assert acc and t == 0.05 and 0 < norm < 1
    f = 0.9 * pow_m02_exact(norm)
    assert h == 0.05 * min(f, 10.0)


def test_zero_rhs_growth_capped_at_10_and_hmax(olib):
    """norm = 0 -> pow = inf -> factor clamped to 10 (dp45.cu:133-136); hmax
    applies after the growth."""
    prob = scalar_problem([])
    with Integrator(olib, prob, np.array([1.0, 2.0])) as it:
        t, h, acc = it.step(1.0, 0.125)
        assert acc and t == 1.125 and h == 1.25 and olib.oracle_last_norm(it.h) == 0.0
        t, h, acc = it.step(t, h, hmax=3.0)
        assert acc and t == 2.375 and h == 3.0


def test_reject_branch(olib):
    """Rejected: t unchanged, h *= max(0.2, 0.9*norm^-0.2) (dp45.cu:140-145)."""
    prob = scalar_problem([Term(-50.0, 0, [(0, 1)])])
    with Integrator(olib, prob, np.array([1.0, 1.0])) as it:
        y0 = it.state().copy()
        t, h, acc = it.step(0.0, 0.08)
        norm = olib.oracle_last_norm(it.h)
        assert not acc and t == 0.0 and norm >= 1
        assert h == 0.08 * max(0.2, 0.9 * pow_m02_exact(norm))
        assert np.array_equal(it.state(), y0)
        with pytest.raises(Exception):
            it.eval(0.0)  # no accepted step to interpolate
        t, h, acc = it.step(0.0, 10.0)  # huge error: factor clamps at 0.2
        assert not acc and h == 2.0


def test_nan_norm_rejects_with_factor_0_2(olib):
    """A NaN error norm gives factor 0.2 and a rejection (dp45.cu:121-123):
    here u^-1 of a zero field."""
    prob = scalar_problem([Term(1.0, 0, [(0, -1)])])
    with Integrator(olib, prob, np.array([0.0, 0.0])) as it:
        t, h, acc = it.step(0.0, 0.5)
        assert np.isnan(olib.oracle_last_norm(it.h))
        assert not acc and t == 0.0 and h == 0.1


def test_fixed_step_always_accepts(olib):
    prob = scalar_problem([Term(-50.0, 0, [(0, 1)])])
    with Integrator(olib, prob, np.array([1.0, 1.0]), fixed=True) as it:
        t, h, acc = it.step(0.0, 0.5)
        assert acc and t == 0.5 and h == 0.5


def test_run_loop_no_clip_and_hmin_exit(olib):
    """The run loop has no clip to t1 (dp45.cu:160-161 are commented out), so
    the last step overshoots; it also stops once h <= hmin."""
    prob = scalar_problem([Term(-1.0, 0, [(0, 1)])])
    with Integrator(olib, prob, np.array([1.0, 1.0])) as it:
        seen = []
        t, h = it.run(0.0, 0.01, 1.0, callback=lambda t, h, ok, cur: seen.append((t, ok)))
        assert t >= 1.0 and t - 1.0 < h and seen[-1][0] == t
    prob = scalar_problem([Term(-1e4, 0, [(0, 1)])])
    with Integrator(olib, prob, np.array([1.0, 1.0])) as it:
        t, h = it.run(0.0, 0.5, 1.0, hmin=0.2)
        assert t == 0.0 and h <= 0.2


def test_fsal_run_equals_repeated_steps(olib):
    """rdcu_run (k1 computed once, then k1 <- k7) and repeated rdcu_step calls
    give the same sequence: the RHS is deterministic."""
    prob = Problem(2, [64], [30.0], __import__("rdcu_amd").cgle_terms(1))
    y0 = __import__("rdcu_amd").glibc_rand_ic(2, 64, seed=7, A=1.0)
    with Integrator(olib, prob, y0) as a, Integrator(olib, prob, y0) as b:
        log = []
        a.run(0.0, 0.01, 0.5, callback=lambda t, h, ok, cur: log.append((t, h, ok)))
        t, h = 0.0, 0.01
        log2 = []
        while t < 0.5:
            t, h, ok = b.step(t, h)
            log2.append((t, h, ok))
        assert log == log2
        assert np.array_equal(a.state(), b.state())


# Dormand-Prince RK5(4)7M tableau as exact fractions, from the method's
# publication (Dormand & Prince 1980; Hairer, Norsett & Wanner, Table II.5.2):
# stage matrix, 5th-order weights b and 4th-order weights bs.
_F = __import__("fractions").Fraction
_DP_A = [
    [],
    [_F(1, 5)],
    [_F(3, 40), _F(9, 40)],
    [_F(44, 45), _F(-56, 15), _F(32, 9)],
    [_F(19372, 6561), _F(-25360, 2187), _F(64448, 6561), _F(-212, 729)],
    [_F(9017, 3168), _F(-355, 33), _F(46732, 5247), _F(49, 176), _F(-5103, 18656)],
    [_F(35, 384), _F(0), _F(500, 1113), _F(125, 192), _F(-2187, 6784), _F(11, 84)],
]
_DP_BS = [_F(5179, 57600), _F(0), _F(7571, 16695), _F(393, 640), _F(-92097, 339200),
          _F(187, 2100), _F(1, 40)]


def _linear_step_exact(z):
    """Stage values S_j (Y_j = S_j*y0) of one DP45 step of y' = lam*y with
    z = lam*h as an exact Fraction, the 5th-order result R = S_7 (FSAL stage)
    and the embedded error E = z*sum_j (b_j - bs_j) S_j (error = E*y0)."""
    S = []
    for s in range(7):
        S.append(1 + z * sum((a * Sj for a, Sj in zip(_DP_A[s], S)), _F(0)))
    b = _DP_A[6] + [_F(0)]
    E = z * sum(((bj - bsj) * Sj for bj, bsj, Sj in zip(b, _DP_BS, S)), _F(0))
    return S[6], E


def test_error_norm_closed_form(olib):
    """The step-control norm is the scaled RMS of dp45.cu:79-84, 117-119:
        err_i = h*sum_j e_j k_j,i,  w_i = atol + rtol*max(|y_i|, |ynew_i|),
        norm = sqrt(sum_i (err_i/w_i)^2) / sqrt(n*N).
    For y' = lam*y every quantity is a polynomial in z = lam*h times y0, so the
    norm follows exactly from the tableau (the step may be rejected; the norm
    is formed either way).  The cases are chosen so that each plausible
    mistake moves the norm far outside the tolerance in at least one case:
    growth (z > 0, |ynew| > |y|) catches a max that drops ynew, decay the one
    that drops y; atol and rtol differ by orders of magnitude at |y| ~ 1
    (a swap); points of unequal size (dividing by n*N instead of its root);
    n = 2 fields (normalising by N alone)."""
    cases = [(3.0, 0.2, [1.0, -0.25], 1e-3, 1e-2, 1),
             (-4.0, 0.25, [0.75, -2.0], 1e-2, 1e-4, 1),
             (2.5, 0.3, [1.5, 0.125, -0.5, 2.0], 1e-6, 1e-3, 2)]

    def norm(E, R, y0, atol, rtol, weight, div):
        tot = sum((E * yi / weight(_F(atol), _F(rtol), abs(yi), abs(R * yi))) ** 2
                  for yi in map(_F, y0))
        return float(tot / div) ** 0.5

    right = lambda a, r, y, yn: a + r * max(y, yn)
    mistakes = {"max drops ynew": (lambda a, r, y, yn: a + r * y, None),
                "max drops y": (lambda a, r, y, yn: a + r * yn, None),
                "atol/rtol swapped": (lambda a, r, y, yn: r + a * max(y, yn), None),
                "divided by (nN)^2": (right, "sq"),
                "divided by N only": (right, "N")}
    caught = set()
    for lam, h, y0, atol, rtol, n in cases:
        terms = [Term(lam, f, [(f, 1)]) for f in range(n)]
        prob = scalar_problem(terms, n=n)
        with Integrator(olib, prob, np.array(y0), atol=atol, rtol=rtol) as it:
            it.step(0.0, h)
            got = olib.oracle_last_norm(it.h)
        R, E = _linear_step_exact(_F(lam) * _F(h))
        want = norm(E, R, y0, atol, rtol, right, len(y0))
        # the computed error sum cancels O(z) stage terms down to O(z^5):
        # rounding leaves about 1e-13 relative at these z
        assert abs(got - want) <= 1e-11 * want, (lam, h, got, want)
        for name, (wfun, div) in mistakes.items():
            d = {None: len(y0), "sq": len(y0) ** 2, "N": len(y0) // n}[div]
            if abs(norm(E, R, y0, atol, rtol, wfun, d) - want) > 1e-6 * want:
                caught.add(name)
    assert caught == set(mistakes), set(mistakes) - caught
