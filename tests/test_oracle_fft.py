"""Pins the oracle's transforms: brute-force DFT, numpy.fft (pocketfft),
Parseval, round trips, the radix plan and the twiddle table."""
import ctypes

import numpy as np
import pytest

D = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(D)


def fft(olib, x, direction=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    assert olib.oracle_fft(len(x), direction, _p(x.view(np.float64)), _p(out.view(np.float64))) == 0
    return out


def r2c(olib, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(len(x) // 2 + 1, dtype=np.complex128)
    assert olib.oracle_r2c(len(x), _p(x), _p(out.view(np.float64))) == 0
    return out


def c2r(olib, X, N):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    out = np.empty(N)
    assert olib.oracle_c2r(N, _p(X.view(np.float64)), _p(out)) == 0
    return out


def plan(olib, L):
    rad = (ctypes.c_int * 16)()
    n = olib.oracle_plan(L, rad, 16)
    return None if n < 0 else list(rad[:n])


def root_decimal(m, L):
    """cos(2 pi m/L), -sin(2 pi m/L) from Taylor series in 40-digit Decimal."""
    from decimal import Decimal, getcontext
    getcontext().prec = 40
    pi = Decimal("3.141592653589793238462643383279502884197")
    x = 2 * pi * m / L
    c, s, term, k = Decimal(0), Decimal(0), Decimal(1), 0
    while True:
        if k % 4 == 0:
            c += term
        elif k % 4 == 1:
            s += term
        elif k % 4 == 2:
            c -= term
        else:
            s -= term
        k += 1
        term = term * x / k
        if abs(term) < Decimal(10) ** -38:
            break
    return float(c), float(-s)


def dft(x, sign=-1):
    L = len(x)
    j = np.arange(L)
    W = np.exp(sign * 2j * np.pi * np.outer(j, j) / L)
    return W @ x


@pytest.mark.parametrize("L,expect", [(1, []), (2, [2]), (4, [4]), (8, [8]), (32, [4, 8]),
                                      (64, [8, 8]), (96, [3, 4, 8]), (128, [2, 8, 8]),
                                      (192, [3, 8, 8]), (256, [4, 8, 8]), (512, [8, 8, 8]),
                                      (1000, [5, 5, 5, 8]), (7, None), (22, None)])
def test_radix_plan(olib, L, expect):
    assert plan(olib, L) == expect


@pytest.mark.parametrize("L", [2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 24, 30, 32, 40, 48, 60, 64, 96])
def test_fft_matches_bruteforce_dft(olib, rng, L):
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    for sign in (-1, 1):
        got = fft(olib, x, sign)
        ref = dft(x, sign)
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max() * max(1, np.log2(L))


@pytest.mark.parametrize("L", [128, 160, 192, 256, 384, 512, 768, 1024, 1280, 2048])
def test_fft_matches_numpy(olib, rng, L):
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    got = fft(olib, x, -1)
    ref = np.fft.fft(x)
    assert np.abs(got - ref).max() <= 2e-15 * np.log2(L) * np.abs(ref).max()
    back = fft(olib, got, +1) / L
    assert np.abs(back - x).max() <= 4e-15 * np.log2(L) * np.abs(x).max()
    # Parseval
    assert abs(np.vdot(got, got).real / L - np.vdot(x, x).real) <= 1e-12 * np.vdot(x, x).real


def test_fft_impulse_and_constant_exact(olib):
    """Impulse -> all ones and constant -> L*delta are exact for radix-2^k plans."""
    for L in (8, 64, 512):
        e = np.zeros(L, complex)
        e[0] = 1
        assert np.array_equal(fft(olib, e), np.ones(L, complex))
        assert np.array_equal(fft(olib, np.ones(L, complex)), L * e)


@pytest.mark.parametrize("N", [2, 4, 6, 8, 10, 16, 24, 48, 64, 96, 128, 192, 256, 512, 1024])
def test_real_transforms_match_numpy(olib, rng, N):
    x = rng.standard_normal(N)
    X = r2c(olib, x)
    ref = np.fft.rfft(x)
    assert np.abs(X - ref).max() <= 4e-15 * max(1, np.log2(N)) * np.abs(ref).max()
    # DC and Nyquist bins of a real signal are real
    assert X[0].imag == 0.0
    back = c2r(olib, X, N)
    assert np.abs(back / N - x).max() <= 8e-15 * max(1, np.log2(N)) * np.abs(x).max()


def test_c2r_ignores_dc_and_nyquist_imaginary(olib, rng):
    """Like numpy.fft.irfft, the imaginary parts of bins 0 and N/2 are ignored."""
    N = 64
    X = np.fft.rfft(rng.standard_normal(N))
    Y = X.copy()
    Y[0] += 3.0j
    Y[-1] -= 5.0j
    assert np.array_equal(c2r(olib, X, N), c2r(olib, Y, N))
    assert np.abs(c2r(olib, Y, N) / N - np.fft.irfft(Y, N)).max() < 1e-14


def test_unit_root_table(olib):
    out = (ctypes.c_double * 2)()
    for L in (8, 12, 96, 512):
        for q in range(4):
            olib.oracle_unit_root(q * L // 4, L, out)
            assert (out[0], out[1]) == [(1.0, 0.0), (0.0, -1.0), (-1.0, 0.0), (0.0, 1.0)][q]
        for m in range(L):
            olib.oracle_unit_root(m, L, out)
            c, s = root_decimal(m, L)
            # within half an ulp of the true value (values in [-1, 1])
            assert abs(out[0] - c) <= 1.12e-16 and abs(out[1] - s) <= 1.12e-16
