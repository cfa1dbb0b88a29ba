"""Monomial dictionaries of the benchmark problems.

Factor indices follow rdcu's map (rdcu.cu:769-785): index = block*n + field
with block 0 = u, 1+a = d/dx_a, 1+ndim+a*ndim+b = d/dx_b d/dx_a.
"""
from __future__ import annotations

from typing import List

from .abi import Term


def lap_index(ndim: int, n: int, axis: int, field: int) -> int:
    """Index of d^2 u_field / dx_axis^2."""
    blk = 1 + ndim + axis * ndim + axis
    return blk * n + field


def cgle_terms(ndim: int, b: float = 2.0, c: float = 0.8) -> List[Term]:
    """Complex Ginzburg-Landau A_t = A + (1+ib) lap A - (1-ic)|A|^2 A as two
    real fields u0 = Re A, u1 = Im A, in the convention and term order of the
    reference's example (README.md:67; makelines.py:248-252 for 3D):
        u0' = u0 + lap u0 - b lap u1 - u0^3 - u0 u1^2 - c u0^2 u1 - c u1^3
        u1' = u1 + lap u1 + b lap u0 - u0^2 u1 - u1^3 + c u0^3 + c u0 u1^2
    2 + 2*ndim + 8 terms (14 in 1D, 18 in 2D, 22 in 3D)."""
    n = 2
    L0 = [lap_index(ndim, n, a, 0) for a in range(ndim)]
    L1 = [lap_index(ndim, n, a, 1) for a in range(ndim)]
    t = [Term(1.0, 0, [(0, 1)])]
    t += [Term(1.0, 0, [(i, 1)]) for i in L0]
    t += [Term(-b, 0, [(i, 1)]) for i in L1]
    t += [Term(-1.0, 0, [(0, 3)]), Term(-1.0, 0, [(0, 1), (1, 2)]),
          Term(-c, 0, [(0, 2), (1, 1)]), Term(-c, 0, [(1, 3)])]
    t += [Term(1.0, 1, [(1, 1)])]
    t += [Term(1.0, 1, [(i, 1)]) for i in L1]
    t += [Term(b, 1, [(i, 1)]) for i in L0]
    t += [Term(-1.0, 1, [(0, 2), (1, 1)]), Term(-1.0, 1, [(1, 3)]),
          Term(c, 1, [(0, 3)]), Term(c, 1, [(0, 1), (1, 2)])]
    return t


def fhn_terms(ndim: int, delta: float, eps: float, a1: float) -> List[Term]:
    """FitzHugh-Nagumo u_t = lap u + u - u^3 - v, v_t = delta lap v + eps (u - a1 v)."""
    n = 2
    t = [Term(1.0, 0, [(lap_index(ndim, n, a, 0), 1)]) for a in range(ndim)]
    t += [Term(1.0, 0, [(0, 1)]), Term(-1.0, 0, [(0, 3)]), Term(-1.0, 0, [(1, 1)])]
    t += [Term(delta, 1, [(lap_index(ndim, n, a, 1), 1)]) for a in range(ndim)]
    t += [Term(eps, 1, [(0, 1)]), Term(-(eps * a1), 1, [(1, 1)])]
    return t


def to_cli(terms: List[Term]) -> List[str]:
    """``-c`` argument strings of a dictionary."""
    out = []
    for t in terms:
        s = [repr(float(t.coef)), str(t.field)]
        for i, p in t.factors:
            s += [str(i), str(p)]
        out.append(",".join(s))
    return out
