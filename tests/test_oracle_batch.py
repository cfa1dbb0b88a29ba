"""Batches on the oracle: a batch is its instances stepped side by side, so
every instance matches the same problem integrated alone, bit for bit."""
import numpy as np
import pytest

from rdcu_amd import Batch, Integrator, Problem, RdcuError, cgle_terms, fhn_terms, glibc_rand_ic
from rdcu_amd.sweep import benchmark_manifest


def small_set():
    return [Problem(2, [32, 32], [30.0, 30.0], cgle_terms(2, 1.5, 0.4)),
            Problem(2, [48, 32], [40.0, 25.0], fhn_terms(2, 3.0, 0.05, 1.0)),
            Problem(2, [32, 32], [20.0, 20.0], fhn_terms(2, 6.0, 0.08, 0.7)),
            Problem(2, [256], [60.0], cgle_terms(1, 2.0, 0.8))]


def test_batch_step_and_run_equal_single(olib):
    probs = small_set()
    ys = [glibc_rand_ic(2, p.N, seed=1000 + i, A=1.0) for i, p in enumerate(probs)]
    with Batch(olib, probs, ys) as b:
        tb, hb, ab = b.step(np.zeros(4), np.full(4, 0.01))
        tb, hb, nb = b.run(tb, hb, 0.3)
        states = [b.member(i).state() for i in range(4)]
    for i, (p, y) in enumerate(zip(probs, ys)):
        with Integrator(olib, p, y) as it:
            t, h, a = it.step(0.0, 0.01)
            assert a == ab[i]
            t, h, n = it.advance(t, h, 0.3, max_attempts=1 << 40)
            assert (t, h, n) == (tb[i], hb[i], nb[i])
            assert np.array_equal(it.state(), states[i])


def test_batch_members_step_only_through_batch(olib):
    p = small_set()[0]
    with Batch(olib, [p], [glibc_rand_ic(2, p.N, seed=1, A=1.0)]) as b:
        m = b.member(0)
        with pytest.raises(RdcuError):
            m.step(0.0, 0.01)
        with pytest.raises(RdcuError):
            m.advance(0.0, 0.01, 1.0)


def test_batch_rejects_3d(olib):
    p = Problem(2, [8, 8, 8], [5.0, 5.0, 5.0], cgle_terms(3))
    with pytest.raises(RdcuError):
        Batch(olib, [p], [np.zeros(2 * 512)])


def test_manifest_is_deterministic_and_balanced():
    a = benchmark_manifest(64, seed=0)
    b = benchmark_manifest(64, seed=0)
    assert [(x.kind, x.grid, x.L, x.params, x.seed) for x in a] == \
           [(x.kind, x.grid, x.L, x.params, x.seed) for x in b]
    assert sum(x.kind == "cgle" for x in a) == 32
    assert {x.grid for x in a} <= {64, 96, 128, 192, 256}
    for x in a:
        assert 50 <= x.L <= 200
        if x.kind == "cgle":
            assert 0.5 <= x.params["b"] <= 3 and -1.5 <= x.params["c"] <= 0
            # standard-convention c enters the README dictionary as -c
            assert x.prob.terms[-1].coef == -x.params["c"]
        else:
            assert 2 <= x.params["delta"] <= 10 and 0.01 <= x.params["eps"] <= 0.1


def test_batch_time_on_the_oracle(olib):
    """rdcu_batch_time on the oracle: the wall clock of every instance's RHS,
    no per-class entries."""
    probs = [Problem(2, [32, 32], [20.0, 20.0], cgle_terms(2)),
             Problem(2, [64], [30.0], cgle_terms(1))]
    ys = [glibc_rand_ic(2, p.N, seed=3 + i, A=1.0) for i, p in enumerate(probs)]
    with Batch(olib, probs, ys) as b:
        ms, classes = b.time_rhs(reps=2)
    assert ms > 0 and classes == []
