"""GPU batches: every instance of a batch (mixed grids, dictionaries and step
histories in one graph) follows the single-instance arithmetic, so it equals
the oracle bit for bit, and equals the same problem run alone on the GPU."""
import numpy as np
import pytest

from rdcu_amd import Batch, Integrator, Problem, cgle_terms, fhn_terms, glibc_rand_ic
from rdcu_amd.sweep import benchmark_manifest

pytestmark = pytest.mark.gpu


def mixed_set():
    return [Problem(2, [64, 64], [50.0, 50.0], cgle_terms(2, 1.5, 0.4)),
            Problem(2, [96, 96], [80.0, 80.0], fhn_terms(2, 3.0, 0.05, 1.0)),
            Problem(2, [64, 64], [40.0, 40.0], fhn_terms(2, 6.0, 0.08, 0.7)),
            Problem(2, [128, 64], [90.0, 45.0], cgle_terms(2, 2.5, -0.3)),
            Problem(2, [96, 96], [60.0, 60.0], cgle_terms(2, 0.7, 1.1)),
            Problem(2, [256], [60.0], cgle_terms(1, 2.0, 0.8))]


@pytest.mark.parametrize("variant", [None, 0, 1, 2, 3],
                         ids=["tuned", "xinv-plain", "xinv-twreg", "xinv-lean", "xinv-lean-twreg"])
def test_batch_steps_bitwise_vs_oracle(hlib, olib, monkeypatch, variant):
    """With the tuner's choice, and with each generated x-inverse variant
    (the tuner restricted to it)."""
    if variant is not None:
        monkeypatch.setenv("RDCU_BATCH_XINV", str(variant))
    probs = mixed_set()
    ys = [glibc_rand_ic(2, p.N, seed=1000 + i, A=1.0) for i, p in enumerate(probs)]
    n = len(probs)
    print(f"case batch steps, x-inverse variant {variant}", flush=True)
    with Batch(hlib, probs, ys) as g, Batch(olib, probs, ys) as o:
        tg, hg = np.zeros(n), np.full(n, 0.01)
        to, ho = tg.copy(), hg.copy()
        for k in range(12):
            tg, hg, ag = g.step(tg, hg)
            to, ho, ao = o.step(to, ho)
            assert np.array_equal(tg, to) and np.array_equal(hg, ho), k
            assert np.array_equal(ag, ao), k
        for i in range(n):
            assert np.array_equal(g.member(i).state(), o.member(i).state()), i
        # run to different end times per instance
        t1 = np.linspace(0.3, 0.6, n)
        tg, hg, ng = g.run(tg, hg, t1)
        to, ho, no = o.run(to, ho, t1)
        assert np.array_equal(tg, to) and np.array_equal(hg, ho)
        assert np.array_equal(ng, no)
        for i in range(n):
            assert np.array_equal(g.member(i).state(), o.member(i).state()), i
            mg, mo = g.member(i), o.member(i)
            f = mg.rhs(mg.state())
            assert np.array_equal(f, mo.rhs(mo.state())), i


def test_batch_rhs_timing_reports_classes_and_leaves_results_unchanged(hlib, olib):
    """rdcu_batch_time: one entry per shape class (by dimension, then grid) with its
    instance count, positive times and the algorithmic bytes of the stated
    formula; timing between attempts leaves the steps equal to the oracle's."""
    probs = mixed_set()
    ys = [glibc_rand_ic(2, p.N, seed=1000 + i, A=1.0) for i, p in enumerate(probs)]
    n = len(probs)
    print("case batch timing", flush=True)
    with Batch(hlib, probs, ys) as g, Batch(olib, probs, ys) as o:
        tg, hg = np.zeros(n), np.full(n, 0.01)
        to, ho = tg.copy(), hg.copy()
        tg, hg, ag = g.step(tg, hg)
        to, ho, ao = o.step(to, ho)
        ms, classes = g.time_rhs(reps=5)
        assert ms > 0
        grids = [c["grid"] for c in classes]
        assert grids == [[256, 1], [64, 64], [96, 96], [128, 64]]
        assert [c["instances"] for c in classes] == [1, 2, 2, 1]
        for c in classes:
            assert c["ms"]["row_fwd"] > 0 and c["ms"]["row_inv"] > 0
            assert (c["ms"]["col_mult"] > 0) == (c["ndim"] == 2)
        # 2 fields; the CGLE and FitzHugh-Nagumo dictionaries reference the
        # second derivatives of both fields (4 factors in 2-D, 2 in 1-D)
        N = 64 * 64
        assert classes[1]["bytes"] == {"row_fwd": 2 * 16 * 2 * N, "col_mult": 2 * (16 * N + 32 * N),
                                       "row_inv": 2 * (32 * N + 32 * N)}
        assert classes[0]["bytes"] == {"row_fwd": 16 * 2 * 256, "col_mult": 0.0,
                                       "row_inv": 24 * 2 * 256}
        ms_o, cls_o = o.time_rhs(reps=1)
        assert ms_o > 0 and cls_o == []
        for k in range(6):
            tg, hg, ag = g.step(tg, hg)
            to, ho, ao = o.step(to, ho)
            assert np.array_equal(tg, to) and np.array_equal(hg, ho) and np.array_equal(ag, ao), k
        for i in range(n):
            assert np.array_equal(g.member(i).state(), o.member(i).state()), i


def test_batch_equals_single_gpu_runs(hlib):
    """The sweep manifest's first instances: batched run == each alone."""
    insts = benchmark_manifest(512, seed=0)[::40]  # 13 instances, both kinds, mixed grids
    n = len(insts)
    t1 = 0.2
    print("case batch vs single", n, flush=True)
    with Batch(hlib, [x.prob for x in insts], [x.y0() for x in insts]) as b:
        tb, hb, nb = b.run(np.zeros(n), np.full(n, 0.01), t1)
        yb = [b.member(i).state() for i in range(n)]
    for i, x in enumerate(insts):
        with Integrator(hlib, x.prob, x.y0()) as it:
            t, h, na = it.advance(0.0, 0.01, t1, max_attempts=1 << 40)
            assert (t, h, na) == (tb[i], hb[i], nb[i]), i
            assert np.array_equal(it.state(), yb[i]), i
