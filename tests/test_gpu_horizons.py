"""GPU parity at the benchmark configurations' own horizons.

  2-D CGLE 512x512 to t = 100 (outputs every 1): every attempt's (t, h,
      accepted) and every output's state and RHS frame bit for bit against
      the oracle's run of the same workload (fixture tests/golden/
      horizon_c2.npz, written by tests/golden/horizon_c2.py with the oracle).
  3-D CGLE 256^3 to t = 20: local parity along the GPU's own trajectory --
      at an early, a middle and the final accepted step the GPU's (t, h, y)
      is handed to the oracle and one attempt from it is compared bit for bit
      with the GPU's next logged attempt and state.
  Batched sweep, 512 instances to t = 0.5: a stratified subset (two instances
      of every kind x grid class, 20 in all) against the oracle run alone,
      bit for bit (t, h, attempt count, final state).
  Adjoint, 2-D CGLE 256x256 to t = 10: the gradient of the whole schedule
      (20 outputs) against the oracle's tangent-linear run in two random
      directions, to 1e-6 relative.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, loss_tl
from rdcu_amd import Batch, Integrator, Problem, cgle_terms, glibc_rand_ic
from rdcu_amd.sweep import GRIDS, benchmark_manifest, run_serial

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import horizon_c2  # noqa: E402

pytestmark = pytest.mark.gpu


def test_c2_full_horizon_bitwise(hlib):
    fx = np.load(horizon_c2.FIXTURE)
    t1 = float(fx["t1"])
    print(f"case c2 512x512 to t={t1}", flush=True)
    t, h, acc, frames, final = horizon_c2.run_horizon(hlib, t1)
    print(f"{t.size} attempts, {int(acc.sum())} accepted", flush=True)
    assert t.size == fx["t"].size
    assert np.array_equal(t, fx["t"]) and np.array_equal(h, fx["h"])
    assert np.array_equal(acc, fx["acc"])
    assert [f[0] for f in frames] == list(fx["y_sha"])
    assert [f[1] for f in frames] == list(fx["f_sha"])
    assert final == str(fx["final_sha"])


def test_c3_local_parity_along_the_horizon(hlib, olib):
    prob = Problem(2, [256, 256, 256], [100.0, 100.0, 100.0], cgle_terms(3, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=1, A=0.1)
    stops = [0.5, 10.0, 20.0]
    print("case c3 256^3 to t=20", flush=True)
    log, cps = [], []
    with Integrator(hlib, prob, y0) as g:
        def cb(t, h, ok, cur):
            log.append((t, h, ok))
            if ok and cur and len(cps) < len(stops) and t >= stops[len(cps)]:
                cps.append((len(log), t, h, g.state()))
            return stops[len(cps)] if len(cps) < len(stops) else np.inf

        g.run(0.0, 0.01, 20.0, stop=stops[0], callback=cb)
        print(f"{len(log)} attempts to t={log[-1][0]:.4f}", flush=True)
        assert len(cps) == 3 and log[-1][0] >= 20.0
        # the final checkpoint has no logged successor: continue the GPU run
        # by one attempt from it
        t_end, h_end = log[-1][0], log[-1][1]
        nxt_final = g.step(t_end, h_end)
        y_final_next = g.state()
    with Integrator(olib, prob, y0) as o:
        for k, (i, t, h, y) in enumerate(cps):
            o.set_state(y)
            got = o.step(t, h)
            if k < 2:
                assert got == log[i], (k, got, log[i])
            else:
                assert got == nxt_final, (got, nxt_final)
                assert np.array_equal(o.state(), y_final_next)
            print(f"checkpoint t={t:.4f}: oracle attempt {got} matches", flush=True)


def test_c4_stratified_subset_vs_oracle(hlib, olib):
    insts = benchmark_manifest(512, seed=0)
    picks = []
    for kind in ("cgle", "fhn"):
        for gsz in GRIDS:
            idx = [i for i, x in enumerate(insts) if x.kind == kind and x.grid == gsz]
            picks += [idx[0], idx[len(idx) // 2]]
    assert len(picks) == 20
    n = len(insts)
    print("case c4 512 instances to t=0.5", flush=True)
    with Batch(hlib, [x.prob for x in insts], [x.y0() for x in insts]) as b:
        t, h, na = b.run(np.zeros(n), np.full(n, 0.01), 0.5)
        yb = {i: b.member(i).state() for i in picks}
    for i in picks:
        to, ho, no, yo = run_serial(olib, insts[i], 0.5)
        assert (to, ho, no) == (t[i], h[i], na[i]), (i, insts[i].kind, insts[i].grid)
        assert np.array_equal(yo, yb[i]), i
        print(f"instance {i} ({insts[i].kind} {insts[i].grid}^2, {no} attempts) matches",
              flush=True)


def test_c5_tangent_linear_over_the_whole_schedule(hlib, olib):
    from test_gpu_adjoint import fit_problem
    prob = Problem(2, [256, 256], [100.0, 100.0], cgle_terms(2, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=1, A=1.0)
    t_out = np.arange(1, 21) * 0.5
    print("case c5 tangent-linear t=10", flush=True)
    g, coef, target, hs = fit_problem(prob, y0, t_out, 10.0, hlib, 1e-10)
    L, gc, gy = g.loss_grad(y0, 0.0, hs, t_out, target)
    g.close()
    with Integrator(olib, prob, y0, atol=1e-10, rtol=1e-10) as o:
        rng = np.random.default_rng(17)
        for k in range(2):
            dc, dy = rng.standard_normal(coef.size), rng.standard_normal(y0.size)
            Lt, dL = loss_tl(olib, o, y0, 0.0, hs, t_out, target, dc, dy)
            assert Lt == pytest.approx(L, rel=1e-12)
            got = float(gc @ dc + gy @ dy)
            print(f"direction {k}: adjoint {got!r} tangent-linear {dL!r}", flush=True)
            assert got == pytest.approx(dL, rel=1e-6)
