"""The oracle's discrete adjoint (rdcu_loss_grad) against its tangent-linear
run and against central finite differences, on frozen step schedules; the
frozen forward against the adaptive run whose schedule it replays."""
import numpy as np
import pytest

from conftest import loss_tl
from rdcu_amd import Integrator, Problem, cgle_terms, fhn_terms, glibc_rand_ic


def setup(olib, prob, seed=7, t1=0.4, nout=4, A=1.0):
    y0 = glibc_rand_ic(prob.n, prob.N, seed=seed, A=A)
    it = Integrator(olib, prob, y0, atol=1e-10, rtol=1e-10)
    hs = it.schedule(0.0, 0.01, t1)
    it.set_state(y0)
    t_out = np.linspace(t1 / nout, t1, nout)
    return it, y0, hs, t_out


CASES = [("1d", Problem(2, [64], [30.0], cgle_terms(1, 2.0, 0.8))),
         ("2d-cgle", Problem(2, [16, 24], [12.0, 18.0], cgle_terms(2, 2.0, 0.8))),
         ("2d-fhn", Problem(2, [16, 16], [14.0, 14.0], fhn_terms(2, 4.0, 0.07, 1.3)))]


@pytest.mark.parametrize("name,prob", CASES, ids=[c[0] for c in CASES])
def test_frozen_forward_replays_adaptive_run(olib, name, prob):
    """Outputs of the adaptive run (rdcu_run with stops) are the targets: the
    frozen schedule reproduces them exactly, so the loss is exactly 0."""
    y0 = glibc_rand_ic(prob.n, prob.N, seed=3, A=1.0)
    t_out = np.array([0.1, 0.2, 0.3, 0.4])
    with Integrator(olib, prob, y0, atol=1e-10, rtol=1e-10) as it:
        outs, nxt = [], [0]

        def cb(t, h, ok, cur):
            while ok and nxt[0] < len(t_out) and t >= t_out[nxt[0]]:
                outs.append(it.eval(t_out[nxt[0]]))
                nxt[0] += 1
            return t_out[nxt[0]] if nxt[0] < len(t_out) else np.inf

        it.run(0.0, 0.01, 0.4, stop=t_out[0], callback=cb)
        it.set_state(y0)
        hs = it.schedule(0.0, 0.01, 0.4)
        assert it.loss_grad(y0, 0.0, hs, t_out, np.concatenate(outs), grad=False) == 0.0


@pytest.mark.parametrize("name,prob", CASES, ids=[c[0] for c in CASES])
def test_adjoint_matches_tangent_linear(olib, name, prob):
    it, y0, hs, t_out = setup(olib, prob)
    rng = np.random.default_rng(1)
    coef = np.array([t.coef for t in prob.terms])
    target = y0[None, :] * 0.9 + 0.01 * rng.standard_normal((len(t_out), y0.size))
    L, gc, gy = it.loss_grad(y0, 0.0, hs, t_out, target)
    for trial in range(3):
        dc = rng.standard_normal(coef.size)
        dy = rng.standard_normal(y0.size)
        Lt, dL = loss_tl(olib, it, y0, 0.0, hs, t_out, target, dc, dy)
        assert Lt == pytest.approx(L, rel=1e-14)
        adj = float(gc @ dc + gy @ dy)
        assert adj == pytest.approx(dL, rel=1e-9), (trial, adj, dL)
    it.close()


def test_adjoint_matches_central_differences(olib):
    prob = Problem(2, [16, 16], [12.0, 12.0], cgle_terms(2, 2.0, 0.8))
    it, y0, hs, t_out = setup(olib, prob)
    rng = np.random.default_rng(2)
    coef = np.array([t.coef for t in prob.terms])
    target = 0.5 * y0[None, :] + 0.02 * rng.standard_normal((len(t_out), y0.size))
    L, gc, gy = it.loss_grad(y0, 0.0, hs, t_out, target)
    eps = 1e-6
    for t in range(len(coef)):
        cp, cm = coef.copy(), coef.copy()
        cp[t] += eps
        cm[t] -= eps
        it.set_coef(cp)
        Lp = it.loss_grad(y0, 0.0, hs, t_out, target, grad=False)
        it.set_coef(cm)
        Lm = it.loss_grad(y0, 0.0, hs, t_out, target, grad=False)
        fd = (Lp - Lm) / (2 * eps)
        assert fd == pytest.approx(gc[t], rel=1e-6, abs=1e-6 * np.abs(gc).max()), t
    it.set_coef(coef)
    for i in (0, 17, 300, 511):
        yp, ym = y0.copy(), y0.copy()
        yp[i] += eps
        ym[i] -= eps
        fd = (it.loss_grad(yp, 0.0, hs, t_out, target, grad=False) -
              it.loss_grad(ym, 0.0, hs, t_out, target, grad=False)) / (2 * eps)
        assert fd == pytest.approx(gy[i], rel=1e-6, abs=1e-6 * np.abs(gy).max()), i
    it.close()


def _batch_instances(lib, prob, count, t1=0.3, nout=3):
    """count instances of one problem with perturbed coefficients and their
    own initial states, schedules and targets."""
    rng = np.random.default_rng(11)
    base = np.array([t.coef for t in prob.terms])
    its, y0s, hss, touts, tgs = [], [], [], [], []
    for i in range(count):
        y0 = glibc_rand_ic(prob.n, prob.N, seed=20 + i, A=1.0)
        it = Integrator(lib, prob, y0, atol=1e-10, rtol=1e-10)
        it.set_coef(base * (1 + 0.05 * rng.standard_normal(base.size)))
        hs = it.schedule(0.0, 0.01, t1)
        it.set_state(y0)
        t_out = np.linspace(t1 / nout, t1, nout)
        its.append(it)
        y0s.append(y0)
        hss.append(hs)
        touts.append(t_out)
        tgs.append(y0[None, :] * 0.9 + 0.01 * rng.standard_normal((nout, y0.size)))
    return its, y0s, hss, touts, tgs


def test_batch_gradients_equal_single_calls(olib):
    """rdcu_loss_grad_batch gives each instance what rdcu_loss_grad gives it."""
    from rdcu_amd import loss_grad_batch
    prob = CASES[1][1]
    its, y0s, hss, touts, tgs = _batch_instances(olib, prob, 3)
    got = loss_grad_batch(its, y0s, [0.0] * 3, hss, touts, tgs, concurrency=2)
    for i, it in enumerate(its):
        L, gc, gy = it.loss_grad(y0s[i], 0.0, hss[i], touts[i], tgs[i])
        assert got[i][0] == L
        assert np.array_equal(got[i][1], gc) and np.array_equal(got[i][2], gy)
    losses = loss_grad_batch(its, y0s, [0.0] * 3, hss, touts, tgs, grad=False)
    assert losses == [g[0] for g in got]
    for it in its:
        it.close()


def test_batch_gradients_reject_shared_handles_and_report_failures(olib):
    from rdcu_amd import RdcuError, loss_grad_batch
    prob = CASES[1][1]
    its, y0s, hss, touts, tgs = _batch_instances(olib, prob, 2)
    with pytest.raises(ValueError):
        loss_grad_batch([its[0], its[0]], y0s, [0.0] * 2, hss, touts, tgs)
    bad = [touts[0], touts[1] + 100.0]  # second instance: output beyond its schedule
    with pytest.raises(RdcuError) as e:
        loss_grad_batch(its, y0s, [0.0] * 2, hss, bad, tgs)
    assert e.value.code == 1 and "beyond" in str(e.value)
    for it in its:
        it.close()
