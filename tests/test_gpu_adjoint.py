"""GPU discrete adjoint (rdcu_loss_grad) against the oracle's adjoint and
tangent-linear runs and against central finite differences, and the frozen
forward against the adaptive run it replays."""
import numpy as np
import pytest

from conftest import loss_tl
from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic

pytestmark = pytest.mark.gpu


def outputs_of_run(it, t_out, t1):
    outs, nxt = [], [0]

    def cb(t, h, ok, cur):
        while ok and nxt[0] < len(t_out) and t >= t_out[nxt[0]]:
            outs.append(it.eval(t_out[nxt[0]]))
            nxt[0] += 1
        return t_out[nxt[0]] if nxt[0] < len(t_out) else np.inf

    it.run(0.0, 0.01, t1, stop=t_out[0], callback=cb)
    return np.concatenate(outs)


def fit_problem(prob, y0, t_out, t1, lib, rtol, noise=0.05, seed=0):
    """Target from perturbed coefficients; schedule of the unperturbed model."""
    rng = np.random.default_rng(seed)
    coef = np.array([t.coef for t in prob.terms])
    pert = coef * (1 + noise * rng.standard_normal(coef.size))
    with Integrator(lib, prob, y0, atol=rtol, rtol=rtol) as it:
        it.set_coef(pert)
        target = outputs_of_run(it, t_out, t1)
    it = Integrator(lib, prob, y0, atol=rtol, rtol=rtol)
    hs = it.schedule(0.0, 0.01, t1)
    return it, coef, target, hs


def test_schedule_and_frozen_forward_bitwise(hlib, olib):
    prob = Problem(2, [64, 64], [50.0, 50.0], cgle_terms(2, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=4, A=1.0)
    t_out = np.array([0.1, 0.2, 0.3])
    print("case frozen forward", flush=True)
    with Integrator(hlib, prob, y0, atol=1e-10, rtol=1e-10) as g, \
            Integrator(olib, prob, y0, atol=1e-10, rtol=1e-10) as o:
        outs = outputs_of_run(g, t_out, 0.3)
        g.set_state(y0)
        hg = g.schedule(0.0, 0.01, 0.3)
        ho = o.schedule(0.0, 0.01, 0.3)
        assert np.array_equal(hg, ho) and hg.size > 10
        assert g.loss_grad(y0, 0.0, hg, t_out, outs, grad=False) == 0.0
        tgt = outs * 0.9 + 0.01
        Lg = g.loss_grad(y0, 0.0, hg, t_out, tgt, grad=False)
        Lo = o.loss_grad(y0, 0.0, ho, t_out, tgt, grad=False)
        assert Lg == pytest.approx(Lo, rel=1e-13)


def test_gradients_match_oracle_adjoint_and_tangent_linear(hlib, olib):
    prob = Problem(2, [48, 64], [40.0, 55.0], cgle_terms(2, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=5, A=1.0)
    t_out = np.array([0.1, 0.2, 0.3])
    print("case gradients vs oracle", flush=True)
    g, coef, target, hs = fit_problem(prob, y0, t_out, 0.3, hlib, 1e-10)
    Lg, gcg, gyg = g.loss_grad(y0, 0.0, hs, t_out, target)
    g.close()
    with Integrator(olib, prob, y0, atol=1e-10, rtol=1e-10) as o:
        Lo, gco, gyo = o.loss_grad(y0, 0.0, hs, t_out, target)
        assert Lg == pytest.approx(Lo, rel=1e-12)
        assert np.abs(gcg - gco).max() <= 1e-10 * np.abs(gco).max()
        assert np.abs(gyg - gyo).max() <= 1e-10 * np.abs(gyo).max()
        rng = np.random.default_rng(3)
        for _ in range(2):
            dc, dy = rng.standard_normal(coef.size), rng.standard_normal(y0.size)
            _, dL = loss_tl(olib, o, y0, 0.0, hs, t_out, target, dc, dy)
            assert float(gcg @ dc + gyg @ dy) == pytest.approx(dL, rel=1e-8)


@pytest.mark.parametrize("store", ["1", "0"])
def test_3d_gradients_match_oracle_adjoint(hlib, olib, monkeypatch, store):
    """3-D CGLE (16x24x32): the reverse sweep's seven stage spectra run as one
    problem of 7n fields, whose factor list is split into the factors without
    k_z (the y pass) and the others (the z pass); gradients against the
    oracle's adjoint, with the stages stored and recomputed."""
    monkeypatch.setenv("RDCU_ADJ_STORE", store)
    prob = Problem(2, [16, 24, 32], [10.0, 20.0, 30.0], cgle_terms(3, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=6, A=1.0)
    t_out = np.array([0.05, 0.1])
    print(f"case 3-D gradients vs oracle store={store}", flush=True)
    g, coef, target, hs = fit_problem(prob, y0, t_out, 0.1, hlib, 1e-10)
    Lg, gcg, gyg = g.loss_grad(y0, 0.0, hs, t_out, target)
    g.close()
    assert hs.size > 3
    with Integrator(olib, prob, y0, atol=1e-10, rtol=1e-10) as o:
        Lo, gco, gyo = o.loss_grad(y0, 0.0, hs, t_out, target)
    assert Lg == pytest.approx(Lo, rel=1e-12)
    assert np.abs(gcg - gco).max() <= 1e-10 * np.abs(gco).max()
    assert np.abs(gyg - gyo).max() <= 1e-10 * np.abs(gyo).max()


def test_adjoint_config_vs_central_differences(hlib, olib):
    """2-D CGLE 256x256, L=100, t=10, 20 outputs, rtol=atol=1e-10, target from
    coefficients with 5% relative noise: every coefficient gradient against
    central differences (h=1e-6) on the frozen schedule; the full gradient
    against the oracle's tangent-linear run over the first half time unit."""
    prob = Problem(2, [256, 256], [100.0, 100.0], cgle_terms(2, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=1, A=1.0)
    t_out = np.arange(1, 21) * 0.5
    print("case adjoint 256x256 t=10", flush=True)
    g, coef, target, hs = fit_problem(prob, y0, t_out, 10.0, hlib, 1e-10)
    L, gc, gy = g.loss_grad(y0, 0.0, hs, t_out, target)
    assert L > 0 and np.all(np.isfinite(gc)) and np.all(np.isfinite(gy))
    eps = 1e-6
    for t in range(coef.size):
        cp, cm = coef.copy(), coef.copy()
        cp[t] += eps
        cm[t] -= eps
        g.set_coef(cp)
        Lp = g.loss_grad(y0, 0.0, hs, t_out, target, grad=False)
        g.set_coef(cm)
        Lm = g.loss_grad(y0, 0.0, hs, t_out, target, grad=False)
        fd = (Lp - Lm) / (2 * eps)
        assert fd == pytest.approx(gc[t], rel=1e-6, abs=1e-6 * np.abs(gc).max()), (t, fd, gc[t])
    g.set_coef(coef)
    # tangent-linear check over the steps up to t = 0.5 (one output)
    k = int(np.searchsorted(np.cumsum(hs), 0.5)) + 1
    L1, gc1, gy1 = g.loss_grad(y0, 0.0, hs[:k], t_out[:1], target[:y0.size])
    g.close()
    with Integrator(olib, prob, y0, atol=1e-10, rtol=1e-10) as o:
        rng = np.random.default_rng(9)
        dc, dy = rng.standard_normal(coef.size), rng.standard_normal(y0.size)
        Lt, dL = loss_tl(olib, o, y0, 0.0, hs[:k], t_out[:1], target[:y0.size], dc, dy)
    assert Lt == pytest.approx(L1, rel=1e-12)
    assert float(gc1 @ dc + gy1 @ dy) == pytest.approx(dL, rel=1e-6)


def test_batched_gradients_equal_single_calls(hlib, olib):
    """rdcu_loss_grad_batch: 6 instances (own coefficients, states, schedules
    and targets) three at a time on the GPU give bit for bit what one
    rdcu_loss_grad call gives each; the first also agrees with the oracle."""
    from rdcu_amd import loss_grad_batch
    prob = Problem(2, [32, 48], [30.0, 40.0], cgle_terms(2, 2.0, 0.8))
    rng = np.random.default_rng(21)
    base = np.array([t.coef for t in prob.terms])
    its, y0s, hss, touts, tgs, coefs = [], [], [], [], [], []
    print("case batched gradients", flush=True)
    for i in range(6):
        y0 = glibc_rand_ic(2, prob.N, seed=40 + i, A=1.0)
        coef = base * (1 + 0.05 * rng.standard_normal(base.size))
        it = Integrator(hlib, prob, y0, atol=1e-10, rtol=1e-10)
        it.set_coef(coef)
        hs = it.schedule(0.0, 0.01, 0.2)
        t_out = np.array([0.1, 0.2])
        its.append(it)
        y0s.append(y0)
        hss.append(hs)
        touts.append(t_out)
        coefs.append(coef)
        tgs.append(y0[None, :] * 0.9 + 0.01 * rng.standard_normal((2, y0.size)))
    got = loss_grad_batch(its, y0s, [0.0] * 6, hss, touts, tgs, concurrency=3)
    for i, it in enumerate(its):
        L, gc, gy = it.loss_grad(y0s[i], 0.0, hss[i], touts[i], tgs[i])
        assert got[i][0] == L, i
        assert np.array_equal(got[i][1], gc) and np.array_equal(got[i][2], gy), i
    with Integrator(olib, prob, y0s[0], atol=1e-10, rtol=1e-10) as o:
        o.set_coef(coefs[0])
        Lo, gco, gyo = o.loss_grad(y0s[0], 0.0, hss[0], touts[0], tgs[0])
        assert got[0][0] == pytest.approx(Lo, rel=1e-12)
        assert np.abs(got[0][1] - gco).max() <= 1e-10 * np.abs(gco).max()
        assert np.abs(got[0][2] - gyo).max() <= 1e-10 * np.abs(gyo).max()
    for it in its:
        it.close()


def test_stored_stages_equal_recomputed_stages(hlib, monkeypatch):
    """The reverse sweep with the forward pass's stage values kept (the
    default when they fit) and with them recomputed from the checkpoints
    (RDCU_ADJ_STORE=0) gives the same bits."""
    prob = Problem(2, [32, 64], [30.0, 50.0], cgle_terms(2, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=8, A=1.0)
    t_out = np.array([0.1, 0.25])
    print("case stored vs recomputed stages", flush=True)
    g, coef, target, hs = fit_problem(prob, y0, t_out, 0.25, hlib, 1e-10)
    a = g.loss_grad(y0, 0.0, hs, t_out, target)
    monkeypatch.setenv("RDCU_ADJ_STORE", "0")
    b = g.loss_grad(y0, 0.0, hs, t_out, target)
    g.close()
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
