"""torch.autograd wrapper of rdcu_loss_grad (on the oracle library)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic
from rdcu_amd.adjoint import dictionary_loss


def test_backward_gives_adjoint_gradients_and_passes_gradcheck(olib):
    prob = Problem(2, [16], [8.0], cgle_terms(1, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=2, A=1.0)
    it = Integrator(olib, prob, y0, atol=1e-9, rtol=1e-9)
    hs = it.schedule(0.0, 0.01, 0.2)
    t_out = np.array([0.1, 0.2])
    rng = np.random.default_rng(0)
    target = rng.standard_normal(2 * y0.size) * 0.3
    coef0 = np.array([t.coef for t in prob.terms])
    c = torch.tensor(coef0, requires_grad=True)
    y = torch.tensor(y0, requires_grad=True)
    L = dictionary_loss(c, y, it, 0.0, hs, t_out, target)
    L.backward()
    it.set_coef(coef0)
    L2, gc, gy = it.loss_grad(y0, 0.0, hs, t_out, target)
    assert float(L) == L2
    assert np.array_equal(c.grad.numpy(), gc) and np.array_equal(y.grad.numpy(), gy)
    assert torch.autograd.gradcheck(
        lambda cc, yy: dictionary_loss(cc, yy, it, 0.0, hs, t_out, target),
        (torch.tensor(coef0, requires_grad=True), torch.tensor(y0, requires_grad=True)),
        eps=1e-6, atol=1e-7, rtol=1e-5)
    it.close()
