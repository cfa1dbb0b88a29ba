"""The torch.autograd wrapper of the discrete adjoint on the GPU: with CUDA
(ROCm) tensors the state, target and state gradient stay on the device
(rdcu_loss_grad_device); backward gives the adjoint gradients, equal bit for
bit to the host-array call, and torch.autograd.gradcheck passes on a 32x32
CGLE problem."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic  # noqa: E402
from rdcu_amd.adjoint import dictionary_loss  # noqa: E402

pytestmark = pytest.mark.gpu


def test_cuda_tensors_gradients_and_gradcheck(hlib):
    assert torch.cuda.is_available()
    dev = torch.device("cuda:0")
    prob = Problem(2, [32, 32], [20.0, 20.0], cgle_terms(2, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=2, A=1.0)
    print("case torch adjoint 32x32 on cuda", flush=True)
    with Integrator(hlib, prob, y0, atol=1e-9, rtol=1e-9) as it:
        hs = it.schedule(0.0, 0.01, 0.1)
        t_out = np.array([0.05, 0.1])
        rng = np.random.default_rng(0)
        target = rng.standard_normal(2 * y0.size) * 0.3
        coef0 = np.array([t.coef for t in prob.terms])
        c = torch.tensor(coef0, device=dev, requires_grad=True)
        y = torch.tensor(y0, device=dev, requires_grad=True)
        tg = torch.tensor(target, device=dev)
        L = dictionary_loss(c, y, it, 0.0, hs, t_out, tg)
        L.backward()
        assert c.grad.is_cuda and y.grad.is_cuda
        it.set_coef(coef0)
        L2, gc, gy = it.loss_grad(y0, 0.0, hs, t_out, target)
        assert float(L.detach()) == L2
        assert np.array_equal(c.grad.cpu().numpy(), gc)
        assert np.array_equal(y.grad.cpu().numpy(), gy)
        ok = torch.autograd.gradcheck(
            lambda cc, yy: dictionary_loss(cc, yy, it, 0.0, hs, t_out, tg),
            (torch.tensor(coef0, device=dev, requires_grad=True),
             torch.tensor(y0, device=dev, requires_grad=True)),
            eps=1e-6, atol=1e-7, rtol=1e-5)
        assert ok
