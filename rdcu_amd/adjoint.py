"""Gradients of trajectory-matching losses for the Python analysis side.

``DictionaryLoss`` wraps rdcu_loss_grad (the discrete adjoint of DP45 on a
frozen step schedule) as a ``torch.autograd.Function``: the loss is a function
of the dictionary coefficients and the initial state, and ``backward``
returns the adjoint gradients.  The integrator (GPU library, or the oracle in
tests) does the work on host arrays; torch only carries the values.
"""
from __future__ import annotations

import numpy as np

try:
    import torch
except ImportError:  # the C ABI path works without torch
    torch = None


def _np(x):
    return np.ascontiguousarray(x.detach().cpu().numpy(), dtype=np.float64)


if torch is not None:
    class DictionaryLoss(torch.autograd.Function):
        """loss = mean over outputs and grid of (y(t_o) - target_o)^2 for the
        problem of ``integrator`` with coefficients ``coef`` and initial state
        ``y0``, integrated on the frozen schedule ``hs`` from ``t0``."""

        @staticmethod
        def forward(ctx, coef, y0, integrator, t0, hs, t_out, target):
            integrator.set_coef(_np(coef))
            need = coef.requires_grad or y0.requires_grad
            r = integrator.loss_grad(_np(y0), float(t0), hs, t_out, target, grad=need)
            if need:
                L, gc, gy = r
                ctx.save_for_backward(torch.as_tensor(gc, dtype=coef.dtype, device=coef.device),
                                      torch.as_tensor(gy, dtype=y0.dtype, device=y0.device))
            else:
                L = r
            return torch.tensor(L, dtype=torch.float64, device=coef.device)

        @staticmethod
        def backward(ctx, gout):
            gc, gy = ctx.saved_tensors
            return gout * gc, gout * gy.reshape(gy.shape), None, None, None, None, None

    def dictionary_loss(coef, y0, integrator, t0, hs, t_out, target):
        return DictionaryLoss.apply(coef, y0, integrator, t0, hs, t_out, target)
