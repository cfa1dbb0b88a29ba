"""Gradients of trajectory-matching losses for the Python analysis side.

``DictionaryLoss`` wraps rdcu_loss_grad (the discrete adjoint of DP45 on a
frozen step schedule) as a ``torch.autograd.Function``: the loss is a function
of the dictionary coefficients and the initial state, and ``backward``
returns the adjoint gradients.

On the HIP library with CUDA (ROCm) tensors the initial state, the target and
the state gradient stay on the device: their device pointers go to
rdcu_loss_grad_device, and only the coefficients (a few numbers) and the loss
cross to the host.  Host tensors, or the CPU oracle in tests, go through
rdcu_loss_grad on host arrays; both give the same values.
"""
from __future__ import annotations

import ctypes

import numpy as np

try:
    import torch
except ImportError:  # the C ABI path works without torch
    torch = None


def _np(x):
    return np.ascontiguousarray(x.detach().cpu().numpy(), dtype=np.float64)


def _device_path(integrator, y0):
    """True when y0 lives on a GPU and the integrator is the HIP library."""
    if not (torch is not None and isinstance(y0, torch.Tensor) and y0.is_cuda):
        return False
    return integrator.lib.rdcu_backend() == b"hip-gfx950"


def _loss_grad_device(integrator, y0, t0, hs, t_out, target, need):
    """rdcu_loss_grad_device on tensors of y0's device: returns (loss, grad
    of the coefficients as numpy, grad of y0 as a tensor) or the loss."""
    from .abi import _ptr
    y0d = y0.detach().to(torch.float64).contiguous().reshape(-1)
    tg = torch.as_tensor(target, dtype=torch.float64, device=y0d.device).contiguous().reshape(-1)
    hs = np.ascontiguousarray(hs, dtype=np.float64).reshape(-1)
    to = np.ascontiguousarray(t_out, dtype=np.float64).reshape(-1)
    if y0d.numel() != integrator.nN or tg.numel() != to.size * integrator.nN:
        raise ValueError("y0 needs n*N values and target nout*n*N")
    nt = len(integrator.prob.terms)
    gc = np.empty(max(1, nt)) if need else None
    gy = torch.empty_like(y0d) if need else None
    # the library's stream does not order against torch's: finish torch's
    # writes of y0 and target first (the call itself returns complete)
    torch.cuda.current_stream(y0d.device).synchronize()
    loss = ctypes.c_double()
    integrator._check(integrator.lib.rdcu_loss_grad_device(
        integrator.h, ctypes.c_void_p(y0d.data_ptr()), float(t0), hs.size, _ptr(hs), to.size,
        _ptr(to), ctypes.c_void_p(tg.data_ptr()), ctypes.byref(loss),
        _ptr(gc) if need else None, ctypes.c_void_p(gy.data_ptr()) if need else None))
    if not need:
        return loss.value
    return loss.value, gc[:nt], gy


if torch is not None:
    class DictionaryLoss(torch.autograd.Function):
        """loss = mean over outputs and grid of (y(t_o) - target_o)^2 for the
        problem of ``integrator`` with coefficients ``coef`` and initial state
        ``y0``, integrated on the frozen schedule ``hs`` from ``t0``."""

        @staticmethod
        def forward(ctx, coef, y0, integrator, t0, hs, t_out, target):
            integrator.set_coef(_np(coef))
            need = coef.requires_grad or y0.requires_grad
            if _device_path(integrator, y0):
                r = _loss_grad_device(integrator, y0, t0, hs, t_out, target, need)
            else:
                r = integrator.loss_grad(_np(y0), float(t0), hs, t_out, target, grad=need)
            if need:
                L, gc, gy = r
                ctx.save_for_backward(torch.as_tensor(gc, dtype=coef.dtype, device=coef.device),
                                      torch.as_tensor(gy, dtype=y0.dtype,
                                                      device=y0.device).reshape(y0.shape))
            else:
                L = r
            return torch.tensor(L, dtype=torch.float64, device=coef.device)

        @staticmethod
        def backward(ctx, gout):
            gc, gy = ctx.saved_tensors
            return gout * gc, gout * gy, None, None, None, None, None

    def dictionary_loss(coef, y0, integrator, t0, hs, t_out, target):
        return DictionaryLoss.apply(coef, y0, integrator, t0, hs, t_out, target)
