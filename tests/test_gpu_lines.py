"""Defect-line analysis on the GPU (rdcu_amd.lines on cuda:0) against the CPU
restatement (oracle/lines_ref.py): a smooth random field with several lines,
and the phase of a 3-D CGLE state integrated on the GPU."""
import os
import sys

import numpy as np
import pytest
import torch

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import lines_ref as ref  # noqa: E402

from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic  # noqa: E402
from rdcu_amd import lines as gl  # noqa: E402
from test_lines import smooth_field  # noqa: E402

pytestmark = pytest.mark.gpu


def _check(th_np, Ns, Ls):
    th = torch.from_numpy(th_np).to("cuda:0")
    for a1, a2 in ((0, 1), (1, 2), (2, 0)):
        assert np.array_equal(gl.wloop(th, a1, a2).cpu().numpy(), ref.wloop(th_np, a1, a2))
    lg, lr = gl.findlines(th, Ns, Ls), ref.findlines(th_np, Ns, Ls)
    assert len(lg) == len(lr) >= 1
    assert all(np.array_equal(a, b) for a, b in zip(lg, lr))
    return lr


def test_gpu_lines_smooth_field():
    Ns, Ls = (20, 12, 14), (30.0, 20.0, 25.0)
    print("case lines smooth field", flush=True)
    st = smooth_field(Ns, 2)
    lr = _check(np.arctan2(st[1], st[0]) + np.pi, Ns, Ls)
    st2 = st + 0.05 * smooth_field(Ns, 12)
    lr2 = _check(np.arctan2(st2[1], st2[0]) + np.pi, Ns, Ls)
    prev, cur = [l for l in lr if len(l)], [l for l in lr2 if len(l)]
    assert np.array_equal(gl.correlate(prev, cur, Ns, Ls, device="cuda:0"),
                          ref.correlate(prev, cur, Ns, Ls))


def test_gpu_lines_of_cgle_state(hlib):
    Ns, Ls = (32, 32, 32), (40.0, 40.0, 40.0)
    prob = Problem(2, list(Ns), list(Ls), cgle_terms(3, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=3, A=1.0)
    print("case lines of a 3-D CGLE state", flush=True)
    with Integrator(hlib, prob, y0) as it:
        it.advance(0.0, 0.01, t1=4.0, max_attempts=5000)
        y = it.state()
    st = torch.from_numpy(y.reshape((2,) + Ns)).to("cuda:0")
    th_np = gl.phase(st).cpu().numpy()
    lines = _check(th_np, Ns, Ls)
    assert sum(len(l) for l in lines) > 0
