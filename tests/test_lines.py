"""Phase-defect line analysis (rdcu_amd.lines, torch) against the CPU
restatement of makelines.py:19-57, 112-144 (oracle/lines_ref.py), which is
pinned by fields with known vortex lines and by the zero divergence of the
winding flux.  These run torch on the CPU; tests/test_gpu_lines.py runs the
same on the GPU."""
import math
import os
import sys

import numpy as np
import pytest
import torch

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import lines_ref as ref  # noqa: E402

from rdcu_amd import lines as gl  # noqa: E402


def four_vortices(Ns, ax=(0.25, 0.35)):
    """u = sin(2 pi (x+a)/N0) + i sin(2 pi (y+b)/N1): four straight vortex
    lines along the last axis, zeros between grid points, charges +-1."""
    x = np.arange(Ns[0])[:, None, None]
    y = np.arange(Ns[1])[None, :, None]
    z = np.zeros(Ns[2])[None, None, :]
    re = np.sin(2 * np.pi * (x + ax[0]) / Ns[0]) + z + 0 * y
    im = np.sin(2 * np.pi * (y + ax[1]) / Ns[1]) + z + 0 * x
    return np.arctan2(im, re) + np.pi


def smooth_field(Ns, seed, modes=3):
    """A smooth random complex field (a few Fourier modes): many defect lines."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(n) / n for n in Ns], indexing="ij")
    u = np.zeros(Ns, dtype=complex)
    for _ in range(12):
        k = rng.integers(-modes, modes + 1, size=3)
        u += (rng.standard_normal() + 1j * rng.standard_normal()) * np.exp(
            2j * np.pi * sum(ki * g for ki, g in zip(k, grids)))
    return np.stack([u.real, u.imag])


def test_oracle_straight_vortex_lines():
    Ns = (16, 12, 10)
    th = four_vortices(Ns)
    w01 = ref.wloop(th, 0, 1)
    assert np.count_nonzero(w01) == 4 * Ns[2]
    assert w01.sum() == 0 and set(np.unique(w01)) == {-1.0, 0.0, 1.0}
    assert not ref.wloop(th, 1, 2).any() and not ref.wloop(th, 2, 0).any()
    # every vortex is a ring of Ns[2] cells along z closed through the boundary
    lines = ref.findlines(th, Ns, (15.0, 11.0, 9.0))
    assert len(lines) == 4 and all(len(l) == Ns[2] for l in lines)
    for l in lines:
        xy = l[:, :, :2].reshape(-1, 2)
        assert len(np.unique(xy, axis=0)) == 1  # a straight line along z
    # distances: identical frames are at distance 0 on the diagonal
    d = ref.correlate(lines, lines, Ns, (15.0, 11.0, 9.0))
    assert np.all(np.diag(d) == 0) and np.all(d + np.eye(4, dtype=int) > 0)


def test_oracle_winding_flux_is_divergence_free():
    th = np.arctan2(*smooth_field((14, 12, 10), 3)[::-1]) + np.pi
    w01, w12, w20 = ref.wloop(th, 0, 1), ref.wloop(th, 1, 2), ref.wloop(th, 2, 0)
    assert np.count_nonzero(w01) + np.count_nonzero(w12) + np.count_nonzero(w20) > 20
    div = (np.roll(w01, -1, 2) - w01) + (np.roll(w12, -1, 0) - w12) + (np.roll(w20, -1, 1) - w20)
    assert not div.any()


def test_oracle_uniform_field_has_no_lines():
    th = np.full((8, 8, 8), 1.0)
    assert ref.findlines(th, (8, 8, 8), (1.0, 1.0, 1.0)) == []


@pytest.mark.parametrize("seed,Ns", [(1, (16, 16, 16)), (2, (20, 12, 14)), (4, (9, 10, 11))])
def test_torch_matches_oracle(seed, Ns):
    Ls = (30.0, 20.0, 25.0)
    st = smooth_field(Ns, seed)
    th_np = np.arctan2(st[1], st[0]) + np.pi
    assert torch.allclose(gl.phase(torch.from_numpy(st)), torch.from_numpy(th_np), rtol=0,
                          atol=1e-14)
    th = torch.from_numpy(th_np)  # both pipelines from the same phases
    for a1, a2 in ((0, 1), (1, 2), (2, 0)):
        assert np.array_equal(gl.wloop(th, a1, a2).numpy(), ref.wloop(th_np, a1, a2))
    lg, lr = gl.findlines(th, Ns, Ls), ref.findlines(th_np, Ns, Ls)
    assert len(lg) == len(lr) >= 1
    for a, b in zip(lg, lr):
        assert a.shape == b.shape and np.array_equal(a, b)
    st2 = st + 0.05 * smooth_field(Ns, seed + 10)
    th2 = np.arctan2(st2[1], st2[0]) + np.pi
    lr2 = ref.findlines(th2, Ns, Ls)
    lg2 = gl.findlines(torch.from_numpy(th2), Ns, Ls)
    assert len(lg2) == len(lr2) and all(np.array_equal(a, b) for a, b in zip(lg2, lr2))
    prev, cur = [l for l in lr if len(l)], [l for l in lr2 if len(l)]
    dg = gl.correlate(prev, cur, Ns, Ls)
    dr = ref.correlate(prev, cur, Ns, Ls)
    assert dg.shape == (len(cur), len(prev)) and np.array_equal(dg, dr)


def test_torch_matches_oracle_separate_lines():
    Ns, Ls = (16, 12, 10), (15.0, 11.0, 9.0)
    th_np = four_vortices(Ns)
    lg, lr = gl.findlines(torch.from_numpy(th_np), Ns, Ls), ref.findlines(th_np, Ns, Ls)
    assert len(lg) == len(lr) == 4
    assert all(np.array_equal(a, b) for a, b in zip(lg, lr))
    th2 = four_vortices(Ns, (0.75, 0.15))  # the lines moved by about half a cell
    lr2 = ref.findlines(th2, Ns, Ls)
    assert np.array_equal(gl.correlate(lr, lr2, Ns, Ls), ref.correlate(lr, lr2, Ns, Ls))


def test_frame_reader(tmp_path):
    a = np.arange(3 * 2 * 4 * 3 * 2, dtype=np.float64)
    a.tofile(tmp_path / "xstates.dat")
    f = gl.frame(str(tmp_path / "xstates.dat"), (4, 3, 2), 1)
    assert f.shape == (2, 4, 3, 2) and f[0, 0, 0, 0].item() == 48.0
    with pytest.raises(ValueError):
        gl.frame(str(tmp_path / "xstates.dat"), (4, 3, 2), 3)
