"""Phase-defect (vortex) lines of a two-field state on the GPU (torch-ROCm).

The analysis step of the reference's makelines.py, run on the device that
holds the frames:
  * wloop      -- winding of the phase around every plaquette (makelines.py:19-24);
  * findlines  -- cells sharing a face with a defect plaquette, joined into
                  lines where cells are at periodic L1 distance 1, each line as
                  its edges in physical coordinates (makelines.py:26-57);
  * correlate  -- minimum periodic L1 distance between the cells of each line
                  of one frame and each line of the next (makelines.py:112-144).

The reference builds a dense pairwise-distance adjacency between defect cells
(chunked on the GPU) and hands it to networkx.  Here the cells are a mask on
the grid: an edge joins two masked cells that are neighbours along an axis
(periodic), and lines are the connected components of that graph, found on
the device by label propagation (each cell takes the smallest label among
itself and its masked neighbours, then jumps to its label's label, until
nothing changes).  Lines come in the order of their smallest flat cell index
(the order networkx yields them in); each line's edges are listed once as
(u, v), u < v in flat cell index, sorted.
"""
from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np
import torch

TWO_PI = 2 * math.pi


def phase(states: torch.Tensor) -> torch.Tensor:
    """Phase in [0, 2*pi) of a frame [2, N0, N1, N2] (u0 = Re, u1 = Im), as
    makelines.py:75 (atan2(u1, u0) + pi)."""
    return torch.atan2(states[1], states[0]) + math.pi


def wloop(th: torch.Tensor, axis1: int = 0, axis2: int = 1) -> torch.Tensor:
    """Integer winding of th around the (axis1, axis2) plaquette at each point."""
    def wrap(d):
        return torch.remainder(d + math.pi, TWO_PI) - math.pi
    r1 = torch.roll(th, -1, dims=axis1)
    r12 = torch.roll(r1, -1, dims=axis2)
    r2 = torch.roll(th, -1, dims=axis2)
    return torch.round((wrap(r1 - th) + wrap(r12 - r1) + wrap(r2 - r12) + wrap(th - r2)) / TWO_PI)


def line_mask(th: torch.Tensor) -> torch.Tensor:
    """Boolean grid of the cells sharing a face with a defect plaquette: each
    nonzero plaquette marks its own cell and the cell behind it along the
    third axis (makelines.py:28-42)."""
    m = torch.zeros(th.shape, dtype=torch.bool, device=th.device)
    for a1, a2, shift in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        w = wloop(th, a1, a2) != 0
        m |= w
        m |= torch.roll(w, -1, dims=shift)
    return m


def _labels(mask: torch.Tensor) -> torch.Tensor:
    """Connected-component labels of the masked cells (periodic axis
    neighbours): each label is the smallest flat index in the component;
    unmasked cells get -1."""
    shape = mask.shape
    n = mask.numel()
    big = torch.iinfo(torch.int64).max
    idx = torch.arange(n, device=mask.device, dtype=torch.int64).reshape(shape)
    lab = torch.where(mask, idx, torch.full_like(idx, big))
    while True:
        new = lab
        for ax in range(3):
            if shape[ax] < 2:
                continue
            for s in (1, -1):
                nb = torch.roll(lab, s, dims=ax)
                new = torch.where(mask, torch.minimum(new, nb), new)
        flat = new.reshape(-1)
        jump = torch.where(mask.reshape(-1), flat[flat.clamp(max=n - 1)], flat)
        new = torch.minimum(flat, jump).reshape(shape)
        if torch.equal(new, lab):
            break
        lab = new
    return torch.where(mask, lab, torch.full_like(lab, -1))


def findlines(th: torch.Tensor, Ns: Sequence[int], Ls: Sequence[float]) -> List[np.ndarray]:
    """Lines of the phase field th [N0, N1, N2]: a list of arrays [nedges, 2, 3]
    of edge end points in physical coordinates (cell index * L/(N-1))."""
    Ns_t = tuple(int(x) for x in Ns)
    mask = line_mask(th)
    lab = _labels(mask)
    n = mask.numel()
    flat_lab = lab.reshape(-1)
    idx = torch.arange(n, device=th.device, dtype=torch.int64).reshape(Ns_t)
    us, vs = [], []
    for ax in range(3):
        if Ns_t[ax] < 2:
            continue
        nb_idx = torch.roll(idx, -1, dims=ax)
        both = mask & torch.roll(mask, -1, dims=ax)
        u = idx[both]
        v = nb_idx[both]
        us.append(torch.minimum(u, v))
        vs.append(torch.maximum(u, v))
    if us:
        u = torch.cat(us)
        v = torch.cat(vs)
        keep = u != v
        key = torch.unique(u[keep] * n + v[keep])  # sorted, each undirected edge once
        u, v = key // n, key % n
    else:
        u = v = torch.zeros(0, dtype=torch.int64, device=th.device)
    comps = torch.unique(flat_lab[flat_lab >= 0])  # sorted: smallest cell first
    el = flat_lab[u]
    scale = torch.tensor([float(L) / (N - 1) for L, N in zip(Ls, Ns_t)], dtype=torch.float64,
                         device=th.device)

    def coords(f):
        c = []
        rem = f
        for N in reversed(Ns_t):
            c.append(rem % N)
            rem = rem // N
        return torch.stack(list(reversed(c)), dim=-1).to(torch.float64) * scale

    order = torch.argsort(el, stable=True)  # edges grouped by line, sorted within
    u, v, el = u[order], v[order], el[order]
    counts = torch.searchsorted(el, comps, right=True) - torch.searchsorted(el, comps)
    pu, pv = coords(u), coords(v)
    pos = torch.stack([pu, pv], dim=1).cpu().numpy()  # [nedges, 2, 3]
    out, start = [], 0
    for c in counts.cpu().tolist():
        out.append(pos[start:start + c])
        start += c
    return out


def correlate(prev: List[np.ndarray], cur: List[np.ndarray], Ns: Sequence[int],
              Ls: Sequence[float], device="cpu", chunk: int = 4096) -> np.ndarray:
    """diffs[j, i]: minimum periodic L1 distance between the cells of line i of
    the previous frame and line j of the current one.  As makelines.py:114-123,
    previous-frame positions are rounded back to cells and current-frame ones
    truncated."""
    Ns_a = np.asarray(Ns)
    scale = np.asarray(Ls, dtype=float) / (Ns_a - 1)

    def cells_of(lines, rnd):
        out = []
        for l in lines:
            if len(l) == 0:  # the reference's concatenate fails here too
                raise ValueError("a line without edges (an isolated cell) has no cells")
            q = l / scale
            arr = np.concatenate(np.round(q).astype(int) if rnd else q.astype(int), axis=0)
            flat = np.unique(np.ravel_multi_index(tuple(arr.T), tuple(Ns_a)))
            out.append(np.array(np.unravel_index(flat, tuple(Ns_a))).T)
        return out

    c1, c2 = cells_of(prev, True), cells_of(cur, False)
    diffs = np.zeros((len(cur), len(prev)), dtype=int)
    if not c1 or not c2:
        return diffs
    e1 = np.concatenate([[0], np.cumsum([len(a) for a in c1])])
    e2 = np.concatenate([[0], np.cumsum([len(b) for b in c2])])
    t1 = torch.tensor(np.vstack(c1), device=device)
    t2 = torch.tensor(np.vstack(c2), device=device)
    half = torch.tensor(Ns_a // 2, device=device)
    Nt = torch.tensor(Ns_a, device=device)
    # per (cell of prev, line of cur): the minimum over that line's cells
    seg = torch.repeat_interleave(torch.arange(len(c2), device=device),
                                  torch.tensor([len(b) for b in c2], device=device))
    best = torch.empty((len(t1), len(c2)), dtype=torch.int64, device=device)
    for m in range(0, len(t1), chunk):
        d = t1[m:m + chunk, None, :] - t2[None, :, :]
        d = torch.remainder(d + half, Nt) - half
        dist = d.abs().sum(-1)  # [chunk, ncells2]
        blk = torch.full((dist.shape[0], len(c2)), torch.iinfo(torch.int64).max,
                         dtype=torch.int64, device=device)
        blk.scatter_reduce_(1, seg.expand(dist.shape[0], -1), dist, reduce="amin")
        best[m:m + chunk] = blk
    best = best.cpu().numpy()
    for i in range(len(c1)):
        diffs[:, i] = best[e1[i]:e1[i + 1]].min(axis=0)
    return diffs


def frame(states_path: str, Ns: Sequence[int], k: int, device="cpu") -> torch.Tensor:
    """Frame k of FILEBASEstates.dat (2 fields, float64) as a tensor
    [2, N0, N1, N2] on `device` (makelines.py:59-68 reads it the same way)."""
    n = 2 * int(np.prod(Ns))
    a = np.fromfile(states_path, dtype=np.float64, count=n, offset=8 * n * k)
    if a.size != n:
        raise ValueError(f"{states_path} has no frame {k}")
    return torch.from_numpy(a.reshape([2] + [int(x) for x in Ns])).to(device)
