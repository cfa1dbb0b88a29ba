"""Parameter sweeps as one batched run (rdcu_batch_*).

The reference runs a sweep as a SLURM job array, one process per parameter
set (sweep.sh:21-41), each calling makelines.py, which shells out to ./rdcu
(makelines.py:260-268).  Here a sweep is a manifest of instances -- grid,
domain, dictionary, initial-condition seed -- integrated side by side on one
GPU by a single Batch.

The benchmark manifest (``benchmark_manifest``) draws, from numpy's PCG64
generator seeded with ``seed``, for each of ``count`` instances:
  * kind: the first half complex Ginzburg-Landau, the second half
    FitzHugh-Nagumo;
  * grid: uniform over {64^2, 96^2, 128^2, 192^2, 256^2};
  * L: uniform in [50, 200] (both axes);
  * CGLE: b ~ U[0.5, 3], c ~ U[-1.5, 0] in the standard A + (1+ib) lap A -
    (1+ic)|A|^2 A convention, i.e. the reference's README coefficient -c;
  * FHN:  delta ~ U[2, 10], eps ~ U[0.01, 0.1], a1 ~ U[0.5, 2];
  * initial condition: glibc rand() after srand(1000 + i), amplitude 1
    (rdcu -A 1 -s 1000+i).
Draws happen in that order per instance, instance by instance.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from .abi import Batch, Integrator, Problem, glibc_rand_ic
from .dictionaries import cgle_terms, fhn_terms

GRIDS = (64, 96, 128, 192, 256)


@dataclass
class Instance:
    kind: str
    grid: int
    L: float
    params: dict
    seed: int
    prob: Problem

    def y0(self) -> np.ndarray:
        return glibc_rand_ic(self.prob.n, self.prob.N, seed=self.seed, A=1.0)


def benchmark_manifest(count: int = 512, seed: int = 0) -> List[Instance]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = "cgle" if i < count // 2 else "fhn"
        grid = int(GRIDS[rng.integers(len(GRIDS))])
        L = float(rng.uniform(50.0, 200.0))
        if kind == "cgle":
            b = float(rng.uniform(0.5, 3.0))
            c = float(rng.uniform(-1.5, 0.0))
            params = {"b": b, "c": c}
            terms = cgle_terms(2, b, -c)
        else:
            d = float(rng.uniform(2.0, 10.0))
            e = float(rng.uniform(0.01, 0.1))
            a1 = float(rng.uniform(0.5, 2.0))
            params = {"delta": d, "eps": e, "a1": a1}
            terms = fhn_terms(2, d, e, a1)
        prob = Problem(2, [grid, grid], [L, L], terms)
        out.append(Instance(kind, grid, L, params, 1000 + i, prob))
    return out


def run_batched(lib, insts: Sequence[Instance], t1: float, h0: float = 0.01, atol=1e-8,
                rtol=1e-8, max_attempts=1 << 40):
    """All instances in one Batch from t = 0: returns (batch, t, h, attempts).
    The caller closes the batch."""
    b = Batch(lib, [x.prob for x in insts], [x.y0() for x in insts], atol=atol, rtol=rtol)
    n = len(insts)
    t, h, na = b.run(np.zeros(n), np.full(n, h0), t1, max_attempts=max_attempts)
    return b, t, h, na


def run_serial(lib, inst: Instance, t1: float, h0: float = 0.01, atol=1e-8, rtol=1e-8,
               max_attempts=1 << 40):
    """One instance alone (a plain Integrator): returns (t, h, attempts, state)."""
    with Integrator(lib, inst.prob, inst.y0(), atol=atol, rtol=rtol) as it:
        t, h, na = it.advance(0.0, h0, t1, max_attempts=max_attempts)
        return t, h, na, it.state()
