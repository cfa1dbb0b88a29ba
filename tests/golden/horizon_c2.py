"""The 2-D CGLE 512x512 workload run to its full horizon, and the fixture
that pins it (tests/golden/horizon_c2.npz).

Workload: 2 FP64 fields, 18-term CGLE dictionary (b = 2, c = 0.8), L = 200^2,
initial condition glibc rand() after srand(1), amplitude 1 (rdcu -A 1 -s 1),
rtol = atol = 1e-8, h0 = 0.01, outputs every 1 to t = 100 -- the rdcu command
line `-n 2 -N 512,512 -L 200,200 -t 100 -d 1 -A 1 -s 1 -D2`.

`run_horizon(lib)` integrates it through the C ABI exactly as the CLI does
(rdcu_run with a per-attempt callback; each output time is evaluated in the
accepted step that reaches it) and returns every attempt's (t, h, accepted)
and, per output time (t = 0 first), the SHA-256 of the state frame and of the
RHS frame -- the bytes the CLI writes to FILEBASEstates.dat and FILEBASEf.dat.

Running this file with the oracle built (oracle/_bin/liboracle.so) writes the
fixture: `python tests/golden/horizon_c2.py [t1]`.  It takes a few minutes on
8 host cores.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

FIXTURE = os.path.join(HERE, "horizon_c2.npz")


def problem():
    from rdcu_amd import Problem, cgle_terms, glibc_rand_ic
    prob = Problem(2, [512, 512], [200.0, 200.0], cgle_terms(2, 2.0, 0.8))
    return prob, glibc_rand_ic(2, prob.N, seed=1, A=1.0)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def run_horizon(lib, t1=100.0, dt=1.0, progress=None):
    from rdcu_amd import Integrator
    prob, y0 = problem()
    outs = [dt * k for k in range(1, int(round(t1 / dt)) + 1)]
    log, frames, nxt = [], [], [0]
    with Integrator(lib, prob, y0) as it:
        frames.append((digest(y0), digest(it.rhs(y0))))

        def cb(t, h, ok, cur):
            log.append((t, h, ok))
            if progress and len(log) % progress == 0:
                print(f"{len(log)} attempts, t = {t:.4f}", flush=True)
            while ok and nxt[0] < len(outs) and t >= outs[nxt[0]]:
                y, f, _ = it.frame(outs[nxt[0]], dense=2)
                frames.append((digest(y), digest(f)))
                nxt[0] += 1
            return outs[nxt[0]] if nxt[0] < len(outs) else np.inf

        it.run(0.0, 0.01, t1, stop=outs[0], callback=cb)
        final = digest(it.state())
    t = np.array([x[0] for x in log])
    h = np.array([x[1] for x in log])
    acc = np.array([x[2] for x in log], dtype=np.uint8)
    return t, h, acc, frames, final


def main():
    from rdcu_amd.abi import load
    t1 = float(sys.argv[1]) if len(sys.argv) > 1 else 100.0
    lib = load(os.path.join(ROOT, "oracle", "_bin", "liboracle.so"))
    t, h, acc, frames, final = run_horizon(lib, t1, progress=500)
    np.savez_compressed(FIXTURE, t=t, h=h, acc=acc, t1=np.array(t1),
                        y_sha=np.array([f[0] for f in frames]),
                        f_sha=np.array([f[1] for f in frames]), final_sha=np.array(final))
    print(f"{t.size} attempts, {int(acc.sum())} accepted, {len(frames)} frames -> {FIXTURE}")


if __name__ == "__main__":
    main()
