#!/usr/bin/env python3
"""Benchmark: DP45 steps/s on the 2-D complex Ginzburg-Landau problem, 512x512,
FP64, on one MI355X, with the RHS evaluation measured against the HBM roofline.

Workload (synthetic, as the reference's own runs): two real fields u0 = Re A,
u1 = Im A of A_t = A + (1+2i) lap A - (1-0.8i)|A|^2 A (the 18-term dictionary
of the reference's README example), periodic 512x512 grid, L = 200x200,
random initial condition of amplitude 1 from glibc rand() with seed 1
(rdcu -A 1 -s 1), rtol = atol = 1e-8.  A "step" is one DP45 attempt (six RHS
evaluations plus the error norm and step control, all on the device); the
state and every intermediate stay in HBM for the whole timed region.

Prints one JSON line.  Diagnostics go to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(n, M, Ns, Mz=None):
    """Bytes one RHS evaluation must move through HBM, per kernel, for
    half-spectrum storage (8 B per grid point per field or factor) with one
    read and one write per array per pass:
      x forward   read u (8nN), write n half spectra (8nN)
      y forward   (3-D) read and write the n spectra (16nN), and write the
                  M - Mz factors without k_z, made in this pass (8(M-Mz)N)
      last axis   read n spectra (8nN), write the factor spectra that go
                  through it (2-D: all M; 3-D: the Mz with k_z) (8MN, 8MzN)
      y inverse   (3-D) read and write the Mz factor spectra (16MzN)
      x inverse   read M factor spectra (8MN) and u (8nN), write f (8nN)
    Mz (3-D) defaults to n: the Laplacian's d2/dz2 of each field.  The sum is
    8N[n(2 ndim + 1) + 2M] in 2-D and 8N[7n + 2M + 2Mz] in 3-D."""
    N = 1
    for s in Ns:
        N *= s
    if Mz is None:
        Mz = n
    fwd, inv = 16 * n * N, 8 * M * N + 16 * n * N
    out = {"row_r2c": fwd, "row_fwd": fwd, "row_c2r": inv, "row_inv": inv}
    if len(Ns) == 3:
        out.update({"col_fwd": 16 * n * N + 8 * (M - Mz) * N, "col_mult": 8 * n * N + 8 * Mz * N,
                    "col_inv": 16 * Mz * N})
    else:
        out["col_mult"] = 8 * n * N + 8 * M * N
    return out


def attempt_bytes(n, M, Ns, Mz=None):
    """Algorithmic HBM bytes of one DP45 attempt through the fused kernels,
    one read and one write per array per kernel (A = 8nN: a state-size array
    or the n half spectra; F = 8MN: the M factor spectra):
      x forward + stage-2 sum        read y, k1; write ytemp2, S         4A
      per RHS (stages 2..7)          strided: read S, write F           A + F
                                     (3-D: the y forward pass 2A and the
                                     factors without k_z; z: A and the Mz
                                     factors with k_z; y inverse 2 of those)
      x inverse of stage s = 2..6    read F, ytemp_s, y and the k_j (j < s) of
                                     stage s+1's sum with nonzero a_{s+1,j}
                                     (k_s stays on chip); write k_s,
                                     ytemp_{s+1}, S
      x inverse of stage 7           read F, ytemp7, y and the k_j (j < 7) with
                                     a nonzero error weight; write k7
    The zero pattern is DP45's: a_72 = b_2 = 0 and e_2 = 0 (dp45.cu:11-13);
    every other a_ij (j < i) and e_j is nonzero.  The step control's error
    partials (one double per line) are left out."""
    N = 1
    for s in Ns:
        N *= s
    A, F = 8 * n * N, 8 * M * N
    if Mz is None:
        Mz = n
    Fz = 8 * Mz * N  # 3-D: the factors with k_z
    nz_sum = {2: 1, 3: 2, 4: 3, 5: 4, 6: 4}  # k_j read for stage s+1's sum, j < s
    nz_err = 5                                # e_1, e_3..e_6
    rhs = (3 * A + (F - Fz) + 3 * Fz) if len(Ns) == 3 else A + F
    b = 4 * A + 6 * rhs
    for s in range(2, 7):
        b += F + (2 + nz_sum[s]) * A + 3 * A
    b += F + (2 + nz_err) * A + A
    return b


PROFILE = os.path.join(ROOT, "profiles", "rhs_2d_cgle_512.md")
KERNEL_SYMBOL = {"row_fwd": "k_rowf_reg", "col_mult": "k_col_reg", "row_inv": "rdcu_rowi_jit",
                 "row_r2c": "k_row_r2c", "col": "k_col", "row_c2r": "k_row_c2r"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by the copy-kernel calibration),
    or None when the summary has no row for it."""
    sym = KERNEL_SYMBOL.get(kernel)
    if not sym or not os.path.exists(PROFILE):
        return None
    in_pmc = False  # rows of the counter table only (after its header)
    for line in open(PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[-1] == "scaled total MB":
            in_pmc = True
            continue
        if in_pmc and len(cells) == 6 and sym in cells[0]:
            try:
                return float(cells[5]) * 1e6
            except ValueError:
                return None
    return None


ATTEMPT_PROFILE = os.path.join(ROOT, "profiles", "attempt_2d_cgle_512.md")


def measured_attempt_traffic():
    """HBM bytes of one DP45 attempt at 512^2 from the committed PMC table of
    the attempt profile (scaled MB per kernel position; the strided row
    stands for its six launches), or None without it."""
    if not os.path.exists(ATTEMPT_PROFILE):
        return None
    total, seen = 0.0, False
    for line in open(ATTEMPT_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[4] == "scaled MB":
            seen = True
            continue
        if seen and len(cells) == 7 and cells[0][:1].isdigit():
            try:
                mb = float(cells[4])
            except ValueError:
                continue
            total += 6 * mb if "," in cells[0] else mb
    return total * 1e6 if seen else None


def copy_calibration(mib, reps):
    """A plain device-to-device copy kernel (torch's elementwise copy) of a known
    size: the FETCH_SIZE / WRITE_SIZE counters it reads, against its exact
    byte count, give the scale factor applied to the RHS kernels' counters."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.arange(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"mode": "copy", "bytes_read": n * 8, "bytes_written": n * 8,
                      "ms": ms, "GBps": 2 * n * 8 / (ms * 1e-3) / 1e9}), flush=True)


def batched_rhs(b, reps=20):
    """One plain RHS of all instances of a batch (rdcu_batch_time: HIP events,
    every instance in the launches): the time of the whole batched RHS, its
    algorithmic bytes against the HBM peak, and per shape class the time, bytes
    and bandwidth of each launch."""
    ms, classes = b.time_rhs(reps)
    total = sum(sum(c["bytes"].values()) for c in classes)
    per = []
    for c in classes:
        per.append({"grid": c["grid"], "instances": c["instances"],
                    "ms": c["ms"], "MB": {k: v / 1e6 for k, v in c["bytes"].items()},
                    "GBps": {k: (c["bytes"][k] / (c["ms"][k] * 1e-3) / 1e9 if c["ms"][k] > 0
                                 else None) for k in c["ms"]}})
    out = {"ms": ms, "bytes": total, "achieved_GBps": total / (ms * 1e-3) / 1e9,
           "frac": total / (ms * 1e-3) / PEAK_HBM, "reps": reps, "classes": per,
           "what": "one plain RHS of every instance (all classes back to back), HIP events "
                   "on the batch's stream; bytes: algorithmic (half spectra, one read and "
                   "one write per array per pass)"}
    log("batched rhs:", json.dumps(out))
    return out


def sweep_leg(lib, count, t_end, every):
    """Batched sweep against one-at-a-time plain solves on the same GPU.

    The manifest (rdcu_amd.sweep.benchmark_manifest, PCG64 seed 0): half CGLE,
    half FitzHugh-Nagumo, grids from {64..256}^2, L in [50, 200], IC seeds
    1000+i.  Every instance runs from t=0 to t_end at rtol=atol=1e-8.
    Batched: all instances in one Batch, one rdcu_batch_run.  Serial: every
    `every`-th instance alone in a plain Integrator (rdcu_advance), the
    subset's solves/s standing for the whole sweep's; each of those must
    reproduce its batched t, h, attempt count and final state bit for bit.
    Setup (context creation, kernel compilation, tuning) is timed apart."""
    import numpy as np
    from rdcu_amd import Integrator
    from rdcu_amd.sweep import benchmark_manifest, Batch
    insts = benchmark_manifest(count, seed=0)
    y0s = [x.y0() for x in insts]
    c0 = time.perf_counter()
    b = Batch(lib, [x.prob for x in insts], y0s)
    c1 = time.perf_counter()
    n = len(insts)
    t, h, na = b.run(np.zeros(n), np.full(n, 0.01), t_end)  # returns after a stream sync
    c2 = time.perf_counter()
    sub = list(range(0, n, every))
    yb = {i: b.member(i).state() for i in sub}
    rhs = batched_rhs(b)  # after the run: not part of the timed window
    b.close()
    log(f"batched: {n} instances, setup {c1 - c0:.2f} s, run {c2 - c1:.3f} s, "
        f"attempts {int(na.sum())} (max {int(na.max())})")
    for gsz in sorted({x.grid for x in insts}):
        for kind in ("cgle", "fhn"):
            a = [int(na[i]) for i, x in enumerate(insts) if x.grid == gsz and x.kind == kind]
            if a:
                log(f"  {kind} {gsz}^2: {len(a)} instances, attempts mean {np.mean(a):.0f} "
                    f"max {max(a)}")
    # warm the plain path's per-process caches (kernel compilation per
    # dictionary structure and transform length) before timing
    warm = {}
    for i in sub:
        key = (insts[i].kind, insts[i].grid % 3 == 0)
        warm.setdefault(key, i)
    for i in warm.values():
        Integrator(lib, insts[i].prob, y0s[i]).close()
    setup_s = run_s = 0.0
    same = True
    for i in sub:
        s0 = time.perf_counter()
        it = Integrator(lib, insts[i].prob, y0s[i])
        s1 = time.perf_counter()
        ts, hs, ns = it.advance(0.0, 0.01, t_end, max_attempts=1 << 40)
        s2 = time.perf_counter()
        ys = it.state()
        it.close()
        setup_s += s1 - s0
        run_s += s2 - s1
        same = same and (ts, hs, ns) == (t[i], h[i], na[i]) and np.array_equal(ys, yb[i])
    batched = n / (c2 - c1)
    serial = len(sub) / run_s
    out = {
        "instances": n, "t_end": t_end, "manifest": "rdcu_amd.sweep.benchmark_manifest(seed=0)",
        "batched_solves_per_s": batched, "serial_solves_per_s": serial,
        "speedup": batched / serial,
        "batched_solves_per_s_with_setup": n / (c2 - c0),
        "serial_solves_per_s_with_setup": len(sub) / (setup_s + run_s),
        "batched_attempts": int(na.sum()), "serial_subset": len(sub),
        "identical_step_sequences": bool(same),
        "rhs": rhs,
    }
    log("batched:", json.dumps(out))
    return out


def adjoint_leg(lib, nbatch=0, concurrency=4):
    """Gradient of an output-matching loss by the discrete adjoint.

    2-D CGLE 256x256, L=100, IC amplitude 1 seed 1, rtol=atol=1e-10, outputs
    at t = 0.5, 1, ..., 10; the target comes from the coefficients with 5%
    relative Gaussian noise (seed 0), the gradient is taken at the nominal
    coefficients on the step schedule of their adaptive run.  Reports the
    time of one loss-only evaluation and of loss + gradient (all 18
    coefficients and the initial state), and central differences (h=1e-6) on
    three coefficients against the adjoint.

    With nbatch > 0, also the gradients of nbatch instances at once
    (rdcu_loss_grad_batch, `concurrency` instances on the GPU together): a
    population of coefficient points (nominal with 1% relative Gaussian
    noise, seed 1) fitted to the same target on the nominal schedule, timed
    against nbatch times one instance's loss + gradient; the first instance's
    batched result is checked bit for bit against its own single call."""
    import numpy as np
    from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic
    prob = Problem(2, [256, 256], [100.0, 100.0], cgle_terms(2, 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=1, A=1.0)
    t_out = np.arange(1, 21) * 0.5
    coef = np.array([t.coef for t in prob.terms])
    rng = np.random.default_rng(0)
    pert = coef * (1 + 0.05 * rng.standard_normal(coef.size))
    with Integrator(lib, prob, y0, atol=1e-10, rtol=1e-10) as it:
        it.set_coef(pert)
        outs, nxt = [], [0]

        def cb(t, h, ok, cur):
            while ok and nxt[0] < len(t_out) and t >= t_out[nxt[0]]:
                outs.append(it.eval(t_out[nxt[0]]))
                nxt[0] += 1
            return t_out[nxt[0]] if nxt[0] < len(t_out) else np.inf

        it.run(0.0, 0.01, 10.0, stop=t_out[0], callback=cb)
        target = np.concatenate(outs)
        it.set_coef(coef)
        it.set_state(y0)
        a0 = time.perf_counter()
        hs = it.schedule(0.0, 0.01, 10.0)
        a1 = time.perf_counter()
        it.loss_grad(y0, 0.0, hs, t_out, target)  # warm (companion problem setup)
        a2 = time.perf_counter()
        L = it.loss_grad(y0, 0.0, hs, t_out, target, grad=False)
        a3 = time.perf_counter()
        L2, gc, gy = it.loss_grad(y0, 0.0, hs, t_out, target)
        a4 = time.perf_counter()
        errs = []
        for t in (0, 3, 13):
            cp, cm = coef.copy(), coef.copy()
            cp[t] += 1e-6
            cm[t] -= 1e-6
            it.set_coef(cp)
            Lp = it.loss_grad(y0, 0.0, hs, t_out, target, grad=False)
            it.set_coef(cm)
            Lm = it.loss_grad(y0, 0.0, hs, t_out, target, grad=False)
            errs.append(abs((Lp - Lm) / 2e-6 - gc[t]) / max(abs(gc[t]), 1e-300))
    out = {"workload": "2-D CGLE 256x256, L=100, t=10, 20 outputs, rtol=atol=1e-10, 5% "
                       "coefficient noise target", "steps": int(hs.size),
           "schedule_s": a1 - a0, "loss_s": a3 - a2, "loss_grad_s": a4 - a3,
           "grad_over_loss": (a4 - a3) / (a3 - a2), "loss": L2,
           "fd_max_rel_err_3coef": max(errs), "grad_norm_coef": float(np.linalg.norm(gc))}
    if nbatch > 0:
        from rdcu_amd import loss_grad_batch
        rng1 = np.random.default_rng(1)
        b0 = time.perf_counter()
        its = []
        for i in range(nbatch):
            it_i = Integrator(lib, prob, y0, atol=1e-10, rtol=1e-10)
            it_i.set_coef(coef * (1 + 0.01 * rng1.standard_normal(coef.size)))
            its.append(it_i)
        # companion problems are built by the first gradient call
        loss_grad_batch(its, [y0] * nbatch, [0.0] * nbatch, [hs[:1]] * nbatch,
                        [hs[:1]] * nbatch, [y0] * nbatch, concurrency=concurrency)
        b1 = time.perf_counter()
        got = loss_grad_batch(its, [y0] * nbatch, [0.0] * nbatch, [hs] * nbatch,
                              [t_out] * nbatch, [target] * nbatch, concurrency=concurrency)
        b2 = time.perf_counter()
        L1, gc1, gy1 = its[0].loss_grad(y0, 0.0, hs, t_out, target)
        same = bool(got[0][0] == L1 and np.array_equal(got[0][1], gc1)
                    and np.array_equal(got[0][2], gy1))
        for it_i in its:
            it_i.close()
        out["batch"] = {"instances": nbatch, "concurrency": concurrency, "setup_s": b1 - b0,
                        "batch_s": b2 - b1, "per_instance_s": (b2 - b1) / nbatch,
                        "speedup_vs_single": nbatch * (a4 - a3) / (b2 - b1),
                        "first_identical_to_single": same}
    log("adjoint:", json.dumps(out))
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(prob, y0, gname, warmup, attempts, gpu_steps_s, gpu_attempts_s, budget):
    """The FP64 C++ oracle on the host cores over the GPU's own timed window:
    the same run from the same initial condition, `warmup` attempts untimed,
    then the same `attempts` attempts timed, first with every host core and
    then with one thread, so both sides cover identical attempts (the same
    accepted/rejected mix).  A longer all-core sample of the run's first
    attempts (about `budget` seconds) is reported beside it."""
    import ctypes
    import numpy as np  # noqa: F401
    from rdcu_amd import Integrator, load
    ora = load(os.path.join(ROOT, "oracle", "_bin", "liboracle.so"))
    ora.oracle_get_threads.restype = ctypes.c_int
    cores = ora.oracle_get_threads()

    def window(threads, limit):
        """Attempts warmup+1.. of the run, timed; stops after `limit` seconds."""
        ora.oracle_set_threads(threads)
        with Integrator(ora, prob, y0, atol=1e-8, rtol=1e-8) as o:
            tc, hc = 0.0, 0.01
            for _ in range(warmup):
                tc, hc, _ok = o.step(tc, hc)
            na = nacc = 0
            c0 = time.perf_counter()
            while na < attempts and (na == 0 or time.perf_counter() - c0 < limit):
                tc, hc, ok = o.step(tc, hc)
                na += 1
                nacc += ok
            return na, nacc, time.perf_counter() - c0

    _, nacc, cel = window(cores, float("inf"))
    na1, nacc1, cel1 = window(1, budget)
    ora.oracle_set_threads(cores)
    with Integrator(ora, prob, y0, atol=1e-8, rtol=1e-8) as o:
        tc, hc, na, nl = 0.0, 0.01, 0, 0
        c0 = time.perf_counter()
        while time.perf_counter() - c0 < budget or na < 2:
            tc, hc, ok = o.step(tc, hc)
            na += 1
            nl += ok
        lel = time.perf_counter() - c0
    out = {
        "value": nacc / cel, "unit": "accepted DP45 steps/s", "cores": cores, "kind": "port",
        "sample": f"the bench's own window of the same {gname} CGLE run (attempts "
                  f"{warmup + 1}..{warmup + attempts}: {attempts} attempts, {nacc} accepted), "
                  f"FP64 C++ oracle, OpenMP {cores} threads, {cel:.2f} s",
        "attempts_per_s": attempts / cel,
        "cpu_model": cpu_model(),
        "one_thread": {"value": nacc1 / cel1, "attempts_per_s": na1 / cel1, "attempts": na1,
                       "accepted": nacc1, "seconds": cel1,
                       "what": f"attempts {warmup + 1}..{warmup + na1} of the run, 1 thread"},
        "long_sample": {"value": nl / lel, "attempts_per_s": na / lel, "attempts": na,
                        "accepted": nl, "seconds": lel,
                        "what": f"first {na} attempts of the run, {cores} threads"},
        "gpu_over_cpu": {"steps_per_s": gpu_steps_s / (nacc / cel),
                         "attempts_per_s": gpu_attempts_s / (attempts / cel)},
    }
    log("cpu:", json.dumps(out))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed DP45 attempts (default 600 in 2-D, 20 in 3-D)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed attempts first (default 60 in 2-D, 3 in 3-D)")
    ap.add_argument("--grid", type=int, default=None, help="grid points per axis")
    ap.add_argument("--workload", choices=("cgle2d", "cgle3d"), default="cgle2d",
                    help="cgle2d: 2-D CGLE 512x512, L=200, amplitude 1 (the headline); "
                         "cgle3d: 3-D CGLE 256^3, L=100, amplitude 0.1 (same coefficients)")
    ap.add_argument("--reps", type=int, default=None,
                    help="RHS evaluations per kernel timing (default 200 in 2-D, 10 in 3-D)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="wall-clock budget of the CPU-oracle sample (0 to skip)")
    ap.add_argument("--mode", choices=("bench", "rhs", "copy"), default="bench",
                    help="bench: the JSON line; rhs: only the per-kernel RHS timing (a short "
                         "target for profilers); copy: a plain device copy of --copy-mib MiB, "
                         "the calibration for the HBM byte counters")
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--sweep", type=int, default=512,
                    help="instances of the batched sweep (0 to skip it)")
    ap.add_argument("--sweep-t", type=float, default=50.0,
                    help="end time of every sweep instance (the sweep's stated horizon)")
    ap.add_argument("--serial-every", type=int, default=16,
                    help="serial baseline on every k-th sweep instance")
    ap.add_argument("--adjoint", type=int, default=1,
                    help="1: time the discrete adjoint on 2-D CGLE 256x256, t=10 (0 to skip)")
    ap.add_argument("--adjoint-batch", type=int, default=64,
                    help="instances of the batched adjoint (0 to skip)")
    ap.add_argument("--adjoint-concurrency", type=int, default=4,
                    help="instances of the batched adjoint on the GPU together")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path here (each integrator "
              "instance lives on one GPU). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.mode == "copy":
        return copy_calibration(args.copy_mib, args.reps)

    import numpy as np
    from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic, hip_lib, load

    nd = 2 if args.workload == "cgle2d" else 3
    G = args.grid or (512 if nd == 2 else 256)
    Ns, Ls = [G] * nd, [200.0 if nd == 2 else 100.0] * nd
    amp = 1.0 if nd == 2 else 0.1
    if args.steps is None:
        args.steps = 600 if nd == 2 else 20
    if args.warmup is None:
        args.warmup = 60 if nd == 2 else 3
    if args.reps is None:
        args.reps = 200 if nd == 2 else 10

    prob = Problem(2, Ns, Ls, cgle_terms(nd, 2.0, 0.8))
    gname = "x".join(str(x) for x in Ns)
    workload = (f"{nd}-D CGLE {gname}, L={Ls[0]:g}^{nd}, b=2, c=0.8, 2 FP64 fields, "
                f"{len(prob.terms)}-term dictionary, IC amplitude {amp:g} seed 1, "
                "rtol=atol=1e-8, DP45 pseudospectral")
    y0 = glibc_rand_ic(2, G ** nd, seed=1, A=amp)
    lib = hip_lib()
    it = Integrator(lib, prob, y0, atol=1e-8, rtol=1e-8)

    if args.mode == "rhs":
        km = it.time_rhs(args.reps)
        it.close()
        print(json.dumps({"mode": "rhs", "grid": Ns, "reps": args.reps, "kernels_ms": km}),
              flush=True)
        return

    t, h = 0.0, 0.01  # -d 1: first step dt/100
    t, h, done = it.advance(t, h, max_attempts=args.warmup)
    s0 = it.stats()
    w0 = time.perf_counter()
    t, h, done = it.advance(t, h, max_attempts=args.steps)  # returns after a stream sync
    w1 = time.perf_counter()
    s1 = it.stats()
    el = w1 - w0
    attempts = s1["attempts"] - s0["attempts"]
    accepted = s1["accepted"] - s0["accepted"]
    log(f"timed: {attempts} attempts ({accepted} accepted) in {el:.4f} s, t={t:.6f} h={h:.6f}")

    # RHS kernels, HIP events on the library's stream
    km = it.time_rhs(args.reps)
    log("kernel ms:", json.dumps(km))
    M = 2 * nd  # referenced derivative factors: d2/dx_a^2 of both fields
    kb = {k: v for k, v in algorithmic_bytes(2, M, Ns).items() if k in km}
    B_rhs = sum(kb.values())
    dom = max((k for k in km if k != "rhs"), key=lambda k: km[k])
    ach = kb[dom] / (km[dom] * 1e-3)
    rhs_ach = B_rhs / (km["rhs"] * 1e-3)

    result = {
        "metric": "DP45 steps/s, 2D CGLE 512² FP64 (%HBM roofline); batched vs serial solves/s",
        "value": accepted / el,
        "unit": "accepted DP45 steps/s",
        "n_gpus": 1,
        "steps": attempts,
        "warmup": args.warmup,
        "ms_per_step": el / attempts * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (glibc rand() initial condition, seed 1, amplitude {amp:g})",
        "config": {"workload": workload,
                   "grid": Ns, "attempts_per_s": attempts / el, "accepted": accepted,
                   "t_end": t, "h_end": h, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": ach / PEAK_HBM,
                     "traffic": measured_traffic(dom) if (nd, G) == (2, 512) else None,
                     "bytes_per_launch": kb[dom], "ms_per_launch": km[dom]},
        "rhs": {"ms": km["rhs"], "bytes": B_rhs, "achieved_GBps": rhs_ach / 1e9,
                "frac": rhs_ach / PEAK_HBM, "kernels_ms": km, "kernel_bytes": kb,
                "rhs_per_s": 1e3 / km["rhs"]},
    }
    B_att = attempt_bytes(2, M, Ns)
    result["attempt"] = {"bytes": B_att, "us": el / attempts * 1e6,
                         "achieved_GBps": B_att * attempts / el / 1e9,
                         "frac": B_att * attempts / el / PEAK_HBM,
                         "traffic": measured_attempt_traffic() if (nd, G) == (2, 512) else None}

    if args.cpu_seconds > 0:
        result["cpu_baseline"] = cpu_leg(prob, y0, gname, args.warmup, attempts, accepted / el,
                                         attempts / el, args.cpu_seconds)

    it.close()
    if args.sweep > 0 and nd == 2:
        result["batched"] = sweep_leg(lib, args.sweep, args.sweep_t, args.serial_every)
    if args.adjoint and nd == 2:
        result["adjoint"] = adjoint_leg(lib, args.adjoint_batch, args.adjoint_concurrency)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
