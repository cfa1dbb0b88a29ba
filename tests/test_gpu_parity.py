"""GPU parity: the HIP library against the CPU oracle through the same C ABI.

Both follow one operation sequence (radix plan, twiddle tables, fma
placement, reduction trees, correctly rounded pow), so results are compared
for exact equality (== on doubles), not with a tolerance.
"""
import os
import subprocess

import numpy as np
import pytest

from rdcu_amd import Integrator, Problem, Term, cgle_terms, fhn_terms, glibc_rand_ic

pytestmark = pytest.mark.gpu


def mixed_terms(ndim):
    """Every derivative block of both fields appears (first derivatives and
    mixed second derivatives exercise all Nyquist rules)."""
    n = 2
    nblk = 1 + ndim + ndim * ndim
    t = []
    for blk in range(1, nblk):
        for f in range(n):
            t.append(Term(0.5 + 0.1 * blk, f, [(blk * n + f, 1), ((blk * n + 1 - f) % (n * nblk), 1)]))
    t.append(Term(-1.0, 0, [(0, 3)]))
    t.append(Term(0.3, 1, [(0, 2), (1, -1)]))
    t.append(Term(0.7, 0, [(1, 0)]))
    t.append(Term(1.25, 1, []))
    return t


RHS_CASES = [
    ("1d-256", [256], [100.0], "cgle"),
    ("1d-96", [96], [30.0], "mixed"),
    ("2d-64", [64, 64], [50.0, 50.0], "cgle"),
    ("2d-64x96", [64, 96], [40.0, 70.0], "mixed"),
    ("2d-96x64", [96, 64], [70.0, 40.0], "fhn"),
    ("2d-192x128", [192, 128], [150.0, 90.0], "cgle"),
    ("2d-512", [512, 512], [200.0, 200.0], "cgle"),
    ("3d-16x24x32", [16, 24, 32], [10.0, 20.0, 30.0], "mixed"),
    ("3d-32", [32, 32, 32], [30.0, 30.0, 30.0], "cgle"),
    ("2d-64x192", [64, 192], [40.0, 120.0], "cgle"),      # x length 3*2^k: radix-3 pass in LDS
    ("2d-96x384", [96, 384], [60.0, 240.0], "fhn"),
    ("2d-40x20-lds", [40, 20], [30.0, 15.0], "mixed"),    # radix-5 lengths: LDS kernels
    ("3d-10x12x8-lds", [10, 12, 8], [5.0, 6.0, 4.0], "cgle"),
    ("2d-32x64-empty", [32, 64], [20.0, 30.0], "empty"),      # no terms: f = 0
    ("2d-64x32-onefield", [64, 32], [30.0, 20.0], "onefield"),  # field 1 has no terms
    ("2d-1024x32", [1024, 32], [300.0, 20.0], "cgle"),     # long strided lines: twiddles from
    ("2d-1024x16", [1024, 16], [300.0, 10.0], "mixed"),    # global memory in the strided pass
    ("3d-1024x8x16", [1024, 8, 16], [300.0, 8.0, 12.0], "cgle"),
]


def make_problem(Ns, Ls, kind):
    nd = len(Ns)
    if kind == "cgle":
        terms = cgle_terms(nd)
    elif kind == "fhn":
        terms = fhn_terms(nd, 4.0, 0.05, 1.2)
    elif kind == "empty":
        terms = []
    elif kind == "onefield":
        terms = [t for t in cgle_terms(nd) if t.field == 0]
    else:
        terms = mixed_terms(nd)
    return Problem(2, Ns, Ls, terms)


@pytest.mark.parametrize("name,Ns,Ls,kind", RHS_CASES, ids=[c[0] for c in RHS_CASES])
def test_rhs_and_factors_bitwise(name, Ns, Ls, kind, hlib, olib, rng):
    prob = make_problem(Ns, Ls, kind)
    y = rng.uniform(-1.0, 1.0, prob.n * prob.N)
    if kind == "mixed":
        y += 2.0  # keep away from zero (negative power)
    print(f"case {name}", flush=True)
    with Integrator(hlib, prob, y) as g, Integrator(olib, prob, y) as o:
        fg, Yg = g.rhs(y, factors=True)
        fo, Yo = o.rhs(y, factors=True)
    assert np.array_equal(Yg.real, Yo.real), np.abs(Yg.real - Yo.real).max()
    # imaginary parts: the Nyquist leakage of first and mixed derivatives
    assert np.array_equal(Yg.imag, Yo.imag), np.abs(Yg.imag - Yo.imag).max()
    assert np.array_equal(fg, fo), np.abs(fg - fo).max()


@pytest.mark.parametrize("Ns,Ls,kind", [([64, 96], [40.0, 70.0], "mixed"),
                                         ([16, 24, 32], [10.0, 20.0, 30.0], "mixed"),
                                         ([256], [100.0], "cgle")])
def test_register_and_lds_paths_identical(hlib, rng, monkeypatch, Ns, Ls, kind):
    """The register-resident kernels (with stage fusion) and the LDS kernels
    give identical RHS values and step sequences."""
    prob = make_problem(Ns, Ls, kind)
    y0 = rng.uniform(1.0, 2.0, prob.n * prob.N) if kind == "mixed" else \
        glibc_rand_ic(prob.n, prob.N, seed=5, A=1.0)
    res = []
    for path in ("reg", "lds"):
        monkeypatch.setenv("RDCU_PATH", path)
        with Integrator(hlib, prob, y0) as g:
            f = g.rhs(y0)
            # bounded by attempts: the "mixed" test dictionary blows up in finite time
            t, h, done = g.advance(0.0, 1e-3, max_attempts=30)
            res.append((f, (t, h, done), g.state()))
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1] and res[0][1][2] == 30
    assert np.array_equal(res[0][2], res[1][2], equal_nan=True)


def test_rhs_plain_equals_factor_path(hlib, rng):
    """f from the referenced-factor pipeline equals f recomputed with the
    full factor set materialised (same kernels, different factor list)."""
    prob = make_problem([64, 96], [40.0, 70.0], "mixed")
    y = rng.uniform(1.0, 2.0, prob.n * prob.N)
    with Integrator(hlib, prob, y) as g:
        f1 = g.rhs(y)
        f2, _ = g.rhs(y, factors=True)
        f3 = g.rhs(y)
    assert np.array_equal(f1, f2) and np.array_equal(f1, f3)


STEP_CASES = [
    ("1d-256", [256], [100.0], "cgle", 1.0),
    ("2d-64", [64, 64], [50.0, 50.0], "cgle", 1.0),
    ("2d-96x64-fhn", [96, 64], [70.0, 40.0], "fhn", 1.0),
    ("3d-16x24x32", [16, 24, 32], [10.0, 20.0, 30.0], "cgle", 1.0),
]


@pytest.mark.parametrize("name,Ns,Ls,kind,A", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_step_sequence_bitwise(name, Ns, Ls, kind, A, hlib, olib):
    """rdcu_step attempts (accepted and rejected) give identical t, h, y."""
    prob = make_problem(Ns, Ls, kind)
    y0 = glibc_rand_ic(prob.n, prob.N, seed=1, A=A)
    print(f"case {name}", flush=True)
    with Integrator(hlib, prob, y0) as g, Integrator(olib, prob, y0) as o:
        tg = to = 0.0
        hg = ho = 0.01
        nacc = nrej = 0
        for i in range(25):
            tg, hg, ag = g.step(tg, hg)
            to, ho, ao = o.step(to, ho)
            assert (tg, hg, ag) == (to, ho, ao), (i, tg, to, hg, ho, ag, ao)
            nacc += ag
            nrej += not ag
        assert np.array_equal(g.state(), o.state())
        if ag:
            assert np.array_equal(g.eval(tg - 0.3 * hg), o.eval(to - 0.3 * ho))
    assert nacc > 5


@pytest.mark.parametrize("name,Ns,Ls,kind,A", STEP_CASES[:2], ids=[c[0] for c in STEP_CASES[:2]])
def test_advance_in_groups_bitwise(name, Ns, Ls, kind, A, hlib, olib):
    """rdcu_advance launches whole groups of attempts through one graph and the
    rest one by one; counts that are not multiples of the group, and an end
    time reached inside a group, give the oracle's t, h, attempt count and y."""
    prob = make_problem(Ns, Ls, kind)
    y0 = glibc_rand_ic(prob.n, prob.N, seed=1, A=A)
    print(f"case {name}", flush=True)
    with Integrator(hlib, prob, y0) as g, Integrator(olib, prob, y0) as o:
        tg = to = 0.0
        hg = ho = 0.01
        for k in (13, 8, 21, 3):
            tg, hg, dg = g.advance(tg, hg, max_attempts=k)
            to, ho, do = o.advance(to, ho, max_attempts=k)
            assert (tg, hg, dg) == (to, ho, do), (k, tg, to, hg, ho, dg, do)
            assert dg == k
        t1 = tg + 9.5 * hg  # ends a few attempts into the next group
        tg, hg, dg = g.advance(tg, hg, t1=t1, max_attempts=64)
        to, ho, do = o.advance(to, ho, t1=t1, max_attempts=64)
        assert (tg, hg, dg) == (to, ho, do), (tg, to, hg, ho, dg, do)
        assert 0 < dg < 64 and tg >= t1
        assert np.array_equal(g.state(), o.state())


def test_run_with_outputs_bitwise(hlib, olib):
    """rdcu_run with output stops (device-side halting, chunked replay) gives
    the same accepted/rejected sequence and dense outputs as the oracle."""
    prob = Problem(2, [256], [100.0], cgle_terms(1))
    y0 = glibc_rand_ic(2, 256, seed=1, A=1.0)
    res = {}
    for name, lib in (("gpu", hlib), ("cpu", olib)):
        with Integrator(lib, prob, y0) as it:
            log = []
            outs = []
            nxt = [1.0]

            def cb(t, h, ok, cur):
                log.append((t, h, ok))
                while ok and t >= nxt[0] and nxt[0] <= 10.0:
                    assert cur
                    outs.append(it.eval(nxt[0]))
                    nxt[0] += 1.0
                return nxt[0]

            t, h = it.run(0.0, 0.01, 10.0, stop=1.0, callback=cb)
            res[name] = (log, outs, t, h, it.state())
    lg, og, tg, hg, yg = res["gpu"]
    lc, oc, tc, hc, yc = res["cpu"]
    assert len(lg) == len(lc)
    assert lg == lc
    assert len(og) == len(oc) == 10
    for a, b in zip(og, oc):
        assert np.array_equal(a, b)
    assert (tg, hg) == (tc, hc)
    assert np.array_equal(yg, yc)


def test_fixed_step_bitwise(hlib, olib):
    prob = Problem(2, [64, 64], [50.0, 50.0], cgle_terms(2))
    y0 = glibc_rand_ic(2, 64 * 64, seed=3, A=1.0)
    out = []
    for lib in (hlib, olib):
        with Integrator(lib, prob, y0, fixed=True) as it:
            t, h = it.run(0.0, 0.01, 0.2)
            out.append((t, h, it.state()))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1] == 0.01
    assert np.array_equal(out[0][2], out[1][2])


def _run_cli(exe, args, cwd):
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


CLI_CASES = [(["-N", "256", "-L", "100", "-t", "5", "-d", "0.5"], 1),
             (["-N", "64,96", "-L", "40,70", "-t", "1", "-d", "0.25"], 2),
             (["-N", "16,24,32", "-L", "10,20,30", "-t", "0.2", "-d", "0.1"], 3)]


@pytest.mark.parametrize("grid,nd", CLI_CASES, ids=["1d", "2d", "3d"])
def test_cli_outputs_identical(tmp_path, grid, nd):
    """The drop-in CLI on the GPU writes byte-identical output files to the
    same CLI on the oracle (CGLE, -D3: states, f and Y frames through
    rdcu_frame, restart file, steps and times)."""
    from conftest import HIP_CLI, ORACLE_CLI
    args = ["-n", "2"] + grid + ["-A", "1", "-D3"]
    args += sum((["-c", c] for c in __import__("rdcu_amd").to_cli(cgle_terms(nd))), [])
    for sub, exe in (("g", HIP_CLI), ("c", ORACLE_CLI)):
        d = tmp_path / sub
        d.mkdir()
        _run_cli(exe, args + ["run"], str(d))
    for suf in ("steps.dat", "times.dat", "states.dat", "f.dat", "Y.dat", "fs.dat"):
        a = (tmp_path / "g" / ("run" + suf)).read_bytes()
        b = (tmp_path / "c" / ("run" + suf)).read_bytes()
        assert a == b, suf


# (the mixed dictionary divides by u1, so its steps are kept tiny)
FRAME_CASES = [([256], [100.0], "cgle", 0.01, 0.05), ([64, 96], [40.0, 70.0], "mixed", 1e-6, 3e-6),
               ([16, 24, 32], [10.0, 20.0, 30.0], "mixed", 1e-6, 3e-6),
               ([64, 64], [50.0, 50.0], "cgle", 0.01, 0.05)]


@pytest.mark.parametrize("Ns,Ls,kind,h0,t1", FRAME_CASES, ids=["1d", "2d", "3d", "2d-cgle"])
def test_frame_bitwise(Ns, Ls, kind, h0, t1, hlib, olib):
    """rdcu_frame (state kept on the device, Y packed there) gives the bits of
    rdcu_eval then rdcu_rhs, and the oracle's frame, at every -D level."""
    prob = make_problem(Ns, Ls, kind)
    N = int(np.prod(Ns))
    y0 = glibc_rand_ic(2, N, seed=5, A=1.0)
    out = {}
    for name, lib in (("gpu", hlib), ("cpu", olib)):
        with Integrator(lib, prob, y0) as it:
            t, h = it.run(0.0, h0, t1)
            y = it.eval(t1)
            f, Y = it.rhs(y, factors=True)
            for dense in (3, 1, 2):
                y2, f2, Y2 = it.frame(t1, dense)
                assert np.array_equal(y2, y)
                if dense >= 2:
                    assert np.array_equal(f2, f)
                if dense >= 3:
                    assert np.array_equal(Y2, Y)
            out[name] = (t, h, y, f, Y)
    g, c = out["gpu"], out["cpu"]
    assert g[:2] == c[:2]
    for a, b in zip(g[2:], c[2:]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("Ns,Ls,shapes", [([512, 512], [200.0, 200.0], "2,8,1,4,1,16"),
                                          ([64, 128], [50.0, 90.0], "4,4,2,2,2,16"),
                                          ([32, 32, 32], [30.0, 30.0, 30.0], "2,8,1,4,1,16")],
                         ids=["2d-512", "2d-64x128", "3d-32"])
def test_sixteen_values_per_thread_columns_bitwise(hlib, olib, monkeypatch, Ns, Ls, shapes):
    """Strided passes with 16 values per thread (the same butterflies as with
    8) give the oracle's bits; shapes forced so the case is always covered."""
    prob = make_problem(Ns, Ls, "cgle")
    y0 = glibc_rand_ic(prob.n, prob.N, seed=6, A=1.0)
    monkeypatch.setenv("RDCU_SHAPES", shapes)
    print(f"case E16 {Ns}", flush=True)
    with Integrator(hlib, prob, y0) as g, Integrator(olib, prob, y0) as o:
        assert np.array_equal(g.rhs(y0), o.rhs(y0))
        tg = to = 0.0
        hg = ho = 0.01
        for i in range(3):
            tg, hg, ag = g.step(tg, hg)
            to, ho, ao = o.step(to, ho)
            assert (tg, hg, ag) == (to, ho, ao)
        assert np.array_equal(g.state(), o.state())


@pytest.mark.parametrize("Ns,Ls,shapes", [([512, 512], [200.0, 200.0], "8,4,2,4,2,8"),
                                          ([512, 512], [200.0, 200.0], "8,4,1,4,2,8"),
                                          ([256, 256], [100.0, 100.0], "4,4,2,4,2,16"),
                                          ([256, 16, 32], [60.0, 20.0, 30.0], "4,4,1,4,1,8"),
                                          ([512, 512], [200.0, 200.0], "2,5,1,4,1,8,64"),
                                          ([512, 512], [200.0, 200.0], "2,3,2,4,1,8,128"),
                                          ([96, 48, 64], [40.0, 20.0, 30.0], "4,2,1,4,1,8,11"),
                                          ([512, 512], [200.0, 200.0], "2,4,1,4,1,8,0,1"),
                                          ([32, 64, 256], [30.0, 40.0, 90.0], "4,4,1,4,1,8,0,1"),
                                          ([512, 512], [200.0, 200.0], "2,4,1,4,2,8,0,2"),
                                          ([32, 64, 256], [30.0, 40.0, 90.0], "4,4,1,4,1,8,0,2"),
                                          ([32, 64, 256], [30.0, 40.0, 90.0], "4,4,1,4,1,8,0,3"),
                                          ([96, 192], [40.0, 80.0], "4,4,1,4,1,8,0,3")],
                         ids=["2d-512-fpw2", "2d-512-fpw1", "2d-256-e16", "3d-256z",
                              "2d-512-64tiles", "2d-512-128tiles-fpw2", "3d-11tiles",
                              "2d-512-twreg", "3d-256x-twreg", "2d-512-lean-R2", "3d-256x-lean",
                              "3d-256x-lean-twreg", "2d-96x192-lean-twreg"])
def test_compiled_length_kernels_bitwise(hlib, olib, monkeypatch, Ns, Ls, shapes):
    """Kernels compiled for one transform length (128/256-point x lines,
    256/512-point strided lines), with two factors per block (fpw = 2) or one
    (fpw = 1), 8 or 16 values per thread, column tiles of unequal widths
    (the 257 columns of a 512-point row over 64 or 128 tiles; 33 over 11),
    and the generated x-inverse kernel with its twiddles in registers, and
    both compiled lean for two wavefronts per SIMD: the oracle's bits."""
    prob = make_problem(Ns, Ls, "cgle")
    y0 = glibc_rand_ic(prob.n, prob.N, seed=7, A=1.0)
    monkeypatch.setenv("RDCU_SHAPES", shapes)
    print(f"case compiled lengths {Ns} {shapes}", flush=True)
    with Integrator(hlib, prob, y0) as g, Integrator(olib, prob, y0) as o:
        assert np.array_equal(g.rhs(y0), o.rhs(y0))
        tg = to = 0.0
        hg = ho = 0.01
        for i in range(3):
            tg, hg, ag = g.step(tg, hg)
            to, ho, ao = o.step(to, ho)
            assert (tg, hg, ag) == (to, ho, ao)
        assert np.array_equal(g.state(), o.state())


def test_generated_kernels_compile_and_do_not_spill():
    """The dictionary-specialised x-inverse kernels (and the 1-D attempt
    kernel, and the gradient's row kernel) are generated at run time; a
    generated kernel that keeps values in scratch memory is rejected (the
    library then uses its interpreter).  For the bench shapes, short 3*2^k
    rows and 3-D, the generated kernels must be in use."""
    import sys
    code = (
        "import numpy as np\n"
        "from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic, hip_lib\n"
        "lib = hip_lib()\n"
        "for Ns in ([256], [48], [512, 512], [64, 192], [96, 96], [32, 32, 32]):\n"
        "    nd = len(Ns); N = int(np.prod(Ns))\n"
        "    print('case', Ns, flush=True)\n"
        "    with Integrator(lib, Problem(2, Ns, [50.0] * nd, cgle_terms(nd)),\n"
        "                    glibc_rand_ic(2, N, seed=1, A=1.0)) as it:\n"
        "        if nd == 2 and Ns[0] == 64:  # the gradient kernel is built on first use\n"
        "            y0 = glibc_rand_ic(2, N, seed=1, A=1.0)\n"
        "            it.loss_grad(y0, 0.0, np.full(3, 1e-3), [2e-3], np.zeros(2 * N))\n")
    from conftest import ROOT
    env = dict(os.environ, RDCU_VERBOSE="1", PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       env=env, cwd=ROOT)
    assert r.returncode == 0, r.stdout + r.stderr
    out = (r.stdout + r.stderr).split("\n")
    status = [l for l in out if l.startswith("rdcu jit:")]
    # 6 handles, plus the gradient's companion problem
    assert len(status) == 7 and all(l == "rdcu jit: ok" for l in status), status
    assert [l for l in out if l.startswith("rdcu vjp jit:")] == ["rdcu vjp jit: ok"]
