"""GPU parity at the benchmark configurations' full sizes.

  1-D CGLE, N=256, L=100, 14 terms, amplitude 1, t=100, -D3: the whole run,
      CLI output files byte for byte against the oracle's CLI.
  2-D CGLE 512x512, L=200, 18 terms, amplitude 1: attempts step by step,
      RHS and every referenced factor, and dense output, bit for bit.
  3-D CGLE 256^3, L=100, 22 terms, amplitude 0.1: RHS and the first attempts
      bit for bit (the oracle needs a few seconds per RHS here).

Size-independent properties at full size: dense output at the end of an
accepted step equals the state (x = 1 of the interpolant), and a state
written back with set_state continues identically.
"""
import numpy as np
import pytest

import rdcu_amd
from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic

pytestmark = pytest.mark.gpu


def _run_cli(exe, args, cwd):
    import subprocess
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def test_c1_full_run_files_identical(tmp_path):
    from conftest import HIP_CLI, ORACLE_CLI
    args = ["-n", "2", "-N", "256", "-L", "100", "-t", "100", "-d", "1", "-r", "1e-8", "-a",
            "1e-8", "-A", "1", "-s", "1", "-D3"]
    args += sum((["-c", c] for c in rdcu_amd.to_cli(cgle_terms(1))), [])
    print("case c1 full run", flush=True)
    for sub, exe in (("g", HIP_CLI), ("c", ORACLE_CLI)):
        d = tmp_path / sub
        d.mkdir()
        _run_cli(exe, args + ["run"], str(d))
    steps = (tmp_path / "g" / "runsteps.dat").read_bytes()
    assert len(steps) > 16 * 1000  # thousands of attempts
    for suf in ("steps.dat", "times.dat", "states.dat", "f.dat", "Y.dat", "fs.dat"):
        a = (tmp_path / "g" / ("run" + suf)).read_bytes()
        b = (tmp_path / "c" / ("run" + suf)).read_bytes()
        assert a == b, suf


def test_c2_full_size_bitwise(hlib, olib):
    prob = Problem(2, [512, 512], [200.0, 200.0], cgle_terms(2, 2.0, 0.8))
    y0 = glibc_rand_ic(2, 512 * 512, seed=1, A=1.0)
    print("case c2 512x512", flush=True)
    with Integrator(hlib, prob, y0) as g, Integrator(olib, prob, y0) as o:
        fg, Yg = g.rhs(y0, factors=True)
        fo, Yo = o.rhs(y0, factors=True)
        assert np.array_equal(fg, fo)
        assert np.array_equal(Yg.real, Yo.real) and np.array_equal(Yg.imag, Yo.imag)
        tg = to = 0.0
        hg = ho = 0.01
        for i in range(6):
            tg, hg, ag = g.step(tg, hg)
            to, ho, ao = o.step(to, ho)
            assert (tg, hg, ag) == (to, ho, ao), i
            assert np.array_equal(g.state(), o.state()), i
        assert ag
        # property: the interpolant at x = 1 reproduces the accepted state
        ye = g.eval(tg)
        assert np.array_equal(ye, o.eval(to))
        assert np.allclose(ye, g.state(), rtol=0, atol=1e-12)
        # property: writing the state back and continuing changes nothing
        y = g.state()
        g.set_state(y)
        o.set_state(y)
        for i in range(2):
            tg, hg, ag = g.step(tg, hg)
            to, ho, ao = o.step(to, ho)
            assert (tg, hg, ag) == (to, ho, ao)
        assert np.array_equal(g.state(), o.state())


def test_c3_full_size_bitwise(hlib, olib):
    prob = Problem(2, [256, 256, 256], [100.0, 100.0, 100.0], cgle_terms(3, 2.0, 0.8))
    y0 = glibc_rand_ic(2, 256 ** 3, seed=1, A=0.1)
    print("case c3 256^3", flush=True)
    with Integrator(hlib, prob, y0) as g:
        fg = g.rhs(y0)
        steps = []
        tg, hg = 0.0, 0.01
        for i in range(2):
            tg, hg, ag = g.step(tg, hg)
            steps.append((tg, hg, ag))
        yg = g.state()
    with Integrator(olib, prob, y0) as o:
        assert np.array_equal(fg, o.rhs(y0))
        to, ho = 0.0, 0.01
        for i in range(2):
            to, ho, ao = o.step(to, ho)
            assert steps[i] == (to, ho, ao), i
        assert np.array_equal(yg, o.state())


@pytest.mark.parametrize("Ns,Ls", [([2048], [400.0]), ([2048, 2048], [400.0, 400.0])],
                         ids=["1d-2048", "2d-2048"])
def test_largest_register_rows_bitwise(hlib, olib, Ns, Ls):
    """The longest x lines the register path takes (2048 points, 128 threads
    per line): RHS and two attempts bit for bit."""
    prob = Problem(2, Ns, Ls, cgle_terms(len(Ns), 2.0, 0.8))
    y0 = glibc_rand_ic(2, prob.N, seed=2, A=1.0)
    print(f"case largest rows {Ns}", flush=True)
    with Integrator(hlib, prob, y0) as g, Integrator(olib, prob, y0) as o:
        assert np.array_equal(g.rhs(y0), o.rhs(y0))
        tg = to = 0.0
        hg = ho = 0.005
        for i in range(2):
            tg, hg, ag = g.step(tg, hg)
            to, ho, ao = o.step(to, ho)
            assert (tg, hg, ag) == (to, ho, ao)
        assert np.array_equal(g.state(), o.state())


def test_unsupported_sizes_fail_cleanly(hlib):
    """Lengths no kernel takes are refused at creation (before any launch)."""
    from rdcu_amd import RdcuError
    for Ns in ([8192], [30, 31], [64, 4096]):
        prob = Problem(2, Ns, [10.0] * len(Ns), cgle_terms(len(Ns)))
        with pytest.raises(RdcuError) as e:
            Integrator(hlib, prob, np.zeros(2 * prob.N))
        assert e.value.code == 2  # RDCU_EUNSUPPORTED
