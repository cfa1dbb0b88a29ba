"""The C ABI boundary: both libraries load and export every function that
include/rdcu_dp45.h declares (no GPU call is made here)."""
import ctypes
import os
import re

import pytest

from conftest import ORACLE_LIB, ROOT

HEADER = os.path.join(ROOT, "include", "rdcu_dp45.h")
HIP_LIB = os.path.join(ROOT, "rdcu_amd", "lib", "librdcu_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(rdcu_[a-z_]+)\s*\(", text))
    names.discard("rdcu_step_fn")
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared()
    for n in ("rdcu_create", "rdcu_step", "rdcu_run", "rdcu_eval", "rdcu_destroy", "rdcu_rhs"):
        assert n in names


@pytest.mark.parametrize("path", [ORACLE_LIB, HIP_LIB], ids=["oracle", "hip"])
def test_library_exports_every_declared_symbol(path):
    assert os.path.exists(path), f"{path} not built"
    lib = ctypes.CDLL(path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_symbol():
    from rdcu_amd.abi import SYMBOLS
    assert sorted(SYMBOLS) == declared()


def test_backend_names():
    from rdcu_amd.abi import load
    assert load(ORACLE_LIB).rdcu_backend() == b"cpu-oracle"
    assert load(HIP_LIB).rdcu_backend() == b"hip-gfx950"


def test_hip_cli_help_runs_without_gpu(tmp_path):
    """-h prints the help and index listing before any device call."""
    import subprocess
    exe = os.path.join(ROOT, "rdcu_amd", "bin", "rdcu")
    r = subprocess.run([exe, "-n", "2", "-N", "8,8", "-h"], capture_output=True, text=True,
                       timeout=60, cwd=tmp_path)
    assert r.returncode == 0 and "13: u1_11" in r.stdout


def test_error_codes_on_the_oracle(olib):
    """Statuses of bad calls (the reference returns none; dp45.cu never checks)."""
    import numpy as np
    import pytest as _pt
    from rdcu_amd import Integrator, Problem, RdcuError, Term, cgle_terms
    from rdcu_amd.abi import EINVAL, ESTATE, EUNSUPPORTED

    def code_of(fn):
        with _pt.raises(RdcuError) as e:
            fn()
        return e.value.code

    assert code_of(lambda: Integrator(olib, Problem(2, [31], [1.0], []), np.zeros(62))) == EUNSUPPORTED
    bad = Problem(2, [16], [1.0], [Term(1.0, 0, [(99, 1)])])
    assert code_of(lambda: Integrator(olib, bad, np.zeros(32))) == EINVAL
    badf = Problem(2, [16], [1.0], [Term(1.0, 5, [(0, 1)])])
    assert code_of(lambda: Integrator(olib, badf, np.zeros(32))) == EINVAL
    p = Problem(2, [16], [8.0], cgle_terms(1))
    with Integrator(olib, p, np.full(32, 0.1)) as it:
        assert code_of(lambda: it.eval(0.5)) == ESTATE  # no accepted step yet
        assert code_of(lambda: it.frame(0.5)) == ESTATE
        hs = np.full(4, 0.01)
        tgt = np.zeros(32)
        assert code_of(lambda: it.loss_grad(np.zeros(32), 0.0, hs, [1.0], tgt)) == EINVAL
        assert code_of(lambda: it.loss_grad(np.zeros(32), 0.0, hs, [0.0], tgt)) == EINVAL
        with _pt.raises(ValueError):
            it.set_coef([1.0])


def test_oracle_frame_is_eval_then_rhs(olib):
    """rdcu_frame on the oracle: the dense output, then the RHS and factor
    blocks at it, at each -D level (the CLI's output path)."""
    import numpy as np
    from rdcu_amd import Integrator, Problem, cgle_terms, glibc_rand_ic
    prob = Problem(2, [32, 16], [20.0, 10.0], cgle_terms(2))
    with Integrator(olib, prob, glibc_rand_ic(2, 512, seed=2, A=1.0)) as it:
        t, h = it.run(0.0, 0.01, 0.05)
        te = 0.05
        y = it.eval(te)
        f, Y = it.rhs(y, factors=True)
        for dense in (1, 2, 3):
            y2, f2, Y2 = it.frame(te, dense)
            assert np.array_equal(y2, y)
            assert (f2 is None) == (dense < 2) and (Y2 is None) == (dense < 3)
            if dense >= 2:
                assert np.array_equal(f2, f)
            if dense >= 3:
                assert Y2.shape == (7, 2, 512) and np.array_equal(Y2, Y)


def test_host_alloc_round_trip(olib):
    vp = ctypes.c_void_p()
    assert olib.rdcu_host_alloc(4096, ctypes.byref(vp)) == 0 and vp.value
    ctypes.memset(vp, 0, 4096)
    assert olib.rdcu_host_free(vp) == 0
