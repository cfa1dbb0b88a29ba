"""The batched sweep front end (rdcu_sweep MANIFEST) through its oracle build
on the CPU: every manifest line's files equal those of the same line run
alone by the CLI, byte for byte, including a -R continuation phase; manifest
errors are reported with their line."""
import pytest

from conftest import ORACLE_CLI, ORACLE_SWEEP
from sweep_cases import LINES, run_exe, sweep_vs_single


def test_sweep_files_equal_single_runs(tmp_path):
    bad, sw, si = sweep_vs_single(ORACLE_SWEEP, ORACLE_CLI, tmp_path)
    assert not bad, bad
    # the runs wrote something: several frames and many attempts
    assert (sw / "s_atimes.dat").stat().st_size == 8 * 7  # t = 0, 0.5, ..., 3
    assert (sw / "s_asteps.dat").stat().st_size > 8 * 20
    # fixed h = 0.02: ten steps sum to just under 0.2, so an eleventh runs
    # (the reference loop has no clip to t1)
    assert (sw / "s_fsteps.dat").stat().st_size == 8 * 11


def test_sweep_manifest_errors(tmp_path):
    import subprocess
    (tmp_path / "m.txt").write_text("-n 2 -N 64 -L 30 -S x\n")
    r = subprocess.run([ORACLE_SWEEP, "m.txt"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "line 1" in r.stderr
    (tmp_path / "m2.txt").write_text("# nothing but flags\n-n 2 -N 64\n")
    r = subprocess.run([ORACLE_SWEEP, "m2.txt"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "line 2" in r.stderr and "no FILEBASE" in r.stderr
    r = subprocess.run([ORACLE_SWEEP], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
