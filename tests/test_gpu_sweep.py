"""rdcu_sweep on the GPU: every manifest line's files equal, byte for byte,
those of the same line run alone by rdcu (HIP) and by the oracle CLI; a group
the batch cannot hold (3-D) runs its instances one at a time."""
import pytest

from conftest import HIP_CLI, HIP_SWEEP, ORACLE_CLI
from sweep_cases import CGLE2, SUFFIXES, run_exe, sweep_vs_single, write_manifest

pytestmark = pytest.mark.gpu


def test_gpu_sweep_files_equal_single_runs(tmp_path):
    print("case sweep vs single (hip)", flush=True)
    bad, sw, si = sweep_vs_single(HIP_SWEEP, HIP_CLI, tmp_path)
    assert not bad, bad
    # and the single HIP runs equal the oracle's
    oc = tmp_path / "oracle"
    oc.mkdir()
    from sweep_cases import LINES, RESTART
    for name, args in LINES.items():
        run_exe(ORACLE_CLI, args.split() + [f"s_{name}"], str(oc))
        if name in RESTART:
            run_exe(ORACLE_CLI, args.split() + RESTART[name].split() + [f"s_{name}"], str(oc))
    for name in LINES:
        for suf in SUFFIXES:
            a, b = si / f"s_{name}{suf}", oc / f"s_{name}{suf}"
            assert a.exists() == b.exists(), (name, suf)
            if a.exists():
                assert a.read_bytes() == b.read_bytes(), (name, suf)


def test_gpu_sweep_unbatchable_group_runs_singly(tmp_path):
    from rdcu_amd import cgle_terms, to_cli
    c3 = " ".join("-c " + c for c in to_cli(cgle_terms(3, 2.0, 0.8)))
    lines = {"p": f"-n 2 -N 16,16,16 -L 20,20,20 -t 0.5 -d 0.25 -A 1 -s 2 -r 1e-7 -a 1e-7 -D1 {c3}",
             "q": f"-n 2 -N 32,32 -L 30,30 -t 0.5 -d 0.25 -A 1 -s 4 -D1 {CGLE2}"}
    print("case sweep with a 3-D group", flush=True)
    write_manifest(tmp_path / "m.txt", lines, "u_")
    r = run_exe(HIP_SWEEP, ["m.txt"], str(tmp_path))
    assert "one at a time" in r.stderr
    single = tmp_path / "single"
    single.mkdir()
    for name, args in lines.items():
        run_exe(HIP_CLI, args.split() + [f"u_{name}"], str(single))
        for suf in ("steps.dat", "times.dat", "states.dat", "fs.dat"):
            assert (tmp_path / f"u_{name}{suf}").read_bytes() == \
                (single / f"u_{name}{suf}").read_bytes(), (name, suf)
