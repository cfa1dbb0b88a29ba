"""Manifests for the batched sweep front end (rdcu_sweep): each line is an
rdcu command line; the files every line's FILEBASE receives from rdcu_sweep
must equal, byte for byte, those of the same line run alone by rdcu."""
import os
import subprocess

from rdcu_amd import cgle_terms, fhn_terms, to_cli

SUFFIXES = ("steps.dat", "times.dat", "states.dat", "f.dat", "Y.dat", "fs.dat")


def _terms(ts):
    return " ".join("-c " + c for c in to_cli(ts))


CGLE1 = _terms(cgle_terms(1, 2.0, 0.8))
CGLE2 = _terms(cgle_terms(2, 2.0, 0.8))
FHN2 = _terms(fhn_terms(2, 4.0, 0.05, 1.2))

# name -> arguments (FILEBASE appended); mixed grids, dictionaries, output
# densities, tolerances (a second batch) and a fixed-step run (a third)
LINES = {
    "a": f"-n 2 -N 64 -L 30 -t 2 -d 0.5 -A 1 -s 3 -D3 {CGLE1}",
    "b": f"-n 2 -N 32,48 -L 20,30 -t 1 -d 0.25 -A 1 -s 5 -D2 {CGLE2}",
    "c": f"-n 2 -N 32,32 -L 40,40 -t 1 -d 0.5 -A 1 -s 7 -D1 {FHN2}",
    "d": f"-n 2 -N 96 -L 40 -t 1.5 -d 0.5 -A 1 -s 9 -D1 {CGLE1}",
    "e": f"-n 2 -N 64 -L 30 -t 1 -d 0.25 -A 1 -s 11 -r 1e-6 -a 1e-6 -D1 {CGLE1}",
    "f": f"-n 2 -N 64 -L 30 -t 0.2 -d 0.1,0.02 -A 1 -s 13 -F -D2 {CGLE1}",
}
# second phase: -R continues from FILEBASEfs.dat and appends
RESTART = {"a": "-t 3 -R", "b": "-t 1.5 -R"}


def write_manifest(path, lines, prefix, extra=None):
    with open(path, "w") as f:
        f.write("# rdcu_sweep manifest\n\n")
        for name, args in lines.items():
            more = (" " + extra[name]) if extra and name in extra else ""
            if extra is not None and name not in extra:
                continue
            f.write(f"{args}{more} {prefix}{name}   # instance {name}\n")


def run_exe(exe, args, cwd):
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def sweep_vs_single(sweep_exe, single_exe, tmp_path):
    """Runs LINES (and the RESTART phase) through the sweep and one by one;
    returns the list of (name, suffix) file pairs that differ."""
    sw, si = tmp_path / "sweep", tmp_path / "single"
    sw.mkdir()
    si.mkdir()
    write_manifest(sw / "m1.txt", LINES, "s_")
    run_exe(sweep_exe, ["m1.txt"], str(sw))
    write_manifest(sw / "m2.txt", LINES, "s_", RESTART)
    run_exe(sweep_exe, ["m2.txt"], str(sw))
    for name, args in LINES.items():
        run_exe(single_exe, args.split() + [f"s_{name}"], str(si))
        if name in RESTART:
            run_exe(single_exe, args.split() + RESTART[name].split() + [f"s_{name}"], str(si))
    bad = []
    for name in LINES:
        for suf in SUFFIXES:
            a, b = sw / f"s_{name}{suf}", si / f"s_{name}{suf}"
            if a.exists() != b.exists() or (a.exists() and a.read_bytes() != b.read_bytes()):
                bad.append((name, suf))
    return bad, sw, si
