"""Host-side pieces of bench.py (no GPU): the algorithmic byte model and the
measured-traffic figure read from the committed rocprof summary."""
import os
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_algorithmic_bytes_match_the_closed_form():
    """B_rhs = 8N[n(2 ndim + 1) + 2M(ndim - 1)]: 37,748,736 B at 512^2 (n=2, M=4)
    and 5.10 GB at 256^3 (n=2, M=6)."""
    kb = bench.algorithmic_bytes(2, 4, [512, 512])
    assert kb["row_fwd"] + kb["col_mult"] + kb["row_inv"] == 37748736
    N = 256 ** 3
    kb3 = bench.algorithmic_bytes(2, 6, [256] * 3)
    total = sum(kb3[k] for k in ("row_fwd", "col_fwd", "col_mult", "col_inv", "row_inv"))
    assert total == 8 * N * (2 * 7 + 2 * 6 * 2)


def test_measured_traffic_comes_from_the_counter_table():
    """The traffic figure is the PMC table's scaled total for the kernel, not a
    number from the kernel-trace table above it (both tables have six
    columns and name the same kernels)."""
    assert os.path.exists(bench.PROFILE)
    rows = {}
    seen = False
    for line in open(bench.PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[-1] == "scaled total MB":
            seen = True
        elif seen and len(cells) == 6 and cells[0].startswith("`"):
            rows[cells[0]] = float(cells[5]) * 1e6
    assert seen
    col = [v for k, v in rows.items() if "k_col_reg" in k]
    assert bench.measured_traffic("col_mult") == col[0]
    # within 10% of the kernel's algorithmic bytes (no re-reads)
    alg = bench.algorithmic_bytes(2, 4, [512, 512])["col_mult"]
    assert 1.0 <= bench.measured_traffic("col_mult") / alg < 1.1


def test_attempt_bytes_match_the_per_kernel_listing():
    """The whole-attempt byte model against the attempt profile's table, which
    lists each kernel's arrays one by one (the strided row covers its six
    launches): 339.7 MB at 512^2 (n=2, M=4)."""
    path = os.path.join(ROOT, "profiles", "attempt_2d_cgle_512.md")
    total = 0.0
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 8 and cells[0][:1].isdigit():
            mb = float(cells[4])
            total += 6 * mb if "," in cells[0] else mb
    assert abs(total - 339.7) < 0.2
    assert abs(bench.attempt_bytes(2, 4, [512, 512]) / 1e6 - total) < 0.2
    # 3-D adds the y forward (2A) and y inverse (2F) passes to each of the six RHS
    N = 256 ** 3
    A, F = 16 * N, 48 * N
    assert bench.attempt_bytes(2, 6, [256] * 3) - bench.attempt_bytes(2, 6, [256, 256 * 256]) \
        == 6 * (2 * A + 2 * F)
