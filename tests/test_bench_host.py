"""Host-side pieces of bench.py (no GPU): the algorithmic byte model and the
measured-traffic figure read from the committed rocprof summary."""
import os
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_algorithmic_bytes_match_the_closed_form():
    """B_rhs = 8N[n(2 ndim + 1) + 2M] in 2-D: 37,748,736 B at 512^2 (n=2, M=4).
    In 3-D the factors without k_z skip the z pass and the inverse y pass:
    8N[7n + 2M + 2Mz], 4.03 GB at 256^3 (n=2, M=6, Mz=2), against 5.10 GB
    when every factor goes through both (Mz = M)."""
    kb = bench.algorithmic_bytes(2, 4, [512, 512])
    assert kb["row_fwd"] + kb["col_mult"] + kb["row_inv"] == 37748736
    N = 256 ** 3
    keys = ("row_fwd", "col_fwd", "col_mult", "col_inv", "row_inv")
    kb3 = bench.algorithmic_bytes(2, 6, [256] * 3)
    assert sum(kb3[k] for k in keys) == 8 * N * (7 * 2 + 2 * 6 + 2 * 2)
    kb3_all = bench.algorithmic_bytes(2, 6, [256] * 3, Mz=6)
    assert sum(kb3_all[k] for k in keys) == 8 * N * (2 * 7 + 2 * 6 * 2)


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
    # 3-D adds to each of the six RHS the y forward pass (2A, and the writes of
    # the factors without k_z moved out of the last-axis pass) and the inverse
    # y pass of the factors with k_z (2Fz)
    N = 256 ** 3
    A, Fz = 16 * N, 16 * N
    assert bench.attempt_bytes(2, 6, [256] * 3) - bench.attempt_bytes(2, 6, [256, 256 * 256]) \
        == 6 * (2 * A + 2 * Fz)


def test_attempt_traffic_comes_from_the_counter_table():
    """The attempt's measured HBM bytes are the sum of the PMC table's scaled
    MB (the strided row six times), within 5% of the algorithmic bytes."""
    t = bench.measured_attempt_traffic()
    assert t is not None
    assert abs(t / 1e6 - 350.6) < 0.5
    assert 1.0 <= t / bench.attempt_bytes(2, 4, [512, 512]) < 1.05
