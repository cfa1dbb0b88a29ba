"""The committed full-horizon fixture of the 2-D CGLE 512x512 workload
(tests/golden/horizon_c2.npz) is the oracle's own run: its first time unit
is re-run live on the oracle here and must match attempt for attempt and
frame for frame (the GPU test compares the whole horizon with the fixture)."""
import os
import sys

import numpy as np

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import horizon_c2  # noqa: E402


def test_fixture_prefix_is_the_oracle_run(olib):
    fx = np.load(horizon_c2.FIXTURE)
    t, h, acc, frames, _ = horizon_c2.run_horizon(olib, 1.0)
    k = t.size
    assert 100 < k < fx["t"].size and t[-1] >= 1.0
    assert np.array_equal(t, fx["t"][:k]) and np.array_equal(h, fx["h"][:k])
    assert np.array_equal(acc, fx["acc"][:k])
    assert [f[0] for f in frames] == list(fx["y_sha"][:2])
    assert [f[1] for f in frames] == list(fx["f_sha"][:2])
