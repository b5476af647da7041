"""The autotuner: measured rates are positive and finite, a cache file of
the same device is read back exactly, and plans made with the measured
rates still give the oracle's results."""
import math
import os
import sys

import numpy as np
import pytest

import oracle
from tests.test_gpu_dense import planted

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_autotune_measures_and_caches(dim3, tmp_path):
    with dim3.Context(0) as ctx:
        assert ctx.tuning()["source"] == 0
        path = str(tmp_path / "tuning.txt")
        t = ctx.autotune(path)
        say("measured", t)
        assert t["source"] == 1
        for k in ("path_ns", "seed_bit_ns", "seed_word_ns", "mfma_tops", "bitor_gbs", "hbm_gbs"):
            assert t[k] > 0 and math.isfinite(t[k]), k
        assert os.path.exists(path)
    with dim3.Context(0) as ctx2:
        t2 = ctx2.autotune(path)
        assert t2["source"] == 2
        for k in ("path_ns", "mfma_tops", "hbm_gbs"):
            assert abs(t2[k] - t[k]) <= 1e-5 * abs(t[k]), k
        # plans made with the measured rates give the oracle's answer
        R = planted(31, 6000, 200, 0.35, 3000, 900)
        S = planted(32, 6000, 200, 0.35, 3000, 900)
        got = ctx2.join_project(R, S)
        want, _, _ = oracle.join_project(R, S)
        assert np.array_equal(oracle.sort_pairs(got.pairs), oracle.sort_pairs(want))


def test_set_tuning_round_trip_and_rejects_nonpositive(dim3):
    with dim3.Context(0) as ctx:
        d = ctx.tuning()
        d2 = dict(d, mfma_tops=1234.5, source=1)
        ctx.set_tuning(d2)
        assert ctx.tuning()["mfma_tops"] == 1234.5
        with pytest.raises(dim3.Dim3Error):
            ctx.set_tuning(dict(d2, path_ns=0.0))
