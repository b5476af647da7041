"""GPU EstimateJoinCardinality (dim3.hpp:190-270): the device estimate is
the oracle's (itself pinned to the compiled reference) on sampled inputs,
the exact |S| <= 10000 case, the 1000-tuple-prefix fallback and the swap."""
import sys

import numpy as np
import pytest

import oracle
from tests.golden import cases
from tests.test_gpu_batch import rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def want(R, S):
    return oracle.estimate(R, S) if R.shape[0] >= S.shape[0] else oracle.estimate(S, R)


def test_estimate_matches_oracle(ctx):
    rng = np.random.default_rng(31)
    R, S = oracle.gen_rand_data(300_000, 200_000, 50_000, 20_000, 50_000, 77)
    say("gen_rand_data 300k x 200k")
    assert ctx.estimate_join_cardinality(R, S) == want(R, S)
    # the swap rule: the smaller side samples
    assert ctx.estimate_join_cardinality(S, R) == want(R, S)
    # |S| <= 10000: every key sampled, the estimate is OUT_J
    R3, S3 = oracle.gen_rand_data(20_000, 8_000, 3_000, 500, 3_000, 5)
    assert ctx.estimate_join_cardinality(R3, S3) == oracle.outj(R3, S3) == want(R3, S3)
    # Zipf keys (the sample holds the hot key 0) and negative keys
    spec = cases.MEDIUM["aggregate_scaled"]
    Rz, Sz = cases.make_medium(spec)
    assert ctx.estimate_join_cardinality(Rz, Sz) == want(Rz, Sz)
    Rn = rel(rng, 100_000, 1000, 5000, amin=-1000) - [0, 2500]
    Sn = rel(rng, 60_000, 1000, 5000, amin=-1000) - [0, 2500]
    assert ctx.estimate_join_cardinality(Rn.astype(np.int32), Sn.astype(np.int32)) == want(Rn, Sn)


def test_estimate_fallback_and_empty(ctx):
    # sampled keys (low 7 bits zero) never joined, prefix keys do: the fallback
    n = 1_000_000
    R = np.stack([np.arange(n), (np.arange(n) % 1000) * 128 + 1], 1).astype(np.int32)
    S = np.stack([np.arange(n), (np.arange(n) % 500) * 128 + 1], 1).astype(np.int32)
    w = want(R, S)
    say("fallback", w)
    assert w > 0 and ctx.estimate_join_cardinality(R, S) == w
    # no join at all
    S2 = (S + [0, 1]).astype(np.int32)
    assert ctx.estimate_join_cardinality(R, S2) == want(R, S2)
    # fewer than 1000 tuples: the prefixes are clamped
    assert ctx.estimate_join_cardinality(R[:300], (S[:200] + [0, 1]).astype(np.int32)) == 0
    assert ctx.estimate_join_cardinality(np.zeros((0, 2), np.int32), S) == 0


def test_estimate_dev(ctx):
    import torch
    R, S = oracle.gen_rand_data(300_000, 200_000, 50_000, 20_000, 50_000, 78)
    dR = torch.from_numpy(R).cuda()
    dS = torch.from_numpy(S).cuda()
    torch.cuda.synchronize()
    say("device estimate 300k x 200k")
    w = want(R, S)
    assert ctx.estimate_join_cardinality_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0]) == w
    assert ctx.estimate_join_cardinality_dev(dS.data_ptr(), S.shape[0], dR.data_ptr(), R.shape[0]) == w
    assert ctx.estimate_join_cardinality_dev(dR.data_ptr(), R.shape[0], 0, 0) == 0
