"""GPU parity of the value aggregates (SUM / MIN / MAX / AVG of vR + vS per
(x, z) group, int64): materialised rows and fingerprints against the
reference's rows (golden fixtures written by the compiled reference) and the
oracle, across z tiles, listed and swept tiles, hashed and dense domains,
the swap rule, the u != w filter and empty inputs."""
import glob
import os
import sys

import numpy as np
import pytest

import oracle
from tests.golden import cases
from tests.test_gpu_batch import rel

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
AGGS = ["sum", "min", "max", "avg"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def vals(rng, n, lo=-(1 << 31), hi=(1 << 31) - 1):
    return rng.integers(lo, hi, n, endpoint=True).astype(np.int32)


def check(ctx, R, vR, S, vS, agg, **kw):
    ex = kw.get("exclude_equal", False)
    wp, wv, wfp, sw = oracle.join_aggregate(R, vR, S, vS, agg, exclude_equal=ex)
    want = oracle.sort_pairs(wp, wv)
    g = ctx.join_aggregate(R, vR, S, vS, agg, **kw)
    got = oracle.sort_pairs(g.xz, g.counts)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), agg
    assert g.swapped == sw and g.n == wfp.count, agg
    f = ctx.join_aggregate(R, vR, S, vS, agg, output_mode=1, **kw)
    assert f.fp.as_tuple() == wfp.as_tuple() and f.n == wfp.count, agg
    return g


@pytest.mark.parametrize("path", sorted(set(glob.glob(os.path.join(GOLDEN, "values_*.npz")))
                                       - set(glob.glob(os.path.join(GOLDEN, "values_f32_*.npz")))),
                         ids=lambda p: os.path.basename(p))
def test_values_match_reference_golden(ctx, path):
    d = np.load(path)
    say("golden", os.path.basename(path), d["R"].shape, d["S"].shape)
    for agg in AGGS:
        g = ctx.join_aggregate(d["R"], d["vR"], d["S"], d["vS"], agg)
        gp, gv = oracle.sort_pairs(g.xz, g.counts)
        assert np.array_equal(gp, d["pairs"]) and np.array_equal(gv, d[agg]), agg


@pytest.mark.parametrize("agg", AGGS)
def test_values_zipf_tiles(ctx, agg):
    spec = cases.MEDIUM["aggregate_scaled"]
    R, S = cases.make_medium(spec)
    R = R[R[:, 0] % 8 == 0]  # a slice keeps the oracle quick
    rng = np.random.default_rng(11)
    vR, vS = vals(rng, R.shape[0]), vals(rng, S.shape[0])
    say("zipf", agg, R.shape, S.shape)
    g = check(ctx, R, vR, S, vS, agg)
    say("plan", g.stats.as_dict())
    # small tiles: many z tiles per row, touched lists that overflow
    check(ctx, R, vR, S, vS, agg, tile_bits=1024)
    check(ctx, R, vR, S, vS, agg, tile_bits=1024, block=1024)


@pytest.mark.parametrize("agg", AGGS)
def test_values_domains_and_edges(ctx, agg):
    rng = np.random.default_rng(12)
    # hashed x / z (values spread over the int32 range), swapped sides
    R = rel(rng, 6000, 10 ** 9, 300, amin=-10 ** 9)
    S = rel(rng, 9000, 10 ** 9, 300, amin=-10 ** 9)
    say("hashed", agg)
    check(ctx, R, vals(rng, R.shape[0]), S, vals(rng, S.shape[0]), agg)
    # a hub key: every row sweeps its tile (more paths than the list holds)
    R = np.stack([np.arange(3000), np.zeros(3000)], 1).astype(np.int32)
    S = np.stack([rng.integers(0, 20000, 12000), np.zeros(12000)], 1).astype(np.int32)
    say("hub", agg)
    check(ctx, R, vals(rng, 3000, -5, 5), S, vals(rng, 12000), agg)
    # self-join with the u != w filter
    E = rel(rng, 20_000, 3000, 3000)
    E = np.concatenate([E, E[:, ::-1]])
    vE = vals(rng, E.shape[0], -1000, 1000)
    say("self-join", agg)
    check(ctx, E, vE, E, vE, agg, exclude_equal=True)
    # empty inputs and no joined key
    e = np.zeros((0, 2), np.int32)
    ev = np.zeros(0, np.int32)
    g = ctx.join_aggregate(e, ev, S, vals(rng, S.shape[0]), agg)
    assert g.n == 0 and g.rows.shape[0] == 0
    g = ctx.join_aggregate(R, vals(rng, 3000), (S + [0, 7]).astype(np.int32), vals(rng, S.shape[0]), agg)
    assert g.n == 0


def test_values_against_count(ctx, dim3):
    """vR = 1, vS = 0 at the 10M x 10M COUNT workload's z slice (5.8e8
    groups): the SUM fingerprint is the reference's COUNT fingerprint, and
    AVG / MIN / MAX (all 1) give the join-project's (n, sum, sum2) with
    csum = n."""
    import json

    with open(os.path.join(GOLDEN, "large.json")) as f:
        rec = json.load(f)["aggregate_zslice"]
    R, S = cases.make_large_full(rec["spec"])
    R, S = cases.slice_large(rec["spec"], R, S)
    want = tuple(int(v) for v in rec["aggregate"]["fp"])
    ones, zeros = np.ones(R.shape[0], np.int32), np.zeros(S.shape[0], np.int32)
    say("count property", R.shape, S.shape)
    proj = ctx.join_project(R, S, output_mode=1).fp.as_tuple()
    for agg in AGGS:
        g = ctx.join_aggregate(R, ones, S, zeros, agg, output_mode=1)
        say(agg, g.fp.as_tuple(), "expand ms", g.stats.ms_expand)
        if agg == "sum":
            assert g.fp.as_tuple() == want
        else:
            assert g.fp.as_tuple() == proj[:3] + (proj[0],)


def test_values_dev_capacity(ctx, dim3, torch):
    rng = np.random.default_rng(13)
    R, S = rel(rng, 20_000, 2000, 500), rel(rng, 15_000, 2000, 500)
    vR, vS = vals(rng, R.shape[0]), vals(rng, S.shape[0])
    wp, wv, _, _ = oracle.join_aggregate(R, vR, S, vS, "avg")
    n = wp.shape[0]
    dR, dS = torch.from_numpy(R).cuda(), torch.from_numpy(S).cuda()
    dvR, dvS = torch.from_numpy(vR).cuda(), torch.from_numpy(vS).cuda()
    args = (dR.data_ptr(), dvR.data_ptr(), R.shape[0], dS.data_ptr(), dvS.data_ptr(), S.shape[0], "avg")
    rc, r0 = ctx.join_aggregate_dev(*args, 0, 0)
    assert rc == dim3.ERR_CAPACITY and r0.n == n
    out = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
    say("dev materialise", n)
    rc, r = ctx.join_aggregate_dev(*args, out.data_ptr(), n)
    assert rc == dim3.OK and r.n == n
    rows = np.ascontiguousarray(out.cpu().numpy()).view(dim3.COUNT_ROW_DTYPE).reshape(-1)
    gp, gv = oracle.sort_pairs(np.stack([rows["x"], rows["z"]], 1), rows["count"])
    want = oracle.sort_pairs(wp, wv)
    assert np.array_equal(gp, want[0]) and np.array_equal(gv, want[1])
    # a bad aggregate code is an error, not a launch
    with pytest.raises(dim3.Dim3Error):
        ctx.join_aggregate(R, vR, S, vS, 9)


# ---- float values ----

def check_f32(ctx, R, vR, S, vS, agg, **kw):
    ex = kw.get("exclude_equal", False)
    wp, wv, wfp, sw = oracle.join_aggregate_f32(R, vR, S, vS, agg, exclude_equal=ex)
    want = oracle.sort_pairs(wp, wv.view(np.int32).astype(np.int64))
    g = ctx.join_aggregate_f32(R, vR, S, vS, agg, **kw)
    got = oracle.sort_pairs(g.xz, g.counts.view(np.int32).astype(np.int64))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), agg  # bit for bit
    assert g.swapped == sw and g.n == wfp.count, agg
    f = ctx.join_aggregate_f32(R, vR, S, vS, agg, output_mode=1, **kw)
    assert f.fp.as_tuple() == wfp.as_tuple(), agg
    return g


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "values_f32_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_values_f32_match_reference_golden(ctx, path):
    d = np.load(path)
    say("golden f32", os.path.basename(path), d["R"].shape, d["S"].shape)
    for agg in AGGS:
        g = ctx.join_aggregate_f32(d["R"], d["vR"], d["S"], d["vS"], agg)
        gp, gb = oracle.sort_pairs(g.xz, g.counts.view(np.int32).astype(np.int64))
        assert np.array_equal(gp, d["pairs"]) and gb.astype(np.int32).tobytes() == d[agg].tobytes(), agg


@pytest.mark.parametrize("agg", AGGS)
def test_values_f32_zipf(ctx, agg):
    spec = cases.MEDIUM["aggregate_scaled"]
    R, S = cases.make_medium(spec)
    R = R[R[:, 0] % 8 == 0]
    rng = np.random.default_rng(14)
    # multiples of 1/8 below 2^20: float path values and double sums are exact
    vR = (rng.integers(-(1 << 22), 1 << 22, R.shape[0]) / 8).astype(np.float32)
    vS = (rng.integers(-(1 << 22), 1 << 22, S.shape[0]) / 8).astype(np.float32)
    say("f32 zipf", agg, R.shape, S.shape)
    check_f32(ctx, R, vR, S, vS, agg)
    check_f32(ctx, R, vR, S, vS, agg, tile_bits=1024)
    # the reference generator's values (tuple index) above 2^24: the float
    # path value rounds (one IEEE add on both sides), the double sums stay exact
    vR = (np.arange(R.shape[0]) + (1 << 24)).astype(np.float32)
    vS = (np.arange(S.shape[0]) * 3 + (1 << 24)).astype(np.float32)
    check_f32(ctx, R, vR, S, vS, agg)
    E = rel(rng, 20_000, 3000, 3000)
    E = np.concatenate([E, E[:, ::-1]])
    vE = rng.integers(-1000, 1000, E.shape[0]).astype(np.float32)
    check_f32(ctx, E, vE, E, vE, agg, exclude_equal=True)


def test_values_f32_dev(ctx, dim3, torch):
    rng = np.random.default_rng(15)
    R, S = rel(rng, 20_000, 2000, 500), rel(rng, 15_000, 2000, 500)
    vR = rng.integers(0, 10 ** 6, R.shape[0]).astype(np.float32)
    vS = rng.integers(0, 10 ** 6, S.shape[0]).astype(np.float32)
    wp, wv, _, _ = oracle.join_aggregate_f32(R, vR, S, vS, "avg")
    n = wp.shape[0]
    dR, dS = torch.from_numpy(R).cuda(), torch.from_numpy(S).cuda()
    dvR, dvS = torch.from_numpy(vR).cuda(), torch.from_numpy(vS).cuda()
    args = (dR.data_ptr(), dvR.data_ptr(), R.shape[0], dS.data_ptr(), dvS.data_ptr(), S.shape[0], "avg")
    rc, r0 = ctx.join_aggregate_f32_dev(*args, 0, 0)
    assert rc == dim3.ERR_CAPACITY and r0.n == n
    out = torch.zeros((n, 3), dtype=torch.int32, device="cuda")
    say("f32 dev materialise", n)
    rc, r = ctx.join_aggregate_f32_dev(*args, out.data_ptr(), n)
    assert rc == dim3.OK and r.n == n
    rows = np.ascontiguousarray(out.cpu().numpy()).view(dim3.F32_ROW_DTYPE).reshape(-1)
    gp, gb = oracle.sort_pairs(np.stack([rows["x"], rows["z"]], 1), rows["v"].view(np.int32).astype(np.int64))
    want = oracle.sort_pairs(wp, wv.view(np.int32).astype(np.int64))
    assert np.array_equal(gp, want[0]) and np.array_equal(gb, want[1])


# ---- TUPLE-array ingestion ----

def tuples(dt, pairs, v):
    t = np.zeros(pairs.shape[0], dtype=dt)
    t["x"], t["y"], t["v"] = pairs[:, 0], pairs[:, 1], v
    return t


def test_values_from_tuple_arrays(ctx, dim3):
    rng = np.random.default_rng(16)
    R, S = rel(rng, 30_000, 3000, 800), rel(rng, 20_000, 3000, 800)
    # TUPLE<int,int,float>, AVG (join_aggregate.cpp's layout)
    vR = rng.integers(0, 10 ** 6, R.shape[0]).astype(np.float32)
    vS = rng.integers(0, 10 ** 6, S.shape[0]).astype(np.float32)
    g = ctx.join_aggregate_tuples(tuples(dim3.TUPLE_F32, R, vR), tuples(dim3.TUPLE_F32, S, vS), "avg")
    wp, wv, _, _ = oracle.join_aggregate_f32(R, vR, S, vS, "avg")
    got = oracle.sort_pairs(g.xz, g.counts.view(np.int32).astype(np.int64))
    want = oracle.sort_pairs(wp, wv.view(np.int32).astype(np.int64))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    # TUPLE<int,int,long long> (16 bytes, padded), SUM and COUNT
    iR = rng.integers(-(1 << 31), 1 << 31, R.shape[0]).astype(np.int64)
    iS = rng.integers(-(1 << 31), 1 << 31, S.shape[0]).astype(np.int64)
    tR, tS = tuples(dim3.TUPLE_I64, R, iR), tuples(dim3.TUPLE_I64, S, iS)
    g = ctx.join_aggregate_tuples(tR, tS, "sum")
    wp, wv, _, _ = oracle.join_aggregate(R, iR.astype(np.int32), S, iS.astype(np.int32), "sum")
    got, want = oracle.sort_pairs(g.xz, g.counts), oracle.sort_pairs(wp, wv)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    g = ctx.join_aggregate_tuples(tR, tS, "count")
    wp, wc, _, _ = oracle.join_aggregate_count(R, S)
    got, want = oracle.sort_pairs(g.xz, g.counts), oracle.sort_pairs(wp, wc)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    # a 64-bit value outside int32 is an error, not a wrong answer
    tR["v"][7] = 1 << 40
    with pytest.raises(dim3.Dim3Error):
        ctx.join_aggregate_tuples(tR, tS, "max")
