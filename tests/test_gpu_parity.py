"""GPU parity: libdim3 (HIP, through the C-ABI) against the reference's
golden fixtures and the CPU oracle, bit-exact.  Project results are compared
as sorted (x, z) sets, COUNT results as sorted (x, z, count) rows, large
results by count + order-independent fingerprint."""
import glob
import json
import os
import sys

import numpy as np
import pytest

import oracle
from tests.golden import cases

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_rel(rng, n, amax, bmax, amin=0, bmin=0):
    return np.stack([rng.integers(amin, amax, n), rng.integers(bmin, bmax, n)], 1).astype(np.int32)


def check_project(ctx, R, S, **kw):
    say("project", R.shape[0], S.shape[0], kw)
    got = ctx.join_project(R, S, **kw)
    want, fp, sw = oracle.join_project(R, S, exclude_equal=kw.get("exclude_equal", False))
    if kw.get("no_swap"):
        # the oracle always applies the swap rule; with NO_SWAP compare in R,S orientation
        if R.shape[0] < S.shape[0]:
            want, fp, sw = oracle.join_project(S, R, exclude_equal=kw.get("exclude_equal", False))
            want = want[:, ::-1]
            fp = oracle.fingerprint(want)
        sw = False
    assert got.swapped == sw
    assert got.n == want.shape[0]
    assert np.array_equal(oracle.sort_pairs(got.pairs), oracle.sort_pairs(want))
    return got


def check_count(ctx, R, S, **kw):
    say("count", R.shape[0], S.shape[0], kw)
    got = ctx.join_aggregate_count(R, S, **kw)
    wp, wc, wfp, sw = oracle.join_aggregate_count(R, S, exclude_equal=kw.get("exclude_equal", False))
    assert got.swapped == sw
    gp, gc = oracle.sort_pairs(got.xz, got.counts)
    ep, ec = oracle.sort_pairs(wp, wc)
    assert np.array_equal(gp, ep) and np.array_equal(gc, ec)
    return got


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "small_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_golden_small(ctx, path):
    d = np.load(path)
    R, S = d["R"], d["S"]
    say("golden", os.path.basename(path))
    got = ctx.join_project(R, S)
    assert got.swapped == bool(d["swapped"][0])
    assert np.array_equal(oracle.sort_pairs(got.pairs), d["project"])
    agg = ctx.join_aggregate_count(R, S)
    gp, gc = oracle.sort_pairs(agg.xz, agg.counts)
    assert np.array_equal(gp, d["agg_pairs"]) and np.array_equal(gc, d["agg_counts"])


@pytest.mark.parametrize("seed", range(12))
def test_random_small_vs_oracle(ctx, seed):
    rng = np.random.default_rng(seed)
    nR, nS = int(rng.integers(1, 5000)), int(rng.integers(1, 5000))
    amax = int(rng.integers(2, 3000))
    bmax = int(rng.integers(1, 500))
    R = rand_rel(rng, nR, amax, bmax, -int(rng.integers(0, 3)), -int(rng.integers(0, 3)))
    S = rand_rel(rng, nS, amax, bmax, -int(rng.integers(0, 3)), -int(rng.integers(0, 3)))
    kw = {"exclude_equal": bool(seed % 3 == 0), "block": [0, 256, 512, 1024][seed % 4],
          "hash_only": bool(seed % 2)}
    check_project(ctx, R, S, **kw)
    check_count(ctx, R, S, **kw)


@pytest.mark.parametrize("seed", range(4))
def test_sparse_value_domains_use_hashing(ctx, seed):
    # values drawn from a few thousand random int32s: ranges far wider than the
    # tuple counts, so every attribute goes through the hash tables
    rng = np.random.default_rng(50 + seed)
    dom = rng.integers(-2**31, 2**31 - 1, 3000).astype(np.int32)
    dom[0], dom[1] = -1, -2**31
    R = dom[rng.integers(0, 3000, (6000, 2))]
    S = dom[rng.integers(0, 3000, (5000, 2))]
    R[:, 1] = dom[rng.integers(0, 200, 6000)]
    S[:, 1] = dom[rng.integers(0, 200, 5000)]
    check_project(ctx, R, S, exclude_equal=bool(seed % 2))
    check_count(ctx, R, S)


def test_hash_and_range_mapping_agree(ctx):
    rng = np.random.default_rng(77)
    R = rand_rel(rng, 40000, 3000, 800, -100, -5)
    S = rand_rel(rng, 30000, 3000, 800, -100, -5)
    a = ctx.join_project(R, S, output_mode=1).fp.as_tuple()
    b = ctx.join_project(R, S, output_mode=1, hash_only=True).fp.as_tuple()
    assert a == b == oracle.join_project(R, S, fingerprint_only=True)[1].as_tuple()
    a = ctx.join_aggregate_count(R, S, output_mode=1).fp.as_tuple()
    b = ctx.join_aggregate_count(R, S, output_mode=1, hash_only=True).fp.as_tuple()
    assert a == b == oracle.join_aggregate_count(R, S, fingerprint_only=True)[2].as_tuple()


@pytest.mark.parametrize("tile_bits", [32, 64, 1000, 4096])
def test_z_tiling(ctx, tile_bits):
    rng = np.random.default_rng(tile_bits)
    R = rand_rel(rng, 20000, 500, 300)
    S = rand_rel(rng, 20000, 20000, 300)
    for h in (False, True):
        check_project(ctx, R, S, tile_bits=tile_bits, hash_only=h)
        check_count(ctx, R, S, tile_bits=tile_bits, hash_only=h)
        check_project(ctx, R, S, tile_bits=tile_bits, exclude_equal=True, hash_only=h)


def test_edge_cases(ctx):
    e = np.zeros((0, 2), np.int32)
    R = np.array([[1, 2], [3, 4]], np.int32)
    for a, b in ((e, R), (R, e), (e, e)):
        assert ctx.join_project(a, b).n == 0
        assert ctx.join_aggregate_count(a, b).n == 0
    assert ctx.join_project(R, np.array([[9, 99]], np.int32)).n == 0
    check_project(ctx, np.array([[5, 1]], np.int32), np.array([[6, 1]], np.int32))
    # key -1 and extreme values
    Rm = np.array([[1, -1], [2, -1], [3, 0], [2**31 - 1, -2**31], [-2**31, 2**31 - 1]], np.int32)
    Sm = np.array([[7, -1], [8, 0], [-1, -2**31], [9, 2**31 - 1]], np.int32)
    check_project(ctx, Rm, Sm)
    check_count(ctx, Rm, Sm)
    # |R| < |S| swaps; NO_SWAP keeps (R.x, S.z)
    rng = np.random.default_rng(9)
    Rs, Ss = rand_rel(rng, 100, 50, 10), rand_rel(rng, 400, 50, 10)
    assert check_project(ctx, Rs, Ss).swapped
    assert not check_project(ctx, Rs, Ss, no_swap=True).swapped
    # self pairs dropped for a self-join
    E = rand_rel(rng, 3000, 400, 400)
    g = check_project(ctx, E, E, exclude_equal=True)
    assert not np.any(g.pairs[:, 0] == g.pairs[:, 1])


def test_wide_rows_and_overflowing_touch_list(ctx):
    # rows with far more outputs than the LDS stage holds (several passes)
    R = np.stack([np.repeat(np.arange(10), 3), np.tile(np.arange(3), 10)], 1).astype(np.int32)
    S = np.stack([np.arange(15000), np.arange(15000) % 3], 1).astype(np.int32)
    check_project(ctx, R, S)
    check_count(ctx, R, S)
    # a sparse row that touches more words than the first-touch list holds
    R2 = np.stack([np.zeros(6000), np.arange(6000)], 1).astype(np.int32)
    S2 = np.stack([(np.arange(6000) * 97) % 400000, np.arange(6000)], 1).astype(np.int32)
    S2 = np.concatenate([S2, np.stack([np.arange(400000), np.full(400000, 10**6)], 1).astype(np.int32)])
    R2 = np.concatenate([R2, np.array([[1, 10**6]], np.int32)])
    check_project(ctx, R2, S2)
    check_count(ctx, R2, S2)


@pytest.mark.parametrize("tile_bits", [0, 50000])
def test_dense_z_cluster_in_a_swept_row(ctx, tile_bits):
    """Every row reaches all 2048 z IDs of a hub cluster (32 or 64 fully set
    tile words, more rows per 64 words than a wavefront's stage holds) plus a
    sparse tail, in rows with too many paths for the first-touch list: the
    swept emission's dense-chunk rounds.  tile_bits 0: blocked bit layout,
    50000 (not a multiple of 1024): plain layout over two tiles."""
    rng = np.random.default_rng(77)
    ys = np.arange(40)
    S = np.concatenate([np.stack([rng.permutation(2048), np.full(2048, y)], 1) for y in ys] +
                       [np.stack([rng.integers(2048, 100000, 300), np.full(300, y)], 1) for y in ys]).astype(np.int32)
    R = np.stack([np.repeat(np.arange(50), 40), np.tile(ys, 50)], 1).astype(np.int32)
    got = check_project(ctx, R, S, no_swap=True, tile_bits=tile_bits)
    say("  rows", got.n, "tiles", got.stats.n_tiles, "tile bits", got.stats.tile_bits)
    assert got.n > 50 * 2048


def test_duplicates_count_multiplicity(ctx):
    rng = np.random.default_rng(4)
    R = rand_rel(rng, 30000, 50, 40)
    S = rand_rel(rng, 20000, 60, 40)
    g = check_count(ctx, R, S)
    assert int(g.counts.sum()) == oracle.outj(R, S)
    p = check_project(ctx, R, S)
    assert p.n == g.n


def _medium():
    with open(os.path.join(GOLDEN, "medium.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["kat", "project_scaled", "graph_scaled", "dense_block", "project_slice",
                                  "aggregate_scaled"])
def test_medium_fingerprints_vs_reference(ctx, name):
    rec = _medium()[name]
    say("medium", name)
    R, S = cases.make_medium(rec["spec"])
    from tests.golden.make_golden import sha

    assert sha(R, S) == rec["sha256"], "generator drift"
    if "project" in rec:
        g = ctx.join_project(R, S, output_mode=1)
        assert [str(v) for v in g.fp.as_tuple()] == rec["project"]["fp"]
        if rec["project"]["n"] < 150_000_000:
            m = ctx.join_project(R, S)
            assert [str(v) for v in oracle.fingerprint(m.pairs).as_tuple()] == rec["project"]["fp"]
    if "aggregate" in rec:
        g = ctx.join_aggregate_count(R, S, output_mode=1)
        assert [str(v) for v in g.fp.as_tuple()] == rec["aggregate"]["fp"]
        m = ctx.join_aggregate_count(R, S)
        assert [str(v) for v in oracle.fingerprint(m.xz, m.counts).as_tuple()] == rec["aggregate"]["fp"]


def test_device_api_capacity_and_stream(ctx, dim3):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    R = rand_rel(rng, 50000, 5000, 2000)
    S = rand_rel(rng, 40000, 5000, 2000)
    want, _, _ = oracle.join_project(R, S)
    dR = torch.from_numpy(R).cuda()
    dS = torch.from_numpy(S).cuda()
    small = torch.empty((1000, 2), dtype=torch.int32, device="cuda")
    rc, res = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], small.data_ptr(), 1000)
    assert rc == dim3.ERR_CAPACITY and res.n == want.shape[0]
    out = torch.empty((res.n, 2), dtype=torch.int32, device="cuda")
    st = torch.cuda.Stream()
    rc, res2 = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], out.data_ptr(), res.n,
                                    stream=st.cuda_stream)
    assert rc == dim3.OK and res2.n == res.n
    assert np.array_equal(oracle.sort_pairs(out.cpu().numpy()), oracle.sort_pairs(want))
    # COUNT rows on the device
    wp, wc, _, _ = oracle.join_aggregate_count(R, S)
    rows = torch.empty((wp.shape[0] * 16,), dtype=torch.uint8, device="cuda")
    rc, r3 = ctx.join_aggregate_count_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], rows.data_ptr(),
                                          wp.shape[0])
    assert rc == dim3.OK
    arr = np.frombuffer(rows.cpu().numpy().tobytes(), dtype=dim3.COUNT_ROW_DTYPE)
    gp, gc = oracle.sort_pairs(np.stack([arr["x"], arr["z"]], 1), arr["count"])
    ep, ec = oracle.sort_pairs(wp, wc)
    assert np.array_equal(gp, ep) and np.array_equal(gc, ec)


def test_fingerprint_mode_equals_materialized(ctx):
    rng = np.random.default_rng(12)
    R = rand_rel(rng, 200000, 20000, 3000)
    S = rand_rel(rng, 150000, 20000, 3000)
    for ex in (False, True):
        m = ctx.join_project(R, S, exclude_equal=ex)
        f = ctx.join_project(R, S, exclude_equal=ex, output_mode=1)
        assert f.fp.as_tuple() == oracle.fingerprint(m.pairs).as_tuple()
        ma = ctx.join_aggregate_count(R, S, exclude_equal=ex)
        fa = ctx.join_aggregate_count(R, S, exclude_equal=ex, output_mode=1)
        assert fa.fp.as_tuple() == oracle.fingerprint(ma.xz, ma.counts).as_tuple()


def test_count_join_paths(ctx):
    rng = np.random.default_rng(13)
    R = rand_rel(rng, 100000, 1000, 5000)
    S = rand_rel(rng, 70000, 1000, 5000)
    assert ctx.count_join_paths(R, S) == oracle.outj(R, S)


def test_repeated_queries_reuse_workspace(ctx):
    rng = np.random.default_rng(14)
    R = rand_rel(rng, 100000, 4000, 1000)
    S = rand_rel(rng, 100000, 4000, 1000)
    a = ctx.join_project(R, S, output_mode=1).fp.as_tuple()
    held = ctx.device_bytes()
    for _ in range(3):
        assert ctx.join_project(R, S, output_mode=1).fp.as_tuple() == a
    assert ctx.device_bytes() == held


def test_low_outj_and_short_rows(ctx, dim3):
    """Join sizes near n (the reference's f1 picks its classical hash join
    here, dim3.hpp:41-50, 273-391) and rows of a few paths each, which go
    one per wavefront: rows and counts equal the oracle's, in materialised
    and fingerprint modes, with and without the u != w filter."""
    from dim3_amd import workloads

    R, _, S, _ = workloads.gen_rand_tuples(1_000_000, 800_000, 100_000, 1_000_000, 100_000, 11)
    want, wfp, sw = oracle.join_project(R, S)
    g = ctx.join_project(R, S)
    assert g.swapped == sw and np.array_equal(oracle.sort_pairs(g.pairs), oracle.sort_pairs(want))
    assert ctx.join_project(R, S, output_mode=1).fp.as_tuple() == wfp.as_tuple()
    wp, wc, cfp, _ = oracle.join_aggregate_count(R, S)
    a = ctx.join_aggregate_count(R, S)
    gp, gc = oracle.sort_pairs(a.xz, a.counts)
    ep, ec = oracle.sort_pairs(wp, wc)
    assert np.array_equal(gp, ep) and np.array_equal(gc, ec)
    assert ctx.join_aggregate_count(R, S, output_mode=1).fp.as_tuple() == cfp.as_tuple()
    # a self-join of short rows with the filter
    E = np.concatenate([R[:200_000], R[:200_000, ::-1]]).astype(np.int32)
    for count in (False, True):
        if count:
            wp, wc, cfp, _ = oracle.join_aggregate_count(E, E, exclude_equal=True)
            f = ctx.join_aggregate_count(E, E, exclude_equal=True, output_mode=1)
            assert f.fp.as_tuple() == cfp.as_tuple()
        else:
            want, wfp, _ = oracle.join_project(E, E, exclude_equal=True)
            g = ctx.join_project(E, E, exclude_equal=True)
            assert np.array_equal(oracle.sort_pairs(g.pairs), oracle.sort_pairs(want))


# ---- the classical path (f1, dim3.hpp:41-50 / 273-391) ----

def classical_cases():
    rng = np.random.default_rng(2024)
    yield "uniform_low_outj", rand_rel(rng, 20000, 20000, 20000), rand_rel(rng, 15000, 20000, 20000)
    yield "duplicates", np.repeat(rand_rel(rng, 3000, 500, 800), 3, 0), rand_rel(rng, 4000, 500, 800)
    yield "negative_and_extremes", np.array([[-5, 1], [2**31 - 1, 1], [-2**31, 2], [0, 2], [7, -1]], np.int32), \
        np.array([[2**31 - 1, 1], [-2**31, 1], [3, 2], [-1, -1], [9, 5]], np.int32)
    yield "skewed_key", np.stack([np.arange(5000), np.zeros(5000)], 1).astype(np.int32), \
        np.stack([np.arange(3), np.zeros(3)], 1).astype(np.int32)
    E = rand_rel(rng, 6000, 3000, 3000)
    yield "self_join", E, E


@pytest.mark.parametrize("case", [c[0] for c in classical_cases()])
def test_classical_path_vs_oracle(ctx, case):
    R, S = next((r, s) for n, r, s in classical_cases() if n == case)
    for ex in (False, True):
        got = check_project(ctx, R, S, classical=True, exclude_equal=ex)
        assert got.stats.n_tiles == 0  # the classical path ran
        check_count(ctx, R, S, classical=True, exclude_equal=ex)
        # fingerprints equal the general plan's
        a = ctx.join_project(R, S, classical=True, exclude_equal=ex, output_mode=1)
        b = ctx.join_project(R, S, classical=False, exclude_equal=ex, output_mode=1)
        assert a.fp.as_tuple() == b.fp.as_tuple() and b.stats.n_tiles > 0
        a = ctx.join_aggregate_count(R, S, classical=True, exclude_equal=ex, output_mode=1)
        b = ctx.join_aggregate_count(R, S, classical=False, exclude_equal=ex, output_mode=1)
        assert a.fp.as_tuple() == b.fp.as_tuple()
    check_project(ctx, R, S, classical=True, no_swap=True)


def test_classical_path_choice_and_device_api(ctx, dim3, torch):
    rng = np.random.default_rng(9)
    big = rand_rel(rng, 5_000_000, 5_000_000, 5_000_000)
    big2 = big[::-1].copy()
    auto = ctx.join_project(big, big2, output_mode=1)
    say("low OUT_J auto (1e7 tuples): tiles", auto.stats.n_tiles, "paths", auto.stats.out_j)
    assert auto.stats.n_tiles == 0  # f1 picks the classical path
    gen = ctx.join_project(big, big2, output_mode=1, classical=False)
    assert gen.stats.n_tiles > 0 and gen.fp.as_tuple() == auto.fp.as_tuple()
    dense = rand_rel(rng, 4_100_000, 4_100_000, 40_000)  # about 100 paths per tuple
    assert ctx.join_project(dense, dense, output_mode=1).stats.n_tiles > 0
    R, S = rand_rel(rng, 200000, 200000, 200000), rand_rel(rng, 150000, 200000, 200000)
    small = ctx.join_project(R, S, output_mode=1)
    assert small.stats.n_tiles > 0  # small: the general plan
    # device API: capacity shortfall reports the row count, then fills
    dR, dS = torch.from_numpy(R).cuda(), torch.from_numpy(S).cuda()
    rc, r = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, classical=True)
    assert rc == dim3.ERR_CAPACITY and r.n == small.n
    out = torch.empty((r.n, 2), dtype=torch.int32, device="cuda")
    rc, r2 = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], out.data_ptr(), r.n,
                                  classical=True)
    assert rc == dim3.OK and r2.n == r.n
    want, _, _ = oracle.join_project(R, S)
    assert np.array_equal(oracle.sort_pairs(out.cpu().numpy()), oracle.sort_pairs(want))


def _zipf_values(rng, n, domain, s, lo=0):
    """n values of [lo, lo + domain) whose ranks (over a shuffled domain) are
    Zipf(s): a few values take percents of the tuples."""
    p = np.arange(1, domain + 1, dtype=np.float64) ** -s
    p /= p.sum()
    return (rng.permutation(domain)[rng.choice(domain, size=n, p=p)] + lo).astype(np.int32)


def _count_rows_no_swap(R, S):
    """The oracle's COUNT rows of R x S in (R.x, S.z) orientation."""
    if R.shape[0] < S.shape[0]:
        wp, wc, _, _ = oracle.join_aggregate_count(S, R)
        return oracle.sort_pairs(np.ascontiguousarray(wp[:, ::-1]), wc)
    wp, wc, _, _ = oracle.join_aggregate_count(R, S)
    return oracle.sort_pairs(wp, wc)


@pytest.mark.parametrize("hash_only", [False, True])
def test_hot_value_tallies_vs_oracle(ctx, hash_only):
    """Relations of at least 2^20 tuples take samples of their most frequent
    key and x / z values, whose tuples the y and x mapping, S count,
    heaviness and heavy-degree passes count per workgroup in LDS.  A large
    side with hub keys and hub x (as R) or hub z (as S) values joined with a
    small side, in both orientations (no swap), with range and with hashed
    slots, with and without a forced dense core: rows and COUNT rows against
    the oracle."""
    rng = np.random.default_rng(2026)
    big, small = 1_300_000, 60_000
    B = np.stack([_zipf_values(rng, big, 200_000, 1.1, lo=-50_000), _zipf_values(rng, big, 20_000, 1.0)], 1)
    T = np.stack([_zipf_values(rng, small, 100_000, 1.1), rng.integers(0, 20_000, small).astype(np.int32)], 1)
    for R, S in ((B, T), (T, B)):
        for core in (-1, 1):
            kw = dict(no_swap=True, hash_only=hash_only, dense_core=core, classical=False)
            check_project(ctx, R, S, **kw)
            got = ctx.join_aggregate_count(R, S, **kw)
            gp, gc = oracle.sort_pairs(got.xz, got.counts)
            ep, ec = _count_rows_no_swap(R, S)
            say("count rows", gp.shape[0], ep.shape[0])
            assert np.array_equal(gp, ep) and np.array_equal(gc, ec)
