"""Pins the CPU oracle: against the reference's golden fixtures, the
gen_rand_data known answer, brute force on tiny inputs, closed forms and,
where it is built, the compiled reference itself."""
import glob
import json
import os
from collections import defaultdict

import numpy as np
import pytest

import oracle
from tests.golden import cases

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def brute_project(R, S, exclude_equal=False):
    byy = defaultdict(list)
    for z, y in S:
        byy[int(y)].append(int(z))
    out = set()
    for x, y in R:
        for z in byy.get(int(y), ()):
            if exclude_equal and int(x) == z:
                continue
            out.add((int(x), z))
    return out


def brute_count(R, S):
    byy = defaultdict(list)
    for z, y in S:
        byy[int(y)].append(int(z))
    out = defaultdict(int)
    for x, y in R:
        for z in byy.get(int(y), ()):
            out[(int(x), z)] += 1
    return dict(out)


def as_set(p):
    return set(map(tuple, np.asarray(p).tolist()))


def rand_rel(rng, n, amax, bmax, amin=0, bmin=0):
    return np.stack([rng.integers(amin, amax, n), rng.integers(bmin, bmax, n)], 1).astype(np.int32)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_project_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    nR, nS = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    R = rand_rel(rng, nR, 40, 25, -3, -2)
    S = rand_rel(rng, nS, 40, 25, -3, -2)
    pairs, fp, sw = oracle.join_project(R, S, exclude_equal=bool(seed % 2))
    big, small = (R, S) if nR >= nS else (S, R)
    assert sw == (nR < nS)
    want = brute_project(big, small, exclude_equal=bool(seed % 2))
    assert as_set(pairs) == want
    assert len(pairs) == len(want)  # distinct
    assert fp.count == len(want)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_count_matches_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    R = rand_rel(rng, int(rng.integers(1, 400)), 30, 20, -2, -1)
    S = rand_rel(rng, int(rng.integers(1, 400)), 30, 20, -2, -1)
    p, c, fp, sw = oracle.join_aggregate_count(R, S)
    big, small = (R, S) if R.shape[0] >= S.shape[0] else (S, R)
    want = brute_count(big, small)
    got = {(int(a), int(b)): int(k) for (a, b), k in zip(p.tolist(), c.tolist())}
    assert got == want
    # closed forms: sum of counts = OUT_J, groups = |join-project|
    assert int(c.sum()) == oracle.outj(big, small) == fp.csum
    assert len(got) == len(oracle.join_project(R, S)[0])


def test_oracle_special_cases():
    e = np.zeros((0, 2), np.int32)
    R = np.array([[1, 2], [3, 4]], np.int32)
    assert oracle.join_project(e, R)[0].shape[0] == 0
    assert oracle.join_project(R, e)[0].shape[0] == 0
    # no matching key
    assert oracle.join_project(R, np.array([[9, 99]], np.int32))[0].shape[0] == 0
    # single shared key: cross product of the touched x and z
    Rx = np.stack([np.arange(50), np.full(50, 5)], 1).astype(np.int32)
    Sx = np.stack([np.arange(30) + 100, np.full(30, 5)], 1).astype(np.int32)
    p = oracle.join_project(Rx, Sx)[0]
    assert as_set(p) == {(a, b) for a in range(50) for b in range(100, 130)}
    # the key -1 is an ordinary key (LS_mapper empty marker is internal)
    Rm = np.array([[1, -1], [2, -1], [3, 0]], np.int32)
    Sm = np.array([[7, -1], [8, 0]], np.int32)
    assert as_set(oracle.join_project(Rm, Sm)[0]) == {(1, 7), (2, 7), (3, 8)}


def test_fingerprint_is_order_independent_and_sensitive():
    rng = np.random.default_rng(3)
    p = rand_rel(rng, 1000, 1 << 30, 1 << 30)
    a = oracle.fingerprint(p).as_tuple()
    b = oracle.fingerprint(p[rng.permutation(1000)]).as_tuple()
    assert a == b
    q = p.copy()
    q[17, 1] += 1
    assert oracle.fingerprint(q).as_tuple()[1:3] != a[1:3]
    c = np.ones(1000, np.int64)
    c2 = c.copy()
    c2[5] = 2
    assert oracle.fingerprint(p, c).as_tuple() != oracle.fingerprint(p, c2).as_tuple()


def test_fingerprint_factorises_over_a_cross_product():
    """Every term is k(x) g(z): a full cross product X x Z therefore has
    sum = (sum_x k1(x)) (sum_z g(z)) mod 2^64, and disjoint sets add."""
    M = (1 << 64) - 1
    X = np.array([-7, 0, 3, 1 << 30, 99], np.int64)
    Z = np.array([5, -1, 2 ** 31 - 1, 8], np.int64)
    p = np.array([(x, z) for x in X for z in Z], np.int32)
    fp = oracle.fingerprint(p)
    u = lambda v: int(v) & 0xFFFFFFFF
    k1 = sum(oracle.mix64(u(x) ^ 0x243F6A8885A308D3) | 1 for x in X) & M
    # the z factors of the row set, read off single-row fingerprints with x = 0
    k0 = oracle.mix64(0x243F6A8885A308D3) | 1
    inv = pow(k0, -1, 1 << 64)
    g1 = sum((oracle.fingerprint(np.array([[0, z]], np.int32)).sum * inv) & M for z in Z) & M
    assert all((oracle.fingerprint(np.array([[0, z]], np.int32)).sum * inv) & M == oracle.zfactor(z) for z in Z)
    assert fp.sum == (k1 * g1) & M and fp.count == len(p)
    a, b = oracle.fingerprint(p[:7]), oracle.fingerprint(p[7:])
    assert fp.sum == (a.sum + b.sum) & M and fp.sum2 == (a.sum2 + b.sum2) & M
    # COUNT: a row with count c weighs its second term by c
    c1, c2 = oracle.fingerprint(p[:1], np.array([1])), oracle.fingerprint(p[:1], np.array([3]))
    assert c2.sum == c1.sum and c2.sum2 == (3 * c1.sum2) & M and c2.csum == 3


def test_zfactor_word_structure():
    """g(z) = a(u(z) >> 5) c(u(z) & 31): within one 32-aligned word of z
    values g(z) / c(b) is one odd constant, c(b) being the splitmix64
    outputs for seed 0 (forced odd), so the z factors of a word sum to
    a(w) times a sum of position factors; other words have other constants."""
    M = (1 << 64) - 1
    c = [oracle.mix64((0x9E3779B97F4A7C15 * (b + 1)) & M) | 1 for b in range(32)]
    assert c[0] == 0xE220A8397B1DCDAF and c[2] == 0x06C45D188009454F  # published splitmix64 outputs
    words = {}
    for base in (0, 32 * 12345, -64, (1 << 31) - 32, -(1 << 31)):
        a = {(oracle.zfactor(base + b) * pow(c[b], -1, 1 << 64)) & M for b in range(32)}
        assert len(a) == 1
        a = a.pop()
        assert a & 1 and a < (1 << 32)
        words[base] = a
        zs = [base + b for b in (0, 3, 17, 31)]
        total = sum(oracle.zfactor(z) for z in zs) & M
        assert total == (a * sum(c[b] for b in (0, 3, 17, 31))) & M
    assert len(set(words.values())) == len(words)


def test_mix64_known_values():
    # splitmix64 finaliser (Steele, Lea, Flood 2014): the outputs of the
    # splitmix64 generator seeded with 0 are mix64(k * 0x9e3779b97f4a7c15), k = 1, 2, 3
    golden = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    for k, g in enumerate(golden, start=1):
        assert oracle.mix64((k * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)) == g


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "small_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_oracle_matches_reference_golden(path):
    d = np.load(path)
    R, S = d["R"], d["S"]
    p, fp, sw = oracle.join_project(R, S)
    assert sw == bool(d["swapped"][0])
    assert np.array_equal(oracle.sort_pairs(p), d["project"])
    cp, cc, _, _ = oracle.join_aggregate_count(R, S)
    gp, gc = oracle.sort_pairs(cp, cc)
    assert np.array_equal(gp, d["agg_pairs"]) and np.array_equal(gc, d["agg_counts"])


def _medium():
    with open(os.path.join(GOLDEN, "medium.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["project_scaled", "dense_block", "aggregate_scaled"])
def test_oracle_matches_reference_medium(name):
    rec = _medium()[name]
    R, S = cases.make_medium(rec["spec"])
    from tests.golden.make_golden import sha

    assert sha(R, S) == rec["sha256"], "generator drift"
    if "project" in rec:
        _, fp, _ = oracle.join_project(R, S, fingerprint_only=True)
        assert [str(v) for v in fp.as_tuple()] == rec["project"]["fp"]
    if "aggregate" in rec:
        _, _, fp, _ = oracle.join_aggregate_count(R, S, fingerprint_only=True)
        assert [str(v) for v in fp.as_tuple()] == rec["aggregate"]["fp"]


@pytest.mark.slow
def test_oracle_kat_gen_rand_data():
    """|Result| = 98,931,829 for gen_rand_data defaults (main.cpp, dim3.hpp:1556-1568)."""
    rec = _medium()["kat"]
    R, S = oracle.gen_rand_data()
    from tests.golden.make_golden import sha

    assert sha(R, S) == rec["sha256"]
    _, fp, _ = oracle.join_project(R, S, fingerprint_only=True)
    assert fp.count == 98_931_829
    assert [str(v) for v in fp.as_tuple()] == rec["project"]["fp"]


def test_gen_rand_data_prefix():
    # minstd_rand0 from seed 492: x_1 = 492*16807 mod (2^31-1) = 8269044
    R, S = oracle.gen_rand_data(2, 1, 1_000_000, 1_000_000, 1_000_000, 492)
    m = 2147483647
    s = 492
    seq = []
    for _ in range(6):
        s = s * 16807 % m
        seq.append(s % 1_000_000)
    assert R.tolist() == [[seq[0], seq[1]], [seq[2], seq[3]]]
    assert S.tolist() == [[seq[4], seq[5]]]


needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference checker not built")


@needs_ref
@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_compiled_reference(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(2000, 30000))
    R = rand_rel(rng, n, 3000, 2000)
    S = rand_rel(rng, int(rng.integers(2000, 30000)), 3000, 2000)
    p = oracle.join_project(R, S)[0]
    q = oracle.ref_join_project(R, S, parallel=bool(seed % 2))
    assert np.array_equal(oracle.sort_pairs(p), oracle.sort_pairs(q))
    cp, cc, _, _ = oracle.join_aggregate_count(R, S)
    rp, rc = oracle.ref_join_aggregate_count(R, S)
    a, b = oracle.sort_pairs(cp, cc), oracle.sort_pairs(rp, rc)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@needs_ref
def test_oracle_estimate_and_generator_match_reference():
    R, S = oracle.gen_rand_data(300_000, 200_000, 50_000, 20_000, 50_000, 77)
    R2, S2 = oracle.ref_gen_rand_data(300_000, 200_000, 50_000, 20_000, 50_000, 77)
    assert np.array_equal(R, R2) and np.array_equal(S, S2)
    assert oracle.estimate(R, S) == oracle.ref_estimate(R, S)
    # |S| <= 10000: the sample keeps every key and the estimate is exact
    R3, S3 = oracle.gen_rand_data(20_000, 8_000, 3_000, 500, 3_000, 5)
    assert oracle.estimate(R3, S3) == oracle.ref_estimate(R3, S3) == oracle.outj(R3, S3)


def brute_values(R, vR, S, vS, agg):
    byy = defaultdict(list)
    for (z, y), v in zip(S.tolist(), vS.tolist()):
        byy[y].append((z, v))
    acc = defaultdict(list)
    for (x, y), v in zip(R.tolist(), vR.tolist()):
        for z, w in byy.get(y, ()):
            acc[(x, z)].append(v + w)
    f = {"sum": sum, "min": min, "max": max,
         "avg": lambda l: int(sum(l) / len(l)) if abs(sum(l)) < 2 ** 52 else None}
    return {k: f[agg](l) for k, l in acc.items()}


def rand_vals(rng, n, lo=-(1 << 31), hi=(1 << 31) - 1):
    return rng.integers(lo, hi, n, endpoint=True).astype(np.int32)


@pytest.mark.parametrize("agg", ["sum", "min", "max", "avg"])
@pytest.mark.parametrize("seed", range(3))
def test_oracle_values_match_brute_force(agg, seed):
    rng = np.random.default_rng(200 + seed)
    R = rand_rel(rng, int(rng.integers(1, 400)), 30, 20, -2, -1)
    S = rand_rel(rng, int(rng.integers(1, 400)), 30, 20, -2, -1)
    vR, vS = rand_vals(rng, R.shape[0], -10 ** 6, 10 ** 6), rand_vals(rng, S.shape[0])
    p, v, fp, sw = oracle.join_aggregate(R, vR, S, vS, agg, exclude_equal=bool(seed == 2))
    big, vb, small, vs = (R, vR, S, vS) if R.shape[0] >= S.shape[0] else (S, vS, R, vR)
    assert sw == (R.shape[0] < S.shape[0])
    want = brute_values(big, vb, small, vs, agg)
    if seed == 2:
        want = {k: w for k, w in want.items() if k[0] != k[1]}
    got = {(int(a), int(b)): int(k) for (a, b), k in zip(p.tolist(), v.tolist())}
    assert got == want
    assert fp.count == len(want) and fp.csum == sum(want.values()) % (1 << 64)


def test_oracle_values_closed_forms():
    rng = np.random.default_rng(7)
    R = rand_rel(rng, 3000, 200, 60)
    S = rand_rel(rng, 2000, 200, 60)
    p, c, _, _ = oracle.join_aggregate_count(R, S)
    cnt = dict(zip(map(tuple, p.tolist()), c.tolist()))
    # constant values: SUM = 5 * count, AVG = MIN = MAX = 5
    five, zero = np.full(R.shape[0], 5, np.int32), np.zeros(S.shape[0], np.int32)
    for agg in ["sum", "min", "max", "avg"]:
        q, v, _, _ = oracle.join_aggregate(R, five, S, zero, agg)
        got = dict(zip(map(tuple, q.tolist()), v.tolist()))
        assert got.keys() == cnt.keys()
        assert all(got[k] == (5 * cnt[k] if agg == "sum" else 5) for k in got)
    # MIN <= AVG <= MAX, and AVG truncates toward zero (negative sums)
    vR, vS = rand_vals(rng, R.shape[0], -100, 50), rand_vals(rng, S.shape[0], -100, 50)
    out = {a: dict(zip(map(tuple, q.tolist()), v.tolist()))
           for a in ["sum", "min", "max", "avg"] for q, v, _, _ in [oracle.join_aggregate(R, vR, S, vS, a)]}
    for k, n in cnt.items():
        s = out["sum"][k]
        assert out["min"][k] <= out["avg"][k] <= out["max"][k]
        assert out["avg"][k] == int(s / n) and (s >= 0 or out["avg"][k] * n >= s)


@needs_ref
@pytest.mark.parametrize("agg", ["sum", "min", "max", "avg"])
def test_oracle_values_match_compiled_reference(agg):
    rng = np.random.default_rng(300)
    for trial in range(3):
        R = rand_rel(rng, int(rng.integers(2000, 20000)), 2000, 800)
        S = rand_rel(rng, int(rng.integers(2000, 20000)), 2000, 800)
        vR, vS = rand_vals(rng, R.shape[0]), rand_vals(rng, S.shape[0])
        p, v, _, _ = oracle.join_aggregate(R, vR, S, vS, agg)
        rp, rv = oracle.ref_join_aggregate(R, vR, S, vS, agg)
        a, b = oracle.sort_pairs(p, v), oracle.sort_pairs(rp, rv)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("path", sorted(set(glob.glob(os.path.join(GOLDEN, "values_*.npz")))
                                       - set(glob.glob(os.path.join(GOLDEN, "values_f32_*.npz")))),
                         ids=lambda p: os.path.basename(p))
def test_oracle_values_match_reference_golden(path):
    d = np.load(path)
    for agg in ["sum", "min", "max", "avg"]:
        p, v, _, _ = oracle.join_aggregate(d["R"], d["vR"], d["S"], d["vS"], agg)
        gp, gv = oracle.sort_pairs(p, v)
        assert np.array_equal(gp, d["pairs"]) and np.array_equal(gv, d[agg]), agg


def brute_values_f32(R, vR, S, vS, agg):
    byy = defaultdict(list)
    for (z, y), v in zip(S.tolist(), vS.tolist()):
        byy[y].append((z, np.float32(v)))
    acc = defaultdict(list)
    for (x, y), v in zip(R.tolist(), vR.tolist()):
        for z, w in byy.get(y, ()):
            acc[(x, z)].append(np.float32(v) + w)  # float add of the path value
    out = {}
    for k, l in acc.items():
        s = sum(float(t) for t in l)  # exact here: the values are multiples of 1/8 below 2^20
        out[k] = {"sum": np.float32(s), "avg": np.float32(s / len(l)), "min": min(l), "max": max(l)}[agg]
    return out


@pytest.mark.parametrize("agg", ["sum", "min", "max", "avg"])
def test_oracle_values_f32_match_brute_force(agg):
    rng = np.random.default_rng(400)
    R = rand_rel(rng, 300, 30, 20, -2, -1)
    S = rand_rel(rng, 250, 30, 20, -2, -1)
    vR = (rng.integers(-8000, 8000, R.shape[0]) / 8).astype(np.float32)
    vS = (rng.integers(-8000, 8000, S.shape[0]) / 8).astype(np.float32)
    p, v, fp, _ = oracle.join_aggregate_f32(R, vR, S, vS, agg)
    want = brute_values_f32(R, vR, S, vS, agg)
    got = {(int(a), int(b)): np.float32(k) for (a, b), k in zip(p.tolist(), v.tolist())}
    assert got.keys() == want.keys()
    assert all(got[k].tobytes() == want[k].tobytes() for k in got)
    # the fingerprint counts the result's bit pattern
    assert fp.csum == int(v.view(np.uint32).astype(np.uint64).sum()) % (1 << 64)


@needs_ref
def test_oracle_values_f32_against_reference_workload():
    """The join_aggregate.cpp workload shape (values = tuple index) with
    values above 2^24, where the reference's float accumulation rounds: equal
    groups and AVG within its rounding (paths * 2^-23 relative)."""
    R, S = oracle.gen_rand_data(200_000, 200_000, 20_000, 20_000, 20_000, 492)
    vR = (np.arange(R.shape[0]) + (1 << 24)).astype(np.float32)
    vS = (np.arange(S.shape[0]) + (1 << 24)).astype(np.float32)
    p, v, _, _ = oracle.join_aggregate_f32(R, vR, S, vS, "avg")
    _, c, _, _ = oracle.join_aggregate_count(R, S)
    rp, rv = oracle.ref_join_aggregate_f32(R, vR, S, vS, "avg")
    o = np.lexsort((p[:, 1], p[:, 0]))
    ro = np.lexsort((rp[:, 1], rp[:, 0]))
    assert np.array_equal(p[o], rp[ro])
    a, b = v[o].astype(np.float64), rv[ro].astype(np.float64)
    cnt = oracle.sort_pairs(*oracle.join_aggregate_count(R, S)[:2])[1].astype(np.float64)
    assert np.all(np.abs(a - b) <= cnt * 2.0 ** -23 * np.abs(b))
    assert (a != b).any()  # the reference does round here


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "values_f32_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_oracle_values_f32_match_reference_golden(path):
    d = np.load(path)
    for agg in ["sum", "min", "max", "avg"]:
        p, v, _, _ = oracle.join_aggregate_f32(d["R"], d["vR"], d["S"], d["vS"], agg)
        gp, gb = oracle.sort_pairs(p, v.view(np.int32).astype(np.int64))
        assert np.array_equal(gp, d["pairs"]) and gb.astype(np.int32).tobytes() == d[agg].tobytes(), agg


def test_workload_generator_matches_oracle_generator():
    """The benchmark's numpy minstd_rand0 generator (workloads.gen_rand_tuples)
    draws what the oracle's gen_rand_data draws."""
    import __graft_entry__ as GE

    GE.load_package()
    from dim3_amd import workloads

    R, vR, S, vS = workloads.gen_rand_tuples(300_000, 200_000, 50_000, 20_000, 50_000, 77)
    R2, S2 = oracle.gen_rand_data(300_000, 200_000, 50_000, 20_000, 50_000, 77)
    assert np.array_equal(R, R2) and np.array_equal(S, S2)
    assert np.array_equal(vR, np.arange(300_000, dtype=np.float32)) and vS.shape == (200_000,)


def _bitset_cases():
    rng = np.random.default_rng(11)
    out = []
    # uniform, Zipf keys (heavy keys), negative and wide values, duplicates
    out.append(("uniform", rand(rng, 3000, 400, 300), rand(rng, 2500, 500, 300)))
    zk = lambda n, dom: np.minimum((rng.pareto(0.8, n) * 3).astype(np.int64), dom - 1)
    R = np.stack([rng.integers(-5000, 5000, 20000), zk(20000, 2000)], 1).astype(np.int32)
    S = np.stack([rng.integers(-2**31, 2**31 - 1, 15000), zk(15000, 2000)], 1).astype(np.int32)
    out.append(("zipf_wide", R, S))
    S2 = np.concatenate([S, S[:3000]])  # duplicate tuples
    out.append(("zipf_dups", R, S2))
    out.append(("swap", S[:4000], R))
    return out


def rand(rng, n, xs, ys):
    return np.stack([rng.integers(0, xs, n), rng.integers(0, ys, n)], 1).astype(np.int32)


@pytest.mark.parametrize("heavy", [0, 1, 8, 64, 1 << 40])
def test_bitset_fingerprint_equals_spa(heavy):
    """oracle_project_fp_bitset (heavy keys as z bitsets, x slices) equals the
    SPA restatement on every case, for every heavy threshold (0 and 2^40: no
    bitset; 1: every key a bitset), summed over 1 or 7 x slices."""
    for name, R, S in _bitset_cases():
        _, want, sw = oracle.join_project(R, S, fingerprint_only=True)
        for slices in (1, 7):
            got, gsw = oracle.fingerprint_bitset(R, S, heavy=heavy, slices=slices)
            assert got.as_tuple() == want.as_tuple() and gsw == sw, (name, heavy, slices)


def test_bitset_fingerprint_self_join_and_count():
    """The u != w filter on a symmetric graph, and the COUNT fingerprint
    (groups from the bitsets, count-weighted terms from path sums) against
    the SPA COUNT with per-group counters."""
    rng = np.random.default_rng(5)
    E = rand(rng, 6000, 900, 900)
    E = np.unique(np.concatenate([E, E[:, ::-1]]), axis=0)
    _, want, _ = oracle.join_project(E, E, exclude_equal=True, fingerprint_only=True)
    for heavy in (0, 4, 16):
        got, _ = oracle.fingerprint_bitset(E, E, exclude_equal=True, heavy=heavy, slices=3, workers=3)
        assert got.as_tuple() == want.as_tuple()
    for name, R, S in _bitset_cases():
        _, _, want, _ = oracle.join_aggregate_count(R, S, fingerprint_only=True)
        got, _ = oracle.fingerprint_bitset(R, S, heavy=16, slices=5, count=True)
        assert got.as_tuple() == want.as_tuple(), name


def test_bitset_fingerprint_against_reference():
    """On a medium Zipf input the bitset fingerprint equals the fingerprint
    of the compiled reference's own rows (DIM3_parallel)."""
    if not oracle.ref_available():
        pytest.skip("reference checker not built")
    rng = np.random.default_rng(8)
    zk = np.minimum((rng.pareto(1.0, 60000) * 5).astype(np.int64), 4999)
    R = np.stack([rng.integers(0, 20000, 60000), zk], 1).astype(np.int32)
    zk = np.minimum((rng.pareto(1.0, 50000) * 5).astype(np.int64), 4999)
    S = np.stack([rng.integers(0, 20000, 50000), zk], 1).astype(np.int32)
    pairs = oracle.ref_join_project(R, S, parallel=True)
    got, _ = oracle.fingerprint_bitset(R, S, heavy=200, slices=4, workers=4)
    assert got.as_tuple() == oracle.fingerprint(pairs).as_tuple()
