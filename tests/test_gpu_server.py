"""GPU tests of the serving loop: every request's result in a concurrent,
batched mix equals the same request run alone (and the CPU oracle's);
admission by footprint, unknown handles and dropped relations."""
import sys

import numpy as np
import pytest

import oracle
from tests.test_gpu_batch import mixed_instances

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_mix_equals_alone(ctx, dim3):
    inst = mixed_instances(11) + mixed_instances(12)
    with dim3.Server(0, workers=4, max_batch=8) as srv:
        hs = [(srv.put(R), srv.put(S)) for R, S in inst]
        reqs = []
        for i, (r, s) in enumerate(hs):
            count = i % 5 == 3
            mat = i % 7 == 2
            kw = dict(output_mode=0 if mat else 1, exclude_equal=(i % 4 == 1))
            rc, t = srv.submit(r, s, count=count, **kw)
            assert rc == dim3.OK
            reqs.append((i, count, kw, t))
        say("submitted", len(reqs), "requests")
        for i, count, kw, t in reqs:
            rc, res, q, tot = srv.wait(t)
            assert rc == dim3.OK, i
            assert tot >= q >= 0
            R, S = inst[i]
            alone = ctx.join_aggregate_count(R, S, **kw) if count else ctx.join_project(R, S, **kw)
            assert res.fp.as_tuple() == alone.fp.as_tuple(), i
            assert res.swapped == alone.swapped and res.n == alone.n, i
            if kw["output_mode"] == 0:
                if count:
                    gp, gc = oracle.sort_pairs(res.xz, res.counts)
                    wp, wc, _, _ = oracle.join_aggregate_count(R, S, exclude_equal=kw["exclude_equal"])
                    ep, ec = oracle.sort_pairs(wp, wc)
                    assert np.array_equal(gp, ep) and np.array_equal(gc, ec), i
                else:
                    want, _, _ = oracle.join_project(R, S, exclude_equal=kw["exclude_equal"])
                    assert np.array_equal(oracle.sort_pairs(res.pairs), oracle.sort_pairs(want)), i
        st = srv.stats()
        say("server stats", st)
        assert st["completed"] == len(reqs)
        assert st["batches"] >= 1 and st["batched_requests"] >= 2


def test_admission_and_handles(dim3):
    rng = np.random.default_rng(5)
    big = np.stack([rng.integers(0, 1000, 1_000_000), rng.integers(0, 1000, 1_000_000)], 1).astype(np.int32)
    small = big[:1000].copy()
    with dim3.Server(0, hbm_budget_bytes=64 << 20, workers=2) as srv:
        hb = srv.put(big)
        hs = srv.put(small)
        rc, _ = srv.submit(hb, hb, output_mode=1)
        assert rc == dim3.ERR_ADMISSION  # ~200 MB of workspace never fits 64 MiB
        rc, t = srv.submit(hs, hs, output_mode=1)
        assert rc == dim3.OK
        srv.drop(hs)  # freed once the request is done
        rc, res, _, _ = srv.wait(t)
        assert rc == dim3.OK
        _, want, _ = oracle.join_project(small, small, fingerprint_only=True)
        assert res.fp.as_tuple() == want.as_tuple()
        with pytest.raises(dim3.Dim3Error):
            srv.submit(hs, hs, output_mode=1)
        with pytest.raises(dim3.Dim3Error):
            srv.wait(123456)
        st = srv.stats()
        assert st["rejected"] == 1 and st["completed"] == 1


def test_large_lane_partition(ctx, dim3):
    """Requests above large_tuples go to worker 0 on a CU-masked stream while
    the others run on the remaining CUs; every result equals the same
    request run alone."""
    inst = mixed_instances(31)
    with dim3.Server(0, workers=3, max_batch=4, large_share_pct=80, large_tuples=4000) as srv:
        hs = [(srv.put(R), srv.put(S)) for R, S in inst]
        reqs = []
        for i, (r, s) in enumerate(hs):
            count = i % 3 == 1
            rc, t = srv.submit(r, s, count=count, output_mode=1)
            assert rc == dim3.OK
            reqs.append((i, count, t))
        n_large = sum(1 for R, S in inst if R.shape[0] + S.shape[0] >= 4000)
        say("submitted", len(reqs), "requests,", n_large, "large")
        assert n_large >= 2
        for i, count, t in reqs:
            rc, res, _, _ = srv.wait(t)
            assert rc == dim3.OK, i
            R, S = inst[i]
            alone = ctx.join_aggregate_count(R, S, output_mode=1) if count else ctx.join_project(R, S, output_mode=1)
            assert res.fp.as_tuple() == alone.fp.as_tuple(), i
        assert srv.stats()["completed"] == len(reqs)


def test_server_value_requests(dim3):
    """Value aggregates through the server (int and float values, materialised
    and fingerprint) equal the oracle's; mixed with project requests."""
    import oracle

    rng = np.random.default_rng(77)
    R = np.stack([rng.integers(0, 2000, 30_000), rng.integers(0, 500, 30_000)], 1).astype(np.int32)
    S = np.stack([rng.integers(0, 2000, 20_000), rng.integers(0, 500, 20_000)], 1).astype(np.int32)
    vi_r, vi_s = rng.integers(-10 ** 6, 10 ** 6, 30_000).astype(np.int32), rng.integers(-10 ** 6, 10 ** 6, 20_000).astype(np.int32)
    vf_r, vf_s = (rng.integers(-4000, 4000, 30_000) / 4).astype(np.float32), (rng.integers(-4000, 4000, 20_000) / 4).astype(np.float32)
    with dim3.Server(0, workers=3) as srv:
        ri, si, rf, sf = srv.put(R), srv.put(S), srv.put(R), srv.put(S)
        srv.put_values(ri, vi_r)
        srv.put_values(si, vi_s)
        srv.put_values(rf, vf_r)
        srv.put_values(sf, vf_s)
        with pytest.raises(dim3.Dim3Error):
            srv.put_values(ri, vi_r)  # once per relation
        tickets = []
        for agg in ("sum", "min", "max", "avg"):
            tickets.append(("i", agg, 0, srv.submit(ri, si, agg=agg)[1]))
            tickets.append(("f", agg, 0, srv.submit(rf, sf, agg=agg)[1]))
            tickets.append(("f", agg, 1, srv.submit(rf, sf, agg=agg, output_mode=1)[1]))
        tp = srv.submit(ri, si)[1]
        # values of two different types: refused
        with pytest.raises(dim3.Dim3Error):
            srv.submit(ri, sf, agg="sum")
        for kind, agg, fpm, t in tickets:
            rc, res, _, _ = srv.wait(t)
            assert rc == dim3.OK, (kind, agg)
            if kind == "i":
                wp, wv, wfp, _ = oracle.join_aggregate(R, vi_r, S, vi_s, agg)
                got, want = oracle.sort_pairs(res.xz, res.counts), oracle.sort_pairs(wp, wv)
            else:
                wp, wv, wfp, _ = oracle.join_aggregate_f32(R, vf_r, S, vf_s, agg)
                if fpm:
                    assert res.fp.as_tuple() == wfp.as_tuple(), agg
                    continue
                got = oracle.sort_pairs(res.xz, res.counts.view(np.int32).astype(np.int64))
                want = oracle.sort_pairs(wp, wv.view(np.int32).astype(np.int64))
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), (kind, agg)
        rc, res, _, _ = srv.wait(tp)
        want, _, _ = oracle.join_project(R, S)
        assert rc == dim3.OK and np.array_equal(oracle.sort_pairs(res.pairs), oracle.sort_pairs(want))


def test_materialize_admission_by_estimate(dim3):
    """The join-size estimate only sizes a materialising request's output
    pre-allocation (capped at a quarter of the free budget); it never refuses
    a request.  One key shared by 4000 tuples on each side is 1.6e7 rows
    (128 MB): under a 64 MiB budget the request is admitted and ends with
    DIM3_ERR_OOM once the query reports its real size; the fingerprint of
    the same query fits."""
    n = 4000
    R = np.stack([np.arange(n), np.zeros(n)], 1).astype(np.int32)
    S = np.stack([np.arange(n) + 7, np.zeros(n)], 1).astype(np.int32)
    with dim3.Server(0, hbm_budget_bytes=64 << 20, workers=1) as srv:
        hr, hs = srv.put(R), srv.put(S)
        rc, t = srv.submit(hr, hs, output_mode=1)
        assert rc == dim3.OK
        rc, res, _, _ = srv.wait(t)
        assert rc == dim3.OK and res.n == n * n
        rc, t = srv.submit(hr, hs, output_mode=0)
        assert rc == dim3.OK
        rc, res, _, _ = srv.wait(t)
        assert rc == dim3.ERR_OOM
        assert srv.stats()["rejected"] == 0
    with dim3.Server(0, hbm_budget_bytes=512 << 20, workers=1) as srv:
        hr, hs = srv.put(R), srv.put(S)
        rc, t = srv.submit(hr, hs, output_mode=0)
        assert rc == dim3.OK
        rc, res, _, _ = srv.wait(t)
        assert rc == dim3.OK and res.n == n * n
        want, _, _ = oracle.join_project(R, S)
        assert np.array_equal(oracle.sort_pairs(res.pairs), oracle.sort_pairs(want))
    say("materialise admission ok")


def test_materialize_paths_far_above_rows(dim3):
    """1e7 join paths but 1e4 distinct rows (100 x 100 values joined through
    the same 1000 keys): the estimate (one row per path, 80 MB) exceeds a
    64 MiB budget, yet the request is admitted and its rows are exact."""
    x = np.repeat(np.arange(100), 1000)
    y = np.tile(np.arange(1000), 100)
    R = np.stack([x, y], 1).astype(np.int32)
    S = np.stack([x + 50, y], 1).astype(np.int32)
    assert oracle.estimate(R, S) >= 8_000_000
    with dim3.Server(0, hbm_budget_bytes=64 << 20, workers=1) as srv:
        hr, hs = srv.put(R), srv.put(S)
        for kind in ("project", "count"):
            rc, t = srv.submit(hr, hs, count=kind == "count", output_mode=0)
            assert rc == dim3.OK, kind
            rc, res, _, _ = srv.wait(t)
            assert rc == dim3.OK and res.n == 100 * 100, kind
            if kind == "project":
                want, _, _ = oracle.join_project(R, S)
                assert np.array_equal(oracle.sort_pairs(res.pairs), oracle.sort_pairs(want))
            else:
                wp, wc, _, _ = oracle.join_aggregate_count(R, S)
                gp, gc = oracle.sort_pairs(res.xz, res.counts)
                ep, ec = oracle.sort_pairs(wp, wc)
                assert np.array_equal(gp, ep) and np.array_equal(gc, ec)
                assert int(gc.sum()) == 100 * 100 * 1000
        assert srv.stats()["rejected"] == 0
    say("paths far above rows ok")


def test_materialize_under_estimate(dim3):
    """The estimate can miss every path (here: no sampled key joins and the
    1000-tuple prefixes do not meet); the worker then grows its output
    buffer to the size the query reports and runs it again."""
    n = 20_000
    i = np.arange(n)
    R = np.stack([i, 2 * i + 100_001], 1).astype(np.int32)
    S = np.stack([i, 2 * i + 200_001], 1).astype(np.int32)
    R[1000:1100, 1] = 7
    S[1000:1100, 1] = 7
    assert oracle.estimate(R, S) == 0
    with dim3.Server(0, workers=1) as srv:
        hr, hs = srv.put(R), srv.put(S)
        rc, t = srv.submit(hr, hs, output_mode=0)
        assert rc == dim3.OK
        rc, res, _, _ = srv.wait(t)
        assert rc == dim3.OK and res.n == 100 * 100
        want, _, _ = oracle.join_project(R, S)
        assert np.array_equal(oracle.sort_pairs(res.pairs), oracle.sort_pairs(want))
    say("under-estimate ok")
