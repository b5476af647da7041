"""GPU tests of the streamed output mode: the rows of a query delivered to a
host callback in chunks of bounded size equal the oracle's rows (project and
COUNT, dense and sparse plans, the u != w filter); the callback can stop the
query; empty queries deliver nothing."""
import sys

import numpy as np
import pytest

import oracle
from tests.golden import cases
from tests.test_gpu_batch import planted, rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def collect(ctx, R, S, count, chunk_rows, **kw):
    chunks = []

    def on_rows(rows):
        chunks.append(rows.copy())

    fn = ctx.join_aggregate_count_stream if count else ctx.join_project_stream
    rc, res = fn(R, S, on_rows, chunk_rows=chunk_rows, **kw)
    return rc, res, chunks


def check(ctx, dim3, R, S, count, chunk_rows, min_chunks=1, **kw):
    rc, res, chunks = collect(ctx, R, S, count, chunk_rows, **kw)
    assert rc == dim3.OK
    assert len(chunks) >= min_chunks
    assert all(c.shape[0] <= max(chunk_rows, res.stats.nz) for c in chunks)
    ex = kw.get("exclude_equal", False)
    if count:
        rows = np.concatenate(chunks) if chunks else np.zeros((0,), dtype=dim3.COUNT_ROW_DTYPE)
        got_p, got_c = oracle.sort_pairs(np.stack([rows["x"], rows["z"]], 1), rows["count"])
        wp, wc, _, sw = oracle.join_aggregate_count(R, S, exclude_equal=ex)
        want_p, want_c = oracle.sort_pairs(wp, wc)
        assert np.array_equal(got_p, want_p) and np.array_equal(got_c, want_c)
    else:
        rows = np.concatenate(chunks) if chunks else np.zeros((0, 2), np.int32)
        want, _, sw = oracle.join_project(R, S, exclude_equal=ex)
        assert np.array_equal(oracle.sort_pairs(rows), oracle.sort_pairs(want))
    assert res.n == rows.shape[0] and res.swapped == sw
    return len(chunks)


@pytest.mark.parametrize("count", [False, True])
def test_stream_zipf_chunks(ctx, dim3, count):
    spec = cases.MEDIUM["aggregate_scaled" if count else "project_scaled"]
    R, S = cases.make_medium(spec)
    # a 1/8 slice keeps the oracle quick; small chunks force many of them
    R = R[R[:, 0] % 8 == 0]
    say("stream", "count" if count else "project", R.shape, S.shape)
    n = check(ctx, dim3, R, S, count, chunk_rows=50_000, min_chunks=3)
    say("chunks", n)


@pytest.mark.parametrize("dense", [0, 1, 2])
def test_stream_dense_plans(ctx, dim3, dense):
    R = planted(41, 6000, 200, 0.3, 4000, 900)
    S = planted(42, 6000, 200, 0.3, 4000, 900)
    say("stream dense", dense, R.shape, S.shape)
    check(ctx, dim3, R, S, False, chunk_rows=20_000, min_chunks=2, dense_mode=dense)
    if dense != 1:
        check(ctx, dim3, R, S, True, chunk_rows=20_000, min_chunks=2, dense_mode=dense)


def test_stream_self_join_filter(ctx, dim3):
    rng = np.random.default_rng(9)
    E = rel(rng, 20_000, 3000, 3000)
    E = np.concatenate([E, E[:, ::-1]])
    check(ctx, dim3, E, E, False, chunk_rows=100_000, min_chunks=2, exclude_equal=True)
    check(ctx, dim3, E, E, True, chunk_rows=100_000, min_chunks=2, exclude_equal=True)


def test_stream_cancel_and_empty(ctx, dim3):
    rng = np.random.default_rng(10)
    R, S = rel(rng, 20_000, 500, 100), rel(rng, 20_000, 500, 100)
    seen = []

    def stop(rows):
        seen.append(rows.shape[0])
        return False

    rc, _ = ctx.join_project_stream(R, S, stop, chunk_rows=10_000)
    assert rc == dim3.ERR_CANCELLED and len(seen) == 1
    # an ordinary query on the same context still works afterwards
    check(ctx, dim3, R, S, False, chunk_rows=10_000, min_chunks=2)
    calls = []
    rc, res = ctx.join_project_stream(np.zeros((0, 2), np.int32), S, lambda rows: calls.append(1))
    assert rc == dim3.OK and res.n == 0 and not calls
    rc, res = ctx.join_aggregate_count_stream(R, (S + [0, 10_000]).astype(np.int32), lambda rows: calls.append(1))
    assert rc == dim3.OK and res.n == 0 and not calls


@pytest.mark.parametrize("f32", [False, True])
def test_stream_value_aggregates(ctx, dim3, f32):
    """Value aggregates streamed in chunks equal the oracle's rows; chunks of
    short rows (one per wavefront) and of tiled long rows."""
    rng = np.random.default_rng(12)
    spec = cases.MEDIUM["aggregate_scaled"]
    R, S = cases.make_medium(spec)
    R = R[R[:, 0] % 8 == 0]
    if f32:
        vR = (rng.integers(-4000, 4000, R.shape[0]) / 4).astype(np.float32)
        vS = (rng.integers(-4000, 4000, S.shape[0]) / 4).astype(np.float32)
    else:
        vR = rng.integers(-10 ** 6, 10 ** 6, R.shape[0]).astype(np.int32)
        vS = rng.integers(-10 ** 6, 10 ** 6, S.shape[0]).astype(np.int32)
    for agg in ("sum", "avg"):
        chunks = []
        rc, res = ctx.join_aggregate_stream(R, vR, S, vS, agg, lambda rows: chunks.append(rows.copy()),
                                            chunk_rows=40_000, f32=f32)
        assert rc == dim3.OK and len(chunks) >= 3
        rows = np.concatenate(chunks)
        if f32:
            wp, wv, _, _ = oracle.join_aggregate_f32(R, vR, S, vS, agg)
            got = oracle.sort_pairs(np.stack([rows["x"], rows["z"]], 1), rows["v"].view(np.int32).astype(np.int64))
            want = oracle.sort_pairs(wp, wv.view(np.int32).astype(np.int64))
        else:
            wp, wv, _, _ = oracle.join_aggregate(R, vR, S, vS, agg)
            got = oracle.sort_pairs(np.stack([rows["x"], rows["z"]], 1), rows["count"])
            want = oracle.sort_pairs(wp, wv)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), agg
        assert res.n == rows.shape[0]
