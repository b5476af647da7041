"""GPU parity of the COUNT fingerprint whose groups' counts are formed: the
heavy keys' counts as an int8 MFMA product over the (x, z) box, the light
keys' counts per (x, z-tile) task (k_group_count.hip), against the CPU
oracle's fingerprint of its (x, z, count) rows, bit-exact.  The oracle's
count-weighted terms (sum2, csum) change with any single group's count (odd
multipliers), so a wrong count fails the comparison."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def zipf_rel(n, attr, keys, s, seed, attr_lo=0):
    from dim3_amd import workloads

    r = workloads.zipf_relation(n, attr, keys, s, seed)
    r[:, 0] += attr_lo
    return r


def want_fp(R, S):
    _, _, fp, sw = oracle.join_aggregate_count(R, S, fingerprint_only=True)
    return fp.as_tuple(), sw


def check(ctx, dim3, R, S, mfma=True, **kw):
    want, sw = want_fp(R, S)
    say("group count", R.shape[0], S.shape[0], kw, "want", want)
    dense = dim3.DENSE_MFMA if mfma else kw.pop("dense_mode", dim3.DENSE_AUTO)
    g = ctx.join_aggregate_count(R, S, output_mode=dim3.OUT_FINGERPRINT, dense_mode=dense, **kw)
    say("stats", g.stats.as_dict())
    assert g.swapped == sw
    assert g.fp.as_tuple() == want
    assert g.n == want[0]
    return g


def test_zipf_box_forced(ctx, dim3):
    R = zipf_rel(150_000, 15_000, 5_000, 1.2, 11)
    S = zipf_rel(150_000, 15_000, 5_000, 1.2, 12)
    g = check(ctx, dim3, R, S)
    assert g.stats.dense_mode == dim3.DENSE_MFMA and 0 < g.stats.heavy_keys <= 64
    assert g.stats.nx_heavy > 0 and g.stats.nz_heavy > 0
    # the same fingerprint from join-path sums, and from the all-sparse counters
    p = ctx.join_aggregate_count(R, S, output_mode=dim3.OUT_FINGERPRINT, path_sums=True)
    assert p.fp.as_tuple() == g.fp.as_tuple()
    o = ctx.join_aggregate_count(R, S, output_mode=dim3.OUT_FINGERPRINT, dense_mode=dim3.DENSE_OFF)
    assert o.fp.as_tuple() == g.fp.as_tuple() and o.stats.dense_mode == dim3.DENSE_OFF


@pytest.mark.parametrize("heavy_k", [1, 5, 64])
def test_heavy_key_counts(ctx, dim3, heavy_k):
    R = zipf_rel(120_000, 12_000, 4_000, 1.1, 21, attr_lo=-6_000)
    S = zipf_rel(100_000, 12_000, 4_000, 1.1, 22, attr_lo=-6_000)
    g = check(ctx, dim3, R, S, heavy_k=heavy_k)
    assert g.stats.dense_mode == dim3.DENSE_MFMA and 0 < g.stats.heavy_keys <= heavy_k


def test_hashed_attributes(ctx, dim3):
    R = zipf_rel(80_000, 9_000, 3_000, 1.3, 31)
    S = zipf_rel(90_000, 9_000, 3_000, 1.3, 32)
    R[:, 0] = R[:, 0] * 7919 - 123456  # sparse x and z values: hash mapping
    S[:, 0] = S[:, 0] * 104729 + 17
    g = check(ctx, dim3, R, S, hash_only=True)
    assert g.stats.dense_mode == dim3.DENSE_MFMA


def test_swap_rule(ctx, dim3):
    R = zipf_rel(40_000, 6_000, 2_000, 1.2, 41)
    S = zipf_rel(90_000, 6_000, 2_000, 1.2, 42)
    g = check(ctx, dim3, R, S)
    assert g.swapped


def long_rows_case(seed=51):
    """50 x values, each with 3 heavy keys (every x, 5000 z each) and ~100
    light keys out of 2000 (30 z each): R rows of ~103 entries (two 64-entry
    chunks) and ~3000 light paths per row."""
    rng = np.random.default_rng(seed)
    xs = np.arange(50, dtype=np.int64)
    R = [np.stack([np.repeat(xs, 3), np.tile(np.arange(3), 50)], 1)]
    R.append(np.stack([np.repeat(xs, 100), rng.integers(3, 2003, 5000)], 1))
    S = [np.stack([rng.integers(0, 20_000, 15_000), np.repeat(np.arange(3), 5000)], 1)]
    S.append(np.stack([rng.integers(0, 20_000, 60_000), np.repeat(np.arange(3, 2003), 30)], 1))
    R = np.unique(np.concatenate(R), axis=0).astype(np.int32)
    S = np.unique(np.concatenate(S), axis=0).astype(np.int32)
    return rng.permutation(R), rng.permutation(S)


def test_long_rows_and_light_passes(ctx, dim3):
    # one z tile (tile_bits) makes (x, z-tile) tasks of ~3000 light paths,
    # which run in passes over z ranges of 512 IDs
    R, S = long_rows_case()
    for tb in (0, 1 << 20):
        g = check(ctx, dim3, R, S, tile_bits=tb)
        assert g.stats.dense_mode == dim3.DENSE_MFMA and 3 <= g.stats.heavy_keys <= 64
        assert g.stats.paths_sparse > 0


def test_all_keys_heavy(ctx, dim3):
    R = zipf_rel(30_000, 5_000, 40, 0.5, 61)
    S = zipf_rel(30_000, 5_000, 40, 0.5, 62)
    g = check(ctx, dim3, R, S)
    assert g.stats.heavy_keys == 40 and g.stats.paths_sparse == 0


def test_duplicate_heavy_tuples(ctx, dim3):
    # duplicate tuples count with multiplicity; a bit mask cannot hold them,
    # so the plan falls back to counters (same fingerprint)
    R = zipf_rel(50_000, 6_000, 2_000, 1.2, 71)
    S = zipf_rel(50_000, 6_000, 2_000, 1.2, 72)
    R = np.concatenate([R, R[:5_000]])
    check(ctx, dim3, R, S)


def test_auto_chooses_box_for_dense_counts(ctx, dim3):
    # Zipf(1.2) keys on a small a / c domain: the heavy keys fill the box,
    # so AUTO takes the MFMA counts
    R = zipf_rel(400_000, 20_000, 50_000, 1.2, 81)
    S = zipf_rel(400_000, 20_000, 50_000, 1.2, 82)
    g = check(ctx, dim3, R, S, mfma=False)
    assert g.stats.dense_mode == dim3.DENSE_MFMA
