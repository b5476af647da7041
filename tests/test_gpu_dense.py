"""GPU parity of the dense part: heavy join keys routed to the int8 MFMA
product or to OR-ed z bitmaps, against numpy (the product on its own: the
closed form C = A @ Bt.T, C > 0) and against the CPU oracle (whole queries),
bit-exact.  The all-sparse, MFMA, bit-or and planner-chosen plans are also
checked against each other on fingerprints."""
import sys

import numpy as np
import pytest

import oracle
from tests.test_gpu_parity import check_count, check_project

pytestmark = pytest.mark.gpu

MFMA, BITOR, OFF = 2, 1, 0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pack_bits(C):
    """Bit rows of C > 0: bit n % 32 of uint32 word n // 32."""
    M, N = C.shape
    W = (N + 31) // 32
    b = np.zeros((M, W * 32), np.uint8)
    b[:, :N] = C > 0
    return np.packbits(b, axis=1, bitorder="little").view("<u4").reshape(M, W)


@pytest.mark.parametrize("shape", [(1, 1, 1), (5, 7, 3), (64, 64, 64), (130, 200, 70), (257, 129, 300),
                                   (128, 256, 1), (300, 33, 1000), (1000, 1000, 64)])
def test_dense_product_matches_numpy(ctx, shape):
    M, N, K = shape
    say("dense product", shape)
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.integers(-128, 128, (M, K)).astype(np.int8)
    Bt = rng.integers(-128, 128, (N, K)).astype(np.int8)
    C = A.astype(np.int64) @ Bt.astype(np.int64).T
    bits, cnt = ctx.dense_product(A, Bt)
    assert np.array_equal(cnt.astype(np.int64), C)
    assert np.array_equal(bits, pack_bits(C))


def test_dense_product_boolean_and_extremes(ctx):
    rng = np.random.default_rng(5)
    A = (rng.random((200, 4096)) < 0.01).astype(np.int8)
    Bt = (rng.random((300, 4096)) < 0.01).astype(np.int8)
    C = A.astype(np.int64) @ Bt.astype(np.int64).T
    bits, _ = ctx.dense_product(A, Bt, counts=False)
    assert np.array_equal(bits, pack_bits(C))
    # the largest magnitudes: -128 * -128 summed over K
    A = np.full((17, 4096), -128, np.int8)
    Bt = np.full((19, 4096), -128, np.int8)
    _, cnt = ctx.dense_product(A, Bt, bits=False)
    assert (cnt == 128 * 128 * 4096).all()


def planted(seed, n_bg, block, density, dom_a, dom_b, s=1.0, offset=0):
    """Zipf(s) background plus a block of `block` attribute values x `block`
    keys, each cell kept with probability `density`; distinct tuples."""
    rng = np.random.default_rng(seed)
    k = np.arange(dom_b, dtype=np.float64) + 1.0
    p = k ** -s
    p /= p.sum()
    bg = np.stack([rng.integers(0, dom_a, n_bg), rng.choice(dom_b, n_bg, p=p)], 1)
    m = rng.random((block, block)) < density
    xi, yi = np.nonzero(m)
    blk = np.stack([xi + offset, yi], 1)
    R = np.unique(np.concatenate([bg, blk]).astype(np.int64), axis=0).astype(np.int32)
    return R[rng.permutation(R.shape[0])]


def planted_pair(seed=1, n_bg=20000, block=300, density=0.3):
    R = planted(seed, n_bg, block, density, 5000, 2000)
    S = planted(seed + 100, n_bg, block, density, 5000, 2000, offset=7)
    return R, S


@pytest.mark.parametrize("mode,heavy_k", [(MFMA, 0), (MFMA, 8), (MFMA, 400), (BITOR, 0), (BITOR, 50)])
def test_project_dense_vs_oracle(ctx, mode, heavy_k):
    R, S = planted_pair()
    got = check_project(ctx, R, S, dense_mode=mode, heavy_k=heavy_k)
    st = got.stats
    say("  dense", st.dense_mode, "keys", st.heavy_keys, "rows", st.nx_heavy, "cols", st.nz_heavy,
        "sparse paths", st.paths_sparse, "of", st.out_j)
    assert st.dense_mode == mode and st.heavy_keys > 0
    assert st.paths_sparse < st.out_j


@pytest.mark.parametrize("heavy_k", [0, 8, 400])
def test_count_dense_vs_oracle(ctx, heavy_k):
    R, S = planted_pair(seed=3)
    got = check_count(ctx, R, S, dense_mode=MFMA, heavy_k=heavy_k)
    assert got.stats.dense_mode == MFMA and got.stats.heavy_keys > 0
    assert int(got.counts.sum()) == got.stats.out_j


@pytest.mark.parametrize("tile_bits", [64, 1000])
@pytest.mark.parametrize("mode", [MFMA, BITOR])
def test_dense_with_z_tiles(ctx, tile_bits, mode):
    R, S = planted_pair(seed=5, n_bg=8000, block=200)
    got = check_project(ctx, R, S, dense_mode=mode, tile_bits=tile_bits)
    assert got.stats.n_tiles > 1 and got.stats.dense_mode == mode
    agg = check_count(ctx, R, S, dense_mode=MFMA, tile_bits=tile_bits)
    assert agg.stats.dense_mode == MFMA


@pytest.mark.parametrize("mode", [MFMA, BITOR])
def test_dense_hash_mapping_and_self_join(ctx, mode):
    rng = np.random.default_rng(9)
    R, _ = planted_pair(seed=9, n_bg=10000, block=250)
    # spread values so every attribute is hashed, then a self-join with u != w
    R = (R.astype(np.int64) * 7919 - 3_000_000).astype(np.int32)
    got = check_project(ctx, R, R, dense_mode=mode, exclude_equal=True)
    assert got.stats.dense_mode == mode
    check_project(ctx, R, R, dense_mode=mode, exclude_equal=True, hash_only=True)
    check_count(ctx, R, R, dense_mode=MFMA, exclude_equal=True)
    E = planted(int(rng.integers(100)), 6000, 150, 0.4, 3000, 1500)
    check_project(ctx, E, E[:, ::-1].copy(), dense_mode=mode)


def test_count_multiplicity_and_overflow_fallback(ctx):
    R, S = planted_pair(seed=11, n_bg=5000, block=120, density=0.4)
    # duplicates of heavy tuples count with multiplicity in the product
    Rd = np.concatenate([R, R[:3000], R[:500]])
    got = check_count(ctx, Rd, S, dense_mode=MFMA)
    assert got.stats.dense_mode == MFMA
    # a multiplicity above 127 cannot be an int8 operand: the query reruns all-sparse
    hot = R[R[:, 1] == 0][:1]  # key 0 heads the Zipf law and the block: always heavy
    Ro = np.concatenate([R] + [hot] * 200)
    got = check_count(ctx, Ro, S, dense_mode=MFMA, heavy_k=400)
    assert got.stats.dense_mode == OFF


def test_dense_modes_agree_on_fingerprints(ctx):
    R = planted(21, 150000, 1000, 0.3, 20000, 8000)
    S = planted(22, 150000, 1000, 0.3, 20000, 8000, offset=11)
    _, wfp, _ = oracle.join_project(R, S, fingerprint_only=True)
    _, _, wcfp, _ = oracle.join_aggregate_count(R, S, fingerprint_only=True)
    fps = {}
    for mode in (OFF, MFMA, BITOR, -1):
        say("fingerprint mode", mode)
        r = ctx.join_project(R, S, dense_mode=mode, output_mode=1)
        fps[mode] = r.fp.as_tuple()
        assert fps[mode] == wfp.as_tuple(), mode
        if mode != BITOR:
            a = ctx.join_aggregate_count(R, S, dense_mode=mode, output_mode=1)
            assert a.fp.as_tuple() == wcfp.as_tuple(), mode
    auto = ctx.join_project(R, S, output_mode=1)
    say("  auto picked", auto.stats.dense_mode, "keys", auto.stats.heavy_keys)


def test_bitor_rejected_for_count(ctx, dim3):
    R, S = planted_pair(seed=13, n_bg=2000, block=50)
    with pytest.raises(dim3.Dim3Error) as e:
        ctx.join_aggregate_count(R, S, dense_mode=BITOR)
    assert e.value.status == dim3.ERR_UNSUPPORTED


def zipf_with_block(seed, n_bg, block, density, dom_a, dom_b, s=1.0):
    """Zipf(s) background whose hottest keys also carry a planted block."""
    return planted(seed, n_bg, block, density, dom_a, dom_b, s=s)


@pytest.mark.parametrize("mode", [MFMA, BITOR])
@pytest.mark.parametrize("tile_bits", [0, 640])
def test_dense_core_vs_oracle(ctx, mode, tile_bits):
    """Forced dense core: only rows / columns with many heavy keys enter the
    product; the other heavy paths stay in the (split) sparse CSR."""
    R = zipf_with_block(41, 30000, 250, 0.35, 4000, 1500)
    S = zipf_with_block(42, 30000, 250, 0.35, 4000, 1500)
    got = check_project(ctx, R, S, dense_mode=mode, dense_core=1, tile_bits=tile_bits)
    st = got.stats
    say("  core", st.dense_mode, "keys", st.heavy_keys, "rows", st.nx_heavy, "of", st.nx, "cols", st.nz_heavy,
        "of", st.nz, "sparse paths", st.paths_sparse, "of", st.out_j)
    assert st.dense_mode == mode and st.heavy_keys > 0
    assert st.nx_heavy < st.nx  # a real core: not every row with a heavy key
    if mode == MFMA:
        agg = check_count(ctx, R, S, dense_mode=MFMA, dense_core=1, tile_bits=tile_bits)
        assert agg.stats.dense_mode == MFMA


def test_dense_core_self_join_and_hashing(ctx):
    R = zipf_with_block(43, 20000, 200, 0.4, 3000, 1000)
    R = (R.astype(np.int64) * 7919 - 5_000_000).astype(np.int32)
    got = check_project(ctx, R, R, dense_mode=MFMA, dense_core=1, exclude_equal=True)
    assert got.stats.dense_mode == MFMA
    check_count(ctx, R, R, dense_mode=MFMA, dense_core=1, exclude_equal=True, hash_only=True)
