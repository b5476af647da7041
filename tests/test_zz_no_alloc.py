"""dim3_join_project_dev_noalloc against dim3_join_project_dev.

The allocation-free entry point takes every workspace buffer from the
library's per-device scratch array: its rows must be exactly those of the
existing device entry point (sorted pairs, bit for bit, and the
fingerprint), on two calls in a row with the second input smaller (the array
keeps the first call's contents), in materialised and fingerprint mode, at
the 2^17-tuple limit with hash-mapped attributes, and the context that ran
them must hold no device memory of its own."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(rng, n, xs, ys, off=0):
    return np.stack([rng.integers(0, xs, n) + off, rng.integers(0, ys, n)], 1).astype(np.int32)


def run_both(dim3, ctx, nctx, torch, R, S, **kw):
    dR = torch.from_numpy(R).cuda()
    dS = torch.from_numpy(S).cuda()
    rc, r0 = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, **kw)
    n = r0.n
    want = torch.zeros((max(n, 1), 2), dtype=torch.int32, device="cuda")
    rc, a = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], want.data_ptr(), n, **kw)
    assert rc == dim3.OK
    got = torch.full((max(n, 1), 2), -7, dtype=torch.int32, device="cuda")
    rc, b = nctx.join_project_dev_noalloc(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], got.data_ptr(), n, **kw)
    assert rc == dim3.OK, nctx.last_error()
    torch.cuda.synchronize()
    assert b.n == a.n and b.swapped == a.swapped
    assert np.array_equal(oracle.sort_pairs(got[:n].cpu().numpy()), oracle.sort_pairs(want[:n].cpu().numpy()))
    rc, fa = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, output_mode=1, **kw)
    rc2, fb = nctx.join_project_dev_noalloc(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, output_mode=1,
                                            **kw)
    assert rc == rc2 == dim3.OK and fa.fp.as_tuple() == fb.fp.as_tuple()
    return n


def test_noalloc_matches_device_entry_point(dim3, ctx, torch):
    rng = np.random.default_rng(2024)
    nctx = dim3.Context(0)
    try:
        say("noalloc: 120000 x 90000, then 9000 x 12000 (swap)")
        n1 = run_both(dim3, ctx, nctx, torch, rel(rng, 120_000, 20_000, 3_000), rel(rng, 90_000, 25_000, 3_000))
        n2 = run_both(dim3, ctx, nctx, torch, rel(rng, 9_000, 900, 400), rel(rng, 12_000, 1_100, 400))
        assert n1 > n2 > 0
        say("noalloc: the u != w self-join of a small graph")
        E = rel(rng, 30_000, 5_000, 5_000)
        E = np.concatenate([E, E[:, ::-1]])
        run_both(dim3, ctx, nctx, torch, E, E, exclude_equal=True)
        say("noalloc: 2^17 tuples per side, hash-mapped attributes")
        big = 1 << 17
        R = np.stack([rng.integers(-2**31, 2**31 - 1, big), rng.integers(0, 50_000, big) * 40_000], 1).astype(np.int32)
        S = np.stack([rng.integers(-2**31, 2**31 - 1, big), rng.integers(0, 50_000, big) * 40_000], 1).astype(np.int32)
        run_both(dim3, ctx, nctx, torch, R, S, hash_only=True)
        assert nctx.device_bytes() == 0
        say("noalloc: sizes beyond the limit are refused")
        t = torch.zeros((4, 2), dtype=torch.int32, device="cuda")
        rc, _ = nctx.join_project_dev_noalloc(t.data_ptr(), big + 1, t.data_ptr(), 4, 0, 0, output_mode=1)
        assert rc == dim3.ERR_INVALID
        assert "2^17" in nctx.last_error() or "131072" in nctx.last_error()
    finally:
        nctx.close()
