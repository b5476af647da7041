"""GPU parity at the sizes of the two single-query Zipf workloads.

  1M x 1M join-project (a,c ~ U[0,1e5), b ~ Zipf(1.0) on [0,1e4)):
    the fingerprint and the materialised output (6.9e9 pairs, fingerprinted
    on the device) against the oracle's fingerprint of the full query.
  10M x 10M COUNT join-aggregate (a,c ~ U[0,2e6), b ~ Zipf(1.2) on [0,1e6)):
    a z slice (c = 0 mod 4096) against the reference's output, both as a
    fingerprint and as materialised rows; the full query (2.4e12 groups)
    against the sum of the fingerprints of its 16 z slices (rows of
    different z values are disjoint, so the fingerprint is additive), its
    group count against the join-project row count and its count sum
    against the join paths.
"""
import json
import os
import sys

import numpy as np
import pytest

from tests.golden import cases
from tests.golden.make_golden import sha

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _large():
    with open(os.path.join(HERE, "golden", "large.json")) as f:
        return json.load(f)


def _s64(c):
    """A u64 constant as the int64 torch multiplies with (two's complement)."""
    return c - (1 << 64) if c >= (1 << 63) else c


def _mix64(k):
    # splitmix64 finaliser in int64 arithmetic: logical shifts by masking
    k = k ^ ((k >> 30) & ((1 << 34) - 1))
    k = k * _s64(0xBF58476D1CE4E5B9)
    k = k ^ ((k >> 27) & ((1 << 37) - 1))
    k = k * _s64(0x94D049BB133111EB)
    return k ^ ((k >> 31) & ((1 << 33) - 1))


def _mix64_int(k):
    """splitmix64 finaliser on a Python int (mod 2^64)."""
    k ^= k >> 30
    k = (k * 0xBF58476D1CE4E5B9) & M64
    k ^= k >> 27
    k = (k * 0x94D049BB133111EB) & M64
    return k ^ (k >> 31)


def _fmix32(h):
    h = h ^ (h >> 16)
    h = (h * 0x85EBCA6B) & 0xFFFFFFFF
    h = h ^ (h >> 13)
    h = (h * 0xC2B2AE35) & 0xFFFFFFFF
    return h ^ (h >> 16)


def _sum_mod64(v):
    """Exact sum mod 2^64 of an int64 tensor (halves summed separately)."""
    lo = int((v & 0xFFFFFFFF).sum().item())
    hi = int(((v >> 32) & 0xFFFFFFFF).sum().item())
    return (lo + (hi << 32)) & M64


def device_fingerprint(torch, xz, counts=None, chunk=1 << 26):
    """The row fingerprint (count, sum, sum2, csum) of device rows: xz an
    (n, 2) int32 tensor, counts an int64 tensor or None (join-project).
    A row adds k1(x) g(z) to sum and k2(x) c g(z) to sum2 (mod 2^64), with
    g(z) = a(z >> 5) c(z & 31) (word factor times bit-position factor)."""
    n = int(xz.shape[0])
    s1 = s2 = cs = 0
    pos = torch.tensor([_s64(_mix64_int((0x9E3779B97F4A7C15 * (b + 1)) & M64) | 1) for b in range(32)],
                       dtype=torch.int64, device=xz.device)
    for i in range(0, n, chunk):
        x = xz[i:i + chunk, 0].to(torch.int64) & 0xFFFFFFFF
        z = xz[i:i + chunk, 1].to(torch.int64) & 0xFFFFFFFF
        g = (_fmix32((z >> 5) ^ 0x243F6A88) | 1) * pos[z & 31]
        k1 = _mix64(x ^ _s64(0x243F6A8885A308D3)) | 1
        k2 = _mix64(x ^ _s64(0x13198A2E03707344)) | 1
        s1 = (s1 + _sum_mod64(k1 * g)) & M64
        if counts is None:
            s2 = (s2 + _sum_mod64(k2 * g)) & M64
        else:
            c = counts[i:i + chunk]
            s2 = (s2 + _sum_mod64(k2 * c * g)) & M64
            cs = (cs + _sum_mod64(c)) & M64
        del x, z, g, k1, k2
    return (n, s1, s2, cs)


def test_device_fingerprint_matches_oracle(torch):
    """The device fingerprint above against the oracle's on random rows."""
    import oracle

    rng = np.random.default_rng(3)
    xz = rng.integers(-2**31, 2**31, size=(5000, 2), dtype=np.int64).astype(np.int32)
    cnt = rng.integers(1, 1000, size=5000).astype(np.int64)
    d = torch.from_numpy(xz).cuda()
    dc = torch.from_numpy(cnt).cuda()
    assert device_fingerprint(torch, d, chunk=999) == oracle.fingerprint(xz).as_tuple()
    assert device_fingerprint(torch, d, dc, chunk=1024) == oracle.fingerprint(xz, cnt).as_tuple()


def _inputs(name):
    rec = _large()[name]
    spec = rec["spec"]
    R, S = cases.make_large_full(spec)
    Rs, Ss = cases.slice_large(spec, R, S)
    assert sha(Rs, Ss) == rec["sha256"], "generator drift"
    return rec, R, S, Rs, Ss


def test_project_full(ctx, dim3, torch):
    rec, R, S, _, _ = _inputs("project_full")
    want = tuple(int(v) for v in rec["project"]["fp"])
    say("project_full fingerprint", R.shape, S.shape)
    g = ctx.join_project(R, S, output_mode=1)
    say("stats", g.stats.as_dict())
    assert g.fp.as_tuple() == want
    # materialised in HBM, fingerprinted on the device
    dR = torch.from_numpy(R).cuda()
    dS = torch.from_numpy(S).cuda()
    rc, r0 = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0)
    assert rc == dim3.ERR_CAPACITY and r0.n == want[0]
    out = torch.empty((r0.n, 2), dtype=torch.int32, device="cuda")
    say("project_full materialise", r0.n, "rows")
    rc, r = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], out.data_ptr(), r0.n)
    assert rc == dim3.OK and r.n == want[0]
    torch.cuda.synchronize()
    got = device_fingerprint(torch, out)
    assert got[:3] == want[:3]
    del out


def test_aggregate_zslice(ctx, dim3, torch):
    rec, _, _, R, Ss = _inputs("aggregate_zslice")
    want = tuple(int(v) for v in rec["aggregate"]["fp"])
    say("aggregate_zslice", R.shape, Ss.shape)
    g = ctx.join_aggregate_count(R, Ss, output_mode=1)
    say("stats", g.stats.as_dict())
    assert g.fp.as_tuple() == want
    # materialised COUNT rows (per-group counters), fingerprinted on the device
    dR = torch.from_numpy(R).cuda()
    dS = torch.from_numpy(Ss).cuda()
    rc, r0 = ctx.join_aggregate_count_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), Ss.shape[0], 0, 0)
    assert rc == dim3.ERR_CAPACITY and r0.n == want[0]
    rows = torch.empty((r0.n, 4), dtype=torch.int32, device="cuda")  # (x, z, count int64)
    rc, r = ctx.join_aggregate_count_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), Ss.shape[0], rows.data_ptr(),
                                         r0.n)
    assert rc == dim3.OK and r.n == want[0]
    torch.cuda.synchronize()
    counts = rows[:, 2:4].contiguous().view(torch.int64).reshape(-1)
    assert device_fingerprint(torch, rows[:, 0:2], counts) == want


def test_aggregate_full_slices(ctx, dim3, torch):
    rec, R, S, _, _ = _inputs("aggregate_zslice")
    dR = torch.from_numpy(R).cuda()
    dS = torch.from_numpy(S).cuda()
    say("aggregate full", R.shape, S.shape)
    rc, full = ctx.join_aggregate_count_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, output_mode=1)
    assert rc == dim3.OK
    say("full", full.fp.as_tuple(), full.stats.as_dict())
    assert full.fp.csum == full.stats.out_j
    rc, proj = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, output_mode=1)
    assert rc == dim3.OK
    assert proj.n == full.n and proj.fp.sum == full.fp.sum
    acc = [0, 0, 0, 0]
    mod = 16
    for r in range(mod):
        sl = dS[(dS[:, 0] % mod) == r].contiguous()
        rc, part = ctx.join_aggregate_count_dev(dR.data_ptr(), R.shape[0], sl.data_ptr(), sl.shape[0], 0, 0,
                                                output_mode=1)
        assert rc == dim3.OK and not part.swapped
        acc = [(a + v) & M64 for a, v in zip(acc, part.fp.as_tuple())]
        say("slice", r, part.n)
    assert tuple(acc) == full.fp.as_tuple()
