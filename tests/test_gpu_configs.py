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

import oracle
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


def test_aggregate_full_group_counts(ctx, dim3, torch):
    """The whole 10M x 10M COUNT with every group's count formed (MFMA box
    over the heavy keys, light (x, z-tile) tasks) against the bitset oracle's
    fingerprint of all 2.36e12 groups, and against the path-sum shortcut."""
    rec = _large()["aggregate_full"]
    want = tuple(int(v) for v in rec["aggregate"]["fp"])
    R, S = cases.make_large_full(rec["spec"])
    assert sha(R, S) == rec["sha256"], "generator drift"
    dR = torch.from_numpy(R).cuda()
    dS = torch.from_numpy(S).cuda()
    torch.cuda.synchronize()
    say("aggregate full, group counts", R.shape, S.shape)
    rc, g = ctx.join_aggregate_count_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, output_mode=1)
    assert rc == dim3.OK
    say("stats", g.stats.as_dict())
    assert g.stats.dense_mode == dim3.DENSE_MFMA and g.stats.paths_sparse > 0
    assert g.fp.as_tuple() == want
    rc, p = ctx.join_aggregate_count_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, output_mode=1,
                                         path_sums=True)
    assert rc == dim3.OK and p.fp.as_tuple() == want


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
        torch.cuda.synchronize()  # torch's gather runs on its own stream, the query on the context's
        rc, part = ctx.join_aggregate_count_dev(dR.data_ptr(), R.shape[0], sl.data_ptr(), sl.shape[0], 0, 0,
                                                output_mode=1)
        assert rc == dim3.OK and not part.swapped
        acc = [(a + v) & M64 for a, v in zip(acc, part.fp.as_tuple())]
        say("slice", r, part.n)
    assert tuple(acc) == full.fp.as_tuple()


def _row_terms(torch, x, z, pos):
    """Per-row fingerprint terms k1(x) g(z) and k2(x) g(z) (int64, mod 2^64)."""
    x = x.to(torch.int64) & 0xFFFFFFFF
    z = z.to(torch.int64) & 0xFFFFFFFF
    g = (_fmix32((z >> 5) ^ 0x243F6A88) | 1) * pos[z & 31]
    k1 = _mix64(x ^ _s64(0x243F6A8885A308D3)) | 1
    k2 = _mix64(x ^ _s64(0x13198A2E03707344)) | 1
    return x, k1 * g, k2 * g


def test_graph_full(ctx, dim3, torch):
    """The headline workload at full size: the two-hop self-join of the
    Chung-Lu graph (1M vertices, 4M undirected edges, u != w), 15.1e9 rows.
    Fingerprint mode, and the output materialised in HBM and fingerprinted on
    the device, against the oracle's fingerprint of the whole query; the rows
    of two vertex slices (u % 256 = 0, 1) of that same output against the
    compiled reference's rows for those slices; and the slices queried on
    their own (R restricted to the slice, no swap)."""
    rec = _large()["graph_full"]
    spec = rec["spec"]
    E = cases.graph_full()
    assert sha(E, E) == rec["sha256"], "generator drift"
    want = tuple(int(v) for v in rec["project"]["fp"])
    dE = torch.from_numpy(E).cuda()
    n = E.shape[0]
    say("graph_full fingerprint mode")
    rc, f = ctx.join_project_dev(dE.data_ptr(), n, dE.data_ptr(), n, 0, 0, output_mode=1, exclude_equal=True)
    assert rc == dim3.OK and f.fp.as_tuple() == want, (f.fp.as_tuple(), want)
    for r in spec["slices"]:
        ws = tuple(int(v) for v in rec["slices"][str(r)]["fp"])
        sl = dE[(dE[:, 0] % spec["slice_mod"]) == r].contiguous()
        torch.cuda.synchronize()  # torch's gather runs on its own stream, the query on the context's
        rc, g = ctx.join_project_dev(sl.data_ptr(), sl.shape[0], dE.data_ptr(), n, 0, 0, output_mode=1,
                                     exclude_equal=True, no_swap=True)
        say("slice query", r, g.n)
        assert rc == dim3.OK and g.fp.as_tuple() == ws
    say("graph_full materialise", want[0], "rows")
    out = torch.empty((want[0], 2), dtype=torch.int32, device="cuda")
    rc, m = ctx.join_project_dev(dE.data_ptr(), n, dE.data_ptr(), n, out.data_ptr(), want[0], exclude_equal=True)
    assert rc == dim3.OK and m.n == want[0]
    torch.cuda.synchronize()
    pos = torch.tensor([_s64(_mix64_int((0x9E3779B97F4A7C15 * (b + 1)) & M64) | 1) for b in range(32)],
                       dtype=torch.int64, device="cuda")
    s1 = s2 = 0
    sl = {r: [0, 0, 0] for r in spec["slices"]}
    chunk = 1 << 27
    for i in range(0, want[0], chunk):
        x, t1, t2 = _row_terms(torch, out[i:i + chunk, 0], out[i:i + chunk, 1], pos)
        s1 = (s1 + _sum_mod64(t1)) & M64
        s2 = (s2 + _sum_mod64(t2)) & M64
        for r in spec["slices"]:
            keep = (x % spec["slice_mod"]) == r
            a = sl[r]
            a[0] += int(keep.sum().item())
            a[1] = (a[1] + _sum_mod64(t1[keep])) & M64
            a[2] = (a[2] + _sum_mod64(t2[keep])) & M64
        del x, t1, t2
    say("materialised fingerprint", s1, s2)
    assert (want[0], s1, s2) == want[:3]
    for r in spec["slices"]:
        ws = tuple(int(v) for v in rec["slices"][str(r)]["fp"])
        assert tuple(sl[r]) == ws[:3], (r, sl[r], ws)
    del out


def test_batch_largest_instance(ctx, dim3, torch):
    """The largest instance of the 256-instance batch (1.99M tuples per side,
    Zipf s = 1.47 over a 19,923-key domain, 3.97e10 distinct pairs) alone
    (join-project and COUNT fingerprints) and inside a batch with smaller
    instances of the same law, against the oracle."""
    rec = _large()["batch_largest"]
    R, S, idx, p = cases.batch_largest(rec["spec"])
    assert idx == rec["instance"] and sha(R, S) == rec["sha256"], "generator drift"
    want = tuple(int(v) for v in rec["project"]["fp"])
    wagg = tuple(int(v) for v in rec["aggregate"]["fp"])
    dR, dS = torch.from_numpy(R).cuda(), torch.from_numpy(S).cuda()
    say("batch largest alone", R.shape, S.shape)
    rc, g = ctx.join_project_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, output_mode=1)
    assert rc == dim3.OK and g.fp.as_tuple() == want, g.stats.as_dict()
    rc, c = ctx.join_aggregate_count_dev(dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0], 0, 0, output_mode=1)
    assert rc == dim3.OK and c.fp.as_tuple() == wagg
    from dim3_amd import workloads

    small = [workloads.batch_instance(q) for q in workloads.batch_instance_params(6, lo=20_000, hi=200_000, seed=77)]
    dev = [(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()) for a, b in small]
    inst = [(a.data_ptr(), a.shape[0], b.data_ptr(), b.shape[0]) for a, b in dev]
    inst.insert(3, (dR.data_ptr(), R.shape[0], dS.data_ptr(), S.shape[0]))
    say("batch with the largest instance")
    for count in (False, True):
        fn = ctx.batch_join_aggregate_count_dev if count else ctx.batch_join_project_dev
        rc, res = fn(inst, output_mode=1)
        assert rc == dim3.OK
        assert res[3].fp.as_tuple() == (wagg if count else want)
        for k, (a, b) in enumerate(small):
            j = k if k < 3 else k + 1
            fpk, _ = oracle.fingerprint_bitset(a, b, heavy=1024, count=count)
            assert res[j].fp.as_tuple() == fpk.as_tuple(), (count, j)
