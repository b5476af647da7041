"""Synthetic relations for the five benchmark workloads, as numpy int32 (n, 2)
arrays of (x|z, y) pairs with the join key in column 1.

All generators are seeded (numpy PCG64) and deterministic.  "Distinct" means
duplicate tuples are rejected and redrawn; the final tuple order is a seeded
permutation so that input order carries no structure.

  two_hop_graph     Chung-Lu power-law graph, both edge directions
                    (1M vertices, 4M undirected edges, exponent 2.1, cap 50k)
  zipf_relation     (uniform attribute, Zipf join key) relation
  project_small     R, S of 1M distinct tuples, a,c ~ U[0,1e5), b ~ Zipf(1.0) on [0,1e4)
  aggregate_large   R, S of 10M distinct tuples, a,c ~ U[0,2e6), b ~ Zipf(1.2) on [0,1e6)
  batch_instances   256 independent (R, S) with log-uniform sizes in [1e4, 2e6]
  served_stream     Poisson arrival stream of mixed project / COUNT requests
  gen_rand_tuples   the reference's gen_rand_data (dim3.hpp:1556-1591): the
                    std::minstd_rand0 draws from seed 492, (x, y) then (z, y),
                    tuple value = tuple index (join_aggregate.cpp's data)
"""
import math

import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def zipf_sampler(n, s):
    """Inverse-CDF sampler of Zipf(s) on [0, n): P(k) proportional to (k+1)^-s."""
    w = np.arange(1, n + 1, dtype=np.float64) ** (-float(s))
    cdf = np.cumsum(w)
    cdf /= cdf[-1]

    def draw(rng, m):
        u = rng.random(m)
        return np.minimum(np.searchsorted(cdf, u, side="right"), n - 1).astype(np.int64)

    return draw


def _pack(a, b):
    return (a.astype(np.int64) << 32) | (b.astype(np.int64) & 0xFFFFFFFF)


def _unpack(k):
    return np.stack([(k >> 32).astype(np.int32), (k & 0xFFFFFFFF).astype(np.uint32).astype(np.int32)], axis=1)


def distinct_relation(rng, n, draw_attr, draw_key):
    """n distinct (attr, key) tuples; duplicates are redrawn."""
    have = np.zeros(0, dtype=np.int64)
    need = n
    while need > 0:
        m = int(need * 1.1) + 1024
        cand = _pack(draw_attr(rng, m), draw_key(rng, m))
        have = np.unique(np.concatenate([have, cand]))
        need = n - have.shape[0]
    if have.shape[0] > n:
        have = rng.permutation(have)[:n]
    return _unpack(rng.permutation(have))


def zipf_relation(n, attr_domain, key_domain, s, seed):
    rng = _rng(seed)
    zk = zipf_sampler(key_domain, s)
    return distinct_relation(rng, n, lambda r, m: r.integers(0, attr_domain, m), zk)


def project_small(seed_r=1, seed_s=2, n=1_000_000, ac_domain=100_000, b_domain=10_000, s=1.0):
    """join-project of R(a,b), S(c,b): 1M distinct tuples per side, a,c uniform,
    b Zipf(1.0) on [0, 1e4)."""
    return (zipf_relation(n, ac_domain, b_domain, s, seed_r), zipf_relation(n, ac_domain, b_domain, s, seed_s))


def aggregate_large(seed_r=1, seed_s=2, n=10_000_000, ac_domain=2_000_000, b_domain=1_000_000, s=1.2):
    """COUNT join-aggregate: 10M distinct tuples per side, a,c ~ U[0,2e6),
    b ~ Zipf(1.2) on [0, 1e6)."""
    return (zipf_relation(n, ac_domain, b_domain, s, seed_r), zipf_relation(n, ac_domain, b_domain, s, seed_s))


def chung_lu_weights(n_vertices, n_directed, gamma, max_degree):
    """w_i = min(max_degree, C (i+1)^(-1/(gamma-1))) with C solved so that
    sum(w) = n_directed (bisection)."""
    i = np.arange(1, n_vertices + 1, dtype=np.float64)
    base = i ** (-1.0 / (gamma - 1.0))
    lo, hi = 0.0, float(n_directed)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if np.minimum(max_degree, mid * base).sum() < n_directed:
            lo = mid
        else:
            hi = mid
    return np.minimum(max_degree, hi * base)


def two_hop_graph(n_vertices=1_000_000, n_edges=4_000_000, gamma=2.1, max_degree=50_000, seed=7):
    """Chung-Lu graph: endpoints drawn proportional to the weights until
    n_edges distinct undirected non-loop edges; returned with both
    directions as (src, dst) pairs, so R = S = E for the two-hop self-join."""
    rng = _rng(seed)
    w = chung_lu_weights(n_vertices, 2 * n_edges, gamma, max_degree)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    draws = []
    keys_seen = np.zeros(0, dtype=np.int64)
    total = 0
    while True:
        m = int((n_edges - keys_seen.shape[0]) * 1.3) + 65536
        u = np.minimum(np.searchsorted(cdf, rng.random(m), side="right"), n_vertices - 1)
        v = np.minimum(np.searchsorted(cdf, rng.random(m), side="right"), n_vertices - 1)
        keep = u != v
        a = np.minimum(u, v)[keep].astype(np.int64)
        b = np.maximum(u, v)[keep].astype(np.int64)
        draws.append((a << 32) | b)
        total += a.shape[0]
        keys_seen = np.unique(np.concatenate([keys_seen, draws[-1]]))
        if keys_seen.shape[0] >= n_edges:
            break
    allk = np.concatenate(draws)
    uk, first = np.unique(allk, return_index=True)
    order = np.argsort(first, kind="stable")[:n_edges]  # first n_edges distinct edges in draw order
    ek = uk[order]
    src = (ek >> 32).astype(np.int32)
    dst = (ek & 0xFFFFFFFF).astype(np.int32)
    E = np.concatenate([np.stack([src, dst], 1), np.stack([dst, src], 1)])
    return E[rng.permutation(E.shape[0])]


def batch_instance_params(n_instances=256, lo=10_000, hi=2_000_000, seed=1000):
    """Per-instance sizes log-uniform in [lo, hi]; b domain 1% of the size
    with Zipf skew U[0.5, 1.5]; a, c domains 10% of the size."""
    rng = _rng(seed)
    out = []
    for i in range(n_instances):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        s = float(rng.uniform(0.5, 1.5))
        out.append({"n": n, "b_domain": max(1, n // 100), "ac_domain": max(1, n // 10), "s": s,
                    "seed_r": 100_000 + 2 * i, "seed_s": 100_001 + 2 * i})
    return out


def batch_instance(p):
    R = zipf_relation(p["n"], p["ac_domain"], p["b_domain"], p["s"], p["seed_r"])
    S = zipf_relation(p["n"], p["ac_domain"], p["b_domain"], p["s"], p["seed_s"])
    return R, S


def zipf_relation_device(torch, n, attr_domain, key_domain, s, seed, max_rounds=400):
    """zipf_relation drawn on the GPU (torch's Philox generator, seeded): the
    same law (uniform attribute, Zipf(s) key, duplicates redrawn), other
    random streams.  Returns an (n, 2) int32 device tensor."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    w = torch.arange(1, key_domain + 1, dtype=torch.float64, device="cuda") ** (-float(s))
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    have = torch.empty(0, dtype=torch.int64, device="cuda")
    need = n
    for _ in range(max_rounds):
        if need <= 0:
            break
        m = int(need * 1.1) + 1024
        a = torch.randint(0, attr_domain, (m,), device="cuda", generator=g, dtype=torch.int64)
        u = torch.rand(m, device="cuda", generator=g, dtype=torch.float64)
        b = torch.clamp(torch.searchsorted(cdf, u, right=True), max=key_domain - 1)
        have = torch.unique(torch.cat([have, (a << 32) | b]))
        need = n - have.numel()
    if have.numel() > n:
        have = have[torch.randperm(have.numel(), generator=g, device="cuda")[:n]]
    have = have[torch.randperm(have.numel(), generator=g, device="cuda")]
    return torch.stack([(have >> 32).to(torch.int32), (have & 0xFFFFFFFF).to(torch.int32)], 1).contiguous()


def batch_instance_device(torch, p):
    """batch_instance drawn on the GPU: (R, S) int32 device tensors."""
    R = zipf_relation_device(torch, p["n"], p["ac_domain"], p["b_domain"], p["s"], p["seed_r"])
    S = zipf_relation_device(torch, p["n"], p["ac_domain"], p["b_domain"], p["s"], p["seed_s"])
    return R, S


def planted_block_relation(n_background, block_keys, density, s, attr_domain, key_domain, seed):
    """Zipf(s) background plus a planted dense block: the first block_keys
    attribute values x the first block_keys join keys, each cell present with
    probability `density` (added on top of the background)."""
    rng = _rng(seed)
    bg = zipf_relation(n_background, attr_domain, key_domain, s, seed + 17)
    nb = int(block_keys * block_keys * density)
    cells = np.unique(rng.integers(0, block_keys * block_keys, int(nb * 1.2)))
    cells = rng.permutation(cells)[:nb]
    blk = np.stack([(cells // block_keys).astype(np.int32), (cells % block_keys).astype(np.int32)], 1)
    both = np.unique(_pack(np.concatenate([bg[:, 0], blk[:, 0]]), np.concatenate([bg[:, 1], blk[:, 1]])))
    return _unpack(rng.permutation(both))


def bernoulli_block(block_keys, density, seed, attr_offset=0, chunk_rows=2048):
    """Every (attr, key) cell of a block_keys x block_keys grid kept with
    probability `density` (distinct by construction), shuffled; int32."""
    rng = _rng(seed)
    parts = []
    for r0 in range(0, block_keys, chunk_rows):
        rows = min(chunk_rows, block_keys - r0)
        xi, yi = np.nonzero(rng.random((rows, block_keys), dtype=np.float32) < density)
        parts.append(np.stack([(xi + r0 + attr_offset).astype(np.int32), yi.astype(np.int32)], 1))
    B = np.concatenate(parts)
    return B[rng.permutation(B.shape[0])]


def served_stream(n_requests=200, rate_per_s=50.0, large_frac=0.05, agg_frac=0.2, seed=4242, large_n=None):
    """Poisson arrival times and request descriptors: 80% project / 20%
    COUNT, sizes from batch_instance_params, plus large_frac large requests
    (large_n tuples per side with a planted 20k x 20k block at 0.3 density)."""
    rng = _rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, n_requests))
    params = batch_instance_params(n_requests, seed=seed + 1)
    reqs = []
    for i in range(n_requests):
        kind = "aggregate" if rng.random() < agg_frac else "project"
        large = rng.random() < large_frac
        p = dict(params[i])
        if large:
            p.update({"large": True, "n": large_n or 50_000_000})
        reqs.append({"id": i, "t": float(t[i]), "kind": kind, "params": p})
    return reqs


def minstd_draws(seed, n):
    """The first n outputs of std::minstd_rand0 seeded with `seed`:
    x_k = seed * 16807^k mod (2^31 - 1), k = 1..n, from a table of powers
    (16807^(B i) and 16807^j for j < B), all products below 2^62."""
    m, a, B = 2147483647, 16807, 1 << 16
    k = np.arange(1, n + 1, dtype=np.int64)
    lo = np.empty(B, dtype=np.int64)
    lo[0] = 1
    for j in range(1, B):
        lo[j] = lo[j - 1] * a % m
    aB = lo[B - 1] * a % m
    hi = np.empty(n // B + 2, dtype=np.int64)
    hi[0] = 1
    for i in range(1, hi.shape[0]):
        hi[i] = hi[i - 1] * aB % m
    p = hi[k // B] * lo[k % B] % m
    return p * (seed % m) % m


def gen_rand_tuples(nR=10_000_000, nS=10_000_000, modx=1_000_000, mody=1_000_000, modz=1_000_000, seed=492):
    """R (x, y) and S (z, y) int32 pairs of gen_rand_data and their float32
    tuple values (the tuple index, as the TUPLE<int,int,float> overload)."""
    d = minstd_draws(seed, 2 * (nR + nS))
    R = np.stack([d[0:2 * nR:2] % modx, d[1:2 * nR:2] % mody], 1).astype(np.int32)
    S = np.stack([d[2 * nR::2] % modz, d[2 * nR + 1::2] % mody], 1).astype(np.int32)
    return R, np.arange(nR, dtype=np.float32), S, np.arange(nS, dtype=np.float32)
