"""Inputs of the golden fixtures.  Shared by make_golden.py (which writes
the reference's outputs) and the tests (which regenerate the medium inputs
and check their sha256 before comparing)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _wl():
    import __graft_entry__ as GE

    GE.load_package()
    from dim3_amd import workloads

    return workloads


def _rand_pairs(rng, n, a_lo, a_hi, b_lo, b_hi):
    return np.stack([rng.integers(a_lo, a_hi, n), rng.integers(b_lo, b_hi, n)], 1).astype(np.int32)


def bernoulli_block(nx, ny, p, seed):
    """(x, y) for every cell of an nx-by-ny grid kept with probability p."""
    rng = np.random.default_rng(seed)
    m = rng.random((nx, ny)) < p
    x, y = np.nonzero(m)
    R = np.stack([x, y], 1).astype(np.int32)
    return R[rng.permutation(R.shape[0])]


def small_cases():
    import oracle

    wl = _wl()
    rng = np.random.default_rng(20240601)
    cases = {}
    # gen_rand_data shape (dim3.hpp:1556-1568), scaled down
    cases["uniform"] = oracle.gen_rand_data(5000, 5000, 1000, 1000, 1000, 492)
    # Zipf join key
    cases["zipf"] = (wl.zipf_relation(4000, 300, 400, 1.0, 11), wl.zipf_relation(4000, 300, 400, 1.0, 12))
    # |R| < |S|: the reference swaps and emits (S.first, R.first)
    cases["swap"] = (_rand_pairs(rng, 1500, 0, 400, 0, 150), _rand_pairs(rng, 3000, 0, 400, 0, 150))
    # duplicates, negative values and the key -1
    cases["dups_neg"] = (_rand_pairs(rng, 3000, -5, 60, -3, 40), _rand_pairs(rng, 2500, -7, 50, -3, 40))
    # a dense z block the reference's f2 model routes to its bitmap path
    cases["dense"] = (bernoulli_block(300, 1000, 0.1, 5), bernoulli_block(300, 1000, 0.1, 6))
    # one shared key: full cross product of the touched x and z
    cases["cross"] = (np.stack([np.arange(1200), np.full(1200, 7)], 1).astype(np.int32),
                      np.stack([np.arange(1000) * 3, np.full(1000, 7) + (np.arange(1000) >= 40) * 5], 1)
                      .astype(np.int32))
    return cases


def value_cases():
    """Inputs of the value aggregates: some small cases with int32 values per
    tuple (full int32 range, small ranges, negative sums)."""
    cs = small_cases()
    rng = np.random.default_rng(20240602)
    out = {}
    for name, lo, hi, nr in [("zipf", -(1 << 31), (1 << 31) - 1, 800), ("swap", -1000, 1000, 600),
                             ("dups_neg", -50, 10, 3000), ("cross", -(1 << 31), (1 << 31) - 1, 150)]:
        R, S = cs[name]
        R = R[:nr]
        vR = rng.integers(lo, hi, R.shape[0], endpoint=True).astype(np.int32)
        vS = rng.integers(lo, hi, S.shape[0], endpoint=True).astype(np.int32)
        out[name] = (R, vR, S, vS)
    return out


def value_cases_f32():
    """Float-valued inputs whose float sums are exact (small multiples of
    1/4), so the reference's first-touch float accumulation and a double
    accumulation agree bit for bit."""
    cs = small_cases()
    rng = np.random.default_rng(20240603)
    out = {}
    for name, nr in [("zipf", 800), ("dups_neg", 3000)]:
        R, S = cs[name]
        R = R[:nr]
        vR = (rng.integers(-4000, 4000, R.shape[0]) / 4).astype(np.float32)
        vS = (rng.integers(-4000, 4000, S.shape[0]) / 4).astype(np.float32)
        out[name] = (R, vR, S, vS)
    return out


# medium cases: count + fingerprint of the reference output
MEDIUM = {
    "kat": {"gen": "gen_rand_data", "nR": 10_000_000, "nS": 10_000_000, "mod": 1_000_000, "seed": 492,
            "aggregate": True},
    "project_scaled": {"gen": "project_small", "n": 100_000, "ac": 10_000, "b": 1_000, "s": 1.0,
                       "seed_r": 1, "seed_s": 2},
    "graph_scaled": {"gen": "two_hop_graph", "nv": 100_000, "ne": 400_000, "gamma": 2.1, "dmax": 5_000,
                     "seed": 7},
    "dense_block": {"gen": "bernoulli_block", "nx": 2000, "ny": 5000, "p": 0.05, "seed_r": 21, "seed_s": 22},
    "project_slice": {"gen": "project_small_slice", "n": 1_000_000, "ac": 100_000, "b": 10_000, "s": 1.0,
                      "seed_r": 1, "seed_s": 2, "a_hi": 2_000},
    "aggregate_scaled": {"gen": "aggregate_small", "n": 50_000, "ac": 10_000, "b": 5_000, "s": 1.2,
                         "seed_r": 3, "seed_s": 4, "project": False, "aggregate": True},
}


# config-scale cases: the full 1M x 1M join-project (fingerprint from the
# oracle: the reference would materialise 6.9e9 pairs) and a z slice of the
# 10M x 10M COUNT join-aggregate (reference and oracle)
LARGE = {
    "project_full": {"gen": "project_small", "n": 1_000_000, "ac": 100_000, "b": 10_000, "s": 1.0,
                     "seed_r": 1, "seed_s": 2, "checker": "oracle"},
    "aggregate_zslice": {"gen": "aggregate_zslice", "n": 10_000_000, "ac": 2_000_000, "b": 1_000_000, "s": 1.2,
                         "seed_r": 1, "seed_s": 2, "mod": 4096, "r": 0, "checker": "reference",
                         "project": False, "aggregate": True},
}


# config-scale cases checked through the bitset restatement
# (oracle.fingerprint_bitset): the two-hop graph of the headline benchmark
# with u != w (and the reference on two of its 256 vertex slices), and the
# largest instance of the 256-instance batch (project and COUNT)
GRAPH_FULL = {"gen": "two_hop_graph", "nv": 1_000_000, "ne": 4_000_000, "gamma": 2.1, "dmax": 50_000, "seed": 7,
              "exclude_equal": True, "slice_mod": 256, "slices": [0, 1], "heavy": 4096}
BATCH_LARGEST = {"gen": "batch_instance", "n_instances": 256, "seed": 1000, "rank": 0, "heavy": 4096}


def graph_full():
    wl = _wl()
    s = GRAPH_FULL
    return wl.two_hop_graph(s["nv"], s["ne"], s["gamma"], s["dmax"], s["seed"])


def batch_largest(spec=BATCH_LARGEST):
    """The instance of batch_instance_params(n_instances, seed) with the
    rank-th largest size (numpy generator), and its parameters."""
    wl = _wl()
    params = wl.batch_instance_params(spec["n_instances"], seed=spec["seed"])
    order = sorted(range(len(params)), key=lambda i: (-params[i]["n"], i))
    p = params[order[spec["rank"]]]
    R, S = wl.batch_instance(p)
    return R, S, order[spec["rank"]], p


# The served stream of the served-stream test: Poisson arrivals of
# join-project (80%) and COUNT (20%) requests with the batch size law, and
# the requests at positions `large_at` made large: a planted block_keys x
# block_keys block at `density` (keys 0..block_keys-1, the hottest of the
# background) added to a Zipf(1.0) background of large_bg tuples per side
# (the specification's 5e7 reduced; the block is the specified one).
SERVED = {"n_requests": 20, "seed": 4242, "rate_per_s": 40.0, "large_at": [4, 13], "large_bg": 1_000_000,
          "block_keys": 20_000, "density": 0.3, "heavy": 1024}


def served_large_relation(n_bg, block_keys, density, seed):
    """Zipf(1.0) background (attribute domain 10%, key domain 1% of n_bg)
    plus every (attr, key) cell of the block kept with probability
    `density`; the background's tuples inside the block that the block also
    holds are dropped (the union is distinct)."""
    wl = _wl()
    bg = wl.zipf_relation(n_bg, max(1, n_bg // 10), max(1, n_bg // 100), 1.0, seed + 17)
    rng = np.random.Generator(np.random.PCG64(seed))
    parts, mask = [], np.zeros((block_keys, block_keys), dtype=bool)
    for r0 in range(0, block_keys, 2048):
        rows = min(2048, block_keys - r0)
        m = rng.random((rows, block_keys), dtype=np.float32) < density
        mask[r0:r0 + rows] = m
        xi, yi = np.nonzero(m)
        parts.append(np.stack([(xi + r0).astype(np.int32), yi.astype(np.int32)], 1))
    inb = (bg[:, 0] < block_keys) & (bg[:, 1] < block_keys) & (bg[:, 0] >= 0) & (bg[:, 1] >= 0)
    dup = np.zeros(bg.shape[0], dtype=bool)
    dup[inb] = mask[bg[inb, 0], bg[inb, 1]]
    return np.concatenate(parts + [bg[~dup]])


def served_requests(spec=SERVED):
    """[(request, R, S)] of the served stream, request = the served_stream
    descriptor (id, arrival t at rate_per_s, kind, params)."""
    wl = _wl()
    reqs = wl.served_stream(spec["n_requests"], rate_per_s=spec["rate_per_s"], large_frac=0.0, seed=spec["seed"])
    out = []
    for q in reqs:
        p = q["params"]
        if q["id"] in spec["large_at"]:
            p = dict(p, large=True, n=spec["large_bg"])
            q = dict(q, params=p)
            R = served_large_relation(spec["large_bg"], spec["block_keys"], spec["density"], p["seed_r"])
            S = served_large_relation(spec["large_bg"], spec["block_keys"], spec["density"], p["seed_s"])
        else:
            R, S = wl.batch_instance(p)
        out.append((q, R, S))
    return out


def make_large_full(spec):
    """The unsliced relations of a LARGE case."""
    wl = _wl()
    if spec["gen"] == "project_small":
        return wl.project_small(spec["seed_r"], spec["seed_s"], spec["n"], spec["ac"], spec["b"], spec["s"])
    if spec["gen"] == "aggregate_zslice":
        return wl.aggregate_large(spec["seed_r"], spec["seed_s"], spec["n"], spec["ac"], spec["b"], spec["s"])
    raise ValueError(spec["gen"])


def slice_large(spec, R, S):
    """A LARGE case's inputs from its unsliced relations."""
    if spec["gen"] == "aggregate_zslice":
        return R, S[(S[:, 0] % spec["mod"]) == spec["r"]]
    return R, S


def make_medium(spec):
    g = spec["gen"]
    if g == "gen_rand_data":
        import oracle

        return oracle.gen_rand_data(spec["nR"], spec["nS"], spec["mod"], spec["mod"], spec["mod"], spec["seed"])
    wl = _wl()
    if g == "project_small":
        return wl.project_small(spec["seed_r"], spec["seed_s"], spec["n"], spec["ac"], spec["b"], spec["s"])
    if g == "project_small_slice":
        R, S = wl.project_small(spec["seed_r"], spec["seed_s"], spec["n"], spec["ac"], spec["b"], spec["s"])
        return R[R[:, 0] < spec["a_hi"]], S
    if g == "aggregate_small":
        return wl.aggregate_large(spec["seed_r"], spec["seed_s"], spec["n"], spec["ac"], spec["b"], spec["s"])
    if g == "two_hop_graph":
        E = wl.two_hop_graph(spec["nv"], spec["ne"], spec["gamma"], spec["dmax"], spec["seed"])
        return E, E
    if g == "bernoulli_block":
        return (bernoulli_block(spec["nx"], spec["ny"], spec["p"], spec["seed_r"]),
                bernoulli_block(spec["nx"], spec["ny"], spec["p"], spec["seed_s"]))
    raise ValueError(g)
