"""Writes the golden fixtures under tests/golden/ from the compiled reference
(oracle/_ref/libdim3ref.so, built by `make -C oracle ref` from the reference
sources).  Run in the build container:  python tests/golden/make_golden.py

small_*.npz   inputs R, S and the reference's sorted join-project pairs and
              sorted COUNT rows (DIM3 serial and DIM3JoinAggregate + COUNT).
values_*.npz  inputs R, vR, S, vS and the reference's sorted rows of the
              SUM / MIN / MAX / AVG value aggregates (DIM3JoinAggregate with
              value policies and its AggregateFunctionAvg<long long>).
values_f32_*.npz  the same on float values with exact float sums (its
              AggregateFunctionAvg<float> and float SUM / MIN / MAX).
medium.json   inputs regenerated from a named seeded generator (with a
              sha256 of the inputs to detect generator drift), and the count
              and fingerprint of the reference's output.
large.json    the same for the config-scale cases (python make_golden.py large,
              large_scale, aggregate_full).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tests.golden import cases  # noqa: E402


def sha(R, S):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(R, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(S, dtype=np.int32).tobytes())
    return h.hexdigest()


def small():
    for name, (R, S) in cases.small_cases().items():
        p = oracle.ref_join_project(R, S)
        cp, cc = oracle.ref_join_aggregate_count(R, S)
        p = oracle.sort_pairs(p)
        cp, cc = oracle.sort_pairs(cp, cc)
        np.savez_compressed(os.path.join(HERE, "small_%s.npz" % name), R=R, S=S, project=p, agg_pairs=cp,
                            agg_counts=cc, swapped=np.array([R.shape[0] < S.shape[0]]))
        print(name, R.shape[0], S.shape[0], "->", p.shape[0], "pairs,", cp.shape[0], "groups", flush=True)


def values():
    for name, (R, vR, S, vS) in cases.value_cases().items():
        rec = {}
        pairs = None
        for agg in ["sum", "min", "max", "avg"]:
            p, v = oracle.ref_join_aggregate(R, vR, S, vS, agg)
            p, v = oracle.sort_pairs(p, v)
            assert pairs is None or np.array_equal(pairs, p)
            pairs = p
            rec[agg] = v
        np.savez_compressed(os.path.join(HERE, "values_%s.npz" % name), R=R, vR=vR, S=S, vS=vS, pairs=pairs, **rec)
        print(name, R.shape[0], S.shape[0], "->", pairs.shape[0], "groups", flush=True)


def values_f32():
    for name, (R, vR, S, vS) in cases.value_cases_f32().items():
        rec = {}
        pairs = None
        for agg in ["sum", "min", "max", "avg"]:
            p, v = oracle.ref_join_aggregate_f32(R, vR, S, vS, agg)
            bits = v.view(np.int32).astype(np.int64)
            p, bits = oracle.sort_pairs(p, bits)
            assert pairs is None or np.array_equal(pairs, p)
            pairs = p
            rec[agg] = bits.astype(np.int32).view(np.float32)
        np.savez_compressed(os.path.join(HERE, "values_f32_%s.npz" % name), R=R, vR=vR, S=S, vS=vS, pairs=pairs,
                            **rec)
        print("f32", name, R.shape[0], S.shape[0], "->", pairs.shape[0], "groups", flush=True)


def medium():
    out = {}
    for name, spec in cases.MEDIUM.items():
        R, S = cases.make_medium(spec)
        rec = {"spec": spec, "sha256": sha(R, S), "n_r": int(R.shape[0]), "n_s": int(S.shape[0])}
        if spec.get("project", True):
            p = oracle.ref_join_project(R, S, parallel=True)
            fp = oracle.fingerprint(p)
            rec["project"] = {"n": int(p.shape[0]), "fp": [str(v) for v in fp.as_tuple()]}
            del p
        if spec.get("aggregate", False):
            cp, cc = oracle.ref_join_aggregate_count(R, S)
            fp = oracle.fingerprint(cp, cc)
            rec["aggregate"] = {"n": int(cp.shape[0]), "fp": [str(v) for v in fp.as_tuple()]}
            del cp, cc
        out[name] = rec
        print(name, json.dumps({k: v for k, v in rec.items() if k != "spec"}), flush=True)
    with open(os.path.join(HERE, "medium.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def large():
    """Config-scale fingerprints.  checker "oracle": the C restatement in
    fingerprint mode (no materialised output); "reference": the compiled
    reference, cross-checked by the oracle."""
    out = {}
    for name, spec in cases.LARGE.items():
        R, S = cases.slice_large(spec, *cases.make_large_full(spec))
        rec = {"spec": spec, "sha256": sha(R, S), "n_r": int(R.shape[0]), "n_s": int(S.shape[0])}
        if spec.get("project", True):
            _, fp, _ = oracle.join_project(R, S, fingerprint_only=True)
            n = fp.as_tuple()[0]
            if spec["checker"] == "reference":
                p = oracle.ref_join_project(R, S, parallel=True)
                assert oracle.fingerprint(p).as_tuple() == fp.as_tuple(), name
                del p
            rec["project"] = {"n": int(n), "fp": [str(v) for v in fp.as_tuple()]}
        if spec.get("aggregate", False):
            _, _, fp, _ = oracle.join_aggregate_count(R, S, fingerprint_only=True)
            n = fp.as_tuple()[0]
            if spec["checker"] == "reference":
                cp, cc = oracle.ref_join_aggregate_count(R, S)
                assert oracle.fingerprint(cp, cc).as_tuple() == fp.as_tuple(), name
                del cp, cc
            rec["aggregate"] = {"n": int(n), "fp": [str(v) for v in fp.as_tuple()]}
        out[name] = rec
        print(name, json.dumps({k: v for k, v in rec.items() if k != "spec"}), flush=True)
    with open(os.path.join(HERE, "large.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def large_scale():
    """Full-size headline graph and the largest batch instance, added to
    large.json: fingerprints by oracle.fingerprint_bitset (pinned against
    the SPA oracle and the compiled reference in tests/test_oracle.py), and
    the compiled reference's own rows on two vertex slices of the graph
    (src % 256 == r against all of S; the reference swaps, |R_r| < |S|, so
    its (S.first, R.first) rows are turned back to (u, w); u != w applied
    after it, as the reference has no such filter), each cross-checked by
    the bitset oracle on the same x slice."""
    path = os.path.join(HERE, "large.json")
    with open(path) as f:
        out = json.load(f)
    workers = min(8, os.cpu_count() or 1)
    spec = cases.GRAPH_FULL
    E = cases.graph_full()
    rec = {"spec": spec, "sha256": sha(E, E), "n_r": int(E.shape[0]), "n_s": int(E.shape[0])}
    fp, _ = oracle.fingerprint_bitset(E, E, exclude_equal=True, heavy=spec["heavy"], slices=64, workers=workers)
    rec["project"] = {"n": int(fp.count), "fp": [str(v) for v in fp.as_tuple()]}
    print("graph_full", rec["project"], flush=True)
    rec["slices"] = {}
    for r in spec["slices"]:
        Rr = E[(E[:, 0] % spec["slice_mod"]) == r]
        pairs = oracle.ref_join_project(Rr, E, parallel=True)
        if Rr.shape[0] < E.shape[0]:
            pairs = pairs[:, ::-1]
        pairs = np.ascontiguousarray(pairs[pairs[:, 0] != pairs[:, 1]])
        assert np.all(pairs[:, 0] % spec["slice_mod"] == r)
        fr = oracle.fingerprint(pairs).as_tuple()
        sl = _slice_fp(E, spec, r)
        assert sl == fr, (r, sl, fr)
        rec["slices"][str(r)] = {"n": int(fr[0]), "fp": [str(v) for v in fr], "checker": "reference"}
        print("graph slice", r, fr, flush=True)
        del pairs
    out["graph_full"] = rec
    bspec = cases.BATCH_LARGEST
    R, S, idx, p = cases.batch_largest(bspec)
    brec = {"spec": bspec, "instance": idx, "params": p, "sha256": sha(R, S), "n_r": int(R.shape[0]),
            "n_s": int(S.shape[0])}
    fp, sw = oracle.fingerprint_bitset(R, S, heavy=bspec["heavy"], slices=64, workers=workers)
    brec["project"] = {"n": int(fp.count), "fp": [str(v) for v in fp.as_tuple()], "swapped": sw}
    fpc, _ = oracle.fingerprint_bitset(R, S, heavy=bspec["heavy"], slices=64, workers=workers, count=True)
    brec["aggregate"] = {"n": int(fpc.count), "fp": [str(v) for v in fpc.as_tuple()]}
    print("batch_largest", idx, p["n"], brec["project"], brec["aggregate"], flush=True)
    out["batch_largest"] = brec
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def _slice_fp(E, spec, r):
    """The bitset oracle's fingerprint of the rows of x slice r alone."""
    fp, _ = oracle.fingerprint_bitset(E, E, exclude_equal=True, heavy=spec["heavy"], slices=spec["slice_mod"], only=r)
    return fp.as_tuple()


def aggregate_full():
    """The whole 10M x 10M Zipf(1.2) COUNT (the aggregate_zslice case's
    unsliced relations), added to large.json as "aggregate_full": its COUNT
    fingerprint by oracle.fingerprint_bitset (count=True; pinned against the
    SPA oracle's per-group counts in tests/test_oracle.py), 256 x slices over
    forked workers."""
    path = os.path.join(HERE, "large.json")
    with open(path) as f:
        out = json.load(f)
    workers = min(8, os.cpu_count() or 1)
    spec = dict(cases.LARGE["aggregate_zslice"])
    spec.pop("mod")
    spec.pop("r")
    spec["checker"] = "oracle_bitset"
    spec["heavy"] = 4096
    R, S = cases.make_large_full(cases.LARGE["aggregate_zslice"])
    rec = {"spec": spec, "sha256": sha(R, S), "n_r": int(R.shape[0]), "n_s": int(S.shape[0])}
    fp, sw = oracle.fingerprint_bitset(R, S, heavy=spec["heavy"], slices=256, workers=workers, count=True)
    rec["aggregate"] = {"n": int(fp.count), "fp": [str(v) for v in fp.as_tuple()], "swapped": sw}
    print("aggregate_full", rec["aggregate"], flush=True)
    out["aggregate_full"] = rec
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def served():
    """Per-request fingerprints of the served stream (served.json), by the
    bitset oracle (COUNT requests: the COUNT fingerprint)."""
    spec = cases.SERVED
    workers = min(8, os.cpu_count() or 1)
    out = {"spec": spec, "requests": []}
    for q, R, S in cases.served_requests(spec):
        count = q["kind"] == "aggregate"
        big = q["params"].get("large", False)
        fp, sw = oracle.fingerprint_bitset(R, S, heavy=spec["heavy"], slices=16 if big else 1,
                                           workers=workers if big else 1, count=count)
        rec = {"id": q["id"], "t": q["t"], "kind": q["kind"], "large": big, "n_r": int(R.shape[0]),
               "n_s": int(S.shape[0]), "sha256": sha(R, S), "fp": [str(v) for v in fp.as_tuple()], "swapped": sw}
        out["requests"].append(rec)
        print(json.dumps(rec), flush=True)
    with open(os.path.join(HERE, "served.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if not oracle.ref_available():
        sys.exit("reference checker missing: make -C oracle ref")
    which = sys.argv[1:] or ["small", "values", "medium"]
    if "small" in which:
        small()
    if "values" in which:
        values()
        values_f32()
    if "medium" in which:
        medium()
    if "large" in which:
        large()
    if "large_scale" in which:
        large_scale()
    if "served" in which:
        served()
    if "aggregate_full" in which:
        aggregate_full()
