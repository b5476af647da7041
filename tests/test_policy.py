"""Caller-written aggregate policies (the reference's policy concept,
dim3.hpp:1453-1480, used by SpmmAggregate at dim3.hpp:1275-1307).

dim3_join_paths_grouped returns the join paths grouped by (x, z), checked here
against a brute-force nested-loop join (small inputs); tests/policy_driver.cpp
runs three caller-written policies (one order-sensitive, one over double
values) through DIM3JoinAggregate and checks each result against the same
policy folded over a brute-force join."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "joinproject-dim3_amd")


def build_driver(tmp_path):
    if not os.path.isfile(os.path.join(LIBDIR, "libdim3.so")):
        pytest.skip("libdim3.so not built")
    exe = str(tmp_path / "policy_driver")
    cmd = ["g++", "-std=c++14", "-O2", "-I" + os.path.join(ROOT, "include"), "-o", exe,
           os.path.join(ROOT, "tests", "policy_driver.cpp"), "-L" + LIBDIR, "-ldim3", "-Wl,-rpath," + LIBDIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_policy_driver_builds(tmp_path):
    """A caller-written policy compiles against the drop-in header and binds
    the grouped-paths entry point."""
    exe = build_driver(tmp_path)
    syms = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True, check=True).stdout
    assert "dim3_join_paths_grouped" in syms and "dim3_join_aggregate_tuples" in syms


@pytest.mark.gpu
def test_policy_driver_runs(tmp_path):
    exe = build_driver(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr


def brute_grouped(R, S, exclude_equal=False):
    """Groups sorted by (x, z) with their paths' (v1, v2) in (R index, S index)
    order, and the dense z key: the rank of z among the distinct S.x."""
    paths = {}
    for i in range(R.shape[0]):
        for j in np.nonzero(S["y"] == R["y"][i])[0]:
            x, z = int(R["x"][i]), int(S["x"][j])
            if exclude_equal and x == z:
                continue
            paths.setdefault((x, z), []).append((R["v"][i], S["v"][j]))
    uz = np.unique(S["x"])
    keys = sorted(paths)
    start = np.zeros(len(keys) + 1, np.uint64)
    v1, v2 = [], []
    for k, key in enumerate(keys):
        start[k + 1] = start[k] + len(paths[key])
        v1 += [p[0] for p in paths[key]]
        v2 += [p[1] for p in paths[key]]
    return {"x": np.array([k[0] for k in keys], np.int32), "z": np.array([k[1] for k in keys], np.int32),
            "zkey": np.searchsorted(uz, [k[1] for k in keys]).astype(np.uint32), "path_start": start,
            "v1": np.array(v1, R.dtype["v"]), "v2": np.array(v2, R.dtype["v"]), "nz": int(uz.shape[0])}


def tuples(rng, dtype, n, dom, keys, lo):
    t = np.zeros(n, dtype)
    t["x"] = rng.integers(lo, lo + dom, n)
    t["y"] = rng.integers(0, keys, n)
    t["v"] = rng.integers(-50, 50, n)
    return t


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["i64", "f32", "self_exclude", "empty", "no_match"])
def test_grouped_paths_vs_bruteforce(ctx, dim3, case):
    rng = np.random.default_rng(["i64", "f32", "self_exclude", "empty", "no_match"].index(case) + 11)
    dt = dim3.TUPLE_F32 if case == "f32" else dim3.TUPLE_I64
    if case == "empty":
        R, S = tuples(rng, dt, 0, 10, 5, 0), tuples(rng, dt, 40, 10, 5, 0)
    elif case == "no_match":
        R, S = tuples(rng, dt, 50, 10, 5, 0), tuples(rng, dt, 50, 10, 5, 0)
        S["y"] += 100
    elif case == "self_exclude":
        R = tuples(rng, dt, 600, 60, 30, -30)
        S = R.copy()
    else:
        R, S = tuples(rng, dt, 700, 80, 40, -40), tuples(rng, dt, 900, 70, 40, -20)
    ex = case == "self_exclude"
    got = ctx.join_paths_grouped(R, S, exclude_equal=ex)
    want = brute_grouped(R, S, exclude_equal=ex)
    print(case, "groups", want["x"].shape[0], "paths", want["v1"].shape[0], flush=True)
    for k in ("x", "z", "zkey", "path_start", "v1", "v2"):
        assert np.array_equal(got[k], want[k]), k
    assert got["nz"] == want["nz"]
