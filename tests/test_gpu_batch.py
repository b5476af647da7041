"""GPU parity of the batched form: many instances in one persistent,
graph-captured launch must give, per instance, exactly the single-query
result (rows as sorted sets, fingerprints) and the CPU oracle's."""
import sys

import numpy as np
import pytest

import oracle
from tests.test_gpu_dense import planted

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(rng, n, amax, bmax, amin=0):
    return np.stack([rng.integers(amin, amax, n), rng.integers(0, bmax, n)], 1).astype(np.int32)


def mixed_instances(seed=0):
    rng = np.random.default_rng(seed)
    inst = []
    inst.append((rel(rng, 3000, 500, 200), rel(rng, 2500, 500, 200)))            # plain
    inst.append((np.zeros((0, 2), np.int32), rel(rng, 100, 50, 20)))             # empty R
    inst.append((rel(rng, 50, 30, 10), rel(rng, 400, 30, 10)))                   # |R| < |S|: swapped
    inst.append((rel(rng, 1, 5, 5), rel(rng, 1, 5, 5)))                          # one tuple each
    inst.append((rel(rng, 500, 100, 50), (rel(rng, 500, 100, 50) + [0, 1000]).astype(np.int32)))  # no matching key
    big = rel(rng, 4000, 10 ** 9, 300, amin=-10 ** 9)                            # hashed x / z
    inst.append((big, rel(rng, 3000, 10 ** 9, 300, amin=-10 ** 9)))
    inst.append((planted(3, 4000, 150, 0.4, 3000, 800), planted(4, 4000, 150, 0.4, 3000, 800)))  # dense block
    for _ in range(9):                                                           # small skewed
        n = int(rng.integers(10, 3000))
        inst.append((rel(rng, n, int(rng.integers(2, 400)), int(rng.integers(1, 60))),
                     rel(rng, int(rng.integers(10, 3000)), int(rng.integers(2, 400)), int(rng.integers(1, 60)))))
    return inst


def to_dev(torch, inst):
    dev = []
    for R, S in inst:
        R = np.ascontiguousarray(R, dtype=np.int32)
        S = np.ascontiguousarray(S, dtype=np.int32)
        dR = torch.from_numpy(R).cuda() if R.shape[0] else torch.zeros((1, 2), dtype=torch.int32,
                                                                                             device="cuda")
        dS = torch.from_numpy(S).cuda() if S.shape[0] else torch.zeros((1, 2), dtype=torch.int32,
                                                                                             device="cuda")
        dev.append((dR, R.shape[0], dS, S.shape[0]))
    return dev


@pytest.mark.parametrize("count", [False, True])
@pytest.mark.parametrize("exclude_equal", [False, True])
def test_batch_fingerprints_equal_single_and_oracle(ctx, count, exclude_equal):
    import torch

    inst = mixed_instances(1 + int(count) + 2 * int(exclude_equal))
    dev = to_dev(torch, inst)
    kw = dict(output_mode=1, exclude_equal=exclude_equal)
    fn = ctx.batch_join_aggregate_count_dev if count else ctx.batch_join_project_dev
    say("batch", "count" if count else "project", "exclude_equal", exclude_equal, len(inst), "instances")
    rc, res = fn([(d[0].data_ptr(), d[1], d[2].data_ptr(), d[3]) for d in dev], **kw)
    assert rc == 0
    for i, ((R, S), r) in enumerate(zip(inst, res)):
        if count:
            _, _, want, sw = oracle.join_aggregate_count(R, S, exclude_equal=exclude_equal, fingerprint_only=True)
            single = ctx.join_aggregate_count(R, S, **kw)
        else:
            _, want, sw = oracle.join_project(R, S, exclude_equal=exclude_equal, fingerprint_only=True)
            single = ctx.join_project(R, S, **kw)
        assert r.fp.as_tuple() == want.as_tuple(), i
        assert r.fp.as_tuple() == single.fp.as_tuple(), i
        assert r.swapped == sw == single.swapped, i
        assert r.stats.out_j == single.stats.out_j, i
    # a second batch reuses every bank and gives the same answer
    rc2, res2 = fn([(d[0].data_ptr(), d[1], d[2].data_ptr(), d[3]) for d in dev], **kw)
    assert rc2 == 0 and [r.fp.as_tuple() for r in res2] == [r.fp.as_tuple() for r in res]


@pytest.mark.parametrize("count", [False, True])
def test_batch_led_by_empty_instances(ctx, count):
    """A batch whose first instances have no join path (no matching key,
    empty R): the output mode is the batch's, not read off its first plan."""
    import torch

    inst = mixed_instances(21)
    inst = [inst[4], inst[1]] + inst
    dev = to_dev(torch, inst)
    fn = ctx.batch_join_aggregate_count_dev if count else ctx.batch_join_project_dev
    say("batch led by empty instances", "count" if count else "project", len(inst))
    rc, res = fn([(d[0].data_ptr(), d[1], d[2].data_ptr(), d[3]) for d in dev], output_mode=1)
    assert rc == 0
    for i, ((R, S), r) in enumerate(zip(inst, res)):
        if count:
            _, _, want, _ = oracle.join_aggregate_count(R, S, fingerprint_only=True)
        else:
            _, want, _ = oracle.join_project(R, S, fingerprint_only=True)
        assert r.fp.as_tuple() == want.as_tuple(), i


@pytest.mark.parametrize("count", [False, True])
def test_batch_materialized_rows_and_capacity(ctx, count):
    import torch

    inst = mixed_instances(7)
    dev = to_dev(torch, inst)
    wants = []
    for R, S in inst:
        if count:
            p, c, _, _ = oracle.join_aggregate_count(R, S)
            wants.append(oracle.sort_pairs(p, c))
        else:
            p, _, _ = oracle.join_project(R, S)
            wants.append(oracle.sort_pairs(p))
    width = 4 if count else 2  # int32 words per row
    outs = [torch.zeros((max(w[0].shape[0] if count else w.shape[0], 1), width), dtype=torch.int32, device="cuda")
            for w in wants]
    caps = [(w[0].shape[0] if count else w.shape[0]) for w in wants]
    fn = ctx.batch_join_aggregate_count_dev if count else ctx.batch_join_project_dev
    say("batch materialize", "count" if count else "project")
    rc, res = fn([(d[0].data_ptr(), d[1], d[2].data_ptr(), d[3], o.data_ptr(), cap)
                  for d, o, cap in zip(dev, outs, caps)])
    assert rc == 0
    import dim3_amd
    for i, (w, r, o) in enumerate(zip(wants, res, outs)):
        assert r.n == caps[i], i
        h = o.cpu().numpy()[: r.n]
        if count:
            rows = np.ascontiguousarray(h).view(dim3_amd.COUNT_ROW_DTYPE).reshape(-1)
            gp, gc = oracle.sort_pairs(np.stack([rows["x"], rows["z"]], 1), rows["count"])
            assert np.array_equal(gp, w[0]) and np.array_equal(gc, w[1]), i
        else:
            assert np.array_equal(oracle.sort_pairs(h), w), i
    # one instance with too small a buffer: status CAPACITY, its n = rows needed, the others complete
    big = int(np.argmax(caps))
    caps2 = list(caps)
    caps2[big] = caps[big] // 2
    rc, res = fn([(d[0].data_ptr(), d[1], d[2].data_ptr(), d[3], o.data_ptr(), cap)
                  for d, o, cap in zip(dev, outs, caps2)])
    assert rc == dim3_amd.ERR_CAPACITY
    assert res[big].n == caps[big]
    assert all(r.n == c for r, c in zip(res, caps))


def test_batch_of_many_uneven_instances(ctx, dim3):
    import torch

    from dim3_amd import workloads as wl
    params = wl.batch_instance_params(48, lo=2_000, hi=60_000, seed=77)
    inst = [wl.batch_instance(p) for p in params]
    dev = to_dev(torch, inst)
    say("uneven batch of", len(inst), "instances, sizes", [p["n"] for p in params][:8], "...")
    rc, res = ctx.batch_join_project_dev([(d[0].data_ptr(), d[1], d[2].data_ptr(), d[3]) for d in dev],
                                         output_mode=1)
    assert rc == 0
    for i, ((R, S), r) in enumerate(zip(inst, res)):
        _, want, _ = oracle.join_project(R, S, fingerprint_only=True)
        assert r.fp.as_tuple() == want.as_tuple(), i
