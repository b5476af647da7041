"""The served stream, replayed through the request server.

Twenty requests arrive as a Poisson stream (40 per second) in the
specification's mix: join-project (80%) and COUNT join-aggregate (20%),
sizes drawn from the batch size law (1e4 to 2e6 tuples per side, Zipf keys),
and two large requests, each a planted 20,000 x 20,000 key block at density
0.3 (1.2e8 tuples per side, ~7e11 join paths) on a Zipf(1.0) background
(reduced to 1e6 tuples).  The server runs them with four workers (the large
ones on the CU-partitioned lane, the others beside it, fingerprint requests
batched as they wait); every request's fingerprint must equal the oracle's
(tests/golden/served.json, written by make_golden.py served) and the same
request run alone on a plain context."""
import json
import os
import threading
import time

import numpy as np
import pytest

from tests.golden import cases
from tests.golden.make_golden import sha

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)


def test_served_stream_mix_equals_alone(ctx, dim3, torch):
    with open(os.path.join(HERE, "golden", "served.json")) as f:
        gold = json.load(f)
    t0 = time.perf_counter()
    reqs = cases.served_requests(gold["spec"])
    say("served stream: %d requests generated in %.1fs" % (len(reqs), time.perf_counter() - t0))
    assert len(reqs) == len(gold["requests"]) and sum(1 for q, _, _ in reqs if q["params"].get("large")) >= 2
    kinds = [q["kind"] for q, _, _ in reqs]
    assert kinds.count("aggregate") >= 2 and kinds.count("project") >= 10
    dev = []
    for (q, R, S), g in zip(reqs, gold["requests"]):
        assert g["id"] == q["id"] and g["sha256"] == sha(R, S), "generator drift at request %d" % q["id"]
        dev.append((torch.from_numpy(R).cuda(), torch.from_numpy(S).cuda()))
    torch.cuda.synchronize()
    results = [None] * len(reqs)
    with dim3.Server(0, workers=4, max_batch=8) as srv:
        hs = [(srv.put_dev(R.data_ptr(), int(R.shape[0])), srv.put_dev(S.data_ptr(), int(S.shape[0]))) for R, S in dev]
        tickets = [None] * len(reqs)
        errors = []

        def submitter():
            start = time.perf_counter()
            for i, (q, _, _) in enumerate(reqs):
                delay = q["t"] - (time.perf_counter() - start)
                if delay > 0:
                    time.sleep(delay)
                rc, t = srv.submit(hs[i][0], hs[i][1], count=q["kind"] == "aggregate", output_mode=1)
                if rc != dim3.OK:
                    errors.append((i, rc))
                tickets[i] = t

        th = threading.Thread(target=submitter)
        th.start()
        for i in range(len(reqs)):
            while tickets[i] is None and not errors:
                time.sleep(0.0005)
            assert not errors, errors
            rc, res, queued, total = srv.wait(tickets[i])
            assert rc == dim3.OK, (i, rc)
            results[i] = res
            say("request %d %s%s: %d rows, queued %.1f ms, total %.1f ms" % (
                i, reqs[i][0]["kind"], " (large)" if reqs[i][0]["params"].get("large") else "", res.n,
                1e3 * queued, 1e3 * total))
        th.join()
        st = srv.stats()
        say("server stats", st)
    for i, ((q, R, S), g) in enumerate(zip(reqs, gold["requests"])):
        want = tuple(int(v) for v in g["fp"])
        assert results[i].fp.as_tuple() == want, ("served vs oracle", i)
        dR, dS = dev[i]
        fn = ctx.join_aggregate_count_dev if q["kind"] == "aggregate" else ctx.join_project_dev
        rc, alone = fn(dR.data_ptr(), int(dR.shape[0]), dS.data_ptr(), int(dS.shape[0]), 0, 0, output_mode=1)
        assert rc == dim3.OK and alone.fp.as_tuple() == want, ("alone vs oracle", i)
        assert results[i].swapped == alone.swapped == g["swapped"]
