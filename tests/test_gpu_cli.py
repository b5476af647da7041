"""The reference's drivers rebuilt against the drop-in header
(joinproject-dim3_amd/cli/dim3_main.cpp over include/dim3.hpp): main.cpp,
main_parallel.cpp, join_aggregate.cpp (AVG of R.v + S.v over
TUPLE<int,int,float>) and the COUNT aggregate print the reference's output
contract on gen_rand_data's defaults, |Result|=98931829 (main.cpp:23-24,
join_aggregate.cpp:27-28)."""
import os
import re
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "joinproject-dim3_amd", "dim3_main")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin")


@pytest.mark.parametrize("mode", ["project", "parallel", "aggregate", "count"])
def test_reference_driver_output_contract(mode):
    if not os.path.exists(EXE):
        pytest.fail("dim3_main is not built (make -C joinproject-dim3_amd)")
    print("dim3_main", mode, flush=True)
    sys.stdout.flush()
    out = subprocess.run([EXE, mode], capture_output=True, text=True, timeout=240, check=True).stdout
    print(out, flush=True)
    lines = out.strip().splitlines()
    assert re.fullmatch(r"It takes [0-9.]+ seconds\.", lines[-2])
    assert lines[-1] == "|Result|=98931829"


@pytest.mark.parametrize("mode", ["estimate", "estimate_tuples"])
def test_estimate_mode_matches_oracle(mode):
    """EstimateJoinCardinality through the drop-in header on gen_rand_data's
    defaults (pairs, and the TUPLE<int,int,float> overload over the same
    draws) equals the oracle's estimate of the same data."""
    import oracle

    if not os.path.exists(EXE):
        pytest.fail("dim3_main is not built (make -C joinproject-dim3_amd)")
    print("dim3_main", mode, flush=True)
    out = subprocess.run([EXE, mode], capture_output=True, text=True, timeout=240, check=True).stdout
    print(out, flush=True)
    R, S = oracle.gen_rand_data()
    want = oracle.estimate(R, S) if R.shape[0] >= S.shape[0] else oracle.estimate(S, R)
    assert out.strip().splitlines()[-1] == "|Estimate|=%d" % want


@pytest.mark.parametrize("name", ["main", "main_parallel", "join_aggregate"])
def test_unchanged_reference_driver(name):
    """The reference's own main.cpp / main_parallel.cpp / join_aggregate.cpp,
    compiled unchanged against include/ by oracle/Makefile's dropin recipe
    (built in this tree where the reference sources are present), run on the
    GPU through libdim3 and print the reference's |Result|=98931829."""
    exe = os.path.join(DROPIN, name)
    if not os.path.exists(exe):
        pytest.skip("the reference drivers were not built here (no reference sources at build time)")
    print("reference driver", name, flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240, check=True).stdout
    print(out, flush=True)
    lines = out.strip().splitlines()
    assert re.fullmatch(r"It takes [0-9.]+ seconds\.", lines[-2])
    assert lines[-1] == "|Result|=98931829"
