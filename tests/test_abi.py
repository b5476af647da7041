"""The C-ABI library loads and exports every function include/dim3.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    names = set(re.findall(r"\b(dim3_[a-z0-9_]+)\s*\(", src))
    return names


def test_header_declares_the_entry_points():
    names = header_functions(os.path.join(ROOT, "include", "dim3.h"))
    for n in ("dim3_join_project", "dim3_join_aggregate_count", "dim3_ctx_create", "dim3_result_free"):
        assert n in names


def test_library_exports_every_declared_symbol(dim3):
    L = dim3.lib()
    for path in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not path.endswith(".h"):
            continue
        for name in header_functions(os.path.join(ROOT, "include", path)):
            assert hasattr(L, name), "%s declared in %s but not exported" % (name, path)
    assert set(dim3.EXPORTS) <= header_functions(os.path.join(ROOT, "include", "dim3.h"))


def test_version_and_status_strings(dim3):
    assert "gfx950" in dim3.version()
    assert dim3.status_string(dim3.ERR_CAPACITY) == "output capacity too small"


def test_struct_layouts_match_header(dim3):
    # dim3_count_row mirrors TUPLE<int,int,long long> (16 bytes)
    assert ctypes.sizeof(dim3.CountRow) == 16
    assert dim3.COUNT_ROW_DTYPE.itemsize == 16
    assert ctypes.sizeof(dim3.Fingerprint) == 32
    assert ctypes.sizeof(dim3.Opts) == 40


def _gcc_print(tmp_path, body, lang="c"):
    """Compiles and runs a small host program against include/ (no GPU calls)."""
    import subprocess

    src = tmp_path / ("probe." + lang)
    src.write_text(body)
    exe = tmp_path / "probe"
    cc = ["gcc", "-std=c11"] if lang == "c" else ["g++", "-std=c++14"]
    subprocess.run(cc + ["-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()


def test_result_and_row_layouts_match_the_c_compiler(dim3, tmp_path):
    out = _gcc_print(tmp_path, r'''
#include <stddef.h>
#include <stdio.h>
#include "dim3.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu\n", sizeof(dim3_result), offsetof(dim3_result, rows_f32),
           offsetof(dim3_result, stats), sizeof(dim3_value_row_f32), sizeof(dim3_stats), sizeof(dim3_opts));
    return 0;
}
''')
    size, off_f32, off_stats, row_f32, stats, opts = map(int, out)
    assert ctypes.sizeof(dim3.Result) == size
    assert dim3.Result.rows_f32.offset == off_f32 and dim3.Result.stats.offset == off_stats
    assert row_f32 == 12 == dim3.F32_ROW_DTYPE.itemsize
    assert ctypes.sizeof(dim3.Stats) == stats and ctypes.sizeof(dim3.Opts) == opts


def test_cpp_header_accepts_every_policy(tmp_path):
    """include/dim3.hpp compiles for the reference's driver shapes and every
    supported policy / value type (syntax only: nothing runs)."""
    import subprocess

    src = tmp_path / "shim.cpp"
    src.write_text(r'''
#include "dim3.hpp"
template <typename V, typename A>
void agg(myvector<TUPLE<int, int, V>>& R, myvector<TUPLE<int, int, V>>& S, myvector<TUPLE<int, int, V>>& out) {
    DIM3JoinAggregate<int, int, int, V, A, myHasher, myHasher, myHasher, myHasher, uint32_t>::doJoinAggregate(R, S, out);
}
int main() {
    myvector<TUPLE<int, int, float>> Rf, Sf, of;
    gen_rand_data(Rf, Sf, 10, 10);
    agg<float, AggregateFunctionAvg<float>>(Rf, Sf, of);
    agg<float, AggregateFunctionSum<float>>(Rf, Sf, of);
    agg<float, AggregateFunctionMin<float>>(Rf, Sf, of);
    agg<float, AggregateFunctionMax<float>>(Rf, Sf, of);
    myvector<TUPLE<int, int, long long>> Rl, Sl, ol;
    gen_rand_data(Rl, Sl, 10, 10);
    agg<long long, AggregateFunctionAvg<long long>>(Rl, Sl, ol);
    agg<long long, AggregateFunctionCount<long long>>(Rl, Sl, ol);
    myvector<TUPLE<int, int, int>> Ri, Si, oi;
    agg<int, AggregateFunctionSum<int>>(Ri, Si, oi);
    myvector<pair<int, int>> R, S, o;
    gen_rand_data(R, S, 10, 10);
    DIM3<int, int, int, myHasher, myHasher, myHasher, myHasher, uint32_t>::doJoinProject(R, S, o);
    long long est = EstimateJoinCardinality<int, int, int, myHasher, uint32_t>::estimate(R, S);
    (void)est;
    return 0;
}
''')
    subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                   check=True)
