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
