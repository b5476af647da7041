"""Python binding of libdim3 (MI355X DIM3 join-project / join-aggregate).

Host-side plumbing over the C-ABI declared in include/dim3.h, used by the
tests, the benchmark and the server front end.  It mirrors the reference's
query entry points (dim3.hpp:982-1055 DIM3::doJoinProject, dim3.hpp:1313-1373
DIM3JoinAggregate::doJoinAggregate): relations are (n, 2) int32 arrays of
(x|z, y) pairs with the join key in column 1, results are (x, z) pairs or
(x, z, count) rows, and the |R| < |S| swap rule is kept (result.swapped).

The product path is the HIP library; there is no CPU fallback.  Loading
fails loudly when libdim3.so is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdim3.so")

(OK, ERR_INVALID, ERR_OOM, ERR_DEVICE, ERR_CAPACITY, ERR_ADMISSION, ERR_BUSY, ERR_NOT_FOUND, ERR_UNSUPPORTED,
 ERR_CANCELLED) = range(10)
OUT_MATERIALIZE, OUT_FINGERPRINT = 0, 1
OPT_NO_SWAP, OPT_EXCLUDE_EQUAL, OPT_HASH_ONLY, OPT_CLASSICAL, OPT_NO_CLASSICAL, OPT_PATH_SUMS = 1, 2, 4, 8, 16, 32
DENSE_AUTO, DENSE_OFF, DENSE_BITOR, DENSE_MFMA = -1, 0, 1, 2
AGG = {"sum": 1, "min": 2, "max": 3, "avg": 4}  # DIM3_AGG_*


class Opts(ctypes.Structure):
    _fields_ = [("output_mode", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("dense_mode", ctypes.c_int32),
                ("heavy_k", ctypes.c_int32), ("tile_bits", ctypes.c_int32), ("block", ctypes.c_int32),
                ("dense_core", ctypes.c_int32), ("reserved", ctypes.c_uint32 * 3)]


class Fingerprint(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("sum", ctypes.c_uint64), ("sum2", ctypes.c_uint64),
                ("csum", ctypes.c_uint64)]

    def as_tuple(self):
        return (int(self.count), int(self.sum), int(self.sum2), int(self.csum))


class Stats(ctypes.Structure):
    _fields_ = [("n_r", ctypes.c_uint64), ("n_s", ctypes.c_uint64), ("nx", ctypes.c_uint64),
                ("ny", ctypes.c_uint64), ("nz", ctypes.c_uint64), ("out_j", ctypes.c_uint64),
                ("paths_sparse", ctypes.c_uint64), ("heavy_keys", ctypes.c_uint64),
                ("nx_heavy", ctypes.c_uint64), ("nz_heavy", ctypes.c_uint64), ("dense_mode", ctypes.c_uint32),
                ("tile_bits", ctypes.c_uint32), ("n_tiles", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("ms_map", ctypes.c_double), ("ms_csr", ctypes.c_double), ("ms_dense", ctypes.c_double),
                ("ms_expand", ctypes.c_double), ("ms_total", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CountRow(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("z", ctypes.c_int32), ("count", ctypes.c_int64)]


class Result(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("pairs", ctypes.c_void_p), ("rows", ctypes.c_void_p),
                ("fp", Fingerprint), ("swapped", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("stats", Stats), ("rows_f32", ctypes.c_void_p)]


class Grouped(ctypes.Structure):
    _fields_ = [("n_groups", ctypes.c_uint64), ("n_paths", ctypes.c_uint64), ("nz", ctypes.c_uint32),
                ("value_size", ctypes.c_uint32), ("x", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("zkey", ctypes.c_void_p), ("path_start", ctypes.c_void_p), ("v1", ctypes.c_void_p),
                ("v2", ctypes.c_void_p)]


class Instance(ctypes.Structure):
    _fields_ = [("R_xy", ctypes.c_void_p), ("n_r", ctypes.c_uint64), ("S_zy", ctypes.c_void_p),
                ("n_s", ctypes.c_uint64), ("out", ctypes.c_void_p), ("cap", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    _fields_ = [("path_ns", ctypes.c_double), ("seed_bit_ns", ctypes.c_double), ("seed_word_ns", ctypes.c_double),
                ("mfma_tops", ctypes.c_double), ("bitor_gbs", ctypes.c_double), ("hbm_gbs", ctypes.c_double),
                ("source", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class ServerOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_uint64), ("workers", ctypes.c_uint32),
                ("max_batch", ctypes.c_uint32), ("queue_cap", ctypes.c_uint32), ("large_share_pct", ctypes.c_uint32),
                ("large_tuples", ctypes.c_uint64)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("agg", ctypes.c_uint32), ("r", ctypes.c_uint64),
                ("s", ctypes.c_uint64), ("opts", Opts)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("batched_requests", ctypes.c_uint64),
                ("resident_bytes", ctypes.c_uint64), ("workspace_bytes", ctypes.c_uint64),
                ("peak_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


COUNT_ROW_DTYPE = np.dtype([("x", "<i4"), ("z", "<i4"), ("count", "<i8")])
F32_ROW_DTYPE = np.dtype([("x", "<i4"), ("z", "<i4"), ("v", "<f4")])  # dim3_value_row_f32
# the reference's TUPLE<int,int,float> (12 bytes) and TUPLE<int,int,long long> (16 bytes)
TUPLE_F32 = np.dtype({"names": ["x", "y", "v"], "formats": ["<i4", "<i4", "<f4"], "offsets": [0, 4, 8],
                      "itemsize": 12})
TUPLE_I64 = np.dtype({"names": ["x", "y", "v"], "formats": ["<i4", "<i4", "<i8"], "offsets": [0, 4, 8],
                      "itemsize": 16})

# every symbol include/dim3.h declares (checked by the ABI test)
EXPORTS = [
    "dim3_ctx_create", "dim3_ctx_destroy", "dim3_last_error", "dim3_status_string", "dim3_ctx_device_bytes",
    "dim3_opts_default", "dim3_join_project", "dim3_join_aggregate_count", "dim3_result_free",
    "dim3_join_project_dev", "dim3_join_aggregate_count_dev", "dim3_count_join_paths", "dim3_dense_product",
    "dim3_batch_join_project_dev", "dim3_batch_join_aggregate_count_dev", "dim3_autotune", "dim3_get_tuning",
    "dim3_set_tuning", "dim3_ctx_release", "dim3_ctx_set_cu_share", "dim3_server_opts_default", "dim3_server_create", "dim3_server_destroy",
    "dim3_server_put_relation", "dim3_server_put_relation_dev", "dim3_server_drop_relation", "dim3_submit",
    "dim3_poll", "dim3_wait", "dim3_server_get_stats", "dim3_join_project_stream",
    "dim3_join_aggregate_count_stream", "dim3_version", "dim3_join_aggregate", "dim3_join_aggregate_dev",
    "dim3_join_aggregate_f32", "dim3_join_aggregate_f32_dev", "dim3_estimate_join_cardinality",
    "dim3_join_aggregate_tuples", "dim3_server_put_values", "dim3_join_aggregate_stream",
    "dim3_join_aggregate_f32_stream", "dim3_estimate_join_cardinality_dev", "dim3_join_paths_grouped",
    "dim3_grouped_free",
]

_lib = None

# int (*dim3_rows_cb)(void* user, const void* rows, uint64_t n)
ROWS_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64)


class Dim3Error(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("libdim3 status %d (%s): %s" % (status, status_string(status), msg))
        self.status = status


def lib():
    """Load libdim3.so; raises when the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libdim3.so is not built (%s); run __graft_entry__.build() or "
                          "make -C joinproject-dim3_amd" % LIB_PATH)
    # PyTorch-ROCm ships its own HIP runtime (file libamdhip64.so, soname
    # libamdhip64.so.7).  Loaded first, it satisfies libdim3's dependency on
    # that soname and the process keeps one HIP runtime; loaded second, it
    # would bring a second HSA runtime that cannot open the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    L.dim3_ctx_create.argtypes = [i32, u64, ctypes.POINTER(vp)]
    L.dim3_ctx_create.restype = i32
    L.dim3_ctx_destroy.argtypes = [vp]
    L.dim3_ctx_destroy.restype = i32
    L.dim3_last_error.argtypes = [vp]
    L.dim3_last_error.restype = ctypes.c_char_p
    L.dim3_status_string.argtypes = [i32]
    L.dim3_status_string.restype = ctypes.c_char_p
    L.dim3_ctx_device_bytes.argtypes = [vp]
    L.dim3_ctx_device_bytes.restype = u64
    L.dim3_opts_default.argtypes = [ctypes.POINTER(Opts)]
    L.dim3_opts_default.restype = None
    for name in ("dim3_join_project", "dim3_join_aggregate_count"):
        f = getattr(L, name)
        f.argtypes = [vp, vp, u64, vp, u64, ctypes.POINTER(Opts), ctypes.POINTER(Result)]
        f.restype = i32
    for name in ("dim3_join_project_dev", "dim3_join_aggregate_count_dev", "dim3_join_project_dev_noalloc"):
        f = getattr(L, name)
        f.argtypes = [vp, vp, u64, vp, u64, ctypes.POINTER(Opts), vp, u64, vp, ctypes.POINTER(Result)]
        f.restype = i32
    for name in ("dim3_join_project_stream", "dim3_join_aggregate_count_stream"):
        f = getattr(L, name)
        f.argtypes = [vp, vp, u64, vp, u64, ctypes.POINTER(Opts), u64, ROWS_CB, vp, ctypes.POINTER(Result)]
        f.restype = i32
    L.dim3_join_aggregate.argtypes = [vp, vp, vp, u64, vp, vp, u64, i32, ctypes.POINTER(Opts), ctypes.POINTER(Result)]
    L.dim3_join_aggregate.restype = i32
    L.dim3_join_aggregate_dev.argtypes = [vp, vp, vp, u64, vp, vp, u64, i32, ctypes.POINTER(Opts), vp, u64, vp,
                                          ctypes.POINTER(Result)]
    L.dim3_join_aggregate_dev.restype = i32
    L.dim3_join_aggregate_f32.argtypes = L.dim3_join_aggregate.argtypes
    L.dim3_join_aggregate_f32.restype = i32
    L.dim3_join_aggregate_f32_dev.argtypes = L.dim3_join_aggregate_dev.argtypes
    L.dim3_join_aggregate_f32_dev.restype = i32
    L.dim3_join_aggregate_tuples.argtypes = [vp, vp, u64, vp, u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                             i32, ctypes.POINTER(Opts), ctypes.POINTER(Result)]
    L.dim3_join_aggregate_tuples.restype = i32
    L.dim3_join_paths_grouped.argtypes = [vp, vp, u64, vp, u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.POINTER(Opts), ctypes.POINTER(Grouped)]
    L.dim3_join_paths_grouped.restype = i32
    L.dim3_grouped_free.argtypes = [ctypes.POINTER(Grouped)]
    L.dim3_grouped_free.restype = None
    for name in ("dim3_join_aggregate_stream", "dim3_join_aggregate_f32_stream"):
        f = getattr(L, name)
        f.argtypes = [vp, vp, vp, u64, vp, vp, u64, i32, ctypes.POINTER(Opts), u64, ROWS_CB, vp,
                      ctypes.POINTER(Result)]
        f.restype = i32
    L.dim3_result_free.argtypes = [ctypes.POINTER(Result)]
    L.dim3_result_free.restype = None
    L.dim3_estimate_join_cardinality.argtypes = [vp, vp, u64, vp, u64, ctypes.POINTER(ctypes.c_int64)]
    L.dim3_estimate_join_cardinality.restype = i32
    L.dim3_estimate_join_cardinality_dev.argtypes = [vp, vp, u64, vp, u64, vp, ctypes.POINTER(ctypes.c_int64)]
    L.dim3_estimate_join_cardinality_dev.restype = i32
    L.dim3_count_join_paths.argtypes = [vp, vp, u64, vp, u64, ctypes.POINTER(ctypes.c_uint64)]
    L.dim3_count_join_paths.restype = i32
    for f in (L.dim3_batch_join_project_dev, L.dim3_batch_join_aggregate_count_dev):
        f.argtypes = [vp, ctypes.POINTER(Instance), ctypes.c_uint32, ctypes.POINTER(Opts), vp,
                      ctypes.POINTER(Result)]
        f.restype = i32
    L.dim3_autotune.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(Tuning)]
    L.dim3_autotune.restype = i32
    L.dim3_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.dim3_get_tuning.restype = i32
    L.dim3_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.dim3_set_tuning.restype = i32
    L.dim3_ctx_set_cu_share.argtypes = [vp, ctypes.c_double]
    L.dim3_ctx_set_cu_share.restype = i32
    L.dim3_ctx_release.argtypes = [vp]
    L.dim3_ctx_release.restype = i32
    L.dim3_server_opts_default.argtypes = [ctypes.POINTER(ServerOpts)]
    L.dim3_server_opts_default.restype = None
    L.dim3_server_create.argtypes = [ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)]
    L.dim3_server_create.restype = i32
    L.dim3_server_destroy.argtypes = [vp]
    L.dim3_server_destroy.restype = i32
    for f in (L.dim3_server_put_relation, L.dim3_server_put_relation_dev):
        f.argtypes = [vp, vp, u64, ctypes.POINTER(ctypes.c_uint64)]
        f.restype = i32
    L.dim3_server_put_values.argtypes = [vp, u64, vp, ctypes.c_uint32]
    L.dim3_server_put_values.restype = i32
    L.dim3_server_drop_relation.argtypes = [vp, u64]
    L.dim3_server_drop_relation.restype = i32
    L.dim3_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    L.dim3_submit.restype = i32
    L.dim3_poll.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_int)]
    L.dim3_poll.restype = i32
    L.dim3_wait.argtypes = [vp, u64, ctypes.POINTER(Result), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double)]
    L.dim3_wait.restype = i32
    L.dim3_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.dim3_server_get_stats.restype = i32
    L.dim3_dense_product.argtypes = [vp, vp, u64, vp, u64, u64, vp, vp]
    L.dim3_dense_product.restype = i32
    L.dim3_version.argtypes = []
    L.dim3_version.restype = ctypes.c_char_p
    _lib = L
    return L


def status_string(s):
    try:
        return lib().dim3_status_string(s).decode()
    except ImportError:
        return str(s)


def make_opts(output_mode=OUT_MATERIALIZE, no_swap=False, exclude_equal=False, dense_mode=DENSE_AUTO, heavy_k=0,
              tile_bits=0, block=0, hash_only=False, dense_core=-1, classical=None, path_sums=False):
    """classical: None lets f1 choose (AUTO plans), True forces the classical
    path, False forbids it.  path_sums: a COUNT fingerprint from join-path
    sums instead of the groups' counts."""
    o = Opts()
    lib().dim3_opts_default(ctypes.byref(o))
    o.output_mode = output_mode
    o.flags = ((OPT_NO_SWAP if no_swap else 0) | (OPT_EXCLUDE_EQUAL if exclude_equal else 0)
               | (OPT_HASH_ONLY if hash_only else 0)
               | (OPT_CLASSICAL if classical is True else OPT_NO_CLASSICAL if classical is False else 0)
               | (OPT_PATH_SUMS if path_sums else 0))
    o.dense_mode = dense_mode
    o.heavy_k = heavy_k
    o.tile_bits = tile_bits
    o.block = block
    o.dense_core = dense_core
    return o


def _pairs(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a.reshape(-1, 2)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _values(v, n):
    v = np.ascontiguousarray(v, dtype=np.int32).reshape(-1)
    if v.shape[0] != n:
        raise ValueError("one value per tuple: %d values for %d tuples" % (v.shape[0], n))
    return v


def _fvalues(v, n):
    v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1)
    if v.shape[0] != n:
        raise ValueError("one value per tuple: %d values for %d tuples" % (v.shape[0], n))
    return v


def _agg(agg):
    """DIM3_AGG_* of a name ("sum", "min", "max", "avg") or a code."""
    return AGG[agg] if isinstance(agg, str) else int(agg)


class QueryResult:
    """Host-side copy of a dim3_result."""

    def __init__(self, res, pairs=None, rows=None):
        self.n = int(res.n)
        self.pairs = pairs
        self.rows = rows
        self.fp = Fingerprint.from_buffer_copy(res.fp)
        self.swapped = bool(res.swapped)
        self.stats = Stats.from_buffer_copy(res.stats)

    @property
    def counts(self):
        """The row values: counts, int64 aggregates or float32 aggregates."""
        return None if self.rows is None else self.rows[self.rows.dtype.names[2]]

    @property
    def xz(self):
        if self.pairs is not None:
            return self.pairs
        if self.rows is not None:
            return np.stack([self.rows["x"], self.rows["z"]], axis=1)
        return None


class Context:
    """One libdim3 context (one device, one HIP stream, one workspace)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.dim3_ctx_create(device, hbm_budget_bytes, ctypes.byref(h))
        if rc != OK:
            raise Dim3Error(rc, "dim3_ctx_create(device=%d)" % device)
        self._h = h
        self._L = L

    def close(self):
        if self._h is not None and self._h.value:
            self._L.dim3_ctx_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def last_error(self):
        return self._L.dim3_last_error(self._h).decode()

    def device_bytes(self):
        return int(self._L.dim3_ctx_device_bytes(self._h))

    def _check(self, rc, what):
        if rc != OK:
            raise Dim3Error(rc, "%s: %s" % (what, self.last_error()))

    def _host_call(self, fn, R, S, opts):
        R, S = _pairs(R), _pairs(S)
        res = Result()
        rc = fn(self._h, _ptr(R), R.shape[0], _ptr(S), S.shape[0], ctypes.byref(opts), ctypes.byref(res))
        return rc, res

    def join_project(self, R, S, **kw):
        """DIM3::doJoinProject: distinct (x, z) pairs as an (n, 2) int32 array."""
        opts = make_opts(**kw)
        rc, res = self._host_call(self._L.dim3_join_project, R, S, opts)
        self._check(rc, "dim3_join_project")
        pairs = None
        if opts.output_mode == OUT_MATERIALIZE:
            n = int(res.n)
            if n:
                buf = (ctypes.c_int32 * (2 * n)).from_address(res.pairs)
                pairs = np.frombuffer(buf, dtype=np.int32).reshape(n, 2).copy()
            else:
                pairs = np.zeros((0, 2), dtype=np.int32)
        out = QueryResult(res, pairs=pairs)
        self._L.dim3_result_free(ctypes.byref(res))
        return out

    def join_aggregate_count(self, R, S, **kw):
        """DIM3JoinAggregate with COUNT: rows with fields x, z, count."""
        opts = make_opts(**kw)
        rc, res = self._host_call(self._L.dim3_join_aggregate_count, R, S, opts)
        self._check(rc, "dim3_join_aggregate_count")
        rows = None
        if opts.output_mode == OUT_MATERIALIZE:
            n = int(res.n)
            if n:
                buf = (ctypes.c_char * (16 * n)).from_address(res.rows)
                rows = np.frombuffer(buf, dtype=COUNT_ROW_DTYPE).copy()
            else:
                rows = np.zeros((0,), dtype=COUNT_ROW_DTYPE)
        out = QueryResult(res, rows=rows)
        self._L.dim3_result_free(ctypes.byref(res))
        return out

    def join_aggregate(self, R, vR, S, vS, agg, **kw):
        """DIM3JoinAggregate with a value policy: agg ("sum", "min", "max",
        "avg") of vR + vS over the join paths of each (x, z); vR / vS hold one
        int32 value per tuple.  Rows have fields x, z, count (the int64
        aggregate)."""
        opts = make_opts(**kw)
        R, S = _pairs(R), _pairs(S)
        vR = _values(vR, R.shape[0])
        vS = _values(vS, S.shape[0])
        res = Result()
        rc = self._L.dim3_join_aggregate(self._h, _ptr(R), _ptr(vR), R.shape[0], _ptr(S), _ptr(vS), S.shape[0],
                                         _agg(agg), ctypes.byref(opts), ctypes.byref(res))
        self._check(rc, "dim3_join_aggregate")
        rows = None
        if opts.output_mode == OUT_MATERIALIZE:
            n = int(res.n)
            if n:
                buf = (ctypes.c_char * (16 * n)).from_address(res.rows)
                rows = np.frombuffer(buf, dtype=COUNT_ROW_DTYPE).copy()
            else:
                rows = np.zeros((0,), dtype=COUNT_ROW_DTYPE)
        out = QueryResult(res, rows=rows)
        self._L.dim3_result_free(ctypes.byref(res))
        return out

    def join_aggregate_f32(self, R, vR, S, vS, agg, **kw):
        """Value aggregate on float32 values (the reference's
        AggregateFunctionAvg<float> workload): rows with fields x, z, v
        (float32)."""
        opts = make_opts(**kw)
        R, S = _pairs(R), _pairs(S)
        vR = _fvalues(vR, R.shape[0])
        vS = _fvalues(vS, S.shape[0])
        res = Result()
        rc = self._L.dim3_join_aggregate_f32(self._h, _ptr(R), _ptr(vR), R.shape[0], _ptr(S), _ptr(vS), S.shape[0],
                                             _agg(agg), ctypes.byref(opts), ctypes.byref(res))
        self._check(rc, "dim3_join_aggregate_f32")
        rows = None
        if opts.output_mode == OUT_MATERIALIZE:
            n = int(res.n)
            if n:
                buf = (ctypes.c_char * (12 * n)).from_address(res.rows_f32)
                rows = np.frombuffer(buf, dtype=F32_ROW_DTYPE).copy()
            else:
                rows = np.zeros((0,), dtype=F32_ROW_DTYPE)
        out = QueryResult(res, rows=rows)
        self._L.dim3_result_free(ctypes.byref(res))
        return out

    def join_aggregate_f32_dev(self, d_R, d_vR, n_r, d_S, d_vS, n_s, agg, d_out, cap, stream=None, **kw):
        """Device-resident float value aggregate; rows are 12-byte
        (x, z, float) records."""
        opts = make_opts(**kw)
        res = Result()
        vp = ctypes.c_void_p
        rc = self._L.dim3_join_aggregate_f32_dev(self._h, vp(d_R), vp(d_vR), n_r, vp(d_S), vp(d_vS), n_s, _agg(agg),
                                                 ctypes.byref(opts), vp(d_out), cap, vp(stream or 0),
                                                 ctypes.byref(res))
        if rc not in (OK, ERR_CAPACITY):
            self._check(rc, "dim3_join_aggregate_f32_dev")
        return rc, QueryResult(res)

    def join_aggregate_tuples(self, R, S, agg, **kw):
        """Value aggregate (agg "sum" / "min" / "max" / "avg", or "count")
        over TUPLE arrays as the reference lays them out: structured arrays
        with int32 fields x (or z), y and a value field v (int32, int64 or
        float32) at any 4-byte-aligned offset >= 8, e.g. TUPLE_F32 /
        TUPLE_I64.  Rows: COUNT_ROW_DTYPE, or F32_ROW_DTYPE for float values."""
        opts = make_opts(**kw)
        if R.dtype != S.dtype:
            raise ValueError("R and S must share the tuple layout")
        vdt, voff = R.dtype.fields["v"][0], R.dtype.fields["v"][1]
        vtype = {np.dtype("<i4"): 0, np.dtype("<i8"): 1, np.dtype("<f4"): 2}[vdt]
        code = 0 if agg == "count" else _agg(agg)
        R, S = np.ascontiguousarray(R), np.ascontiguousarray(S)
        res = Result()
        rc = self._L.dim3_join_aggregate_tuples(self._h, _ptr(R), R.shape[0], _ptr(S), S.shape[0], R.dtype.itemsize,
                                                voff, vtype, code, ctypes.byref(opts), ctypes.byref(res))
        self._check(rc, "dim3_join_aggregate_tuples")
        f32 = vtype == 2 and code != 0
        rows = None
        if opts.output_mode == OUT_MATERIALIZE:
            n = int(res.n)
            dt, width, ptr = (F32_ROW_DTYPE, 12, res.rows_f32) if f32 else (COUNT_ROW_DTYPE, 16, res.rows)
            rows = (np.frombuffer((ctypes.c_char * (width * n)).from_address(ptr), dtype=dt).copy() if n
                    else np.zeros((0,), dtype=dt))
        out = QueryResult(res, rows=rows)
        self._L.dim3_result_free(ctypes.byref(res))
        return out

    def join_paths_grouped(self, R, S, **kw):
        """dim3_join_paths_grouped over TUPLE arrays (structured, int32 x / z
        and y, a 4- or 8-byte value field v): the join paths grouped by
        (x, z), for a caller-written aggregate policy.  Returns a dict of
        numpy arrays: x, z, zkey (n_groups), path_start (n_groups + 1), v1, v2
        (n_paths, the value dtype), and nz."""
        opts = make_opts(**kw)
        if R.dtype != S.dtype:
            raise ValueError("R and S must share the tuple layout")
        vdt, voff = R.dtype.fields["v"][0], R.dtype.fields["v"][1]
        R, S = np.ascontiguousarray(R), np.ascontiguousarray(S)
        g = Grouped()
        rc = self._L.dim3_join_paths_grouped(self._h, _ptr(R), R.shape[0], _ptr(S), S.shape[0], R.dtype.itemsize, voff,
                                             vdt.itemsize, ctypes.byref(opts), ctypes.byref(g))
        self._check(rc, "dim3_join_paths_grouped")
        G, P = int(g.n_groups), int(g.n_paths)

        def take(ptr, n, dt):
            dt = np.dtype(dt)
            return (np.frombuffer((ctypes.c_char * (dt.itemsize * n)).from_address(ptr), dtype=dt).copy() if n
                    else np.zeros((0,), dtype=dt))
        out = {"x": take(g.x, G, np.int32), "z": take(g.z, G, np.int32), "zkey": take(g.zkey, G, np.uint32),
               "path_start": take(g.path_start, G + 1, np.uint64), "v1": take(g.v1, P, vdt), "v2": take(g.v2, P, vdt),
               "nz": int(g.nz)}
        self._L.dim3_grouped_free(ctypes.byref(g))
        return out

    def join_aggregate_dev(self, d_R, d_vR, n_r, d_S, d_vS, n_s, agg, d_out, cap, stream=None, **kw):
        """Device-resident value aggregate (addresses as ints); rows are
        16-byte (x, z, int64) records.  Returns (status, QueryResult)."""
        opts = make_opts(**kw)
        res = Result()
        vp = ctypes.c_void_p
        rc = self._L.dim3_join_aggregate_dev(self._h, vp(d_R), vp(d_vR), n_r, vp(d_S), vp(d_vS), n_s, _agg(agg),
                                             ctypes.byref(opts), vp(d_out), cap, vp(stream or 0), ctypes.byref(res))
        if rc not in (OK, ERR_CAPACITY):
            self._check(rc, "dim3_join_aggregate_dev")
        return rc, QueryResult(res)

    def _stream(self, fn, name, R, S, on_rows, chunk_rows, width, dtype, kw):
        opts = make_opts(**kw)
        R, S = _pairs(R), _pairs(S)
        res = Result()
        err = []

        def cb(_user, rows, n):
            try:
                buf = (ctypes.c_char * (width * n)).from_address(rows)
                return 1 if on_rows(np.frombuffer(buf, dtype=dtype)) is False else 0
            except BaseException as e:  # stop the query, re-raised below
                err.append(e)
                return 1

        c_cb = ROWS_CB(cb)
        rc = fn(self._h, _ptr(R), R.shape[0], _ptr(S), S.shape[0], ctypes.byref(opts), chunk_rows, c_cb, None,
                ctypes.byref(res))
        if err:
            raise err[0]
        if rc != ERR_CANCELLED:
            self._check(rc, name)
        return rc, QueryResult(res)

    def join_aggregate_stream(self, R, vR, S, vS, agg, on_rows, chunk_rows=0, f32=False, **kw):
        """Streamed value aggregate: on_rows(rows) gets each chunk as a
        COUNT_ROW_DTYPE view (int32 values) or an F32_ROW_DTYPE view (f32)."""
        opts = make_opts(**kw)
        R, S = _pairs(R), _pairs(S)
        vR = _fvalues(vR, R.shape[0]) if f32 else _values(vR, R.shape[0])
        vS = _fvalues(vS, S.shape[0]) if f32 else _values(vS, S.shape[0])
        res = Result()
        err = []
        dtype, width = (F32_ROW_DTYPE, 12) if f32 else (COUNT_ROW_DTYPE, 16)

        def cb(_user, rows, n):
            try:
                buf = (ctypes.c_char * (width * n)).from_address(rows)
                return 1 if on_rows(np.frombuffer(buf, dtype=dtype)) is False else 0
            except BaseException as e:  # stop the query, re-raised below
                err.append(e)
                return 1

        c_cb = ROWS_CB(cb)
        fn = self._L.dim3_join_aggregate_f32_stream if f32 else self._L.dim3_join_aggregate_stream
        rc = fn(self._h, _ptr(R), _ptr(vR), R.shape[0], _ptr(S), _ptr(vS), S.shape[0], _agg(agg), ctypes.byref(opts),
                chunk_rows, c_cb, None, ctypes.byref(res))
        if err:
            raise err[0]
        if rc != ERR_CANCELLED:
            self._check(rc, "dim3_join_aggregate_stream")
        return rc, QueryResult(res)

    def join_project_stream(self, R, S, on_rows, chunk_rows=0, **kw):
        """Streamed join-project: on_rows(pairs) gets each chunk as an (n, 2)
        int32 view valid during the call (copy what you keep); returning
        False stops the query (status ERR_CANCELLED).  Returns (status, result)."""
        pair = np.dtype([("x", np.int32), ("z", np.int32)])
        return self._stream(self._L.dim3_join_project_stream, "dim3_join_project_stream", R, S,
                            lambda a: on_rows(a.view(np.int32).reshape(-1, 2)), chunk_rows, 8, pair, kw)

    def join_aggregate_count_stream(self, R, S, on_rows, chunk_rows=0, **kw):
        """Streamed COUNT join-aggregate: on_rows(rows) gets each chunk as a
        COUNT_ROW_DTYPE view (fields x, z, count) valid during the call."""
        return self._stream(self._L.dim3_join_aggregate_count_stream, "dim3_join_aggregate_count_stream", R, S,
                            on_rows, chunk_rows, 16, COUNT_ROW_DTYPE, kw)

    def join_project_dev(self, d_R, n_r, d_S, n_s, d_out, cap, stream=None, **kw):
        """Device-resident variant: d_R/d_S/d_out are device addresses (ints).
        Returns (status, QueryResult); status ERR_CAPACITY reports res.n."""
        opts = make_opts(**kw)
        res = Result()
        rc = self._L.dim3_join_project_dev(self._h, ctypes.c_void_p(d_R), n_r, ctypes.c_void_p(d_S), n_s,
                                           ctypes.byref(opts), ctypes.c_void_p(d_out), cap,
                                           ctypes.c_void_p(stream or 0), ctypes.byref(res))
        if rc not in (OK, ERR_CAPACITY):
            self._check(rc, "dim3_join_project_dev")
        return rc, QueryResult(res)

    def join_project_dev_noalloc(self, d_R, n_r, d_S, n_s, d_out, cap, stream=None, **kw):
        """dim3_join_project_dev_noalloc: join_project_dev with every workspace
        buffer in the library's per-device scratch array (no allocation; at
        most 2^17 tuples per relation).  Returns (status, QueryResult)."""
        opts = make_opts(**kw)
        res = Result()
        rc = self._L.dim3_join_project_dev_noalloc(self._h, ctypes.c_void_p(d_R), n_r, ctypes.c_void_p(d_S), n_s,
                                                   ctypes.byref(opts), ctypes.c_void_p(d_out), cap,
                                                   ctypes.c_void_p(stream or 0), ctypes.byref(res))
        if rc not in (OK, ERR_CAPACITY, ERR_INVALID):
            self._check(rc, "dim3_join_project_dev_noalloc")
        return rc, QueryResult(res)

    def join_aggregate_count_dev(self, d_R, n_r, d_S, n_s, d_out, cap, stream=None, **kw):
        opts = make_opts(**kw)
        res = Result()
        rc = self._L.dim3_join_aggregate_count_dev(self._h, ctypes.c_void_p(d_R), n_r, ctypes.c_void_p(d_S), n_s,
                                                   ctypes.byref(opts), ctypes.c_void_p(d_out), cap,
                                                   ctypes.c_void_p(stream or 0), ctypes.byref(res))
        if rc not in (OK, ERR_CAPACITY):
            self._check(rc, "dim3_join_aggregate_count_dev")
        return rc, QueryResult(res)

    def _batch(self, fn, name, instances, stream, kw):
        n = len(instances)
        arr = (Instance * max(n, 1))()
        for i, t in enumerate(instances):
            d_R, n_r, d_S, n_s = t[:4]
            d_out, cap = (t[4], t[5]) if len(t) > 4 else (0, 0)
            arr[i] = Instance(d_R, n_r, d_S, n_s, d_out or None, cap)
        res = (Result * max(n, 1))()
        opts = make_opts(**kw)
        rc = fn(self._h, arr, n, ctypes.byref(opts), ctypes.c_void_p(stream or 0), res)
        if rc not in (OK, ERR_CAPACITY):
            self._check(rc, name)
        return rc, [QueryResult(res[i]) for i in range(n)]

    def batch_join_project_dev(self, instances, stream=None, **kw):
        """Many instances in one persistent launch.  instances: tuples
        (d_R, n_r, d_S, n_s[, d_out, cap]) of device addresses.  Returns
        (status, [QueryResult per instance])."""
        return self._batch(self._L.dim3_batch_join_project_dev, "dim3_batch_join_project_dev", instances, stream, kw)

    def batch_join_aggregate_count_dev(self, instances, stream=None, **kw):
        return self._batch(self._L.dim3_batch_join_aggregate_count_dev, "dim3_batch_join_aggregate_count_dev",
                           instances, stream, kw)

    def autotune(self, cache_path=None):
        """Measure the planner's device rates (dim3_autotune); returns them as a dict."""
        t = Tuning()
        rc = self._L.dim3_autotune(self._h, cache_path.encode() if cache_path else None, ctypes.byref(t))
        self._check(rc, "dim3_autotune")
        return t.as_dict()

    def tuning(self):
        t = Tuning()
        self._check(self._L.dim3_get_tuning(self._h, ctypes.byref(t)), "dim3_get_tuning")
        return t.as_dict()

    def set_tuning(self, d):
        t = Tuning()
        for k, v in d.items():
            setattr(t, k, v)
        self._check(self._L.dim3_set_tuning(self._h, ctypes.byref(t)), "dim3_set_tuning")

    def estimate_join_cardinality(self, R, S):
        """EstimateJoinCardinality::estimate: the reference's sampled join size."""
        R, S = _pairs(R), _pairs(S)
        out = ctypes.c_int64()
        rc = self._L.dim3_estimate_join_cardinality(self._h, _ptr(R), R.shape[0], _ptr(S), S.shape[0],
                                                    ctypes.byref(out))
        self._check(rc, "dim3_estimate_join_cardinality")
        return int(out.value)

    def estimate_join_cardinality_dev(self, d_R, n_r, d_S, n_s, stream=None):
        """The same over device addresses (ints) of int32 pair arrays."""
        out = ctypes.c_int64()
        rc = self._L.dim3_estimate_join_cardinality_dev(self._h, ctypes.c_void_p(d_R), n_r, ctypes.c_void_p(d_S),
                                                        n_s, ctypes.c_void_p(stream or 0), ctypes.byref(out))
        self._check(rc, "dim3_estimate_join_cardinality_dev")
        return int(out.value)

    def count_join_paths(self, R, S):
        R, S = _pairs(R), _pairs(S)
        out = ctypes.c_uint64()
        rc = self._L.dim3_count_join_paths(self._h, _ptr(R), R.shape[0], _ptr(S), S.shape[0], ctypes.byref(out))
        self._check(rc, "dim3_count_join_paths")
        return int(out.value)

    def dense_product(self, A, Bt, bits=True, counts=True):
        """C = A @ Bt.T on the MFMA cores (int8 in, int32 accumulate).
        Returns (bit rows as uint32 [M, ceil(N/32)] or None, counts int32 [M, N] or None)."""
        A = np.ascontiguousarray(A, dtype=np.int8)
        Bt = np.ascontiguousarray(Bt, dtype=np.int8)
        if A.ndim != 2 or Bt.ndim != 2 or A.shape[1] != Bt.shape[1]:
            raise ValueError("A (M x K) and Bt (N x K) must share K")
        M, K = A.shape
        N = Bt.shape[0]
        ob = np.zeros((M, (N + 31) // 32), np.uint32) if bits else None
        oc = np.zeros((M, N), np.int32) if counts else None
        rc = self._L.dim3_dense_product(self._h, _ptr(A), M, _ptr(Bt), N, K,
                                        None if ob is None else _ptr(ob), None if oc is None else _ptr(oc))
        self._check(rc, "dim3_dense_product")
        return ob, oc


class Server:
    """The libdim3 request loop (dim3_server_*): resident relations,
    admission by HBM footprint, concurrent workers, continuous batching."""

    def __init__(self, device=0, hbm_budget_bytes=0, workers=4, max_batch=32, queue_cap=4096, large_share_pct=None,
                 large_tuples=None):
        L = lib()
        o = ServerOpts()
        L.dim3_server_opts_default(ctypes.byref(o))
        o.device, o.hbm_budget_bytes, o.workers, o.max_batch, o.queue_cap = (device, hbm_budget_bytes, workers,
                                                                              max_batch, queue_cap)
        if large_share_pct is not None:
            o.large_share_pct = large_share_pct
        if large_tuples is not None:
            o.large_tuples = large_tuples
        h = ctypes.c_void_p()
        rc = L.dim3_server_create(ctypes.byref(o), ctypes.byref(h))
        if rc != OK:
            raise Dim3Error(rc, "dim3_server_create")
        self._h, self._L = h, L

    def close(self):
        if self._h is not None and self._h.value:
            self._L.dim3_server_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def put(self, pairs):
        """Copy a host (n, 2) int32 relation into the server; returns its handle."""
        a = _pairs(pairs)
        h = ctypes.c_uint64()
        rc = self._L.dim3_server_put_relation(self._h, _ptr(a), a.shape[0], ctypes.byref(h))
        if rc != OK:
            raise Dim3Error(rc, "dim3_server_put_relation")
        return int(h.value)

    def put_dev(self, d_ptr, n):
        h = ctypes.c_uint64()
        rc = self._L.dim3_server_put_relation_dev(self._h, ctypes.c_void_p(d_ptr), n, ctypes.byref(h))
        if rc != OK:
            raise Dim3Error(rc, "dim3_server_put_relation_dev")
        return int(h.value)

    def put_values(self, handle, values):
        """Attach one int32 or float32 value per tuple to a relation (once,
        before any request uses it)."""
        v = np.ascontiguousarray(values).reshape(-1)
        vtype = {np.dtype("<i4"): 0, np.dtype("<f4"): 2}[v.dtype]
        rc = self._L.dim3_server_put_values(self._h, handle, _ptr(v), vtype)
        if rc != OK:
            raise Dim3Error(rc, "dim3_server_put_values")

    def drop(self, handle):
        rc = self._L.dim3_server_drop_relation(self._h, handle)
        if rc != OK:
            raise Dim3Error(rc, "dim3_server_drop_relation")

    def submit(self, r, s, count=False, agg=None, **kw):
        """Returns (status, ticket); status OK, ERR_BUSY or ERR_ADMISSION.
        agg ("sum", "min", "max", "avg"): a value aggregate over the
        relations' values (put_values)."""
        req = Request()
        req.kind = 2 if agg is not None else 1 if count else 0
        req.agg = _agg(agg) if agg is not None else 0
        req.r, req.s = r, s
        req.opts = make_opts(**kw)
        t = ctypes.c_uint64()
        rc = self._L.dim3_submit(self._h, ctypes.byref(req), ctypes.byref(t))
        if rc not in (OK, ERR_BUSY, ERR_ADMISSION):
            raise Dim3Error(rc, "dim3_submit")
        return rc, int(t.value)

    def poll(self, ticket):
        d = ctypes.c_int()
        rc = self._L.dim3_poll(self._h, ticket, ctypes.byref(d))
        if rc != OK:
            raise Dim3Error(rc, "dim3_poll")
        return bool(d.value)

    def wait(self, ticket):
        """Returns (status, QueryResult, queued seconds, total seconds)."""
        res = Result()
        q, t = ctypes.c_double(), ctypes.c_double()
        rc = self._L.dim3_wait(self._h, ticket, ctypes.byref(res), ctypes.byref(q), ctypes.byref(t))
        if rc == ERR_NOT_FOUND:
            raise Dim3Error(rc, "dim3_wait")
        pairs = rows = None
        n = int(res.n)
        if rc == OK and res.pairs:
            pairs = np.frombuffer((ctypes.c_int32 * (2 * n)).from_address(res.pairs), dtype=np.int32).reshape(n, 2).copy()
        if rc == OK and res.rows:
            rows = np.frombuffer((CountRow * n).from_address(res.rows), dtype=COUNT_ROW_DTYPE).copy()
        if rc == OK and res.rows_f32:
            rows = np.frombuffer((ctypes.c_char * (12 * n)).from_address(res.rows_f32), dtype=F32_ROW_DTYPE).copy()
        out = QueryResult(res, pairs=pairs, rows=rows)
        self._L.dim3_result_free(ctypes.byref(res))
        return rc, out, q.value, t.value

    def stats(self):
        st = ServerStats()
        self._L.dim3_server_get_stats(self._h, ctypes.byref(st))
        return st.as_dict()


def version():
    return lib().dim3_version().decode()
