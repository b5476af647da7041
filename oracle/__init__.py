"""ctypes access to the CPU oracle (liboracle.so) and, when built, the
reference checker (_ref/libdim3ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, always as the checker, never as the thing
measured or shipped.

Pinning of each oracle function (what would fail if it were wrong):
  oracle_join_project          compiled reference on random/Zipf inputs
                               (tests/test_oracle.py), committed golden
                               fixtures written by the reference
                               (tests/golden/), the gen_rand_data KAT
                               |Result| = 98,931,829, brute force
                               nested-loop + set on tiny inputs.
  oracle_join_aggregate_count  compiled reference with a COUNT policy,
                               brute force via a dict, closed forms
                               sum(count) = OUT_J and #groups = |project|.
  oracle_join_aggregate        compiled reference with SUM / MIN / MAX
                               policies and its AggregateFunctionAvg<long long>,
                               brute force via a dict on tiny inputs, and
                               closed forms (SUM with vS = 0 is vR times the
                               COUNT, AVG of constant values is the constant).
  oracle_outj                  brute force on tiny inputs, sum of counts.
  oracle_estimate              compiled reference estimate on the same inputs.
  oracle_gen_rand_data         compiled reference generator (byte equality)
                               and the KAT above.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libdim3ref.so")


class Fingerprint(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("sum", ctypes.c_uint64),
                ("sum2", ctypes.c_uint64), ("csum", ctypes.c_uint64)]

    def as_tuple(self):
        return (int(self.count), int(self.sum), int(self.sum2), int(self.csum))


_lib = None
_ref = None


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = ctypes.CDLL(ORACLE_SO)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.oracle_join_project.argtypes = [vp, u64, vp, u64, i32, i32, vp, vp, vp, vp]
        L.oracle_join_project.restype = i32
        L.oracle_join_aggregate_count.argtypes = [vp, u64, vp, u64, i32, i32, vp, vp, vp, vp, vp]
        L.oracle_join_aggregate_count.restype = i32
        L.oracle_join_aggregate.argtypes = [vp, vp, u64, vp, vp, u64, i32, i32, i32, vp, vp, vp, vp, vp]
        L.oracle_join_aggregate.restype = i32
        L.oracle_join_aggregate_f32.argtypes = [vp, vp, u64, vp, vp, u64, i32, i32, i32, vp, vp, vp, vp, vp]
        L.oracle_join_aggregate_f32.restype = i32
        L.oracle_outj.argtypes = [vp, u64, vp, u64]
        L.oracle_outj.restype = u64
        L.oracle_estimate.argtypes = [vp, u64, vp, u64]
        L.oracle_estimate.restype = ctypes.c_int64
        L.oracle_gen_rand_data.argtypes = [vp, u64, vp, u64, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_gen_rand_data.restype = None
        L.oracle_fingerprint.argtypes = [vp, vp, u64, vp]
        L.oracle_fingerprint.restype = None
        L.oracle_mix64.argtypes = [u64]
        L.oracle_mix64.restype = u64
        L.oracle_zfactor.argtypes = [ctypes.c_int32]
        L.oracle_zfactor.restype = u64
        L.oracle_free.argtypes = [vp]
        L.oracle_free.restype = None
        L.oracle_project_fp_bitset.argtypes = [vp, u64, vp, u64, i32, u64, ctypes.c_uint32, ctypes.c_uint32, i32,
                                               vp, vp]
        L.oracle_project_fp_bitset.restype = i32
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference checker not built: run make -C oracle ref")
        L = ctypes.CDLL(REF_SO)
        vp, u32 = ctypes.c_void_p, ctypes.c_uint32
        L.ref_join_project.argtypes = [vp, u32, vp, u32, ctypes.c_int, vp, vp]
        L.ref_join_project.restype = ctypes.c_int
        L.ref_join_aggregate_count.argtypes = [vp, u32, vp, u32, vp, vp, vp]
        L.ref_join_aggregate_count.restype = ctypes.c_int
        L.ref_join_aggregate.argtypes = [vp, vp, u32, vp, vp, u32, ctypes.c_int, vp, vp, vp]
        L.ref_join_aggregate.restype = ctypes.c_int
        L.ref_join_aggregate_f32.argtypes = [vp, vp, u32, vp, vp, u32, ctypes.c_int, vp, vp, vp]
        L.ref_join_aggregate_f32.restype = ctypes.c_int
        L.ref_estimate.argtypes = [vp, u32, vp, u32]
        L.ref_estimate.restype = ctypes.c_longlong
        L.ref_gen_rand_data.argtypes = [vp, u32, vp, u32, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int]
        L.ref_gen_rand_data.restype = None
        L.ref_thread_num.argtypes = []
        L.ref_thread_num.restype = ctypes.c_int
        L.ref_free.argtypes = [vp]
        L.ref_free.restype = None
        _ref = L
    return _ref


def _pairs(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    if a.ndim != 2 or a.shape[1] != 2:
        a = a.reshape(-1, 2)
    return a


def _take_pairs(L, ptr, n, free):
    if n == 0 or not ptr.value:
        if ptr.value:
            free(ptr)
        return np.zeros((0, 2), dtype=np.int32)
    buf = (ctypes.c_int32 * (2 * n)).from_address(ptr.value)
    out = np.frombuffer(buf, dtype=np.int32).reshape(n, 2).copy()
    free(ptr)
    return out


def _take_counts(ptr, n, free):
    if n == 0 or not ptr.value:
        if ptr.value:
            free(ptr)
        return np.zeros((0,), dtype=np.int64)
    buf = (ctypes.c_int64 * n).from_address(ptr.value)
    out = np.frombuffer(buf, dtype=np.int64).copy()
    free(ptr)
    return out


def join_project(R, S, exclude_equal=False, fingerprint_only=False):
    """Returns (pairs[n,2] int32 or None, Fingerprint, swapped)."""
    L = lib()
    R, S = _pairs(R), _pairs(S)
    out = ctypes.c_void_p()
    n = ctypes.c_uint64()
    fp = Fingerprint()
    sw = ctypes.c_int()
    rc = L.oracle_join_project(_p(R), R.shape[0], _p(S), S.shape[0], int(exclude_equal),
                               int(fingerprint_only), ctypes.byref(out), ctypes.byref(n),
                               ctypes.byref(fp), ctypes.byref(sw))
    if rc != 0:
        raise MemoryError("oracle_join_project failed rc=%d" % rc)
    pairs = None if fingerprint_only else _take_pairs(L, out, n.value, L.oracle_free)
    return pairs, fp, bool(sw.value)


_FP_JOB = None


def _fp_slice(rem):
    R, S, exclude_equal, heavy, xmod, with_paths = _FP_JOB
    fp = Fingerprint()
    sw = ctypes.c_int()
    rc = lib().oracle_project_fp_bitset(_p(R), R.shape[0], _p(S), S.shape[0], int(exclude_equal), int(heavy),
                                        int(xmod), int(rem), int(with_paths), ctypes.byref(fp), ctypes.byref(sw))
    if rc != 0:
        raise MemoryError("oracle_project_fp_bitset failed rc=%d" % rc)
    return fp.as_tuple(), bool(sw.value)


def fingerprint_bitset(R, S, exclude_equal=False, heavy=1024, slices=1, workers=1, count=False, only=None):
    """Join-project fingerprint (count=True: the COUNT fingerprint) by the
    bitset restatement, oracle_project_fp_bitset: keys with at least `heavy`
    S tuples as z bitsets, rows split into `slices` x slices ((u32)x % slices)
    run by `workers` forked processes and summed mod 2^64 (only=r: slice r
    alone).  Returns (Fingerprint, swapped)."""
    global _FP_JOB
    R, S = _pairs(R), _pairs(S)
    _FP_JOB = (R, S, exclude_equal, heavy, slices, count)
    rems = [only] if only is not None else list(range(max(slices, 1)))
    try:
        if workers > 1 and len(rems) > 1:
            import multiprocessing as mp

            with mp.get_context("fork").Pool(min(workers, len(rems))) as pool:
                parts = pool.map(_fp_slice, rems)
        else:
            parts = [_fp_slice(r) for r in rems]
    finally:
        _FP_JOB = None
    acc = [0, 0, 0, 0]
    for t, _ in parts:
        acc = [(a + v) & 0xFFFFFFFFFFFFFFFF for a, v in zip(acc, t)]
    fp = Fingerprint()
    fp.count, fp.sum, fp.sum2, fp.csum = acc
    return fp, parts[0][1]


def join_aggregate_count(R, S, exclude_equal=False, fingerprint_only=False):
    """Returns (pairs[n,2] int32, counts[n] int64, Fingerprint, swapped)."""
    L = lib()
    R, S = _pairs(R), _pairs(S)
    out = ctypes.c_void_p()
    cnt = ctypes.c_void_p()
    n = ctypes.c_uint64()
    fp = Fingerprint()
    sw = ctypes.c_int()
    rc = L.oracle_join_aggregate_count(_p(R), R.shape[0], _p(S), S.shape[0], int(exclude_equal),
                                       int(fingerprint_only), ctypes.byref(out), ctypes.byref(cnt),
                                       ctypes.byref(n), ctypes.byref(fp), ctypes.byref(sw))
    if rc != 0:
        raise MemoryError("oracle_join_aggregate_count failed rc=%d" % rc)
    if fingerprint_only:
        return None, None, fp, bool(sw.value)
    return (_take_pairs(L, out, n.value, L.oracle_free), _take_counts(cnt, n.value, L.oracle_free),
            fp, bool(sw.value))


AGG = {"sum": 1, "min": 2, "max": 3, "avg": 4}


def _vals(v, n):
    v = np.ascontiguousarray(v, dtype=np.int32).reshape(-1)
    if v.shape[0] != n:
        raise ValueError("one value per tuple")
    return v


def join_aggregate(R, vR, S, vS, agg, exclude_equal=False, fingerprint_only=False):
    """Value aggregate agg ("sum", "min", "max", "avg") of vR + vS over the
    join paths of each (x, z).  Returns (pairs[n,2] int32, values[n] int64,
    Fingerprint, swapped)."""
    L = lib()
    R, S = _pairs(R), _pairs(S)
    vR, vS = _vals(vR, R.shape[0]), _vals(vS, S.shape[0])
    out = ctypes.c_void_p()
    val = ctypes.c_void_p()
    n = ctypes.c_uint64()
    fp = Fingerprint()
    sw = ctypes.c_int()
    rc = L.oracle_join_aggregate(_p(R), _p(vR), R.shape[0], _p(S), _p(vS), S.shape[0], AGG[agg],
                                 int(exclude_equal), int(fingerprint_only), ctypes.byref(out),
                                 ctypes.byref(val), ctypes.byref(n), ctypes.byref(fp), ctypes.byref(sw))
    if rc != 0:
        raise MemoryError("oracle_join_aggregate failed rc=%d" % rc)
    if fingerprint_only:
        return None, None, fp, bool(sw.value)
    return (_take_pairs(L, out, n.value, L.oracle_free), _take_counts(val, n.value, L.oracle_free),
            fp, bool(sw.value))


def _fvals(v, n):
    v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1)
    if v.shape[0] != n:
        raise ValueError("one value per tuple")
    return v


def join_aggregate_f32(R, vR, S, vS, agg, exclude_equal=False, fingerprint_only=False):
    """join_aggregate on float32 values: float path value vR + vS, SUM / AVG
    accumulated in double and rounded once, MIN / MAX by totalOrder.
    Returns (pairs[n,2] int32, values[n] float32, Fingerprint, swapped); the
    fingerprint counts the result's bit pattern."""
    L = lib()
    R, S = _pairs(R), _pairs(S)
    vR, vS = _fvals(vR, R.shape[0]), _fvals(vS, S.shape[0])
    out = ctypes.c_void_p()
    val = ctypes.c_void_p()
    n = ctypes.c_uint64()
    fp = Fingerprint()
    sw = ctypes.c_int()
    rc = L.oracle_join_aggregate_f32(_p(R), _p(vR), R.shape[0], _p(S), _p(vS), S.shape[0], AGG[agg],
                                     int(exclude_equal), int(fingerprint_only), ctypes.byref(out),
                                     ctypes.byref(val), ctypes.byref(n), ctypes.byref(fp), ctypes.byref(sw))
    if rc != 0:
        raise MemoryError("oracle_join_aggregate_f32 failed rc=%d" % rc)
    if fingerprint_only:
        return None, None, fp, bool(sw.value)
    bits = _take_counts(val, n.value, L.oracle_free)
    return (_take_pairs(L, out, n.value, L.oracle_free), bits.astype(np.uint32).view(np.float32), fp,
            bool(sw.value))


def outj(R, S):
    L = lib()
    R, S = _pairs(R), _pairs(S)
    return int(L.oracle_outj(_p(R), R.shape[0], _p(S), S.shape[0]))


def estimate(R, S):
    L = lib()
    R, S = _pairs(R), _pairs(S)
    return int(L.oracle_estimate(_p(R), R.shape[0], _p(S), S.shape[0]))


def gen_rand_data(nR=10_000_000, nS=10_000_000, modx=1_000_000, mody=1_000_000, modz=1_000_000,
                  seed=492):
    L = lib()
    R = np.empty((nR, 2), dtype=np.int32)
    S = np.empty((nS, 2), dtype=np.int32)
    L.oracle_gen_rand_data(_p(R), nR, _p(S), nS, modx, mody, modz, seed)
    return R, S


def fingerprint(pairs, counts=None):
    L = lib()
    pairs = _pairs(pairs)
    fp = Fingerprint()
    if counts is not None:
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        L.oracle_fingerprint(_p(pairs), _p(counts), pairs.shape[0], ctypes.byref(fp))
    else:
        L.oracle_fingerprint(_p(pairs), None, pairs.shape[0], ctypes.byref(fp))
    return fp


def mix64(k):
    return int(lib().oracle_mix64(k & 0xFFFFFFFFFFFFFFFF))


def zfactor(z):
    """The fingerprint's z factor g(z) (mod 2^64)."""
    return int(lib().oracle_zfactor(int(np.int32(z))))


# ---------------- reference checker ----------------

def ref_join_project(R, S, parallel=False):
    L = ref()
    R, S = _pairs(R), _pairs(S)
    out = ctypes.c_void_p()
    n = ctypes.c_uint64()
    L.ref_join_project(_p(R), R.shape[0], _p(S), S.shape[0], int(parallel), ctypes.byref(out),
                       ctypes.byref(n))
    return _take_pairs(L, out, n.value, L.ref_free)


def ref_join_aggregate_count(R, S):
    L = ref()
    R, S = _pairs(R), _pairs(S)
    out = ctypes.c_void_p()
    cnt = ctypes.c_void_p()
    n = ctypes.c_uint64()
    L.ref_join_aggregate_count(_p(R), R.shape[0], _p(S), S.shape[0], ctypes.byref(out),
                               ctypes.byref(cnt), ctypes.byref(n))
    return _take_pairs(L, out, n.value, L.ref_free), _take_counts(cnt, n.value, L.ref_free)


def ref_join_aggregate(R, vR, S, vS, agg):
    L = ref()
    R, S = _pairs(R), _pairs(S)
    vR, vS = _vals(vR, R.shape[0]), _vals(vS, S.shape[0])
    out = ctypes.c_void_p()
    val = ctypes.c_void_p()
    n = ctypes.c_uint64()
    rc = L.ref_join_aggregate(_p(R), _p(vR), R.shape[0], _p(S), _p(vS), S.shape[0], AGG[agg], ctypes.byref(out),
                              ctypes.byref(val), ctypes.byref(n))
    if rc != 0:
        raise ValueError("ref_join_aggregate rc=%d" % rc)
    return _take_pairs(L, out, n.value, L.ref_free), _take_counts(val, n.value, L.ref_free)


def ref_join_aggregate_f32(R, vR, S, vS, agg):
    L = ref()
    R, S = _pairs(R), _pairs(S)
    vR, vS = _fvals(vR, R.shape[0]), _fvals(vS, S.shape[0])
    out = ctypes.c_void_p()
    val = ctypes.c_void_p()
    n = ctypes.c_uint64()
    rc = L.ref_join_aggregate_f32(_p(R), _p(vR), R.shape[0], _p(S), _p(vS), S.shape[0], AGG[agg],
                                  ctypes.byref(out), ctypes.byref(val), ctypes.byref(n))
    if rc != 0:
        raise ValueError("ref_join_aggregate_f32 rc=%d" % rc)
    nn = n.value
    if nn == 0 or not val.value:
        if val.value:
            L.ref_free(val)
        vals = np.zeros(0, np.float32)
    else:
        vals = np.frombuffer((ctypes.c_float * nn).from_address(val.value), dtype=np.float32).copy()
        L.ref_free(val)
    return _take_pairs(L, out, nn, L.ref_free), vals


def ref_estimate(R, S):
    L = ref()
    R, S = _pairs(R), _pairs(S)
    return int(L.ref_estimate(_p(R), R.shape[0], _p(S), S.shape[0]))


def ref_gen_rand_data(nR, nS, modx, mody, modz, seed):
    L = ref()
    R = np.empty((nR, 2), dtype=np.int32)
    S = np.empty((nS, 2), dtype=np.int32)
    L.ref_gen_rand_data(_p(R), nR, _p(S), nS, modx, mody, modz, seed)
    return R, S


# ---------------- helpers shared by tests ----------------

def sort_pairs(p, counts=None):
    """Lexicographic sort of (x,z) rows (and their counts)."""
    p = _pairs(p)
    key = (p[:, 0].astype(np.int64) << 32) | (p[:, 1].astype(np.int64) & 0xFFFFFFFF)
    o = np.argsort(key, kind="stable")
    if counts is None:
        return p[o]
    return p[o], np.asarray(counts)[o]
