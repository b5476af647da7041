// dim3.hpp — drop-in C++ front end of libdim3 for callers of the reference.
//
// Re-exposes the reference's query entry points with the same names,
// template parameters and argument meaning, forwarding to the C-ABI in
// dim3.h (hand-written HIP kernels for gfx950):
//
//   DIM3<_Tx,_Ty,_Tz,_Hx,_Hy,_Hz,_Hxz,_Tcounter>::doJoinProject          dim3.hpp:982-1055
//   DIM3_parallel<...>::doJoinProject_parallel                           dim3_parallel.hpp:137-209
//   DIM3JoinAggregate<_Tx,_Ty,_Tz,_Tv,_Agg,...>::doJoinAggregate         dim3.hpp:1313-1373
//   EstimateJoinCardinality<_Tx,_Ty,_Tz,_H,_Tcounter>::estimate          dim3.hpp:190-270
//   myvector<T,_CountType> (same layout: data, cur_num, max_num)         Include/myvector.hpp:9-86
//   TUPLE<_Tx,_Ty,_Tv> (same layout: x, y, v)                            dim3.hpp:20-33
//   myHasher (placeholder: hashing happens on the device)                Include/structures.hpp:92
//   gen_rand_data (minstd_rand0, same draws; TUPLE value = tuple index)   dim3.hpp:1556-1591
//
// Conventions kept: inputs by value (shallow), result cleared then filled
// with malloc-compatible storage the caller frees, |R| < |S| swaps the roles
// and rows come out as (S.first, R.first), duplicates count with
// multiplicity in COUNT.  Errors print a message and exit(-1), as the
// reference does on allocation failure (e.g. dim3.hpp:303-307); the
// library's own C-ABI returns status codes instead.  A result that does not
// fit _Tcounter is an error here, where the reference's uint32 counters
// would wrap (Include/myvector.hpp:33-46).
//
// Supported instantiations: 4-byte integer _Tx/_Ty/_Tz (int, as every
// reference driver uses); join-aggregate with the policies
// AggregateFunctionAvg<_Tv> (the reference's, dim3.hpp:1453-1480),
// AggregateFunctionSum / Min / Max<_Tv> (value policies over v1 + v2) and
// AggregateFunctionCount<_Tv>, for _Tv float (float results, double
// accumulation) or an integer type whose values fit int32 (int64 results),
// aggregated on the device; and any caller-written _Aggregate_type with the
// reference's policy concept (ctor(int nz), insert(int z, v1, v2),
// update(int z, v1, v2), get(int z)) and a trivially copyable _Tv of 4 or 8
// bytes: the device joins and groups the paths by (x, z)
// (dim3_join_paths_grouped), the policy folds them on the host in the
// reference's call pattern.
// Device: DIM3_DEVICE env var, default 0.
#pragma once

#include <assert.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <random>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "dim3.h"

// What the reference's header gives its callers besides its own names
// (dim3.hpp:6-18): <chrono> and the rest of the standard library, the whole
// std namespace (main.cpp:17 writes chrono::high_resolution_clock), LL and
// FORCE_INLINE (Include/structures.hpp:7).
using namespace std;
typedef long long LL;
#ifndef FORCE_INLINE
#define FORCE_INLINE inline __attribute__((always_inline))
#endif

template <class T, typename _CountType = uint32_t>
struct myvector {
    T* data;
    _CountType cur_num;
    _CountType max_num;

    myvector(_CountType ini_size = 8) { init(ini_size); }
    void init(_CountType ini_size = 8) {
        max_num = ini_size;
        data = (T*)malloc(sizeof(T) * (max_num ? max_num : 1));
        cur_num = 0;
    }
    inline bool isEmpty() { return cur_num == 0; }
    void push_back(T x) {
        if (cur_num == max_num) {
            _CountType nm = max_num ? max_num * 2 : 8;
            T* p = (T*)realloc(data, sizeof(T) * nm);
            if (!p) {
                printf("Failed when resize myvector to %llu\n", (unsigned long long)nm);
                free(data);
                exit(1);
            }
            data = p;
            max_num = nm;
        }
        data[cur_num++] = x;
    }
    void fixSize(_CountType size) {
        if (size == 0) return;
        reserve(size);
        cur_num = size;
    }
    void reserve(_CountType size) {
        if (size < max_num) return;
        T* p = (T*)realloc(data, sizeof(T) * size);
        if (!p) {
            printf("Failed when reserve myvector to %llu\n", (unsigned long long)size);
            free(data);
            exit(1);
        }
        data = p;
        max_num = size;
    }
    inline _CountType size() const { return cur_num; }
    inline void clear() { cur_num = 0; }
};

template <typename _Tx, typename _Ty, typename _Tv>
struct TUPLE {
    _Tx x;
    _Ty y;
    _Tv v;
    TUPLE(_Tx x_, _Ty y_, _Tv v_) : x(x_), y(y_), v(v_) {}
};

struct myHasher {};

// COUNT(*) policy for DIM3JoinAggregate (the reference ships only AVG,
// dim3.hpp:1453-1480; its policy concept is ctor(int)/insert/update/get).
template <typename _Tv>
struct AggregateFunctionCount {
    typedef _Tv value_type;
};

// Value policies over the path value v1 + v2: the reference's AVG
// (dim3.hpp:1453-1480, sum / paths) and SUM / MIN / MAX.  Tags only: the
// aggregation runs on the device.
template <typename _Tv>
struct AggregateFunctionAvg {
    typedef _Tv value_type;
    static const int code = DIM3_AGG_AVG;
};
template <typename _Tv>
struct AggregateFunctionSum {
    typedef _Tv value_type;
    static const int code = DIM3_AGG_SUM;
};
template <typename _Tv>
struct AggregateFunctionMin {
    typedef _Tv value_type;
    static const int code = DIM3_AGG_MIN;
};
template <typename _Tv>
struct AggregateFunctionMax {
    typedef _Tv value_type;
    static const int code = DIM3_AGG_MAX;
};

namespace dim3_shim {

inline dim3_ctx* context() {
    static dim3_ctx* ctx = nullptr;
    if (!ctx) {
        const char* d = getenv("DIM3_DEVICE");
        int rc = dim3_ctx_create(d ? atoi(d) : 0, 0, &ctx);
        if (rc != DIM3_OK) {
            printf("Fail to create the DIM3 device context: %s.\n", dim3_status_string(rc));
            exit(-1);
        }
    }
    return ctx;
}

[[noreturn]] inline void die(const char* what, int rc) {
    printf("%s failed: %s (%s).\n", what, dim3_status_string(rc), dim3_last_error(context()));
    exit(-1);
}

template <typename T>
constexpr bool is_i32() {
    return std::is_integral<T>::value && sizeof(T) == 4;
}

template <typename _Tx, typename _Ty, typename _Tz, typename _Tcounter>
void join_project(const myvector<pair<_Tx, _Ty>, _Tcounter>& R, const myvector<pair<_Tz, _Ty>, _Tcounter>& S,
                  myvector<pair<_Tx, _Tz>, _Tcounter>& result) {
    static_assert(is_i32<_Tx>() && is_i32<_Ty>() && is_i32<_Tz>(), "libdim3 joins 4-byte integer attributes");
    static_assert(sizeof(pair<_Tx, _Ty>) == 8, "pair layout");
    result.clear();
    dim3_opts o;
    dim3_opts_default(&o);
    dim3_result res;
    int rc = dim3_join_project(context(), (const int32_t*)R.data, R.size(), (const int32_t*)S.data, S.size(), &o,
                               &res);
    if (rc != DIM3_OK) die("DIM3 doJoinProject", rc);
    if (res.n > (uint64_t)(_Tcounter)~(_Tcounter)0) {
        dim3_result_free(&res);
        printf("DIM3 doJoinProject: %llu rows do not fit the result counter type.\n", (unsigned long long)res.n);
        exit(-1);
    }
    if (res.n) {
        free(result.data);
        result.data = (pair<_Tx, _Tz>*)res.pairs;  // malloc'd by the library, caller frees
        res.pairs = nullptr;
        result.cur_num = result.max_num = (_Tcounter)res.n;
    }
    dim3_result_free(&res);
}

}  // namespace dim3_shim

template <typename _Tx, typename _Ty, typename _Tz, typename _Hx = myHasher, typename _Hy = myHasher,
          typename _Hz = myHasher, typename _Hxz = myHasher, typename _Tcounter = uint32_t>
struct DIM3 {
    static void doJoinProject(const myvector<pair<_Tx, _Ty>, _Tcounter> _R,
                              const myvector<pair<_Tz, _Ty>, _Tcounter> _S,
                              myvector<pair<_Tx, _Tz>, _Tcounter>& result) {
        dim3_shim::join_project(_R, _S, result);
    }
};

// The reference's sampled join size (keys whose low sample bits are zero
// under the identity hash, the 1000-tuple-prefix fallback), computed on the
// device: the same number.  Where the reference asserts |R| >= |S|, the
// library swaps the sides.  _H defaults to std::hash<_Ty> as the reference's
// (dim3.hpp:190); the sample is taken under the identity hash, which is what
// std::hash of a 4-byte integer (libstdc++) and myHasher compute, so other
// hashers are refused at compile time.
template <typename _Tx, typename _Ty, typename _Tz, typename _H = std::hash<_Ty>, typename _Tcounter = uint32_t>
struct EstimateJoinCardinality {
    static_assert(std::is_same<_H, myHasher>::value ||
                      (std::is_same<_H, std::hash<_Ty>>::value && dim3_shim::is_i32<_Ty>()),
                  "libdim3 samples join keys under the identity hash (myHasher or std::hash of a 4-byte integer)");

    static long long estimate(const myvector<pair<_Tx, _Ty>, _Tcounter> R,
                              const myvector<pair<_Tz, _Ty>, _Tcounter> S) {
        static_assert(dim3_shim::is_i32<_Tx>() && dim3_shim::is_i32<_Ty>() && dim3_shim::is_i32<_Tz>(),
                      "libdim3 joins 4-byte integer attributes");
        int64_t est = 0;
        const int rc = dim3_estimate_join_cardinality(dim3_shim::context(), (const int32_t*)R.data, R.size(),
                                                      (const int32_t*)S.data, S.size(), &est);
        if (rc != DIM3_OK) dim3_shim::die("EstimateJoinCardinality::estimate", rc);
        return (long long)est;
    }

    // the TUPLE overload (dim3.hpp:226-260): the key is TUPLE::y
    template <typename _Tv>
    static long long estimate(const myvector<TUPLE<_Tx, _Ty, _Tv>, _Tcounter> R,
                              const myvector<TUPLE<_Tz, _Ty, _Tv>, _Tcounter> S) {
        static_assert(dim3_shim::is_i32<_Tx>() && dim3_shim::is_i32<_Ty>() && dim3_shim::is_i32<_Tz>(),
                      "libdim3 joins 4-byte integer attributes");
        static_assert(sizeof(TUPLE<_Tx, _Ty, _Tv>) == sizeof(TUPLE<_Tz, _Ty, _Tv>), "tuple layouts");
        int64_t est = 0;
        const int rc = dim3_estimate_join_cardinality_tuples(dim3_shim::context(), R.data, R.size(), S.data, S.size(),
                                                             (uint32_t)sizeof(TUPLE<_Tx, _Ty, _Tv>), &est);
        if (rc != DIM3_OK) dim3_shim::die("EstimateJoinCardinality::estimate", rc);
        return (long long)est;
    }
};

template <typename _Tx, typename _Ty, typename _Tz, typename _Hx = myHasher, typename _Hy = myHasher,
          typename _Hz = myHasher, typename _Hxz = myHasher, typename _Tcounter = uint32_t>
struct DIM3_parallel : public DIM3<_Tx, _Ty, _Tz, _Hx, _Hy, _Hz, _Hxz, _Tcounter> {
    static void doJoinProject_parallel(const myvector<pair<_Tx, _Ty>, _Tcounter> _R,
                                       const myvector<pair<_Tz, _Ty>, _Tcounter> _S,
                                       myvector<pair<_Tx, _Tz>, _Tcounter>& result) {
        dim3_shim::join_project(_R, _S, result);
    }
};

namespace dim3_shim {

template <typename A, typename _Tv>
constexpr bool is_value_policy() {
    return std::is_same<A, AggregateFunctionAvg<_Tv>>::value || std::is_same<A, AggregateFunctionSum<_Tv>>::value ||
           std::is_same<A, AggregateFunctionMin<_Tv>>::value || std::is_same<A, AggregateFunctionMax<_Tv>>::value;
}

template <typename A, typename _Tv>
struct policy_code {
    static const int value = A::code;
};
template <typename _Tv>
struct policy_code<AggregateFunctionCount<_Tv>, _Tv> {
    static const int value = 0;
};

}  // namespace dim3_shim

namespace dim3_shim {

template <typename A, typename _Tv>
constexpr bool is_library_policy() {
    return std::is_same<A, AggregateFunctionCount<_Tv>>::value || is_value_policy<A, _Tv>();
}

// A policy the library does not implement is the caller's host code (the
// reference's concept, dim3.hpp:1453-1480: ctor(nz), insert / update / get
// with a dense z key): the device joins and groups the join paths by (x, z)
// (dim3_join_paths_grouped) and the policy folds them as SpmmAggregate does
// (dim3.hpp:1275-1307): per x row, insert on a group's first path, update on
// the others, get for each of the row's groups at its end.
template <typename A, typename _Tx, typename _Ty, typename _Tz, typename _Tv, typename _Tcounter>
void join_aggregate_policy(myvector<TUPLE<_Tx, _Ty, _Tv>, _Tcounter>& R, myvector<TUPLE<_Tz, _Ty, _Tv>, _Tcounter>& S,
                           myvector<TUPLE<_Tx, _Tz, _Tv>, _Tcounter>& result) {
    static_assert(is_i32<_Tx>() && is_i32<_Ty>() && is_i32<_Tz>(), "libdim3 joins 4-byte integer attributes");
    static_assert(sizeof(_Tv) == 4 || sizeof(_Tv) == 8, "tuple values of 4 or 8 bytes");
    static_assert(std::is_trivially_copyable<_Tv>::value, "trivially copyable tuple values");
    typedef TUPLE<_Tx, _Ty, _Tv> InT;
    typedef TUPLE<_Tx, _Tz, _Tv> OutT;
    const InT probe((_Tx)0, (_Ty)0, _Tv());
    const uint32_t voff = (uint32_t)((const char*)&probe.v - (const char*)&probe);
    dim3_opts o;
    dim3_opts_default(&o);
    dim3_grouped g;
    const int rc = dim3_join_paths_grouped(context(), R.data, R.size(), S.data, S.size(), (uint32_t)sizeof(InT), voff,
                                           (uint32_t)sizeof(_Tv), &o, &g);
    if (rc != DIM3_OK) die("DIM3JoinAggregate doJoinAggregate", rc);
    if (g.n_groups > (uint64_t)(_Tcounter)~(_Tcounter)0) {
        printf("DIM3JoinAggregate: %llu rows do not fit the result counter type.\n", (unsigned long long)g.n_groups);
        exit(-1);
    }
    result.clear();
    if (g.n_groups) result.reserve((_Tcounter)g.n_groups);
    A aggregate_function((int)g.nz);
    const char* v1 = (const char*)g.v1;
    const char* v2 = (const char*)g.v2;
    auto value = [](const char* base, uint64_t i) {
        _Tv v;
        memcpy(&v, base + i * sizeof(_Tv), sizeof(_Tv));
        return v;
    };
    uint64_t row0 = 0;
    for (uint64_t i = 0; i < g.n_groups; i++) {
        const int key = (int)g.zkey[i];
        uint64_t p = g.path_start[i];
        const uint64_t pe = g.path_start[i + 1];
        aggregate_function.insert(key, value(v1, p), value(v2, p));
        for (++p; p < pe; ++p) aggregate_function.update(key, value(v1, p), value(v2, p));
        if (i + 1 == g.n_groups || g.x[i + 1] != g.x[i]) {
            for (uint64_t k = row0; k <= i; k++)
                result.push_back(OutT((_Tx)g.x[k], (_Tz)g.z[k], aggregate_function.get((int)g.zkey[k])));
            row0 = i + 1;
        }
    }
    dim3_grouped_free(&g);
}

}  // namespace dim3_shim

template <typename _Tx, typename _Ty, typename _Tz, typename _Tv, typename _Aggregate_type,
          typename _Hx = myHasher, typename _Hy = myHasher, typename _Hz = myHasher, typename _Hxz = myHasher,
          typename _Tcounter = uint32_t>
struct DIM3JoinAggregate {
    // The library's policies (COUNT, SUM, MIN, MAX, AVG) aggregate on the
    // device; any other _Aggregate_type is folded on the host over the
    // device's grouped join paths.
    static void doJoinAggregate(myvector<TUPLE<_Tx, _Ty, _Tv>, _Tcounter>& _R,
                                myvector<TUPLE<_Tz, _Ty, _Tv>, _Tcounter>& _S,
                                myvector<TUPLE<_Tx, _Tz, _Tv>, _Tcounter>& result) {
        run(_R, _S, result, std::integral_constant<bool, dim3_shim::is_library_policy<_Aggregate_type, _Tv>()>());
    }

private:
    static void run(myvector<TUPLE<_Tx, _Ty, _Tv>, _Tcounter>& _R, myvector<TUPLE<_Tz, _Ty, _Tv>, _Tcounter>& _S,
                    myvector<TUPLE<_Tx, _Tz, _Tv>, _Tcounter>& result, std::false_type) {
        dim3_shim::join_aggregate_policy<_Aggregate_type>(_R, _S, result);
    }

    static void run(myvector<TUPLE<_Tx, _Ty, _Tv>, _Tcounter>& _R, myvector<TUPLE<_Tz, _Ty, _Tv>, _Tcounter>& _S,
                    myvector<TUPLE<_Tx, _Tz, _Tv>, _Tcounter>& result, std::true_type) {
        static_assert(dim3_shim::is_i32<_Tx>() && dim3_shim::is_i32<_Ty>() && dim3_shim::is_i32<_Tz>(),
                      "libdim3 joins 4-byte integer attributes");
        static_assert(std::is_same<_Tv, float>::value || std::is_integral<_Tv>::value,
                      "tuple values are float or integers");
        static_assert(std::is_same<_Tv, float>::value || sizeof(_Tv) == 4 || sizeof(_Tv) == 8,
                      "integer tuple values of 4 or 8 bytes");
        typedef TUPLE<_Tx, _Ty, _Tv> InT;
        typedef TUPLE<_Tx, _Tz, _Tv> OutT;
        const int agg = dim3_shim::policy_code<_Aggregate_type, _Tv>::value;
        const bool f32 = std::is_same<_Tv, float>::value;
        // the tuples go to the device as they are (the value's offset from a
        // probe object: TUPLE has no default constructor)
        const InT probe((_Tx)0, (_Ty)0, (_Tv)0);
        const uint32_t voff = (uint32_t)((const char*)&probe.v - (const char*)&probe);
        const uint32_t vtype = f32 ? DIM3_VALUE_F32 : sizeof(_Tv) == 8 ? DIM3_VALUE_I64 : DIM3_VALUE_I32;
        result.clear();
        dim3_opts o;
        dim3_opts_default(&o);
        dim3_result res;
        int rc = dim3_join_aggregate_tuples(dim3_shim::context(), _R.data, _R.size(), _S.data, _S.size(),
                                            (uint32_t)sizeof(InT), voff, vtype, agg, &o, &res);
        if (rc != DIM3_OK) dim3_shim::die("DIM3JoinAggregate doJoinAggregate", rc);
        if (res.n > (uint64_t)(_Tcounter)~(_Tcounter)0) {
            dim3_result_free(&res);
            printf("DIM3JoinAggregate: %llu rows do not fit the result counter type.\n", (unsigned long long)res.n);
            exit(-1);
        }
        const OutT oprobe((_Tx)0, (_Tz)0, (_Tv)0);
        const uint32_t ooff = (uint32_t)((const char*)&oprobe.v - (const char*)&oprobe);
        // rows whose layout is the library's row record are handed over as they are
        const bool same_f32 = agg && f32 && sizeof(OutT) == sizeof(dim3_value_row_f32) && ooff == 8;
        const bool same_i64 = !f32 && std::is_integral<_Tv>::value && sizeof(_Tv) == 8 &&
                              sizeof(OutT) == sizeof(dim3_count_row) && ooff == 8;
        if (res.n && (same_f32 || same_i64)) {
            free(result.data);
            result.data = same_f32 ? (OutT*)(void*)res.rows_f32 : (OutT*)(void*)res.rows;
            res.rows_f32 = nullptr;
            res.rows = nullptr;
            result.cur_num = result.max_num = (_Tcounter)res.n;
        } else if (res.n) {
            result.reserve((_Tcounter)res.n);
            for (uint64_t i = 0; i < res.n; i++) {
                if (agg && f32)
                    result.data[i] = OutT((_Tx)res.rows_f32[i].x, (_Tz)res.rows_f32[i].z, (_Tv)res.rows_f32[i].v);
                else
                    result.data[i] = OutT((_Tx)res.rows[i].x, (_Tz)res.rows[i].z, (_Tv)res.rows[i].count);
            }
            result.cur_num = (_Tcounter)res.n;
        }
        dim3_result_free(&res);
    }
};

// gen_rand_data (dim3.hpp:1556-1568): std::minstd_rand0 (the libstdc++
// default_random_engine), R gets (x, y) draws then S gets (z, y) draws.
inline void gen_rand_data(myvector<pair<int, int>>& R, myvector<pair<int, int>>& S, int nR = 1e7, int nS = 1e7,
                          int MODX = 1e6, int MODY = 1e6, int MODZ = 1e6, int seed = 492) {
    R.clear();
    S.clear();
    std::minstd_rand0 e(seed);
    for (int i = 0; i < nR; i++) {
        int x = (int)(e() % MODX);
        int y = (int)(e() % MODY);
        R.push_back(pair<int, int>(x, y));
    }
    for (int i = 0; i < nS; i++) {
        int z = (int)(e() % MODZ);
        int y = (int)(e() % MODY);
        S.push_back(pair<int, int>(z, y));
    }
}

// TUPLE overloads (dim3.hpp:1570-1591): the same draws, value = tuple index.
template <typename _Tv>
inline void gen_rand_tuples(myvector<TUPLE<int, int, _Tv>>& R, myvector<TUPLE<int, int, _Tv>>& S, int nR, int nS,
                            int MODX, int MODY, int MODZ, int seed) {
    R.clear();
    S.clear();
    std::minstd_rand0 e(seed);
    for (int i = 0; i < nR; i++) {
        int x = (int)(e() % MODX);
        int y = (int)(e() % MODY);
        R.push_back(TUPLE<int, int, _Tv>(x, y, (_Tv)i));
    }
    for (int i = 0; i < nS; i++) {
        int z = (int)(e() % MODZ);
        int y = (int)(e() % MODY);
        S.push_back(TUPLE<int, int, _Tv>(z, y, (_Tv)i));
    }
}

inline void gen_rand_data(myvector<TUPLE<int, int, float>>& R, myvector<TUPLE<int, int, float>>& S, int nR = 1e7,
                          int nS = 1e7, int MODX = 1e6, int MODY = 1e6, int MODZ = 1e6, int seed = 492) {
    gen_rand_tuples(R, S, nR, nS, MODX, MODY, MODZ, seed);
}

inline void gen_rand_data(myvector<TUPLE<int, int, long long>>& R, myvector<TUPLE<int, int, long long>>& S,
                          int nR = 1e7, int nS = 1e7, int MODX = 1e6, int MODY = 1e6, int MODZ = 1e6,
                          int seed = 492) {
    gen_rand_tuples(R, S, nR, nS, MODX, MODY, MODZ, seed);
}
