// Thin extern "C" driver around the reference DIM3 headers, compiled from
// the reference sources where they lie (see oracle/Makefile) into
// oracle/_ref/libdim3ref.so.  TEST INFRASTRUCTURE ONLY: it is the checker of
// the CPU oracle and the timed CPU baseline, never part of the product.
//
// Entry points it wraps:
//   DIM3<int,int,int,myHasher x4,uint32_t>::doJoinProject          dim3.hpp:987
//   DIM3_parallel<...>::doJoinProject_parallel (OpenMP, 8 threads)  dim3_parallel.hpp:142
//   DIM3JoinAggregate<int,int,int,long long,AggCount,...>           dim3.hpp:1318
//   DIM3JoinAggregate<int,int,int,long long,AggSum|AggMin|AggMax,...> and
//   DIM3JoinAggregate<...,AggregateFunctionAvg<long long>,...>      dim3.hpp:1318, 1453
//   EstimateJoinCardinality<...>::estimate                          dim3.hpp:193
//
// Guards (the reference's own behaviour there is undefined, see the
// estimate fallback at dim3.hpp:215-221 which reads 1000 tuples even from a
// shorter input, and the classical branch of the parallel build which
// asserts at dim3_parallel.hpp:159-160):
//   - empty inputs return an empty result without calling the reference,
//   - inputs under 1000 tuples with no join path return an empty result,
//   - the parallel entry point runs the serial DIM3 when f1 picks the
//     classical plan.
#include "dim3_parallel.hpp"

#include <cstdint>
#include <cstdlib>
#include <cstring>

// COUNT policy for the aggregate concept (ctor(int), insert, update, get),
// as used at dim3.hpp:1275,1295,1299,1305.
struct AggCount {
    long long* c;
    explicit AggCount(int n) { c = (long long*)malloc((size_t)(n > 0 ? n : 1) * sizeof(long long)); }
    ~AggCount() { free(c); }
    inline void insert(const int key, const long long, const long long) { c[key] = 1; }
    inline void update(const int key, const long long, const long long) { c[key] += 1; }
    inline long long get(const int key) const { return c[key]; }
};

// Value policies over the path value v1 + v2 (the combination the shipped
// AggregateFunctionAvg uses, dim3.hpp:1466-1473): SUM, MIN, MAX.
template <int OP, typename T = long long>
struct AggValue {
    T* a;
    explicit AggValue(int n) { a = (T*)malloc((size_t)(n > 0 ? n : 1) * sizeof(T)); }
    ~AggValue() { free(a); }
    inline void insert(const int key, const T v1, const T v2) { a[key] = v1 + v2; }
    inline void update(const int key, const T v1, const T v2) {
        const T v = v1 + v2;
        if (OP == 0) a[key] += v;
        else if (OP == 1) a[key] = v < a[key] ? v : a[key];
        else a[key] = v > a[key] ? v : a[key];
    }
    inline T get(const int key) const { return a[key]; }
};

typedef DIM3<int, int, int, myHasher, myHasher, myHasher, myHasher, uint32_t> RefDIM3;
typedef DIM3_parallel<int, int, int, myHasher, myHasher, myHasher, myHasher, uint32_t> RefDIM3P;
typedef DIM3JoinAggregate<int, int, int, long long, AggCount, myHasher, myHasher, myHasher, myHasher, uint32_t>
    RefAgg;

static myvector<pair<int, int>> wrap_pairs(const int32_t* p, uint32_t n) {
    myvector<pair<int, int>> v(1);
    free(v.data);
    v.data = (pair<int, int>*)p;
    v.cur_num = n;
    v.max_num = n;
    return v;
}

static bool has_paths_small(const int32_t* R, uint32_t nR, const int32_t* S, uint32_t nS) {
    std::unordered_map<int, int> m;
    for (uint32_t i = 0; i < nR; i++) m[R[2 * i + 1]]++;
    for (uint32_t i = 0; i < nS; i++)
        if (m.count(S[2 * i + 1])) return true;
    return false;
}

static bool guard_empty(const int32_t* R, uint32_t nR, const int32_t* S, uint32_t nS) {
    if (nR == 0 || nS == 0) return true;
    if ((nR < 1000 || nS < 1000) && !has_paths_small(R, nR, S, nS)) return true;
    return false;
}

template <typename Pol, typename T = long long>
static void run_agg(myvector<TUPLE<int, int, T>>& vr, myvector<TUPLE<int, int, T>>& vs,
                    myvector<TUPLE<int, int, T>>& result) {
    DIM3JoinAggregate<int, int, int, T, Pol, myHasher, myHasher, myHasher, myHasher, uint32_t>::doJoinAggregate(
        vr, vs, result);
}

extern "C" {

int ref_join_project(const int32_t* R, uint32_t nR, const int32_t* S, uint32_t nS, int parallel,
                     int32_t** out, uint64_t* nout) {
    *out = NULL;
    *nout = 0;
    if (guard_empty(R, nR, S, nS)) return 0;
    myvector<pair<int, int>> vr = wrap_pairs(R, nR), vs = wrap_pairs(S, nS);
    myvector<pair<int, int>> result;
    if (parallel) {
        // mirror the plan choice of doJoinProject_parallel (dim3_parallel.hpp:151-162)
        const myvector<pair<int, int>>& big = nR >= nS ? vr : vs;
        const myvector<pair<int, int>>& small = nR >= nS ? vs : vr;
        LL est = EstimateJoinCardinality<int, int, int, myHasher, uint32_t>::estimate(big, small);
        if (f1_threshold::is_classical_batter(big.size(), small.size(), est))
            RefDIM3::doJoinProject(vr, vs, result);
        else
            RefDIM3P::doJoinProject_parallel(vr, vs, result);
    } else {
        RefDIM3::doJoinProject(vr, vs, result);
    }
    *out = (int32_t*)result.data;
    *nout = result.cur_num;
    return 0;
}

// Rows are (x, z, count) with the int64 count; returned as two arrays.
int ref_join_aggregate_count(const int32_t* R, uint32_t nR, const int32_t* S, uint32_t nS, int32_t** out_pairs,
                             int64_t** out_counts, uint64_t* nout) {
    *out_pairs = NULL;
    *out_counts = NULL;
    *nout = 0;
    if (guard_empty(R, nR, S, nS)) return 0;
    myvector<TUPLE<int, int, long long>> vr(nR), vs(nS), result;
    for (uint32_t i = 0; i < nR; i++) vr.push_back(TUPLE<int, int, long long>(R[2 * i], R[2 * i + 1], 0));
    for (uint32_t i = 0; i < nS; i++) vs.push_back(TUPLE<int, int, long long>(S[2 * i], S[2 * i + 1], 0));
    RefAgg::doJoinAggregate(vr, vs, result);
    uint32_t n = result.cur_num;
    int32_t* p = (int32_t*)malloc((size_t)(n ? n : 1) * 2 * sizeof(int32_t));
    int64_t* c = (int64_t*)malloc((size_t)(n ? n : 1) * sizeof(int64_t));
    for (uint32_t i = 0; i < n; i++) {
        p[2 * i] = result.data[i].x;
        p[2 * i + 1] = result.data[i].y;
        c[i] = result.data[i].v;
    }
    free(result.data);
    free(vr.data);
    free(vs.data);
    *out_pairs = p;
    *out_counts = c;
    *nout = n;
    return 0;
}

// Value aggregate agg (1 SUM, 2 MIN, 3 MAX, 4 AVG through the reference's own
// AggregateFunctionAvg<long long>) of R(x, y, vR) join S(z, y, vS).
int ref_join_aggregate(const int32_t* R, const int32_t* vR, uint32_t nR, const int32_t* S, const int32_t* vS,
                       uint32_t nS, int agg, int32_t** out_pairs, int64_t** out_vals, uint64_t* nout) {
    *out_pairs = NULL;
    *out_vals = NULL;
    *nout = 0;
    if (agg < 1 || agg > 4) return 1;
    if (guard_empty(R, nR, S, nS)) return 0;
    myvector<TUPLE<int, int, long long>> vr(nR), vs(nS), result;
    for (uint32_t i = 0; i < nR; i++) vr.push_back(TUPLE<int, int, long long>(R[2 * i], R[2 * i + 1], vR[i]));
    for (uint32_t i = 0; i < nS; i++) vs.push_back(TUPLE<int, int, long long>(S[2 * i], S[2 * i + 1], vS[i]));
    if (agg == 1) run_agg<AggValue<0>>(vr, vs, result);
    else if (agg == 2) run_agg<AggValue<1>>(vr, vs, result);
    else if (agg == 3) run_agg<AggValue<2>>(vr, vs, result);
    else run_agg<AggregateFunctionAvg<long long>>(vr, vs, result);
    uint32_t n = result.cur_num;
    int32_t* p = (int32_t*)malloc((size_t)(n ? n : 1) * 2 * sizeof(int32_t));
    int64_t* c = (int64_t*)malloc((size_t)(n ? n : 1) * sizeof(int64_t));
    for (uint32_t i = 0; i < n; i++) {
        p[2 * i] = result.data[i].x;
        p[2 * i + 1] = result.data[i].y;
        c[i] = result.data[i].v;
    }
    free(result.data);
    free(vr.data);
    free(vs.data);
    *out_pairs = p;
    *out_vals = c;
    *nout = n;
    return 0;
}

// Float values: the reference's own AggregateFunctionAvg<float> (the
// join_aggregate.cpp workload) for AVG, float SUM / MIN / MAX policies else;
// float accumulation in first-touch order.
int ref_join_aggregate_f32(const int32_t* R, const float* vR, uint32_t nR, const int32_t* S, const float* vS,
                           uint32_t nS, int agg, int32_t** out_pairs, float** out_vals, uint64_t* nout) {
    *out_pairs = NULL;
    *out_vals = NULL;
    *nout = 0;
    if (agg < 1 || agg > 4) return 1;
    if (guard_empty(R, nR, S, nS)) return 0;
    myvector<TUPLE<int, int, float>> vr(nR), vs(nS), result;
    for (uint32_t i = 0; i < nR; i++) vr.push_back(TUPLE<int, int, float>(R[2 * i], R[2 * i + 1], vR[i]));
    for (uint32_t i = 0; i < nS; i++) vs.push_back(TUPLE<int, int, float>(S[2 * i], S[2 * i + 1], vS[i]));
    if (agg == 1) run_agg<AggValue<0, float>, float>(vr, vs, result);
    else if (agg == 2) run_agg<AggValue<1, float>, float>(vr, vs, result);
    else if (agg == 3) run_agg<AggValue<2, float>, float>(vr, vs, result);
    else run_agg<AggregateFunctionAvg<float>, float>(vr, vs, result);
    uint32_t n = result.cur_num;
    int32_t* p = (int32_t*)malloc((size_t)(n ? n : 1) * 2 * sizeof(int32_t));
    float* c = (float*)malloc((size_t)(n ? n : 1) * sizeof(float));
    for (uint32_t i = 0; i < n; i++) {
        p[2 * i] = result.data[i].x;
        p[2 * i + 1] = result.data[i].y;
        c[i] = result.data[i].v;
    }
    free(result.data);
    free(vr.data);
    free(vs.data);
    *out_pairs = p;
    *out_vals = c;
    *nout = n;
    return 0;
}

long long ref_estimate(const int32_t* R, uint32_t nR, const int32_t* S, uint32_t nS) {
    myvector<pair<int, int>> vr = wrap_pairs(R, nR), vs = wrap_pairs(S, nS);
    return EstimateJoinCardinality<int, int, int, myHasher, uint32_t>::estimate(vr, vs);
}

void ref_gen_rand_data(int32_t* R, uint32_t nR, int32_t* S, uint32_t nS, int modx, int mody, int modz, int seed) {
    myvector<pair<int, int>> vr, vs;
    gen_rand_data(vr, vs, (int)nR, (int)nS, modx, mody, modz, seed);
    memcpy(R, vr.data, (size_t)nR * 8);
    memcpy(S, vs.data, (size_t)nS * 8);
    free(vr.data);
    free(vs.data);
}

int ref_thread_num(void) { return THREAD_NUM; }

void ref_free(void* p) { free(p); }
}
