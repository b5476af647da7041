// Counterpart of the reference drivers main.cpp, main_parallel.cpp and
// join_aggregate.cpp (the same data and the same two output lines), built
// against the drop-in header include/dim3.hpp:
//   dim3_main [project|parallel|aggregate|count|estimate|estimate_tuples]
// aggregate is join_aggregate.cpp's query, AVG(R.v + S.v) over
// TUPLE<int,int,float> (join_aggregate.cpp:6-23); count is the COUNT
// aggregate over TUPLE<int,int,long long>; estimate prints
// EstimateJoinCardinality::estimate of the same data (|Estimate|=...), and
// estimate_tuples its TUPLE overload over the TUPLE<int,int,float> data.
#include <chrono>
#include <cstring>

#include "dim3.hpp"

int main(int argc, char* argv[]) {
    const char* mode = argc > 1 ? argv[1] : "project";
    std::chrono::high_resolution_clock::time_point t0, t1;
    unsigned n = 0;
    if (!strcmp(mode, "estimate")) {
        myvector<pair<int, int>> R, S;
        gen_rand_data(R, S);
        t0 = std::chrono::high_resolution_clock::now();
        const long long est = EstimateJoinCardinality<int, int, int, myHasher, uint32_t>::estimate(R, S);
        t1 = std::chrono::high_resolution_clock::now();
        printf("It takes %f seconds.\n", std::chrono::duration<double, std::ratio<1, 1>>(t1 - t0).count());
        printf("|Estimate|=%lld\n", est);
        return 0;
    }
    if (!strcmp(mode, "estimate_tuples")) {  // the TUPLE overload, the reference's default hasher
        myvector<TUPLE<int, int, float>> R, S;
        gen_rand_data(R, S);
        t0 = std::chrono::high_resolution_clock::now();
        const long long est = EstimateJoinCardinality<int, int, int>::estimate(R, S);
        t1 = std::chrono::high_resolution_clock::now();
        printf("It takes %f seconds.\n", std::chrono::duration<double, std::ratio<1, 1>>(t1 - t0).count());
        printf("|Estimate|=%lld\n", est);
        return 0;
    }
    if (!strcmp(mode, "aggregate")) {
        myvector<TUPLE<int, int, float>> R, S, result;
        gen_rand_data(R, S);
        result.clear();
        t0 = std::chrono::high_resolution_clock::now();
        DIM3JoinAggregate<int, int, int, float, AggregateFunctionAvg<float>, myHasher, myHasher, myHasher, myHasher,
                          uint32_t>::doJoinAggregate(R, S, result);
        t1 = std::chrono::high_resolution_clock::now();
        n = result.size();
    } else if (!strcmp(mode, "count")) {
        myvector<TUPLE<int, int, long long>> R, S, result;
        gen_rand_data(R, S);
        result.clear();
        t0 = std::chrono::high_resolution_clock::now();
        DIM3JoinAggregate<int, int, int, long long, AggregateFunctionCount<long long>, myHasher, myHasher, myHasher,
                          myHasher, uint32_t>::doJoinAggregate(R, S, result);
        t1 = std::chrono::high_resolution_clock::now();
        n = result.size();
    } else {
        myvector<pair<int, int>> R, S, result;
        gen_rand_data(R, S);
        result.clear();
        t0 = std::chrono::high_resolution_clock::now();
        if (!strcmp(mode, "parallel"))
            DIM3_parallel<int, int, int, myHasher, myHasher, myHasher, myHasher, uint32_t>::doJoinProject_parallel(
                R, S, result);
        else
            DIM3<int, int, int, myHasher, myHasher, myHasher, myHasher, uint32_t>::doJoinProject(R, S, result);
        t1 = std::chrono::high_resolution_clock::now();
        n = result.size();
    }
    printf("It takes %f seconds.\n", std::chrono::duration<double, std::ratio<1, 1>>(t1 - t0).count());
    printf("|Result|=%u\n", n);
    return 0;
}
