// Caller-written aggregate policies through DIM3JoinAggregate (the
// reference's policy concept, dim3.hpp:1453-1480): each result is checked
// against a brute-force nested-loop evaluation of the same policy.
// Built by tests/test_policy.py with g++ -I include against libdim3.so.
#include <map>
#include <tuple>

#include "dim3.hpp"

// sum of (v1 + v2)^2 over a group's join paths, as int64
template <typename _Tv>
struct SumSquares {
    std::vector<_Tv> s;
    explicit SumSquares(int n) : s(n) {}
    void insert(const int key, const _Tv a, const _Tv b) { s[key] = (a + b) * (a + b); }
    void update(const int key, const _Tv a, const _Tv b) { s[key] += (a + b) * (a + b); }
    _Tv get(const int key) const { return s[key]; }
};

// order-sensitive: the value of the group's first path times 1000 plus its
// path count (the reference visits a row's paths in R tuple, then S tuple
// order)
template <typename _Tv>
struct FirstAndCount {
    std::vector<std::pair<_Tv, _Tv>> s;
    explicit FirstAndCount(int n) : s(n) {}
    void insert(const int key, const _Tv a, const _Tv b) { s[key] = {a * 100 + b, 1}; }
    void update(const int key, const _Tv, const _Tv) { s[key].second++; }
    _Tv get(const int key) const { return s[key].first * 1000 + s[key].second; }
};

// double values: max |v1 - v2|
struct MaxGap {
    std::vector<double> s;
    explicit MaxGap(int n) : s(n) {}
    void insert(const int key, const double a, const double b) { s[key] = a > b ? a - b : b - a; }
    void update(const int key, const double a, const double b) {
        const double d = a > b ? a - b : b - a;
        if (d > s[key]) s[key] = d;
    }
    double get(const int key) const { return s[key]; }
};

template <typename Policy, typename _Tv>
int check(const char* name, int nr, int ns, int dom, int keys, unsigned seed) {
    myvector<TUPLE<int, int, _Tv>> R, S;
    std::mt19937 rng(seed);
    for (int i = 0; i < nr; i++)
        R.push_back(TUPLE<int, int, _Tv>((int)(rng() % dom) - dom / 3, (int)(rng() % keys), (_Tv)(int)(rng() % 50)));
    for (int i = 0; i < ns; i++)
        S.push_back(TUPLE<int, int, _Tv>((int)(rng() % dom) - dom / 2, (int)(rng() % keys), (_Tv)(int)(rng() % 50)));
    myvector<TUPLE<int, int, _Tv>> out;
    DIM3JoinAggregate<int, int, int, _Tv, Policy>::doJoinAggregate(R, S, out);
    // brute force: per x in first-R-appearance order is not needed, only
    // per (x, z) the paths in (R index, S index) order
    std::map<std::pair<int, int>, std::vector<std::pair<_Tv, _Tv>>> paths;
    for (int i = 0; i < nr; i++)
        for (int j = 0; j < ns; j++)
            if (R.data[i].y == S.data[j].y) paths[{R.data[i].x, S.data[j].x}].push_back({R.data[i].v, S.data[j].v});
    std::map<int, int> zkey;
    for (int j = 0; j < ns; j++) zkey[S.data[j].x] = 0;
    int k = 0;
    for (auto& kv : zkey) kv.second = k++;
    Policy ref((int)zkey.size());
    std::map<std::pair<int, int>, _Tv> want;
    for (auto& kv : paths) {
        const int key = zkey[kv.first.second];
        ref.insert(key, kv.second[0].first, kv.second[0].second);
        for (size_t p = 1; p < kv.second.size(); p++) ref.update(key, kv.second[p].first, kv.second[p].second);
        want[kv.first] = ref.get(key);
    }
    if ((size_t)out.size() != want.size()) {
        printf("%s: %u rows, want %zu\n", name, (unsigned)out.size(), want.size());
        return 1;
    }
    for (unsigned i = 0; i < out.size(); i++) {
        auto it = want.find({out.data[i].x, out.data[i].y});
        if (it == want.end() || !(it->second == out.data[i].v)) {
            printf("%s: row (%d, %d) wrong\n", name, out.data[i].x, out.data[i].y);
            return 1;
        }
    }
    printf("%s: %u rows ok\n", name, (unsigned)out.size());
    return 0;
}

int main() {
    int bad = 0;
    bad += check<SumSquares<long long>, long long>("sum_squares_i64", 3000, 2500, 400, 60, 1);
    bad += check<SumSquares<int>, int>("sum_squares_i32", 2000, 4000, 300, 40, 2);
    bad += check<FirstAndCount<long long>, long long>("first_and_count", 2500, 2500, 200, 30, 3);
    bad += check<MaxGap, double>("max_gap_f64", 1500, 1500, 150, 25, 4);
    bad += check<SumSquares<long long>, long long>("empty_join", 100, 100, 50, 1000000, 5);
    // the library's own policy still runs on the device
    myvector<TUPLE<int, int, int>> R, S, out;
    for (int i = 0; i < 200; i++) {
        R.push_back(TUPLE<int, int, int>(i % 17, i % 11, i));
        S.push_back(TUPLE<int, int, int>(i % 13, i % 11, 2 * i));
    }
    DIM3JoinAggregate<int, int, int, int, AggregateFunctionAvg<int>>::doJoinAggregate(R, S, out);
    printf("avg: %u rows\n", (unsigned)out.size());
    if (out.size() == 0) bad++;
    printf(bad ? "FAILED\n" : "ALL OK\n");
    return bad ? 1 : 0;
}
