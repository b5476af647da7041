// EstimateJoinCardinality::estimate (dim3.hpp:190-270) on the device: the
// reference's sampled join size, the same number it returns.
//
// With the identity hash myHasher (Include/structures.hpp:92-95) the sample
// is every tuple whose key has its low `bits` bits zero, bits being the
// number of halvings that bring |S| to at most 10000 (dim3.hpp:262-269);
// the estimate is the sampled join-path count shifted back up.  The count
// is a key histogram of the sampled S tuples in a device hash table and a
// probe of the sampled R tuples.  When the sample has no path the reference
// counts the paths between the first 1000 tuples of each side and scales by
// |R| / 1e3 * |S| / 1e3 (dim3.hpp:215-221); here the prefixes are clamped to
// the input sizes (the reference reads past the end of shorter inputs).
#include "dim3_internal.h"

namespace {

using namespace dim3d;

constexpr uint32_t kThreads = 256;
// grid-stride launches: each workgroup folds many tuples before touching
// global counters
constexpr uint32_t kMaxBlocks = 2048;
// workgroup-private key counts: a heavy sampled key (the identity sample
// takes every key with its low bits zero, key 0 included, often a hub) would
// otherwise put one global atomic per tuple on a single counter
constexpr uint32_t kLdsSlots = 1024;
constexpr uint32_t kLdsProbes = 8;

__global__ __launch_bounds__(kThreads) void k_est_count_s(const int2* __restrict__ S, uint64_t nS, uint32_t smask,
                                                          uint32_t* keys, uint32_t hmask, uint32_t* cnt) {
    __shared__ uint32_t lk[kLdsSlots], lc[kLdsSlots];
    for (uint32_t j = threadIdx.x; j < kLdsSlots; j += kThreads) {
        lk[j] = DIM3_EMPTY_KEY;
        lc[j] = 0;
    }
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < nS; i += (uint64_t)gridDim.x * kThreads) {
        const uint32_t y = (uint32_t)S[i].y;
        if ((y & smask) != 0u) continue;
        bool done = false;
        if (y != DIM3_EMPTY_KEY) {
            uint32_t j = hash32(y) & (kLdsSlots - 1u);
            for (uint32_t p = 0; p < kLdsProbes && !done; p++, j = (j + 1u) & (kLdsSlots - 1u)) {
                const uint32_t cur = atomicCAS(&lk[j], DIM3_EMPTY_KEY, y);  // claims, or names the key there
                if (cur == DIM3_EMPTY_KEY || cur == y) {
                    atomicAdd(&lc[j], 1u);
                    done = true;
                }
            }
        }
        if (!done) atomicAdd(&cnt[ht_insert(keys, hmask, y)], 1u);  // a full neighbourhood: straight to the table
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < kLdsSlots; j += kThreads)
        if (lc[j] != 0u) atomicAdd(&cnt[ht_insert(keys, hmask, lk[j])], lc[j]);
}

__global__ __launch_bounds__(kThreads) void k_est_probe_r(const int2* __restrict__ R, uint64_t nR, uint32_t smask,
                                                          const uint32_t* keys, uint32_t hmask,
                                                          const uint32_t* __restrict__ cnt,
                                                          unsigned long long* out) {
    __shared__ uint64_t sh[kThreads / 64 + 1];
    uint64_t v = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < nR; i += (uint64_t)gridDim.x * kThreads) {
        const uint32_t y = (uint32_t)R[i].y;
        if ((y & smask) == 0u) {
            const int32_t s = ht_find(keys, hmask, y);
            if (s >= 0) v += cnt[s];
        }
    }
    v = block_reduce_sum<kThreads, uint64_t>(v, sh);
    if (threadIdx.x == 0 && v) atomicAdd(out, (unsigned long long)v);
}

// paths between R[0, a) and S[0, b) (a, b <= 1000): one workgroup
__global__ __launch_bounds__(1024) void k_est_prefix(const int2* __restrict__ R, uint32_t a,
                                                     const int2* __restrict__ S, uint32_t b,
                                                     unsigned long long* out) {
    __shared__ int32_t ry[1000];
    __shared__ uint64_t sh[1024 / 64 + 1];
    for (uint32_t j = threadIdx.x; j < a; j += 1024) ry[j] = R[j].y;
    __syncthreads();
    uint64_t c = 0;
    if (threadIdx.x < b) {
        const int32_t y = S[threadIdx.x].y;
        for (uint32_t j = 0; j < a; j++) c += ry[j] == y ? 1u : 0u;
    }
    c = block_reduce_sum<1024, uint64_t>(c, sh);
    if (threadIdx.x == 0) *out = c;
}

}  // namespace

namespace dim3h {

void estimate_sample(const int2* R, uint64_t nR, const int2* S, uint64_t nS, uint32_t smask, uint32_t* keys,
                     uint32_t hmask, uint32_t* cnt, unsigned long long* out, hipStream_t st) {
    auto blocks = [](uint64_t n) { return (unsigned)std::min<uint64_t>((n + kThreads - 1) / kThreads, kMaxBlocks); };
    if (nS)
        hipLaunchKernelGGL(k_est_count_s, dim3(blocks(nS)), dim3(kThreads), 0, st, S, nS, smask, keys, hmask, cnt);
    if (nR)
        hipLaunchKernelGGL(k_est_probe_r, dim3(blocks(nR)), dim3(kThreads), 0, st, R, nR, smask,
                           (const uint32_t*)keys, hmask, (const uint32_t*)cnt, out);
}

void estimate_prefix(const int2* R, uint32_t a, const int2* S, uint32_t b, unsigned long long* out, hipStream_t st) {
    hipLaunchKernelGGL(k_est_prefix, dim3(1), dim3(1024), 0, st, R, a, S, b, out);
}

}  // namespace dim3h
