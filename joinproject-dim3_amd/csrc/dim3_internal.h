// Internal declarations shared by the libdim3 translation units.
//
// Device data layout (all in HBM, one set per context, reused across calls):
//   relations        int2[n]        (x|z, y) tuples exactly as the caller passes them
//   hash tables      u32 keys[cap+1] open addressing, linear probing, cap = 2^k >= 2n;
//                    0xFFFFFFFF marks an empty slot, slot `cap` stands for the key -1
//   degree arrays    u32 deg[cap+1]  per slot
//   prefix arrays    u64 ex[cap+1]   exclusive scan of (occupied << 32 | degree):
//                                    high word = dense ID, low word = CSR row start
//   CSR of R by x    u32 rowptr[nx+1], i32 col[nnz] (y IDs)
//   CSR of S by (y, z-tile)  u32 rowptr[ny*ntiles+1], i32 col[nnz] (z IDs)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <mutex>
#include <shared_mutex>
#include <string>

#include "../../include/dim3.h"

#define DIM3_EMPTY_KEY 0xFFFFFFFFu

namespace dim3d {

// ---------------------------------------------------------------------------
// hashing (murmur3 fmix32 finaliser; the reference uses the identity hash on
// int plus radix partitioning, Include/structures.hpp:92-95 — the hash only
// places keys, it never changes a result)
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t hash32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// splitmix64 finaliser, used for the order-independent row fingerprint
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t k) {
    k ^= k >> 30;
    k *= 0xbf58476d1ce4e5b9ULL;
    k ^= k >> 27;
    k *= 0x94d049bb133111ebULL;
    k ^= k >> 31;
    return k;
}

// Row fingerprint factors.  A row (x, z) with count c adds k1(x) g(z) to
// sum and k2(x) c g(z) to sum2 (mod 2^64, u(v) = (u32)v):
//   k1 = mix64(u(x) ^ 0x243F6A8885A308D3) | 1, k2 = mix64(u(x) ^ 0x13198A2E03707344) | 1,
//   g(z) = a(u(z) >> 5) * c(u(z) & 31), a(w) = fmix32(w ^ 0x243F6A88) | 1,
//   c(b) = mix64(0x9E3779B97F4A7C15 * (b + 1)) | 1.
// The z factors of one 32-aligned word of z values share a(w), so a word of
// a dense z range sums as a(w) times the position factors of its set bits.
__host__ __device__ __forceinline__ uint32_t zword_factor(uint32_t w) { return hash32(w ^ 0x243F6A88u) | 1u; }
__host__ __device__ __forceinline__ uint64_t zpos_factor(uint32_t b) {
    return mix64(0x9E3779B97F4A7C15ull * (uint64_t)(b + 1u)) | 1ull;
}
__host__ __device__ __forceinline__ uint64_t zfactor(int32_t z) {
    const uint32_t u = (uint32_t)z;
    return (uint64_t)zword_factor(u >> 5) * zpos_factor(u & 31u);
}
// LDS table of the fingerprint emission: [32] position factors, then
// [4][256] sums of the position factors of the set bits of byte j of a word
constexpr uint32_t kFpTableU64 = 32 + 4 * 256;
__host__ __device__ __forceinline__ uint64_t xfactor1(int32_t x) {
    return mix64((uint64_t)(uint32_t)x ^ 0x243F6A8885A308D3ull) | 1ull;
}
__host__ __device__ __forceinline__ uint64_t xfactor2(int32_t x) {
    return mix64((uint64_t)(uint32_t)x ^ 0x13198A2E03707344ull) | 1ull;
}

__device__ __forceinline__ uint64_t row_key(int32_t x, int32_t z) {
    return ((uint64_t)(uint32_t)x << 32) | (uint64_t)(uint32_t)z;
}

// Insert k, return its slot.  Keys only ever go EMPTY -> key, so a stale
// plain read is either EMPTY (then the CAS settles it) or the final key.
__device__ __forceinline__ uint32_t ht_insert(uint32_t* keys, uint32_t mask, uint32_t k) {
    if (k == DIM3_EMPTY_KEY) {
        if (keys[mask + 1] == 0) keys[mask + 1] = 1;  // benign race, every writer stores 1
        return mask + 1;
    }
    uint32_t s = hash32(k) & mask;
    for (;;) {
        uint32_t cur = keys[s];
        if (cur == k) return s;
        if (cur == DIM3_EMPTY_KEY) {
            uint32_t old = atomicCAS(&keys[s], DIM3_EMPTY_KEY, k);
            if (old == DIM3_EMPTY_KEY || old == k) return s;
        }
        s = (s + 1) & mask;
    }
}

// Slot of k, or -1.
__device__ __forceinline__ int32_t ht_find(const uint32_t* keys, uint32_t mask, uint32_t k) {
    if (k == DIM3_EMPTY_KEY) return keys[mask + 1] ? (int32_t)(mask + 1) : -1;
    uint32_t s = hash32(k) & mask;
    for (;;) {
        uint32_t cur = keys[s];
        if (cur == k) return (int32_t)s;
        if (cur == DIM3_EMPTY_KEY) return -1;
        s = (s + 1) & mask;
    }
}

__device__ __forceinline__ int32_t ht_key(const uint32_t* keys, uint32_t cap, uint32_t s) {
    return s < cap ? (int32_t)keys[s] : -1;
}

// ---------------------------------------------------------------------------
// wave / block primitives (64-wide wavefronts)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// Inclusive u32 scan over the wavefront with DPP moves (VALU only, no LDS
// round trip): Hillis-Steele inside each row of 16 lanes, then the row
// totals are carried with row_bcast:15 (into rows 1, 3) and row_bcast:31
// (into rows 2, 3).
__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_u32(v), 63);
}

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_reduce_xor(T v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v ^= __shfl_xor(v, d, 64);
    return v;
}

// Index of v in a hot-key list {count, keys...} (the most frequent values
// of an attribute in a sample, hot_keys), or -1; hot may be null.
__device__ __forceinline__ int hot_index(const int32_t* hot, int32_t v) {
    if (hot == nullptr) return -1;
    const int n = hot[0];
    int h = -1;
    for (int k = 0; k < n; k++)
        if (v == hot[1 + k]) h = k;
    return h;
}

// Workgroup tallies of the hot keys.  Same-address global atomics
// serialise, about 8 ns apart on MI355X, so a per-tuple pass over a
// relation in which one key holds 50,000 tuples (a graph hub, a Zipf head)
// takes ~0.4 ms on that key alone.  The counting passes therefore loop over
// many tuples per workgroup (grid capped at kTallyBlocks) and count each hot
// key's tuples in LDS, with the key's slot, adding each tally to its global
// counter once per workgroup (tally_flush); the other keys keep the
// wavefront-aggregated atomics below.
constexpr unsigned kTallyBlocks = 2048;
constexpr int kTallyKeys = 15;  // = the hot-key lists' length (dim3h::kHotKeys)

template <int BLOCK>
__device__ __forceinline__ void tally_init(uint32_t* tally, uint32_t* tslot, int n) {
    for (int k = (int)threadIdx.x; k < n; k += BLOCK) {
        tally[k] = 0;
        if (tslot != nullptr) tslot[k] = 0;
    }
    __syncthreads();
}

// counter adds (MAX = false) or running maxima (MAX = true) of n tallies
template <int BLOCK, bool MAX = false>
__device__ __forceinline__ void tally_flush(uint32_t* ctr, const uint32_t* tally, const uint32_t* tslot, int n) {
    __syncthreads();
    for (int k = (int)threadIdx.x; k < n; k += BLOCK) {
        if (tally[k] == 0u) continue;
        if (MAX) atomicMax(&ctr[tslot[k]], tally[k]);
        else atomicAdd(&ctr[tslot[k]], tally[k]);
    }
}

// Wavefront-aggregated counter updates for skewed keys (Zipf join keys, hub
// vertices): in two rounds, the lanes whose key equals the first pending
// lane's key share one atomic; the remaining lanes add on their own.  Every
// lane still running must call these (valid = false lanes take part without
// counting).
// hot: the lane's index among the side's most frequent keys (hot_index), or
// -1; the lanes of each hot key present share one atomic first, then the
// rounds above take the rest.
__device__ __forceinline__ void wave_count_add(uint32_t* ctr, uint32_t key, bool valid, int hot = -1) {
    uint64_t pending = __ballot(valid);
    uint64_t hp = __ballot(valid && hot >= 0);
    pending &= ~hp;
    while (hp != 0ull) {  // at most kHotKeys rounds
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)hp) - 1u;
        const int lh = __builtin_amdgcn_readlane(hot, (int)leader);
        const uint64_t same = __ballot(hot == lh) & hp;
        if (lane_id() == leader) atomicAdd(&ctr[key], (uint32_t)__popcll(same));
        hp &= ~same;
    }
#pragma unroll
    for (int r = 0; r < 2; r++) {
        if (pending == 0) return;
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)leader);
        const uint64_t same = __ballot(valid && key == lk) & pending;
        if (lane_id() == leader) atomicAdd(&ctr[lk], (uint32_t)__popcll(same));
        pending &= ~same;
    }
    if ((pending >> lane_id()) & 1ull) atomicAdd(&ctr[key], 1u);
}

// As wave_count_add, returning each valid lane's reserved position (the
// counter's old value plus the lane's rank among the lanes sharing its atomic).
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* ctr, uint32_t key, bool valid) {
    uint64_t pending = __ballot(valid);
    uint32_t pos = 0;
#pragma unroll
    for (int r = 0; r < 2; r++) {
        if (pending == 0) break;
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)leader);
        const uint64_t same = __ballot(valid && key == lk) & pending;
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(&ctr[lk], (uint32_t)__popcll(same));
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, (int)leader);
        if ((same >> lane_id()) & 1ull) pos = base + mbcnt64(same);
        pending &= ~same;
    }
    if ((pending >> lane_id()) & 1ull) pos = atomicAdd(&ctr[key], 1u);
    return pos;
}

// As wave_count_add for a running maximum: the lanes sharing the first
// pending lane's key (two rounds) send one atomicMax of their largest value.
__device__ __forceinline__ void wave_max_update(uint32_t* arr, uint32_t key, uint32_t v, bool valid) {
    uint64_t pending = __ballot(valid);
#pragma unroll
    for (int r = 0; r < 2; r++) {
        if (pending == 0) return;
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)leader);
        const bool mine = valid && key == lk && ((pending >> lane_id()) & 1ull);
        uint32_t m = mine ? v : 0u;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d, 64));
        if (lane_id() == leader && m > arr[lk]) atomicMax(&arr[lk], m);
        pending &= ~__ballot(mine);
    }
    if (((pending >> lane_id()) & 1ull) && v > arr[key]) atomicMax(&arr[key], v);
}

// As wave_count_add for 64-bit sums (mod 2^64): the lanes sharing the first
// pending lane's key (two rounds) send one atomicAdd of their total.
__device__ __forceinline__ void wave_add_u64(uint64_t* arr, uint32_t key, uint64_t v, bool valid, int hot = -1) {
    uint64_t pending = __ballot(valid);
    uint64_t hp = __ballot(valid && hot >= 0);
    pending &= ~hp;
    while (hp != 0ull) {  // the hot keys' lanes first (see wave_count_add)
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)hp) - 1u;
        const int lh = __builtin_amdgcn_readlane(hot, (int)leader);
        const uint64_t same = __ballot(hot == lh) & hp;
        uint64_t m = ((same >> lane_id()) & 1ull) ? v : 0ull;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) m += (uint64_t)__shfl_xor((unsigned long long)m, d, 64);
        if (lane_id() == leader) atomicAdd((unsigned long long*)&arr[key], (unsigned long long)m);
        hp &= ~same;
    }
#pragma unroll
    for (int r = 0; r < 2; r++) {
        if (pending == 0) return;
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)leader);
        const bool mine = valid && key == lk && ((pending >> lane_id()) & 1ull);
        uint64_t m = mine ? v : 0ull;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) m += (uint64_t)__shfl_xor((unsigned long long)m, d, 64);
        if (lane_id() == leader) atomicAdd((unsigned long long*)&arr[lk], (unsigned long long)m);
        pending &= ~__ballot(mine);
    }
    if ((pending >> lane_id()) & 1ull) atomicAdd((unsigned long long*)&arr[key], (unsigned long long)v);
}

// Exclusive scan over the block; sh needs BLOCK/64+1 entries.  Ends with a
// barrier so sh can be reused right away.
template <typename T>
__device__ __forceinline__ T wave_scan_any(T v) {
    return wave_inclusive_scan(v);
}
template <>
__device__ __forceinline__ uint32_t wave_scan_any<uint32_t>(uint32_t v) {
    return wave_scan_u32(v);
}

template <int BLOCK, typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T& total, T* sh) {
    constexpr int NW = BLOCK / 64;
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
    T inc = wave_scan_any(v);
    if (lane == 63) sh[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            T t = sh[w];
            sh[w] = acc;
            acc += t;
        }
        sh[NW] = acc;
    }
    __syncthreads();
    T res = sh[wid] + inc - v;
    total = sh[NW];
    __syncthreads();
    return res;
}

template <int BLOCK, typename T>
__device__ __forceinline__ T block_reduce_sum(T v, T* sh) {
    constexpr int NW = BLOCK / 64;
    v = wave_reduce_sum(v);
    if (lane_id() == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    T r = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) r += sh[w];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// bit layout of project tiles (LDS bitmaps over the local z offsets zl of a
// z tile)
// ---------------------------------------------------------------------------
// Plain (m = 0, sh = 0): offset zl is bit zl & 31 of word zl >> 5.
// Blocked by B = 2^sh words (m = B - 1, sh >= 5; tile words a multiple of
// B): every run of B words holds 32 B consecutive offsets, offset j of a run
// being bit j >> sh of the run's word j & m, so B neighbouring z IDs (the
// hubs of a graph numbered by degree) fall in B different words and LDS
// banks.  One formula serves all: word = ((zl >> 5) & ~m) | (zl & m),
// bit = (zl >> sh) & 31, and back zl = tl_base(word) + (bit << sh).  A tile
// base that is a multiple of 32 B (blocked) or 32 (plain) offsets commutes
// with both maps, so the maps can take the z ID itself and subtract the
// base's word afterwards.
__host__ __device__ __forceinline__ uint32_t tl_word(uint32_t zl, uint32_t m) { return ((zl >> 5) & ~m) | (zl & m); }
__host__ __device__ __forceinline__ uint32_t tl_bit(uint32_t zl, uint32_t sh) { return 1u << ((zl >> sh) & 31u); }
__host__ __device__ __forceinline__ uint32_t tl_base(uint32_t word, uint32_t m) {
    return ((word & ~m) << 5) | (word & m);
}

}  // namespace dim3d

// ---------------------------------------------------------------------------
// host-side launch wrappers (defined in k_mapping.hip / k_expand.hip)
// ---------------------------------------------------------------------------
namespace dim3h {

// HIP refuses memory (de)allocation while any stream of the process is
// capturing a graph, so batch captures (exclusive) and every allocation or
// free of the library (shared) exclude each other through one lock.
std::shared_mutex& capture_mutex();
inline hipError_t dev_malloc(void** p, size_t n) {
    std::shared_lock<std::shared_mutex> lk(capture_mutex());
    return hipMalloc(p, n);
}
inline hipError_t dev_free(void* p) {
    std::shared_lock<std::shared_mutex> lk(capture_mutex());
    return hipFree(p);
}
inline hipError_t dev_malloc_async(void** p, size_t n, hipStream_t st) {
    std::shared_lock<std::shared_mutex> lk(capture_mutex());
    return hipMallocAsync(p, n, st);
}
inline hipError_t dev_free_async(void* p, hipStream_t st) {
    std::shared_lock<std::shared_mutex> lk(capture_mutex());
    return hipFreeAsync(p, st);
}
// A copy on stream st, waited for.  Never hipMemcpy: a legacy-stream copy
// waits for every blocking stream, and fails while another thread captures a
// graph on one (the server's CU-masked streams are blocking).
inline hipError_t copy_wait(void* dst, const void* src, size_t n, hipMemcpyKind k, hipStream_t st) {
    hipError_t e = hipMemcpyAsync(dst, src, n, k, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}
inline hipError_t host_malloc(void** p, size_t n) {
    std::shared_lock<std::shared_mutex> lk(capture_mutex());
    return hipHostMalloc(p, n, hipHostMallocDefault);
}
inline hipError_t host_free(void* p) {
    std::shared_lock<std::shared_mutex> lk(capture_mutex());
    return hipHostFree(p);
}


// Device scalar block, zeroed at the start of every query.
constexpr int kHotKeys = dim3d::kTallyKeys, kHotSlots = kHotKeys + 1;
// hot-key lists in Scalars::hot: R.y, S.y, R.x, S.x (S's x is its z)
constexpr int kHotRY = 0, kHotSY = kHotSlots, kHotRX = 2 * kHotSlots, kHotSZ = 3 * kHotSlots;
constexpr uint64_t kHotMinTuples = 1ull << 20;  // smaller relations take no hot-key sample

struct Scalars {
    uint64_t tot_y;      // scan total of the y table: (ny << 32) | 0
    uint64_t tot_x;      // (nx << 32) | nnzR
    uint64_t tot_z;      // (nz << 32) | nnzS
    uint64_t out_j;      // sum degR*degS
    uint64_t tot_s;      // total of the tiled S count scan
    uint64_t out_n;      // rows emitted (may exceed capacity)
    uint64_t fp_sum;
    uint64_t fp_sum2;
    uint64_t fp_csum;
    uint64_t paths;      // sparse join paths expanded
    uint64_t heavy_n;    // heavy keys selected
    uint64_t nxh, nzh;
    uint64_t s_surv;     // S tuples whose y joined
    int32_t mm[8];       // min/max of R.x, R.y, S.z, S.y
    uint64_t agg_sum2;   // COUNT fingerprint from the join paths: sum of k2(x) g(z) over paths
    uint64_t n_hash;     // value aggregates: rows through the LDS hash table
    uint64_t n_wave;     // value aggregates: rows expanded one per wavefront
    // the most frequent values in a sample of R.y, S.y, R.x and S.x (at
    // kHotRY, kHotSY, kHotRX, kHotSZ): a count, then the values
    int32_t hot[4 * kHotSlots];
    // the hot R.x and S.x values' slots in the x and z maps (hot_slots)
    int32_t hot_xs[kHotSlots], hot_zs[kHotSlots];
    uint64_t spare[4];
    // work counters on a cache line of their own: atomics on one line
    // serialise, and every task fetch would queue behind the row reservations
    // on out_n
    alignas(128) uint32_t task;  // expansion work counter
    uint32_t task2;              // second work counter (value expansion: wave rows)
};

size_t scan_partials_needed(uint64_t n);

// Heaviness bins of the dense split: w < 16 exact, then four bins per octave
// (w = min(degR(y), degS(y)) for a key; x and z carry the largest w of their
// tuples).  128 bins cover every u32.
constexpr uint32_t kDenseBins = 128;
__host__ __device__ __forceinline__ uint32_t dense_bin(uint32_t w) {
    if (w < 16u) return w;
    const uint32_t e = 31u - (uint32_t)__builtin_clz(w);
    return 16u + (e - 4u) * 4u + ((w >> (e - 2u)) & 3u);
}
// smallest w of bin b
__host__ __device__ __forceinline__ uint64_t dense_bin_lo(uint32_t b) {
    if (b < 16u) return b;
    const uint32_t e = (b - 16u) / 4u + 4u, m = (b - 16u) % 4u;
    return (uint64_t)(4u + m) << (e - 2u);
}

// Slot space of one attribute: an open-addressing table (hash mode) or the
// value range itself (range mode: slot = (value - lo) * odd mod 2^k, or
// value - lo when mask == 0).
struct KeyMap {
    uint32_t* keys;   // hash mode: keys[mask + 2] (last slot stands for the key -1)
    uint32_t mask;    // hash mode: capacity - 1; range mode: 2^k - 1 (0 = identity slots)
    int32_t lo;       // range mode: smallest value
    uint32_t nslots;  // slots to scan: mask + 2 (hash), 2^k or span (range)
    uint32_t range;   // 1 = range mode
    uint32_t span;    // range mode: hi - lo + 1
};

void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st);
// complete bipartite relation (i / Y, i % Y), i < X * Y
void grid_relation(int2* out, uint32_t X, uint32_t Y, hipStream_t st);
// Y keys of D scattered z values each: tuple i = (z, y), y = i / D,
// z = (a_y (i % D) + b_y) mod 2^zbits with a_y odd (distinct z per key),
// D <= 2^zbits
void scattered_relation(int2* out, uint32_t D, uint32_t Y, uint32_t zbits, hipStream_t st);
void fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t st);
void map_init(const KeyMap& m, hipStream_t st);
// mm[0..3] = min/max of R.x, R.y; mm[4..7] = min/max of S.z, S.y
void minmax(const int2* R, uint64_t nR, const int2* S, uint64_t nS, int32_t* mm, hipStream_t st);

// y mapping: insert R.y, count degR, record slots; probe S.y, count degS and survivors
// hot (nullable): {R's most frequent key, valid, S's, valid} (hot_keys)
// The kHotKeys most frequent values of R.y, S.y, R.x and S.x among 4096
// evenly spaced tuples (those seen at least 8 times), as lists {count,
// values...} at hot + kHotRY, kHotSY, kHotRX, kHotSZ.  Returns false, writing
// nothing, when neither side has kHotMinTuples tuples.
bool hot_keys(const int2* R, uint64_t nR, const int2* S, uint64_t nS, int32_t* hot, hipStream_t st);
void build_y(const int2* R, uint64_t nR, const KeyMap& m, uint32_t* degR, uint32_t* slotR, hipStream_t st,
             const int32_t* hot = nullptr);
void probe_y(const int2* S, uint64_t nS, const KeyMap& m, const uint32_t* degR, uint32_t* degS, int32_t* slotS,
             uint64_t* surv, hipStream_t st, const int32_t* hot = nullptr);
// scan of the y slots: ex[s] = (#joined y before s) << 32; out_j accumulated
void scan_y(const uint32_t* degR, const uint32_t* degS, uint32_t nslots, uint64_t* ex, uint64_t* partials,
            uint64_t* tot, uint64_t* out_j, hipStream_t st);
void outj_only(const uint32_t* degR, const uint32_t* degS, uint32_t nslots, uint64_t* out_j, hipStream_t st);
// *out += sum over y slots of KR(y) GS(y): the path-sum term of the COUNT
// fingerprint (kr, gs: nslots u64 scratch each)
void path_sums(const int2* R, uint64_t nR, const uint32_t* slotR, const int2* S, uint64_t nS, const int32_t* slotS,
               uint32_t nslots, uint64_t* kr, uint64_t* gs, uint64_t* out, hipStream_t st, const int32_t* hot = nullptr);
// x / z mapping over the tuples whose y joined
void build_x(const int2* R, uint64_t nR, const uint32_t* slotR, const uint32_t* degS_y, const KeyMap& m,
             uint32_t* degX, int32_t* slotRx, hipStream_t st, const int32_t* hot = nullptr);
void build_z(const int2* S, uint64_t nS, const int32_t* slotS, const KeyMap& m, uint32_t* degZ, int32_t* slotSz,
             hipStream_t st);
// scan of a degree array: ex[s] = (#occupied before s) << 32 | (sum deg before s)
void scan_deg(const uint32_t* deg, uint32_t nslots, uint64_t* ex, uint64_t* partials, uint64_t* tot,
              hipStream_t st);
// id -> value and CSR row starts from a scanned slot space
void compact_ids(const KeyMap& m, const uint32_t* deg, const uint64_t* ex, int32_t* id2v, uint32_t* rowptr,
                 hipStream_t st);
void finish_rowptr(uint32_t* rowptr, uint32_t n, const uint64_t* tot, hipStream_t st);
// CSR of R by x: col = y id, with bit 31 set for the heavy keys of product
// rows (hid[y] >= 0 and hx[x] >= 0; hid nullable)
// vin / vout (nullable): the tuple values, written beside the columns
void scatter_r(const uint32_t* slotR, const int32_t* slotRx, uint64_t nR, const uint64_t* ex_y,
               const uint64_t* ex_x, const int32_t* hid, const int32_t* hx, uint32_t* cursor, int32_t* col,
               const int32_t* vin, int32_t* vout, hipStream_t st);
// tiled CSR of S by (y, z-tile); ex_z == nullptr means z identity (zid = z - zlo)
// hid (nullable): heavy index per y ID; a heavy key's tuple with a product
// column (zflag) is left out, or, with cntD (dense core: rows outside the
// core still need it), counted in cntD and placed at the front of its row
void count_s(const int2* S, const int32_t* slotS, const int32_t* slotSz, uint64_t nS, const uint64_t* ex_y,
             const uint64_t* ex_z, int32_t zlo, uint32_t ntiles, uint32_t tile_size, const int32_t* hid,
             const uint32_t* zflag, uint32_t* cnt, uint32_t* cntD, uint32_t* keyS, int32_t* zidS, hipStream_t st,
             const int32_t* hot = nullptr);
void split_cursors(const uint32_t* rowptr, const uint32_t* cntD, uint32_t n, uint32_t* curL, uint32_t* curD,
                   hipStream_t st);
void scan_u32(const uint32_t* cnt, uint32_t n, uint32_t* out_excl, uint64_t* partials, uint64_t* tot,
              hipStream_t st);
// TUPLE arrays (dim3_join_aggregate_tuples) into pairs and int32 values
void split_tuples(const uint8_t* raw, uint64_t n, uint32_t stride, uint32_t voff, uint32_t vtype, int2* pairs,
                  int32_t* vals, uint32_t* bad, hipStream_t st);
// the same with 64-bit offsets
void scan_u32_u64(const uint32_t* cnt, uint32_t n, uint64_t* out_excl, uint64_t* partials, uint64_t* tot,
                  hipStream_t st);
// vin / vout (nullable): the tuple values, written beside the columns
void scatter_s(const uint32_t* keyS, const int32_t* zidS, uint64_t nS, uint32_t* cursor, uint32_t* cursorD,
               int32_t* col, const int32_t* vin, int32_t* vout, hipStream_t st);
void x2z_self(const int32_t* id2x, uint32_t nx, const KeyMap& mz, const uint32_t* degZ, const uint64_t* ex_z,
              int32_t* x2z, hipStream_t st);

struct ExpandArgs {
    const uint32_t* rowptrR;
    const int32_t* colR;
    const uint32_t* rowptrS;
    const int32_t* colS;
    const int32_t* id2x;
    const int32_t* id2z;  // nullptr: z identity, value = zlo + z ID
    int32_t zlo;
    const int32_t* x2z;  // nullable: z ID of each x value (exclude_equal)
    uint32_t nx;
    uint32_t ntiles;
    uint32_t tile_size;   // z IDs per tile (bits for project, counters for COUNT)
    uint32_t tile_words;  // u32 words of the tile in LDS
    uint32_t touched_cap;
    uint32_t stage_cap;
    uint32_t fp_mode;
    uint32_t swz;  // project tile bit layout: log2 of the block width in words (tile_words a multiple of it), 0 plain
    // fingerprint of a dense z range (nullable): the factor of z word
    // (zlo >> 5) + j, zword_factor((zlo >> 5) + j), for the words the tiles
    // cover (fp_word_factors)
    const uint32_t* fp_zw;
    // the (x, z-tile) tasks of this launch: x * ntiles + t in [task_lo, task_hi)
    // (all rows, or one chunk of rows of a streamed query)
    uint32_t task_lo, task_hi;
    Scalars* sc;
    void* out;  // int2* (project) or dim3_count_row* (aggregate)
    uint64_t cap;
    // dense seed (hx == nullptr: no dense part)
    const int32_t* hx;        // product row of each x ID, -1 = none
    const uint32_t* splitS;   // dense core: product-column tuples at the front of each S row (skipped by core rows)
    const uint32_t* dbits;    // project: Mh x dwords bit rows over hz
    const int32_t* cint;      // COUNT: Mh x Nh counts
    const uint32_t* hz2zid;   // z ID of product column hz
    const uint32_t* hz_tile;  // first column of each z tile, ntiles + 1 entries
    uint32_t dwords;
    uint32_t Nh;
    // fused bit-or seed (project, every z ID a product column): a product
    // row's tile is the OR of its heavy keys' bitmaps, read here instead of
    // a materialised product row
    const uint32_t* bkeys;    // K x dwords key bitmaps over z IDs (nullptr: seed from dbits / cint)
    const int32_t* hid;       // heavy index of each y ID
    uint32_t tile_major;      // fused seed: take the tasks tile by tile (key bitmaps larger than an L2)
    // value aggregates: the tuple value of each R / S CSR entry
    const int32_t* valR;
    const int32_t* valS;
    // value task lists (hash_x == nullptr: every row tiled): *n_wave rows
    // wave_x[] one per wavefront, *n_hash rows hash_x[] through the
    // 2^hash_bits-slot LDS hash table, the others big_x[] in z tiles
    const uint32_t* rowptrY;  // S rows by y over all tiles (ny + 1 entries)
    const uint32_t* wave_x;
    const uint32_t* hash_x;
    const uint32_t* big_x;
    const uint64_t* n_wave;
    const uint64_t* n_hash;
    uint32_t hash_bits;
    uint32_t row_lo, row_hi;  // the lists cover rows [row_lo, row_hi) (a streamed chunk, or all)
    // materialised wave rows: group count and output offset per wave row,
    // and the scan's partials (nx entries each)
    uint32_t* row_n;
    uint64_t* row_off;
    uint64_t* scan_part;
};

// returns 0 or a hipError_t
// dense part (k_dense.hip)
int gemm_i8(const int8_t* A, const int8_t* Bt, uint32_t M, uint32_t N, uint32_t Kp, uint32_t* dbits,
            uint32_t dwords, int32_t* cint, hipStream_t st);
int bitor_rows(const uint32_t* rowptrR, const int32_t* colR, const int32_t* hid, const uint32_t* hx2x, uint32_t Mh,
               const uint32_t* bbits, uint32_t dwords, uint32_t* dbits, hipStream_t st);
// largest key heaviness over the tuples of each x slot / z slot (slotSz == nullptr: z identity)
void dense_weights(const int2* R, uint64_t nR, const uint32_t* slotR, const int32_t* slotRx, const int2* S,
                   uint64_t nS, const int32_t* slotS, const int32_t* slotSz, int32_t zlo, const uint32_t* degR,
                   const uint32_t* degS, uint32_t* xw, uint32_t nx1, uint32_t* zw, uint32_t nz1, hipStream_t st,
                   const int32_t* hsx = nullptr, const int32_t* hsz = nullptr);
// A hot-value list (hot_keys) as slots of map m (ident: value - zlo when in
// [0, span), the z identity mapping), in the same {count, ...} format; kernels
// that work in slot space tally those slots (hsx, hsz above and below).
void hot_slots(const int32_t* hot, const KeyMap& m, bool ident, int32_t zlo, uint64_t span, int32_t* out,
               hipStream_t st);
// hist[6][kDenseBins]: keys, paths, sum degR, sum degS, x rows, z columns per heaviness bin
void dense_hist(const uint32_t* degR, const uint32_t* degS, uint32_t ny1, const uint32_t* xw, uint32_t nx1,
                const uint32_t* zw, uint32_t nz1, unsigned long long* hist, hipStream_t st);
// heavy index per y ID (+ inverse); counters[0] = keys
void dense_mark_keys(const uint32_t* degR, const uint32_t* degS, uint32_t ny1, const uint64_t* ex_y, uint32_t bmin,
                     int32_t* hid, uint32_t* h2y, uint32_t* counters, hipStream_t st);
// heavy entries per x slot (hdx) and heavy tuples per z slot (hdz), and
// hist[4][kDenseBins] of rows / entries / columns / tuples by heavy-degree bin
void dense_heavy_degrees(const uint32_t* slotR, const int32_t* slotRx, uint64_t nR, const int2* S,
                         const int32_t* slotS, const int32_t* slotSz, int32_t zlo, uint64_t nS, const uint64_t* ex_y,
                         const int32_t* hid, uint32_t* hdx, uint32_t nx1, uint32_t* hdz, uint32_t nz1,
                         unsigned long long* hist, hipStream_t st, const int32_t* hsx = nullptr,
                         const int32_t* hsz = nullptr);
// product row per x ID (+ inverse) and column flag per z ID: by heaviness
// bin (hdx / hdz null) or by heavy degree >= thx / thz (the dense core);
// counters[1] = rows
void dense_mark_rows_cols(const uint32_t* degX, const uint64_t* ex_x, uint32_t nx1, const uint32_t* xw,
                          const uint32_t* degZ, const uint64_t* ex_z, uint32_t nz1, const uint32_t* zw, uint32_t bmin,
                          const uint32_t* hdx, uint32_t thx, const uint32_t* hdz, uint32_t thz, bool all_z,
                          int32_t* hx, uint32_t* hx2x, uint32_t* zflag, uint32_t* counters, hipStream_t st);
void dense_columns(const uint32_t* zflag, const uint32_t* hzpos, uint32_t nz, uint32_t* hz2zid, uint32_t ntiles,
                   uint32_t tile_size, uint32_t Nh, uint32_t* hz_tile, hipStream_t st);
// A rows from the R CSR (heavy entries carry bit 31); Kp bytes per row
void dense_build_a(bool count, const uint32_t* rowptrR, const int32_t* colR, const int32_t* hid,
                   const uint32_t* hx2x, uint32_t Mh, uint32_t Kp, uint32_t* A, uint64_t* overflow, hipStream_t st);
// from the heavy S tuples: kind 0: Bt [Nh][Kp] 0/1 bytes, 1: Bt byte
// multiplicities, 2: bits of per-key z bitmaps [K][dwords] (tile_words > 0:
// in the layout of project tiles of tile_words words, blocked by 2^blk
// words when blk > 0, else plain)
void dense_build_b(int kind, const int2* S, const int32_t* slotS, const int32_t* slotSz, const uint64_t* ex_z,
                   int32_t zlo, uint64_t nS, const uint64_t* ex_y, const int32_t* hid, const uint32_t* zflag,
                   const uint32_t* hzpos, uint32_t Kp, uint32_t dwords, uint32_t* B, uint64_t* overflow,
                   hipStream_t st, uint32_t tile_words = 0, uint32_t blk = 0);

// bound[x] = min(cap, the S paths of row x plus the product columns of a
// product row) >= min(cap, output rows of row x); cap <= nz
void row_bounds(const ExpandArgs& a, uint32_t cap, uint32_t* bound, hipStream_t st);
int expand_project(const ExpandArgs& a, int block, int cu_count, hipStream_t st);
int expand_count(const ExpandArgs& a, int block, int cu_count, hipStream_t st);
// value aggregates (k_value.hip): agg 1 SUM, 2 MIN, 3 MAX, 4 AVG of vR + vS
// (f32: the values are float bit patterns and the rows dim3_value_row_f32)
int expand_value(const ExpandArgs& a, int agg, bool f32, int block, int cu_count, hipStream_t st);
// the short rows (wave_x) of an all-sparse COUNT / join-project query, one
// per wavefront; k_expand then takes the other rows (big_x)
int expand_short_rows(const ExpandArgs& a, bool count, int cu_count, hipStream_t st);
// dynamic LDS of one value-aggregate workgroup (tile_size groups of 12 bytes)
size_t value_lds_bytes(uint32_t tile_size, uint32_t touched_cap, uint32_t block);
// EstimateJoinCardinality (k_estimate.hip): *out += the join paths of the
// tuples whose key has no bit of smask set (keys / cnt: a hash table of
// hmask + 2 slots, keys initialised empty, cnt zero); estimate_prefix:
// *out = the join paths between R[0, a) and S[0, b), a, b <= 1000
void estimate_sample(const int2* R, uint64_t nR, const int2* S, uint64_t nS, uint32_t smask, uint32_t* keys,
                     uint32_t hmask, uint32_t* cnt, unsigned long long* out, hipStream_t st);
void estimate_prefix(const int2* R, uint32_t a, const int2* S, uint32_t b, unsigned long long* out, hipStream_t st);
// rowptrY[y] = rowptrS[y * ntiles] for y <= ny
void rowptr_y(const uint32_t* rowptrS, uint32_t ny, uint32_t ntiles, uint32_t* rowptrY, hipStream_t st);
// the value expansion's task lists (see ExpandArgs): rows with at most 256
// join paths are wave rows, then rows with at most `limit` hash rows;
// bound / fw / pw / fh / ph: nx words each
void value_tasks(const ExpandArgs& a, uint32_t nz, uint32_t limit, uint32_t* bound, uint32_t* fw, uint32_t* pw,
                 uint32_t* fh, uint32_t* ph, uint64_t* partials, uint64_t* n_wave, uint64_t* n_hash,
                 uint32_t* wave_x, uint32_t* hash_x, uint32_t* big_x, hipStream_t st);
// dynamic LDS of one expansion workgroup (fp: plus the fingerprint table)
// out[k] = zword_factor(((zlo >> 5) + k) mod 2^27), k < n (zlo a multiple of 32)
void fp_word_factors(uint32_t* out, uint32_t n, int32_t zlo, hipStream_t st);
size_t expand_lds_bytes(uint32_t tile_words, uint32_t touched_cap, uint32_t stage_cap, uint32_t block, bool fp);
// Join paths grouped by (x, z) into host arrays (dim3_join_paths_grouped);
// host tuple arrays in, temporaries allocated and freed on st.  Returns a
// DIM3_* status, with *err set on failure.
int grouped_paths(const void* R, uint64_t nR, const void* S, uint64_t nS, uint32_t stride, uint32_t voff, uint32_t vsize,
                  bool exclude_equal, hipStream_t st, dim3_grouped* out, std::string* err);
// The classical join-project / COUNT (low OUT_J, f1): sort-based, one call;
// get_out(rows, &buf, &cap) supplies the output once the row count is known
// (not called in fingerprint mode).  Fills sc->out_n, sc->paths and the
// fingerprint sums (sc zeroed by the caller).
int classical_join(const int2* R, uint64_t nR, const int2* S, uint64_t nS, bool count_mode, bool exclude_equal,
                   bool fp, const std::function<int(uint64_t, void**, uint64_t*)>& get_out, Scalars* sc,
                   hipStream_t st, std::string* err);

// COUNT fingerprint with every group's count formed (k_group_count.hip):
// heavy keys hid[y] in [0, 64) as bit masks hm[x] / zm[z] (xflag / zflag:
// mask != 0; the R CSR's heavy entries get bit 31), the box rows / columns
// as lists, their operand bytes A8 [Mp][64] / B8 [Np][64] with the x factors
// xk [Mp][2], the z factors gz [Np] and the reduction operand G (Np / 64
// steps of 64 lanes x 16 bytes), where Mp is Mh rounded up to
// gc_rows_per_block() and Np is gc_cols_padded(Nh) (zero rows and columns
// past the box); then the dense part's terms (gc_dense, box columns per task
// a multiple of 64) and the light part's terms with their corrections
// (gc_light, over the light S CSR by (z tile, key)).  gc_masks sets *dup (zeroed by the
// caller) when a heavy tuple occurs twice, which a mask cannot count: the
// plan is then dropped.  Both add to sc's out_n and fingerprint sums;
// gc_light also counts the light paths in sc->paths and uses sc->task.
uint32_t gc_rows_per_block();
uint32_t gc_cols_padded(uint32_t Nh);
void gc_masks(const uint32_t* rowptrR, int32_t* colR, uint32_t nx, const int2* S, uint64_t nS, const int32_t* slotS,
              const int32_t* slotSz, const uint64_t* ex_y, const uint64_t* ex_z, int32_t zlo, const int32_t* hid,
              uint32_t nz, uint64_t* hm, uint32_t* xflag, uint64_t* zm, uint32_t* zflag, uint64_t* dup,
              hipStream_t st);
void gc_list(const uint32_t* flag, const uint32_t* pos, uint32_t n, uint32_t* list, hipStream_t st);
void gc_build(const uint32_t* rlist, uint32_t Mh, uint32_t Mp, const uint64_t* hm, const int32_t* id2x,
              const uint32_t* clist, uint32_t Nh, const uint64_t* zm, const int32_t* id2z, int32_t zlo, int8_t* A8,
              uint64_t* xk, int8_t* B8, uint64_t* gz, void* G, hipStream_t st);
int gc_dense(const int8_t* A8, const uint64_t* xk, uint32_t Mp, const int8_t* B8, const void* G, uint32_t Np,
             uint32_t cols, Scalars* sc, hipStream_t st);
// zinfo [nz] x 16 bytes: (zm, g) of every z ID, read by gc_light
void gc_zinfo(const uint64_t* zm, uint32_t nz, const int32_t* id2z, int32_t zlo, void* zinfo, hipStream_t st);
// the light S tuples (joined, not heavy) by (z tile, key): keys t * ny + y
// counted into cnt (zeroed by the caller), then, with cursor = the
// exclusive scan of cnt, their z IDs placed into colL
void gc_light_csr_count(const int2* S, uint64_t nS, const int32_t* slotS, const int32_t* slotSz, const uint64_t* ex_y,
                        const uint64_t* ex_z, int32_t zlo, const int32_t* hid, uint32_t ny, uint32_t tile_size,
                        uint32_t* cnt, uint32_t* keyL, uint32_t* zidL, hipStream_t st);
void gc_light_csr_scatter(const uint32_t* keyL, const uint32_t* zidL, uint64_t nS, uint32_t* cursor, int32_t* colL,
                          hipStream_t st);
int gc_light(const uint32_t* rowptrR, const int32_t* colR, const uint32_t* rowL, const int32_t* colL, uint32_t ny,
             const uint64_t* hm, const void* zinfo, const int32_t* id2x, uint32_t nx, uint32_t ntiles,
             uint32_t tile_size, uint32_t nz, int cu_count, Scalars* sc, hipStream_t st);

// Batched expansion: one persistent launch over the tasks of n_inst
// instances (args[i], tasks task_start[i] .. task_start[i + 1]).
// setup: sets the LDS attribute and returns the grid (no stream work, so it
// runs before a graph capture); launch: scalar reset, expansion, gather of
// every instance's scalars into sc_out[n_inst] (kernels only).
int expand_batch_setup(bool count, int block, size_t lds, int cu_count, uint64_t total_tasks, unsigned* grid);
void expand_batch_launch(bool count, bool fp, int block, size_t lds, unsigned grid, const ExpandArgs* args,
                         const uint32_t* task_start, uint32_t n_inst, uint32_t* counter, Scalars* const* sc_list,
                         Scalars* sc_out, hipStream_t st);

}  // namespace dim3h
