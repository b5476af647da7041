// Mapping (subsystem a) and CSR construction on the GPU.
//
// Replaces the reference's L+S mapping pipeline (dim3.hpp:399-455
// mapping_step::mapping, 600-731 partition_*, 733-882 mapping_xz /
// mapping_y, Include/L+S_mapper.hpp LS_mapper) and the CSR builds of
// divide_R / divide_S (dim3.hpp:1058-1086, 1186-1253).
//
// Instead of radix partitions sized for a CPU L2, every attribute gets one
// device-wide slot space (a KeyMap):
//   - hash mode: an open-addressing table (u32 keys, CAS insert, linear
//     probing, load <= 1/2), for sparse value domains;
//   - range mode: slot = value - min, for dense integer domains (the value
//     range is at most twice the tuple count), with no table at all.
// Dense IDs are the rank of the occupied slot, obtained by an exclusive scan
// over the slots, so no ID counter is contended, and the same scan of
// (occupied << 32 | degree) yields the CSR row starts.  Output z values can
// also use the range slot itself as their ID (z identity), so emission needs
// no id -> value lookup.  Degree histograms come from atomics on the slot
// counters.  All kernels here are HBM-bound streaming passes over the tuples
// or the slots.
#include <algorithm>
#include <climits>

#include "dim3_internal.h"

namespace {

using namespace dim3d;

constexpr int kThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;  // 2048 elements per scan block

// the counting passes: grid-stride over at most kTallyBlocks workgroups
inline unsigned tally_grid(uint64_t n) {
    const uint64_t b = (n + kThreads - 1) / kThreads;
    return (unsigned)(b < 1 ? 1 : b < kTallyBlocks ? b : kTallyBlocks);
}

inline unsigned grid_for(uint64_t n, int threads = kThreads) {
    uint64_t g = (n + threads - 1) / threads;
    if (g == 0) g = 1;
    return (unsigned)g;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_ht_init(uint32_t* keys, uint32_t cap) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) keys[i] = DIM3_EMPTY_KEY;
    else if (i == cap) keys[i] = 0;
}


// Range slots are scattered by an odd multiplier (a bijection modulo the
// power-of-two slot count), so that skewed keys with neighbouring values do
// not pile their degree atomics onto the same cache lines.  With
// m.span == 0 (z identity) the slot is the offset itself.
constexpr uint32_t kScatterMul = 0x9E3779B1u;
constexpr uint32_t kScatterInv = 0x0E8B2F51u;
static_assert(kScatterMul * kScatterInv == 1u, "kScatterInv must invert kScatterMul modulo 2^32");

__device__ __forceinline__ uint32_t slot_insert(const dim3h::KeyMap& m, int32_t v) {
    if (m.range) {
        const uint32_t off = (uint32_t)v - (uint32_t)m.lo;
        return m.mask ? (off * kScatterMul) & m.mask : off;
    }
    return ht_insert(m.keys, m.mask, (uint32_t)v);
}

__device__ __forceinline__ int32_t slot_find(const dim3h::KeyMap& m, int32_t v) {
    if (m.range) {
        const uint32_t off = (uint32_t)v - (uint32_t)m.lo;
        if (off >= m.span) return -1;
        return (int32_t)(m.mask ? (off * kScatterMul) & m.mask : off);
    }
    return ht_find(m.keys, m.mask, (uint32_t)v);
}

// The slots of a hot-value list ({count, values...}, hot_keys) in a key map,
// or, ident, value - zlo when it falls in [0, span): the same list format in
// slot space (a value without a slot gets -1, which no valid slot equals).
__global__ void k_hot_slots(const int32_t* __restrict__ hot, dim3h::KeyMap m, bool ident, int32_t zlo, uint64_t span,
                            int32_t* out) {
    const int t = (int)threadIdx.x;
    const int n = hot[0];
    if (t == 0) out[0] = n;
    if (t >= n) return;
    const int32_t v = hot[1 + t];
    int32_t s = -1;
    if (ident) {
        const int64_t d = (int64_t)v - (int64_t)zlo;
        s = d >= 0 && (uint64_t)d < span ? (int32_t)d : -1;
    } else {
        s = slot_find(m, v);
    }
    out[1 + t] = s;
}

__device__ __forceinline__ int32_t key_of(const dim3h::KeyMap& m, uint32_t s) {
    if (m.range) return (int32_t)((uint32_t)m.lo + (m.mask ? (s * kScatterInv) & m.mask : s));
    return ht_key(m.keys, m.mask + 1u, s);
}

__global__ void k_minmax_init(int32_t* mm) {
    if (threadIdx.x < 8) mm[threadIdx.x] = (threadIdx.x & 1) ? INT32_MIN : INT32_MAX;
}

// min / max of R.x, R.y, S.z, S.y (decides hash or range mode per attribute)
__global__ __launch_bounds__(kThreads) void k_minmax(const int2* __restrict__ V, uint64_t n, int32_t* mm) {
    int32_t a0 = INT32_MAX, a1 = INT32_MIN, b0 = INT32_MAX, b1 = INT32_MIN;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const int2 t = V[i];
        a0 = min(a0, t.x);
        a1 = max(a1, t.x);
        b0 = min(b0, t.y);
        b1 = max(b1, t.y);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        a0 = min(a0, __shfl_xor(a0, d, 64));
        a1 = max(a1, __shfl_xor(a1, d, 64));
        b0 = min(b0, __shfl_xor(b0, d, 64));
        b1 = max(b1, __shfl_xor(b1, d, 64));
    }
    __shared__ int32_t sm[4][kThreads / 64];
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        sm[0][w] = a0;
        sm[1][w] = a1;
        sm[2][w] = b0;
        sm[3][w] = b1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kThreads / 64; k++) {
            a0 = min(a0, sm[0][k]);
            a1 = max(a1, sm[1][k]);
            b0 = min(b0, sm[2][k]);
            b1 = max(b1, sm[3][k]);
        }
        atomicMin(&mm[0], a0);
        atomicMax(&mm[1], a1);
        atomicMin(&mm[2], b0);
        atomicMax(&mm[3], b1);
    }
}

// mapping_y insert side (dim3.hpp:822-828): every R.y gets a slot; degR counts it.
// hot (nullable): R.y's hot-key list, tallied per workgroup.
__global__ __launch_bounds__(kThreads) void k_build_y(const int2* __restrict__ R, uint64_t nR, dim3h::KeyMap m,
                                                      uint32_t* degR, uint32_t* __restrict__ slotR,
                                                      const int32_t* __restrict__ hot) {
    __shared__ uint32_t tally[kTallyKeys], tslot[kTallyKeys];
    tally_init<kThreads>(tally, tslot, kTallyKeys);
    for (uint64_t i0 = (uint64_t)blockIdx.x * kThreads; i0 < nR; i0 += (uint64_t)gridDim.x * kThreads) {
        const uint64_t i = i0 + threadIdx.x;
        const bool ok = i < nR;
        const int32_t y = ok ? R[i].y : 0;
        const uint32_t s = ok ? slot_insert(m, y) : 0u;
        const int h = ok ? hot_index(hot, y) : -1;
        if (h >= 0) {
            tslot[h] = s;  // every writer stores the key's one slot
            atomicAdd(&tally[h], 1u);
        }
        wave_count_add(degR, s, ok && h < 0);
        if (ok) slotR[i] = s;
    }
    tally_flush<kThreads>(degR, tally, tslot, kTallyKeys);
}

// The most frequent values of T.y (X: of T.x) among 4096 evenly spaced
// tuples of T (one workgroup): out = {count, values...}, up to kHotKeys
// values seen at least kMinSeen times, most frequent first.
template <bool X>
__global__ __launch_bounds__(1024) void k_hot_key(const int2* __restrict__ T, uint64_t n, int32_t* out) {
    constexpr uint32_t kSample = 4096, kSlots = 8192, kMinSeen = 8;
    __shared__ uint32_t keys[kSlots];
    __shared__ uint32_t cnt[kSlots];
    __shared__ uint64_t cand[kSample / kMinSeen];
    __shared__ uint32_t ncand;
    for (uint32_t k = threadIdx.x; k < kSlots; k += 1024) {
        keys[k] = 0xFFFFFFFFu;
        cnt[k] = 0;
    }
    if (threadIdx.x == 0) ncand = 0;
    __syncthreads();
    const uint64_t m = n < kSample ? n : kSample;
    const uint64_t step = m ? n / m : 0;  // index k * step < m * step <= n
    for (uint32_t k = threadIdx.x; k < m; k += 1024) {
        const int2 tk = T[(uint64_t)k * step];
        const uint32_t y = (uint32_t)(X ? tk.x : tk.y);
        // open addressing; the key 0xFFFFFFFF (-1) is not a candidate
        if (y == 0xFFFFFFFFu) continue;
        uint32_t s = (y * 0x9E3779B1u) >> 19;  // 13 bits: kSlots
        for (;;) {
            const uint32_t old = atomicCAS(&keys[s], 0xFFFFFFFFu, y);
            if (old == 0xFFFFFFFFu || old == y) break;
            s = (s + 1u) & (kSlots - 1u);
        }
        atomicAdd(&cnt[s], 1u);
    }
    __syncthreads();
    // candidates: slots seen at least kMinSeen times (at most
    // kSample / kMinSeen), packed as count << 32 | slot (distinct)
    for (uint32_t k = threadIdx.x; k < kSlots; k += 1024)
        if (cnt[k] >= kMinSeen) cand[atomicAdd(&ncand, 1u)] = ((uint64_t)cnt[k] << 32) | k;
    __syncthreads();
    const uint32_t nc = ncand;
    // a candidate's rank is the number of larger ones: the first kHotKeys go
    // out, most frequent first
    if (threadIdx.x < nc) {
        const uint64_t me = cand[threadIdx.x];
        uint32_t rank = 0;
        for (uint32_t k = 0; k < nc; k++) rank += cand[k] > me ? 1u : 0u;
        if (rank < (uint32_t)dim3h::kHotKeys) out[1 + rank] = (int32_t)keys[(uint32_t)me];
    }
    if (threadIdx.x == 0) out[0] = (int32_t)min(nc, (uint32_t)dim3h::kHotKeys);
}

// mapping_y probe side (dim3.hpp:830-836): S tuples whose y is not in R are
// dropped (slot -1); degS counts the survivors, *surv their number.
__global__ __launch_bounds__(kThreads) void k_probe_y(const int2* __restrict__ S, uint64_t nS, dim3h::KeyMap m,
                                                      const uint32_t* __restrict__ degR, uint32_t* degS,
                                                      int32_t* __restrict__ slotS, uint64_t* surv,
                                                      const int32_t* __restrict__ hot) {
    __shared__ uint64_t sh[kThreads / 64 + 1];
    __shared__ uint32_t tally[kTallyKeys], tslot[kTallyKeys];
    tally_init<kThreads>(tally, tslot, kTallyKeys);
    uint64_t ok = 0;
    for (uint64_t i0 = (uint64_t)blockIdx.x * kThreads; i0 < nS; i0 += (uint64_t)gridDim.x * kThreads) {
        const uint64_t i = i0 + threadIdx.x;
        int32_t s = -1;
        int h = -1;
        if (i < nS) {
            const int32_t y = S[i].y;
            s = slot_find(m, y);
            if (s >= 0 && m.range && degR[s] == 0) s = -1;
            slotS[i] = s;
            ok += s >= 0 ? 1 : 0;
            h = s >= 0 ? hot_index(hot, y) : -1;
        }
        if (h >= 0) {
            tslot[h] = (uint32_t)s;
            atomicAdd(&tally[h], 1u);
        }
        wave_count_add(degS, (uint32_t)s, s >= 0 && h < 0);
    }
    tally_flush<kThreads>(degS, tally, tslot, kTallyKeys);
    ok = block_reduce_sum<kThreads>(ok, sh);
    if (threadIdx.x == 0 && ok) atomicAdd((unsigned long long*)surv, (unsigned long long)ok);
}

// mapping_xz over R.x (dim3.hpp:749-755), restricted to tuples whose y joined.
// hot (nullable): R.x's hot-key list, tallied per workgroup.
__global__ __launch_bounds__(kThreads) void k_build_x(const int2* __restrict__ R, uint64_t nR,
                                                      const uint32_t* __restrict__ slotR,
                                                      const uint32_t* __restrict__ degS_y, dim3h::KeyMap m,
                                                      uint32_t* degX, int32_t* __restrict__ slotRx,
                                                      const int32_t* __restrict__ hot) {
    __shared__ uint32_t tally[kTallyKeys], tslot[kTallyKeys];
    tally_init<kThreads>(tally, tslot, kTallyKeys);
    for (uint64_t i0 = (uint64_t)blockIdx.x * kThreads; i0 < nR; i0 += (uint64_t)gridDim.x * kThreads) {
        const uint64_t i = i0 + threadIdx.x;
        int32_t out = -1;
        int h = -1;
        if (i < nR && degS_y[slotR[i]] > 0) {
            const int32_t x = R[i].x;
            out = (int32_t)slot_insert(m, x);
            h = hot_index(hot, x);
        }
        if (h >= 0) {
            tslot[h] = (uint32_t)out;
            atomicAdd(&tally[h], 1u);
        }
        wave_count_add(degX, (uint32_t)out, out >= 0 && h < 0);
        if (i < nR) slotRx[i] = out;
    }
    tally_flush<kThreads>(degX, tally, tslot, kTallyKeys);
}

// mapping_xz over the surviving S.z.
__global__ void k_build_z(const int2* __restrict__ S, uint64_t nS, const int32_t* __restrict__ slotS,
                          dim3h::KeyMap m, uint32_t* degZ, int32_t* __restrict__ slotSz) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nS) return;
    int32_t out = -1;
    if (slotS[i] >= 0) out = (int32_t)slot_insert(m, S[i].x);
    wave_count_add(degZ, (uint32_t)out, out >= 0);
    slotSz[i] = out;
}

// Degree counting through a workgroup-private LDS histogram, for range-mode
// slot spaces of at most kLdsSlots slots (small key / attribute domains,
// where the per-tuple atomics would otherwise pile onto a few hot lines):
// LDS atomics per tuple, then one global add per nonzero slot and workgroup.
// KIND 0: build_y, 1: probe_y, 2: build_x, 3: build_z (same outputs as the
// per-tuple kernels above).
constexpr uint32_t kLdsSlots = 32768;
constexpr int kHistThreads = 1024;

template <int KIND>
__global__ __launch_bounds__(kHistThreads) void k_count_lds(const int2* __restrict__ V, uint64_t n, dim3h::KeyMap m,
                                                            const uint32_t* __restrict__ aux,
                                                            const uint32_t* __restrict__ slot_r,
                                                            const int32_t* __restrict__ slot_s, uint32_t* deg,
                                                            int32_t* __restrict__ slot_out, uint64_t* surv) {
    extern __shared__ uint32_t hist[];
    __shared__ uint64_t sh[kHistThreads / 64 + 1];
    for (uint32_t s = threadIdx.x; s < m.nslots; s += kHistThreads) hist[s] = 0;
    __syncthreads();
    uint64_t ok = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kHistThreads + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kHistThreads) {
        int32_t s;
        if (KIND == 0) {
            s = (int32_t)slot_insert(m, V[i].y);
        } else if (KIND == 1) {
            s = slot_find(m, V[i].y);
            if (s >= 0 && aux[s] == 0) s = -1;
        } else if (KIND == 2) {
            s = aux[slot_r[i]] > 0 ? (int32_t)slot_insert(m, V[i].x) : -1;
        } else {
            s = slot_s[i] >= 0 ? (int32_t)slot_insert(m, V[i].x) : -1;
        }
        slot_out[i] = s;
        if (s >= 0) {
            atomicAdd(&hist[s], 1u);
            ok++;
        }
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < m.nslots; s += kHistThreads)
        if (hist[s]) atomicAdd(&deg[s], hist[s]);
    if (KIND == 1) {
        ok = block_reduce_sum<kHistThreads>(ok, sh);
        if (threadIdx.x == 0 && ok) atomicAdd((unsigned long long*)surv, (unsigned long long)ok);
    }
}

// true (and launched) when the LDS histogram applies
template <int KIND>
bool count_lds(const int2* V, uint64_t n, const dim3h::KeyMap& m, const uint32_t* aux, const uint32_t* slot_r,
               const int32_t* slot_s, uint32_t* deg, int32_t* slot_out, uint64_t* surv, hipStream_t st) {
    if (!m.range || m.nslots > kLdsSlots || n < 8ull * m.nslots || n == 0) return false;
    const size_t lds = (size_t)m.nslots * 4u;
    if (hipFuncSetAttribute((const void*)k_count_lds<KIND>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    uint64_t g = (n + 32767) / 32768;
    if (g > 256) g = 256;
    hipLaunchKernelGGL(k_count_lds<KIND>, dim3((unsigned)g), dim3(kHistThreads), lds, st, V, n, m, aux, slot_r, slot_s,
                       deg, slot_out, surv);
    return true;
}

// ---------------------------------------------------------------------------
// three-phase exclusive scan (reduce tiles, scan tile sums, rescan tiles)
// ---------------------------------------------------------------------------
struct ValJoinedY {  // slot of a y that occurs in both R and S
    const uint32_t* degS;
    __device__ uint64_t operator()(uint64_t s) const { return degS[s] ? (1ull << 32) : 0ull; }
};
struct ValDeg {  // occupied slot with its degree
    const uint32_t* deg;
    __device__ uint64_t operator()(uint64_t s) const {
        uint32_t d = deg[s];
        return d ? ((1ull << 32) | d) : 0ull;
    }
};
struct ValU32 {
    const uint32_t* a;
    __device__ uint64_t operator()(uint64_t s) const { return a[s]; }
};
struct OutU64 {
    uint64_t* o;
    __device__ void operator()(uint64_t i, uint64_t v) const { o[i] = v; }
};
struct OutU32 {
    uint32_t* o;
    __device__ void operator()(uint64_t i, uint64_t v) const { o[i] = (uint32_t)v; }
};

template <typename F>
__global__ __launch_bounds__(kThreads) void k_scan_reduce(F f, uint64_t n, uint64_t* partials) {
    __shared__ uint64_t sh[kThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) s += f(base + k);
    s = block_reduce_sum<kThreads>(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_scan_partials(uint64_t* partials, uint64_t nb, uint64_t* tot) {
    __shared__ uint64_t sh[1024 / 64 + 1];
    constexpr int IT = 8;
    uint64_t carry = 0;
    for (uint64_t c0 = 0; c0 < nb; c0 += 1024 * IT) {
        const uint64_t base = c0 + (uint64_t)threadIdx.x * IT;
        uint64_t v[IT];
        uint64_t s = 0;
#pragma unroll
        for (int k = 0; k < IT; k++) {
            v[k] = base + k < nb ? partials[base + k] : 0;
            s += v[k];
        }
        uint64_t total;
        uint64_t ex = block_exclusive_scan<1024>(s, total, sh) + carry;
#pragma unroll
        for (int k = 0; k < IT; k++) {
            if (base + k < nb) partials[base + k] = ex;
            ex += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0 && tot) *tot = carry;
}

template <typename F, typename W>
__global__ __launch_bounds__(kThreads) void k_scan_down(F f, W w, uint64_t n, const uint64_t* partials) {
    __shared__ uint64_t sh[kThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < n ? f(base + k) : 0;
        s += v[k];
    }
    uint64_t total;
    uint64_t ex = block_exclusive_scan<kThreads>(s, total, sh) + partials[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) w(base + k, ex);
        ex += v[k];
    }
}

template <typename F, typename W>
void scan_exclusive(F f, W w, uint64_t n, uint64_t* partials, uint64_t* tot, hipStream_t st) {
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0) {
        hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, st, partials, (uint64_t)0, tot);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce<F>, dim3((unsigned)nb), dim3(kThreads), 0, st, f, n, partials);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, st, partials, nb, tot);
    hipLaunchKernelGGL((k_scan_down<F, W>), dim3((unsigned)nb), dim3(kThreads), 0, st, f, w, n,
                       (const uint64_t*)partials);
}

// OUT_J = sum over joined y of degR * degS
__global__ __launch_bounds__(kThreads) void k_outj(const uint32_t* __restrict__ degR,
                                                   const uint32_t* __restrict__ degS, uint32_t cap1,
                                                   uint64_t* out_j) {
    __shared__ uint64_t sh[kThreads / 64 + 1];
    uint64_t acc = 0;
    for (uint64_t s = (uint64_t)blockIdx.x * kThreads + threadIdx.x; s < cap1; s += (uint64_t)gridDim.x * kThreads)
        acc += (uint64_t)degR[s] * (uint64_t)degS[s];
    acc = block_reduce_sum<kThreads>(acc, sh);
    if (threadIdx.x == 0 && acc) atomicAdd((unsigned long long*)out_j, (unsigned long long)acc);
}


__global__ void k_compact_ids(dim3h::KeyMap m, const uint32_t* __restrict__ deg, const uint64_t* __restrict__ ex,
                              int32_t* __restrict__ id2v, uint32_t* __restrict__ rowptr) {
    uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m.nslots) return;
    if (deg[s] == 0) return;
    const uint64_t e = ex[s];
    const uint32_t id = (uint32_t)(e >> 32);
    id2v[id] = key_of(m, (uint32_t)s);
    if (rowptr) rowptr[id] = (uint32_t)e;
}

__global__ void k_finish_rowptr(uint32_t* rowptr, uint32_t n, const uint64_t* tot) {
    if (threadIdx.x == 0 && blockIdx.x == 0) rowptr[n] = (uint32_t)(*tot);
}

// divide_R (dim3.hpp:1058-1086) as an atomic-cursor scatter: rows are
// contiguous, the order inside a row is not kept (a row is a set).
// Entries of heavy keys (hid >= 0, dense part) carry bit 31.
__global__ void k_scatter_r(const uint32_t* __restrict__ slotR, const int32_t* __restrict__ slotRx, uint64_t nR,
                            const uint64_t* __restrict__ ex_y, const uint64_t* __restrict__ ex_x,
                            const int32_t* __restrict__ hid, const int32_t* __restrict__ hx, uint32_t* cursor,
                            int32_t* __restrict__ col, const int32_t* __restrict__ vin, int32_t* __restrict__ vout) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nR) return;
    const int32_t sx = slotRx[i];
    const uint32_t xid = sx >= 0 ? (uint32_t)(ex_x[sx] >> 32) : 0u;
    const uint32_t pos = wave_reserve(cursor, xid, sx >= 0);
    if (sx < 0) return;
    const uint32_t yid = (uint32_t)(ex_y[slotR[i]] >> 32);
    // a dense-core row's heavy entry: its product-column paths come from the product
    const uint32_t flag = (hid != nullptr && hid[yid] >= 0 && hx[xid] >= 0) ? 0x80000000u : 0u;
    col[pos] = (int32_t)(yid | flag);
    if (vout != nullptr) vout[pos] = vin[i];
}

// z ID of a surviving S tuple: the range slot itself (z identity) or the
// rank of its slot.
__device__ __forceinline__ uint32_t z_id(const int2* S, uint64_t i, const int32_t* slotSz, const uint64_t* ex_z,
                                         int32_t zlo) {
    if (ex_z == nullptr) return (uint32_t)S[i].x - (uint32_t)zlo;
    return (uint32_t)(ex_z[slotSz[i]] >> 32);
}

// divide_S (dim3.hpp:1186-1253): CSR of S keyed by (y, z-tile) so that the
// expansion of one z tile reads only its own column range.
// hot (nullable): S.y's hot-key list; with at most kTallyTiles tiles and no
// dense core, each (hot key, tile) counter is tallied per workgroup.
constexpr uint32_t kTallyTiles = 8;
__global__ __launch_bounds__(kThreads) void k_count_s(const int2* __restrict__ S, const int32_t* __restrict__ slotS,
                                                      const int32_t* __restrict__ slotSz, uint64_t nS,
                                                      const uint64_t* __restrict__ ex_y,
                                                      const uint64_t* __restrict__ ex_z, int32_t zlo, uint32_t ntiles,
                                                      uint32_t tile_size, const int32_t* __restrict__ hid,
                                                      const uint32_t* __restrict__ zflag, uint32_t* cnt,
                                                      uint32_t* cntD, uint32_t* __restrict__ keyS,
                                                      int32_t* __restrict__ zidS, const int32_t* __restrict__ hot) {
    constexpr int kT = kTallyKeys * (int)kTallyTiles;
    __shared__ uint32_t tally[kT], tslot[kT];
    tally_init<kThreads>(tally, tslot, kT);
    const int32_t* th = (ntiles <= kTallyTiles && cntD == nullptr) ? hot : nullptr;
    for (uint64_t i0 = (uint64_t)blockIdx.x * kThreads; i0 < nS; i0 += (uint64_t)gridDim.x * kThreads) {
        const uint64_t i = i0 + threadIdx.x;
        bool counted = false, dcol = false;
        uint32_t key = 0;
        int h = -1;
        if (i < nS) {
            const int32_t sy = slotS[i];
            uint32_t kraw = 0xFFFFFFFFu;  // no match
            if (sy >= 0) {
                const uint32_t yid = (uint32_t)(ex_y[sy] >> 32);
                const uint32_t zid = z_id(S, i, slotSz, ex_z, zlo);
                // a heavy key's tuple with a product column: only rows outside
                // the dense core (cntD != nullptr) still need it, at the front
                // of its CSR row
                dcol = hid != nullptr && hid[yid] >= 0 && zflag[zid] != 0u;
                if (!dcol || cntD != nullptr) {
                    const uint32_t tile = ntiles == 1 ? 0u : zid / tile_size;
                    key = ntiles == 1 ? yid : yid * ntiles + tile;
                    counted = true;
                    h = hot_index(th, S[i].y);
                    if (h >= 0) h = h * (int)ntiles + (int)tile;
                    kraw = key | (dcol ? 0x80000000u : 0u);
                    zidS[i] = (int32_t)zid;
                }
            }
            keyS[i] = kraw;
        }
        if (h >= 0) {
            tslot[h] = key;
            atomicAdd(&tally[h], 1u);
        }
        wave_count_add(cnt, key, counted && h < 0);
        if (cntD != nullptr) wave_count_add(cntD, key, counted && dcol);
    }
    tally_flush<kThreads>(cnt, tally, tslot, kT);
}

// cursors of the split S rows: product-column tuples from the row start,
// the others after them
__global__ void k_split_cursors(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ cntD, uint32_t n,
                                uint32_t* __restrict__ curL, uint32_t* __restrict__ curD) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    curD[i] = rowptr[i];
    curL[i] = rowptr[i] + cntD[i];
}

__global__ void k_scatter_s(const uint32_t* __restrict__ keyS, const int32_t* __restrict__ zidS, uint64_t nS,
                            uint32_t* cursor, uint32_t* cursorD, int32_t* __restrict__ col,
                            const int32_t* __restrict__ vin, int32_t* __restrict__ vout) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nS) return;
    const uint32_t raw = keyS[i];
    const bool valid = raw != 0xFFFFFFFFu;
    const bool dcol = valid && (raw >> 31) != 0u;
    const uint32_t key = raw & 0x7FFFFFFFu;
    uint32_t pos = wave_reserve(cursor, key, valid && !dcol);
    if (cursorD != nullptr) {
        const uint32_t pd = wave_reserve(cursorD, key, dcol);
        if (dcol) pos = pd;
    }
    if (!valid) return;
    col[pos] = zidS[i];
    if (vout != nullptr) vout[pos] = vin[i];
}

// The reference's TUPLE arrays (stride bytes per tuple: int32 first, int32
// key, the value at voff) split into pairs and int32 values; a 64-bit value
// outside int32 sets *bad.
__global__ void k_split_tuples(const uint8_t* __restrict__ raw, uint64_t n, uint32_t stride, uint32_t voff,
                               uint32_t vtype, int2* __restrict__ pairs, int32_t* __restrict__ vals, uint32_t* bad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* t = reinterpret_cast<const uint32_t*>(raw + i * stride);  // stride, voff: multiples of 4
    pairs[i] = make_int2((int32_t)t[0], (int32_t)t[1]);
    const uint32_t lo = t[voff / 4u];
    if (vtype == DIM3_VALUE_I64) {
        const uint32_t hi = t[voff / 4u + 1u];
        if (hi != ((int32_t)lo < 0 ? 0xFFFFFFFFu : 0u)) atomicOr(bad, 1u);
    }
    vals[i] = (int32_t)lo;
}

// z ID of each x value, or -1 (self-pair filter of a self-join)
__global__ void k_x2z_self(const int32_t* __restrict__ id2x, uint32_t nx, dim3h::KeyMap mz,
                           const uint32_t* __restrict__ degZ, const uint64_t* __restrict__ ex_z,
                           int32_t* __restrict__ x2z) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx) return;
    const int32_t s = slot_find(mz, id2x[i]);
    int32_t r = -1;
    if (s >= 0) {
        if (ex_z == nullptr) r = s;  // z identity
        else if (degZ[s] > 0) r = (int32_t)(ex_z[s] >> 32);
    }
    x2z[i] = r;
}

// Complete bipartite relation: tuple i = (i / Y, i % Y) for i < X * Y
// (synthetic calibration queries).
__global__ void k_grid_rel(int2* __restrict__ out, uint32_t X, uint32_t Y) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)X * Y) return;
    out[i] = make_int2((int32_t)(i / Y), (int32_t)(i % Y));
}

// COUNT fingerprint terms that are sums over join paths: sum2 = sum over
// paths (x, y, z) of k2(x) g(z) = sum over y of KR(y) GS(y), with
// KR(y) = sum of k2 over R(y) and GS(y) = sum of g over S(y) (mod 2^64;
// tuples counted with multiplicity, as the COUNT is).
__global__ __launch_bounds__(kThreads) void k_path_kr(const int2* __restrict__ R, uint64_t nR,
                                                      const uint32_t* __restrict__ slotR, uint64_t* kr,
                                                      const int32_t* __restrict__ hot) {
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    const bool ok = i < nR;
    const uint32_t s = ok ? slotR[i] : 0u;
    const uint64_t v = ok ? dim3d::xfactor2(R[i].x) : 0ull;
    wave_add_u64(kr, s, v, ok, ok ? hot_index(hot, R[i].y) : -1);
}
__global__ __launch_bounds__(kThreads) void k_path_gs(const int2* __restrict__ S, uint64_t nS,
                                                      const int32_t* __restrict__ slotS, uint64_t* gs,
                                                      const int32_t* __restrict__ hot) {
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    const int32_t s = i < nS ? slotS[i] : -1;
    const uint64_t g = s >= 0 ? dim3d::zfactor(S[i].x) : 0ull;
    wave_add_u64(gs, s >= 0 ? (uint32_t)s : 0u, g, s >= 0, s >= 0 ? hot_index(hot, S[i].y) : -1);
}
__global__ __launch_bounds__(kThreads) void k_path_dot(const uint64_t* __restrict__ kr,
                                                       const uint64_t* __restrict__ gs, uint32_t n, uint64_t* out) {
    __shared__ uint64_t sh[kThreads / 64 + 1];
    uint64_t acc = 0;
    for (uint64_t s = (uint64_t)blockIdx.x * kThreads + threadIdx.x; s < n; s += (uint64_t)gridDim.x * kThreads)
        acc += kr[s] * gs[s];
    acc = block_reduce_sum<kThreads>(acc, sh);
    if (threadIdx.x == 0 && acc) atomicAdd((unsigned long long*)out, (unsigned long long)acc);
}

}  // namespace

namespace dim3h {

__global__ void k_scattered_rel(int2* __restrict__ out, uint32_t D, uint32_t Y, uint32_t zmask) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)D * Y) return;
    const uint32_t y = (uint32_t)(i / D), j = (uint32_t)(i - (uint64_t)y * D);
    const uint64_t h = mix64(0x9E3779B97F4A7C15ull * (y + 1u));
    const uint32_t a = (uint32_t)h | 1u, b = (uint32_t)(h >> 32);
    out[i] = make_int2((int32_t)((a * j + b) & zmask), (int32_t)y);
}

void scattered_relation(int2* out, uint32_t D, uint32_t Y, uint32_t zbits, hipStream_t st) {
    const uint64_t n = (uint64_t)D * Y;
    if (n) hipLaunchKernelGGL(k_scattered_rel, dim3(grid_for(n)), dim3(kThreads), 0, st, out, D, Y, (1u << zbits) - 1u);
}

void grid_relation(int2* out, uint32_t X, uint32_t Y, hipStream_t st) {
    const uint64_t n = (uint64_t)X * Y;
    if (n) hipLaunchKernelGGL(k_grid_rel, dim3(grid_for(n)), dim3(kThreads), 0, st, out, X, Y);
}

size_t scan_partials_needed(uint64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(kThreads), 0, st, p, n, v);
}
void fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(kThreads), 0, st, p, n, v);
}
void map_init(const KeyMap& m, hipStream_t st) {
    if (!m.range) hipLaunchKernelGGL(k_ht_init, dim3(grid_for((uint64_t)m.mask + 2)), dim3(kThreads), 0, st, m.keys,
                                     m.mask + 1u);
}
void minmax(const int2* R, uint64_t nR, const int2* S, uint64_t nS, int32_t* mm, hipStream_t st) {
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, st, mm);
    unsigned gr = std::min<unsigned>(grid_for(nR), 512u), gs = std::min<unsigned>(grid_for(nS), 512u);
    if (nR) hipLaunchKernelGGL(k_minmax, dim3(gr), dim3(kThreads), 0, st, R, nR, mm);
    if (nS) hipLaunchKernelGGL(k_minmax, dim3(gs), dim3(kThreads), 0, st, S, nS, mm + 4);
}
bool hot_keys(const int2* R, uint64_t nR, const int2* S, uint64_t nS, int32_t* hot, hipStream_t st) {
    // below 2^20 tuples a side's hottest keys serialise for well under a
    // millisecond: no sample
    if (nR < kHotMinTuples && nS < kHotMinTuples) return false;
    hipLaunchKernelGGL(k_hot_key<false>, dim3(1), dim3(1024), 0, st, R, nR, hot + kHotRY);
    hipLaunchKernelGGL(k_hot_key<false>, dim3(1), dim3(1024), 0, st, S, nS, hot + kHotSY);
    hipLaunchKernelGGL(k_hot_key<true>, dim3(1), dim3(1024), 0, st, R, nR, hot + kHotRX);
    hipLaunchKernelGGL(k_hot_key<true>, dim3(1), dim3(1024), 0, st, S, nS, hot + kHotSZ);
    return true;
}
void hot_slots(const int32_t* hot, const KeyMap& m, bool ident, int32_t zlo, uint64_t span, int32_t* out,
               hipStream_t st) {
    hipLaunchKernelGGL(k_hot_slots, dim3(1), dim3(64), 0, st, hot, m, ident, zlo, span, out);
}

void build_y(const int2* R, uint64_t nR, const KeyMap& m, uint32_t* degR, uint32_t* slotR, hipStream_t st,
             const int32_t* hot) {
    if (count_lds<0>(R, nR, m, nullptr, nullptr, nullptr, degR, (int32_t*)slotR, nullptr, st)) return;
    if (nR)
        hipLaunchKernelGGL(k_build_y, dim3(tally_grid(nR)), dim3(kThreads), 0, st, R, nR, m, degR, slotR,
                           hot != nullptr ? hot + kHotRY : nullptr);
}
void probe_y(const int2* S, uint64_t nS, const KeyMap& m, const uint32_t* degR, uint32_t* degS, int32_t* slotS,
             uint64_t* surv, hipStream_t st, const int32_t* hot) {
    if (count_lds<1>(S, nS, m, degR, nullptr, nullptr, degS, slotS, surv, st)) return;
    if (nS)
        hipLaunchKernelGGL(k_probe_y, dim3(tally_grid(nS)), dim3(kThreads), 0, st, S, nS, m, degR, degS, slotS,
                           surv, hot != nullptr ? hot + kHotSY : nullptr);
}
void scan_y(const uint32_t* degR, const uint32_t* degS, uint32_t nslots, uint64_t* ex, uint64_t* partials,
            uint64_t* tot, uint64_t* out_j, hipStream_t st) {
    scan_exclusive(ValJoinedY{degS}, OutU64{ex}, nslots, partials, tot, st);
    outj_only(degR, degS, nslots, out_j, st);
}
void outj_only(const uint32_t* degR, const uint32_t* degS, uint32_t nslots, uint64_t* out_j, hipStream_t st) {
    unsigned g = grid_for(nslots);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_outj, dim3(g), dim3(kThreads), 0, st, degR, degS, nslots, out_j);
}
void path_sums(const int2* R, uint64_t nR, const uint32_t* slotR, const int2* S, uint64_t nS, const int32_t* slotS,
               uint32_t nslots, uint64_t* kr, uint64_t* gs, uint64_t* out, hipStream_t st, const int32_t* hot) {
    fill_u64(kr, nslots, 0, st);
    fill_u64(gs, nslots, 0, st);
    if (nR) hipLaunchKernelGGL(k_path_kr, dim3(grid_for(nR)), dim3(kThreads), 0, st, R, nR, slotR, kr, hot);
    if (nS)
        hipLaunchKernelGGL(k_path_gs, dim3(grid_for(nS)), dim3(kThreads), 0, st, S, nS, slotS, gs,
                           hot != nullptr ? hot + kHotSY : nullptr);
    unsigned g = grid_for(nslots);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_path_dot, dim3(g), dim3(kThreads), 0, st, kr, gs, nslots, out);
}
void build_x(const int2* R, uint64_t nR, const uint32_t* slotR, const uint32_t* degS_y, const KeyMap& m,
             uint32_t* degX, int32_t* slotRx, hipStream_t st, const int32_t* hot) {
    if (count_lds<2>(R, nR, m, degS_y, slotR, nullptr, degX, slotRx, nullptr, st)) return;
    if (nR)
        hipLaunchKernelGGL(k_build_x, dim3(tally_grid(nR)), dim3(kThreads), 0, st, R, nR, slotR, degS_y, m, degX,
                           slotRx, hot != nullptr ? hot + kHotRX : nullptr);
}
void build_z(const int2* S, uint64_t nS, const int32_t* slotS, const KeyMap& m, uint32_t* degZ, int32_t* slotSz,
             hipStream_t st) {
    if (count_lds<3>(S, nS, m, nullptr, nullptr, slotS, degZ, slotSz, nullptr, st)) return;
    if (nS) hipLaunchKernelGGL(k_build_z, dim3(grid_for(nS)), dim3(kThreads), 0, st, S, nS, slotS, m, degZ, slotSz);
}
void scan_deg(const uint32_t* deg, uint32_t nslots, uint64_t* ex, uint64_t* partials, uint64_t* tot,
              hipStream_t st) {
    scan_exclusive(ValDeg{deg}, OutU64{ex}, nslots, partials, tot, st);
}
void compact_ids(const KeyMap& m, const uint32_t* deg, const uint64_t* ex, int32_t* id2v, uint32_t* rowptr,
                 hipStream_t st) {
    hipLaunchKernelGGL(k_compact_ids, dim3(grid_for(m.nslots)), dim3(kThreads), 0, st, m, deg, ex, id2v, rowptr);
}
void finish_rowptr(uint32_t* rowptr, uint32_t n, const uint64_t* tot, hipStream_t st) {
    hipLaunchKernelGGL(k_finish_rowptr, dim3(1), dim3(64), 0, st, rowptr, n, tot);
}
void scatter_r(const uint32_t* slotR, const int32_t* slotRx, uint64_t nR, const uint64_t* ex_y, const uint64_t* ex_x,
               const int32_t* hid, const int32_t* hx, uint32_t* cursor, int32_t* col, const int32_t* vin,
               int32_t* vout, hipStream_t st) {
    if (nR)
        hipLaunchKernelGGL(k_scatter_r, dim3(grid_for(nR)), dim3(kThreads), 0, st, slotR, slotRx, nR, ex_y, ex_x,
                           hid, hx, cursor, col, vin, vout);
}
void count_s(const int2* S, const int32_t* slotS, const int32_t* slotSz, uint64_t nS, const uint64_t* ex_y,
             const uint64_t* ex_z, int32_t zlo, uint32_t ntiles, uint32_t tile_size, const int32_t* hid,
             const uint32_t* zflag, uint32_t* cnt, uint32_t* cntD, uint32_t* keyS, int32_t* zidS, hipStream_t st,
             const int32_t* hot) {
    if (nS)
        hipLaunchKernelGGL(k_count_s, dim3(tally_grid(nS)), dim3(kThreads), 0, st, S, slotS, slotSz, nS, ex_y, ex_z,
                           zlo, ntiles, tile_size, hid, zflag, cnt, cntD, keyS, zidS,
                           hot != nullptr ? hot + kHotSY : nullptr);
}
void split_cursors(const uint32_t* rowptr, const uint32_t* cntD, uint32_t n, uint32_t* curL, uint32_t* curD,
                   hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_split_cursors, dim3(grid_for(n)), dim3(kThreads), 0, st, rowptr, cntD, n, curL, curD);
}
void scan_u32(const uint32_t* cnt, uint32_t n, uint32_t* out_excl, uint64_t* partials, uint64_t* tot,
              hipStream_t st) {
    scan_exclusive(ValU32{cnt}, OutU32{out_excl}, n, partials, tot, st);
}
void split_tuples(const uint8_t* raw, uint64_t n, uint32_t stride, uint32_t voff, uint32_t vtype, int2* pairs,
                  int32_t* vals, uint32_t* bad, hipStream_t st) {
    if (n)
        hipLaunchKernelGGL(k_split_tuples, dim3(grid_for(n)), dim3(kThreads), 0, st, raw, n, stride, voff, vtype, pairs,
                           vals, bad);
}
void scan_u32_u64(const uint32_t* cnt, uint32_t n, uint64_t* out_excl, uint64_t* partials, uint64_t* tot,
                  hipStream_t st) {
    scan_exclusive(ValU32{cnt}, OutU64{out_excl}, n, partials, tot, st);
}
void scatter_s(const uint32_t* keyS, const int32_t* zidS, uint64_t nS, uint32_t* cursor, uint32_t* cursorD,
               int32_t* col, const int32_t* vin, int32_t* vout, hipStream_t st) {
    if (nS)
        hipLaunchKernelGGL(k_scatter_s, dim3(grid_for(nS)), dim3(kThreads), 0, st, keyS, zidS, nS, cursor, cursorD,
                           col, vin, vout);
}
void x2z_self(const int32_t* id2x, uint32_t nx, const KeyMap& mz, const uint32_t* degZ, const uint64_t* ex_z,
              int32_t* x2z, hipStream_t st) {
    if (nx)
        hipLaunchKernelGGL(k_x2z_self, dim3(grid_for(nx)), dim3(kThreads), 0, st, id2x, nx, mz, degZ, ex_z, x2z);
}

}  // namespace dim3h
