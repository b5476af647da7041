// Sparse expansion with per-row dedup in LDS (subsystem d) and its COUNT
// counterpart.
//
// Replaces the reference's SPA loops: Hybrid_solution sparse part
// (dim3.hpp:952-973), its OpenMP copy (dim3_parallel.hpp:89-117) and
// SpmmAggregate (dim3.hpp:1267-1311).
//
// The reference keeps a "last writer" int per z (SPAw) in DRAM-sized arrays
// and pushes each first touch into a growable vector.  Here one workgroup owns
// one (x, z-tile) task at a time, taken from a device work counter:
//   - the tile of the row is a bitmap (project) or u32 counters (COUNT) in
//     LDS; each join path (x, y, z) is one ds_or / ds_add with return,
//   - the first touch of a word (entry) is recorded in an LDS list with one
//     LDS atomic per wavefront (ballot + mbcnt), so emission visits only the
//     touched words of sparse rows and falls back to a full sweep when the
//     list overflows,
//   - the row's rows are counted, one global atomic reserves their output
//     range, and they are staged in LDS and written with contiguous,
//     coalesced stores (or folded into the fingerprint),
//   - the words are cleared on the way out, so the tile is zero for the next
//     task without a separate sweep.
// Work per join path: one 4-byte column read (mostly L2/MALL hits: S rows of
// heavy keys are re-read by every x that reaches them) and one LDS atomic.
#include "dim3_internal.h"

namespace {

using namespace dim3d;
using dim3h::ExpandArgs;

__device__ __forceinline__ uint32_t word_at(const uint32_t* touched, bool listed, uint32_t i) {
    return listed ? touched[i] : i;
}

template <int BLOCK>
__device__ __forceinline__ uint64_t block_reduce_xor64(uint64_t v, uint64_t* sh) {
    constexpr int NW = BLOCK / 64;
    v = wave_reduce_xor(v);
    if (lane_id() == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t r = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) r ^= sh[w];
    __syncthreads();
    return r;
}

// Records the first touches of one wavefront instruction in the LDS list.
__device__ __forceinline__ void push_first_touch(bool first, uint32_t item, uint32_t* s_nt, uint32_t* touched,
                                                 uint32_t cap) {
    const uint64_t ft = __ballot(first);
    if (ft) {
        const uint32_t lane = lane_id();
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)ft) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(s_nt, (uint32_t)__popcll(ft));
        base = __shfl(base, (int)leader, 64);
        if (first) {
            const uint32_t p = base + mbcnt64(ft);
            if (p < cap) touched[p] = item;
        }
    }
}

// ---------------------------------------------------------------------------
// join-project: bitmap tile
// LDS: bits[tile_words] | touched[touched_cap] | (pad) | stage[stage_cap] int2
// ---------------------------------------------------------------------------
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_expand_project(ExpandArgs a) {
    extern __shared__ __align__(16) uint32_t smem[];
    constexpr uint32_t NW = BLOCK / 64;
    uint32_t* bits = smem;
    uint32_t* touched = smem + a.tile_words;
    int2* stage = reinterpret_cast<int2*>(smem + ((a.tile_words + a.touched_cap + 1u) & ~1u));
    __shared__ uint64_t sh[NW + 1];
    __shared__ uint32_t s_task, s_nt;
    __shared__ uint64_t s_gbase;

    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t ntasks = a.nx * a.ntiles;
    const uint32_t tile_size = 1u << a.tile_shift;
    int2* out = reinterpret_cast<int2*>(a.out);

    for (uint32_t w = tid; w < a.tile_words; w += BLOCK) bits[w] = 0;
    uint64_t fsum = 0, fxr = 0, npaths = 0;

    for (;;) {
        if (tid == 0) {
            s_task = atomicAdd(&a.sc->task, 1u);
            s_nt = 0;
        }
        __syncthreads();
        const uint32_t task = s_task;
        if (task >= ntasks) break;
        const uint32_t x = task / a.ntiles;
        const uint32_t t = task - x * a.ntiles;
        const uint32_t zbase = t << a.tile_shift;
        const uint32_t rb = a.rowptrR[x], re = a.rowptrR[x + 1];

        // expansion: wave w takes the y's of the row round robin, lanes walk S[y]
        for (uint32_t i = rb + wid; i < re; i += NW) {
            const uint32_t key = (uint32_t)a.colR[i] * a.ntiles + t;
            const uint32_t sb = a.rowptrS[key], se = a.rowptrS[key + 1];
            npaths += (lane == 0) ? (uint64_t)(se - sb) : 0ull;
            for (uint32_t j = sb + lane; j < se; j += 64) {
                const uint32_t zl = (uint32_t)a.colS[j] - zbase;
                const uint32_t old = atomicOr(&bits[zl >> 5], 1u << (zl & 31u));
                push_first_touch(old == 0u, zl >> 5, &s_nt, touched, a.touched_cap);
            }
        }
        __syncthreads();
        if (a.x2z != nullptr && tid == 0) {  // u != w filter of the self-join
            const int32_t zs = a.x2z[x];
            if (zs >= 0) {
                const uint32_t zl = (uint32_t)zs - zbase;
                if (zl < tile_size && (zl >> 5) < a.tile_words) bits[zl >> 5] &= ~(1u << (zl & 31u));
            }
        }
        __syncthreads();

        // emission: each thread owns a contiguous slice of the word list
        const uint32_t nt = s_nt;
        const bool listed = nt <= a.touched_cap;
        const uint32_t count = listed ? nt : a.tile_words;
        const uint32_t per = (count + BLOCK - 1) / BLOCK;
        const uint32_t lo = min(tid * per, count), hi = min(lo + per, count);
        uint32_t c = 0;
        for (uint32_t i = lo; i < hi; i++) c += __popc(bits[word_at(touched, listed, i)]);
        uint64_t T;
        const uint64_t o = block_exclusive_scan<BLOCK, uint64_t>((uint64_t)c, T, sh);
        const int32_t xv = a.id2x[x];
        const int32_t* id2z = a.id2z + zbase;

        if (a.fp_mode) {
            for (uint32_t i = lo; i < hi; i++) {
                const uint32_t w = word_at(touched, listed, i);
                uint32_t v = bits[w];
                bits[w] = 0;
                while (v) {
                    const uint32_t b = (uint32_t)__ffs(v) - 1u;
                    v &= v - 1u;
                    const uint64_t h = mix64(row_key(xv, id2z[w * 32u + b]));
                    fsum += h;
                    fxr ^= h;
                }
            }
            if (tid == 0 && T) atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)T);
            __syncthreads();
            continue;
        }

        if (tid == 0) s_gbase = T ? atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)T) : 0ull;
        __syncthreads();
        const uint64_t gbase = s_gbase;
        uint32_t ci = lo, cw = 0, cv = 0;
        if (ci < hi) {
            cw = word_at(touched, listed, ci);
            cv = bits[cw];
        }
        uint64_t r = o;
        const uint64_t rend = o + c;
        for (uint64_t q = 0; q < T; q += a.stage_cap) {
            const uint64_t qe = min(q + (uint64_t)a.stage_cap, T);
            while (r < qe && r < rend) {
                while (cv == 0u) {
                    ++ci;
                    cw = word_at(touched, listed, ci);
                    cv = bits[cw];
                }
                const uint32_t b = (uint32_t)__ffs(cv) - 1u;
                cv &= cv - 1u;
                stage[r - q] = make_int2(xv, id2z[cw * 32u + b]);
                ++r;
            }
            __syncthreads();
            const uint32_t nq = (uint32_t)(qe - q);
            for (uint32_t k = tid; k < nq; k += BLOCK) {
                const uint64_t g = gbase + q + k;
                if (g < a.cap) out[g] = stage[k];
            }
            __syncthreads();
        }
        for (uint32_t i = lo; i < hi; i++) bits[word_at(touched, listed, i)] = 0;
        __syncthreads();
    }

    npaths = block_reduce_sum<BLOCK, uint64_t>(npaths, sh);
    if (tid == 0 && npaths) atomicAdd((unsigned long long*)&a.sc->paths, (unsigned long long)npaths);
    if (a.fp_mode) {
        fsum = block_reduce_sum<BLOCK, uint64_t>(fsum, sh);
        fxr = block_reduce_xor64<BLOCK>(fxr, sh);
        if (tid == 0) {
            atomicAdd((unsigned long long*)&a.sc->fp_sum, (unsigned long long)fsum);
            atomicXor((unsigned long long*)&a.sc->fp_xr, (unsigned long long)fxr);
        }
    }
}

// ---------------------------------------------------------------------------
// join-aggregate COUNT: u32 counter tile
// LDS: cnt[tile_words] | touched[touched_cap] | (pad to 16 B) | stage[stage_cap] rows
// ---------------------------------------------------------------------------
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_expand_count(ExpandArgs a) {
    extern __shared__ __align__(16) uint32_t smem[];
    constexpr uint32_t NW = BLOCK / 64;
    uint32_t* cnt = smem;
    uint32_t* touched = smem + a.tile_words;
    dim3_count_row* stage = reinterpret_cast<dim3_count_row*>(smem + ((a.tile_words + a.touched_cap + 3u) & ~3u));
    __shared__ uint64_t sh[NW + 1];
    __shared__ uint32_t s_task, s_nt;
    __shared__ uint64_t s_gbase;

    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t ntasks = a.nx * a.ntiles;
    const uint32_t tile_size = 1u << a.tile_shift;
    dim3_count_row* out = reinterpret_cast<dim3_count_row*>(a.out);

    for (uint32_t w = tid; w < a.tile_words; w += BLOCK) cnt[w] = 0;
    uint64_t fsum = 0, fxr = 0, fcs = 0, npaths = 0;

    for (;;) {
        if (tid == 0) {
            s_task = atomicAdd(&a.sc->task, 1u);
            s_nt = 0;
        }
        __syncthreads();
        const uint32_t task = s_task;
        if (task >= ntasks) break;
        const uint32_t x = task / a.ntiles;
        const uint32_t t = task - x * a.ntiles;
        const uint32_t zbase = t << a.tile_shift;
        const uint32_t rb = a.rowptrR[x], re = a.rowptrR[x + 1];

        for (uint32_t i = rb + wid; i < re; i += NW) {
            const uint32_t key = (uint32_t)a.colR[i] * a.ntiles + t;
            const uint32_t sb = a.rowptrS[key], se = a.rowptrS[key + 1];
            npaths += (lane == 0) ? (uint64_t)(se - sb) : 0ull;
            for (uint32_t j = sb + lane; j < se; j += 64) {
                const uint32_t zl = (uint32_t)a.colS[j] - zbase;
                const uint32_t old = atomicAdd(&cnt[zl], 1u);
                push_first_touch(old == 0u, zl, &s_nt, touched, a.touched_cap);
            }
        }
        __syncthreads();
        if (a.x2z != nullptr && tid == 0) {
            const int32_t zs = a.x2z[x];
            if (zs >= 0) {
                const uint32_t zl = (uint32_t)zs - zbase;
                if (zl < tile_size && zl < a.tile_words) cnt[zl] = 0;
            }
        }
        __syncthreads();

        const uint32_t nt = s_nt;
        const bool listed = nt <= a.touched_cap;
        const uint32_t count = listed ? nt : a.tile_words;
        const uint32_t per = (count + BLOCK - 1) / BLOCK;
        const uint32_t lo = min(tid * per, count), hi = min(lo + per, count);
        uint32_t c = 0;
        for (uint32_t i = lo; i < hi; i++) c += cnt[word_at(touched, listed, i)] != 0u;
        uint64_t T;
        const uint64_t o = block_exclusive_scan<BLOCK, uint64_t>((uint64_t)c, T, sh);
        const int32_t xv = a.id2x[x];
        const int32_t* id2z = a.id2z + zbase;

        if (a.fp_mode) {
            for (uint32_t i = lo; i < hi; i++) {
                const uint32_t e = word_at(touched, listed, i);
                const uint32_t v = cnt[e];
                cnt[e] = 0;
                if (v) {
                    const uint64_t h = mix64(mix64(row_key(xv, id2z[e])) + (uint64_t)v);
                    fsum += h;
                    fxr ^= h;
                    fcs += v;
                }
            }
            if (tid == 0 && T) atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)T);
            __syncthreads();
            continue;
        }

        if (tid == 0) s_gbase = T ? atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)T) : 0ull;
        __syncthreads();
        const uint64_t gbase = s_gbase;
        uint32_t ci = lo;
        uint64_t r = o;
        const uint64_t rend = o + c;
        for (uint64_t q = 0; q < T; q += a.stage_cap) {
            const uint64_t qe = min(q + (uint64_t)a.stage_cap, T);
            while (r < qe && r < rend) {
                uint32_t e = word_at(touched, listed, ci);
                uint32_t v = cnt[e];
                while (v == 0u) {
                    ++ci;
                    e = word_at(touched, listed, ci);
                    v = cnt[e];
                }
                ++ci;
                dim3_count_row row;
                row.x = xv;
                row.z = id2z[e];
                row.count = (int64_t)v;
                stage[r - q] = row;
                ++r;
            }
            __syncthreads();
            const uint32_t nq = (uint32_t)(qe - q);
            for (uint32_t k = tid; k < nq; k += BLOCK) {
                const uint64_t g = gbase + q + k;
                if (g < a.cap) out[g] = stage[k];
            }
            __syncthreads();
        }
        for (uint32_t i = lo; i < hi; i++) cnt[word_at(touched, listed, i)] = 0;
        __syncthreads();
    }

    npaths = block_reduce_sum<BLOCK, uint64_t>(npaths, sh);
    if (tid == 0 && npaths) atomicAdd((unsigned long long*)&a.sc->paths, (unsigned long long)npaths);
    if (a.fp_mode) {
        fsum = block_reduce_sum<BLOCK, uint64_t>(fsum, sh);
        fcs = block_reduce_sum<BLOCK, uint64_t>(fcs, sh);
        fxr = block_reduce_xor64<BLOCK>(fxr, sh);
        if (tid == 0) {
            atomicAdd((unsigned long long*)&a.sc->fp_sum, (unsigned long long)fsum);
            atomicAdd((unsigned long long*)&a.sc->fp_csum, (unsigned long long)fcs);
            atomicXor((unsigned long long*)&a.sc->fp_xr, (unsigned long long)fxr);
        }
    }
}

template <typename K>
int launch_persistent(K kern, int block, size_t lds, const ExpandArgs& a, int cu_count, hipStream_t st) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, lds);
    if (e != hipSuccess) return (int)e;
    if (occ < 1) occ = 1;
    const uint64_t ntasks = (uint64_t)a.nx * a.ntiles;
    uint64_t grid = (uint64_t)cu_count * (uint64_t)occ;
    if (grid > ntasks) grid = ntasks;
    if (grid == 0) return 0;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), lds, st, a);
    return (int)hipGetLastError();
}

}  // namespace

namespace dim3h {

size_t expand_project_lds(uint32_t tile_words, uint32_t touched_cap, uint32_t stage_cap) {
    return (size_t)((tile_words + touched_cap + 1u) & ~1u) * 4u + (size_t)stage_cap * 8u;
}

size_t expand_count_lds(uint32_t tile_entries, uint32_t touched_cap, uint32_t stage_cap) {
    return (size_t)((tile_entries + touched_cap + 3u) & ~3u) * 4u + (size_t)stage_cap * 16u;
}

int expand_project(const ExpandArgs& a, int block, int cu_count, hipStream_t st) {
    const size_t lds = expand_project_lds(a.tile_words, a.touched_cap, a.stage_cap);
    switch (block) {
        case 256: return launch_persistent(k_expand_project<256>, 256, lds, a, cu_count, st);
        case 512: return launch_persistent(k_expand_project<512>, 512, lds, a, cu_count, st);
        case 1024: return launch_persistent(k_expand_project<1024>, 1024, lds, a, cu_count, st);
        default: return (int)hipErrorInvalidValue;
    }
}

int expand_count(const ExpandArgs& a, int block, int cu_count, hipStream_t st) {
    const size_t lds = expand_count_lds(a.tile_words, a.touched_cap, a.stage_cap);
    switch (block) {
        case 256: return launch_persistent(k_expand_count<256>, 256, lds, a, cu_count, st);
        case 512: return launch_persistent(k_expand_count<512>, 512, lds, a, cu_count, st);
        case 1024: return launch_persistent(k_expand_count<1024>, 1024, lds, a, cu_count, st);
        default: return (int)hipErrorInvalidValue;
    }
}

}  // namespace dim3h
