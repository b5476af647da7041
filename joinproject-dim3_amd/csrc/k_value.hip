// Value aggregates of the join: SUM, MIN, MAX or AVG of vR + vS over the join
// paths of each (x, z) group, with int32 tuple values and int64 results, or
// float tuple values and float results.
//
// Replaces SpmmAggregate (dim3.hpp:1267-1311) run with a value policy.  The
// reference's policy concept is insert(z, v1, v2) on a group's first path,
// update(z, v1, v2) on the others and get(z) at the row's end; its shipped
// policy AggregateFunctionAvg (dim3.hpp:1453-1480) keeps v1 + v2 summed and
// a path count and returns sum / count.  The reference's join-aggregate has
// no dense part (divide_S, dim3.hpp:1410-1450), and neither has this one.
//
// Same task scheme as the COUNT expansion: a persistent workgroup takes one
// (x, z-tile) task at a time; the tile's groups live in LDS as an int64
// accumulator (one 64-bit ds_add / ds_min / ds_max per path) and a u32 path
// counter (one ds_add per path; its first touch lists the group for
// emission).  Accumulators start at the operation's identity, so the order
// of the paths does not matter and the integer results are exact; AVG
// divides with truncation toward zero, as the reference's get() does on
// integral values.
// Float values (F32): the path value is one float add vR + vS, as the
// policy's val1 + val2 on float; SUM and AVG accumulate it in double (LDS
// f64 atomics) and round once at the end, SUM = (float)sum and
// AVG = (float)(sum / paths); the double sums are exact, hence independent
// of the path order, whenever every partial sum is a double (integer-valued
// data below 2^53, e.g. the reference's generated values).  MIN / MAX keep
// an order-preserving integer key of the float (IEEE totalOrder, -0 < +0).
// Per join path: two 4-byte column reads (z and vS) and two LDS atomics.
// Per output row: one 16-byte store (x, z, int64 value), 12 bytes for F32.
#include "dim3_internal.h"

namespace {

using namespace dim3d;
using dim3h::ExpandArgs;

constexpr uint32_t kUnroll = 4;
constexpr uint32_t kPiece = 64 * kUnroll;  // join paths per wavefront work item
constexpr uint32_t kNone = 0xFFFFFFFFu;

template <int AGG>
__device__ __forceinline__ int64_t agg_identity() {
    return AGG == DIM3_AGG_MIN ? INT64_MAX : AGG == DIM3_AGG_MAX ? INT64_MIN : 0;
}

template <int AGG>
__device__ __forceinline__ void agg_apply(int64_t* acc, int64_t v) {
    if (AGG == DIM3_AGG_MIN) atomicMin(reinterpret_cast<long long*>(acc), (long long)v);
    else if (AGG == DIM3_AGG_MAX) atomicMax(reinterpret_cast<long long*>(acc), (long long)v);
    else atomicAdd(reinterpret_cast<unsigned long long*>(acc), (unsigned long long)v);
}

// order-preserving int32 key of a float and back (an involution on the
// negative keys)
__device__ __forceinline__ int32_t float_key(float f) {
    const int32_t b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7FFFFFFF;
}
__device__ __forceinline__ uint32_t key_bits(int32_t k) { return (uint32_t)(k >= 0 ? k : k ^ 0x7FFFFFFF); }

// one join path with tuple values vr, vs (bit patterns for F32)
template <int AGG, bool F32>
__device__ __forceinline__ void agg_path(int64_t* acc, int32_t vr, int32_t vs) {
    if (F32) {
        const float v = __int_as_float(vr) + __int_as_float(vs);
        if (AGG == DIM3_AGG_SUM || AGG == DIM3_AGG_AVG) atomicAdd(reinterpret_cast<double*>(acc), (double)v);
        else agg_apply<AGG>(acc, (int64_t)float_key(v));
    } else {
        agg_apply<AGG>(acc, (int64_t)vr + (int64_t)vs);
    }
}

// the group's result: the int64 value, or the result float's bit pattern
template <int AGG, bool F32>
__device__ __forceinline__ int64_t agg_get(int64_t acc, uint32_t paths) {
    if (F32) {
        if (AGG == DIM3_AGG_SUM) return (int64_t)__float_as_uint((float)__longlong_as_double(acc));
        if (AGG == DIM3_AGG_AVG) return (int64_t)__float_as_uint((float)(__longlong_as_double(acc) / (double)paths));
        return (int64_t)key_bits((int32_t)acc);
    }
    return AGG == DIM3_AGG_AVG ? acc / (int64_t)paths : acc;
}

// LDS: acc[tile_size] (int64) | cnt[tile_size] (u32) | touched[touched_cap] |
// area[4 * BLOCK + 1] (the row's segment table: start, length, piece
// prefix, R value).
struct VLds {
    int64_t* acc;
    uint32_t* cnt;
    uint32_t* touched;
    uint32_t* area;
    uint32_t* sh32;
    uint32_t* s_wtot;
    uint32_t* s_nt;
    uint64_t* s_gbase;
};

template <int BLOCK, int AGG, bool ZID, bool F32>
__device__ __forceinline__ void value_task(const ExpandArgs& a, uint32_t x, uint32_t t, const VLds& L,
                                           uint64_t& fsum, uint64_t& fsum2, uint64_t& fcs, uint64_t& npaths) {
    constexpr uint32_t NW = BLOCK / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t TS = a.tile_size;
    uint32_t* s_sb = L.area;
    uint32_t* s_len = L.area + BLOCK;
    uint32_t* s_pre = L.area + 2 * BLOCK;
    int32_t* s_vr = reinterpret_cast<int32_t*>(L.area + 3 * BLOCK);
    // rows with more paths than this sweep their tile instead of listing first touches
    const uint64_t track_limit = min(4ull * a.touched_cap, 4ull * TS);
    const uint32_t zbase = t * TS;
    const uint32_t rb = a.rowptrR[x], re = a.rowptrR[x + 1];
    const int32_t xv = a.id2x[x];

    // ---- expansion ----
    bool track = true;
    for (uint32_t c0 = rb; c0 < re; c0 += BLOCK) {
        const uint32_t i = c0 + tid;
        const bool in_row = i < re;
        const uint32_t ic = in_row ? i : re - 1u;  // clamped, unconditional loads
        const uint32_t key = ((uint32_t)a.colR[ic] & 0x7FFFFFFFu) * a.ntiles + t;
        const int32_t vr = a.valR[ic];
        const uint32_t sb = a.rowptrS[key], se = a.rowptrS[key + 1];
        uint32_t np = 0;
        if (in_row) {
            s_sb[tid] = sb;
            s_len[tid] = se - sb;
            s_vr[tid] = vr;
            np = (se - sb + kPiece - 1) / kPiece;
            npaths += se - sb;
        }
        uint32_t NP;
        const uint32_t pre = block_exclusive_scan<BLOCK, uint32_t>(np, NP, L.sh32);
        s_pre[tid] = pre;
        if (c0 == rb) track = (uint64_t)NP * kPiece <= track_limit;
        __syncthreads();
        const uint32_t nseg = min((uint32_t)BLOCK, re - c0);
        // pieces come in increasing order: each wavefront gallops forward from
        // the segment of its previous piece (the last segment whose prefix is
        // <= k, which holds piece k)
        uint32_t seg = 0;
        for (uint32_t k = wid; k < NP; k += NW) {
            uint32_t step = 1;
            while (seg + step < nseg && s_pre[seg + step] <= k) {
                seg += step;
                step <<= 1;
            }
            while (step > 1) {
                step >>= 1;
                if (seg + step < nseg && s_pre[seg + step] <= k) seg += step;
            }
            const uint32_t off = (k - s_pre[seg]) * kPiece;
            const uint32_t base = s_sb[seg] + off;  // piece k holds >= 1 path: base is in the S row
            const uint32_t len = min(s_len[seg] - off, kPiece);
            const int32_t vseg = s_vr[seg];  // the R value of the piece's segment
            uint32_t zl[kUnroll];
            int32_t vs[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; u++) {
                const uint32_t idx = lane + 64u * u;
                const bool ok = idx < len;
                const uint32_t j = ok ? base + idx : base;
                zl[u] = ok ? (uint32_t)a.colS[j] - zbase : kNone;
                vs[u] = a.valS[j];
            }
            uint64_t fm[kUnroll];
            uint32_t ftot = 0;
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; u++) {
                uint32_t old = 1u;
                if (zl[u] != kNone) {
                    old = atomicAdd(&L.cnt[zl[u]], 1u);
                    agg_path<AGG, F32>(&L.acc[zl[u]], vseg, vs[u]);
                }
                fm[u] = __ballot(old == 0u);
                ftot += (uint32_t)__popcll(fm[u]);
            }
            if (track && ftot) {
                // first touches go to the list with one reservation per wavefront
                uint32_t fbase = 0;
                if (lane == 0u) fbase = atomicAdd(L.s_nt, ftot);
                fbase = (uint32_t)__builtin_amdgcn_readfirstlane((int)fbase);
#pragma unroll
                for (uint32_t u = 0; u < kUnroll; u++) {
                    if ((fm[u] >> lane) & 1ull) {
                        const uint32_t p = fbase + mbcnt64(fm[u]);
                        if (p < a.touched_cap) L.touched[p] = zl[u];
                    }
                    fbase += (uint32_t)__popcll(fm[u]);
                }
            }
        }
        __syncthreads();
    }

    // u != w filter of a self-join: drop the group z = x
    if (a.x2z != nullptr && tid == 0) {
        const int32_t zs = a.x2z[x];
        const uint32_t zl = (uint32_t)zs - zbase;
        if (zs >= 0 && zl < TS) L.cnt[zl] = 0;  // still listed: the clear below resets its accumulator
    }
    __syncthreads();

    // ---- emission ----
    const uint32_t nt = *L.s_nt;
    const bool listed = track && nt <= a.touched_cap;
    const uint32_t count = listed ? nt : TS;
    const uint32_t wper = (count + NW - 1) / NW;
    const uint32_t wlo = min(wid * wper, count), whi = min(wlo + wper, count);
    uint32_t c = 0;
    for (uint32_t i = wlo + lane; i < whi; i += 64u) c += L.cnt[listed ? L.touched[i] : i] != 0u ? 1u : 0u;
    const uint32_t wtot = wave_sum_u32(c);
    if (lane == 0) L.s_wtot[wid] = wtot;
    __syncthreads();
    uint32_t woff = 0, T = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; w++) {
        const uint32_t v = L.s_wtot[w];
        woff += w < wid ? v : 0u;
        T += v;
    }
    const uint32_t zoff = (uint32_t)a.zlo + zbase;
    const int32_t* id2z = a.id2z + zbase;
    const int64_t ident = agg_identity<AGG>();

    if (a.fp_mode) {
        // fingerprint row (x, z, v): sum += k1(x) g(z), sum2 += k2(x) v g(z),
        // csum += v (v: the int64 value, or the result float's bit pattern)
        uint64_t S1 = 0, S2 = 0;
        for (uint32_t i = wlo + lane; i < whi; i += 64u) {
            const uint32_t w = listed ? L.touched[i] : i;
            const uint32_t n = L.cnt[w];
            if (n) {
                const int64_t v = agg_get<AGG, F32>(L.acc[w], n);
                const uint64_t g = zfactor(ZID ? (int32_t)(zoff + w) : id2z[w]);
                S1 += g;
                S2 += (uint64_t)v * g;
                fcs += (uint64_t)v;
            }
            L.cnt[w] = 0;
            L.acc[w] = ident;
        }
        if (S1 | S2) {
            fsum += xfactor1(xv) * S1;
            fsum2 += xfactor2(xv) * S2;
        }
        if (tid == 0 && T) atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)T);
        __syncthreads();
        return;
    }

    if (tid == 0) *L.s_gbase = T ? atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)T) : 0ull;
    __syncthreads();
    uint64_t pos0 = *L.s_gbase + woff;
    dim3_count_row* out = reinterpret_cast<dim3_count_row*>(a.out);
    dim3_value_row_f32* outf = reinterpret_cast<dim3_value_row_f32*>(a.out);
    for (uint32_t i0 = wlo; i0 < whi; i0 += 64u) {
        const uint32_t i = i0 + lane;
        uint32_t n = 0, w = 0;
        if (i < whi) {
            w = listed ? L.touched[i] : i;
            n = L.cnt[w];
        }
        const uint64_t m = __ballot(n != 0u);
        const uint64_t pos = pos0 + mbcnt64(m);
        pos0 += (uint64_t)__popcll(m);
        if (n != 0u && pos < a.cap) {
            const int32_t zv = ZID ? (int32_t)(zoff + w) : id2z[w];
            const int64_t v = agg_get<AGG, F32>(L.acc[w], n);
            if (F32) {
                dim3_value_row_f32 row;
                row.x = xv;
                row.z = zv;
                row.v = __uint_as_float((uint32_t)v);
                outf[pos] = row;
            } else {
                dim3_count_row row;
                row.x = xv;
                row.z = zv;
                row.count = v;
                out[pos] = row;
            }
        }
        if (i < whi) {
            L.cnt[w] = 0;
            L.acc[w] = ident;
        }
    }
    __syncthreads();
}

template <int BLOCK, int AGG, bool ZID, bool F32>
__global__ __launch_bounds__(BLOCK) void k_value(ExpandArgs a) {
    extern __shared__ __align__(16) uint32_t smem[];
    constexpr uint32_t NW = BLOCK / 64;
    __shared__ uint64_t sh[NW + 1];
    __shared__ uint32_t sh32[NW + 1];
    __shared__ uint32_t s_wtot[NW];
    __shared__ uint32_t s_task, s_nt;
    __shared__ uint64_t s_gbase;
    const uint32_t TS = a.tile_size;
    const VLds L{reinterpret_cast<int64_t*>(smem), smem + 2u * TS, smem + 3u * TS, smem + 3u * TS + a.touched_cap,
                 sh32, s_wtot, &s_nt, &s_gbase};

    const uint32_t tid = threadIdx.x;
    const int64_t ident = agg_identity<AGG>();
    for (uint32_t w = tid; w < TS; w += BLOCK) {
        L.acc[w] = ident;
        L.cnt[w] = 0;
    }
    uint64_t fsum = 0, fsum2 = 0, fcs = 0, npaths = 0;

    for (;;) {
        if (tid == 0) {
            s_task = atomicAdd(&a.sc->task, 1u);
            s_nt = 0;
        }
        __syncthreads();
        const uint32_t task = a.task_lo + __builtin_amdgcn_readfirstlane(s_task);
        if (task >= a.task_hi) break;
        const uint32_t x = task / a.ntiles;
        value_task<BLOCK, AGG, ZID, F32>(a, x, task - x * a.ntiles, L, fsum, fsum2, fcs, npaths);
    }

    const uint64_t np = block_reduce_sum<BLOCK, uint64_t>(npaths, sh);
    if (tid == 0 && np) atomicAdd((unsigned long long*)&a.sc->paths, (unsigned long long)np);
    if (a.fp_mode) {
        const uint64_t s1 = block_reduce_sum<BLOCK, uint64_t>(fsum, sh);
        const uint64_t s2 = block_reduce_sum<BLOCK, uint64_t>(fsum2, sh);
        const uint64_t s3 = block_reduce_sum<BLOCK, uint64_t>(fcs, sh);
        if (tid == 0) {
            atomicAdd((unsigned long long*)&a.sc->fp_sum, (unsigned long long)s1);
            atomicAdd((unsigned long long*)&a.sc->fp_sum2, (unsigned long long)s2);
            atomicAdd((unsigned long long*)&a.sc->fp_csum, (unsigned long long)s3);
        }
    }
}

template <int BLOCK, int AGG, bool ZID, bool F32>
int launch_one(const ExpandArgs& a, size_t lds, int cu_count, hipStream_t st) {
    const auto kern = k_value<BLOCK, AGG, ZID, F32>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, BLOCK, lds);
    if (e != hipSuccess) return (int)e;
    if (occ < 1) occ = 1;
    const uint64_t ntasks = a.task_hi > a.task_lo ? a.task_hi - a.task_lo : 0;
    uint64_t grid = (uint64_t)cu_count * (uint64_t)occ;
    if (grid > ntasks) grid = ntasks;
    if (grid == 0) return 0;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(BLOCK), lds, st, a);
    return (int)hipGetLastError();
}

template <int AGG, bool ZID, bool F32>
int launch_block(const ExpandArgs& a, int block, size_t lds, int cu_count, hipStream_t st) {
    switch (block) {
        case 256: return launch_one<256, AGG, ZID, F32>(a, lds, cu_count, st);
        case 512: return launch_one<512, AGG, ZID, F32>(a, lds, cu_count, st);
        case 1024: return launch_one<1024, AGG, ZID, F32>(a, lds, cu_count, st);
        default: return (int)hipErrorInvalidValue;
    }
}

template <int AGG, bool F32>
int launch_agg(const ExpandArgs& a, int block, size_t lds, int cu_count, hipStream_t st) {
    if (a.id2z == nullptr) return launch_block<AGG, true, F32>(a, block, lds, cu_count, st);
    return launch_block<AGG, false, F32>(a, block, lds, cu_count, st);
}

template <bool F32>
int launch_value(const ExpandArgs& a, int agg, int block, size_t lds, int cu_count, hipStream_t st) {
    switch (agg) {
        case DIM3_AGG_SUM: return launch_agg<DIM3_AGG_SUM, F32>(a, block, lds, cu_count, st);
        case DIM3_AGG_MIN: return launch_agg<DIM3_AGG_MIN, F32>(a, block, lds, cu_count, st);
        case DIM3_AGG_MAX: return launch_agg<DIM3_AGG_MAX, F32>(a, block, lds, cu_count, st);
        case DIM3_AGG_AVG: return launch_agg<DIM3_AGG_AVG, F32>(a, block, lds, cu_count, st);
        default: return (int)hipErrorInvalidValue;
    }
}

}  // namespace

namespace dim3h {

size_t value_lds_bytes(uint32_t tile_size, uint32_t touched_cap, uint32_t block) {
    return (size_t)tile_size * 12u + ((size_t)touched_cap + 4u * block + 1u) * 4u;
}

int expand_value(const ExpandArgs& a, int agg, bool f32, int block, int cu_count, hipStream_t st) {
    if (a.valR == nullptr || a.valS == nullptr || a.tile_size == 0) return (int)hipErrorInvalidValue;
    const size_t lds = value_lds_bytes(a.tile_size, a.touched_cap, (uint32_t)block);
    return f32 ? launch_value<true>(a, agg, block, lds, cu_count, st)
               : launch_value<false>(a, agg, block, lds, cu_count, st);
}

}  // namespace dim3h
