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
// Rows are split three ways by their join-path count (a row holds far
// fewer groups than there are z values in the common case):
//   - up to 256 paths: one row per wavefront, its groups hashed by z ID into
//     a private LDS table (k_value_rows; no workgroup barriers);
//   - up to half a workgroup-sized LDS hash table: one workgroup task per
//     row, groups hashed into that table (k_value);
//   - longer rows: cut into z tiles as in the COUNT expansion, one
//     (x, z-tile) task each, the tile's groups addressed directly (k_value).
// A persistent grid takes its tasks from a device counter; hash rows come
// before tiled ones, so a workgroup switches its LDS from the table to the
// tile layout at most once.  Either way a group is an int64 accumulator
// (one 64-bit ds_add / ds_min / ds_max per path) and a u32 path counter
// (one ds_add per path; its first touch lists the group for emission).
// Accumulators start at the operation's identity, so the order of the
// paths does not matter and the integer results are exact; AVG divides
// with truncation toward zero, as the reference's get() does on integral
// values.
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

// Besides DIM3_AGG_SUM .. DIM3_AGG_AVG, the wave-row kernel serves the short
// rows of all-sparse COUNT and join-project queries: no values, a group's
// result is its path counter (COUNT) or nothing (project: the distinct z).
constexpr int kAggCount = 5;
constexpr int kAggProject = 6;

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
    if (AGG >= kAggCount) return;  // the path counter is the whole state
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
    if (AGG >= kAggCount) return (int64_t)paths;
    if (F32) {
        if (AGG == DIM3_AGG_SUM) return (int64_t)__float_as_uint((float)__longlong_as_double(acc));
        if (AGG == DIM3_AGG_AVG) return (int64_t)__float_as_uint((float)(__longlong_as_double(acc) / (double)paths));
        return (int64_t)key_bits((int32_t)acc);
    }
    return AGG == DIM3_AGG_AVG ? acc / (int64_t)paths : acc;
}

// LDS: groups[3 * tile_size words] | touched[touched_cap] | area[4 * BLOCK + 1]
// (the row's segment table: start, length, piece prefix, R value).  The
// group words hold, in the tile layout, acc[tile_size] (int64) and
// cnt[tile_size] (u32); in the hash layout acc[H] | keys[H] | cnt[H] with
// H = 2^hash_bits <= 3 tile_size / 4 (keys: z ID or kNone).
struct VLds {
    int64_t* acc;
    uint32_t* cnt;
    uint32_t* keys;  // hash layout only
    uint32_t* touched;
    uint32_t* area;
    uint32_t* sh32;
    uint32_t* s_wtot;
    uint32_t* s_nt;
    uint64_t* s_gbase;
};

// Slot of z ID z in the hash table (inserted if absent): multiplicative
// hash, linear probing.  A row holds at most half as many groups as the
// table has slots, so a free slot is always found.  One compare-and-swap
// per probe claims an empty slot or returns the key already there (z itself
// ends the probe too), so the probe loop has no inner branch.
__device__ __forceinline__ uint32_t hash_slot(uint32_t* keys, uint32_t z, uint32_t bits) {
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t s = (z * 0x9E3779B1u) >> (32u - bits);
    for (;;) {
        const uint32_t k = atomicCAS(&keys[s], kNone, z);
        if (k == kNone || k == z) return s;
        s = (s + 1u) & mask;
    }
}

// Slot holding z ID z, or kNone.
__device__ __forceinline__ uint32_t hash_find(const uint32_t* keys, uint32_t z, uint32_t bits) {
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t s = (z * 0x9E3779B1u) >> (32u - bits);
    for (uint32_t i = 0; i <= mask; i++, s = (s + 1u) & mask) {
        const uint32_t k = keys[s];
        if (k == z) return s;
        if (k == kNone) break;
    }
    return kNone;
}

// One task: row x, z tile t (HASH: the whole row through the hash table).
template <int BLOCK, int AGG, bool ZID, bool F32, bool HASH>
__device__ __forceinline__ void value_task(const ExpandArgs& a, uint32_t x, uint32_t t, const VLds& L,
                                           uint64_t& fsum, uint64_t& fsum2, uint64_t& fcs, uint64_t& npaths) {
    constexpr uint32_t NW = BLOCK / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t TS = a.tile_size;
    const uint32_t HB = a.hash_bits;
    uint32_t* s_sb = L.area;
    uint32_t* s_len = L.area + BLOCK;
    uint32_t* s_pre = L.area + 2 * BLOCK;
    int32_t* s_vr = reinterpret_cast<int32_t*>(L.area + 3 * BLOCK);
    // tiles of rows with more paths than this are swept instead of listing
    // first touches (a hash row's groups always fit the list)
    const uint64_t track_limit = min(4ull * a.touched_cap, 4ull * TS);
    const uint32_t zbase = HASH ? 0u : t * TS;
    const uint32_t rb = a.rowptrR[x], re = a.rowptrR[x + 1];
    const int32_t xv = a.id2x[x];

    // ---- expansion ----
    bool track = true;
    for (uint32_t c0 = rb; c0 < re; c0 += BLOCK) {
        const uint32_t i = c0 + tid;
        const bool in_row = i < re;
        const uint32_t ic = in_row ? i : re - 1u;  // clamped, unconditional loads
        // the y's S row in tile t, or over all tiles
        const uint32_t y = (uint32_t)a.colR[ic] & 0x7FFFFFFFu;
        const int32_t vr = a.valR[ic];
        const uint32_t key = y * a.ntiles + t;
        const uint32_t sb = HASH ? a.rowptrY[y] : a.rowptrS[key];
        const uint32_t se = HASH ? a.rowptrY[y + 1] : a.rowptrS[key + 1];
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
        if (c0 == rb) track = HASH || (uint64_t)NP * kPiece <= track_limit;
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
                    if (HASH) zl[u] = hash_slot(L.keys, zl[u], HB);  // the group's slot
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
        if (HASH) {
            const uint32_t sl = zs >= 0 ? hash_find(L.keys, (uint32_t)zs, HB) : kNone;
            if (sl != kNone) L.cnt[sl] = 0;  // still listed: the clear below frees its slot
        } else {
            const uint32_t zl = (uint32_t)zs - zbase;
            if (zs >= 0 && zl < TS) L.cnt[zl] = 0;  // still listed: the clear below resets its accumulator
        }
    }
    __syncthreads();

    // ---- emission ----
    const uint32_t nt = *L.s_nt;
    const bool listed = track && nt <= a.touched_cap;
    const uint32_t count = listed ? nt : HASH ? (1u << HB) : TS;
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
    const uint32_t zlo = (uint32_t)a.zlo;
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
                const uint32_t zid = HASH ? L.keys[w] : zbase + w;
                const uint64_t g = zfactor(ZID ? (int32_t)(zlo + zid) : a.id2z[zid]);
                S1 += g;
                S2 += (uint64_t)v * g;
                fcs += (uint64_t)v;
            }
            L.cnt[w] = 0;
            L.acc[w] = ident;
            if (HASH) L.keys[w] = kNone;
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
            const uint32_t zid = HASH ? L.keys[w] : zbase + w;
            const int32_t zv = ZID ? (int32_t)(zlo + zid) : a.id2z[zid];
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
            if (HASH) L.keys[w] = kNone;
        }
    }
    __syncthreads();
}

// LDS group words in the tile layout (hash = false) or the hash layout.
template <int BLOCK, int AGG>
__device__ __forceinline__ VLds value_layout(const ExpandArgs& a, uint32_t* smem, bool hash, uint32_t* sh32,
                                             uint32_t* s_wtot, uint32_t* s_nt, uint64_t* s_gbase) {
    const uint32_t TS = a.tile_size, H = 1u << a.hash_bits;
    uint32_t* touched = smem + 3u * TS;
    VLds L{reinterpret_cast<int64_t*>(smem), hash ? smem + 3u * H : smem + 2u * TS, hash ? smem + 2u * H : nullptr,
           touched, touched + a.touched_cap, sh32, s_wtot, s_nt, s_gbase};
    const int64_t ident = agg_identity<AGG>();
    const uint32_t n = hash ? H : TS;
    for (uint32_t w = threadIdx.x; w < n; w += BLOCK) {
        L.acc[w] = ident;
        L.cnt[w] = 0;
        if (hash) L.keys[w] = kNone;
    }
    __syncthreads();
    return L;
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
    const uint32_t tid = threadIdx.x;
    // tasks: the hash rows hash_x[0 .. nh), then (big_x[j], t) for every
    // tiled row j and tile t (without task lists: every row is tiled)
    const bool lists = a.hash_x != nullptr;
    const uint32_t nh = lists ? (uint32_t)*a.n_hash : 0u;
    const uint32_t nw = lists ? (uint32_t)*a.n_wave : 0u;  // rows of k_value_rows
    const uint32_t total = lists ? nh + ((a.row_hi - a.row_lo) - nh - nw) * a.ntiles : a.nx * a.ntiles;
    bool hash = nh > 0;
    VLds L = value_layout<BLOCK, AGG>(a, smem, hash, sh32, s_wtot, &s_nt, &s_gbase);
    uint64_t fsum = 0, fsum2 = 0, fcs = 0, npaths = 0;

    for (;;) {
        if (tid == 0) {
            s_task = atomicAdd(&a.sc->task, 1u);
            s_nt = 0;
        }
        __syncthreads();
        const uint32_t task = __builtin_amdgcn_readfirstlane(s_task);
        if (task >= total) break;
        if (task < nh) {
            value_task<BLOCK, AGG, ZID, F32, true>(a, a.hash_x[task], 0, L, fsum, fsum2, fcs, npaths);
            continue;
        }
        if (hash) {  // first tiled task of this workgroup: the group words change layout
            hash = false;
            L = value_layout<BLOCK, AGG>(a, smem, false, sh32, s_wtot, &s_nt, &s_gbase);
        }
        const uint32_t j = task - nh;
        const uint32_t r = j / a.ntiles;
        const uint32_t x = lists ? a.big_x[r] : r;
        value_task<BLOCK, AGG, ZID, F32, false>(a, x, j - r * a.ntiles, L, fsum, fsum2, fcs, npaths);
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

// ---- small rows: one row per wavefront ----
// A row with at most kWaveRow join paths is expanded by one wavefront with
// a private LDS hash table of kWaveSlots slots: no workgroup barriers, and
// sixteen rows in flight per CU hide the row's chain of dependent loads
// (y list -> S bounds -> S columns).  Four wavefronts per workgroup.
constexpr uint32_t kWaveBits = 9;
constexpr uint32_t kWaveSlots = 1u << kWaveBits;
constexpr uint32_t kWaveRow = kWaveSlots / 2;  // paths (hence groups) per wave row: the table stays half full
// rows a wavefront claims per fetch: same-address atomics on the device
// counter serialise (~10 ns each), so one per row would bound a million
// short rows to ~10 ms
constexpr uint32_t kWaveChunk = 16;

// LDS order of a wavefront's own words: the table, the listed slots and
// the current 64 R entries' segment table (start, path prefix, R value).
struct WaveLds {
    int64_t acc[kWaveSlots];
    uint32_t keys[kWaveSlots];
    uint32_t cnt[kWaveSlots];
    uint32_t touched[kWaveRow];
    uint32_t sb[64];
    uint32_t pre[64];
    int32_t vr[64];
};

// orders a wavefront's LDS writes before its other lanes' reads
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// PASS 0: fingerprint; 1: count each row's groups into row_n[task];
// 2: write each row's groups at row_off[task] (the exclusive scan of the
// counts).  Rows are reserved without atomics: same-address atomics on the
// output counter serialise (~10 ns each), a floor of ~10 ms per million
// rows.
template <int AGG, bool ZID, bool F32, int PASS>
__global__ __launch_bounds__(256) void k_value_rows(ExpandArgs a) {
    __shared__ WaveLds wl[4];
    __shared__ uint64_t sh[5];
    const uint32_t lane = threadIdx.x & 63u;
    WaveLds& W = wl[threadIdx.x >> 6];
    const int64_t ident = agg_identity<AGG>();
    for (uint32_t s = lane; s < kWaveSlots; s += 64u) {
        W.acc[s] = ident;
        W.keys[s] = kNone;
        W.cnt[s] = 0;
    }
    wave_sync();
    const uint32_t nw = (uint32_t)*a.n_wave;
    const uint32_t zlo = (uint32_t)a.zlo;
    uint64_t fsum = 0, fsum2 = 0, fcs = 0, npaths = 0, nrows = 0;
    uint32_t task = 0, task_end = 0;
    for (;;) {
        if (task == task_end) {
            uint32_t t0 = 0;
            if (lane == 0) t0 = atomicAdd(&a.sc->task2, kWaveChunk);
            task = (uint32_t)__builtin_amdgcn_readfirstlane((int)t0);
            if (task >= nw) break;
            task_end = min(task + kWaveChunk, nw);
        }
        const uint32_t ti = task++;
        const uint32_t x = a.wave_x[ti];
        const uint32_t rb = a.rowptrR[x], re = a.rowptrR[x + 1];
        const int32_t xv = a.id2x[x];
        uint32_t nt = 0;  // listed slots (wave-uniform)
        for (uint32_t c0 = rb; c0 < re; c0 += 64u) {
            const uint32_t i = c0 + lane;
            const bool in_row = i < re;
            const uint32_t ic = in_row ? i : re - 1u;
            const uint32_t y = (uint32_t)a.colR[ic] & 0x7FFFFFFFu;
            const int32_t vr = AGG >= kAggCount ? 0 : a.valR[ic];
            const uint32_t sb = a.rowptrY[y], se = a.rowptrY[y + 1];  // the y's S row over all tiles
            const uint32_t len = in_row ? se - sb : 0u;
            if (PASS != 1) npaths += len;
            const uint32_t inc = wave_scan_u32(len);
            const uint32_t P = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
            W.sb[lane] = sb;
            W.pre[lane] = inc - len;
            W.vr[lane] = vr;
            wave_sync();
            for (uint32_t p0 = 0; p0 < P; p0 += 64u) {
                const uint32_t p = p0 + lane;
                const bool ok = p < P;
                // the last segment whose path prefix is <= p holds path p
                uint32_t seg = 0;
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1)
                    if (W.pre[seg + step] <= p) seg += step;
                const uint32_t j = ok ? W.sb[seg] + (p - W.pre[seg]) : W.sb[0];
                const uint32_t z = (uint32_t)a.colS[j];
                const int32_t vs = AGG >= kAggCount ? 0 : a.valS[j];
                uint32_t old = 1u, slot = 0;
                if (ok) {
                    slot = hash_slot(W.keys, z, kWaveBits);
                    old = atomicAdd(&W.cnt[slot], 1u);
                    agg_path<AGG, F32>(&W.acc[slot], W.vr[seg], vs);
                }
                const uint64_t m = __ballot(old == 0u);
                if (old == 0u) W.touched[nt + mbcnt64(m)] = slot;
                nt += (uint32_t)__popcll(m);
            }
            wave_sync();  // the segment table is rewritten by the next 64 R entries
        }
        // u != w filter of a self-join: drop the group z = x
        if (a.x2z != nullptr && lane == 0) {
            const int32_t zs = a.x2z[x];
            const uint32_t sl = zs >= 0 ? hash_find(W.keys, (uint32_t)zs, kWaveBits) : kNone;
            if (sl != kNone) W.cnt[sl] = 0;
        }
        wave_sync();
        // emission: the listed slots hold the row's groups
        uint32_t T = 0;
        for (uint32_t i = lane; i < nt; i += 64u) T += W.cnt[W.touched[i]] != 0u ? 1u : 0u;
        T = wave_sum_u32(T);
        if (PASS == 1) {
            if (lane == 0) a.row_n[ti] = T;
            for (uint32_t i = lane; i < nt; i += 64u) {
                const uint32_t w = W.touched[i];
                W.cnt[w] = 0;
                W.acc[w] = ident;
                W.keys[w] = kNone;
            }
        } else if (PASS == 0) {
            uint64_t S1 = 0, S2 = 0;
            for (uint32_t i = lane; i < nt; i += 64u) {
                const uint32_t w = W.touched[i];
                const uint32_t n = W.cnt[w];
                if (n) {
                    const int64_t v = agg_get<AGG, F32>(W.acc[w], n);
                    const uint32_t zid = W.keys[w];
                    const uint64_t g = zfactor(ZID ? (int32_t)(zlo + zid) : a.id2z[zid]);
                    S1 += g;
                    if (AGG != kAggProject) {
                        S2 += (uint64_t)v * g;
                        fcs += (uint64_t)v;
                    }
                }
                W.cnt[w] = 0;
                W.acc[w] = ident;
                W.keys[w] = kNone;
            }
            // a project row (x, z) adds k2(x) g(z) to sum2
            fsum += xfactor1(xv) * S1;
            fsum2 += xfactor2(xv) * (AGG == kAggProject ? S1 : S2);
            if (lane == 0) nrows += T;
        } else {
            uint64_t base = a.row_off[ti];
            dim3_count_row* out = reinterpret_cast<dim3_count_row*>(a.out);
            dim3_value_row_f32* outf = reinterpret_cast<dim3_value_row_f32*>(a.out);
            for (uint32_t i0 = 0; i0 < nt; i0 += 64u) {
                const uint32_t i = i0 + lane;
                uint32_t w = 0, n = 0;
                if (i < nt) {
                    w = W.touched[i];
                    n = W.cnt[w];
                }
                const uint64_t m = __ballot(n != 0u);
                const uint64_t pos = base + mbcnt64(m);
                base += (uint64_t)__popcll(m);
                if (n != 0u && pos < a.cap) {
                    const uint32_t zid = W.keys[w];
                    const int32_t zv = ZID ? (int32_t)(zlo + zid) : a.id2z[zid];
                    const int64_t v = agg_get<AGG, F32>(W.acc[w], n);
                    if (AGG == kAggProject) {
                        reinterpret_cast<int2*>(a.out)[pos] = make_int2(xv, zv);
                    } else if (F32) {
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
                if (i < nt) {
                    W.cnt[w] = 0;
                    W.acc[w] = ident;
                    W.keys[w] = kNone;
                }
            }
        }
        wave_sync();
    }
    if (PASS == 1) return;
    const uint64_t np = block_reduce_sum<256, uint64_t>(npaths, sh);
    if (threadIdx.x == 0 && np) atomicAdd((unsigned long long*)&a.sc->paths, (unsigned long long)np);
    if (PASS == 0) {
        const uint64_t s1 = block_reduce_sum<256, uint64_t>(fsum, sh);
        const uint64_t s2 = block_reduce_sum<256, uint64_t>(fsum2, sh);
        const uint64_t s3 = block_reduce_sum<256, uint64_t>(fcs, sh);
        const uint64_t nr = block_reduce_sum<256, uint64_t>(nrows, sh);
        if (threadIdx.x == 0) {
            atomicAdd((unsigned long long*)&a.sc->fp_sum, (unsigned long long)s1);
            atomicAdd((unsigned long long*)&a.sc->fp_sum2, (unsigned long long)s2);
            atomicAdd((unsigned long long*)&a.sc->fp_csum, (unsigned long long)s3);
            if (nr) atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)nr);
        }
    }
}

// rowptrY[y] = rowptrS[y * ntiles]: the S rows over all tiles, compact
__global__ void k_rowptr_y(const uint32_t* __restrict__ rowptrS, uint32_t ny, uint32_t ntiles,
                           uint32_t* __restrict__ rowptrY) {
    const uint32_t y = blockIdx.x * 256u + threadIdx.x;
    if (y <= ny) rowptrY[y] = rowptrS[(size_t)y * ntiles];
}

// Task lists: rows with at most kWaveRow paths are wave rows, rows with at
// most `limit` hash rows, the others tiled.
__global__ void k_value_split(const uint32_t* __restrict__ bound, uint32_t nx, uint32_t lo, uint32_t hi,
                              uint32_t limit, uint32_t* __restrict__ fw, uint32_t* __restrict__ fh) {
    const uint32_t x = blockIdx.x * 256u + threadIdx.x;
    if (x >= nx) return;
    const bool in = x >= lo && x < hi;
    const uint32_t b = bound[x];
    fw[x] = in && b <= kWaveRow ? 1u : 0u;
    fh[x] = in && b > kWaveRow && b <= limit ? 1u : 0u;
}

// rows outside [lo, hi) are in no list; the tiled rows of the range fill
// big_x in order (x - lo rows of the range precede x)
__global__ void k_value_lists(const uint32_t* __restrict__ fw, const uint32_t* __restrict__ pw,
                              const uint32_t* __restrict__ fh, const uint32_t* __restrict__ ph, uint32_t lo,
                              uint32_t hi, uint32_t* __restrict__ wave_x, uint32_t* __restrict__ hash_x,
                              uint32_t* __restrict__ big_x) {
    const uint32_t x = lo + blockIdx.x * 256u + threadIdx.x;
    if (x >= hi) return;
    if (fw[x]) wave_x[pw[x]] = x;
    else if (fh[x]) hash_x[ph[x]] = x;
    else big_x[(x - lo) - pw[x] - ph[x]] = x;
}

template <int AGG, bool ZID, bool F32, int PASS>
int launch_wave_pass(const ExpandArgs& a, int cu_count, hipStream_t st) {
    const auto kern = k_value_rows<AGG, ZID, F32, PASS>;
    int occ = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 256, 0);
    if (e != hipSuccess) return (int)e;
    if (occ < 1) occ = 1;
    uint64_t grid = (uint64_t)cu_count * (uint64_t)occ;
    const uint64_t waves = ((uint64_t)a.nx + 3u) / 4u;  // at most nx wave rows
    if (grid > waves) grid = waves;
    if (grid == 0) return 0;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

// The wave rows: one fingerprint pass, or a count pass, the scan of the
// row counts into output offsets (its total is the wave rows' output, the
// start of the workgroup tasks' rows) and the write pass.
template <int AGG, bool ZID, bool F32>
int launch_wave_rows(const ExpandArgs& a, int cu_count, hipStream_t st) {
    if (a.fp_mode) return launch_wave_pass<AGG, ZID, F32, 0>(a, cu_count, st);
    dim3h::fill_u32(a.row_n, a.nx, 0, st);
    int e = launch_wave_pass<AGG, ZID, F32, 1>(a, cu_count, st);
    if (e != 0) return e;
    dim3h::scan_u32_u64(a.row_n, a.nx, a.row_off, a.scan_part, &a.sc->out_n, st);
    dim3h::fill_u32(&a.sc->task2, 1, 0, st);
    return launch_wave_pass<AGG, ZID, F32, 2>(a, cu_count, st);
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
    const bool zid = a.id2z == nullptr;
    if (a.wave_x != nullptr) {
        const int e = zid ? launch_wave_rows<AGG, true, F32>(a, cu_count, st)
                          : launch_wave_rows<AGG, false, F32>(a, cu_count, st);
        if (e != 0) return e;
    }
    if (zid) return launch_block<AGG, true, F32>(a, block, lds, cu_count, st);
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

void rowptr_y(const uint32_t* rowptrS, uint32_t ny, uint32_t ntiles, uint32_t* rowptrY, hipStream_t st) {
    hipLaunchKernelGGL(k_rowptr_y, dim3(ny / 256u + 1u), dim3(256), 0, st, rowptrS, ny, ntiles, rowptrY);
}

void value_tasks(const ExpandArgs& a, uint32_t nz, uint32_t limit, uint32_t* bound, uint32_t* fw, uint32_t* pw,
                 uint32_t* fh, uint32_t* ph, uint64_t* partials, uint64_t* n_wave, uint64_t* n_hash,
                 uint32_t* wave_x, uint32_t* hash_x, uint32_t* big_x, hipStream_t st) {
    // the classification needs the bounds up to the larger threshold only
    row_bounds(a, std::min<uint32_t>(nz, std::max<uint32_t>(limit, kWaveRow) + 1u), bound, st);
    const unsigned g = (a.nx + 255u) / 256u;
    if (a.nx)
        hipLaunchKernelGGL(k_value_split, dim3(g), dim3(256), 0, st, bound, a.nx, a.row_lo, a.row_hi, limit, fw, fh);
    scan_u32(fw, a.nx, pw, partials, n_wave, st);
    scan_u32(fh, a.nx, ph, partials, n_hash, st);
    if (a.row_hi > a.row_lo)
        hipLaunchKernelGGL(k_value_lists, dim3((a.row_hi - a.row_lo + 255u) / 256u), dim3(256), 0, st, fw, pw, fh, ph,
                           a.row_lo, a.row_hi, wave_x, hash_x, big_x);
}

size_t value_lds_bytes(uint32_t tile_size, uint32_t touched_cap, uint32_t block) {
    return (size_t)tile_size * 12u + ((size_t)touched_cap + 4u * block + 1u) * 4u;
}

int expand_short_rows(const ExpandArgs& a, bool count, int cu_count, hipStream_t st) {
    const bool zid = a.id2z == nullptr;
    if (count)
        return zid ? launch_wave_rows<kAggCount, true, false>(a, cu_count, st)
                   : launch_wave_rows<kAggCount, false, false>(a, cu_count, st);
    return zid ? launch_wave_rows<kAggProject, true, false>(a, cu_count, st)
               : launch_wave_rows<kAggProject, false, false>(a, cu_count, st);
}

int expand_value(const ExpandArgs& a, int agg, bool f32, int block, int cu_count, hipStream_t st) {
    if (a.valR == nullptr || a.valS == nullptr || a.tile_size == 0) return (int)hipErrorInvalidValue;
    const size_t lds = value_lds_bytes(a.tile_size, a.touched_cap, (uint32_t)block);
    return f32 ? launch_value<true>(a, agg, block, lds, cu_count, st)
               : launch_value<false>(a, agg, block, lds, cu_count, st);
}

}  // namespace dim3h
