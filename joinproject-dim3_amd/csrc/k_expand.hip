// Sparse expansion with per-row dedup in LDS (subsystem d) and its COUNT
// counterpart.
//
// Replaces the reference's SPA loops: Hybrid_solution sparse part
// (dim3.hpp:952-973), its OpenMP copy (dim3_parallel.hpp:89-117) and
// SpmmAggregate (dim3.hpp:1267-1311).
//
// The reference keeps a "last writer" int per z (SPAw) and pushes every first
// touch into a growable vector.  Here a persistent workgroup takes one
// (x, z-tile) task at a time from a device work counter:
//   - the z tile of the row lives in LDS as a bitmap (project) or as u32
//     counters (COUNT); one join path (x, y, z) is one ds_or / ds_add with
//     return;
//   - the row's S segments are cut into pieces of 256 paths, the pieces are
//     dealt to the wavefronts round robin (so a hub segment is spread over
//     the whole workgroup) and each lane keeps four column loads in flight;
//   - for rows with few paths the first touch of a word / counter is recorded
//     in an LDS list with one LDS atomic per wavefront (ballot + mbcnt), so
//     sparse rows emit from the touched words only; rows with many paths (or
//     an overflowing list) sweep their tile instead;
//   - emission counts the row, reserves its output range with one global
//     atomic, stages the row's local z offsets in LDS and writes (x, z) rows
//     with coalesced stores; z values are zlo + ID for dense domains (no
//     lookup), else one id -> value read; the tile is cleared on the way out.
// Per join path: one 4-byte column read (mostly L2/MALL hits, as heavy S rows
// are re-read by every x that reaches them) and one LDS atomic.  Per output
// row: one 8-byte (16-byte for COUNT) store.
#include "dim3_internal.h"

namespace {

using namespace dim3d;
using dim3h::ExpandArgs;

constexpr uint32_t kUnroll = 4;
constexpr uint32_t kPiece = 64 * kUnroll;  // join paths per wavefront work item
constexpr uint32_t kPairs = 2;             // pieces in flight per wavefront
constexpr uint32_t kSweepWords = 2;        // tile words per lane and step of a swept row's emission
// word 3 of a raw (untyped, unswizzled) buffer descriptor on CDNA: 32-bit
// data format, so loads past num_records return 0
constexpr int kRawBufferWord3 = 0x00020000;

__device__ __forceinline__ uint32_t list_at(const uint32_t* touched, bool listed, uint32_t i) {
    return listed ? touched[i] : i;
}

// Exclusive scan over the block with two barriers; sh needs NW + 1 words and
// must not be reused before the next barrier.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_scan2_u32(uint32_t v, uint32_t& total, uint32_t* sh) {
    constexpr uint32_t NW = BLOCK / 64;
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t inc = wave_scan_u32(v);
    if (lane_id() == 63) sh[wid] = inc;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; w++) {
        const uint32_t x = sh[w];
        before += w < wid ? x : 0u;
        tot += x;
    }
    total = tot;
    return before + inc - v;
}

// Sum over the block (two barriers, sh reusable afterwards).
template <int BLOCK>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* sh) {
    constexpr uint32_t NW = BLOCK / 64;
    v = wave_reduce_sum(v);
    if (lane_id() == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t r = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; w++) r += sh[w];
    __syncthreads();
    return r;
}

// One output row (plain store: non-temporal stores measured slower here).
__device__ __forceinline__ void put_pair(int2* p, int32_t x, int32_t z) { *p = make_int2(x, z); }

// Project tiles use the blocked bit layout (tl_word / tl_bit / tl_base) when
// ExpandArgs::swz is set, else the plain one; COUNT tiles are one counter per
// local z offset.
template <bool COUNT>
__device__ __forceinline__ uint32_t list_item(uint32_t zl, uint32_t m) {
    return COUNT ? zl : tl_word(zl, m);
}

template <bool COUNT>
__device__ __forceinline__ uint32_t units_of(uint32_t v) {
    return COUNT ? (v != 0u) : (uint32_t)__popc(v);
}

// LDS layout (u32 words):
//   [fingerprint table, 2 * kFpTableU64 words, fingerprint mode only] |
//   tile[tile_words] | touched[touched_cap] | area[max(stage_cap, 3*BLOCK+1)]
// `area` holds the segment table of the row during expansion (and the heavy
// key list of a fused bit-or seed) and the staged local z offsets during
// emission.
// ZID: z IDs are value - zlo (dense domain), so no id -> value lookup.
// One (x, z-tile) task: dense seed, expansion of the light paths into the
// LDS tile, then emission (rows or fingerprint) and the clear of the tile.
struct Lds {
    const uint64_t* fpt;  // fingerprint table (fp mode): kFpTableU64 entries
    uint32_t* tile;
    uint32_t* touched;
    uint32_t* area;
    uint64_t* sh;
    uint32_t* sh32;
    uint32_t* s_wtot;
    uint32_t* s_nt;
    uint64_t* s_gbase;
    uint32_t* s_cursor;  // rows of the task handed out so far (single-pass emission)
};

// Emission of a swept join-project task (rows too dense for the touched
// list) in one pass over the tile after a popcount of it: the row count is
// reserved with one global atomic; then every wavefront walks its share of
// the tile 64 words at a time, clears them, places the chunk's rows by a
// wavefront prefix of the word popcounts into its own LDS stage region (the
// touched list and the area, free here) and flushes full regions with
// coalesced 8-byte stores at offsets handed out by an LDS cursor.  The stage
// holds z values (dense z range) or local z offsets.  Returns false,
// reserving nothing, for a row denser than a quarter of its tile: its rows go
// through the block-staged emission instead (balanced per-thread shares
// written block-wide, measured faster there: 1M x 1M Zipf join-project, 69%
// dense rows).
template <int BLOCK, bool ZID>
__device__ __forceinline__ bool emit_swept(const ExpandArgs& a, const Lds& L, int32_t xv, uint32_t zbase,
                                           uint32_t self_item, uint32_t self_keep) {
    constexpr uint32_t NW = BLOCK / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    uint32_t* tile = L.tile;
    const uint32_t W = a.tile_words;
    const uint32_t lsh = a.swz, lm = lsh ? (1u << lsh) - 1u : 0u;
    {
        // the row count: popcount of the tile, 16 bytes per lane and step
        // (the tile starts 16-byte aligned)
        uint32_t c = 0;
        const uint32_t W4 = W >> 2;
        const uint4* t4 = reinterpret_cast<const uint4*>(tile);
        for (uint32_t i = tid; i < W4; i += BLOCK) {
            const uint4 q = t4[i];
            c += __popc(q.x) + __popc(q.y) + __popc(q.z) + __popc(q.w);
        }
        for (uint32_t i = 4u * W4 + tid; i < W; i += BLOCK) c += __popc(tile[i]);
        const uint32_t wsum = wave_sum_u32(c);
        if (lane == 0) L.s_wtot[wid] = wsum;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t T = 0;
#pragma unroll
        for (uint32_t w = 0; w < NW; w++) T += L.s_wtot[w];
        if (self_item != 0xFFFFFFFFu) {  // the u != w filter: drop the row's own z
            const uint32_t v = tile[self_item];
            if (v & ~self_keep) {
                tile[self_item] = v & self_keep;
                T -= 1u;
            }
        }
        const bool dense = (uint64_t)T * 4u >= (uint64_t)a.tile_size;
        *L.s_gbase = dense ? ~0ull : T ? atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)T) : 0ull;
        *L.s_cursor = 0u;
    }
    __syncthreads();
    const uint64_t gbase = *L.s_gbase;
    if (gbase == ~0ull) return false;
    const uint32_t region = a.touched_cap + max(a.stage_cap, 3u * (uint32_t)BLOCK + 1u);
    const uint32_t SW = region / NW;
    uint32_t* stg = L.touched + wid * SW;
    const uint32_t zoff = ZID ? (uint32_t)a.zlo + zbase : 0u;  // stage entry = zoff + local offset
    const int32_t* id2z = a.id2z + zbase;
    int2* out = reinterpret_cast<int2*>(a.out);
    uint32_t fill = 0;  // wavefront-uniform
    auto flush = [&](uint32_t n) {
        uint32_t off = 0;
        if (lane == 0) off = atomicAdd(L.s_cursor, n);
        off = (uint32_t)__builtin_amdgcn_readfirstlane((int)off);
        __builtin_amdgcn_wave_barrier();
        const uint64_t g0 = gbase + off;
        if (g0 + n <= a.cap) {
            // four rows per lane per round: the four LDS reads are in flight
            // together, then four coalesced stores
            int2* o = out + g0;
            uint32_t k = lane;
            for (; k + 192u < n; k += 256u) {
                const uint32_t e0 = stg[k], e1 = stg[k + 64u], e2 = stg[k + 128u], e3 = stg[k + 192u];
                put_pair(o + k, xv, ZID ? (int32_t)e0 : id2z[e0]);
                put_pair(o + k + 64u, xv, ZID ? (int32_t)e1 : id2z[e1]);
                put_pair(o + k + 128u, xv, ZID ? (int32_t)e2 : id2z[e2]);
                put_pair(o + k + 192u, xv, ZID ? (int32_t)e3 : id2z[e3]);
            }
            for (; k < n; k += 64u) {
                const uint32_t e = stg[k];
                put_pair(o + k, xv, ZID ? (int32_t)e : id2z[e]);
            }
        } else {
            for (uint32_t k = lane; k < n; k += 64u) {
                const uint32_t e = stg[k];
                if (g0 + k < a.cap) put_pair(out + g0 + k, xv, ZID ? (int32_t)e : id2z[e]);
            }
        }
        __builtin_amdgcn_wave_barrier();
    };
    // Chunks of 64 kSweepWords words dealt to the wavefronts round robin (a
    // dense z range, the hubs, spreads over all of them instead of loading
    // one); lane l takes the chunk's words kSweepWords l and on, with one
    // LDS read and clear, one popcount scan for all of them, and a bit loop
    // per word.  The words of a lane lie in one 32-word block, so word q's
    // entries start q (blocked layout) or 32 q (plain) after word 0's.
    constexpr uint32_t Q = kSweepWords;
    const uint32_t qstep = lm ? 1u : 32u;
    // A chunk denser than the stage region (the hub z IDs of a graph
    // numbered by degree fill whole words): rounds of at most SW / 64 rows
    // per lane, each staged and flushed with coalesced stores.  The stage is
    // empty when this runs (flushed before).
    auto rounds = [&](uint32_t v, uint32_t zw) {
        const uint32_t m = SW / 64u;
        for (;;) {
            uint32_t cr = (uint32_t)__popc(v);
            cr = cr < m ? cr : m;
            const uint32_t ir = wave_scan_u32(cr);
            const uint32_t tr = (uint32_t)__builtin_amdgcn_readlane((int)ir, 63);
            if (tr == 0u) break;
            uint32_t p = ir - cr;
            for (uint32_t j = 0; j < cr; j++) {
                const uint32_t b = (uint32_t)__ffs(v) - 1u;
                v &= v - 1u;
                stg[p++] = zw + (b << lsh);
            }
            __builtin_amdgcn_wave_barrier();
            flush(tr);
        }
    };
    for (uint32_t i0 = wid * 64u * Q; i0 < W; i0 += 64u * Q * NW) {
        const uint32_t w = i0 + Q * lane;
        uint32_t v[Q];
#pragma unroll
        for (uint32_t q = 0; q < Q; q++) v[q] = 0u;
        if (w + Q <= W) {
            if (Q == 4u) {
                uint4* t4 = reinterpret_cast<uint4*>(tile + w);
                const uint4 t = *t4;
                v[0] = t.x; v[Q > 1 ? 1 : 0] = t.y; v[Q > 2 ? 2 : 0] = t.z; v[Q > 3 ? 3 : 0] = t.w;
                if (t.x | t.y | t.z | t.w) *t4 = make_uint4(0u, 0u, 0u, 0u);
            } else if (Q == 2u) {
                uint2* t2 = reinterpret_cast<uint2*>(tile + w);
                const uint2 t = *t2;
                v[0] = t.x; v[Q > 1 ? 1 : 0] = t.y;
                if (t.x | t.y) *t2 = make_uint2(0u, 0u);
            } else {
                v[0] = tile[w];
                if (v[0]) tile[w] = 0u;
            }
        } else {
#pragma unroll
            for (uint32_t q = 0; q < Q; q++)
                if (w + q < W) {
                    v[q] = tile[w + q];
                    if (v[q]) tile[w + q] = 0u;
                }
        }
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t q = 0; q < Q; q++) cnt += (uint32_t)__popc(v[q]);
        const uint32_t incl = wave_scan_u32(cnt);
        const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        if (tot == 0u) continue;
        if (fill + tot > SW) {
            flush(fill);
            fill = 0;
        }
        const uint32_t zw = zoff + tl_base(w, lm);  // entry of word 0's bit 0
        if (tot > SW) {
#pragma unroll
            for (uint32_t q = 0; q < Q; q++) rounds(v[q], zw + q * qstep);
            continue;
        }
        uint32_t p = fill + incl - cnt;
#pragma unroll
        for (uint32_t q = 0; q < Q; q++) {
            uint32_t u = v[q];
            const uint32_t zq = zw + q * qstep;
            while (u) {
                const uint32_t b = (uint32_t)__ffs(u) - 1u;
                u &= u - 1u;
                stg[p++] = zq + (b << lsh);
            }
        }
        fill += tot;
        __builtin_amdgcn_wave_barrier();
    }
    if (fill) flush(fill);
    __syncthreads();
    return true;
}

template <int BLOCK, bool COUNT, bool ZID>
__device__ __forceinline__ void row_task(const ExpandArgs& a, uint32_t x, uint32_t t, const Lds& L, uint64_t& fsum,
                                         uint64_t& fsum2, uint64_t& fcs, uint64_t& npaths, uint64_t& frows) {
    constexpr uint32_t NW = BLOCK / 64;
    uint32_t* tile = L.tile;
    uint32_t* touched = L.touched;
    uint32_t* s_sb = L.area;
    uint32_t* s_len = L.area + BLOCK;
    uint32_t* s_pre = L.area + 2 * BLOCK;
    uint32_t* stage = L.area;
    uint32_t* sh32 = L.sh32;
    uint32_t* s_wtot = L.s_wtot;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t W = a.tile_words;
    const uint32_t lsh = COUNT ? 0u : a.swz, lm = lsh ? (1u << lsh) - 1u : 0u;  // tile bit layout
    const uint32_t dummy = min(lane, W - 1u);             // an empty path slot's own word (it adds 0)
    // rows with more paths than this sweep their tile instead of listing first touches
    const uint64_t track_limit = min(4ull * a.touched_cap, 4ull * a.tile_words);
    const uint32_t zbase = t * a.tile_size;
    // word of the tile base: a path's z ID maps straight to a word of tb
    const uint32_t wbase = COUNT ? zbase : zbase >> 5;
    uint32_t* const tb = tile - wbase;
    const uint32_t rb = a.rowptrR[x], re = a.rowptrR[x + 1];
    const int32_t xv = a.id2x[x];
    const int32_t zs_self = a.x2z != nullptr ? a.x2z[x] : -1;

    // ---- dense seed: a row with a heavy key starts from its product row ----
    const int32_t hxr = a.hx != nullptr ? a.hx[x] : -1;
    if (hxr >= 0) {
        const uint32_t h0 = a.hz_tile[t], h1 = a.hz_tile[t + 1];
        if (!COUNT && a.bkeys != nullptr) {
            // fused bit-or: the OR of the row's heavy key bitmaps, kept over
            // every z ID in the tile's own bit layout (plain or swizzled), so
            // tile word q is bitmap word t W + q.  The heavy keys are listed
            // BLOCK entries at a time.
            uint32_t* s_hk = L.area;
            const uint32_t wb = t * W;
            const uint32_t wn = wb < a.dwords ? min(W, a.dwords - wb) : 0u;
            for (uint32_t c0 = rb; c0 < re; c0 += BLOCK) {
                const uint32_t i = c0 + tid;
                int32_t h = -1;
                if (i < re) {
                    const int32_t craw = a.colR[i];
                    if (craw < 0) h = a.hid[(uint32_t)craw & 0x7FFFFFFFu];
                }
                uint32_t nh;
                const uint32_t pos = block_exclusive_scan<BLOCK, uint32_t>(h >= 0 ? 1u : 0u, nh, sh32);
                if (h >= 0) s_hk[pos] = (uint32_t)h;
                __syncthreads();
                if (nh == 0u) continue;  // block-uniform; nobody reads s_hk
                if ((W & 3u) == 0u && (a.dwords & 3u) == 0u) {
                    // 16-byte loads: tile rows and bitmap rows start on
                    // 16-byte boundaries (W and dwords multiples of 4 words).
                    // Four 16-byte positions per thread and two keys per
                    // round: eight loads in flight (L2 / MALL latency, not
                    // bandwidth, bounds the seed of a few keys).
                    const uint32_t wn4 = wn >> 2;
                    uint4* tile4 = reinterpret_cast<uint4*>(tile);
                    const uint4* base4 = reinterpret_cast<const uint4*>(a.bkeys + wb);
                    const size_t kstride4 = a.dwords >> 2;
                    for (uint32_t q0 = tid; q0 < wn4; q0 += 4u * BLOCK) {
                        uint32_t qq[4];
                        bool okq[4];
#pragma unroll
                        for (uint32_t u = 0; u < 4; u++) {
                            const uint32_t q = q0 + u * BLOCK;
                            okq[u] = q < wn4;
                            qq[u] = okq[u] ? q : q0;  // clamped, unconditional
                        }
                        uint4 acc[4];
#pragma unroll
                        for (uint32_t u = 0; u < 4; u++) acc[u] = make_uint4(0u, 0u, 0u, 0u);
                        for (uint32_t j = 0; j < nh; j += 2u) {
                            const uint4* s0 = base4 + (size_t)s_hk[j] * kstride4;
                            const uint4* s1 = base4 + (size_t)s_hk[j + 1u < nh ? j + 1u : j] * kstride4;
                            uint4 v[8];
#pragma unroll
                            for (uint32_t u = 0; u < 4; u++) {
                                v[u] = s0[qq[u]];
                                v[4 + u] = s1[qq[u]];
                            }
#pragma unroll
                            for (uint32_t u = 0; u < 4; u++) {
                                acc[u].x |= v[u].x | v[4 + u].x;
                                acc[u].y |= v[u].y | v[4 + u].y;
                                acc[u].z |= v[u].z | v[4 + u].z;
                                acc[u].w |= v[u].w | v[4 + u].w;
                            }
                        }
#pragma unroll
                        for (uint32_t u = 0; u < 4; u++) {
                            if (okq[u] && (acc[u].x | acc[u].y | acc[u].z | acc[u].w) != 0u) {
                                uint4 t4 = tile4[qq[u]];
                                t4.x |= acc[u].x; t4.y |= acc[u].y; t4.z |= acc[u].z; t4.w |= acc[u].w;
                                tile4[qq[u]] = t4;
                            }
                        }
                    }
                    __syncthreads();
                    continue;
                }
                for (uint32_t q0 = tid; q0 < wn; q0 += 4u * BLOCK) {
                    uint32_t acc[4] = {0u, 0u, 0u, 0u};
                    for (uint32_t j = 0; j < nh; j++) {
                        const uint32_t* src = a.bkeys + (size_t)s_hk[j] * a.dwords + wb;
                        // clamped, unconditional loads: four words in flight
#pragma unroll
                        for (uint32_t u = 0; u < 4; u++) {
                            const uint32_t q = q0 + u * BLOCK;
                            acc[u] |= src[q < wn ? q : q0];
                        }
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 4; u++) {
                        const uint32_t q = q0 + u * BLOCK;
                        if (q < wn && acc[u] != 0u) {
                            tile[q] |= acc[u];
                        }
                    }
                }
                __syncthreads();
            }
        } else if (COUNT) {
            // distinct columns are distinct counters of a zeroed tile
            const int32_t* crow = a.cint + (size_t)hxr * a.Nh;
            for (uint32_t h = h0 + tid; h < h1; h += BLOCK) {
                const int32_t v = crow[h];
                if (v != 0) tile[a.hz2zid[h] - zbase] = (uint32_t)v;
            }
        } else if (h0 < h1 && lm == 0u && a.hz2zid[h1 - 1u] - a.hz2zid[h0] == h1 - 1u - h0) {
            // the tile's product columns are consecutive z IDs: shifted word
            // copies (plain bit layout), tile bit p <- column p - off
            const uint32_t* brow = a.dbits + (size_t)hxr * a.dwords;
            const int64_t off = (int64_t)a.hz2zid[h0] - (int64_t)zbase - (int64_t)h0;
            const uint32_t ql = (uint32_t)(((int64_t)h0 + off) >> 5), qh = (uint32_t)(((int64_t)h1 - 1 + off) >> 5);
            for (uint32_t q = ql + tid; q <= qh; q += BLOCK) {
                const int64_t c0 = (int64_t)q * 32 - off;  // column of tile bit 32 q
                const int64_t wlo = c0 >= 0 ? c0 >> 5 : -1;
                const uint32_t sh = (uint32_t)(c0 - wlo * 32);
                const uint32_t lo = (wlo >= 0 && wlo < (int64_t)a.dwords) ? brow[wlo] : 0u;
                const uint32_t hi = (sh != 0u && wlo + 1 < (int64_t)a.dwords) ? brow[wlo + 1] : 0u;
                uint32_t v = sh ? (lo >> sh) | (hi << (32u - sh)) : lo;
                // keep columns h0 .. h1 - 1 only
                const int64_t first = (int64_t)h0 - c0, last = (int64_t)h1 - 1 - c0;  // bit range in v
                if (first > 0) v &= first >= 32 ? 0u : ~0u << first;
                if (last < 31) v &= last < 0 ? 0u : (~0u >> (31 - last));
                tile[q] = v;  // one writer per word of a zeroed tile
            }
        } else if (h0 < h1) {
            const uint32_t* brow = a.dbits + (size_t)hxr * a.dwords;
            const uint32_t wl = h0 >> 5, wh = (h1 - 1u) >> 5;
            for (uint32_t wd = wl + tid; wd <= wh; wd += BLOCK) {
                uint32_t v = brow[wd];
                const uint32_t base = wd * 32u;
                if (base < h0) v &= ~0u << (h0 - base);
                if (h1 - base < 32u) v &= (1u << (h1 - base)) - 1u;
                // four columns per round: their z-ID loads are in flight together
                while (v) {
                    uint32_t zl[4];
                    const uint32_t b0 = (uint32_t)__ffs(v) - 1u;
#pragma unroll
                    for (uint32_t u = 0; u < 4; u++) {
                        const uint32_t b = v ? (uint32_t)__ffs(v) - 1u : 32u;
                        v &= v - 1u;
                        const uint32_t z = a.hz2zid[base + (b < 32u ? b : b0)];  // clamped, unconditional
                        zl[u] = b < 32u ? z - zbase : 0xFFFFFFFFu;
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 4; u++)
                        if (zl[u] != 0xFFFFFFFFu) atomicOr(&tile[tl_word(zl[u], lm)], tl_bit(zl[u], lsh));
                }
            }
        }
        __syncthreads();
    }

    // ---- expansion ----
    bool track = true;
    for (uint32_t c0 = rb; c0 < re; c0 += BLOCK) {
        // clamped, unconditional loads: the three loads of a y stay in flight
        const uint32_t i = c0 + tid;
        const bool in_row = i < re;
        const uint32_t craw = (uint32_t)a.colR[in_row ? i : re - 1];
        const uint32_t key = (craw & 0x7FFFFFFFu) * a.ntiles + t;
        uint32_t sb = a.rowptrS[key];
        const uint32_t se = a.rowptrS[key + 1];
        // a product row's heavy key: its product-column paths are in the dense
        // part (left out of the S row, or at its front in a dense core)
        if (a.splitS != nullptr) {
            const uint32_t sp = a.splitS[key];
            if (craw >> 31) sb += sp;
        }
        uint32_t np = 0;
        if (in_row) {
            s_sb[tid] = sb;
            s_len[tid] = se - sb;
            np = (se - sb + kPiece - 1) / kPiece;
            npaths += se - sb;
        }
        uint32_t NP;
        const uint32_t pre = block_exclusive_scan<BLOCK, uint32_t>(np, NP, sh32);
        s_pre[tid] = pre;
        if (c0 == rb) track = hxr < 0 && (uint64_t)NP * kPiece <= track_limit;
        __syncthreads();
        // A wavefront's j-th piece is wid + NW j.  Each wavefront resolves
        // the segments of its next 64 pieces at once: lane l binary-searches
        // the piece prefix table for piece j0 + l (the last segment whose
        // prefix is <= it; entries past the chunk's last y hold NP), so the
        // piece loop takes bases and lengths by readlane and never waits on
        // LDS between pieces.
        const uint32_t npw = NP > wid ? (NP - wid + NW - 1) / NW : 0u;
        for (uint32_t j0 = 0; j0 < npw; j0 += 64u) {
            uint32_t pb = 0, pl = 0;
            const uint32_t kl = wid + NW * (j0 + lane);
            if (kl < NP) {
                uint32_t seg = 0;
#pragma unroll
                for (uint32_t step = BLOCK / 2; step > 0; step >>= 1)
                    if (s_pre[seg + step] <= kl) seg += step;
                const uint32_t off = (kl - s_pre[seg]) * kPiece;
                pb = s_sb[seg] + off;
                pl = min(s_len[seg] - off, kPiece);
            }
            const uint32_t nbat = min(npw - j0, 64u);
            for (uint32_t j = 0; j < nbat; j += kPairs) {
                uint32_t col[kPairs * kUnroll], lens[kPairs];
#pragma unroll
                for (uint32_t h = 0; h < kPairs; h++) {
                    const uint32_t jj = j + h;
                    uint32_t base = (uint32_t)__builtin_amdgcn_readlane((int)pb, (int)(jj & 63u));
                    uint32_t len = jj < nbat ? (uint32_t)__builtin_amdgcn_readlane((int)pl, (int)(jj & 63u)) : 0u;
                    // The piece's columns through a buffer descriptor over exactly
                    // its len words (wavefront-uniform base and size, no per-lane
                    // address arithmetic): a lane past the end reads 0 without a
                    // memory access and is masked below.  All loads are
                    // unconditional, so they stay in flight together.
                    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
                    lens[h] = (uint32_t)__builtin_amdgcn_readfirstlane((int)len);
                    const __amdgpu_buffer_rsrc_t rs =
                        __builtin_amdgcn_make_buffer_rsrc((void*)(a.colS + base), (short)0, (int)(lens[h] * 4u), kRawBufferWord3);
#pragma unroll
                    for (uint32_t u = 0; u < kUnroll; u++)
                        col[h * kUnroll + u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)((lane + 64u * u) * 4u), 0, 0);
                }
                // Slot -> (LDS word, bit / increment).  The tile base is a
                // multiple of 1024 (blocked) or 32 (plain) z IDs, so the layout
                // maps the z ID itself, to a word of tb (the tile shifted back by
                // the base's words).
                uint32_t addr[kPairs * kUnroll], val[kPairs * kUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kPairs * kUnroll; u++) {
                    const uint32_t v = col[u];
                    addr[u] = COUNT ? v : tl_word(v, lm);
                    val[u] = COUNT ? 1u : tl_bit(v, lsh);
                }
                bool full = true;
#pragma unroll
                for (uint32_t h = 0; h < kPairs; h++) full = full && lens[h] == kPiece;
                if (!full || track) {
                    // an empty slot ORs / adds 0 into a word of its own, so the
                    // atomics issue back to back with no branch
#pragma unroll
                    for (uint32_t u = 0; u < kPairs * kUnroll; u++) {
                        const bool ok = lane + 64u * (u % kUnroll) < lens[u / kUnroll];
                        addr[u] = ok ? addr[u] : dummy + wbase;
                        val[u] = ok ? val[u] : 0u;
                    }
                }
                if (track) {
                    // first touches (the word / counter was zero) of all slots go
                    // to the list with one reservation per wavefront
                    uint64_t fm[kPairs * kUnroll];
                    uint32_t ftot = 0;
#pragma unroll
                    for (uint32_t u = 0; u < kPairs * kUnroll; u++) {
                        const uint32_t old = COUNT ? atomicAdd(&tb[addr[u]], val[u]) : atomicOr(&tb[addr[u]], val[u]);
                        fm[u] = __ballot(val[u] != 0u && old == 0u);
                        ftot += (uint32_t)__popcll(fm[u]);
                    }
                    if (ftot) {
                        uint32_t fbase = 0;
                        if (lane == 0u) fbase = atomicAdd(L.s_nt, ftot);
                        fbase = (uint32_t)__builtin_amdgcn_readfirstlane((int)fbase);
#pragma unroll
                        for (uint32_t u = 0; u < kPairs * kUnroll; u++) {
                            if ((fm[u] >> lane) & 1ull) {
                                const uint32_t p = fbase + mbcnt64(fm[u]);
                                if (p < a.touched_cap) touched[p] = addr[u] - wbase;
                            }
                            fbase += (uint32_t)__popcll(fm[u]);
                        }
                    }
                } else {
                    // untracked: no returns, so the wavefront never waits on its
                    // atomics (the emission counts the swept tile itself)
#pragma unroll
                    for (uint32_t u = 0; u < kPairs * kUnroll; u++) {
                        if (COUNT) atomicAdd(&tb[addr[u]], val[u]);
                        else atomicOr(&tb[addr[u]], val[u]);
                    }
                }
            }
        }
        __syncthreads();
    }

    // ---- emission ----
    // The list (touched words, or the whole tile) is dealt to the wavefronts
    // in chunks of 64 entries round robin: lane l of wavefront w owns entries
    // 64 (w + NW g) + l, so a wavefront reads 64 consecutive entries per step
    // (no bank conflicts) and a dense z range (the hubs of a graph numbered
    // by degree) spreads over all wavefronts instead of loading one.
    const uint32_t nt = *L.s_nt;
    const bool listed = track && nt <= a.touched_cap;
    const uint32_t count = listed ? nt : a.tile_words;
    uint32_t self_item = 0xFFFFFFFFu, self_keep = 0xFFFFFFFFu;  // u != w filter of the self-join
    {
        const int32_t zs = zs_self;
        const uint32_t zl = (uint32_t)zs - zbase;
        if (zs >= 0 && zl < a.tile_size) {
            self_item = list_item<COUNT>(zl, lm);
            self_keep = COUNT ? 0u : ~tl_bit(zl, lsh);
        }
    }
    const uint32_t cstep = 64u * NW, cfirst = wid * 64u + lane;
    if (!COUNT && !a.fp_mode && !listed && emit_swept<BLOCK, ZID>(a, L, xv, zbase, self_item, self_keep))
        return;

    if (a.fp_mode) {
        // Fingerprint, in one pass over the list (or tile) that also counts
        // the rows (summed per thread across tasks): the z factors of the
        // row's rows, then one multiply by each x factor.  A plain-layout
        // word of a dense z range is one 32-aligned word of z values (zlo and
        // the tile base are multiples of 32): its word factor (from the
        // query's table when there is one) times the byte-table sums of its
        // position factors.  Otherwise (swizzled layout, hashed z, counters)
        // each row takes its own z factor.
        const uint64_t* fpt = L.fpt;
        const uint32_t zoff = (uint32_t)a.zlo + zbase;
        const int32_t* id2z = a.id2z + zbase;
        uint64_t S1 = 0, S2 = 0;
        auto take = [&](uint32_t i) {  // word i of the list, filtered (u != w) and cleared
            const uint32_t w = list_at(touched, listed, i);
            uint32_t v = tile[w];
            tile[w] = 0;
            if (w == self_item) v &= self_keep;
            frows += units_of<COUNT>(v);
            return make_uint2(w, v);
        };
        // the table's word j is z word (zlo >> 5) + j: local z offset e is in
        // word (zbase + e) >> 5 (a project tile's zbase is a multiple of 32)
        const uint32_t* fzw = ZID ? a.fp_zw : nullptr;
        if (!COUNT && ZID && lm == 0u) {
            if (fzw != nullptr) {
                fzw += zbase >> 5;
                for (uint32_t i = cfirst; i < count; i += cstep) {
                    const uint2 wv = take(i);
                    const uint32_t v = wv.y;
                    if (v) {
                        const uint64_t ps = fpt[32u + (v & 255u)] + fpt[288u + ((v >> 8) & 255u)] +
                                            fpt[544u + ((v >> 16) & 255u)] + fpt[800u + (v >> 24)];
                        S1 += (uint64_t)fzw[wv.x] * ps;
                    }
                }
            } else {
                const uint32_t wv0 = zoff >> 5;
                for (uint32_t i = cfirst; i < count; i += cstep) {
                    const uint2 wv = take(i);
                    const uint32_t v = wv.y;
                    if (v) {
                        const uint64_t ps = fpt[32u + (v & 255u)] + fpt[288u + ((v >> 8) & 255u)] +
                                            fpt[544u + ((v >> 16) & 255u)] + fpt[800u + (v >> 24)];
                        S1 += (uint64_t)zword_factor((wv0 + wv.x) & 0x07FFFFFFu) * ps;  // (z >> 5) of a 32-bit z: 27 bits
                    }
                }
            }
            S2 = S1;
        } else {
            for (uint32_t i = cfirst; i < count; i += cstep) {
                const uint2 wv = take(i);
                const uint32_t w = wv.x;
                uint32_t v = wv.y;
                if (COUNT) {
                    if (v) {
                        const uint32_t u = ZID ? zoff + w : (uint32_t)id2z[w];
                        const uint32_t fw = fzw != nullptr ? fzw[(zbase + w) >> 5] : zword_factor(u >> 5);
                        const uint64_t g = (uint64_t)fw * fpt[u & 31u];
                        S1 += g;
                        S2 += (uint64_t)v * g;
                        fcs += v;
                    }
                } else {
                    while (v) {
                        const uint32_t b = (uint32_t)__ffs(v) - 1u;
                        v &= v - 1u;
                        const uint32_t e = tl_base(w, lm) + (b << lsh);
                        const uint32_t u = ZID ? zoff + e : (uint32_t)id2z[e];
                        const uint32_t fw = fzw != nullptr ? fzw[(zbase + e) >> 5] : zword_factor(u >> 5);
                        S1 += (uint64_t)fw * fpt[u & 31u];
                    }
                }
            }
            if (!COUNT) S2 = S1;
        }
        if (S1 | S2) {
            fsum += xfactor1(xv) * S1;
            fsum2 += xfactor2(xv) * S2;
        }
        __syncthreads();
        return;
    }
    uint32_t c = 0;
    uint64_t nzmask = 0;  // bit g: the lane's word of step g holds rows
    {
        uint32_t g = 0;
        for (uint32_t i = cfirst; i < count; i += cstep, g++) {
            const uint32_t w = list_at(touched, listed, i);
            uint32_t v = tile[w];
            if (w == self_item) {
                v &= self_keep;
                tile[w] = v;
            }
            c += units_of<COUNT>(v);
            if (v != 0u && g < 64u) nzmask |= 1ull << g;
        }
    }
    const uint32_t wtot = wave_sum_u32(c);
    if (lane == 0) s_wtot[wid] = wtot;
    __syncthreads();
    uint32_t woff = 0, T = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; w++) {
        const uint32_t v = s_wtot[w];
        woff += w < wid ? v : 0u;
        T += v;
    }
    const uint32_t zoff = (uint32_t)a.zlo + zbase;
    const int32_t* id2z = a.id2z + zbase;


    if (COUNT || (uint64_t)T * 16u <= count) {
        // Direct (COUNT, or at most one row per 16 listed words): each lane
        // writes its word's rows, and the first rows of 64 neighbouring
        // words land on neighbouring addresses.  Denser rows go through the
        // staging below, whose balanced per-thread shares beat the per-word
        // bit loops (two-hop graph: 71 -> 65 ms per expansion).
        if (tid == 0) *L.s_gbase = T ? atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)T) : 0ull;
        __syncthreads();
        uint64_t pos0 = *L.s_gbase + woff;
        for (uint32_t i0 = wid * 64u; i0 < count; i0 += cstep) {
            const uint32_t i = i0 + lane;
            uint32_t v = 0, w = 0;
            if (i < count) {
                w = list_at(touched, listed, i);
                v = tile[w];
            }
            // prefix of the per-word row counts from one ballot per count
            // level (rows are sparse here: usually one to three levels)
            const uint32_t cc = units_of<COUNT>(v);
            uint32_t before = 0, gtot = 0;
            for (uint32_t lvl = 1;; lvl++) {
                const uint64_t m = __ballot(cc >= lvl);
                if (m == 0) break;
                before += mbcnt64(m);
                gtot += (uint32_t)__popcll(m);
                if (COUNT) break;
            }
            uint64_t pos = pos0 + before;
            pos0 += gtot;
            if (COUNT) {
                if (v != 0u && pos < a.cap) {
                    dim3_count_row row;
                    row.x = xv;
                    row.z = ZID ? (int32_t)(zoff + w) : id2z[w];
                    row.count = (int64_t)v;
                    reinterpret_cast<dim3_count_row*>(a.out)[pos] = row;
                }
            } else {
                int2* out = reinterpret_cast<int2*>(a.out);
                while (v) {
                    const uint32_t b = (uint32_t)__ffs(v) - 1u;
                    v &= v - 1u;
                    const uint32_t e = tl_base(w, lm) + (b << lsh);
                    if (pos < a.cap) put_pair(out + pos, xv, ZID ? (int32_t)(zoff + e) : id2z[e]);
                    ++pos;
                }
            }
            if (i < count) tile[w] = 0;
        }
        __syncthreads();
        return;
    }

    // Staged (dense rows): the row's output range is reserved once and its
    // result consumed only after the first fill.  P passes; every thread
    // spreads its c rows evenly over them (q or q + 1 per pass), so all
    // wavefronts fill in every pass and a pass holds at most
    // T / P + BLOCK <= stage_cap rows.
    uint64_t my_base = 0;
    if (tid == 0 && T) my_base = atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)T);
    const uint32_t P = (T + (a.stage_cap - BLOCK) - 1) / (a.stage_cap - BLOCK);
    const uint32_t cq = c / P, cr = c - cq * P;
    const bool use_mask = count <= wid * 64u + 64u * cstep;  // at most 64 steps per lane
    uint32_t g = 0, cw = 0, cv = 0;
    if (!use_mask && cfirst < count) {
        cw = list_at(touched, listed, cfirst);
        cv = tile[cw];
    }
    uint64_t pass_base = 0;
    for (uint32_t p = 0; p < P; p++) {
        const uint32_t mine = cq + (p < cr ? 1u : 0u);
        uint32_t Tp;
        const uint32_t op = block_scan2_u32<BLOCK>(mine, Tp, sh32 + (p & 1) * (NW + 1));
        for (uint32_t j = 0; j < mine; j++) {
            if (use_mask) {
                // jump straight to the next word that holds rows
                if (cv == 0u) {
                    const uint32_t gg = (uint32_t)__ffsll((unsigned long long)nzmask) - 1u;
                    nzmask &= nzmask - 1ull;
                    cw = list_at(touched, listed, cfirst + cstep * gg);
                    cv = tile[cw];
                }
            } else {
                while (cv == 0u) {
                    ++g;
                    cw = list_at(touched, listed, cfirst + cstep * g);
                    cv = tile[cw];
                }
            }
            if (COUNT) {
                stage[op + j] = cw;
                cv = 0;
            } else {
                const uint32_t b = (uint32_t)__ffs(cv) - 1u;
                cv &= cv - 1u;
                stage[op + j] = tl_base(cw, lm) + (b << lsh);
            }
        }
        if (tid == 0 && p == 0) *L.s_gbase = my_base;
        __syncthreads();
        const uint64_t g0 = *L.s_gbase + pass_base;
        const uint32_t nq = Tp;
        int2* out = reinterpret_cast<int2*>(a.out) + g0;
        if (g0 + nq <= a.cap) {
            uint32_t k = tid;
            for (; k + 3u * BLOCK < nq; k += 4u * BLOCK) {
                const uint32_t e0 = stage[k], e1 = stage[k + BLOCK];
                const uint32_t e2 = stage[k + 2u * BLOCK], e3 = stage[k + 3u * BLOCK];
                const int32_t z0 = ZID ? (int32_t)(zoff + e0) : id2z[e0];
                const int32_t z1 = ZID ? (int32_t)(zoff + e1) : id2z[e1];
                const int32_t z2 = ZID ? (int32_t)(zoff + e2) : id2z[e2];
                const int32_t z3 = ZID ? (int32_t)(zoff + e3) : id2z[e3];
                put_pair(out + k, xv, z0);
                put_pair(out + k + BLOCK, xv, z1);
                put_pair(out + k + 2u * BLOCK, xv, z2);
                put_pair(out + k + 3u * BLOCK, xv, z3);
            }
            for (; k < nq; k += BLOCK) {
                const uint32_t e = stage[k];
                put_pair(out + k, xv, ZID ? (int32_t)(zoff + e) : id2z[e]);
            }
        } else {
            for (uint32_t k = tid; k < nq; k += BLOCK) {
                const uint32_t e = stage[k];
                if (g0 + k < a.cap) put_pair(out + k, xv, ZID ? (int32_t)(zoff + e) : id2z[e]);
            }
        }
        pass_base += nq;
        __syncthreads();
    }
    for (uint32_t i = cfirst; i < count; i += cstep) tile[list_at(touched, listed, i)] = 0;
    __syncthreads();
}

// Sums of the per-lane row statistics into the instance's scalars (block-uniform call).
template <int BLOCK, bool COUNT>
__device__ __forceinline__ void flush_stats(const ExpandArgs& a, uint64_t& fsum, uint64_t& fsum2, uint64_t& fcs,
                                            uint64_t& npaths, uint64_t& frows, uint64_t* sh) {
    const uint64_t np = block_reduce_sum<BLOCK, uint64_t>(npaths, sh);
    if (threadIdx.x == 0 && np) atomicAdd((unsigned long long*)&a.sc->paths, (unsigned long long)np);
    if (a.fp_mode) {
        const uint64_t s1 = block_reduce_sum<BLOCK, uint64_t>(fsum, sh);
        const uint64_t s2 = block_reduce_sum<BLOCK, uint64_t>(fsum2, sh);
        const uint64_t s3 = block_reduce_sum<BLOCK, uint64_t>(fcs, sh);
        const uint64_t nr = block_reduce_sum<BLOCK, uint64_t>(frows, sh);
        if (threadIdx.x == 0) {
            atomicAdd((unsigned long long*)&a.sc->fp_sum, (unsigned long long)s1);
            atomicAdd((unsigned long long*)&a.sc->fp_sum2, (unsigned long long)s2);
            if (COUNT) atomicAdd((unsigned long long*)&a.sc->fp_csum, (unsigned long long)s3);
            if (nr) atomicAdd((unsigned long long*)&a.sc->out_n, (unsigned long long)nr);
        }
    }
    fsum = fsum2 = fcs = npaths = frows = 0;
}

// The fingerprint table in LDS: the 32 position factors, then for each byte
// j of a word and each byte value m the sum of the position factors of m's
// set bits (positions 8 j .. 8 j + 7).  Ends with a barrier.
template <int BLOCK>
__device__ void fp_table_init(uint64_t* T) {
    for (uint32_t k = threadIdx.x; k < 32u; k += BLOCK) T[k] = zpos_factor(k);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 1024u; k += BLOCK) {
        const uint32_t j = k >> 8, m = k & 255u;
        uint64_t sum = 0;
#pragma unroll
        for (uint32_t i = 0; i < 8u; i++)
            if ((m >> i) & 1u) sum += T[8u * j + i];
        T[32u + k] = sum;
    }
    __syncthreads();
}

template <int BLOCK, bool COUNT, bool ZID, bool TM>
__global__ __launch_bounds__(BLOCK) void k_expand(ExpandArgs a) {
    extern __shared__ __align__(16) uint32_t smem[];
    constexpr uint32_t NW = BLOCK / 64;
    __shared__ uint64_t sh[NW + 1];
    __shared__ uint32_t sh32[2 * (NW + 1)];
    __shared__ uint32_t s_wtot[NW];
    __shared__ uint32_t s_task, s_nt, s_cursor;
    __shared__ uint64_t s_gbase;
    uint32_t* dyn = smem + (a.fp_mode ? 2u * kFpTableU64 : 0u);
    const Lds L{reinterpret_cast<const uint64_t*>(smem), dyn, dyn + a.tile_words,
                dyn + a.tile_words + a.touched_cap, sh, sh32, s_wtot, &s_nt, &s_gbase, &s_cursor};

    const uint32_t tid = threadIdx.x;
    if (a.fp_mode) fp_table_init<BLOCK>(reinterpret_cast<uint64_t*>(smem));
    for (uint32_t w = tid; w < a.tile_words; w += BLOCK) L.tile[w] = 0;
    uint64_t fsum = 0, fsum2 = 0, fcs = 0, npaths = 0, frows = 0;
    // with a row list the short rows went one per wavefront (k_value_rows):
    // the tasks are (big_x[r], t) for the other rows
    const bool lists = a.big_x != nullptr;
    const uint32_t lo = lists ? 0u : a.task_lo;
    const uint32_t hi = lists ? ((a.row_hi - a.row_lo) - (uint32_t)*a.n_wave) * a.ntiles : a.task_hi;
    // TM (a fused bit-or seed whose key bitmaps exceed an XCD's L2): tasks
    // are taken tile-major, so the workgroups running at one time share one
    // or two z tiles and the heavy-key bitmap words of those tiles stay in
    // every XCD's L2 (10M x 10M Zipf COUNT fingerprint: 327 -> 319 ms).
    // Otherwise row-major: a row's tiles run side by side and share its y
    // list.  A template argument: the branch alone cost the two-hop graph 1.3%.
    const uint32_t nrows = (hi - lo) / a.ntiles;

    // The next task is claimed while the current one runs (the claim's
    // latency overlaps the task; its value reaches LDS after the task).
    uint32_t next = 0;
    if (tid == 0) next = atomicAdd(&a.sc->task, 1u);
    for (;;) {
        if (tid == 0) {
            s_task = next;
            s_nt = 0;
        }
        __syncthreads();
        // task and x are wavefront-uniform: scalar loads
        uint32_t task = lo + __builtin_amdgcn_readfirstlane(s_task);
        if (task >= hi) break;
        if (tid == 0) next = atomicAdd(&a.sc->task, 1u);
        if (TM) {  // the k-th claim is row k % nrows of tile k / nrows
            const uint32_t k = task - lo, t = k / nrows;
            task = lo + (k - t * nrows) * a.ntiles + t;
        }
        const uint32_t r = task / a.ntiles;
        const uint32_t x = lists ? a.big_x[r] : r;
        row_task<BLOCK, COUNT, ZID>(a, x, task - r * a.ntiles, L, fsum, fsum2, fcs, npaths, frows);
    }

    flush_stats<BLOCK, COUNT>(a, fsum, fsum2, fcs, npaths, frows, sh);
}

// Many join instances in one persistent launch fed by one device work
// counter: global task g belongs to the instance i with
// task_start[i] <= g < task_start[i + 1] (tasks of instance i are its
// (x, z-tile) pairs).  A workgroup keeps one instance's arguments while its
// consecutive tasks stay in that instance, and flushes the instance's row
// statistics when it moves on.  LDS is sized for the largest tile of the batch.
template <int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_expand_batch(const ExpandArgs* __restrict__ args,
                                                        const uint32_t* __restrict__ task_start, uint32_t n_inst,
                                                        uint32_t* counter, uint32_t fp_mode) {
    extern __shared__ __align__(16) uint32_t smem[];
    constexpr uint32_t NW = BLOCK / 64;
    __shared__ uint64_t sh[NW + 1];
    __shared__ uint32_t sh32[2 * (NW + 1)];
    __shared__ uint32_t s_wtot[NW];
    __shared__ uint32_t s_task, s_nt, s_cursor;
    __shared__ uint64_t s_gbase;

    const uint32_t tid = threadIdx.x;
    const uint32_t total = task_start[n_inst];
    uint64_t fsum = 0, fsum2 = 0, fcs = 0, npaths = 0, frows = 0;
    int32_t cur = -1;
    ExpandArgs a;
    // every instance of a batch has the batch's output mode
    const bool fp = fp_mode != 0u;
    uint32_t* dyn = smem + (fp ? 2u * kFpTableU64 : 0u);
    if (fp) fp_table_init<BLOCK>(reinterpret_cast<uint64_t*>(smem));
    uint32_t next = 0;  // the next task, claimed while the current one runs
    if (tid == 0) next = atomicAdd(counter, 1u);
    for (;;) {
        if (tid == 0) {
            s_task = next;
            s_nt = 0;
        }
        __syncthreads();
        const uint32_t task = __builtin_amdgcn_readfirstlane(s_task);
        if (task >= total) break;
        if (tid == 0) next = atomicAdd(counter, 1u);
        // largest i with task_start[i] <= task (task_start[n_inst] = total > task)
        uint32_t lo = 0, hi = n_inst;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (task_start[mid] <= task) lo = mid;
            else hi = mid;
        }
        if ((int32_t)lo != cur) {
            if (cur >= 0) flush_stats<BLOCK, COUNT>(a, fsum, fsum2, fcs, npaths, frows, sh);
            cur = (int32_t)lo;
            a = args[lo];
            // the previous instance's touched list / staging area may overlap this tile
            for (uint32_t w = tid; w < a.tile_words; w += BLOCK) dyn[w] = 0;
            __syncthreads();
        }
        const uint32_t local = task - task_start[lo];
        const uint32_t x = local / a.ntiles;
        const Lds L{reinterpret_cast<const uint64_t*>(smem), dyn, dyn + a.tile_words,
                    dyn + a.tile_words + a.touched_cap, sh, sh32, s_wtot, &s_nt, &s_gbase, &s_cursor};
        if (a.id2z == nullptr)
            row_task<BLOCK, COUNT, true>(a, x, local - x * a.ntiles, L, fsum, fsum2, fcs, npaths, frows);
        else row_task<BLOCK, COUNT, false>(a, x, local - x * a.ntiles, L, fsum, fsum2, fcs, npaths, frows);
    }
    if (cur >= 0) flush_stats<BLOCK, COUNT>(a, fsum, fsum2, fcs, npaths, frows, sh);
}

// Upper bound of each row's output: its S paths over all tiles (a product
// row also its product columns), capped at `cap` (<= nz; the loop stops
// there, so a hub row costs no more than a short one).
__global__ __launch_bounds__(256) void k_row_bound(ExpandArgs a, uint32_t cap, uint32_t* __restrict__ bound) {
    const uint32_t x = blockIdx.x * 256u + threadIdx.x;
    if (x >= a.nx) return;
    uint64_t b = (a.hx != nullptr && a.hx[x] >= 0) ? a.Nh : 0u;
    for (uint32_t i = a.rowptrR[x], e = a.rowptrR[x + 1]; i < e && b < cap; i++) {
        const uint32_t y = (uint32_t)a.colR[i] & 0x7FFFFFFFu;
        b += a.rowptrY != nullptr ? a.rowptrY[y + 1u] - a.rowptrY[y]
                                  : a.rowptrS[(y + 1u) * a.ntiles] - a.rowptrS[y * a.ntiles];
    }
    bound[x] = (uint32_t)min<uint64_t>(b, cap);
}

__global__ void k_fp_word_factors(uint32_t* __restrict__ out, uint32_t n, uint32_t w0) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = zword_factor((w0 + k) & 0x07FFFFFFu);
}

// Zero the expansion-owned scalars of every instance of a batch (a kernel, so
// a captured graph holds no memset node).
__global__ void k_reset_scalars(dim3h::Scalars* const* __restrict__ list, uint32_t n, uint32_t* counter) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *counter = 0;
    if (i >= n) return;
    dim3h::Scalars* s = list[i];
    s->out_n = 0;
    s->fp_sum = 0;
    s->fp_sum2 = 0;
    s->fp_csum = 0;
    s->paths = 0;
    s->task = 0;
}

__global__ void k_gather_scalars(dim3h::Scalars* const* __restrict__ list, uint32_t n, dim3h::Scalars* out) {
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const uint64_t* src = reinterpret_cast<const uint64_t*>(list[i]);
    uint64_t* dst = reinterpret_cast<uint64_t*>(out + i);
    for (uint32_t w = threadIdx.x; w < sizeof(dim3h::Scalars) / 8; w += blockDim.x) dst[w] = src[w];
}


template <typename K>
int launch_persistent(K kern, int block, size_t lds, const ExpandArgs& a, int cu_count, hipStream_t st) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, lds);
    if (e != hipSuccess) return (int)e;
    if (occ < 1) occ = 1;
    const uint64_t ntasks = a.task_hi > a.task_lo ? a.task_hi - a.task_lo : 0;
    uint64_t grid = (uint64_t)cu_count * (uint64_t)occ;
    if (grid > ntasks) grid = ntasks;
    if (grid == 0) return 0;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), lds, st, a);
    return (int)hipGetLastError();
}

template <bool COUNT, bool ZID, bool TM = false>
int launch_z(const ExpandArgs& a, int block, size_t lds, int cu_count, hipStream_t st) {
    switch (block) {
        case 256: return launch_persistent(k_expand<256, COUNT, ZID, TM>, 256, lds, a, cu_count, st);
        case 512: return launch_persistent(k_expand<512, COUNT, ZID, TM>, 512, lds, a, cu_count, st);
        case 1024: return launch_persistent(k_expand<1024, COUNT, ZID, TM>, 1024, lds, a, cu_count, st);
        default: return (int)hipErrorInvalidValue;
    }
}

template <bool COUNT>
int launch(const ExpandArgs& a, int block, int cu_count, hipStream_t st) {
    const size_t lds = dim3h::expand_lds_bytes(a.tile_words, a.touched_cap, a.stage_cap, (uint32_t)block,
                                               a.fp_mode != 0u);
    // a fused bit-or seed whose key bitmaps outgrow an L2: tile-major tasks
    if (!COUNT && a.bkeys != nullptr && a.id2z == nullptr && a.ntiles > 1u && a.tile_major)
        return launch_z<false, true, true>(a, block, lds, cu_count, st);
    if (a.id2z == nullptr) return launch_z<COUNT, true>(a, block, lds, cu_count, st);
    return launch_z<COUNT, false>(a, block, lds, cu_count, st);
}

}  // namespace

namespace dim3h {

void fp_word_factors(uint32_t* out, uint32_t n, int32_t zlo, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fp_word_factors, dim3((n + 255u) / 256u), dim3(256), 0, st, out, n, (uint32_t)zlo >> 5);
}

size_t expand_lds_bytes(uint32_t tile_words, uint32_t touched_cap, uint32_t stage_cap, uint32_t block, bool fp) {
    const uint32_t area = stage_cap > 3u * block + 1u ? stage_cap : 3u * block + 1u;
    return ((size_t)tile_words + touched_cap + area) * 4u + (fp ? (size_t)kFpTableU64 * 8u : 0u);
}

void row_bounds(const ExpandArgs& a, uint32_t cap, uint32_t* bound, hipStream_t st) {
    if (a.nx) hipLaunchKernelGGL(k_row_bound, dim3((a.nx + 255u) / 256u), dim3(256), 0, st, a, cap, bound);
}

int expand_project(const ExpandArgs& a, int block, int cu_count, hipStream_t st) {
    return launch<false>(a, block, cu_count, st);
}

int expand_count(const ExpandArgs& a, int block, int cu_count, hipStream_t st) {
    return launch<true>(a, block, cu_count, st);
}

template <int BLOCK, bool COUNT>
int batch_prepare_one(size_t lds, int cu_count, uint64_t total_tasks, unsigned* grid) {
    const void* kern = (const void*)k_expand_batch<BLOCK, COUNT>;
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_expand_batch<BLOCK, COUNT>, BLOCK, lds);
    if (e != hipSuccess) return (int)e;
    if (occ < 1) occ = 1;
    uint64_t g = (uint64_t)cu_count * (uint64_t)occ;
    if (g > total_tasks) g = total_tasks;
    *grid = (unsigned)(g ? g : 1);
    return 0;
}

int expand_batch_setup(bool count, int block, size_t lds, int cu_count, uint64_t total_tasks, unsigned* grid) {
    switch (block) {
        case 256: return count ? batch_prepare_one<256, true>(lds, cu_count, total_tasks, grid)
                               : batch_prepare_one<256, false>(lds, cu_count, total_tasks, grid);
        case 512: return count ? batch_prepare_one<512, true>(lds, cu_count, total_tasks, grid)
                               : batch_prepare_one<512, false>(lds, cu_count, total_tasks, grid);
        case 1024: return count ? batch_prepare_one<1024, true>(lds, cu_count, total_tasks, grid)
                                : batch_prepare_one<1024, false>(lds, cu_count, total_tasks, grid);
        default: return (int)hipErrorInvalidValue;
    }
}

void expand_batch_launch(bool count, bool fp, int block, size_t lds, unsigned grid, const ExpandArgs* args,
                         const uint32_t* task_start, uint32_t n_inst, uint32_t* counter, Scalars* const* sc_list,
                         Scalars* sc_out, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_scalars, dim3((n_inst + 255) / 256), dim3(256), 0, st, sc_list, n_inst, counter);
#define DIM3_BATCH_LAUNCH(B, CNT) \
    hipLaunchKernelGGL((k_expand_batch<B, CNT>), dim3(grid), dim3(B), lds, st, args, task_start, n_inst, counter, \
                       fp ? 1u : 0u)
    if (block == 256) {
        if (count) DIM3_BATCH_LAUNCH(256, true);
        else DIM3_BATCH_LAUNCH(256, false);
    } else if (block == 512) {
        if (count) DIM3_BATCH_LAUNCH(512, true);
        else DIM3_BATCH_LAUNCH(512, false);
    } else {
        if (count) DIM3_BATCH_LAUNCH(1024, true);
        else DIM3_BATCH_LAUNCH(1024, false);
    }
#undef DIM3_BATCH_LAUNCH
    hipLaunchKernelGGL(k_gather_scalars, dim3(n_inst), dim3(64), 0, st, sc_list, n_inst, sc_out);
}

}  // namespace dim3h
