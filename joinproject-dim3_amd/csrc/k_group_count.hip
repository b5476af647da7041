// COUNT fingerprint with every group's count formed: the heavy keys' counts
// as an int8 0/1 product on the MFMA cores, the light keys' counts in
// per-wavefront LDS tables.
//
// Replaces SpmmAggregate with a COUNT policy (dim3.hpp:1267-1311) for a
// query whose heavy join keys make the (x, z) box dense (the 10M x 10M
// Zipf(1.2) COUNT has 2.4e12 groups in a 2e6 x 2e6 box).  The reference
// counts each group in its SPA and hands (x, z, count) to the policy; here
// the count of group (x, z) is
//   c(x, z) = c_d(x, z) + c_l(x, z),
//   c_d = popcount(hm(x) & zm(z))   (hm / zm: bit h set when x / z has a
//                                    tuple with heavy key h, K <= 64 keys)
//   c_l = join paths of (x, z) through light keys.
// The query's fingerprint (count, sum, sum2, csum) is a sum of per-group
// terms (1, k1(x) g(z), k2(x) c g(z), c) over the groups with c > 0.  The
// last two are linear in c, so they are sums over the join paths: sum2 =
// sum over keys y of (sum of k2(x) over y's R tuples) x (sum of g(z) over
// y's S tuples) and csum = the path count, both formed in the planning
// (path_sums).  The first two need each group's presence, and split
// exactly into
//   dense part (k_gc_dense): the terms of [c_d > 0] over every cell of the
//     box of rows with hm != 0 and columns with zm != 0, where c_d is
//     A[x][h] * B[z][h] summed over h on v_mfma_i32_16x16x64_i8 (A, B the
//     0/1 bytes of hm, zm: exact, c_d <= 64);
//   light part (k_gc_light): per light group (c_l > 0) found in an LDS
//     table, c_d = popcount(hm(x) & zm(z)) looked up, and the terms
//     (1, k1 g) when c_d == 0: a group the dense part already counted adds
//     nothing.
// The per-cell presence sum of the dense part is itself a matrix product:
// with g(z) = sum_b 2^(7 b) g_b(z) in 7-bit chunks (b < 10),
//   sum_{x,z} k1(x) [c(x,z) > 0] g(z) = sum_x k1(x) sum_b 2^(7 b) E1[b][x],
//   E1[b][x] = sum_z g_b(z) [c(x, z) > 0],
// and E1 is a second int8 MFMA whose k dimension is z: its B operand is the
// flag bytes (-1 where c > 0) of the count tiles of the first product
// (computed transposed and negated, -C^T[z][x], so the lane holding z =
// 4 q .. 4 q + 3 of a tile holds the operand bytes of k = 16 q .. 16 q + 15
// once the sign bits of the four z tiles of a step are spread to bytes in
// registers) and its A operand the chunk bytes of the step's 64 z factors,
// laid out per lane in advance.  E1 accumulates in the MFMA over the whole
// z sweep of a task (minus its value); k1(x) multiplies once per row at the
// end.  Row b = 10 of
// the chunk operand is all ones, so E1[10][x] counts the row's groups.  Per
// 4096 cells a wavefront issues 16 product and 4 reduction MFMAs and 3 VALU
// instructions per four counts; no count leaves the registers.
#include "dim3_internal.h"

namespace {

using namespace dim3d;
using dim3h::Scalars;

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr uint32_t kGcRowsW = 64;                 // box rows (x) per wavefront: 4 groups of 16
constexpr uint32_t kGcRowsB = 4 * kGcRowsW;       // box rows per workgroup (4 wavefronts)
constexpr uint32_t kGcColsStep = 64;              // box columns (z) per step: the reduction's k
constexpr uint32_t kGcK = 64;                     // heavy keys (bytes per operand row)
// word 3 of a raw (untyped, unswizzled) buffer descriptor on CDNA: 32-bit
// data format, loads past num_records return 0
constexpr int kRawBufferWord3 = 0x00020000;

__device__ __forceinline__ uint32_t key_chunk(uint64_t k, uint32_t b) {
    return b < 9u ? (uint32_t)(k >> (7u * b)) & 127u : b == 9u ? (uint32_t)(k >> 63) : b == 10u ? 1u : 0u;
}

// hm[x] = OR of 1 << hid[y] over the heavy keys y of row x (R CSR by x),
// and bit 31 set on the row's heavy entries (the light part skips them).
// A heavy (x, y) tuple seen twice sets *dup: a bit mask cannot count it.
__global__ void k_gc_row_masks(const uint32_t* __restrict__ rowptrR, int32_t* __restrict__ colR, uint32_t nx,
                               const int32_t* __restrict__ hid, uint64_t* __restrict__ hm,
                               uint32_t* __restrict__ flag, unsigned long long* dup) {
    const uint32_t x = blockIdx.x * 256u + threadIdx.x;
    if (x >= nx) return;
    uint64_t m = 0;
    bool twice = false;
    for (uint32_t i = rowptrR[x], e = rowptrR[x + 1]; i < e; i++) {
        const uint32_t y = (uint32_t)colR[i] & 0x7FFFFFFFu;
        const int32_t h = hid[y];
        if (h >= 0) {
            twice = twice || ((m >> h) & 1ull);
            m |= 1ull << h;
            colR[i] = (int32_t)(y | 0x80000000u);
        }
    }
    hm[x] = m;
    flag[x] = m != 0ull ? 1u : 0u;
    if (twice) atomicOr(dup, 1ull);
}

// zm[z] |= 1 << hid[y] for every S tuple (z, y) with a heavy key (a
// duplicate sets *dup)
__global__ void k_gc_col_masks(const int2* __restrict__ S, uint64_t nS, const int32_t* __restrict__ slotS,
                               const int32_t* __restrict__ slotSz, const uint64_t* __restrict__ ex_y,
                               const uint64_t* __restrict__ ex_z, int32_t zlo, const int32_t* __restrict__ hid,
                               uint64_t* __restrict__ zm, unsigned long long* dup) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= nS) return;
    const int32_t s = slotS[i];
    if (s < 0) return;
    const int32_t h = hid[(uint32_t)(ex_y[s] >> 32)];
    if (h < 0) return;
    const uint32_t zid = ex_z != nullptr ? (uint32_t)(ex_z[slotSz[i]] >> 32) : (uint32_t)S[i].x - (uint32_t)zlo;
    const uint64_t old = atomicOr((unsigned long long*)&zm[zid], 1ull << h);
    if ((old >> h) & 1ull) atomicOr(dup, 1ull);
}

__global__ void k_gc_col_flags(const uint64_t* __restrict__ zm, uint32_t nz, uint32_t* __restrict__ flag) {
    const uint32_t z = blockIdx.x * 256u + threadIdx.x;
    if (z < nz) flag[z] = zm[z] != 0ull ? 1u : 0u;
}

// The light S tuples (joined key, not heavy) by (z tile, key): key
// t * ny + y of tuple i in keyL[i] (~0u for the others), its z ID in zidL[i],
// counted in cnt; then placed by cursors (a row is a set: no order kept).
// Tile-major rows keep one tile's rows, columns and z info together, for the
// light tasks that run tile by tile.
__global__ void k_gc_light_count(const int2* __restrict__ S, uint64_t nS, const int32_t* __restrict__ slotS,
                                 const int32_t* __restrict__ slotSz, const uint64_t* __restrict__ ex_y,
                                 const uint64_t* __restrict__ ex_z, int32_t zlo, const int32_t* __restrict__ hid,
                                 uint32_t ny, uint32_t tile_size, uint32_t* __restrict__ cnt,
                                 uint32_t* __restrict__ keyL, uint32_t* __restrict__ zidL) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= nS) return;
    uint32_t key = 0xFFFFFFFFu;
    const int32_t s = slotS[i];
    if (s >= 0) {
        const uint32_t y = (uint32_t)(ex_y[s] >> 32);
        if (hid[y] < 0) {
            const uint32_t zid = ex_z != nullptr ? (uint32_t)(ex_z[slotSz[i]] >> 32) : (uint32_t)S[i].x - (uint32_t)zlo;
            key = (zid / tile_size) * ny + y;
            zidL[i] = zid;
            atomicAdd(&cnt[key], 1u);
        }
    }
    keyL[i] = key;
}

__global__ void k_gc_light_scatter(const uint32_t* __restrict__ keyL, const uint32_t* __restrict__ zidL, uint64_t nS,
                                   uint32_t* __restrict__ cursor, int32_t* __restrict__ colL) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= nS) return;
    const uint32_t key = keyL[i];
    if (key == 0xFFFFFFFFu) return;
    colL[atomicAdd(&cursor[key], 1u)] = (int32_t)zidL[i];
}

// zinfo[z] = (zm[z], g(z value)) for every z ID: one 16-byte read per light group
__global__ void k_gc_zinfo(const uint64_t* __restrict__ zm, uint32_t nz, const int32_t* __restrict__ id2z, int32_t zlo,
                           uint4* __restrict__ zinfo) {
    const uint32_t z = blockIdx.x * 256u + threadIdx.x;
    if (z >= nz) return;
    const uint64_t m = zm[z];
    const uint64_t g = zfactor(id2z != nullptr ? id2z[z] : (int32_t)((uint32_t)zlo + z));
    zinfo[z] = make_uint4((uint32_t)m, (uint32_t)(m >> 32), (uint32_t)g, (uint32_t)(g >> 32));
}

// compact list of the flagged IDs (pos: exclusive scan of flag)
__global__ void k_gc_list(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, uint32_t n,
                          uint32_t* __restrict__ list) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n && flag[i]) list[pos[i]] = i;
}

// 64 bits -> 64 bytes of 0 / 1 (16-byte stores)
// bit i of m -> byte i of dst: `one` (1, or 0xFF for the int8 -1) where set
__device__ __forceinline__ void bits_to_bytes(uint64_t m, int8_t* dst, uint32_t one = 1u) {
#pragma unroll
    for (uint32_t q = 0; q < 4; q++) {
        uint32_t w[4];
#pragma unroll
        for (uint32_t d = 0; d < 4; d++) {
            const uint32_t nib = (uint32_t)(m >> (16u * q + 4u * d)) & 15u;
            // bit i of the nibble -> byte i
            w[d] = ((nib * 0x00204081u) & 0x01010101u) * one;
        }
        *reinterpret_cast<uint4*>(dst + 16u * q) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Box rows r < Mp: A8[r] = bytes of hm(list[r]), xk[r] = (k1, k2) of its
// x value; padding rows (r >= Mh) zero.
__global__ void k_gc_build_rows(const uint32_t* __restrict__ rlist, uint32_t Mh, uint32_t Mp,
                                const uint64_t* __restrict__ hm, const int32_t* __restrict__ id2x,
                                int8_t* __restrict__ A8, uint64_t* __restrict__ xk) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= Mp) return;
    uint64_t m = 0, k1 = 0, k2 = 0;
    if (r < Mh) {
        const uint32_t x = rlist[r];
        m = hm[x];
        const int32_t xv = id2x[x];
        k1 = xfactor1(xv);
        k2 = xfactor2(xv);
    }
    bits_to_bytes(m, A8 + (size_t)r * kGcK);
    xk[2u * r] = k1;
    xk[2u * r + 1u] = k2;
}

// Box columns n < Np (Np: Nh rounded up to 64): B8[n] = bytes of
// zm(list[n]) as int8 -1 (so the product gives -c_d, whose sign bit flags
// c_d > 0), gz[n] = g(z value); padding columns zero.
__global__ void k_gc_build_cols(const uint32_t* __restrict__ clist, uint32_t Nh, uint32_t Np,
                                const uint64_t* __restrict__ zm, const int32_t* __restrict__ id2z, int32_t zlo,
                                int8_t* __restrict__ B8, uint64_t* __restrict__ gz) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= Np) return;
    uint64_t m = 0, g = 0;
    if (n < Nh) {
        const uint32_t z = clist[n];
        m = zm[z];
        g = zfactor(id2z != nullptr ? id2z[z] : (int32_t)((uint32_t)zlo + z));
    }
    bits_to_bytes(m, B8 + (size_t)n * kGcK, 0xFFu);
    gz[n] = g;
}

// The reduction's A operand per 64-column step s, in lane order: lane l of
// step s holds 16 bytes, byte j = chunk (l % 16) of g(z) for the step's
// column z = 64 s + 16 (j / 4) + 4 (l / 16) + (j % 4) (the k index of byte j
// of a packed count tile, see k_gc_dense); chunk 10 is 1, chunks 11-15 0.
__global__ void k_gc_build_gchunks(const uint64_t* __restrict__ gz, uint32_t nsteps, uint4* __restrict__ G) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;  // step * 64 + lane
    if (i >= nsteps * 64u) return;
    const uint32_t s = i >> 6, l = i & 63u, b = l & 15u, q = l >> 4;
    uint32_t w[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) {
        uint32_t v = 0;
#pragma unroll
        for (uint32_t e = 0; e < 4; e++) v |= key_chunk(gz[64u * s + 16u * u + 4u * q + e], b) << (8u * e);
        w[u] = v;
    }
    G[i] = make_uint4(w[0], w[1], w[2], w[3]);
}

// The dense part's fingerprint terms.  One workgroup = 4 wavefronts = one
// task: box rows [256 xb, 256 xb + 256) (64 per wavefront, their operand
// bytes held in 16 VGPRs) times box columns [cols * zc, cols * zc + cols),
// 64 columns per step; the next step's column operands are loaded while the
// current step computes.  Per step and 16-row group: four product MFMAs give
// the negated transposed count tiles -C^T[z][x] (z = 16 u + 4 q + e in
// lane (x, q); B8 holds -1 bytes), whose sign bits become flag bytes (-1
// where count > 0) in registers, then a reduction MFMA adds E1 += G * flags
// (the count terms
// sum2 and csum are sums over the join paths, formed from the key sums
// instead).  Tasks are
// numbered column-chunk major and dealt to the 8 XCDs in contiguous runs (a
// block goes to XCD blockIdx % 8), so an XCD's workgroups share a column
// chunk's operands in its L2.  Padding: A8 / xk cover Mp = 256 n_xb rows,
// B8 / G cover Np columns (a multiple of 64), zero (G: chunk 10) past the box.
__global__ __launch_bounds__(256, 2) void k_gc_dense(const int8_t* __restrict__ A8, const uint64_t* __restrict__ xk,
                                                      uint32_t n_xb, const int8_t* __restrict__ B8,
                                                      const uint4* __restrict__ G, uint32_t Np, uint32_t cols,
                                                      uint32_t n_tasks, uint32_t per_xcd, Scalars* sc) {
    __shared__ uint64_t red[4][2];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t q = lane >> 4, li = lane & 15u;
    const uint32_t task = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
    if (task >= n_tasks) return;  // workgroup-uniform, before any barrier
    const uint32_t zc = task / n_xb, xb = task - zc * n_xb;
    const uint32_t x0 = xb * kGcRowsB + wid * kGcRowsW;

    // row operands: group gx, lane (li, q) = row x0 + 16 gx + li, keys 16 q .. 16 q + 15
    v4i xa[4];
#pragma unroll
    for (uint32_t gx = 0; gx < 4; gx++)
        xa[gx] = *reinterpret_cast<const v4i*>(A8 + (size_t)(x0 + 16u * gx + li) * kGcK + 16u * q);
    v4i e1[4];
#pragma unroll
    for (uint32_t gx = 0; gx < 4; gx++) e1[gx] = v4i{0, 0, 0, 0};
    const uint32_t z0 = zc * cols;
    const uint32_t z1 = min(z0 + cols, Np);
    // column operands of a step: tile u, lane (li, q) = column zs + 16 u + li,
    // keys 16 q ..  Buffer loads over the chunk's columns: a scalar step
    // offset plus a lane offset fixed for the task, no per-load address math.
    const __amdgpu_buffer_rsrc_t rb8 = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(B8 + (size_t)z0 * kGcK), (short)0, (int)((z1 - z0) * kGcK), kRawBufferWord3);
    const __amdgpu_buffer_rsrc_t rg = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(G + (size_t)(z0 / kGcColsStep) * 64u), (short)0, (int)((z1 - z0) / kGcColsStep * 1024u),
        kRawBufferWord3);
    const int vb = (int)(li * kGcK + 16u * q), vg = (int)(lane * 16u);
    // The step's column operands are read from L2 once per workgroup and
    // shared through LDS (two buffers): thread (u = wid, lane) fetches tile
    // u's bytes for lane position `lane`, wavefront 0 also the reduction
    // operand; after one barrier per step every wavefront reads its lanes'
    // 16-byte pieces.
    __shared__ v4i sB[2][4][64];
    __shared__ v4i sG[2][64];
    const uint32_t nsteps = (z1 - z0) / kGcColsStep;
    v4i nb, ng;
    auto fetch = [&](uint32_t s) {
        const int so = (int)(s * kGcColsStep * kGcK);
        nb = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rb8, vb + (int)(wid * 16u * kGcK), so, 0));
        if (wid == 0) ng = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rg, vg, (int)(s * 1024u), 0));
    };
    auto stage = [&](uint32_t k) {
        sB[k][wid][lane] = nb;
        if (wid == 0) sG[k][lane] = ng;
    };
    auto compute = [&](const v4i* zf, const v4i& gf) {
#pragma unroll
        for (uint32_t gx = 0; gx < 4; gx++) {
            v4i c[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; u++)
                c[u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(zf[u], xa[gx], v4i{0, 0, 0, 0}, 0, 0, 0);
            v4i fl;
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
                // c = -c_d: flag byte 0xFF (int8 -1) where c_d > 0, from the
                // sign bits (v_perm_b32 selector 9 replicates bit 31 of the
                // second operand, 11 of the first, 0x0C gives zero)
                const uint32_t lo = __builtin_amdgcn_perm((uint32_t)c[u][1], (uint32_t)c[u][0], 0x0C0C0B09u);
                const uint32_t hi = __builtin_amdgcn_perm((uint32_t)c[u][3], (uint32_t)c[u][2], 0x0B090C0Cu);
                fl[u] = (int)(lo | hi);
            }
            e1[gx] = __builtin_amdgcn_mfma_i32_16x16x64_i8(gf, fl, e1[gx], 0, 0, 0);
        }
    };
    fetch(0);
    stage(0);
    __syncthreads();
    for (uint32_t s = 0; s < nsteps; s++) {
        const uint32_t k = s & 1u;
        if (s + 1u < nsteps) fetch(s + 1u);  // workgroup-uniform
        v4i zf[4], gf;
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) zf[u] = sB[k][u][lane];
        gf = sG[k][lane];
        compute(zf, gf);
        if (s + 1u < nsteps) stage(k ^ 1u);  // buffer k ^ 1 was last read before the previous barrier
        __syncthreads();
    }
    // lane (li, q) of group gx: row x0 + 16 gx + li, chunk rows b = 4 q + e
    uint64_t S1 = 0, CNT = 0;
#pragma unroll
    for (uint32_t gx = 0; gx < 4; gx++) {
        const uint64_t k1 = xk[2u * (x0 + 16u * gx + li)];
#pragma unroll
        for (uint32_t e = 0; e < 4; e++) {
            const uint32_t b = 4u * q + e;
            // e1 = minus a sum below 2^21
            const uint64_t v1 = (uint64_t)(uint32_t)(-e1[gx][e]);
            if (b < 10u) S1 += k1 * (v1 << (7u * b));
            else if (b == 10u) CNT += v1;
        }
    }
    S1 = wave_reduce_sum(S1);
    CNT = wave_reduce_sum(CNT);
    if (lane == 0) {
        red[wid][0] = S1;
        red[wid][1] = CNT;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const uint64_t v = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
        unsigned long long* dst = (unsigned long long*)(threadIdx.x == 0 ? &sc->fp_sum : &sc->out_n);
        if (v) atomicAdd(dst, (unsigned long long)v);
    }
}

// ---- light part: one (x, z-tile) task per wavefront ----
// The task's light join paths (its R row's entries without the heavy flag,
// each key's row in tile t of the light S CSR) are counted per z ID in
// a private LDS hash table of kLSlots slots.  A task of at most kLCap paths
// runs in one pass.  A longer one runs in passes over z ranges whose width
// is guessed from its path count (7 kLCap / 8 paths per pass if spread
// evenly: every pass enumerates all the task's paths again, so fewer passes
// beat the retries they cause); a pass whose groups pass kLCap stops, clears
// its slots and is retried at half the width, so a pass never lists more than kLCap + 64
// groups (the table stays under two thirds full) and a width of kLCap IDs
// always fits.  A row of at most 64 R entries keeps its segment table in
// LDS from the first sweep to the last pass.
constexpr uint32_t kLBits = 9;
constexpr uint32_t kLSlots = 1u << kLBits;
constexpr uint32_t kLCap = kLSlots / 2;
constexpr uint32_t kLChunk = 16;    // tasks per segment table (one lane each, and their row ends)
constexpr uint32_t kLClaimMax = 128;  // tasks per claim of the work counter, at most
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct LightLds {
    uint32_t keys[kLSlots];
    uint32_t touched[kLCap + 64];
    uint32_t sb[64];
    uint32_t pre[64];
};

__device__ __forceinline__ void lds_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// z's slot: one compare-and-swap per probe claims an empty slot or returns
// the key already there (z itself ends the probe too), so the probe loop has
// no inner branch; `fresh` tells the one lane that claimed it.
__device__ __forceinline__ uint32_t light_slot(uint32_t* keys, uint32_t z, bool& fresh) {
    uint32_t s = (z * 0x9E3779B1u) >> (32u - kLBits);
    for (;;) {
        const uint32_t k = atomicCAS(&keys[s], kNone, z);
        if (k == kNone || k == z) {
            fresh = k == kNone;
            return s;
        }
        s = (s + 1u) & (kLSlots - 1u);
    }
}

// Segment (light S start, path count) of R entry i in tile t: empty for a
// heavy entry (bit 31; a heavy key has no light row either).
__device__ __forceinline__ void light_segment(const int32_t* colR, const uint32_t* rowL, uint32_t ny, uint32_t t,
                                              uint32_t i, uint32_t& sb, uint32_t& len) {
    sb = 0;
    len = 0;
    const int32_t craw = colR[i];
    if (craw >= 0) {
        const uint64_t k = (uint64_t)t * ny + (uint32_t)craw;
        sb = rowL[k];
        len = rowL[k + 1] - sb;
    }
}

__global__ __launch_bounds__(256) void k_gc_light(const uint32_t* __restrict__ rowptrR, const int32_t* __restrict__ colR,
                                                  const uint32_t* __restrict__ rowL, const int32_t* __restrict__ colL,
                                                  uint32_t ny, const uint64_t* __restrict__ hm,
                                                  const uint4* __restrict__ zinfo, const int32_t* __restrict__ id2x,
                                                  uint32_t nx, uint32_t ntiles, uint32_t tile_size, uint32_t nz,
                                                  Scalars* sc, uint32_t claim) {
    __shared__ LightLds wl[4];
    __shared__ uint64_t sh[5];
    const uint32_t lane = threadIdx.x & 63u;
    LightLds& W = wl[threadIdx.x >> 6];
    for (uint32_t s = lane; s < kLSlots; s += 64u) {
        W.keys[s] = kNone;
    }
    lds_wave_sync();
    const uint32_t ntask = nx * ntiles;
    uint64_t fsum = 0, npaths = 0, nrows = 0;
    // claims of `claim` tasks (a multiple of kLChunk): the work counter is
    // one address, whose atomics are served one at a time
    uint32_t task = 0, task_end = 0, claim_end = 0;
    // A chunk's tasks in one tile have consecutive x, so their R entries are
    // one range: its row bounds (lane i: rowptrR of the chunk's i-th x) and,
    // for at most 128 entries, their light segments (entry e in lane e & 63 of
    // register e >> 6) are read once per chunk, and a task takes its own by
    // lane permutes instead of a chain of dependent loads.
    uint32_t cfirst = 0, ct = 0, cx = 0, crp = 0, crb0 = 0, csb0 = 0, clen0 = 0, csb1 = 0, clen1 = 0;
    bool ctab = false;  // wave-uniform
    for (;;) {
        if (task == task_end) {
            if (task == claim_end) {
                uint32_t t0 = 0;
                if (lane == 0) t0 = atomicAdd(&sc->task, claim);
                task = (uint32_t)__builtin_amdgcn_readfirstlane((int)t0);
                if (task >= ntask) break;
                claim_end = min(task + claim, ntask);
            }
            task_end = min(task + kLChunk, claim_end);
            const uint32_t tc = task / nx, xc = task - tc * nx, n = task_end - task;
            ctab = (task_end - 1u) / nx == tc;  // then xc + n <= nx
            if (ctab) {
                crp = lane <= n ? rowptrR[xc + lane] : 0u;
                crb0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)crp);
                const uint32_t E = (uint32_t)__builtin_amdgcn_readlane((int)crp, (int)n) - crb0;
                ctab = E <= 128u;
                if (ctab) {
                    csb0 = clen0 = csb1 = clen1 = 0;
                    if (lane < E) light_segment(colR, rowL, ny, tc, crb0 + lane, csb0, clen0);
                    if (lane + 64u < E) light_segment(colR, rowL, ny, tc, crb0 + 64u + lane, csb1, clen1);
                    cfirst = task;
                    ct = tc;
                    cx = xc;
                }
            }
        }
        // tile-major: the waves running at one time share a tile's light rows
        // and z info in L2
        const uint32_t ti = task++;
        uint32_t t, x, rb, re;
        if (ctab) {
            t = ct;
            x = cx + (ti - cfirst);
            rb = (uint32_t)__builtin_amdgcn_readlane((int)crp, (int)(ti - cfirst));
            re = (uint32_t)__builtin_amdgcn_readlane((int)crp, (int)(ti - cfirst + 1u));
        } else {
            t = ti / nx;
            x = ti - t * nx;
            rb = rowptrR[x];
            re = rowptrR[x + 1];
        }
        const bool short_row = re - rb <= 64u;  // wave-uniform
        // the task's light paths; a short row's segment table stays in LDS
        uint32_t P = 0;
        if (short_row) {
            uint32_t sb = 0, len = 0;
            if (ctab) {
                const uint32_t e = rb - crb0 + lane;  // < 128 for the row's lanes
                const int src = (int)((e & 63u) << 2);
                const uint32_t s0 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)csb0);
                const uint32_t s1 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)csb1);
                const uint32_t l0 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)clen0);
                const uint32_t l1 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)clen1);
                if (rb + lane < re) {
                    sb = e < 64u ? s0 : s1;
                    len = e < 64u ? l0 : l1;
                }
            } else if (rb + lane < re) {
                light_segment(colR, rowL, ny, t, rb + lane, sb, len);
            }
            const uint32_t inc = wave_scan_u32(len);
            P = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
            W.sb[lane] = sb;
            W.pre[lane] = inc - len;
            lds_wave_sync();
        } else {
            for (uint32_t c0 = rb; c0 < re; c0 += 64u) {
                uint32_t sb = 0, len = 0;
                if (c0 + lane < re) light_segment(colR, rowL, ny, t, c0 + lane, sb, len);
                P += wave_sum_u32(len);
            }
        }
        if (P == 0) continue;  // wave-uniform
        npaths += lane == 0 ? P : 0u;
        const uint64_t hmx = hm[x];
        const int32_t xv = id2x[x];
        const uint64_t k1 = xfactor1(xv);
        const uint32_t zt0 = t * tile_size, zt1 = min(zt0 + tile_size, nz);
        uint32_t width = zt1 - zt0;
        if (P > kLCap) width = (uint32_t)max<uint64_t>(1, (uint64_t)(zt1 - zt0) * (kLCap * 7 / 8) / P);
        uint32_t zcur = zt0;
        while (zcur < zt1) {
            const uint32_t zhi = zt1 - zcur <= width ? zt1 : zcur + width;
            uint32_t nt = 0;  // listed slots (wave-uniform)
            bool over = false;
            for (uint32_t c0 = rb; c0 < re && !over; c0 += 64u) {
                uint32_t PC = P;
                if (!short_row) {  // this chunk's segment table
                    uint32_t sb = 0, len = 0;
                    if (c0 + lane < re) light_segment(colR, rowL, ny, t, c0 + lane, sb, len);
                    const uint32_t inc = wave_scan_u32(len);
                    PC = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
                    W.sb[lane] = sb;
                    W.pre[lane] = inc - len;
                    lds_wave_sync();
                }
                // search steps: the highest power of two below the table's entries
                const uint32_t nseg = min(64u, re - c0);
                const uint32_t sbit = nseg > 1u ? 1u << (31 - __builtin_clz(nseg - 1u)) : 0u;
                for (uint32_t p0 = 0; p0 < PC; p0 += 64u) {
                    if (nt > kLCap) {  // too many groups for this width: retry narrower
                        over = true;
                        break;
                    }
                    const uint32_t p = p0 + lane;
                    const bool in = p < PC;
                    // the last segment whose path prefix is <= p holds path p
                    // (entries past the table's nseg have prefix PC)
                    uint32_t seg = 0;
                    for (uint32_t step = sbit; step > 0; step >>= 1)
                        if (W.pre[seg + step] <= p) seg += step;
                    const uint32_t z = in ? (uint32_t)colL[W.sb[seg] + (p - W.pre[seg])] : 0u;
                    const bool ok = in && z >= zcur && z < zhi;
                    uint32_t slot = 0;
                    bool fresh = false;
                    if (ok) slot = light_slot(W.keys, z, fresh);
                    const uint64_t m = __ballot(fresh);
                    if (fresh) W.touched[nt + mbcnt64(m)] = slot;
                    nt += (uint32_t)__popcll(m);
                }
                if (!short_row) lds_wave_sync();  // the table is rewritten by the next 64 R entries
            }
            if (over) {
                for (uint32_t j = lane; j < nt; j += 64u) W.keys[W.touched[j]] = kNone;
                width = max(1u, width >> 1);
                lds_wave_sync();
                continue;
            }
            // emission: the groups of light paths only (the others are
            // counted in the box)
            uint64_t S1 = 0;
            uint32_t T = 0;
            for (uint32_t j = lane; j < nt; j += 64u) {
                const uint32_t s = W.touched[j];
                const uint4 zi = zinfo[W.keys[s]];
                const uint64_t zmz = ((uint64_t)zi.y << 32) | zi.x, g = ((uint64_t)zi.w << 32) | zi.z;
                if ((hmx & zmz) == 0ull) {
                    S1 += g;
                    T += 1u;
                }
                W.keys[s] = kNone;
            }
            fsum += k1 * S1;
            nrows += T;
            lds_wave_sync();
            zcur = zhi;
        }
    }
    const uint64_t np = block_reduce_sum<256, uint64_t>(npaths, sh);
    const uint64_t s1 = block_reduce_sum<256, uint64_t>(fsum, sh);
    const uint64_t nr = block_reduce_sum<256, uint64_t>(nrows, sh);
    if (threadIdx.x == 0) {
        if (np) atomicAdd((unsigned long long*)&sc->paths, (unsigned long long)np);
        if (s1) atomicAdd((unsigned long long*)&sc->fp_sum, (unsigned long long)s1);
        if (nr) atomicAdd((unsigned long long*)&sc->out_n, (unsigned long long)nr);
    }
}

inline unsigned blocks256(uint64_t n) { return (unsigned)(n == 0 ? 1 : (n + 255) / 256); }

}  // namespace

namespace dim3h {

uint32_t gc_rows_per_block() { return kGcRowsB; }

void gc_masks(const uint32_t* rowptrR, int32_t* colR, uint32_t nx, const int2* S, uint64_t nS, const int32_t* slotS,
              const int32_t* slotSz, const uint64_t* ex_y, const uint64_t* ex_z, int32_t zlo, const int32_t* hid,
              uint32_t nz, uint64_t* hm, uint32_t* xflag, uint64_t* zm, uint32_t* zflag, uint64_t* dup,
              hipStream_t st) {
    fill_u64(zm, nz, 0, st);
    if (nx)
        hipLaunchKernelGGL(k_gc_row_masks, dim3(blocks256(nx)), dim3(256), 0, st, rowptrR, colR, nx, hid, hm, xflag,
                           (unsigned long long*)dup);
    if (nS)
        hipLaunchKernelGGL(k_gc_col_masks, dim3(blocks256(nS)), dim3(256), 0, st, S, nS, slotS, slotSz, ex_y, ex_z, zlo,
                           hid, zm, (unsigned long long*)dup);
    if (nz) hipLaunchKernelGGL(k_gc_col_flags, dim3(blocks256(nz)), dim3(256), 0, st, zm, nz, zflag);
}

void gc_zinfo(const uint64_t* zm, uint32_t nz, const int32_t* id2z, int32_t zlo, void* zinfo, hipStream_t st) {
    if (nz) hipLaunchKernelGGL(k_gc_zinfo, dim3(blocks256(nz)), dim3(256), 0, st, zm, nz, id2z, zlo, (uint4*)zinfo);
}

void gc_list(const uint32_t* flag, const uint32_t* pos, uint32_t n, uint32_t* list, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_gc_list, dim3(blocks256(n)), dim3(256), 0, st, flag, pos, n, list);
}

uint32_t gc_cols_padded(uint32_t Nh) { return (Nh + kGcColsStep - 1) / kGcColsStep * kGcColsStep; }

void gc_build(const uint32_t* rlist, uint32_t Mh, uint32_t Mp, const uint64_t* hm, const int32_t* id2x,
              const uint32_t* clist, uint32_t Nh, const uint64_t* zm, const int32_t* id2z, int32_t zlo, int8_t* A8,
              uint64_t* xk, int8_t* B8, uint64_t* gz, void* G, hipStream_t st) {
    const uint32_t Np = gc_cols_padded(Nh);
    if (Mp) hipLaunchKernelGGL(k_gc_build_rows, dim3(blocks256(Mp)), dim3(256), 0, st, rlist, Mh, Mp, hm, id2x, A8, xk);
    if (Np) {
        hipLaunchKernelGGL(k_gc_build_cols, dim3(blocks256(Np)), dim3(256), 0, st, clist, Nh, Np, zm, id2z, zlo, B8,
                           gz);
        const uint32_t nsteps = Np / kGcColsStep;
        hipLaunchKernelGGL(k_gc_build_gchunks, dim3(blocks256(64ull * nsteps)), dim3(256), 0, st, gz, nsteps,
                           (uint4*)G);
    }
}

int gc_dense(const int8_t* A8, const uint64_t* xk, uint32_t Mp, const int8_t* B8, const void* G, uint32_t Np,
             uint32_t cols, Scalars* sc, hipStream_t st) {
    if (Mp == 0 || Np == 0) return 0;
    if (Mp % kGcRowsB != 0 || Np % kGcColsStep != 0 || cols == 0 || cols % kGcColsStep != 0)
        return (int)hipErrorInvalidValue;
    const uint32_t n_xb = Mp / kGcRowsB;
    const uint64_t n_zc = ((uint64_t)Np + cols - 1) / cols;
    const uint64_t n_tasks = (uint64_t)n_xb * n_zc;
    if (n_tasks >= (1ull << 31)) return (int)hipErrorInvalidValue;
    const uint32_t per = (uint32_t)((n_tasks + 7) / 8);
    hipLaunchKernelGGL(k_gc_dense, dim3(8u * per), dim3(256), 0, st, A8, xk, n_xb, B8, (const uint4*)G, Np, cols,
                       (uint32_t)n_tasks, per, sc);
    return (int)hipGetLastError();
}

void gc_light_csr_count(const int2* S, uint64_t nS, const int32_t* slotS, const int32_t* slotSz, const uint64_t* ex_y,
                        const uint64_t* ex_z, int32_t zlo, const int32_t* hid, uint32_t ny, uint32_t tile_size,
                        uint32_t* cnt, uint32_t* keyL, uint32_t* zidL, hipStream_t st) {
    if (nS)
        hipLaunchKernelGGL(k_gc_light_count, dim3(blocks256(nS)), dim3(256), 0, st, S, nS, slotS, slotSz, ex_y, ex_z,
                           zlo, hid, ny, tile_size, cnt, keyL, zidL);
}

void gc_light_csr_scatter(const uint32_t* keyL, const uint32_t* zidL, uint64_t nS, uint32_t* cursor, int32_t* colL,
                          hipStream_t st) {
    if (nS) hipLaunchKernelGGL(k_gc_light_scatter, dim3(blocks256(nS)), dim3(256), 0, st, keyL, zidL, nS, cursor, colL);
}

int gc_light(const uint32_t* rowptrR, const int32_t* colR, const uint32_t* rowL, const int32_t* colL, uint32_t ny,
             const uint64_t* hm, const void* zinfo, const int32_t* id2x, uint32_t nx, uint32_t ntiles,
             uint32_t tile_size, uint32_t nz, int cu_count, Scalars* sc, hipStream_t st) {
    if (nx == 0) return 0;
    int occ = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_gc_light, 256, 0);
    if (e != hipSuccess) return (int)e;
    if (occ < 1) occ = 1;
    const uint64_t ntask = (uint64_t)nx * ntiles;
    uint64_t grid = (uint64_t)cu_count * (uint64_t)occ;
    const uint64_t waves = (ntask + kLChunk - 1) / kLChunk;
    if (grid * 4 > waves) grid = (waves + 3) / 4;
    if (grid == 0) grid = 1;
    // about 16 claims per wavefront at least
    uint32_t claim = (uint32_t)std::min<uint64_t>(kLClaimMax, ntask / (grid * 4 * 16) / kLChunk * kLChunk);
    claim = std::max(claim, kLChunk);
    // the counter ends below ntask + one claim per wavefront
    if (ntask + grid * 4 * claim >= (1ull << 32)) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gc_light, dim3((unsigned)grid), dim3(256), 0, st, rowptrR, colR, rowL, colL, ny, hm,
                       (const uint4*)zinfo, id2x, nx, ntiles, tile_size, nz, sc, claim);
    return (int)hipGetLastError();
}

}  // namespace dim3h
