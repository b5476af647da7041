// Dense part of the DIM3 split (subsystems b, c): heavy join keys as one
// 0/1 (or multiplicity) int8 matrix product on the MFMA cores, or as OR-ed
// z bitmaps.
//
// The reference's dense part (Hybrid_solution DENSE, dim3.hpp:894-950) keeps
// one bitvector over y per dense z and tests each x's y list against it (AVX
// testz sweep, or bit probes).  Here the split is by join key instead: keys
// whose R and S degrees are both at or above a planner threshold are heavy;
// every join path through a heavy key goes to
//   C = A * B^T,  A [Mh rows x with a heavy key][Kp] int8 (A[hx][h] = multiplicity of (x, y_h))
//                 Bt[Nh columns z with a heavy key][Kp] int8 (Bt[hz][h] = multiplicity of (z, y_h))
// both K-contiguous, K padded to a multiple of 64, int32 accumulation (exact:
// entries are at most 127 * 127 * Kp < 2^31 for Kp <= 131072).  Epilogues:
//   project: bit rows Dbits[Mh][ceil(Nh / 32)] of C > 0, packed with ballots;
//   COUNT:   int32 counts Cint[Mh][Nh].
// The bit-packed alternative (project only) keeps one z bitmap per heavy key
// and ORs the bitmaps of x's heavy keys into Dbits row hx.
// Columns hz are numbered in z-ID order, so a z tile of the expansion is one
// contiguous hz range; the expansion seeds each row's LDS tile from Dbits /
// Cint before adding the light paths.  Heavy S tuples leave the sparse CSR.
//
// GEMM: 256-thread workgroup, 128 x 128 output tile, K step 64 staged in
// LDS (row stride 80 bytes so the 16-byte fragment reads spread over banks),
// each wavefront a 64 x 64 quadrant = 4 x 4 v_mfma_i32_16x16x64_i8.
// Operand layout of that instruction (checked on gfx950): lane l holds row
// (l % 16) of A and column (l % 16) of B for the 16 k values of group l / 16;
// the result lane l holds column l % 16, rows 4 (l / 16) .. 4 (l / 16) + 3.
#include <algorithm>

#include "dim3_internal.h"

namespace {

using namespace dim3d;
using dim3h::kDenseBins;

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kTM = 128, kTN = 128, kTK = 64;
constexpr int kLdsStride = 80;  // bytes per staged row (64 data + 16 pad)

__global__ __launch_bounds__(256) void k_gemm_i8(const int8_t* __restrict__ A, const int8_t* __restrict__ Bt,
                                                 uint32_t M, uint32_t N, uint32_t Kp, uint32_t* __restrict__ dbits,
                                                 uint32_t dwords, int32_t* __restrict__ cint) {
    __shared__ __align__(16) int8_t sA[kTM * kLdsStride];
    __shared__ __align__(16) int8_t sB[kTN * kLdsStride];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const uint32_t wm = (wid >> 1) * 64, wn = (wid & 1) * 64;  // wavefront quadrant

    v4i acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = v4i{0, 0, 0, 0};

    for (uint32_t k0 = 0; k0 < Kp; k0 += kTK) {
        // stage 128 rows x 64 bytes of A and of Bt: 512 16-byte chunks each
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const uint32_t chunk = tid + (uint32_t)c * 256u;  // 0..511
            const uint32_t r = chunk >> 2, q = chunk & 3u;
            v4i va = v4i{0, 0, 0, 0}, vb = v4i{0, 0, 0, 0};
            if (m0 + r < M) va = *reinterpret_cast<const v4i*>(A + (size_t)(m0 + r) * Kp + k0 + q * 16u);
            if (n0 + r < N) vb = *reinterpret_cast<const v4i*>(Bt + (size_t)(n0 + r) * Kp + k0 + q * 16u);
            *reinterpret_cast<v4i*>(sA + r * kLdsStride + q * 16u) = va;
            *reinterpret_cast<v4i*>(sB + r * kLdsStride + q * 16u) = vb;
        }
        __syncthreads();
        v4i fa[4], fb[4];
        const uint32_t fr = lane & 15u, fk = (lane >> 4) * 16u;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            fa[i] = *reinterpret_cast<const v4i*>(sA + (wm + i * 16 + fr) * kLdsStride + fk);
            fb[i] = *reinterpret_cast<const v4i*>(sB + (wn + i * 16 + fr) * kLdsStride + fk);
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    // result lane l: column (l % 16), rows 4 (l / 16) + e
    const uint32_t col_in = lane & 15u, rgrp = (lane >> 4) * 4u;
    if (cint != nullptr) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t col = n0 + wn + j * 16 + col_in;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const uint32_t row = m0 + wm + i * 16 + rgrp + e;
                    if (row < M && col < N) cint[(size_t)row * N + col] = acc[i][j][e];
                }
            }
    }
    if (dbits != nullptr) {
        // a ballot over one (i, j, e) gives rows 4g + e (g = lane group) x 16
        // columns; two neighbouring j tiles make a 32-column word
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int jp = 0; jp < 2; jp++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const uint64_t b0 = __ballot(acc[i][2 * jp][e] > 0);
                    const uint64_t b1 = __ballot(acc[i][2 * jp + 1][e] > 0);
                    if (lane < 4u) {
                        const uint32_t g = lane;
                        const uint32_t word = (uint32_t)((b0 >> (16 * g)) & 0xFFFFull) |
                                              ((uint32_t)((b1 >> (16 * g)) & 0xFFFFull) << 16);
                        const uint32_t row = m0 + wm + i * 16 + 4 * g + e;
                        const uint32_t wcol = (n0 + wn + jp * 32) >> 5;
                        if (row < M && wcol < dwords) dbits[(size_t)row * dwords + wcol] = word;
                    }
                }
    }
}

// Large-tile product for M, N >= 256: 256 threads, each wavefront a
// (16 TI) x (16 TJ) quadrant of TI x TJ MFMA tiles, workgroup tile
// (32 TI) x (32 TJ); every 16-byte fragment read from LDS feeds TJ (A) or
// TI (B) MFMAs.  K steps of 64 are double-buffered in LDS with the next
// step's global loads issued before the current step's MFMAs (one barrier
// per step).  Tiles are numbered so that each XCD (workgroups are dealt to
// the 8 XCDs round robin) works through a contiguous run of 8-row tile
// groups, sharing A and B panels in its own L2.
template <int TI, int TJ>
__global__ __launch_bounds__(256, 2) void k_gemm_i8_big(const int8_t* __restrict__ A, const int8_t* __restrict__ Bt,
                                                     uint32_t M, uint32_t N, uint32_t Kp,
                                                     uint32_t* __restrict__ dbits, uint32_t dwords,
                                                     int32_t* __restrict__ cint) {
    constexpr uint32_t BM = 32 * TI, BN = 32 * TJ;   // workgroup tile
    constexpr uint32_t CA = BM * 4 / 256, CB = BN * 4 / 256;  // 16-byte chunks per thread per step
    __shared__ __align__(16) int8_t sA[2][BM * kLdsStride];
    __shared__ __align__(16) int8_t sB[2][BN * kLdsStride];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t gm = (M + BM - 1) / BM, gn = (N + BN - 1) / BN, total = gm * gn;
    const uint32_t per_xcd = (total + 7u) / 8u;
    const uint32_t seq = (blockIdx.x % 8u) * per_xcd + blockIdx.x / 8u;
    if (seq >= total) return;
    const uint32_t grp = seq / (8u * gn), r = seq - grp * 8u * gn;
    const uint32_t gh = min(8u, gm - grp * 8u);
    const uint32_t m0 = (grp * 8u + r % gh) * BM, n0 = (r / gh) * BN;
    const uint32_t wm = (wid >> 1) * (16u * TI), wn = (wid & 1u) * (16u * TJ);

    v4i acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++) acc[i][j] = v4i{0, 0, 0, 0};

    v4i pa[CA], pb[CB];
    auto load = [&](uint32_t k0) {
#pragma unroll
        for (uint32_t c = 0; c < CA; c++) {
            const uint32_t chunk = tid + c * 256u;
            const uint32_t row = chunk >> 2, q = chunk & 3u;
            pa[c] = m0 + row < M ? *reinterpret_cast<const v4i*>(A + (size_t)(m0 + row) * Kp + k0 + q * 16u)
                                 : v4i{0, 0, 0, 0};
        }
#pragma unroll
        for (uint32_t c = 0; c < CB; c++) {
            const uint32_t chunk = tid + c * 256u;
            const uint32_t row = chunk >> 2, q = chunk & 3u;
            pb[c] = n0 + row < N ? *reinterpret_cast<const v4i*>(Bt + (size_t)(n0 + row) * Kp + k0 + q * 16u)
                                 : v4i{0, 0, 0, 0};
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (uint32_t c = 0; c < CA; c++) {
            const uint32_t chunk = tid + c * 256u;
            *reinterpret_cast<v4i*>(&sA[buf][(chunk >> 2) * kLdsStride + (chunk & 3u) * 16u]) = pa[c];
        }
#pragma unroll
        for (uint32_t c = 0; c < CB; c++) {
            const uint32_t chunk = tid + c * 256u;
            *reinterpret_cast<v4i*>(&sB[buf][(chunk >> 2) * kLdsStride + (chunk & 3u) * 16u]) = pb[c];
        }
    };
    load(0);
    store(0);
    __syncthreads();
    const uint32_t fr = lane & 15u, fk = (lane >> 4) * 16u;
    int buf = 0;
    for (uint32_t k0 = 0; k0 < Kp; k0 += kTK) {
        const bool more = k0 + kTK < Kp;
        if (more) load(k0 + kTK);
        // all fragments of the step first, so the MFMAs issue back to back
        v4i fb[TJ], fa[TI];
#pragma unroll
        for (int j = 0; j < TJ; j++)
            fb[j] = *reinterpret_cast<const v4i*>(&sB[buf][(wn + j * 16 + fr) * kLdsStride + fk]);
#pragma unroll
        for (int i = 0; i < TI; i++)
            fa[i] = *reinterpret_cast<const v4i*>(&sA[buf][(wm + i * 16 + fr) * kLdsStride + fk]);
#pragma unroll
        for (int i = 0; i < TI; i++)
#pragma unroll
            for (int j = 0; j < TJ; j++)
                acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

    const uint32_t col_in = lane & 15u, rgrp = (lane >> 4) * 4u;
    if (cint != nullptr) {
#pragma unroll
        for (int i = 0; i < TI; i++)
#pragma unroll
            for (int j = 0; j < TJ; j++) {
                const uint32_t col = n0 + wn + j * 16 + col_in;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const uint32_t row = m0 + wm + i * 16 + rgrp + e;
                    if (row < M && col < N) cint[(size_t)row * N + col] = acc[i][j][e];
                }
            }
    }
    if (dbits != nullptr) {
        // ballots over (i, j, e) are wavefront-uniform: for each 16-row tile
        // i, lane L < 16 assembles the words of row 16 i + L (g = L / 4,
        // e = L % 4) and stores them, one instruction per 32-column word
        const uint32_t g = (lane >> 2) & 3u, ee = lane & 3u;
#pragma unroll
        for (int i = 0; i < TI; i++) {
            const uint32_t row = m0 + wm + i * 16 + lane;
#pragma unroll
            for (int jp = 0; jp < TJ / 2; jp++) {
                uint64_t b0[4], b1[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    b0[e] = __ballot(acc[i][2 * jp][e] > 0);
                    b1[e] = __ballot(acc[i][2 * jp + 1][e] > 0);
                }
                const uint64_t lo = ee == 0 ? b0[0] : ee == 1 ? b0[1] : ee == 2 ? b0[2] : b0[3];
                const uint64_t hi = ee == 0 ? b1[0] : ee == 1 ? b1[1] : ee == 2 ? b1[2] : b1[3];
                const uint32_t word = (uint32_t)((lo >> (16 * g)) & 0xFFFFull) |
                                      ((uint32_t)((hi >> (16 * g)) & 0xFFFFull) << 16);
                // lane L holds row 4 g + e = L of the tile
                const uint32_t wcol = (n0 + wn + jp * 32) >> 5;
                if (lane < 16u && row < M && wcol < dwords) dbits[(size_t)row * dwords + wcol] = word;
            }
        }
    }
}

// Bit-packed product: Dbits row r = OR of the z bitmaps of the heavy keys of
// x = hx2x[r].  One workgroup per row; each thread owns words tid + 256 j of
// a 2048-word chunk of the row.
constexpr uint32_t kOrWords = 8;
__global__ __launch_bounds__(256) void k_bitor(const uint32_t* __restrict__ rowptrR, const int32_t* __restrict__ colR,
                                               const int32_t* __restrict__ hid, const uint32_t* __restrict__ hx2x,
                                               uint32_t Mh, const uint32_t* __restrict__ bbits, uint32_t dwords,
                                               uint32_t* __restrict__ dbits) {
    const uint32_t r = blockIdx.x;
    if (r >= Mh) return;
    const uint32_t x = hx2x[r];
    const uint32_t rb = rowptrR[x], re = rowptrR[x + 1];
    for (uint32_t w0 = 0; w0 < dwords; w0 += 256u * kOrWords) {
        uint32_t acc[kOrWords];
#pragma unroll
        for (uint32_t j = 0; j < kOrWords; j++) acc[j] = 0;
        for (uint32_t i = rb; i < re; i++) {
            const int32_t c = colR[i];
            if (c >= 0) continue;  // light key
            const int32_t h = hid[(uint32_t)c & 0x7FFFFFFFu];
            const uint32_t* src = bbits + (size_t)h * dwords;
#pragma unroll
            for (uint32_t j = 0; j < kOrWords; j++) {
                const uint32_t w = w0 + threadIdx.x + 256u * j;
                if (w < dwords) acc[j] |= src[w];
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kOrWords; j++) {
            const uint32_t w = w0 + threadIdx.x + 256u * j;
            if (w < dwords) dbits[(size_t)r * dwords + w] = acc[j];
        }
    }
}

// ---------------------------------------------------------------------------
// planning: heaviness w(y) = min(degR(y), degS(y)); x and z carry the largest
// w over their tuples.  Bins: w < 16 exact, then four bins per octave.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t key_w(const uint32_t* degR, const uint32_t* degS, uint32_t s) {
    return min(degR[s], degS[s]);
}

// hsx / hsz (nullable): the hot x / z values' slots, tallied per workgroup.
__global__ __launch_bounds__(256) void k_wmax_x(const uint32_t* __restrict__ slotR,
                                                const int32_t* __restrict__ slotRx, uint64_t nR,
                                                const uint32_t* __restrict__ degR, const uint32_t* __restrict__ degS,
                                                uint32_t* xw, const int32_t* __restrict__ hsx) {
    __shared__ uint32_t tally[kTallyKeys], tslot[kTallyKeys];
    tally_init<256>(tally, tslot, kTallyKeys);
    for (uint64_t i0 = (uint64_t)blockIdx.x * 256; i0 < nR; i0 += (uint64_t)gridDim.x * 256) {
        const uint64_t i = i0 + threadIdx.x;
        const int32_t sx = i < nR ? slotRx[i] : -1;
        const uint32_t w = sx >= 0 ? key_w(degR, degS, slotR[i]) : 0u;
        const int h = sx >= 0 ? hot_index(hsx, sx) : -1;
        if (h >= 0) {
            tslot[h] = (uint32_t)sx;
            atomicMax(&tally[h], w);
        }
        wave_max_update(xw, (uint32_t)sx, w, sx >= 0 && h < 0);
    }
    tally_flush<256, true>(xw, tally, tslot, kTallyKeys);
}
__global__ __launch_bounds__(256) void k_wmax_z(const int2* __restrict__ S, const int32_t* __restrict__ slotS,
                                                const int32_t* __restrict__ slotSz, int32_t zlo, uint64_t nS,
                                                const uint32_t* __restrict__ degR, const uint32_t* __restrict__ degS,
                                                uint32_t* zw, const int32_t* __restrict__ hsz) {
    __shared__ uint32_t tally[kTallyKeys], tslot[kTallyKeys];
    tally_init<256>(tally, tslot, kTallyKeys);
    for (uint64_t i0 = (uint64_t)blockIdx.x * 256; i0 < nS; i0 += (uint64_t)gridDim.x * 256) {
        const uint64_t i = i0 + threadIdx.x;
        const int32_t sy = i < nS ? slotS[i] : -1;
        const uint32_t zs = sy < 0 ? 0u : (slotSz == nullptr ? (uint32_t)S[i].x - (uint32_t)zlo : (uint32_t)slotSz[i]);
        const uint32_t w = sy >= 0 ? key_w(degR, degS, (uint32_t)sy) : 0u;
        const int h = sy >= 0 ? hot_index(hsz, (int32_t)zs) : -1;
        if (h >= 0) {
            tslot[h] = zs;
            atomicMax(&tally[h], w);
        }
        wave_max_update(zw, zs, w, sy >= 0 && h < 0);
    }
    tally_flush<256, true>(zw, tally, tslot, kTallyKeys);
}

// hist layout: [6][kDenseBins] u64 = keys, paths, sum degR, sum degS, x rows, z columns
__global__ __launch_bounds__(256) void k_dense_hist(const uint32_t* __restrict__ degR,
                                                    const uint32_t* __restrict__ degS, uint32_t ny1,
                                                    const uint32_t* __restrict__ xw, uint32_t nx1,
                                                    const uint32_t* __restrict__ zw, uint32_t nz1,
                                                    unsigned long long* hist) {
    __shared__ unsigned long long lh[6 * kDenseBins];
    for (uint32_t i = threadIdx.x; i < 6 * kDenseBins; i += 256) lh[i] = 0;
    __syncthreads();
    const uint32_t n = max(ny1, max(nx1, nz1));
    for (uint32_t s = blockIdx.x * 256 + threadIdx.x; s < n; s += gridDim.x * 256) {
        if (s < ny1) {
            const uint32_t r = degR[s], q = degS[s];
            if (r != 0 && q != 0) {
                const uint32_t b = dim3h::dense_bin(min(r, q));
                atomicAdd(&lh[b], 1ull);
                atomicAdd(&lh[kDenseBins + b], (unsigned long long)r * q);
                atomicAdd(&lh[2 * kDenseBins + b], (unsigned long long)r);
                atomicAdd(&lh[3 * kDenseBins + b], (unsigned long long)q);
            }
        }
        if (s < nx1 && xw[s] != 0) atomicAdd(&lh[4 * kDenseBins + dim3h::dense_bin(xw[s])], 1ull);
        if (s < nz1 && zw[s] != 0) atomicAdd(&lh[5 * kDenseBins + dim3h::dense_bin(zw[s])], 1ull);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 6 * kDenseBins; i += 256)
        if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

// heavy index of every joined y ID (-1 = light) and its inverse
__global__ void k_mark_hid(const uint32_t* __restrict__ degR, const uint32_t* __restrict__ degS, uint32_t ny1,
                           const uint64_t* __restrict__ ex_y, uint32_t bmin, int32_t* __restrict__ hid,
                           uint32_t* __restrict__ h2y, uint32_t* kcount) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ny1) return;
    const uint32_t r = degR[s], q = degS[s];
    if (r == 0 || q == 0) return;
    const uint32_t yid = (uint32_t)(ex_y[s] >> 32);
    int32_t h = -1;
    if (dim3h::dense_bin(min(r, q)) >= bmin) {
        h = (int32_t)atomicAdd(kcount, 1u);
        h2y[h] = yid;
    }
    hid[yid] = h;
}

// row index hx of every x ID (-1 = no heavy key) and its inverse
// (hdx != nullptr: a row needs at least thx heavy entries, the dense core)
__global__ void k_mark_hx(const uint32_t* __restrict__ degX, const uint64_t* __restrict__ ex_x, uint32_t nx1,
                          const uint32_t* __restrict__ xw, uint32_t bmin, const uint32_t* __restrict__ hdx,
                          uint32_t thx, int32_t* __restrict__ hx, uint32_t* __restrict__ hx2x, uint32_t* mcount) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nx1 || degX[s] == 0) return;
    const uint32_t xid = (uint32_t)(ex_x[s] >> 32);
    int32_t v = -1;
    const bool dense = hdx != nullptr ? hdx[s] >= thx : (xw[s] != 0 && dim3h::dense_bin(xw[s]) >= bmin);
    if (dense) {
        v = (int32_t)atomicAdd(mcount, 1u);
        hx2x[v] = xid;
    }
    hx[xid] = v;
}

// column flag of every z ID (hash mode: occupied slots; identity: every ID)
// (bmin == 0: every z ID is a column, so the columns of a tile are its z IDs)
// (hdz != nullptr: a column needs at least thz heavy tuples, the dense core)
__global__ void k_mark_zflag(const uint32_t* __restrict__ degZ, const uint64_t* __restrict__ ex_z, uint32_t nz1,
                             const uint32_t* __restrict__ zw, uint32_t bmin, const uint32_t* __restrict__ hdz,
                             uint32_t thz, uint32_t* __restrict__ zflag) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nz1) return;
    uint32_t zid = s;
    if (ex_z != nullptr) {
        if (degZ[s] == 0) return;
        zid = (uint32_t)(ex_z[s] >> 32);
    }
    bool dense;
    if (bmin == 0u) dense = true;
    else if (hdz != nullptr) dense = hdz[s] >= thz;
    else dense = zw[s] != 0 && dim3h::dense_bin(zw[s]) >= bmin;
    zflag[zid] = dense ? 1u : 0u;
}

// Heavy entries per x slot and heavy tuples per z slot (dense-core selection).
__global__ __launch_bounds__(256) void k_hdeg_x(const uint32_t* __restrict__ slotR,
                                                const int32_t* __restrict__ slotRx, uint64_t nR,
                                                const uint64_t* __restrict__ ex_y, const int32_t* __restrict__ hid,
                                                uint32_t* hdx, const int32_t* __restrict__ hsx) {
    __shared__ uint32_t tally[kTallyKeys], tslot[kTallyKeys];
    tally_init<256>(tally, tslot, kTallyKeys);
    for (uint64_t i0 = (uint64_t)blockIdx.x * 256; i0 < nR; i0 += (uint64_t)gridDim.x * 256) {
        const uint64_t i = i0 + threadIdx.x;
        const int32_t sx = i < nR ? slotRx[i] : -1;
        const bool heavy = sx >= 0 && hid[(uint32_t)(ex_y[slotR[i]] >> 32)] >= 0;
        const int h = heavy ? hot_index(hsx, sx) : -1;
        if (h >= 0) {
            tslot[h] = (uint32_t)sx;
            atomicAdd(&tally[h], 1u);
        }
        wave_count_add(hdx, (uint32_t)sx, heavy && h < 0);
    }
    tally_flush<256>(hdx, tally, tslot, kTallyKeys);
}
__global__ __launch_bounds__(256) void k_hdeg_z(const int2* __restrict__ S, const int32_t* __restrict__ slotS,
                                                const int32_t* __restrict__ slotSz, int32_t zlo, uint64_t nS,
                                                const uint64_t* __restrict__ ex_y, const int32_t* __restrict__ hid,
                                                uint32_t* hdz, const int32_t* __restrict__ hsz) {
    __shared__ uint32_t tally[kTallyKeys], tslot[kTallyKeys];
    tally_init<256>(tally, tslot, kTallyKeys);
    for (uint64_t i0 = (uint64_t)blockIdx.x * 256; i0 < nS; i0 += (uint64_t)gridDim.x * 256) {
        const uint64_t i = i0 + threadIdx.x;
        const int32_t sy = i < nS ? slotS[i] : -1;
        const bool heavy = sy >= 0 && hid[(uint32_t)(ex_y[sy] >> 32)] >= 0;
        const uint32_t zs = heavy ? (slotSz == nullptr ? (uint32_t)S[i].x - (uint32_t)zlo : (uint32_t)slotSz[i]) : 0u;
        const int h = heavy ? hot_index(hsz, (int32_t)zs) : -1;
        if (h >= 0) {
            tslot[h] = zs;
            atomicAdd(&tally[h], 1u);
        }
        wave_count_add(hdz, zs, heavy && h < 0);
    }
    tally_flush<256>(hdz, tally, tslot, kTallyKeys);
}

// The same counts (KIND 0: heavy entries per x slot, 1: heavy tuples per z
// slot) and the same maxima (KIND 2 / 3: key heaviness per x / z slot)
// through a workgroup-private LDS array, for slot spaces of at most 32768
// slots (small attribute domains, where per-tuple atomics would pile onto a
// few hot lines); one global update per nonzero slot and workgroup.
constexpr uint32_t kLdsSlotsD = 32768;
template <int KIND>
__global__ __launch_bounds__(1024) void k_slot_lds(const int2* __restrict__ S, const uint32_t* __restrict__ slotR,
                                                   const int32_t* __restrict__ slotRx,
                                                   const int32_t* __restrict__ slotS,
                                                   const int32_t* __restrict__ slotSz, int32_t zlo, uint64_t n,
                                                   const uint64_t* __restrict__ ex_y, const int32_t* __restrict__ hid,
                                                   const uint32_t* __restrict__ degR,
                                                   const uint32_t* __restrict__ degS, uint32_t nslots,
                                                   uint32_t* out) {
    extern __shared__ uint32_t lds[];
    for (uint32_t s = threadIdx.x; s < nslots; s += 1024) lds[s] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * 1024 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 1024) {
        int32_t slot = -1;
        uint32_t ys = 0;
        if (KIND == 0 || KIND == 2) {
            slot = slotRx[i];
            if (slot >= 0) ys = slotR[i];
        } else {
            const int32_t sy = slotS[i];
            if (sy >= 0) {
                ys = (uint32_t)sy;
                slot = slotSz == nullptr ? (int32_t)((uint32_t)S[i].x - (uint32_t)zlo) : slotSz[i];
            }
        }
        if (slot < 0) continue;
        if (KIND <= 1) {
            if (hid[(uint32_t)(ex_y[ys] >> 32)] >= 0) atomicAdd(&lds[slot], 1u);
        } else {
            const uint32_t w = key_w(degR, degS, ys);
            if (w > lds[slot]) atomicMax(&lds[slot], w);
        }
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < nslots; s += 1024) {
        const uint32_t v = lds[s];
        if (v == 0) continue;
        if (KIND <= 1) atomicAdd(&out[s], v);
        else if (v > out[s]) atomicMax(&out[s], v);
    }
}

template <int KIND>
bool slot_lds(const int2* S, const uint32_t* slotR, const int32_t* slotRx, const int32_t* slotS, const int32_t* slotSz,
              int32_t zlo, uint64_t n, const uint64_t* ex_y, const int32_t* hid, const uint32_t* degR,
              const uint32_t* degS, uint32_t nslots, uint32_t* out, hipStream_t st) {
    if (nslots > kLdsSlotsD || n < 8ull * nslots || n == 0) return false;
    const size_t lds = (size_t)nslots * 4u;
    if (hipFuncSetAttribute((const void*)k_slot_lds<KIND>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    uint64_t g = (n + 32767) / 32768;
    if (g > 256) g = 256;
    hipLaunchKernelGGL(k_slot_lds<KIND>, dim3((unsigned)g), dim3(1024), lds, st, S, slotR, slotRx, slotS, slotSz, zlo,
                       n, ex_y, hid, degR, degS, nslots, out);
    return true;
}

// hist[4][kDenseBins]: x rows, their heavy entries, z columns, their heavy tuples, by heavy-degree bin
__global__ __launch_bounds__(256) void k_hdeg_hist(const uint32_t* __restrict__ hdx, uint32_t nx1,
                                                   const uint32_t* __restrict__ hdz, uint32_t nz1,
                                                   unsigned long long* hist) {
    __shared__ unsigned long long lh[4 * kDenseBins];
    for (uint32_t i = threadIdx.x; i < 4 * kDenseBins; i += 256) lh[i] = 0;
    __syncthreads();
    const uint32_t n = max(nx1, nz1);
    for (uint32_t s = blockIdx.x * 256 + threadIdx.x; s < n; s += gridDim.x * 256) {
        if (s < nx1 && hdx[s] != 0) {
            const uint32_t b = dim3h::dense_bin(hdx[s]);
            atomicAdd(&lh[b], 1ull);
            atomicAdd(&lh[kDenseBins + b], (unsigned long long)hdx[s]);
        }
        if (s < nz1 && hdz[s] != 0) {
            const uint32_t b = dim3h::dense_bin(hdz[s]);
            atomicAdd(&lh[2 * kDenseBins + b], 1ull);
            atomicAdd(&lh[3 * kDenseBins + b], (unsigned long long)hdz[s]);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 4 * kDenseBins; i += 256)
        if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

// hz -> z ID, and the first column of every z tile (ntiles + 1 entries)
__global__ void k_hz_fill(const uint32_t* __restrict__ zflag, const uint32_t* __restrict__ hzpos, uint32_t nz,
                          uint32_t* __restrict__ hz2zid, uint32_t ntiles, uint32_t tile_size, uint32_t Nh,
                          uint32_t* __restrict__ hz_tile) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nz && zflag[i]) hz2zid[hzpos[i]] = i;
    if (i <= ntiles) {
        const uint64_t zt = (uint64_t)i * tile_size;
        hz_tile[i] = zt < nz ? hzpos[zt] : Nh;
    }
}

// One byte of a u32 word array: project stores 1, COUNT adds the
// multiplicity (a byte reaching 128 flags the overflow; the query then runs
// without the int8 product).
template <bool COUNT>
__device__ __forceinline__ void put_byte(uint32_t* words, uint64_t byte, uint64_t* overflow) {
    const uint32_t sh = (uint32_t)(byte & 3u) * 8u;
    if (COUNT) {
        const uint32_t old = atomicAdd(&words[byte >> 2], 1u << sh);
        if (((old >> sh) & 0xFFu) >= 127u) atomicOr((unsigned long long*)overflow, 1ull);
    } else {
        atomicOr(&words[byte >> 2], 1u << sh);
    }
}

// A row hx from x's CSR row: the heavy entries (bit 31) set or count their
// key's byte.  One workgroup per row, so the atomics stay in one row.
template <bool COUNT>
__global__ __launch_bounds__(256) void k_build_a_rows(const uint32_t* __restrict__ rowptrR,
                                                      const int32_t* __restrict__ colR, const int32_t* __restrict__ hid,
                                                      const uint32_t* __restrict__ hx2x, uint32_t Kp, uint32_t* A,
                                                      uint64_t* overflow) {
    const uint32_t r = blockIdx.x;
    const uint32_t x = hx2x[r];
    const uint32_t rb = rowptrR[x], re = rowptrR[x + 1];
    for (uint32_t i = rb + threadIdx.x; i < re; i += 256) {
        const int32_t c = colR[i];
        if (c >= 0) continue;
        const int32_t h = hid[(uint32_t)c & 0x7FFFFFFFu];
        put_byte<COUNT>(A, (uint64_t)r * Kp + (uint32_t)h, overflow);
    }
}

// A row built in LDS (Kp <= kLdsRowBytes): byte atomics on LDS, one
// coalesced row store.
constexpr uint32_t kLdsRowBytes = 65536;
template <bool COUNT>
__global__ __launch_bounds__(256) void k_build_a_lds(const uint32_t* __restrict__ rowptrR,
                                                     const int32_t* __restrict__ colR, const int32_t* __restrict__ hid,
                                                     const uint32_t* __restrict__ hx2x, uint32_t Kp,
                                                     uint32_t* __restrict__ A, uint64_t* overflow) {
    extern __shared__ uint32_t row[];
    const uint32_t r = blockIdx.x;
    const uint32_t words = Kp / 4u;
    for (uint32_t w = threadIdx.x; w < words; w += 256) row[w] = 0;
    __syncthreads();
    const uint32_t x = hx2x[r];
    const uint32_t rb = rowptrR[x], re = rowptrR[x + 1];
    for (uint32_t i = rb + threadIdx.x; i < re; i += 256) {
        const int32_t c = colR[i];
        if (c >= 0) continue;
        const uint32_t h = (uint32_t)hid[(uint32_t)c & 0x7FFFFFFFu];
        const uint32_t sh = (h & 3u) * 8u;
        if (COUNT) {
            const uint32_t old = atomicAdd(&row[h >> 2], 1u << sh);
            if (((old >> sh) & 0xFFu) >= 127u) atomicOr((unsigned long long*)overflow, 1ull);
        } else {
            atomicOr(&row[h >> 2], 1u << sh);
        }
    }
    __syncthreads();
    uint32_t* dst = A + (uint64_t)r * words;
    for (uint32_t w = threadIdx.x; w < words; w += 256) dst[w] = row[w];
}

__device__ __forceinline__ uint32_t z_of(const int2* S, uint64_t i, const int32_t* slotSz, const uint64_t* ex_z,
                                         int32_t zlo) {
    return ex_z == nullptr ? (uint32_t)S[i].x - (uint32_t)zlo : (uint32_t)(ex_z[slotSz[i]] >> 32);
}

// Bt[hz][h] for each S tuple with a heavy key (MODE 0: 0/1, 1: multiplicity),
// or bit hz of the z bitmap of key h (MODE 2, bit-or).  MODE 2 with tw > 0:
// the bitmap in the layout of project tiles of tw words (tile t = hz / 32 tw
// holds words t tw .. t tw + tw - 1, local offset zl in the layout blocked
// by 2^blk words when blk > 0, else plain), so a fused seed ORs whole words
// into its tile.
template <int MODE>
__global__ void k_build_b(const int2* __restrict__ S, const int32_t* __restrict__ slotS,
                          const int32_t* __restrict__ slotSz, const uint64_t* __restrict__ ex_z, int32_t zlo,
                          uint64_t nS, const uint64_t* __restrict__ ex_y, const int32_t* __restrict__ hid,
                          const uint32_t* __restrict__ zflag, const uint32_t* __restrict__ hzpos, uint32_t Kp,
                          uint32_t dwords, uint32_t* B, uint64_t* overflow, uint32_t tw, uint32_t blk) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nS) return;
    const int32_t sy = slotS[i];
    if (sy < 0) return;
    const int32_t h = hid[(uint32_t)(ex_y[sy] >> 32)];
    if (h < 0) return;
    const uint32_t zid = z_of(S, i, slotSz, ex_z, zlo);
    if (!zflag[zid]) return;  // a light column: its paths stay sparse
    const uint32_t hz = hzpos[zid];
    if (MODE == 2) {
        uint32_t word = hz >> 5, bit = 1u << (hz & 31u);
        if (tw != 0u && blk != 0u) {
            const uint32_t t = hz / (32u * tw), zl = hz - t * 32u * tw;
            word = t * tw + tl_word(zl, (1u << blk) - 1u);
            bit = tl_bit(zl, blk);
        }
        atomicOr(&B[(uint64_t)h * dwords + word], bit);
    } else {
        put_byte<MODE == 1>(B, (uint64_t)hz * Kp + (uint32_t)h, overflow);
    }
}

inline unsigned grid1(uint64_t n) { return (unsigned)(n == 0 ? 1 : (n + 255) / 256); }
// the counting passes: grid-stride over at most kTallyBlocks workgroups
inline unsigned tally_grid(uint64_t n) { return std::min<unsigned>(grid1(n), kTallyBlocks); }

}  // namespace

namespace dim3h {

int gemm_i8(const int8_t* A, const int8_t* Bt, uint32_t M, uint32_t N, uint32_t Kp, uint32_t* dbits,
            uint32_t dwords, int32_t* cint, hipStream_t st) {
    if (M == 0 || N == 0) return 0;
    if (Kp % kTK != 0) return (int)hipErrorInvalidValue;
    if (M >= 256u && N >= 128u) {
        constexpr int TI = 8, TJ = 4;  // 256 x 128 workgroup tiles, 128 x 64 per wavefront
        const uint64_t tiles = (uint64_t)((M + 32 * TI - 1) / (32 * TI)) * ((N + 32 * TJ - 1) / (32 * TJ));
        const uint64_t grid = (tiles + 7) / 8 * 8;
        if (grid > 0x7FFFFFFFull) return (int)hipErrorInvalidValue;
        hipLaunchKernelGGL((k_gemm_i8_big<TI, TJ>), dim3((unsigned)grid), dim3(256), 0, st, A, Bt, M, N, Kp, dbits,
                           dwords, cint);
        return (int)hipGetLastError();
    }
    dim3 grid((N + kTN - 1) / kTN, (M + kTM - 1) / kTM);
    if (grid.y > 65535) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gemm_i8, grid, dim3(256), 0, st, A, Bt, M, N, Kp, dbits, dwords, cint);
    return (int)hipGetLastError();
}

int bitor_rows(const uint32_t* rowptrR, const int32_t* colR, const int32_t* hid, const uint32_t* hx2x, uint32_t Mh,
               const uint32_t* bbits, uint32_t dwords, uint32_t* dbits, hipStream_t st) {
    if (Mh == 0) return 0;
    hipLaunchKernelGGL(k_bitor, dim3(Mh), dim3(256), 0, st, rowptrR, colR, hid, hx2x, Mh, bbits, dwords, dbits);
    return (int)hipGetLastError();
}

void dense_weights(const int2* R, uint64_t nR, const uint32_t* slotR, const int32_t* slotRx, const int2* S,
                   uint64_t nS, const int32_t* slotS, const int32_t* slotSz, int32_t zlo, const uint32_t* degR,
                   const uint32_t* degS, uint32_t* xw, uint32_t nx1, uint32_t* zw, uint32_t nz1, hipStream_t st,
                   const int32_t* hsx, const int32_t* hsz) {
    (void)R;
    if (nR && !slot_lds<2>(S, slotR, slotRx, slotS, slotSz, zlo, nR, nullptr, nullptr, degR, degS, nx1, xw, st))
        hipLaunchKernelGGL(k_wmax_x, dim3(tally_grid(nR)), dim3(256), 0, st, slotR, slotRx, nR, degR, degS, xw, hsx);
    if (nS && !slot_lds<3>(S, slotR, slotRx, slotS, slotSz, zlo, nS, nullptr, nullptr, degR, degS, nz1, zw, st))
        hipLaunchKernelGGL(k_wmax_z, dim3(tally_grid(nS)), dim3(256), 0, st, S, slotS, slotSz, zlo, nS, degR, degS, zw,
                           hsz);
}

void dense_hist(const uint32_t* degR, const uint32_t* degS, uint32_t ny1, const uint32_t* xw, uint32_t nx1,
                const uint32_t* zw, uint32_t nz1, unsigned long long* hist, hipStream_t st) {
    uint64_t n = std::max(ny1, std::max(nx1, nz1));
    unsigned g = grid1(n);
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(k_dense_hist, dim3(g), dim3(256), 0, st, degR, degS, ny1, xw, nx1, zw, nz1, hist);
}

void dense_mark_keys(const uint32_t* degR, const uint32_t* degS, uint32_t ny1, const uint64_t* ex_y, uint32_t bmin,
                     int32_t* hid, uint32_t* h2y, uint32_t* counters, hipStream_t st) {
    hipLaunchKernelGGL(k_mark_hid, dim3(grid1(ny1)), dim3(256), 0, st, degR, degS, ny1, ex_y, bmin, hid, h2y,
                       counters);
}

void dense_heavy_degrees(const uint32_t* slotR, const int32_t* slotRx, uint64_t nR, const int2* S,
                         const int32_t* slotS, const int32_t* slotSz, int32_t zlo, uint64_t nS, const uint64_t* ex_y,
                         const int32_t* hid, uint32_t* hdx, uint32_t nx1, uint32_t* hdz, uint32_t nz1,
                         unsigned long long* hist, hipStream_t st, const int32_t* hsx, const int32_t* hsz) {
    if (nR && !slot_lds<0>(S, slotR, slotRx, slotS, slotSz, zlo, nR, ex_y, hid, nullptr, nullptr, nx1, hdx, st))
        hipLaunchKernelGGL(k_hdeg_x, dim3(tally_grid(nR)), dim3(256), 0, st, slotR, slotRx, nR, ex_y, hid, hdx, hsx);
    if (nS && !slot_lds<1>(S, slotR, slotRx, slotS, slotSz, zlo, nS, ex_y, hid, nullptr, nullptr, nz1, hdz, st))
        hipLaunchKernelGGL(k_hdeg_z, dim3(tally_grid(nS)), dim3(256), 0, st, S, slotS, slotSz, zlo, nS, ex_y, hid, hdz,
                           hsz);
    unsigned g = grid1(std::max(nx1, nz1));
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(k_hdeg_hist, dim3(g), dim3(256), 0, st, hdx, nx1, hdz, nz1, hist);
}

void dense_mark_rows_cols(const uint32_t* degX, const uint64_t* ex_x, uint32_t nx1, const uint32_t* xw,
                          const uint32_t* degZ, const uint64_t* ex_z, uint32_t nz1, const uint32_t* zw, uint32_t bmin,
                          const uint32_t* hdx, uint32_t thx, const uint32_t* hdz, uint32_t thz, bool all_z,
                          int32_t* hx, uint32_t* hx2x, uint32_t* zflag, uint32_t* counters, hipStream_t st) {
    hipLaunchKernelGGL(k_mark_hx, dim3(grid1(nx1)), dim3(256), 0, st, degX, ex_x, nx1, xw, bmin, hdx, thx, hx, hx2x,
                       counters + 1);
    hipLaunchKernelGGL(k_mark_zflag, dim3(grid1(nz1)), dim3(256), 0, st, degZ, ex_z, nz1, zw, all_z ? 0u : bmin, hdz,
                       thz, zflag);
}

void dense_columns(const uint32_t* zflag, const uint32_t* hzpos, uint32_t nz, uint32_t* hz2zid, uint32_t ntiles,
                   uint32_t tile_size, uint32_t Nh, uint32_t* hz_tile, hipStream_t st) {
    const uint64_t n = std::max<uint64_t>(nz, (uint64_t)ntiles + 1);
    hipLaunchKernelGGL(k_hz_fill, dim3(grid1(n)), dim3(256), 0, st, zflag, hzpos, nz, hz2zid, ntiles, tile_size, Nh,
                       hz_tile);
}

void dense_build_a(bool count, const uint32_t* rowptrR, const int32_t* colR, const int32_t* hid,
                   const uint32_t* hx2x, uint32_t Mh, uint32_t Kp, uint32_t* A, uint64_t* overflow, hipStream_t st) {
    if (!Mh) return;
    if (Kp <= kLdsRowBytes) {
        hipFuncSetAttribute((const void*)k_build_a_lds<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kLdsRowBytes);
        hipFuncSetAttribute((const void*)k_build_a_lds<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kLdsRowBytes);
        if (count)
            hipLaunchKernelGGL(k_build_a_lds<true>, dim3(Mh), dim3(256), Kp, st, rowptrR, colR, hid, hx2x, Kp, A,
                               overflow);
        else
            hipLaunchKernelGGL(k_build_a_lds<false>, dim3(Mh), dim3(256), Kp, st, rowptrR, colR, hid, hx2x, Kp, A,
                               overflow);
        return;
    }
    if (count)
        hipLaunchKernelGGL(k_build_a_rows<true>, dim3(Mh), dim3(256), 0, st, rowptrR, colR, hid, hx2x, Kp, A,
                           overflow);
    else
        hipLaunchKernelGGL(k_build_a_rows<false>, dim3(Mh), dim3(256), 0, st, rowptrR, colR, hid, hx2x, Kp, A,
                           overflow);
}

void dense_build_b(int kind, const int2* S, const int32_t* slotS, const int32_t* slotSz, const uint64_t* ex_z,
                   int32_t zlo, uint64_t nS, const uint64_t* ex_y, const int32_t* hid, const uint32_t* zflag,
                   const uint32_t* hzpos, uint32_t Kp, uint32_t dwords, uint32_t* B, uint64_t* overflow,
                   hipStream_t st, uint32_t tile_words, uint32_t blk) {
    if (!nS) return;
    const unsigned g = grid1(nS);
    if (kind == 2)
        hipLaunchKernelGGL(k_build_b<2>, dim3(g), dim3(256), 0, st, S, slotS, slotSz, ex_z, zlo, nS, ex_y, hid, zflag,
                           hzpos, Kp, dwords, B, overflow, tile_words, blk);
    else if (kind == 1)
        hipLaunchKernelGGL(k_build_b<1>, dim3(g), dim3(256), 0, st, S, slotS, slotSz, ex_z, zlo, nS, ex_y, hid, zflag,
                           hzpos, Kp, dwords, B, overflow, 0u, 0u);
    else
        hipLaunchKernelGGL(k_build_b<0>, dim3(g), dim3(256), 0, st, S, slotS, slotSz, ex_z, zlo, nS, ex_y, hid, zflag,
                           hzpos, Kp, dwords, B, overflow, 0u, 0u);
}

}  // namespace dim3h
