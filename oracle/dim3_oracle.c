/*
 * CPU oracle for DIM3 join-project and join-aggregate (COUNT).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this file; only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark load it, and only as the checker.
 *
 * It restates, in plain C with 64-bit counts, the algorithm of the reference
 * (schencoding/JoinProject-DIM3):
 *
 *   oracle_join_project     DIM3::doJoinProject          dim3.hpp:987-1055
 *       swap rule                                         dim3.hpp:993-994
 *       y-ID mapping + semi-join filter of S              dim3.hpp:794-836 (mapping_y)
 *       x / z ID mapping                                  dim3.hpp:733-761 (mapping_xz)
 *       early return when no S tuple survives             dim3.hpp:1027
 *       CSR of R by x                                     dim3.hpp:1058-1086 (divide_R)
 *       CSR of S by y (all-sparse branch)                 dim3.hpp:1230-1253 (divide_S)
 *       SPA expansion with last-writer dedup              dim3.hpp:952-973
 *   oracle_join_aggregate_count
 *                           DIM3JoinAggregate + SpmmAggregate with a COUNT
 *                           policy                        dim3.hpp:1267-1373
 *       first touch -> insert (count = 1), later -> update (+1); rows are
 *       emitted in first-touch order                      dim3.hpp:1286-1307
 *   oracle_join_aggregate   SpmmAggregate with a SUM / MIN / MAX / AVG policy
 *                           over the path value vR + vS   dim3.hpp:1267-1311,
 *                           (the value AggregateFunctionAvg combines) 1453-1480
 *   oracle_join_aggregate_f32  the same on float values (float path value,
 *                           double accumulation, see run_value_f32)
 *   oracle_estimate         EstimateJoinCardinality       dim3.hpp:193-224, 262-269
 *   oracle_outj             exact sum_b degR(b)*degS(b)   (closed form of the above)
 *   oracle_gen_rand_data    gen_rand_data (minstd_rand0)  dim3.hpp:1556-1568
 *   oracle_project_fp_bitset join-project (and COUNT) fingerprints with heavy
 *                           keys as z bitsets, by x slices (config scale):
 *                           the dense/sparse hybrid of dim3.hpp:894-973 per key
 *
 * Deliberate differences, none of which is observable in the output set:
 *   - no radix partitioning (dim3.hpp:600-731 is a cache optimisation),
 *   - one open-addressing table per attribute instead of LS_mapper + ska
 *     overflow cache (Include/L+S_mapper.hpp), IDs in first-seen order,
 *   - no dense/sparse split: the reference's f2 split (dim3.hpp:1118-1126)
 *     partitions z values and the dense and sparse outputs are disjoint, so
 *     the all-sparse SPA yields the same set,
 *   - counts are 64-bit (the reference's uint32 counters overflow above
 *     ~2.1e9 rows),
 *   - the estimate fallback (dim3.hpp:215-221) reads min(1000, n) tuples
 *     instead of reading past the end of small inputs.
 *
 * Fingerprint (order independent, shared with the GPU library), all
 * arithmetic mod 2^64, u(v) = (u32)v:
 *   k1(x) = mix64(u(x) ^ 0x243F6A8885A308D3) | 1, k2(x) = mix64(u(x) ^ 0x13198A2E03707344) | 1
 *   g(z)  = a(u(z) >> 5) * c(u(z) & 31), with the word factor
 *           a(w) = fmix32(w ^ 0x243F6A88) | 1 and the bit-position factor
 *           c(b) = mix64(0x9E3779B97F4A7C15 * (b + 1)) | 1 (the splitmix64
 *           outputs for seed 0)
 *   project row (x, z):          count += 1, sum += k1(x) g(z), sum2 += k2(x) g(z)
 *   aggregate row (x, z, c):     count += 1, sum += k1(x) g(z), sum2 += k2(x) c g(z), csum += c
 * mix64 is the splitmix64 finaliser (Steele, Lea, Flood 2014), fmix32 the
 * MurmurHash3 32-bit finaliser.  Each term is an x factor times a z factor,
 * so a row's share is k(x) times the sum of its z factors, and the z factors
 * of one 32-aligned word of z values are one word factor times a sum of
 * position factors (the GPU sums a dense word with four byte-table lookups).
 * All factors are odd, so z values of one word never share a factor.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_OOM 2
#define ORC_INVALID 1

typedef struct {
    uint64_t count;
    uint64_t sum;
    uint64_t sum2;
    uint64_t csum;
} orc_fp;

static inline uint64_t orc_mix64(uint64_t k) {
    k ^= k >> 30;
    k *= 0xbf58476d1ce4e5b9ULL;
    k ^= k >> 27;
    k *= 0x94d049bb133111ebULL;
    k ^= k >> 31;
    return k;
}

uint64_t oracle_mix64(uint64_t k) { return orc_mix64(k); }

static inline uint32_t orc_fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

/* z factor of the fingerprint (see the header comment) */
static inline uint64_t orc_zfactor(int32_t z) {
    const uint32_t u = (uint32_t)z;
    const uint64_t a = (uint64_t)(orc_fmix32((u >> 5) ^ 0x243F6A88u) | 1u);
    return a * (orc_mix64(0x9E3779B97F4A7C15ULL * (uint64_t)((u & 31u) + 1u)) | 1u);
}

uint64_t oracle_zfactor(int32_t z) { return orc_zfactor(z); }

/* one row's terms of the fingerprint (see the header comment) */
static inline void orc_fp_add(orc_fp* fp, int32_t x, int32_t z, const int64_t* count) {
    const uint64_t ux = (uint64_t)(uint32_t)x;
    const uint64_t k1 = orc_mix64(ux ^ 0x243F6A8885A308D3ULL) | 1u, k2 = orc_mix64(ux ^ 0x13198A2E03707344ULL) | 1u;
    const uint64_t g = orc_zfactor(z);
    fp->count += 1;
    fp->sum += k1 * g;
    if (count) {
        fp->sum2 += k2 * ((uint64_t)*count * g);
        fp->csum += (uint64_t)*count;
    } else {
        fp->sum2 += k2 * g;
    }
}

/* ------------------------------------------------------------------ */
/* int32 -> dense ID map, open addressing with linear probing.          */
/* Every int32 is a legal key (the reference allows -1 too,             */
/* Include/L+S_mapper.hpp:67-104), so occupancy is a separate byte.     */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t* keys;
    int32_t* ids;
    uint8_t* used;
    uint64_t mask;
    int32_t n;
} omap;

static inline uint32_t omap_hash(int32_t k) {
    uint32_t h = (uint32_t)k;
    h ^= h >> 16;
    h *= 0x85ebca6bU;
    h ^= h >> 13;
    h *= 0xc2b2ae35U;
    h ^= h >> 16;
    return h;
}

static int omap_init(omap* m, uint64_t expected) {
    uint64_t cap = 16;
    while (cap < 2 * expected + 16) cap <<= 1;
    m->keys = (int32_t*)malloc(cap * sizeof(int32_t));
    m->ids = (int32_t*)malloc(cap * sizeof(int32_t));
    m->used = (uint8_t*)calloc(cap, 1);
    m->mask = cap - 1;
    m->n = 0;
    if (!m->keys || !m->ids || !m->used) return ORC_OOM;
    return ORC_OK;
}

static void omap_free(omap* m) {
    free(m->keys);
    free(m->ids);
    free(m->used);
    memset(m, 0, sizeof(*m));
}

/* Returns the dense ID of key, assigning the next one on first sight
 * (LS_mapper::insert, Include/L+S_mapper.hpp:58-114). */
static int32_t omap_insert(omap* m, int32_t key) {
    uint64_t s = omap_hash(key) & m->mask;
    for (;;) {
        if (!m->used[s]) {
            m->used[s] = 1;
            m->keys[s] = key;
            m->ids[s] = m->n;
            return m->n++;
        }
        if (m->keys[s] == key) return m->ids[s];
        s = (s + 1) & m->mask;
    }
}

/* Returns the dense ID of key or -1 (LS_mapper::find, L+S_mapper.hpp:116-160). */
static int32_t omap_find(const omap* m, int32_t key) {
    uint64_t s = omap_hash(key) & m->mask;
    for (;;) {
        if (!m->used[s]) return -1;
        if (m->keys[s] == key) return m->ids[s];
        s = (s + 1) & m->mask;
    }
}

/* ------------------------------------------------------------------ */
/* Output sinks                                                          */
/* ------------------------------------------------------------------ */
typedef struct {
    int fingerprint_only;
    int with_counts;
    int32_t* pairs;  /* 2*cap int32: x, z */
    int64_t* counts; /* cap */
    uint64_t n, cap;
    orc_fp fp;
    int oom;
} orc_sink;

static void sink_push(orc_sink* k, int32_t x, int32_t z, int64_t c) {
    orc_fp_add(&k->fp, x, z, k->with_counts ? &c : NULL);
    if (k->fingerprint_only || k->oom) return;
    if (k->n == k->cap) {
        uint64_t nc = k->cap ? 2 * k->cap : 1024;
        int32_t* p = (int32_t*)realloc(k->pairs, nc * 2 * sizeof(int32_t));
        if (!p) { k->oom = 1; return; }
        k->pairs = p;
        if (k->with_counts) {
            int64_t* q = (int64_t*)realloc(k->counts, nc * sizeof(int64_t));
            if (!q) { k->oom = 1; return; }
            k->counts = q;
        }
        k->cap = nc;
    }
    k->pairs[2 * k->n] = x;
    k->pairs[2 * k->n + 1] = z;
    if (k->with_counts) k->counts[k->n] = c;
    k->n++;
}

/* ------------------------------------------------------------------ */
/* Shared mapping + CSR build (dim3.hpp:399-455, 1058-1086, 1230-1253)  */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t nx, ny, nz;
    int32_t* id2x; /* nx */
    int32_t* id2z; /* nz */
    uint64_t* rjr; /* nx+1 : R CSR by x, columns = y id */
    int32_t* ric;
    uint64_t* sjr; /* ny+1 : S CSR by y, columns = z id */
    int32_t* sic;
    uint64_t* sidx; /* S CSR slot -> original S tuple index (for the row's values) */
    int32_t* rval;  /* value of each R CSR entry (value aggregates only) */
    int32_t* sval;  /* value of each S CSR entry */
} orc_plan;

static void plan_free(orc_plan* p) {
    free(p->id2x); free(p->id2z); free(p->rjr); free(p->ric);
    free(p->sjr); free(p->sic); free(p->sidx); free(p->rval); free(p->sval);
    memset(p, 0, sizeof(*p));
}

/* R and S are interleaved (first, second) int32 pairs with the join key in
 * .second (main.cpp:6-11).  Returns ORC_OK; p->nx == 0 means empty output. */
static int build_plan_v(const int32_t* R, const int32_t* vR, uint64_t nR, const int32_t* S, const int32_t* vS,
                        uint64_t nS, orc_plan* p);
static int build_plan(const int32_t* R, uint64_t nR, const int32_t* S, uint64_t nS, orc_plan* p) {
    return build_plan_v(R, NULL, nR, S, NULL, nS, p);
}

/* As build_plan; with vR / vS given, the CSR entries also carry the value
 * of their tuple (the VAL arrays of dim3.hpp:1376-1450). */
static int build_plan_v(const int32_t* R, const int32_t* vR, uint64_t nR, const int32_t* S, const int32_t* vS,
                        uint64_t nS, orc_plan* p) {
    memset(p, 0, sizeof(*p));
    omap my, mx, mz;
    int rc = ORC_OK;
    int32_t* ry = NULL; int32_t* rx = NULL; int32_t* sy = NULL; int32_t* sz = NULL; int32_t* sv = NULL;
    uint64_t nSm = 0;
    if (omap_init(&my, nR) || omap_init(&mx, nR) || omap_init(&mz, nS)) return ORC_OOM;

    ry = (int32_t*)malloc((nR ? nR : 1) * sizeof(int32_t));
    rx = (int32_t*)malloc((nR ? nR : 1) * sizeof(int32_t));
    sy = (int32_t*)malloc((nS ? nS : 1) * sizeof(int32_t));
    sz = (int32_t*)malloc((nS ? nS : 1) * sizeof(int32_t));
    sv = (int32_t*)malloc((nS ? nS : 1) * sizeof(int32_t));
    if (!ry || !rx || !sy || !sz || !sv) { rc = ORC_OOM; goto done; }

    /* mapping_y: y IDs from R, S probes them and drops misses (dim3.hpp:822-836) */
    for (uint64_t i = 0; i < nR; i++) ry[i] = omap_insert(&my, R[2 * i + 1]);
    for (uint64_t i = 0; i < nS; i++) {
        int32_t t = omap_find(&my, S[2 * i + 1]);
        if (t != -1) { sy[nSm] = t; sz[nSm] = S[2 * i]; sv[nSm] = vS ? vS[i] : 0; nSm++; }
    }
    /* dim3.hpp:1027: nothing survives -> empty result */
    if (nR == 0 || nSm == 0) goto done;

    /* mapping_xz over R.x and the surviving S.z (dim3.hpp:733-761) */
    for (uint64_t i = 0; i < nR; i++) rx[i] = omap_insert(&mx, R[2 * i]);
    for (uint64_t i = 0; i < nSm; i++) sz[i] = omap_insert(&mz, sz[i]);
    p->nx = mx.n; p->ny = my.n; p->nz = mz.n;
    p->id2x = (int32_t*)malloc((size_t)p->nx * sizeof(int32_t));
    p->id2z = (int32_t*)malloc((size_t)p->nz * sizeof(int32_t));
    if (!p->id2x || !p->id2z) { rc = ORC_OOM; goto done; }
    for (uint64_t s = 0; s <= mx.mask; s++) if (mx.used[s]) p->id2x[mx.ids[s]] = mx.keys[s];
    for (uint64_t s = 0; s <= mz.mask; s++) if (mz.used[s]) p->id2z[mz.ids[s]] = mz.keys[s];

    /* divide_R: stable counting sort of R by x (dim3.hpp:1058-1086) */
    p->rjr = (uint64_t*)calloc((size_t)p->nx + 1, sizeof(uint64_t));
    p->ric = (int32_t*)malloc((nR ? nR : 1) * sizeof(int32_t));
    if (!p->rjr || !p->ric) { rc = ORC_OOM; goto done; }
    for (uint64_t i = 0; i < nR; i++) p->rjr[rx[i] + 1]++;
    for (int32_t i = 0; i < p->nx; i++) p->rjr[i + 1] += p->rjr[i];
    {
        uint64_t* cur = (uint64_t*)malloc(((size_t)p->nx + 1) * sizeof(uint64_t));
        if (!cur) { rc = ORC_OOM; goto done; }
        memcpy(cur, p->rjr, ((size_t)p->nx + 1) * sizeof(uint64_t));
        if (vR) {
            p->rval = (int32_t*)malloc((nR ? nR : 1) * sizeof(int32_t));
            if (!p->rval) { free(cur); rc = ORC_OOM; goto done; }
        }
        for (uint64_t i = 0; i < nR; i++) {
            if (vR) p->rval[cur[rx[i]]] = vR[i];
            p->ric[cur[rx[i]]++] = ry[i];
        }
        free(cur);
    }

    /* divide_S, all-sparse branch: stable counting sort of S by y
     * (dim3.hpp:1230-1253) */
    p->sjr = (uint64_t*)calloc((size_t)p->ny + 1, sizeof(uint64_t));
    p->sic = (int32_t*)malloc(nSm * sizeof(int32_t));
    if (!p->sjr || !p->sic) { rc = ORC_OOM; goto done; }
    for (uint64_t i = 0; i < nSm; i++) p->sjr[sy[i] + 1]++;
    for (int32_t i = 0; i < p->ny; i++) p->sjr[i + 1] += p->sjr[i];
    {
        uint64_t* cur = (uint64_t*)malloc(((size_t)p->ny + 1) * sizeof(uint64_t));
        if (!cur) { rc = ORC_OOM; goto done; }
        memcpy(cur, p->sjr, ((size_t)p->ny + 1) * sizeof(uint64_t));
        if (vS) {
            p->sval = (int32_t*)malloc((nSm ? nSm : 1) * sizeof(int32_t));
            if (!p->sval) { free(cur); rc = ORC_OOM; goto done; }
        }
        for (uint64_t i = 0; i < nSm; i++) {
            if (vS) p->sval[cur[sy[i]]] = sv[i];
            p->sic[cur[sy[i]]++] = sz[i];
        }
        free(cur);
    }

done:
    free(ry); free(rx); free(sy); free(sz); free(sv);
    omap_free(&my); omap_free(&mx); omap_free(&mz);
    if (rc != ORC_OK) plan_free(p);
    return rc;
}

/* Hybrid_solution sparse part (dim3.hpp:952-973): SPAw[z] holds the last
 * row that emitted z. */
static int run_project(const orc_plan* p, int exclude_equal, orc_sink* k) {
    if (p->nx == 0) return ORC_OK;
    int32_t* spaw = (int32_t*)malloc((size_t)p->nz * sizeof(int32_t));
    if (!spaw) return ORC_OOM;
    for (int32_t i = 0; i < p->nz; i++) spaw[i] = -1;
    for (int32_t x = 0; x < p->nx; x++) {
        int32_t xv = p->id2x[x];
        for (uint64_t i = p->rjr[x]; i < p->rjr[x + 1]; i++) {
            int32_t y = p->ric[i];
            for (uint64_t j = p->sjr[y]; j < p->sjr[y + 1]; j++) {
                int32_t z = p->sic[j];
                if (spaw[z] != x) {
                    spaw[z] = x;
                    int32_t zv = p->id2z[z];
                    if (exclude_equal && zv == xv) continue;
                    sink_push(k, xv, zv, 1);
                }
            }
        }
    }
    free(spaw);
    return k->oom ? ORC_OOM : ORC_OK;
}

/* SpmmAggregate with a COUNT policy (dim3.hpp:1286-1307): insert sets the
 * count to 1, update adds 1, rows are emitted in first-touch order.  A
 * duplicate (x,y) or (z,y) tuple contributes one join path per copy. */
static int run_count(const orc_plan* p, int exclude_equal, orc_sink* k) {
    if (p->nx == 0) return ORC_OK;
    int32_t* spaw = (int32_t*)malloc((size_t)p->nz * sizeof(int32_t));
    int64_t* cnt = (int64_t*)malloc((size_t)p->nz * sizeof(int64_t));
    int32_t* list = (int32_t*)malloc((size_t)p->nz * sizeof(int32_t));
    if (!spaw || !cnt || !list) { free(spaw); free(cnt); free(list); return ORC_OOM; }
    for (int32_t i = 0; i < p->nz; i++) spaw[i] = -1;
    for (int32_t x = 0; x < p->nx; x++) {
        int32_t xv = p->id2x[x];
        int32_t nl = 0;
        for (uint64_t i = p->rjr[x]; i < p->rjr[x + 1]; i++) {
            int32_t y = p->ric[i];
            for (uint64_t j = p->sjr[y]; j < p->sjr[y + 1]; j++) {
                int32_t z = p->sic[j];
                if (spaw[z] != x) { spaw[z] = x; cnt[z] = 1; list[nl++] = z; }
                else cnt[z] += 1;
            }
        }
        for (int32_t i = 0; i < nl; i++) {
            int32_t zv = p->id2z[list[i]];
            if (exclude_equal && zv == xv) continue;
            sink_push(k, xv, zv, cnt[list[i]]);
        }
    }
    free(spaw); free(cnt); free(list);
    return k->oom ? ORC_OOM : ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Public entry points (ctypes)                                          */
/* ------------------------------------------------------------------ */

/* Join-project.  R_xy / S_zy hold interleaved int32 pairs.  When nR < nS
 * the roles are swapped (dim3.hpp:993-994) and rows come out as
 * (S.first, R.first); *swapped reports that.  With fingerprint_only the
 * rows are not stored.  Output pairs are malloc'd; free with oracle_free. */
int oracle_join_project(const int32_t* R_xy, uint64_t nR, const int32_t* S_zy, uint64_t nS,
                        int exclude_equal, int fingerprint_only,
                        int32_t** out_pairs, uint64_t* n_out, orc_fp* fp, int* swapped) {
    const int32_t* R = R_xy; const int32_t* S = S_zy;
    int sw = 0;
    if (nR < nS) { const int32_t* t = R; R = S; S = t; uint64_t u = nR; nR = nS; nS = u; sw = 1; }
    if (swapped) *swapped = sw;
    orc_plan p;
    int rc = build_plan(R, nR, S, nS, &p);
    if (rc) return rc;
    orc_sink k;
    memset(&k, 0, sizeof(k));
    k.fingerprint_only = fingerprint_only;
    rc = run_project(&p, exclude_equal, &k);
    plan_free(&p);
    if (fp) *fp = k.fp;
    if (n_out) *n_out = k.fp.count;
    if (rc) { free(k.pairs); return rc; }
    if (out_pairs) *out_pairs = k.pairs; else free(k.pairs);
    return ORC_OK;
}

/* Join-project fingerprint with the heavy join keys as z bitsets, for
 * config-scale checks (1e10-1e12 join paths).  It restates the reference's
 * hybrid (dim3.hpp:894-973: a dense part of bitvectors, a sparse SPA part)
 * per join key: a key whose S row holds at least `heavy` tuples keeps one
 * bitset over the z IDs (numbered here in z value order, so a key whose z
 * values cluster spans few words), OR-ed into the row's bitset word by word;
 * the other keys set their paths' bits one by one.  The row's set bits are
 * its distinct z.  Only rows whose x value has (u32)x % xmod == xrem are
 * expanded (xmod <= 1: every row), so disjoint slices add up to the whole
 * fingerprint.  with_paths: the COUNT fingerprint, whose count-weighted
 * terms are sums over the join paths: sum2 = sum over paths of k2(x) g(z) =
 * sum over the row's keys of k2(x) GS(y), GS(y) the sum of g over S[y], and
 * csum = the row's paths (not with exclude_equal).  Swap rule as
 * oracle_join_project. */
int oracle_project_fp_bitset(const int32_t* R_xy, uint64_t nR, const int32_t* S_zy, uint64_t nS, int exclude_equal,
                             uint64_t heavy, uint32_t xmod, uint32_t xrem, int with_paths, orc_fp* fp, int* swapped) {
    const int32_t* R = R_xy; const int32_t* S = S_zy;
    int sw = 0;
    if (with_paths && exclude_equal) return ORC_INVALID;
    if (nR < nS) { const int32_t* t = R; R = S; S = t; uint64_t u = nR; nR = nS; nS = u; sw = 1; }
    if (swapped) *swapped = sw;
    memset(fp, 0, sizeof(*fp));
    orc_plan p;
    int rc = build_plan(R, nR, S, nS, &p);
    if (rc) return rc;
    if (p.nx == 0) { plan_free(&p); return ORC_OK; }
    const int32_t nz = p.nz, ny = p.ny;
    const uint64_t nw = ((uint64_t)nz + 63) / 64;
    /* z IDs renumbered by value: rank[id] */
    int32_t* order = (int32_t*)malloc((size_t)nz * sizeof(int32_t));
    int32_t* rank = (int32_t*)malloc((size_t)nz * sizeof(int32_t));
    uint64_t* gz = (uint64_t*)malloc((size_t)nz * sizeof(uint64_t));
    int32_t* zval = (int32_t*)malloc((size_t)nz * sizeof(int32_t));
    int32_t* hid = (int32_t*)malloc((size_t)(ny ? ny : 1) * sizeof(int32_t));
    uint64_t* gs = with_paths ? (uint64_t*)calloc((size_t)(ny ? ny : 1), sizeof(uint64_t)) : NULL;
    uint64_t* row = (uint64_t*)calloc(nw, sizeof(uint64_t));
    uint64_t* bits = NULL; uint32_t* hlo = NULL; uint32_t* hhi = NULL;
    if (!order || !rank || !gz || !zval || !hid || (with_paths && !gs) || !row) { rc = ORC_OOM; goto out; }
    {
        /* sort z IDs by value (counting is not possible over int32: a
         * radix sort of (value, id) pairs in two 16-bit passes) */
        int32_t* tmp = (int32_t*)malloc((size_t)nz * sizeof(int32_t));
        uint64_t* cnt = (uint64_t*)calloc(65537, sizeof(uint64_t));
        if (!tmp || !cnt) { free(tmp); free(cnt); rc = ORC_OOM; goto out; }
        for (int32_t i = 0; i < nz; i++) order[i] = i;
        for (int pass = 0; pass < 2; pass++) {
            const int shift = pass * 16;
            memset(cnt, 0, 65537 * sizeof(uint64_t));
            for (int32_t i = 0; i < nz; i++) {
                uint32_t key = ((uint32_t)p.id2z[order[i]] ^ 0x80000000u) >> shift & 0xFFFFu;
                cnt[key + 1]++;
            }
            for (int k = 0; k < 65536; k++) cnt[k + 1] += cnt[k];
            for (int32_t i = 0; i < nz; i++) {
                uint32_t key = ((uint32_t)p.id2z[order[i]] ^ 0x80000000u) >> shift & 0xFFFFu;
                tmp[cnt[key]++] = order[i];
            }
            memcpy(order, tmp, (size_t)nz * sizeof(int32_t));
        }
        free(tmp); free(cnt);
        for (int32_t r = 0; r < nz; r++) {
            rank[order[r]] = r;
            zval[r] = p.id2z[order[r]];
            gz[r] = orc_zfactor(zval[r]);
        }
    }
    /* heavy keys and their bitsets (over value-ordered z IDs) */
    int32_t K = 0;
    for (int32_t y = 0; y < ny; y++) hid[y] = (p.sjr[y + 1] - p.sjr[y] >= heavy && heavy > 0) ? K++ : -1;
    if (K) {
        bits = (uint64_t*)calloc((size_t)K * nw, sizeof(uint64_t));
        hlo = (uint32_t*)malloc((size_t)K * sizeof(uint32_t));
        hhi = (uint32_t*)malloc((size_t)K * sizeof(uint32_t));
        if (!bits || !hlo || !hhi) { rc = ORC_OOM; goto out; }
    }
    for (int32_t y = 0; y < ny; y++) {
        const int32_t h = hid[y];
        uint64_t acc = 0;
        uint32_t lo = (uint32_t)nw, hi = 0;
        for (uint64_t j = p.sjr[y]; j < p.sjr[y + 1]; j++) {
            const int32_t z = rank[p.sic[j]];
            p.sic[j] = z;  /* the plan's S columns in value order from here on */
            if (with_paths) acc += gz[z];
            if (h >= 0) {
                bits[(size_t)h * nw + (uint32_t)z / 64] |= 1ull << (z & 63);
                if ((uint32_t)z / 64 < lo) lo = (uint32_t)z / 64;
                if ((uint32_t)z / 64 > hi) hi = (uint32_t)z / 64;
            }
        }
        if (with_paths) gs[y] = acc;
        if (h >= 0) { hlo[h] = lo; hhi[h] = hi; }
    }
    for (int32_t x = 0; x < p.nx; x++) {
        const int32_t xv = p.id2x[x];
        if (xmod > 1 && (uint32_t)xv % xmod != xrem) continue;
        uint32_t lo = (uint32_t)nw, hi = 0;
        uint64_t s2 = 0, cs = 0;
        for (uint64_t i = p.rjr[x]; i < p.rjr[x + 1]; i++) {
            const int32_t y = p.ric[i];
            if (with_paths) { s2 += gs[y]; cs += p.sjr[y + 1] - p.sjr[y]; }
            const int32_t h = hid[y];
            if (h >= 0) {
                const uint64_t* b = bits + (size_t)h * nw;
                for (uint32_t w = hlo[h]; w <= hhi[h]; w++) row[w] |= b[w];
                if (hlo[h] < lo) lo = hlo[h];
                if (hhi[h] > hi) hi = hhi[h];
            } else {
                for (uint64_t j = p.sjr[y]; j < p.sjr[y + 1]; j++) {
                    const uint32_t z = (uint32_t)p.sic[j];
                    row[z / 64] |= 1ull << (z & 63);
                    if (z / 64 < lo) lo = z / 64;
                    if (z / 64 > hi) hi = z / 64;
                }
            }
        }
        uint64_t s1 = 0, n = 0;
        for (uint32_t w = lo; w <= hi && lo <= hi; w++) {
            uint64_t v = row[w];
            row[w] = 0;
            while (v) {
                const uint32_t z = w * 64 + (uint32_t)__builtin_ctzll(v);
                v &= v - 1;
                if (exclude_equal && zval[z] == xv) continue;
                s1 += gz[z];
                n++;
            }
        }
        const uint64_t ux = (uint64_t)(uint32_t)xv;
        const uint64_t k1 = orc_mix64(ux ^ 0x243F6A8885A308D3ULL) | 1u, k2 = orc_mix64(ux ^ 0x13198A2E03707344ULL) | 1u;
        fp->count += n;
        fp->sum += k1 * s1;
        if (with_paths) { fp->sum2 += k2 * s2; fp->csum += cs; }
        else fp->sum2 += k2 * s1;
    }
out:
    free(order); free(rank); free(gz); free(zval); free(hid); free(gs); free(row);
    free(bits); free(hlo); free(hhi);
    plan_free(&p);
    return rc;
}

/* Join-aggregate COUNT(*) GROUP BY x,z.  Same swap rule (dim3.hpp:1325-1326). */
/* Value aggregates: SpmmAggregate (dim3.hpp:1267-1311) with a policy over
 * the path value v = vR + vS (the combination AggregateFunctionAvg uses,
 * dim3.hpp:1466-1473): 1 SUM of v, 2 MIN, 3 MAX, 4 AVG = SUM / number of
 * paths, truncated toward zero (its get() on integral values).  All int64. */
static int run_value(const orc_plan* p, int agg, int exclude_equal, orc_sink* k) {
    if (p->nx == 0) return ORC_OK;
    int32_t* spaw = (int32_t*)malloc((size_t)p->nz * sizeof(int32_t));
    int64_t* cnt = (int64_t*)malloc((size_t)p->nz * sizeof(int64_t));
    int64_t* acc = (int64_t*)malloc((size_t)p->nz * sizeof(int64_t));
    int32_t* list = (int32_t*)malloc((size_t)p->nz * sizeof(int32_t));
    if (!spaw || !cnt || !acc || !list) { free(spaw); free(cnt); free(acc); free(list); return ORC_OOM; }
    for (int32_t i = 0; i < p->nz; i++) spaw[i] = -1;
    for (int32_t x = 0; x < p->nx; x++) {
        int32_t xv = p->id2x[x];
        int32_t nl = 0;
        for (uint64_t i = p->rjr[x]; i < p->rjr[x + 1]; i++) {
            int32_t y = p->ric[i];
            for (uint64_t j = p->sjr[y]; j < p->sjr[y + 1]; j++) {
                int32_t z = p->sic[j];
                int64_t v = (int64_t)p->rval[i] + (int64_t)p->sval[j];
                if (spaw[z] != x) { spaw[z] = x; cnt[z] = 1; acc[z] = v; list[nl++] = z; continue; }
                cnt[z] += 1;
                if (agg == 2) { if (v < acc[z]) acc[z] = v; }
                else if (agg == 3) { if (v > acc[z]) acc[z] = v; }
                else acc[z] += v;
            }
        }
        for (int32_t i = 0; i < nl; i++) {
            int32_t zv = p->id2z[list[i]];
            if (exclude_equal && zv == xv) continue;
            int64_t out = agg == 4 ? acc[list[i]] / cnt[list[i]] : acc[list[i]];
            sink_push(k, xv, zv, out);
        }
    }
    free(spaw); free(cnt); free(acc); free(list);
    return k->oom ? ORC_OOM : ORC_OK;
}

/* Float values (the reference's TUPLE<int,int,float> workload,
 * join_aggregate.cpp:21-23): the path value is the float sum vR + vS (one
 * float add, as the policy's val1 + val2 on float); SUM and AVG accumulate
 * the path values in double in first-touch order (exact, hence independent
 * of the order, whenever every partial sum is a double, e.g. for
 * integer-valued data below 2^53) and round once: SUM = (float)sum,
 * AVG = (float)(sum / paths).  MIN / MAX order the path values by IEEE
 * totalOrder (-0 < +0; NaN inputs unsupported).  The result float's bit
 * pattern is the row value (and the fingerprint's count). */
static inline float orc_as_float(int32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline uint32_t orc_float_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }
static inline int32_t orc_fkey(float f) {  /* order-preserving int of a float */
    int32_t b = (int32_t)orc_float_bits(f);
    return b >= 0 ? b : b ^ 0x7FFFFFFF;
}

static int run_value_f32(const orc_plan* p, int agg, int exclude_equal, orc_sink* k) {
    if (p->nx == 0) return ORC_OK;
    int32_t* spaw = (int32_t*)malloc((size_t)p->nz * sizeof(int32_t));
    int64_t* cnt = (int64_t*)malloc((size_t)p->nz * sizeof(int64_t));
    double* acc = (double*)malloc((size_t)p->nz * sizeof(double));
    int32_t* key = (int32_t*)malloc((size_t)p->nz * sizeof(int32_t));
    int32_t* list = (int32_t*)malloc((size_t)p->nz * sizeof(int32_t));
    if (!spaw || !cnt || !acc || !key || !list) {
        free(spaw); free(cnt); free(acc); free(key); free(list);
        return ORC_OOM;
    }
    for (int32_t i = 0; i < p->nz; i++) spaw[i] = -1;
    for (int32_t x = 0; x < p->nx; x++) {
        int32_t xv = p->id2x[x];
        int32_t nl = 0;
        for (uint64_t i = p->rjr[x]; i < p->rjr[x + 1]; i++) {
            int32_t y = p->ric[i];
            for (uint64_t j = p->sjr[y]; j < p->sjr[y + 1]; j++) {
                int32_t z = p->sic[j];
                const float v = orc_as_float(p->rval[i]) + orc_as_float(p->sval[j]);
                const int32_t kv = orc_fkey(v);
                if (spaw[z] != x) { spaw[z] = x; cnt[z] = 1; acc[z] = v; key[z] = kv; list[nl++] = z; continue; }
                cnt[z] += 1;
                acc[z] += v;
                if (agg == 2 ? kv < key[z] : kv > key[z]) key[z] = kv;
            }
        }
        for (int32_t i = 0; i < nl; i++) {
            const int32_t z = list[i];
            int32_t zv = p->id2z[z];
            if (exclude_equal && zv == xv) continue;
            uint32_t bits;
            if (agg == 1) bits = orc_float_bits((float)acc[z]);
            else if (agg == 4) bits = orc_float_bits((float)(acc[z] / (double)cnt[z]));
            else bits = (uint32_t)(key[z] >= 0 ? key[z] : key[z] ^ 0x7FFFFFFF);
            sink_push(k, xv, zv, (int64_t)bits);
        }
    }
    free(spaw); free(cnt); free(acc); free(key); free(list);
    return k->oom ? ORC_OOM : ORC_OK;
}

/* The value aggregate (agg 1-4, see run_value) of R(x, y, vR) join S(z, y, vS);
 * the swap rule moves the values with their relation. */
static int join_aggregate_values(const int32_t* R_xy, const int32_t* vR, uint64_t nR, const int32_t* S_zy,
                                 const int32_t* vS, uint64_t nS, int agg, int f32, int exclude_equal,
                                 int fingerprint_only, int32_t** out_pairs, int64_t** out_vals, uint64_t* n_out,
                                 orc_fp* fp, int* swapped);

int oracle_join_aggregate(const int32_t* R_xy, const int32_t* vR, uint64_t nR, const int32_t* S_zy,
                          const int32_t* vS, uint64_t nS, int agg, int exclude_equal, int fingerprint_only,
                          int32_t** out_pairs, int64_t** out_vals, uint64_t* n_out, orc_fp* fp, int* swapped) {
    return join_aggregate_values(R_xy, vR, nR, S_zy, vS, nS, agg, 0, exclude_equal, fingerprint_only, out_pairs,
                                 out_vals, n_out, fp, swapped);
}

/* As oracle_join_aggregate on float values (bit patterns in vR / vS); each
 * row value is the result float's bit pattern. */
int oracle_join_aggregate_f32(const int32_t* R_xy, const int32_t* vR, uint64_t nR, const int32_t* S_zy,
                              const int32_t* vS, uint64_t nS, int agg, int exclude_equal, int fingerprint_only,
                              int32_t** out_pairs, int64_t** out_vals, uint64_t* n_out, orc_fp* fp, int* swapped) {
    return join_aggregate_values(R_xy, vR, nR, S_zy, vS, nS, agg, 1, exclude_equal, fingerprint_only, out_pairs,
                                 out_vals, n_out, fp, swapped);
}

static int join_aggregate_values(const int32_t* R_xy, const int32_t* vR, uint64_t nR, const int32_t* S_zy,
                                 const int32_t* vS, uint64_t nS, int agg, int f32, int exclude_equal,
                                 int fingerprint_only, int32_t** out_pairs, int64_t** out_vals, uint64_t* n_out,
                                 orc_fp* fp, int* swapped) {
    if (agg < 1 || agg > 4) return ORC_INVALID;
    const int32_t* R = R_xy; const int32_t* S = S_zy;
    int sw = 0;
    if (nR < nS) {
        const int32_t* t = R; R = S; S = t;
        t = vR; vR = vS; vS = t;
        uint64_t u = nR; nR = nS; nS = u; sw = 1;
    }
    if (swapped) *swapped = sw;
    orc_plan p;
    int rc = build_plan_v(R, vR, nR, S, vS, nS, &p);
    if (rc) return rc;
    orc_sink k;
    memset(&k, 0, sizeof(k));
    k.fingerprint_only = fingerprint_only;
    k.with_counts = 1;
    rc = !p.nx ? ORC_OK : f32 ? run_value_f32(&p, agg, exclude_equal, &k) : run_value(&p, agg, exclude_equal, &k);
    plan_free(&p);
    if (fp) *fp = k.fp;
    if (n_out) *n_out = k.fp.count;
    if (rc) { free(k.pairs); free(k.counts); return rc; }
    if (out_pairs) *out_pairs = k.pairs; else free(k.pairs);
    if (out_vals) *out_vals = k.counts; else free(k.counts);
    return ORC_OK;
}

int oracle_join_aggregate_count(const int32_t* R_xy, uint64_t nR, const int32_t* S_zy, uint64_t nS,
                                int exclude_equal, int fingerprint_only,
                                int32_t** out_pairs, int64_t** out_counts, uint64_t* n_out,
                                orc_fp* fp, int* swapped) {
    const int32_t* R = R_xy; const int32_t* S = S_zy;
    int sw = 0;
    if (nR < nS) { const int32_t* t = R; R = S; S = t; uint64_t u = nR; nR = nS; nS = u; sw = 1; }
    if (swapped) *swapped = sw;
    orc_plan p;
    int rc = build_plan(R, nR, S, nS, &p);
    if (rc) return rc;
    orc_sink k;
    memset(&k, 0, sizeof(k));
    k.fingerprint_only = fingerprint_only;
    k.with_counts = 1;
    rc = run_count(&p, exclude_equal, &k);
    plan_free(&p);
    if (fp) *fp = k.fp;
    if (n_out) *n_out = k.fp.count;
    if (rc) { free(k.pairs); free(k.counts); return rc; }
    if (out_pairs) *out_pairs = k.pairs; else free(k.pairs);
    if (out_counts) *out_counts = k.counts; else free(k.counts);
    return ORC_OK;
}

/* Exact number of join paths OUT_J = sum_b degR(b) * degS(b). */
uint64_t oracle_outj(const int32_t* R_xy, uint64_t nR, const int32_t* S_zy, uint64_t nS) {
    omap m;
    if (omap_init(&m, nR)) return 0;
    uint64_t* deg = (uint64_t*)calloc(nR + 1, sizeof(uint64_t));
    if (!deg) { omap_free(&m); return 0; }
    for (uint64_t i = 0; i < nR; i++) deg[omap_insert(&m, R_xy[2 * i + 1])]++;
    uint64_t out = 0;
    for (uint64_t i = 0; i < nS; i++) {
        int32_t t = omap_find(&m, S_zy[2 * i + 1]);
        if (t >= 0) out += deg[t];
    }
    free(deg);
    omap_free(&m);
    return out;
}

/* EstimateJoinCardinality::estimate (dim3.hpp:193-224) with the identity
 * hash myHasher (Include/structures.hpp:92-95): keep keys whose low
 * sample_bits are zero, where sample_bits halves |S| until it is at most
 * 10000 (dim3.hpp:262-269), and scale the sampled path count back up.
 * Requires nR >= nS like the reference (dim3.hpp:194). */
int64_t oracle_estimate(const int32_t* R_xy, uint64_t nR, const int32_t* S_zy, uint64_t nS) {
    int bits = 0;
    int64_t ss = (int64_t)nS;
    while (ss > 10000) { bits++; ss /= 2; }
    uint32_t mask = (1u << bits) - 1u;
    omap m;
    if (omap_init(&m, nS)) return -1;
    int64_t* cnt = (int64_t*)calloc(nS + 1, sizeof(int64_t));
    if (!cnt) { omap_free(&m); return -1; }
    for (uint64_t i = 0; i < nS; i++)
        if (((uint32_t)S_zy[2 * i + 1] & mask) == 0) cnt[omap_insert(&m, S_zy[2 * i + 1])]++;
    int64_t res = 0;
    for (uint64_t i = 0; i < nR; i++) {
        if (((uint32_t)R_xy[2 * i + 1] & mask) == 0) {
            int32_t t = omap_find(&m, R_xy[2 * i + 1]);
            if (t >= 0) res += cnt[t];
        }
    }
    free(cnt);
    omap_free(&m);
    if (res == 0) {
        /* fallback on a 1000x1000 prefix (dim3.hpp:215-221), clamped to the
         * input sizes */
        uint64_t a = nR < 1000 ? nR : 1000, b = nS < 1000 ? nS : 1000;
        omap f;
        if (omap_init(&f, a)) return -1;
        int64_t* c2 = (int64_t*)calloc(a + 1, sizeof(int64_t));
        if (!c2) { omap_free(&f); return -1; }
        for (uint64_t i = 0; i < a; i++) c2[omap_insert(&f, R_xy[2 * i + 1])]++;
        int64_t r2 = 0;
        for (uint64_t i = 0; i < b; i++) {
            int32_t t = omap_find(&f, S_zy[2 * i + 1]);
            if (t >= 0) r2 += c2[t];
        }
        free(c2);
        omap_free(&f);
        return (int64_t)((double)r2 * ((double)nR / 1e3) * ((double)nS / 1e3));
    }
    return res << bits;
}

/* gen_rand_data (dim3.hpp:1556-1568): std::default_random_engine, which
 * libstdc++ defines as minstd_rand0 (x <- 16807 x mod 2^31-1, seeded with
 * seed mod m, or 1 when that is 0).  R gets (x, y) draws, then S gets (z, y)
 * draws; each draw is engine() % MOD. */
void oracle_gen_rand_data(int32_t* R_xy, uint64_t nR, int32_t* S_zy, uint64_t nS,
                          uint32_t modx, uint32_t mody, uint32_t modz, uint32_t seed) {
    const uint64_t m = 2147483647ULL;
    uint64_t st = seed % m;
    if (st == 0) st = 1;
#define ORC_NEXT() (st = (st * 16807ULL) % m)
    for (uint64_t i = 0; i < nR; i++) {
        R_xy[2 * i] = (int32_t)(ORC_NEXT() % modx);
        R_xy[2 * i + 1] = (int32_t)(ORC_NEXT() % mody);
    }
    for (uint64_t i = 0; i < nS; i++) {
        S_zy[2 * i] = (int32_t)(ORC_NEXT() % modz);
        S_zy[2 * i + 1] = (int32_t)(ORC_NEXT() % mody);
    }
#undef ORC_NEXT
}

/* Fingerprint of a row set given as arrays (counts may be NULL). */
void oracle_fingerprint(const int32_t* pairs, const int64_t* counts, uint64_t n, orc_fp* fp) {
    memset(fp, 0, sizeof(*fp));
    for (uint64_t i = 0; i < n; i++) orc_fp_add(fp, pairs[2 * i], pairs[2 * i + 1], counts ? &counts[i] : NULL);
}

void oracle_free(void* p) { free(p); }
