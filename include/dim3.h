/*
 * libdim3 — MI355X (gfx950) DIM3 join-project and join-aggregate (COUNT, and
 * SUM / MIN / MAX / AVG of tuple values).
 *
 * C-ABI drop-in for the reference's query entry points.  Plain pointers and
 * sizes; relations are arrays of interleaved int32 pairs laid out exactly
 * like the reference's myvector<pair<int,int>> payload (Include/myvector.hpp:9-15,
 * main.cpp:6-11): R = (x, y) pairs, S = (z, y) pairs, join key in the second
 * field.  All counts are 64-bit (the reference's uint32 counters overflow
 * above ~2.1e9 rows, Include/myvector.hpp:9-15).
 *
 * Semantics kept from the reference:
 *   - join-project returns the set of distinct (x, z) with (x,y) in R and
 *     (z,y) in S, unordered (dim3.hpp:952-973);
 *   - when |R| < |S| the roles are swapped and rows are (S.first, R.first)
 *     (dim3.hpp:993-994, 1325-1326); dim3_result.swapped reports it and
 *     DIM3_OPT_NO_SWAP turns it off;
 *   - join-aggregate COUNT counts join paths with multiplicity, so duplicate
 *     tuples count once per copy (dim3.hpp:1293-1300);
 *   - empty inputs or no matching key give an empty result (dim3.hpp:1027).
 * Errors are status codes, never exit(): the reference printf+exit()s on
 * allocation failure (e.g. dim3.hpp:303-307).
 *
 * Size limit: each relation holds at most 2^30 - 1 tuples (DIM3_ERR_INVALID
 * beyond), so that every device key table fits 2^31 slots.
 *
 * Thread safety: a context serialises its calls on its own HIP stream; use
 * one context per host thread, or the server API, which is thread safe.
 */
#ifndef DIM3_H
#define DIM3_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define DIM3_OK 0
#define DIM3_ERR_INVALID 1   /* bad argument */
#define DIM3_ERR_OOM 2       /* device or host allocation failed, or over the context budget */
#define DIM3_ERR_DEVICE 3    /* HIP runtime error */
#define DIM3_ERR_CAPACITY 4  /* caller output buffer too small; dim3_result.n holds the size needed */
#define DIM3_ERR_ADMISSION 5 /* server: request larger than the HBM budget can ever admit */
#define DIM3_ERR_BUSY 6      /* server: queue full */
#define DIM3_ERR_NOT_FOUND 7 /* server: unknown ticket or relation handle */
#define DIM3_ERR_UNSUPPORTED 8
#define DIM3_ERR_CANCELLED 9   /* a streamed query stopped by its row callback */

/* ---- output modes ---- */
#define DIM3_OUT_MATERIALIZE 0 /* rows are written out */
#define DIM3_OUT_FINGERPRINT 1 /* only count + order-independent fingerprint */

/* ---- flags ---- */
#define DIM3_OPT_NO_SWAP 1u       /* never swap R and S */
#define DIM3_OPT_EXCLUDE_EQUAL 2u /* drop rows with x == z (self-join u != w filter) */
#define DIM3_OPT_HASH_ONLY 4u     /* hash every attribute (no dense-range mapping) */
/* The classical path (f1, dim3.hpp:41-50 / 273-391): join-project and
 * materialised COUNT queries of at least 8e6 input tuples with an AUTO plan
 * whose estimated OUT_J is at most 4 join paths per input tuple and 3 per x
 * row (rows bounded by R.x's value range) skip the dense mapping and CSRs
 * and sort one (x, z) key per join path instead (stats.n_tiles = 0 reports
 * it; rows in (x, z) order).  These flags force it (project and COUNT, at
 * most 2^31 - 1 join paths) or forbid it. */
#define DIM3_OPT_CLASSICAL 8u
#define DIM3_OPT_NO_CLASSICAL 16u
/* A COUNT fingerprint (without DIM3_OPT_EXCLUDE_EQUAL) is by default a sum
 * over the groups, each group's count formed on the device as the
 * reference's SpmmAggregate forms it (dim3.hpp:1267-1311).  When a few heavy
 * join keys make the (x, z) box dense, the heavy keys' counts are an int8
 * 0/1 matrix product on the MFMA cores whose int32 counts are folded into
 * the fingerprint in registers (stats.dense_mode = DIM3_DENSE_MFMA; forced
 * by dense_mode = DIM3_DENSE_MFMA, at most 64 heavy keys, fewer with
 * heavy_k).  This flag takes the join-project fingerprint of the groups and
 * sum2 / csum as sums over the join paths instead (the same numbers; no
 * group's count is formed). */
#define DIM3_OPT_PATH_SUMS 32u

/* ---- dense (heavy-key) strategy ---- */
#define DIM3_DENSE_AUTO (-1)
#define DIM3_DENSE_OFF 0    /* every key through the sparse CSR expansion */
#define DIM3_DENSE_BITOR 1  /* heavy keys as z-bitmaps OR-ed into each row (project only) */
#define DIM3_DENSE_MFMA 2   /* heavy keys as an int8 0/1 matrix product on the MFMA cores */

typedef struct dim3_opts {
    uint32_t output_mode; /* DIM3_OUT_* */
    uint32_t flags;       /* DIM3_OPT_* */
    int32_t dense_mode;   /* DIM3_DENSE_* */
    int32_t heavy_k;      /* number of heavy join keys for the dense part; 0 = planner */
    int32_t tile_bits;    /* z-tile width of the per-row dedup bitmap; 0 = planner */
    int32_t block;        /* expansion workgroup size (256/512/1024); 0 = planner */
    int32_t dense_core;   /* dense rows / columns: -1 planner, 0 every row and column with a heavy
                             key, 1 only those with many heavy keys (a dense core such as a
                             planted block; the other heavy paths stay sparse) */
    uint32_t reserved[3];
} dim3_opts;

/* Order-independent digest of a row set, with u(v) = (u32)v, mix64 the
 * splitmix64 finaliser and fmix32 the MurmurHash3 32-bit finaliser, all
 * arithmetic mod 2^64:
 *   k1(x) = mix64(u(x) ^ 0x243F6A8885A308D3) | 1, k2(x) = mix64(u(x) ^ 0x13198A2E03707344) | 1,
 *   g(z)  = a(u(z) >> 5) c(u(z) & 31), a(w) = fmix32(w ^ 0x243F6A88) | 1,
 *           c(b) = mix64(0x9E3779B97F4A7C15 (b + 1)) | 1;
 *   per project row (x, z):   count += 1, sum += k1(x) g(z), sum2 += k2(x) g(z);
 *   per COUNT row (x, z, c):  count += 1, sum += k1(x) g(z), sum2 += k2(x) c g(z), csum += c.
 * Every term factors into an x part and a z part, so the device sums the z
 * parts of a row and multiplies once per row; the z parts of one 32-aligned
 * word of z values share a(w), so a bitmap word sums with four byte-table
 * lookups.  A COUNT's sum2 and csum equal sums over join paths, which
 * DIM3_OPT_PATH_SUMS uses instead of the groups' counts. */
typedef struct dim3_fingerprint {
    uint64_t count;
    uint64_t sum;
    uint64_t sum2;
    uint64_t csum;
} dim3_fingerprint;

typedef struct dim3_stats {
    uint64_t n_r, n_s;        /* input sizes after the swap rule */
    uint64_t nx, ny, nz;      /* distinct x, joined y, z */
    uint64_t out_j;           /* join paths sum_b degR(b)*degS(b) */
    uint64_t paths_sparse;    /* join paths handled by the sparse expansion */
    uint64_t heavy_keys;      /* keys routed to the dense part */
    uint64_t nx_heavy, nz_heavy;
    uint32_t dense_mode;      /* DIM3_DENSE_* actually used */
    uint32_t tile_bits;       /* z-tile width actually used */
    uint32_t n_tiles;
    uint32_t block;
    double ms_map, ms_csr, ms_dense, ms_expand, ms_total; /* device time per phase (a COUNT fingerprint
                                                             with MFMA group counts: ms_dense is its box
                                                             kernel, inside ms_expand) */
} dim3_stats;

/* Row record of join-aggregate COUNT; same layout as the reference's
 * TUPLE<int,int,long long> {x, y, v} (dim3.hpp:20-33). */
typedef struct dim3_count_row {
    int32_t x;
    int32_t z;
    int64_t count;
} dim3_count_row;

/* Row record of a float value aggregate; same layout as the reference's
 * TUPLE<int,int,float> {x, y, v} (12 bytes). */
typedef struct dim3_value_row_f32 {
    int32_t x;
    int32_t z;
    float v;
} dim3_value_row_f32;

typedef struct dim3_result {
    uint64_t n;                 /* rows */
    int32_t* pairs;             /* project: 2*n int32 (x,z) pairs, malloc'd by the library */
    dim3_count_row* rows;       /* aggregate: n rows, malloc'd by the library */
    dim3_fingerprint fp;
    uint32_t swapped;           /* 1: rows are (S.first, R.first) */
    uint32_t reserved;
    dim3_stats stats;
    dim3_value_row_f32* rows_f32; /* float value aggregate: n rows, malloc'd by the library */
} dim3_result;

typedef struct dim3_ctx dim3_ctx;

/* Context on HIP device `device`.  hbm_budget_bytes caps the device memory
 * the context may hold (0 = 90% of the device's free memory). */
int dim3_ctx_create(int device, uint64_t hbm_budget_bytes, dim3_ctx** out);
int dim3_ctx_destroy(dim3_ctx* ctx);
/* Text of the last error on this context ("" if none). */
const char* dim3_last_error(const dim3_ctx* ctx);
const char* dim3_status_string(int status);
/* Bytes of device memory the context currently holds. */
uint64_t dim3_ctx_device_bytes(const dim3_ctx* ctx);
/* Frees every workspace buffer the context holds (they grow back on the next
 * query); the context stays usable. */
int dim3_ctx_release(dim3_ctx* ctx);
/* Fraction (0, 1] of the device's compute units the context's persistent
 * expansion grid occupies (default 1), leaving room for other streams'
 * kernels beside a long query. */
int dim3_ctx_set_cu_share(dim3_ctx* ctx, double share);

void dim3_opts_default(dim3_opts* o);

/* Replaces DIM3<int,int,int,...>::doJoinProject (dim3.hpp:982-1055) and
 * DIM3_parallel<...>::doJoinProject_parallel (dim3_parallel.hpp:137-209).
 * Host inputs, host output (res->pairs, free with dim3_result_free). */
int dim3_join_project(dim3_ctx* ctx, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy, uint64_t n_s,
                      const dim3_opts* opts, dim3_result* res);

/* Replaces DIM3JoinAggregate<int,int,int,_Tv,_Agg,...>::doJoinAggregate
 * (dim3.hpp:1313-1373) for the COUNT aggregate.  Host in, host out
 * (res->rows). */
int dim3_join_aggregate_count(dim3_ctx* ctx, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy, uint64_t n_s,
                              const dim3_opts* opts, dim3_result* res);

void dim3_result_free(dim3_result* res);

/* Device-resident variants: d_R / d_S are device pointers, the rows go to a
 * caller device buffer of cap rows (d_pairs: 2*cap int32; d_rows: cap
 * records).  If the rows do not fit, DIM3_ERR_CAPACITY is returned with
 * res->n set to the rows needed and the buffer holding the first cap rows.
 * stream: hipStream_t to order the work on, or NULL for the context's
 * stream; the call returns after the work on it has finished. */
int dim3_join_project_dev(dim3_ctx* ctx, const int32_t* d_R, uint64_t n_r, const int32_t* d_S, uint64_t n_s,
                          const dim3_opts* opts, int32_t* d_pairs, uint64_t cap, void* stream, dim3_result* res);
int dim3_join_aggregate_count_dev(dim3_ctx* ctx, const int32_t* d_R, uint64_t n_r, const int32_t* d_S,
                                  uint64_t n_s, const dim3_opts* opts, dim3_count_row* d_rows, uint64_t cap,
                                  void* stream, dim3_result* res);

/* dim3_join_project_dev without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as dim3_join_project_dev (the rows of the
 * all-sparse plan, opts->dense_mode ignored), with every workspace buffer
 * carved from a 48 MiB device array the library holds per device until the
 * process ends (its address taken at dim3_ctx_create).  At most 2^17 tuples
 * per relation (DIM3_ERR_INVALID beyond).  Calls that overlap on one device
 * share that array and race: serialise them (one context per device, or
 * one thread).  The caller passes inputs and the output buffer. */
int dim3_join_project_dev_noalloc(dim3_ctx* ctx, const int32_t* d_R, uint64_t n_r, const int32_t* d_S, uint64_t n_s,
                                  const dim3_opts* opts, int32_t* d_pairs, uint64_t cap, void* stream,
                                  dim3_result* res);

/* ---- value aggregates: SELECT x, z, AGG(R.v + S.v) GROUP BY x, z ----
 * Replaces DIM3JoinAggregate<int,int,int,_Tv,_Agg,...>::doJoinAggregate
 * (dim3.hpp:1313-1373) with a value policy over the path value
 * R.v + S.v, the combination of the shipped AggregateFunctionAvg
 * (dim3.hpp:1453-1480): R_v[i] / S_v[i] is the int32 value of tuple i.
 * Results are int64 and exact (independent of the path order): SUM, MIN,
 * MAX of R.v + S.v, and AVG = SUM / path count truncated toward zero (the
 * reference's get() on integral values).  Rows are dim3_count_row records
 * whose `count` field holds the aggregate.  The fingerprint treats the
 * aggregate as the row's count (sum2 += k2(x) v g(z), csum += v).  The swap
 * rule moves the values with their relation.  No dense part (the
 * reference's join-aggregate has none): opts->dense_mode is ignored. */
#define DIM3_AGG_SUM 1
#define DIM3_AGG_MIN 2
#define DIM3_AGG_MAX 3
#define DIM3_AGG_AVG 4
int dim3_join_aggregate(dim3_ctx* ctx, const int32_t* R_xy, const int32_t* R_v, uint64_t n_r, const int32_t* S_zy,
                        const int32_t* S_v, uint64_t n_s, int agg, const dim3_opts* opts, dim3_result* res);
int dim3_join_aggregate_dev(dim3_ctx* ctx, const int32_t* d_R, const int32_t* d_Rv, uint64_t n_r, const int32_t* d_S,
                            const int32_t* d_Sv, uint64_t n_s, int agg, const dim3_opts* opts, dim3_count_row* d_rows,
                            uint64_t cap, void* stream, dim3_result* res);
/* Float values (the reference's TUPLE<int,int,float> workload,
 * join_aggregate.cpp:21-23 with AggregateFunctionAvg<float>): the path
 * value is the float sum R.v + S.v; SUM and AVG accumulate the path values
 * in double and round once (SUM = (float)sum, AVG = (float)(sum / paths));
 * MIN / MAX order by IEEE totalOrder (-0 < +0; NaN values unsupported).
 * The double sums are exact, so the result does not depend on the path
 * order, whenever every partial sum is a double (integer-valued data below
 * 2^53, as the reference's generator makes); the reference instead adds in
 * float in first-touch order, so its AVG can differ by its own rounding
 * (relative paths * 2^-24).  Rows: res->rows_f32 / d_rows.  The fingerprint
 * counts each result float's bit pattern. */
int dim3_join_aggregate_f32(dim3_ctx* ctx, const int32_t* R_xy, const float* R_v, uint64_t n_r, const int32_t* S_zy,
                            const float* S_v, uint64_t n_s, int agg, const dim3_opts* opts, dim3_result* res);
int dim3_join_aggregate_f32_dev(dim3_ctx* ctx, const int32_t* d_R, const float* d_Rv, uint64_t n_r,
                                const int32_t* d_S, const float* d_Sv, uint64_t n_s, int agg, const dim3_opts* opts,
                                dim3_value_row_f32* d_rows, uint64_t cap, void* stream, dim3_result* res);

/* Value aggregates straight from the reference's TUPLE arrays (the payload
 * of myvector<TUPLE<int,int,_Tv>>, dim3.hpp:20-33): tuple i of R is the
 * int32 first and the int32 key at R_tuples + i * stride, and its value at
 * byte value_offset, of value_type DIM3_VALUE_I32, DIM3_VALUE_I64 (must fit
 * int32, else DIM3_ERR_INVALID) or DIM3_VALUE_F32.  The tuples are copied to
 * the device as they are and split there.  agg: DIM3_AGG_* on the values,
 * or 0 for COUNT (values ignored).  Rows: res->rows (COUNT, integer
 * values) or res->rows_f32 (float values).  Stride and offset are multiples
 * of 4, the offset at least 8.  TUPLE<int,int,float>: stride 12, offset 8;
 * TUPLE<int,int,long long>: 16, 8. */
#define DIM3_VALUE_I32 0
#define DIM3_VALUE_I64 1
#define DIM3_VALUE_F32 2
int dim3_join_aggregate_tuples(dim3_ctx* ctx, const void* R_tuples, uint64_t n_r, const void* S_tuples, uint64_t n_s,
                               uint32_t stride, uint32_t value_offset, uint32_t value_type, int agg,
                               const dim3_opts* opts, dim3_result* res);

/* ---- join paths grouped by (x, z): caller-written aggregate policies ----
 * The reference's join-aggregate takes its aggregate as a policy class
 * (SpmmAggregate<..., _Aggregate_type>, dim3.hpp:1263-1311, with the policy
 * concept of AggregateFunctionAvg, dim3.hpp:1453-1480: ctor(nz),
 * insert(z, v1, v2) on a group's first path, update(z, v1, v2) on the others,
 * get(z) at the row's end).  The library runs its own policies (COUNT and
 * the value aggregates above) on the device; for any other policy this call
 * does the join on the device and returns every join path's values grouped
 * by (x, z), for the caller's policy code to fold (dim3.hpp's
 * DIM3JoinAggregate does).  Tuples as dim3_join_aggregate_tuples (int32
 * attribute at offset 0, int32 key at 4), the value value_size bytes (4 or
 * 8) at value_offset, both multiples of value_size, value_offset >= 8.
 * Fills g: n_groups groups sorted by (x, z) (x-major, signed order), group i
 * with x[i], z[i], a dense z key zkey[i] in [0, nz) (the rank of z among the
 * distinct S attribute values; nz = their count) and its paths
 * [path_start[i], path_start[i + 1]); the paths' R values v1 and S values v2
 * (value_size bytes each), within a group in (R tuple index, S tuple index)
 * order, as the reference visits them.  opts: exclude_equal drops paths with
 * x == z; at most 2^31 - 1 join paths.  Free with dim3_grouped_free. */
typedef struct {
    uint64_t n_groups, n_paths;
    uint32_t nz, value_size;
    int32_t* x;
    int32_t* z;
    uint32_t* zkey;
    uint64_t* path_start; /* n_groups + 1 entries */
    void* v1;
    void* v2;
} dim3_grouped;
int dim3_join_paths_grouped(dim3_ctx* ctx, const void* R_tuples, uint64_t n_r, const void* S_tuples, uint64_t n_s,
                            uint32_t stride, uint32_t value_offset, uint32_t value_size, const dim3_opts* opts,
                            dim3_grouped* g);
void dim3_grouped_free(dim3_grouped* g);

/* ---- streamed output: rows handed to the caller in chunks ----
 * For outputs larger than any one buffer (the chunked STREAM output mode).
 * The query runs in chunks of consecutive x rows, each producing at most
 * max(chunk_rows, the largest row's output) rows (0 = 2^24); after each chunk
 * `cb` gets n rows in pinned host memory that stays valid during the call:
 * project 2n int32 (x, z) pairs, COUNT n dim3_count_row.  Rows are distinct
 * across chunks and in no particular order (as dim3_join_project's); a
 * nonzero return from `cb` stops the query with DIM3_ERR_CANCELLED.  res
 * carries n (all rows), swapped and the stats; no rows.  opts->output_mode is
 * ignored.  No counterpart in the reference (its result is one in-memory
 * myvector, dim3.hpp:982-989). */
typedef int (*dim3_rows_cb)(void* user, const void* rows, uint64_t n);
int dim3_join_project_stream(dim3_ctx* ctx, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy, uint64_t n_s,
                             const dim3_opts* opts, uint64_t chunk_rows, dim3_rows_cb cb, void* user,
                             dim3_result* res);
int dim3_join_aggregate_count_stream(dim3_ctx* ctx, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy,
                                     uint64_t n_s, const dim3_opts* opts, uint64_t chunk_rows, dim3_rows_cb cb,
                                     void* user, dim3_result* res);

/* Streamed value aggregates (as dim3_join_aggregate / _f32, rows in chunks
 * as above: n dim3_count_row, or n dim3_value_row_f32 for float values). */
int dim3_join_aggregate_stream(dim3_ctx* ctx, const int32_t* R_xy, const int32_t* R_v, uint64_t n_r,
                               const int32_t* S_zy, const int32_t* S_v, uint64_t n_s, int agg, const dim3_opts* opts,
                               uint64_t chunk_rows, dim3_rows_cb cb, void* user, dim3_result* res);
int dim3_join_aggregate_f32_stream(dim3_ctx* ctx, const int32_t* R_xy, const float* R_v, uint64_t n_r,
                                   const int32_t* S_zy, const float* S_v, uint64_t n_s, int agg,
                                   const dim3_opts* opts, uint64_t chunk_rows, dim3_rows_cb cb, void* user,
                                   dim3_result* res);

/* ---- batches: many independent instances in one persistent launch ----
 * No counterpart in the reference (it runs one query per call); the batch
 * returns, per instance, exactly what dim3_join_project_dev /
 * dim3_join_aggregate_count_dev return for that instance alone. */
typedef struct dim3_instance {
    const int32_t* R_xy; /* device (x, y) pairs */
    uint64_t n_r;
    const int32_t* S_zy; /* device (z, y) pairs */
    uint64_t n_s;
    void* out;           /* device rows: 2*cap int32 pairs / cap dim3_count_row; NULL in FINGERPRINT mode */
    uint64_t cap;
} dim3_instance;

/* results[i] is instance i's result (rows written to inst[i].out).  Returns
 * DIM3_ERR_CAPACITY when some instance's rows exceed its cap (its
 * results[i].n holds the rows needed).  The context keeps every instance's
 * workspace (one bank per instance) for reuse by the next batch.
 * With stream == NULL and 16 or more instances, the larger half of the batch
 * (by tuples) expands on a context-owned stream masked to 7/8 of the compute
 * units while the rest is planned; the call returns after both halves end.
 * On a caller's stream the whole batch runs on that stream. */
int dim3_batch_join_project_dev(dim3_ctx* ctx, const dim3_instance* inst, uint32_t n, const dim3_opts* opts,
                                void* stream, dim3_result* results);
int dim3_batch_join_aggregate_count_dev(dim3_ctx* ctx, const dim3_instance* inst, uint32_t n, const dim3_opts* opts,
                                        void* stream, dim3_result* results);

/* Exact OUT_J = sum_b degR(b)*degS(b) on the device (the quantity
 * EstimateJoinCardinality samples, dim3.hpp:190-270). */
int dim3_count_join_paths(dim3_ctx* ctx, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy, uint64_t n_s,
                          uint64_t* out_j);

/* EstimateJoinCardinality<...>::estimate (dim3.hpp:190-270) on the device:
 * the reference's sampled join size, the same number (identity hash
 * myHasher: keys with their low b bits zero, b the halvings that bring |S|
 * to at most 10000; the sampled path count << b; with no sampled path, the
 * paths between the first 1000 tuples of each side scaled by
 * |R|/1e3 * |S|/1e3, prefixes clamped to the input sizes).  The swap rule
 * applies first (the reference asserts |R| >= |S|).  Host inputs. */
int dim3_estimate_join_cardinality(dim3_ctx* ctx, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy,
                                   uint64_t n_s, int64_t* estimate);
/* The same over device-resident relations, on `stream` (NULL: the
 * context's stream); returns once the estimate is on the host.  The server
 * sizes a materialising request's output with it at submit. */
int dim3_estimate_join_cardinality_dev(dim3_ctx* ctx, const int32_t* d_R, uint64_t n_r, const int32_t* d_S,
                                       uint64_t n_s, void* stream, int64_t* estimate);

/* The same over the reference's TUPLE arrays (EstimateJoinCardinality's
 * TUPLE overload, dim3.hpp:226-260): tuple i is the int32 first and the
 * int32 key at R_tuples + i * stride (stride a multiple of 4, at least 8);
 * values are ignored.  Host inputs, copied as they are and split on the
 * device. */
int dim3_estimate_join_cardinality_tuples(dim3_ctx* ctx, const void* R_tuples, uint64_t n_r, const void* S_tuples,
                                          uint64_t n_s, uint32_t stride, int64_t* estimate);

/* The dense part's product on its own (the boolean product that
 * Hybrid_solution's DENSE part evaluates one (x, dense z) pair at a time,
 * dim3.hpp:894-950): C = A * Bt^T over int8 on the MFMA cores with int32
 * accumulation.  A is M x K, Bt is N x K, both row-major host arrays.
 * bits (nullable): M rows of ceil(N / 32) uint32 words, bit (n % 32) of word
 * n / 32 set iff C[m][n] > 0; counts (nullable): M x N int32 row-major.
 * K <= 131072 keeps every entry exact; M < 8388480. */
int dim3_dense_product(dim3_ctx* ctx, const int8_t* A, uint64_t M, const int8_t* Bt, uint64_t N, uint64_t K,
                       uint32_t* bits, int32_t* counts);

/* ---- serving: a thread-safe request loop over resident relations ----
 * No counterpart in the reference (a library of blocking calls).  Relations
 * are copied once into server-owned device memory and named by handles.
 * Requests are admitted by HBM footprint: one whose estimated workspace can
 * never fit next to the resident relations is refused at submit
 * (DIM3_ERR_ADMISSION); others wait until memory and a worker are free.
 * Each worker owns a context (stream and workspace), so admitted requests
 * run concurrently; a worker that finds several fingerprint requests of the
 * same kind and options waiting runs them as one batch
 * (dim3_batch_join_*_dev).  A request's result is the same as the
 * single-query call's. */
typedef struct dim3_server dim3_server;

typedef struct dim3_server_opts {
    int device;
    uint64_t hbm_budget_bytes; /* 0 = 90% of the device's free memory */
    uint32_t workers;          /* concurrent requests (0 = 4) */
    uint32_t max_batch;        /* requests per batch (0 = 32; 1 = no batching) */
    uint32_t queue_cap;        /* pending requests before DIM3_ERR_BUSY (0 = 4096) */
    uint32_t large_share_pct;  /* with 2+ workers, worker 0 is the lane of large requests: they run
                                  on this % of the compute units (a CU-masked stream) and the other
                                  workers on the rest while one runs (0 = 90) */
    uint64_t large_tuples;     /* a request with at least this many input tuples is large (0 = 2e7) */
} dim3_server_opts;

#define DIM3_REQ_PROJECT 0
#define DIM3_REQ_COUNT 1
#define DIM3_REQ_VALUE 2 /* value aggregate `agg` over the relations' values (dim3_server_put_values) */

typedef struct dim3_request {
    uint32_t kind;     /* DIM3_REQ_* */
    uint32_t agg;      /* DIM3_REQ_VALUE: DIM3_AGG_*; otherwise 0 */
    uint64_t r, s;     /* relation handles: R (x, y) and S (z, y) */
    dim3_opts opts;
} dim3_request;

typedef struct dim3_server_stats {
    uint64_t submitted, completed, rejected, batches, batched_requests;
    uint64_t resident_bytes;   /* relations */
    uint64_t workspace_bytes;  /* held by the workers' contexts */
    uint64_t peak_bytes;
} dim3_server_stats;

void dim3_server_opts_default(dim3_server_opts* o);
int dim3_server_create(const dim3_server_opts* opts, dim3_server** out);
/* Waits for the running requests, then frees everything. */
int dim3_server_destroy(dim3_server* srv);
/* Copies n (x|z, y) pairs (host or device memory) into the server. */
int dim3_server_put_relation(dim3_server* srv, const int32_t* pairs, uint64_t n, uint64_t* handle);
int dim3_server_put_relation_dev(dim3_server* srv, const int32_t* d_pairs, uint64_t n, uint64_t* handle);
/* Attaches one value per tuple (host or device memory; DIM3_VALUE_I32 or
 * DIM3_VALUE_F32) to a relation, once, before any request uses it; value
 * requests (DIM3_REQ_VALUE, not batched) need values of one type on both
 * relations; their rows are dim3_count_row (int) or dim3_value_row_f32. */
int dim3_server_put_values(dim3_server* srv, uint64_t handle, const void* values, uint32_t value_type);
/* Frees a relation once no pending request uses it. */
int dim3_server_drop_relation(dim3_server* srv, uint64_t handle);
int dim3_submit(dim3_server* srv, const dim3_request* req, uint64_t* ticket);
/* *done = 1 when the request has finished. */
int dim3_poll(dim3_server* srv, uint64_t ticket, int* done);
/* Blocks until the request finishes; returns its status and result (rows in
 * host memory for MATERIALIZE, freed with dim3_result_free).  The ticket is
 * released.  timing (nullable): seconds from submit to start and to end. */
int dim3_wait(dim3_server* srv, uint64_t ticket, dim3_result* res, double* queued_s, double* total_s);
int dim3_server_get_stats(dim3_server* srv, dim3_server_stats* st);

/* ---- cost-model calibration (autotuner) ----
 * The planner prices the dense part against the sparse expansion with device
 * rates.  dim3_autotune measures the context's device with short synthetic
 * queries (rows reaching 32 keys of scattered z values against rows reaching
 * one key with the same output, all-sparse and MFMA-seeded) plus an int8
 * product and a streaming fill, and installs the rates in the context: the
 * product and fill rates as measured, the path and seeding rates as the
 * defaults (fitted on real workloads on MI355X) scaled by the device's speed
 * on the synthetic queries relative to MI355X's.  It replaces the reference's offline
 * constants (server_parameters.hpp, produced by server_parameter_estimator.cpp).
 * cache_path (nullable): a text file; read when it holds a calibration of the
 * same device model, else written after measuring. */
typedef struct dim3_tuning {
    double path_ns;      /* sparse expansion per join path */
    double seed_bit_ns;  /* dense seeding per product bit */
    double seed_word_ns; /* dense seeding per 32-bit word copied */
    double mfma_tops;    /* int8 product, tera-ops/s */
    double bitor_gbs;    /* bit-or product, GB/s of key-bitmap reads */
    double hbm_gbs;      /* streaming writes, GB/s */
    uint32_t source;     /* 0 static defaults, 1 measured, 2 read from cache_path */
    uint32_t reserved;
} dim3_tuning;

int dim3_autotune(dim3_ctx* ctx, const char* cache_path, dim3_tuning* out);
int dim3_get_tuning(const dim3_ctx* ctx, dim3_tuning* out);
int dim3_set_tuning(dim3_ctx* ctx, const dim3_tuning* in);

/* Library version string. */
const char* dim3_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DIM3_H */
