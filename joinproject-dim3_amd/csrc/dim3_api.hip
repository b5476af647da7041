// libdim3 C-ABI: context, workspace, planner and the query pipeline.
//
// Query pipeline (one HIP stream per context, one host sync after mapping
// to size the CSR arrays and pick the tile):
//   swap rule                         dim3.hpp:993-994 / 1325-1326
//   y / x / z mapping + semi-join     dim3.hpp:399-455      (k_mapping.hip)
//   CSR of R by x, of S by (y,tile)   dim3.hpp:1058-1253    (k_mapping.hip)
//   sparse expansion + dedup/count    dim3.hpp:952-973, 1267-1311 (k_expand.hip)
// Device buffers live in per-context grow-only workspace slots, so repeated
// queries of similar size allocate nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "dim3_internal.h"
#include "dim3_sizes.h"

using dim3h::Scalars;

// Device workspace of the allocation-free entry point
// (dim3_join_project_dev_noalloc): one array per device, held until the
// process ends.  In host code `scratch` only names the symbol; its device
// address comes from hipGetSymbolAddress, once per device (scratch_address).
constexpr size_t kScratchBytes = 48ull << 20;
constexpr uint64_t kNoAllocMaxTuples = 1ull << 17;  // per relation: the workspace then fits the array
// project tiles: blocked bit layout, blocks of 2^kTileBlockLog words (32,
// 64 and 128 measured alike on the two-hop graph)
constexpr uint32_t kTileBlockLog = 5;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

// The device address of `scratch` on the current device (nullptr if HIP
// cannot give it), looked up once per device.
uint8_t* scratch_address(int device) {
    static std::mutex mu;
    static uint8_t* addr[64] = {};
    static bool done[64] = {};
    if (device < 0 || device >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!done[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
        }
        addr[device] = (uint8_t*)p;
        done[device] = true;
    }
    return addr[device];
}

}  // namespace

namespace {

enum SlotId {
    WS_IN_R,
    WS_IN_S,
    WS_KEYS_Y,
    WS_DEGR_Y,
    WS_DEGS_Y,
    WS_EX_Y,
    WS_SLOT_R,
    WS_SLOT_S,
    WS_KEYS_X,
    WS_DEG_X,
    WS_EX_X,
    WS_SLOT_RX,
    WS_KEYS_Z,
    WS_DEG_Z,
    WS_EX_Z,
    WS_SLOT_SZ,
    WS_PARTIALS,
    WS_SCALARS,
    WS_AGG_KR,
    WS_AGG_GS,
    WS_ID2X,
    WS_ID2Z,
    WS_X2Z,
    WS_ROWPTR_R,
    WS_CUR_R,
    WS_COL_R,
    WS_CNT_S,
    WS_ROWPTR_S,
    WS_CUR_S,
    WS_KEY_S,
    WS_ZID_S,
    WS_COL_S,
    WS_IN_VR,
    WS_IN_VS,
    WS_VAL_R,
    WS_VAL_S,
    WS_OUT,
    WS_ROW_BOUND,
    WS_VT_FW,
    WS_VT_PW,
    WS_VT_FH,
    WS_VT_PH,
    WS_VT_WAVE,
    WS_VT_HASH,
    WS_VT_BIG,
    WS_ROWPTR_Y,
    WS_VT_ROWN,
    WS_VT_ROWOFF,
    WS_XW,
    WS_ZW,
    WS_DHIST,
    WS_HID,
    WS_H2Y,
    WS_HDX,
    WS_HDZ,
    WS_CNT_SD,
    WS_CUR_SD,
    WS_HX,
    WS_HX2X,
    WS_ZFLAG,
    WS_HZPOS,
    WS_HZ2ZID,
    WS_HZTILE,
    WS_DA,
    WS_DB,
    WS_DBITS,
    WS_CINT,
    WS_GC_HM,
    WS_GC_ZM,
    WS_GC_XFLAG,
    WS_GC_ZFLAG,
    WS_GC_XPOS,
    WS_GC_ZPOS,
    WS_GC_RLIST,
    WS_GC_CLIST,
    WS_GC_A8,
    WS_GC_B8,
    WS_GC_XK,
    WS_GC_GZ,
    WS_GC_G,
    WS_GC_ZINFO,
    WS_GC_CNTL,
    WS_GC_ROWL,
    WS_GC_CURL,
    WS_GC_COLL,
    WS_FP_ZW,  // z word factors of a fingerprint over a dense z range
    WS_BATCH_ARGS,
    WS_BATCH_START,
    WS_BATCH_SCL,
    WS_BATCH_SCOUT,
    WS_COUNT
};

struct Slot {
    void* p = nullptr;
    size_t bytes = 0;
};

using dim3s::pow2_at_least;

}  // namespace

namespace {

// Plan of one prepared query (the host-side decisions and sizes).
struct QueryPlan {
    const int2* R = nullptr;
    const int2* S = nullptr;
    uint64_t nR = 0, nS = 0;
    bool swapped = false;
    bool empty = true;
    bool count_mode = false;  // expansion with counters (false also for a COUNT fingerprint, see agg_fp)
    bool agg_fp = false;      // COUNT fingerprint: groups from the project expansion, sum2 / csum from path sums
    int agg = 0;              // value aggregate DIM3_AGG_* (0: none) over the tuple values vR, vS
    bool vf32 = false;        // float values
    const int32_t* vR = nullptr;
    const int32_t* vS = nullptr;
    uint32_t flags = 0;
    uint32_t nx = 0, ny = 0, nz = 0;
    uint32_t nnzR = 0, nnzS = 0;
    uint64_t out_j = 0;
    uint32_t ntiles = 1, tile_size = 0, tile_words = 0;
    bool x_range = false, y_range = false, z_ident = false;
    bool hot = false;  // Scalars::hot holds the hot-value samples (large inputs)
    bool fp_zw = false;  // WS_FP_ZW holds the factors of the tiles' z words (fingerprint, dense z)
    int32_t zlo = 0;
    dim3h::KeyMap mx{}, my{}, mz{};
    uint32_t touched_cap = 0, stage_cap = 0;
    int block = 256;
    float ms_map = 0, ms_csr = 0, ms_dense = 0;
    // dense split
    int dense = DIM3_DENSE_OFF;  // mode used
    uint32_t bmin = 0;           // heaviness bin threshold
    uint32_t K = 0, Kp = 0, Mh = 0, Nh = 0, dwords = 0;
    uint64_t z_span = 0;  // dense z IDs (range mapping): the 32-aligned span of z values
    bool all_z = false;  // every z ID is a product column
    bool fused = false;  // bit-or product read from the key bitmaps by the seeding (no product rows)
    bool plain = false;  // project tiles in the plain bit layout (word seeding / word fingerprints), else swizzled
    // dense core: rows / columns with at least thx / thz heavy entries
    bool core_try = false;     // analyse the core of the keys with bin >= bc
    uint32_t bc = 0;
    uint64_t core_K = 0, core_paths = 0;
    double p1_gain = -1e300;   // saving of the row-and-column-complete choice (phase 1)
    bool core = false;         // rows outside the core keep heavy entries (split S rows)
    uint32_t thx = 1, thz = 1;
    // COUNT fingerprint with every group's count: heavy keys (bin >= gc_bmin,
    // gc_K <= 64) as an MFMA product over the box of gc_Mh rows and gc_Nh
    // columns (gc_Mp rows padded), light keys by (x, z-tile) wavefront tasks
    bool gc = false;
    bool gc_dup = false;  // a heavy tuple occurs twice (multiplicity > 1): not this path
    uint32_t gc_bmin = 0, gc_K = 0, gc_Mh = 0, gc_Nh = 0, gc_Mp = 0;
};

// The tuple values of a value-aggregate query (agg 0: none).
struct ValueSpec {
    int agg = 0;
    bool f32 = false;  // float values (bit patterns), float results
    const int32_t* vR = nullptr;
    const int32_t* vS = nullptr;
};

// One query's device workspace and plan.  A context holds one bank per
// query that must stay prepared at the same time (a batch holds one per
// instance); slots grow and are reused across calls.
struct Bank {
    Slot ws[WS_COUNT];
    QueryPlan plan;
    // pinned host mirrors of this query's scalars and dense-split histogram
    // (a batch plans all its instances with one synchronisation per stage)
    Scalars* hsc = nullptr;
    unsigned long long* hhist = nullptr;
    ~Bank() {
        if (hsc) dim3h::host_free(hsc);
        if (hhist) dim3h::host_free(hhist);
    }
};

// Device rates of the planner's cost model.  The defaults are static
// estimates; dim3_autotune replaces them with rates measured on the device.
struct DenseModel {
    double path_s = 0.9e-12;       // sparse expansion, per join path
    double seed_s = 1.0e-12;       // seeding, per product bit or count (a gathered z ID + one LDS op)
    double seed_word_s = 1.0e-12;  // seeding, per 32-bit word copied
    double mfma_ops = 1.5e15;      // int8 ops/s reached by k_gemm_i8
    double hbm_bps = 5.0e12;       // streaming bytes/s
    double l2_bps = 10.0e12;       // bit-or re-reads of key bitmaps (mostly L2 hits)
    double launch_s = 60e-6;       // fixed cost of the dense launches
};

}  // namespace

struct dim3_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t qstream = nullptr;  // stream of the call in progress (workspace allocations are ordered on it)
    uint64_t budget = 0;
    uint64_t held = 0;
    std::vector<std::unique_ptr<Bank>> banks;
    uint32_t cur = 0;  // bank of the query being prepared / expanded
    std::string err;
    int cu_count = 0;
    double cu_share = 1.0;  // fraction of the compute units the expansion grid may occupy
    int grid_cus() const { return std::max(1, (int)(cu_count * cu_share)); }
    size_t max_lds = 0;
    Scalars* h_sc = nullptr;  // pinned mirror of the device scalars
    unsigned long long* h_hist = nullptr;  // pinned mirror of the dense-split histogram
    Scalars* h_batch = nullptr;            // pinned scalars of a batch
    uint32_t h_batch_cap = 0;
    void* h_stream = nullptr;              // pinned chunk buffer of a streamed query
    size_t h_stream_bytes = 0;
    hipEvent_t ev[8] = {};
    std::recursive_mutex mu;
    DenseModel model;           // cost-model rates (defaults or measured)
    uint32_t model_source = 0;  // 0 defaults, 1 measured, 2 read from a cache file
    // allocation-free queries (dim3_join_project_dev_noalloc): this device's
    // copy of the library's scratch array and the bank whose slots are
    // carved from it (pinned mirrors made at creation); arena != nullptr
    // while such a query runs
    uint8_t* scratch = nullptr;
    std::unique_ptr<Bank> nab;
    // the batch launch's instantiated graph, replayed while its shape and
    // buffers are those of the call (captured again when they change)
    hipGraphExec_t batch_exec[2] = {nullptr, nullptr};  // chunk slots of a pipelined batch
    uint64_t batch_key[2][12] = {};
    // a pipelined batch's first chunk runs on this stream, masked to pipe_cus
    // compute units (created on first use)
    hipStream_t pipe_stream = nullptr;
    int pipe_cus = 0;
    uint8_t* arena = nullptr;
    size_t arena_used = 0, arena_size = 0;

    Bank& bank() { return arena ? *nab : *banks[cur]; }
    void use_bank(uint32_t b) {
        while (banks.size() <= b) banks.emplace_back(new Bank());
        cur = b;
    }
};

namespace {

}  // namespace

std::shared_mutex& dim3h::capture_mutex() {
    static std::shared_mutex mu;
    return mu;
}

// EstimateJoinCardinality over device relations (defined below)
static int estimate_on(dim3_ctx* c, const int2* dR, uint64_t nR, const int2* dS, uint64_t nS, hipStream_t st,
                       int64_t* estimate);

namespace {

int fail(dim3_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, x)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail((c), DIM3_ERR_DEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHK(c, what)                                                                       \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) return fail((c), DIM3_ERR_DEVICE, "%s launch: %s", what, hipGetErrorString(e_)); \
    } while (0)

// Grow-only device workspace slot.
int ws_get(dim3_ctx* c, SlotId id, size_t bytes, void** out) {
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~(size_t)255;
    Slot& s = c->bank().ws[id];
    if (s.bytes >= bytes) {
        *out = s.p;
        return DIM3_OK;
    }
    if (c->arena) {
        // an allocation-free query: slots are carved from the scratch array
        // (a slot that must grow within the call takes a new piece)
        if (bytes > c->arena_size - c->arena_used)
            return fail(c, DIM3_ERR_OOM, "workspace slot %d needs %zu bytes: the %zu-byte scratch array is full",
                        (int)id, bytes, c->arena_size);
        s.p = c->arena + c->arena_used;
        s.bytes = bytes;
        c->arena_used += bytes;
        *out = s.p;
        return DIM3_OK;
    }
    const uint64_t need = c->held - s.bytes + bytes;
    if (c->budget && need > c->budget)
        return fail(c, DIM3_ERR_OOM, "workspace slot %d needs %zu bytes: over the %llu-byte budget", (int)id, bytes,
                    (unsigned long long)c->budget);
    // stream-ordered: no device-wide synchronisation, so other contexts'
    // queries keep running while this one grows its workspace
    if (s.p) {
        dim3h::dev_free_async(s.p, c->qstream);
        c->held -= s.bytes;
        s.p = nullptr;
        s.bytes = 0;
    }
    void* p = nullptr;
    hipError_t e = dim3h::dev_malloc_async(&p, bytes, c->qstream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, DIM3_ERR_OOM, "hipMalloc(%zu) for workspace slot %d: %s", bytes, (int)id, hipGetErrorString(e));
    }
    s.p = p;
    s.bytes = bytes;
    c->held += bytes;
    *out = p;
    return DIM3_OK;
}

template <typename T>
int ws(dim3_ctx* c, SlotId id, uint64_t n, T** out) {
    void* p = nullptr;
    int rc = ws_get(c, id, (size_t)n * sizeof(T), &p);
    *out = (T*)p;
    return rc;
}

#define TRY(x)                    \
    do {                          \
        int rc_ = (x);            \
        if (rc_ != DIM3_OK) return rc_; \
    } while (0)

// ---------------------------------------------------------------------------
// planner: z-tile width and workgroup size of the expansion
// ---------------------------------------------------------------------------
// force_wgs > 0: only the candidate with that many workgroups per CU (a
// batch shares one workgroup shape and LDS size across its instances).
void plan_tiles(dim3_ctx* c, const dim3_opts& o, int force_wgs = 0) {
    auto& P = c->bank().plan;
    const bool cnt = P.count_mode;
    const uint32_t per_word = cnt ? 1u : 32u;  // z IDs per LDS word
    // a value tile holds an int64 accumulator and a u32 counter per z ID
    const uint64_t wpu = P.agg ? 3u : 1u;
    P.touched_cap = cnt ? 4096u : 2048u;
    const uint64_t nz_words = ((uint64_t)P.nz + per_word - 1) / per_word;
    // static shared + slack, and the fingerprint table of fingerprint mode
    // (the value kernel computes its z factors instead)
    const uint64_t fp_words = o.output_mode == DIM3_OUT_FINGERPRINT && !P.agg ? 2u * dim3d::kFpTableU64 : 0u;
    const uint64_t lds_words = c->max_lds / 4u - 256u;
    struct Cand {
        int block;
        uint32_t ntiles, tile_words, stage;
        bool ok;
        int wgs;
    };
    // candidates: 1, 2 or 4 workgroups per CU (16 wavefronts per CU in all)
    auto make = [&](int wgs, int block, uint64_t forced_units) {
        Cand k{block, 1, 0, 0, false, wgs};
        const uint64_t budget = lds_words / (uint64_t)wgs - fp_words;
        const uint64_t min_stage =
            P.agg ? 4u * (uint64_t)block + 1u : std::max<uint64_t>(1024u, 3u * (uint64_t)block + 1u);
        if (budget <= P.touched_cap + min_stage + wpu) return k;
        uint64_t max_words = (budget - P.touched_cap - min_stage) / wpu;
        uint64_t tw;
        if (forced_units) {
            tw = (forced_units + per_word - 1) / per_word;
        } else {
            const uint64_t nt = (nz_words + max_words - 1) / max_words;
            tw = (nz_words + nt - 1) / nt;
        }
        if (tw == 0) tw = 1;
        // project tiles: multiples of the block width (the blocked bit
        // layout) where they fit and the caller did not fix a width that is
        // not one, else 16-byte tile rows
        if (!cnt) {
            const uint64_t bw = 1ull << kTileBlockLog;
            const bool blk = ((tw + bw - 1) & ~(bw - 1)) <= max_words && (!forced_units || forced_units % (32u * bw) == 0);
            const uint64_t r = blk ? bw - 1 : 3u;
            if (((tw + r) & ~r) <= max_words) tw = (tw + r) & ~r;
        }
        if (tw > max_words) return k;
        k.tile_words = (uint32_t)tw;
        k.ntiles = (uint32_t)(((uint64_t)P.nz + tw * per_word - 1) / (tw * per_word));
        if (k.ntiles == 0) k.ntiles = 1;
        k.stage = (uint32_t)(budget - P.touched_cap - tw * wpu);
        k.ok = true;
        return k;
    };
    const uint64_t forced = o.tile_bits > 0 ? (uint64_t)o.tile_bits : 0;
    Cand best{0, 0, 0, 0, false, 1};
    const int wg_opts[3] = {4, 2, 1};
    const int blk_opts[3] = {256, 512, 1024};
    for (int i = 0; i < 3; i++) {
        if (force_wgs > 0 && wg_opts[i] != force_wgs) continue;
        const int block = o.block ? o.block : blk_opts[i];
        Cand k = make(wg_opts[i], block, forced);
        if (!k.ok) continue;
        // Two or more workgroups per CU overlap one row's emission with
        // another's expansion, which pays for up to twice the tiles (each
        // tile re-reads the row's y list and its S bounds): score ntiles /
        // min(workgroups per CU, 2), ties to more workgroups.
        const double score = (double)k.ntiles / (double)std::min(wg_opts[i], 2);
        const double best_score = best.ok ? (double)best.ntiles / (double)std::min(best.wgs, 2) : 1e300;
        if (score < best_score) best = k;
    }
    if (!best.ok && force_wgs > 0) {  // forced tile too big: the batch shape with the planner's tile
        best = make(force_wgs, o.block ? o.block : blk_opts[force_wgs == 4 ? 0 : force_wgs == 2 ? 1 : 2], 0);
    }
    if (!best.ok) {  // forced tile too big for the candidates: one workgroup per CU
        best = make(1, o.block ? o.block : 1024, 0);
    }
    P.block = best.block;
    P.ntiles = best.ntiles;
    P.tile_words = best.tile_words;
    P.tile_size = best.tile_words * per_word;
    P.stage_cap = best.stage;
}

// ---------------------------------------------------------------------------
// planner: dense split.  From the heaviness histogram (per bin: keys, join
// paths, sum degR, sum degS, x rows and z columns whose heaviest key falls in
// it) every threshold bin b gives the exact product shape of "keys with
// w >= lo(b) are heavy": K keys, Mh rows, Nh columns, P_h paths.  The cost
// model (device rates) prices the sparse work those paths would take against
// the dense product and its seeding; AUTO takes the best positive saving.
// ---------------------------------------------------------------------------
struct DenseShape {
    uint64_t K = 0, Mh = 0, Nh = 0, paths = 0, sumR = 0, sumS = 0;
    uint64_t nz = 0;  // all z IDs: product columns covering nearly all of them seed whole words
    uint64_t tile_words = 0, ntiles = 0;  // project tile shape: a fused key bitmap holds ntiles * tile_words words
};

// LDS bandwidth of a fused seed's tile updates (one read-or-write of every
// seeded tile word, all compute units)
constexpr double kLdsBps = 50e12;

// Seconds the dense part takes for one shape and mode, <0 if it cannot run.
// Join-project bit-or has two forms: fused (any_fused, or the columns are
// nearly all z IDs) keeps one bitmap per heavy key over every z ID, in the
// expansion tile's layout, and each row tile ORs its keys' words straight
// into LDS (no product rows); the materialised form writes product rows over
// the columns only and seeds them bit by bit.  *fused_out reports the form.
double dense_cost(const DenseModel& m, const DenseShape& d, int mode, bool count, uint64_t mem_room,
                  uint64_t* bytes_out, bool any_fused = false, bool* fused_out = nullptr) {
    if (fused_out) *fused_out = false;
    if (d.K == 0 || d.Mh == 0 || d.Nh == 0) return -1;
    const uint64_t Kp = (d.K + 63) / 64 * 64;
    const uint64_t dwords = (d.Nh + 31) / 32;
    const double cells = (double)d.Mh * (double)d.Nh;
    const double set_bits = std::min(cells, (double)d.paths);
    uint64_t bytes = 0;
    double t = 0;
    if (mode != DIM3_DENSE_MFMA && mode != DIM3_DENSE_BITOR) return -1;
    const bool words = !count && d.Nh * 10 >= d.nz * 9;
    if (mode == DIM3_DENSE_MFMA) {
        if (d.Mh >= 65535ull * 128ull || d.Nh >= (1ull << 31)) return -1;
        if (count && Kp > 131072) return -1;  // int32 accumulator bound for multiplicities up to 127
        bytes = (d.Mh + d.Nh) * Kp + (count ? d.Mh * d.Nh * 4 : d.Mh * dwords * 4);
        t = 2.0 * cells * (double)Kp / m.mfma_ops + (double)((d.Mh + d.Nh) * Kp) / m.hbm_bps;
        if (bytes > mem_room) return -1;
        // product written once and read back by the seeding, plus one LDS op
        // per set entry (per word when the columns are nearly all z IDs: word
        // copies)
        t += 2.0 * (count ? cells * 4.0 : cells / 8.0) / m.hbm_bps;
        t += (words ? cells / 32.0 * m.seed_word_s : set_bits * m.seed_s) + m.launch_s;
        if (bytes_out) *bytes_out = bytes;
        return t;
    }
    if (count) return -1;
    double tf = -1, tm = -1;
    uint64_t bf = 0, bm = 0;
    if (words || any_fused) {
        // fused: every row tile of a row with heavy keys ORs its keys' words
        // (L2 / MALL reads) into its LDS tile
        const uint64_t fw = std::max<uint64_t>(d.ntiles * d.tile_words, (d.nz + 31) / 32);
        bf = d.K * fw * 4;
        tf = (double)d.sumR * (double)fw * 4.0 / m.l2_bps + (double)bf / m.hbm_bps +
             (double)d.Mh * (double)fw * 4.0 / kLdsBps + m.launch_s;
        if (bf > mem_room) tf = -1;
    }
    if (!words) {
        // materialised product rows over the columns, seeded bit by bit
        bm = d.K * dwords * 4 + d.Mh * dwords * 4;
        tm = (double)d.sumR * (double)dwords * 4.0 / m.l2_bps + (double)(d.K * dwords * 4) / m.hbm_bps +
             2.0 * cells / 8.0 / m.hbm_bps + set_bits * m.seed_s + m.launch_s;
        if (bm > mem_room) tm = -1;
    }
    const bool fused = tf >= 0 && (tm < 0 || tf <= tm);
    t = fused ? tf : tm;
    bytes = fused ? bf : bm;
    if (t < 0) return -1;
    if (fused_out) *fused_out = fused;
    if (bytes_out) *bytes_out = bytes;
    return t;
}

void plan_dense(dim3_ctx* c, const dim3_opts& o) {
    auto& P = c->bank().plan;
    P.dense = DIM3_DENSE_OFF;
    if (o.dense_mode == DIM3_DENSE_OFF) return;
    const unsigned long long* h = c->bank().hhist;
    const uint32_t NB = dim3h::kDenseBins;
    const DenseModel& m = c->model;
    const uint64_t room = c->budget > c->held ? (c->budget - c->held) / 3 : 0;
    const bool count = P.count_mode;
    // suffix sums from the heaviest bin down
    DenseShape acc;
    acc.nz = P.nz;
    acc.tile_words = P.tile_words;
    acc.ntiles = P.ntiles;
    int best_mode = DIM3_DENSE_OFF;
    uint32_t best_b = 0;
    double best_gain = -1e300;
    bool best_fused = false;
    DenseShape best_shape;
    for (int b = (int)NB - 1; b >= 1; b--) {
        acc.K += h[b];
        acc.paths += h[NB + b];
        acc.sumR += h[2 * NB + b];
        acc.sumS += h[3 * NB + b];
        acc.Mh += h[4 * NB + b];
        acc.Nh += h[5 * NB + b];
        if (acc.K == 0) continue;
        if (o.heavy_k > 0 && acc.K > (uint64_t)o.heavy_k && best_mode != DIM3_DENSE_OFF) break;
        const int modes[2] = {DIM3_DENSE_MFMA, DIM3_DENSE_BITOR};
        int bm = DIM3_DENSE_OFF;
        double bg = -1e300;
        bool bf = false;
        for (int mode : modes) {
            if (o.dense_mode > 0 && o.dense_mode != mode) continue;
            bool fz = false;
            const double t = dense_cost(m, acc, mode, count, room, nullptr, true, &fz);
            if (t < 0) continue;
            const double gain = (double)acc.paths * m.path_s - t;
            if (gain > bg) {
                bg = gain;
                bm = mode;
                bf = fz;
            }
        }
        if (bm == DIM3_DENSE_OFF) continue;
        // with heavy_k set, the largest key set within it wins
        if (o.heavy_k > 0 || bg > best_gain) {
            best_gain = bg;
            best_mode = bm;
            best_b = (uint32_t)b;
            best_shape = acc;
            best_fused = bf;
        }
    }
    // dense-core candidate: the keys whose R and S degrees are both >= 64, when
    // there are many of them (>= 2048: a block of keys rather than hubs)
    // (large inputs only: the analysis costs a host round trip)
    const bool core_ok = o.dense_core == 1 || (o.dense_core == -1 && P.nR + P.nS >= 8000000ull);
    if (core_ok && !(count && o.dense_mode == DIM3_DENSE_BITOR)) {
        uint64_t k = 0, pth = 0;
        uint32_t bc = 0;
        for (int b = (int)NB - 1; b >= 1; b--) {
            if (dim3h::dense_bin_lo((uint32_t)b) < 64 && o.dense_core != 1) break;
            if (dim3h::dense_bin_lo((uint32_t)b) < 2) break;
            k += h[b];
            pth += h[NB + b];
            if (h[b]) bc = (uint32_t)b;
            if (o.dense_core != 1 && k >= 2048 && dim3h::dense_bin_lo((uint32_t)b) <= 1024) break;
        }
        if (bc && (k >= 2048 || o.dense_core == 1)) {
            P.core_try = true;
            P.bc = bc;
            P.core_K = k;
            P.core_paths = pth;
        }
    }
    P.p1_gain = best_mode == DIM3_DENSE_OFF ? -1e300 : best_gain;
    if (best_mode == DIM3_DENSE_OFF) return;
    // The rates are good to about a factor of two (a fused seed's bitmap
    // word has measured between 0.4 and 1 path's cost), so AUTO takes a
    // dense part only when it promises to halve the sparse work it replaces:
    // marginal ones (the two-hop graph's hubs, whose 3%-dense bitmaps cost
    // what their columns cost) stay sparse.
    if (o.dense_mode == DIM3_DENSE_AUTO && o.heavy_k == 0 && best_gain <= 0.5 * (double)best_shape.paths * m.path_s)
        return;
    P.dense = best_mode;
    P.bmin = best_b;
    P.K = (uint32_t)best_shape.K;
    P.Kp = (P.K + 63u) / 64u * 64u;
    P.Mh = (uint32_t)best_shape.Mh;
    P.Nh = (uint32_t)best_shape.Nh;
    // columns covering nearly every z (or a fused bit-or, whose key bitmaps
    // span every z ID): take them all, so that each tile's columns are its
    // consecutive z IDs and the seeding copies words
    P.all_z = !P.count_mode && ((uint64_t)P.Nh * 10 >= (uint64_t)P.nz * 9 || (best_mode == DIM3_DENSE_BITOR && best_fused));
    if (P.all_z) P.Nh = P.nz;
    P.dwords = ((P.Nh + 31u) / 32u + 3u) & ~3u;  // rows of 16-byte multiples
}

// Dense core of the keys with bin >= P.bc, from the heavy-degree histogram
// in the bank's hhist ([4][bins]: rows, their heavy entries, columns, their heavy
// tuples): rows with >= thx heavy entries times columns with >= thz heavy
// tuples.  The paths kept inside the core are estimated as the key paths
// times the fractions of heavy entries and of heavy tuples inside it.
// Returns true when the core (or, forced, any core) beats the phase-1 plan.
bool plan_core(dim3_ctx* c, const dim3_opts& o) {
    auto& P = c->bank().plan;
    const unsigned long long* h = c->bank().hhist;
    const uint32_t NB = dim3h::kDenseBins;
    const DenseModel& m = c->model;
    const uint64_t room = c->budget > c->held ? (c->budget - c->held) / 3 : 0;
    const bool count = P.count_mode;
    std::vector<uint64_t> xd(NB + 1, 0), sx(NB + 1, 0), zd(NB + 1, 0), sz(NB + 1, 0);
    for (int b = (int)NB - 1; b >= 0; b--) {
        xd[b] = xd[b + 1] + h[b];
        sx[b] = sx[b + 1] + h[NB + b];
        zd[b] = zd[b + 1] + h[2 * NB + b];
        sz[b] = sz[b + 1] + h[3 * NB + b];
    }
    if (sx[1] == 0 || sz[1] == 0) return false;
    double best = -1e300;
    int bm = DIM3_DENSE_OFF;
    uint32_t bxb = 1, bzb = 1;
    for (uint32_t bx = 1; bx < NB; bx++) {
        if (xd[bx] == 0) break;
        for (uint32_t bz = 1; bz < NB; bz++) {
            if (zd[bz] == 0) break;
            if (o.dense_core == 1 && bx == 1) continue;  // forced: a real core of rows
            DenseShape d;
            d.K = P.core_K;
            d.Mh = xd[bx];
            d.Nh = zd[bz];
            d.paths = (uint64_t)((double)P.core_paths * ((double)sx[bx] / (double)sx[1]) *
                                 ((double)sz[bz] / (double)sz[1]));
            d.sumR = sx[bx];
            d.nz = P.nz;
            // a core pays off only where its cells are reached by many paths
            // (a dense block); thinner cores are priced wrongly by the
            // independence estimate above and are left to the sparse part
            if (o.dense_core != 1 && (double)d.paths < 256.0 * (double)d.Mh * (double)d.Nh) continue;
            const int modes[2] = {DIM3_DENSE_MFMA, DIM3_DENSE_BITOR};
            for (int mode : modes) {
                if (o.dense_mode > 0 && o.dense_mode != mode) continue;
                const double t = dense_cost(m, d, mode, count, room, nullptr);
                if (t < 0) continue;
                const double gain = (double)d.paths * m.path_s - t;
                if (gain > best) {
                    best = gain;
                    bm = mode;
                    bxb = bx;
                    bzb = bz;
                }
            }
        }
    }
    if (bm == DIM3_DENSE_OFF) return false;
    // (a core must save at least a millisecond: small cores are within the
    // error of the estimate and cost their own passes)
    if (o.dense_core != 1 && (best <= P.p1_gain || (o.dense_mode == DIM3_DENSE_AUTO && best <= 1e-3))) return false;
    P.dense = bm;
    P.bmin = P.bc;
    P.K = (uint32_t)P.core_K;
    P.Kp = (P.K + 63u) / 64u * 64u;
    P.Mh = (uint32_t)xd[bxb];
    P.Nh = (uint32_t)zd[bzb];
    P.thx = (uint32_t)dim3h::dense_bin_lo(bxb);
    P.thz = (uint32_t)dim3h::dense_bin_lo(bzb);
    P.core = P.thx > 1;
    P.all_z = !count && (uint64_t)P.Nh * 10 >= (uint64_t)P.nz * 9;
    if (P.all_z) P.Nh = P.nz;
    P.dwords = ((P.Nh + 31u) / 32u + 3u) & ~3u;  // rows of 16-byte multiples
    return true;
}

// Tile bit layout and fused key-bitmap width of a planned query.  Project
// Project tiles take the blocked bit layout (hot z IDs spread over words and
// banks) unless the tile width is not a multiple of 32 words or whole words
// pay more: a fingerprint over a dense z range (a plain word sums its
// 32 z factors with one word factor and four byte-table lookups: the two-hop
// graph's fingerprint query 64.8 -> 42.5 ms, all-sparse), or an MFMA product
// whose columns are every z ID (word copies into the tile).  A fused bit-or
// keeps one bitmap per heavy key in the tile layout, ntiles * tile_words
// words, so that a row tile ORs words t * tile_words .. + tile_words of each
// of its keys.
void set_layout(QueryPlan& P, const dim3_opts& o) {
    P.plain = (o.output_mode == DIM3_OUT_FINGERPRINT && P.z_ident) ||
              (P.dense == DIM3_DENSE_MFMA && P.all_z) || P.tile_words % (1u << kTileBlockLog) != 0;
    if (P.fused) P.dwords = P.ntiles * P.tile_words;
}

// ---------------------------------------------------------------------------
// planner: COUNT fingerprint with every group's count (k_group_count.hip).
// The heavy keys are the top heaviness bins holding at most 64 keys (or
// heavy_k); their paths become the MFMA box of the rows and columns that
// reach one (Mh x Nh cells, from the heaviness histogram), the other paths
// light (x, z-tile) tasks.  Taken when the model prices it under half of the
// all-sparse COUNT (or forced by dense_mode = DIM3_DENSE_MFMA).  Rates on one
// MI355X: the box at ~6e-14 s per cell (VALU-bound: the count bytes packed
// and flagged in registers), light paths ~4e-12 s (a hash insert each),
// light tasks ~1e-9 s, the all-sparse COUNT ~2.2e-12 s per join path.
// ---------------------------------------------------------------------------
constexpr uint32_t kGcMaxKeys = 64;
constexpr double kGcCellS = 6e-14, kGcLightPathS = 4e-12, kGcTaskS = 1e-9, kCountPathS = 2.2e-12;
constexpr uint32_t kGcLightPerTask = 96;  // mean light paths per (x, z-tile) task the tiles aim at
constexpr uint32_t kGcCols = 16384;       // box columns per dense task

// internal: prepare() planning again after the masks found a duplicate heavy tuple
constexpr uint32_t kOptNoGc = 1u << 31;

void plan_gc(dim3_ctx* c, const dim3_opts& o, int force_wgs) {
    auto& P = c->bank().plan;
    P.gc = false;
    P.gc_dup = false;
    if ((o.flags & kOptNoGc) || !P.count_mode || P.agg || o.output_mode != DIM3_OUT_FINGERPRINT || (P.flags & DIM3_OPT_EXCLUDE_EQUAL)) return;
    if (force_wgs > 0 || o.dense_mode == DIM3_DENSE_OFF || o.dense_mode == DIM3_DENSE_BITOR) return;
    const unsigned long long* h = c->bank().hhist;
    const uint32_t NB = dim3h::kDenseBins;
    const uint64_t kmax = o.heavy_k > 0 ? std::min<uint64_t>((uint64_t)o.heavy_k, kGcMaxKeys) : kGcMaxKeys;
    uint64_t K = 0, paths = 0, Mh = 0, Nh = 0;
    uint32_t bmin = 0;
    for (int b = (int)NB - 1; b >= 1; b--) {
        if (K + h[b] > kmax) break;
        K += h[b];
        paths += h[NB + b];
        Mh += h[4 * NB + b];
        Nh += h[5 * NB + b];
        if (h[b]) bmin = (uint32_t)b;
    }
    if (K == 0 || Mh == 0 || Nh == 0) return;
    const uint64_t light = P.out_j > paths ? P.out_j - paths : 0;
    uint64_t nt = (light / std::max<uint64_t>(P.nx, 1) + kGcLightPerTask - 1) / kGcLightPerTask;
    if (o.tile_bits > 0) nt = ((uint64_t)P.nz + (uint64_t)o.tile_bits - 1) / (uint64_t)o.tile_bits;  // caller's z-tile width
    nt = std::max<uint64_t>(1, std::min<uint64_t>(nt, 1024));
    nt = std::min<uint64_t>(nt, P.nz);
    while (nt > 1 && ((uint64_t)P.ny * nt >= (1ull << 31) || (uint64_t)P.nx * nt >= (1ull << 31))) nt >>= 1;
    const uint64_t rows = Mh;
    const uint64_t Mp = (rows + dim3h::gc_rows_per_block() - 1) / dim3h::gc_rows_per_block() *
                        dim3h::gc_rows_per_block();
    const uint64_t bytes = Mp * (64 + 16) + (Nh + 64) * (64 + 8 + 16) + ((uint64_t)P.nx + P.nz) * 12 + 4 * (uint64_t)P.nx +
                           4 * (uint64_t)P.nz;
    const uint64_t room = c->budget > c->held ? (c->budget - c->held) / 3 : 0;
    if (bytes > room) return;
    const double t_gc = (double)Mh * (double)Nh * kGcCellS + (double)light * kGcLightPathS +
                        (double)P.nx * (double)nt * kGcTaskS + 1e-4;
    const double t_sparse = (double)P.out_j * kCountPathS;
    if (o.dense_mode != DIM3_DENSE_MFMA && t_gc >= 0.5 * t_sparse) return;
    // forced MFMA without a key count: the box only when its keys carry most
    // paths (a block of many equally heavy keys, which 64 keys cannot hold,
    // goes to the materialised MFMA product instead)
    if (o.dense_mode == DIM3_DENSE_MFMA && o.heavy_k == 0 && 2 * paths < P.out_j) return;
    P.gc = true;
    P.gc_bmin = bmin;
    P.gc_K = (uint32_t)K;
    P.dense = DIM3_DENSE_OFF;
    P.core_try = false;
    P.core = false;
    P.fused = false;
    P.ntiles = (uint32_t)nt;
    P.tile_size = (uint32_t)((P.nz + nt - 1) / nt);
}

// The GC plan's masks, box lists and operands (after the CSRs); one host
// synchronisation reads the box size.
int gc_prepare(dim3_ctx* c, const int32_t* slotS, const int32_t* slotSz, const uint32_t* degR, const uint32_t* degS,
               const uint64_t* exY, const uint64_t* exZ, uint64_t ny1, hipStream_t st) {
    Bank& B = c->bank();
    auto& P = B.plan;
    Slot* W = B.ws;
    Scalars* sc = (Scalars*)W[WS_SCALARS].p;
    int32_t* hid;
    uint32_t *h2y, *xflag, *zflag, *xpos, *zpos, *rlist, *clist;
    uint64_t *hm, *zm, *part;
    TRY(ws(c, WS_HID, P.ny, &hid));
    TRY(ws(c, WS_H2Y, P.ny, &h2y));
    TRY(ws(c, WS_GC_HM, P.nx, &hm));
    TRY(ws(c, WS_GC_ZM, P.nz, &zm));
    TRY(ws(c, WS_GC_XFLAG, P.nx, &xflag));
    TRY(ws(c, WS_GC_ZFLAG, P.nz, &zflag));
    TRY(ws(c, WS_GC_XPOS, P.nx, &xpos));
    TRY(ws(c, WS_GC_ZPOS, P.nz, &zpos));
    TRY(ws(c, WS_GC_RLIST, P.nx, &rlist));
    TRY(ws(c, WS_GC_CLIST, P.nz, &clist));
    TRY(ws(c, WS_PARTIALS, dim3h::scan_partials_needed(std::max<uint64_t>(P.nx, P.nz) + 1), &part));
    uint32_t* counters = (uint32_t*)&sc->spare[0];
    HIPCHK(c, hipMemsetAsync(counters, 0, 8, st));
    HIPCHK(c, hipMemsetAsync(&sc->spare[3], 0, 8, st));
    dim3h::dense_mark_keys(degR, degS, (uint32_t)ny1, exY, P.gc_bmin, hid, h2y, counters, st);
    dim3h::gc_masks((const uint32_t*)W[WS_ROWPTR_R].p, (int32_t*)W[WS_COL_R].p, P.nx, P.S, P.nS, slotS,
                    P.z_ident ? nullptr : slotSz, exY, P.z_ident ? nullptr : exZ, P.zlo, hid, P.nz, hm, xflag, zm,
                    zflag, &sc->spare[3], st);
    // the light S tuples by (z tile, key)
    const uint64_t nkl = (uint64_t)P.ntiles * P.ny;
    uint32_t *cntL, *rowL, *curL, *keyL, *zidL;
    int32_t* colL;
    TRY(ws(c, WS_GC_CNTL, nkl + 1, &cntL));
    TRY(ws(c, WS_GC_ROWL, nkl + 1, &rowL));
    TRY(ws(c, WS_GC_CURL, nkl, &curL));
    TRY(ws(c, WS_KEY_S, P.nS, &keyL));
    TRY(ws(c, WS_ZID_S, P.nS, &zidL));
    TRY(ws(c, WS_GC_COLL, std::max<uint64_t>(P.nS, 1), &colL));
    TRY(ws(c, WS_PARTIALS, dim3h::scan_partials_needed(std::max<uint64_t>(nkl, std::max<uint64_t>(P.nx, P.nz)) + 1),
           &part));
    dim3h::fill_u32(cntL, nkl + 1, 0, st);
    dim3h::gc_light_csr_count(P.S, P.nS, slotS, P.z_ident ? nullptr : slotSz, exY, P.z_ident ? nullptr : exZ, P.zlo,
                              hid, P.ny, P.tile_size, cntL, keyL, zidL, st);
    dim3h::scan_u32(cntL, (uint32_t)nkl, rowL, part, &sc->tot_s, st);
    dim3h::finish_rowptr(rowL, (uint32_t)nkl, &sc->tot_s, st);
    HIPCHK(c, hipMemcpyAsync(curL, rowL, (size_t)nkl * 4, hipMemcpyDeviceToDevice, st));
    dim3h::gc_light_csr_scatter(keyL, zidL, P.nS, curL, colL, st);
    dim3h::scan_u32(xflag, P.nx, xpos, part, &sc->nxh, st);
    dim3h::gc_list(xflag, xpos, P.nx, rlist, st);
    dim3h::scan_u32(zflag, P.nz, zpos, part, &sc->nzh, st);
    dim3h::gc_list(zflag, zpos, P.nz, clist, st);
    LAUNCHCHK(c, "group-count masks");
    HIPCHK(c, dim3h::copy_wait(B.hsc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    P.gc_K = (uint32_t)B.hsc->spare[0];  // the keys marked (the bins' count)
    P.gc_Mh = (uint32_t)B.hsc->nxh;
    P.gc_Nh = (uint32_t)B.hsc->nzh;
    const uint32_t rb = dim3h::gc_rows_per_block();
    P.gc_Mp = (uint32_t)(((uint64_t)P.gc_Mh + rb - 1) / rb * rb);
    if (P.gc_K > kGcMaxKeys) return fail(c, DIM3_ERR_DEVICE, "group count: %u heavy keys marked", P.gc_K);
    if (B.hsc->spare[3] != 0) {  // a heavy tuple twice: prepare() plans again without this path
        P.gc_dup = true;
        return DIM3_OK;
    }
    int8_t *A8, *B8;
    uint64_t *xk, *gz, *G;
    const uint64_t Np = dim3h::gc_cols_padded(P.gc_Nh);
    TRY(ws(c, WS_GC_A8, std::max<uint64_t>((uint64_t)P.gc_Mp * 64, 64), &A8));
    TRY(ws(c, WS_GC_XK, std::max<uint64_t>(2ull * P.gc_Mp, 2), &xk));
    TRY(ws(c, WS_GC_B8, std::max<uint64_t>(Np * 64, 64), &B8));
    TRY(ws(c, WS_GC_GZ, std::max<uint64_t>(Np, 1), &gz));
    TRY(ws(c, WS_GC_G, std::max<uint64_t>(Np * 2, 2), &G));  // 16 bytes per column
    dim3h::gc_build(rlist, P.gc_Mh, P.gc_Mp, hm, (const int32_t*)W[WS_ID2X].p, clist, P.gc_Nh, zm,
                    P.z_ident ? nullptr : (const int32_t*)W[WS_ID2Z].p, P.zlo, A8, xk, B8, gz, G, st);
    uint64_t* zinfo;
    TRY(ws(c, WS_GC_ZINFO, 2ull * P.nz, &zinfo));
    dim3h::gc_zinfo(zm, P.nz, P.z_ident ? nullptr : (const int32_t*)W[WS_ID2Z].p, P.zlo, zinfo, st);
    LAUNCHCHK(c, "group-count operands");
    return DIM3_OK;
}

// Mapping + CSR of one query.  Leaves the plan in c->bank().plan.
// A query is prepared in three stages, each ending with a copy of the
// device scalars to the bank's host mirror: start (swap rule, min/max of
// the attributes), map (dense IDs, degrees, the dense-split histogram) and
// plan (tiles, dense split, CSRs, dense product).  Each stage reads what the
// previous one copied, so the host synchronises between stages; a batch runs
// each stage over all its instances before one synchronisation.
int prepare_start(dim3_ctx* c, bool count_mode, const int2* dR, uint64_t nR, const int2* dS, uint64_t nS,
                  const dim3_opts& o, hipStream_t st, const ValueSpec& vals = ValueSpec()) {
    Bank& B = c->bank();
    auto& P = B.plan;
    P = QueryPlan();
    P.agg = vals.agg;
    P.vf32 = vals.f32;
    P.vR = vals.vR;
    P.vS = vals.vS;
    if (P.agg) count_mode = true;  // per-z counters; never the COUNT fingerprint shortcut
    if (!B.hsc) {
        HIPCHK(c, dim3h::host_malloc((void**)&B.hsc, sizeof(Scalars)));
        HIPCHK(c, dim3h::host_malloc((void**)&B.hhist, 6 * dim3h::kDenseBins * sizeof(unsigned long long)));
    }
    // With DIM3_OPT_PATH_SUMS a COUNT fingerprint forms no counter: its
    // groups are the join-project rows (n, sum) and its sum2 / csum are sums
    // over the join paths, taken from the key degrees.  Otherwise (and always
    // with the u != w filter) every group's count is formed.
    P.agg_fp = count_mode && !P.agg && o.output_mode == DIM3_OUT_FINGERPRINT && !(o.flags & DIM3_OPT_EXCLUDE_EQUAL) &&
               (o.flags & DIM3_OPT_PATH_SUMS);
    if (P.agg_fp) count_mode = false;
    P.count_mode = count_mode;
    P.flags = o.flags;
    if (!(o.flags & DIM3_OPT_NO_SWAP) && nR < nS) {
        std::swap(dR, dS);
        std::swap(nR, nS);
        std::swap(P.vR, P.vS);
        P.swapped = true;
    }
    P.R = dR;
    P.S = dS;
    P.nR = nR;
    P.nS = nS;
    if (nR > dim3s::kMaxTuples || nS > dim3s::kMaxTuples)
        return fail(c, DIM3_ERR_INVALID, "relations are limited to 2^30-1 tuples (got %llu, %llu)",
                    (unsigned long long)nR, (unsigned long long)nS);

    Scalars* sc;
    TRY(ws(c, WS_SCALARS, 1, &sc));
    HIPCHK(c, hipMemsetAsync(sc, 0, sizeof(Scalars), st));
    if (nR == 0 || nS == 0) {  // dim3.hpp:1027
        P.empty = true;
        return DIM3_OK;
    }
    P.empty = false;  // until the mapping finds no joined key

    HIPCHK(c, hipEventRecord(c->ev[0], st));
    // value ranges decide hash or range mapping per attribute
    dim3h::minmax(dR, nR, dS, nS, sc->mm, st);
    HIPCHK(c, hipMemcpyAsync(B.hsc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    return DIM3_OK;
}

int prepare_map(dim3_ctx* c, const dim3_opts& o, hipStream_t st) {
    Bank& B = c->bank();
    auto& P = B.plan;
    if (P.empty) return DIM3_OK;
    const int2 *dR = P.R, *dS = P.S;
    const uint64_t nR = P.nR, nS = P.nS;
    Scalars* sc = (Scalars*)B.ws[WS_SCALARS].p;
    const int32_t* mm = B.hsc->mm;
    const bool hash_only = (o.flags & DIM3_OPT_HASH_ONLY) != 0;
    auto span = [](int32_t lo, int32_t hi) { return (uint64_t)((int64_t)hi - (int64_t)lo + 1); };
    const uint64_t xr = span(mm[0], mm[1]), yr = span(mm[2], mm[3]);
    uint64_t zr = span(mm[4], mm[5]);
    P.x_range = !hash_only && dim3s::range_fits(xr, nR);
    P.y_range = !hash_only && dim3s::range_fits(yr, nR);
    // z IDs are value - zlo (identity slots, no power of two); the 32-aligned
    // start below adds at most 31 values
    P.z_ident = !hash_only && zr <= 2 * nS + 64 && zr + 31 <= dim3s::kMaxSlots;
    // dense z IDs start at a multiple of 32 below min(z), so that a word of
    // a project tile is one 32-aligned word of z values (the fingerprint
    // sums such a word at once); the IDs below min(z) stay empty
    P.zlo = P.z_ident ? (int32_t)((uint32_t)mm[4] & ~31u) : 0;
    if (P.z_ident) zr += (uint64_t)((int64_t)mm[4] - (int64_t)P.zlo);
    P.z_span = zr;

    uint32_t *keysY = nullptr, *degR, *degS, *slotR, *keysX = nullptr, *degX, *keysZ = nullptr, *degZ = nullptr;
    int32_t *slotS, *slotRx, *slotSz = nullptr;
    uint64_t *exY, *exX, *exZ = nullptr, *part;
    auto make_map = [&](bool range, bool identity, int32_t lo, uint64_t span_v, uint64_t n, SlotId keys_slot,
                        uint32_t** keys, dim3h::KeyMap* m) -> int {
        m->range = range ? 1u : 0u;
        m->lo = lo;
        m->keys = nullptr;
        m->mask = 0;
        m->span = range ? (uint32_t)span_v : 0u;
        if (range) {
            const uint32_t p2 = (uint32_t)pow2_at_least(span_v);  // <= 2^31 (range_fits)
            m->mask = identity ? 0u : p2 - 1u;
            m->nslots = identity ? (uint32_t)span_v : p2;
        } else {
            const uint32_t cap = (uint32_t)dim3s::key_table_cap(n);  // <= 2^31 for n <= kMaxTuples
            TRY(ws(c, keys_slot, (uint64_t)cap + 1, keys));
            m->keys = *keys;
            m->mask = cap - 1;
            m->nslots = cap + 1;
        }
        return DIM3_OK;
    };
    TRY(make_map(P.y_range, false, mm[2], yr, nR, WS_KEYS_Y, &keysY, &P.my));
    TRY(make_map(P.x_range, false, mm[0], xr, nR, WS_KEYS_X, &keysX, &P.mx));
    TRY(make_map(P.z_ident, true, P.z_ident ? P.zlo : mm[4], zr, nS, WS_KEYS_Z, &keysZ, &P.mz));
    const uint64_t ny1 = P.my.nslots, nx1 = P.mx.nslots, nz1 = P.mz.nslots;
    TRY(ws(c, WS_DEGR_Y, ny1, &degR));
    TRY(ws(c, WS_DEGS_Y, ny1, &degS));
    TRY(ws(c, WS_EX_Y, ny1, &exY));
    TRY(ws(c, WS_SLOT_R, nR, &slotR));
    TRY(ws(c, WS_SLOT_S, nS, &slotS));
    TRY(ws(c, WS_DEG_X, nx1, &degX));
    TRY(ws(c, WS_EX_X, nx1, &exX));
    TRY(ws(c, WS_SLOT_RX, nR, &slotRx));
    if (!P.z_ident) {
        TRY(ws(c, WS_DEG_Z, nz1, &degZ));
        TRY(ws(c, WS_EX_Z, nz1, &exZ));
        TRY(ws(c, WS_SLOT_SZ, nS, &slotSz));
    }
    const uint64_t maxscan = std::max<uint64_t>(std::max(ny1, std::max(nx1, nz1)), std::max<uint64_t>(nS, nR) + 1);
    TRY(ws(c, WS_PARTIALS, dim3h::scan_partials_needed(maxscan), &part));

    dim3h::map_init(P.my, st);
    dim3h::map_init(P.mx, st);
    dim3h::fill_u32(degR, ny1, 0, st);
    dim3h::fill_u32(degS, ny1, 0, st);
    dim3h::fill_u32(degX, nx1, 0, st);
    // the most frequent key of each side: its tuples share one atomic per
    // wavefront in the key-counting kernels
    const int32_t* hot = dim3h::hot_keys(dR, nR, dS, nS, sc->hot, st) ? sc->hot : nullptr;
    P.hot = hot != nullptr;
    dim3h::build_y(dR, nR, P.my, degR, slotR, st, hot);
    dim3h::probe_y(dS, nS, P.my, degR, degS, slotS, &sc->s_surv, st, hot);
    dim3h::scan_y(degR, degS, (uint32_t)ny1, exY, part, &sc->tot_y, &sc->out_j, st);
    if (P.agg_fp) {
        uint64_t *kr, *gs;
        TRY(ws(c, WS_AGG_KR, ny1, &kr));
        TRY(ws(c, WS_AGG_GS, ny1, &gs));
        dim3h::path_sums(dR, nR, slotR, dS, nS, slotS, (uint32_t)ny1, kr, gs, &sc->agg_sum2, st, hot);
    }
    dim3h::build_x(dR, nR, slotR, degS, P.mx, degX, slotRx, st, hot);
    dim3h::scan_deg(degX, (uint32_t)nx1, exX, part, &sc->tot_x, st);
    if (!P.z_ident) {
        dim3h::map_init(P.mz, st);
        dim3h::fill_u32(degZ, nz1, 0, st);
        dim3h::build_z(dS, nS, slotS, P.mz, degZ, slotSz, st);
        dim3h::scan_deg(degZ, (uint32_t)nz1, exZ, part, &sc->tot_z, st);
    }
    // the hot x and z values as slots, for the passes that count per slot
    if (hot) {
        dim3h::hot_slots(sc->hot + dim3h::kHotRX, P.mx, false, 0, 0, sc->hot_xs, st);
        dim3h::hot_slots(sc->hot + dim3h::kHotSZ, P.mz, P.z_ident, P.zlo, P.z_span, sc->hot_zs, st);
    }
    // heaviness histogram of the dense split
    const bool want_dense = o.dense_mode != DIM3_DENSE_OFF;
    uint32_t *xw = nullptr, *zw = nullptr;
    unsigned long long* dhist = nullptr;
    if (want_dense) {
        TRY(ws(c, WS_XW, nx1, &xw));
        TRY(ws(c, WS_ZW, nz1, &zw));
        TRY(ws(c, WS_DHIST, 6ull * dim3h::kDenseBins, &dhist));
        dim3h::fill_u32(xw, nx1, 0, st);
        dim3h::fill_u32(zw, nz1, 0, st);
        dim3h::fill_u64((uint64_t*)dhist, 6ull * dim3h::kDenseBins, 0, st);
        dim3h::dense_weights(dR, nR, slotR, slotRx, dS, nS, slotS, P.z_ident ? nullptr : slotSz, P.zlo, degR, degS, xw,
                             (uint32_t)nx1, zw, (uint32_t)nz1, st, hot ? sc->hot_xs : nullptr,
                             hot ? sc->hot_zs : nullptr);
        dim3h::dense_hist(degR, degS, (uint32_t)ny1, xw, (uint32_t)nx1, zw, (uint32_t)nz1, dhist, st);
        HIPCHK(c, hipMemcpyAsync(B.hhist, dhist, 6ull * dim3h::kDenseBins * 8, hipMemcpyDeviceToHost, st));
    }
    LAUNCHCHK(c, "mapping");
    HIPCHK(c, hipEventRecord(c->ev[1], st));
    HIPCHK(c, hipMemcpyAsync(B.hsc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    return DIM3_OK;
}

int prepare_plan(dim3_ctx* c, const dim3_opts& o, hipStream_t st, int force_wgs) {
    Bank& B = c->bank();
    auto& P = B.plan;
    if (P.empty) return DIM3_OK;
    const int2* dS = P.S;
    const uint64_t nR = P.nR, nS = P.nS;
    Slot* W = B.ws;
    Scalars* sc = (Scalars*)W[WS_SCALARS].p;
    // the mapping stage's buffers
    auto slot = [&](SlotId id) { return W[id].p; };
    uint32_t* degR = (uint32_t*)slot(WS_DEGR_Y);
    uint32_t* degS = (uint32_t*)slot(WS_DEGS_Y);
    uint64_t* exY = (uint64_t*)slot(WS_EX_Y);
    uint32_t* slotR = (uint32_t*)slot(WS_SLOT_R);
    int32_t* slotS = (int32_t*)slot(WS_SLOT_S);
    uint32_t* degX = (uint32_t*)slot(WS_DEG_X);
    uint64_t* exX = (uint64_t*)slot(WS_EX_X);
    int32_t* slotRx = (int32_t*)slot(WS_SLOT_RX);
    uint32_t* degZ = P.z_ident ? nullptr : (uint32_t*)slot(WS_DEG_Z);
    uint64_t* exZ = P.z_ident ? nullptr : (uint64_t*)slot(WS_EX_Z);
    int32_t* slotSz = P.z_ident ? nullptr : (int32_t*)slot(WS_SLOT_SZ);
    const bool want_dense = o.dense_mode != DIM3_DENSE_OFF;
    uint32_t* xw = want_dense ? (uint32_t*)slot(WS_XW) : nullptr;
    uint32_t* zw = want_dense ? (uint32_t*)slot(WS_ZW) : nullptr;
    uint64_t* part = (uint64_t*)slot(WS_PARTIALS);
    const uint64_t ny1 = P.my.nslots, nx1 = P.mx.nslots, nz1 = P.mz.nslots;
    const uint64_t maxscan = std::max<uint64_t>(std::max(ny1, std::max(nx1, nz1)), std::max<uint64_t>(nS, nR) + 1);

    P.ny = (uint32_t)(B.hsc->tot_y >> 32);
    P.nx = (uint32_t)(B.hsc->tot_x >> 32);
    P.nnzR = (uint32_t)B.hsc->tot_x;
    P.nnzS = (uint32_t)B.hsc->s_surv;
    P.nz = P.z_ident ? (uint32_t)P.z_span : (uint32_t)(B.hsc->tot_z >> 32);
    P.out_j = B.hsc->out_j;
    if (P.ny == 0 || P.nx == 0 || P.nz == 0 || P.nnzS == 0) {
        P.empty = true;
        return DIM3_OK;
    }
    P.empty = false;
    plan_tiles(c, o, force_wgs);
    plan_dense(c, o);
    plan_gc(c, o, force_wgs);
    P.fused = P.dense == DIM3_DENSE_BITOR && P.all_z;
    set_layout(P, o);
    if ((uint64_t)P.ny * P.ntiles >= (1ull << 31))
        return fail(c, DIM3_ERR_INVALID, "tiled S index does not fit 32 bits (ny=%u, tiles=%u)", P.ny, P.ntiles);

    HIPCHK(c, hipEventRecord(c->ev[2], st));
    int32_t *id2x, *id2z = nullptr, *colR, *zidS, *colS, *x2z;
    uint32_t *rowR, *curR, *cntS, *rowS, *curS, *keyS;
    // (the group-count plan builds its own light S CSR, tile-major)
    const uint64_t nkeyS = P.gc ? 1 : (uint64_t)P.ny * P.ntiles;
    TRY(ws(c, WS_ID2X, P.nx, &id2x));
    if (!P.z_ident) TRY(ws(c, WS_ID2Z, (uint64_t)P.ntiles * P.tile_size, &id2z));
    TRY(ws(c, WS_ROWPTR_R, (uint64_t)P.nx + 1, &rowR));
    TRY(ws(c, WS_CUR_R, P.nx, &curR));
    TRY(ws(c, WS_COL_R, P.nnzR, &colR));
    TRY(ws(c, WS_CNT_S, nkeyS + 1, &cntS));
    TRY(ws(c, WS_ROWPTR_S, nkeyS + 1, &rowS));
    TRY(ws(c, WS_CUR_S, nkeyS, &curS));
    TRY(ws(c, WS_KEY_S, nS, &keyS));
    TRY(ws(c, WS_ZID_S, nS, &zidS));
    TRY(ws(c, WS_COL_S, P.nnzS, &colS));
    TRY(ws(c, WS_X2Z, P.nx, &x2z));
    TRY(ws(c, WS_PARTIALS, dim3h::scan_partials_needed(std::max<uint64_t>(maxscan, nkeyS + 1)), &part));

    int32_t *hid = nullptr, *hx = nullptr;
    uint32_t *hx2x = nullptr, *h2y = nullptr, *zflag = nullptr, *hzpos = nullptr, *hz2zid = nullptr,
             *hz_tile = nullptr, *hdx = nullptr, *hdz = nullptr;
    uint32_t* counters = (uint32_t*)&sc->spare[0];
    if (P.dense != DIM3_DENSE_OFF || P.core_try) {
        // heavy keys (of the core candidate when there is one)
        TRY(ws(c, WS_HID, P.ny, &hid));
        TRY(ws(c, WS_H2Y, std::max<uint64_t>(P.core_try ? P.core_K : 0, P.K), &h2y));
        dim3h::dense_mark_keys(degR, degS, (uint32_t)ny1, exY, P.core_try ? P.bc : P.bmin, hid, h2y, counters, st);
        if (P.core_try) {
            unsigned long long* hh;
            TRY(ws(c, WS_HDX, nx1, &hdx));
            TRY(ws(c, WS_HDZ, nz1, &hdz));
            TRY(ws(c, WS_DHIST, 6ull * dim3h::kDenseBins, &hh));
            dim3h::fill_u32(hdx, nx1, 0, st);
            dim3h::fill_u32(hdz, nz1, 0, st);
            dim3h::fill_u64((uint64_t*)hh, 4ull * dim3h::kDenseBins, 0, st);
            dim3h::dense_heavy_degrees(slotR, slotRx, nR, dS, slotS, P.z_ident ? nullptr : slotSz, P.zlo, nS, exY,
                                       hid, hdx, (uint32_t)nx1, hdz, (uint32_t)nz1, hh, st,
                                       P.hot ? sc->hot_xs : nullptr, P.hot ? sc->hot_zs : nullptr);
            HIPCHK(c, hipMemcpyAsync(B.hhist, hh, 4ull * dim3h::kDenseBins * 8, hipMemcpyDeviceToHost, st));
            HIPCHK(c, hipStreamSynchronize(st));
            const bool core = plan_core(c, o);
            P.fused = P.dense == DIM3_DENSE_BITOR && P.all_z;
            set_layout(P, o);
            if (!core && P.dense != DIM3_DENSE_OFF) {  // back to the phase-1 keys
                HIPCHK(c, hipMemsetAsync(counters, 0, 8, st));
                dim3h::dense_mark_keys(degR, degS, (uint32_t)ny1, exY, P.bmin, hid, h2y, counters, st);
            }
        }
    }
    if (P.dense != DIM3_DENSE_OFF) {
        // product rows and columns (columns numbered in z-ID order)
        TRY(ws(c, WS_HX, P.nx, &hx));
        TRY(ws(c, WS_HX2X, P.Mh, &hx2x));
        TRY(ws(c, WS_ZFLAG, P.nz, &zflag));
        TRY(ws(c, WS_HZPOS, (uint64_t)P.nz + 1, &hzpos));
        TRY(ws(c, WS_HZ2ZID, P.Nh, &hz2zid));
        TRY(ws(c, WS_HZTILE, (uint64_t)P.ntiles + 1, &hz_tile));
        const bool deg_rule = P.thx > 1 || P.thz > 1;
        dim3h::dense_mark_rows_cols(degX, exX, (uint32_t)nx1, xw, P.z_ident ? nullptr : degZ, P.z_ident ? nullptr : exZ,
                                    (uint32_t)nz1, zw, P.bmin, deg_rule ? hdx : nullptr, P.thx,
                                    deg_rule ? hdz : nullptr, P.thz, P.all_z, hx, hx2x, zflag, counters, st);
        dim3h::scan_u32(zflag, P.nz, hzpos, part, &sc->nzh, st);
        dim3h::dense_columns(zflag, hzpos, P.nz, hz2zid, P.ntiles, P.tile_size, P.Nh, hz_tile, st);
    } else {
        hid = nullptr;
    }
    dim3h::compact_ids(P.mx, degX, exX, id2x, rowR, st);
    dim3h::finish_rowptr(rowR, P.nx, &sc->tot_x, st);
    if (!P.z_ident) dim3h::compact_ids(P.mz, degZ, exZ, id2z, nullptr, st);
    HIPCHK(c, hipMemcpyAsync(curR, rowR, (size_t)P.nx * 4, hipMemcpyDeviceToDevice, st));
    int32_t *valR = nullptr, *valS = nullptr;
    if (P.agg) {
        TRY(ws(c, WS_VAL_R, P.nnzR, &valR));
        TRY(ws(c, WS_VAL_S, P.nnzS, &valS));
    }
    dim3h::scatter_r(slotR, slotRx, nR, exY, exX, hid, hx, curR, colR, P.vR, valR, st);
    dim3h::fill_u32(cntS, nkeyS + 1, 0, st);
    uint32_t *cntSD = nullptr, *curSD = nullptr;
    if (P.core) {
        TRY(ws(c, WS_CNT_SD, nkeyS + 1, &cntSD));
        TRY(ws(c, WS_CUR_SD, nkeyS, &curSD));
        dim3h::fill_u32(cntSD, nkeyS + 1, 0, st);
    }
    if (!P.gc) {
        dim3h::count_s(dS, slotS, slotSz, nS, exY, P.z_ident ? nullptr : exZ, P.zlo, P.ntiles, P.tile_size, hid,
                       zflag, cntS, cntSD, keyS, zidS, st, P.hot ? sc->hot : nullptr);
        dim3h::scan_u32(cntS, (uint32_t)nkeyS, rowS, part, &sc->tot_s, st);
        dim3h::finish_rowptr(rowS, (uint32_t)nkeyS, &sc->tot_s, st);
        if (P.core) dim3h::split_cursors(rowS, cntSD, (uint32_t)nkeyS, curS, curSD, st);
        else HIPCHK(c, hipMemcpyAsync(curS, rowS, (size_t)nkeyS * 4, hipMemcpyDeviceToDevice, st));
        dim3h::scatter_s(keyS, zidS, nS, curS, curSD, colS, P.vS, valS, st);
    }
    if (o.flags & DIM3_OPT_EXCLUDE_EQUAL)
        dim3h::x2z_self(id2x, P.nx, P.mz, degZ, P.z_ident ? nullptr : exZ, x2z, st);
    // a fingerprint over a dense z range takes each z word's factor from a
    // table (the allocation-free entry computes them)
    P.fp_zw = o.output_mode == DIM3_OUT_FINGERPRINT && P.z_ident && !P.agg && c->arena == nullptr;
    if (P.fp_zw) {
        uint32_t* zw;
        const uint64_t nw = ((uint64_t)P.ntiles * P.tile_size + 31u) / 32u;
        TRY(ws(c, WS_FP_ZW, nw, &zw));
        dim3h::fp_word_factors(zw, (uint32_t)nw, P.zlo, st);
    }
    LAUNCHCHK(c, "csr");
    HIPCHK(c, hipEventRecord(c->ev[3], st));
    if (P.gc) {
        TRY(gc_prepare(c, slotS, slotSz, degR, degS, exY, exZ, ny1, st));
        // sum2 and csum are sums over the join paths (sum over keys of the
        // R-side k2 sum times the S-side g sum; csum = the path count), so
        // the group kernels form only each group's presence terms
        uint64_t *kr, *gs;
        TRY(ws(c, WS_AGG_KR, ny1, &kr));
        TRY(ws(c, WS_AGG_GS, ny1, &gs));
        dim3h::fill_u64(&sc->agg_sum2, 1, 0, st);
        dim3h::path_sums(P.R, nR, slotR, dS, nS, slotS, (uint32_t)ny1, kr, gs, &sc->agg_sum2, st,
                         P.hot ? sc->hot : nullptr);
        LAUNCHCHK(c, "group-count path sums");
    }
    if (P.dense != DIM3_DENSE_OFF) {
        // the product of the heavy keys: Dbits (project) or Cint (COUNT)
        uint64_t* overflow = &sc->spare[1];
        uint32_t *dbits = nullptr, *B = nullptr;
        int32_t* cint = nullptr;
        if (P.count_mode) TRY(ws(c, WS_CINT, (uint64_t)P.Mh * P.Nh, &cint));
        else if (!P.fused) TRY(ws(c, WS_DBITS, (uint64_t)P.Mh * P.dwords, &dbits));
        if (P.dense == DIM3_DENSE_MFMA) {
            // A [Mh][Kp] from R's CSR rows, Bt [Nh][Kp] from the heavy S tuples
            uint32_t* A;
            TRY(ws(c, WS_DA, (uint64_t)P.Mh * P.Kp / 4, &A));
            TRY(ws(c, WS_DB, (uint64_t)P.Nh * P.Kp / 4, &B));
            if (P.Kp > 65536) dim3h::fill_u32(A, (uint64_t)P.Mh * P.Kp / 4, 0, st);
            dim3h::fill_u32(B, (uint64_t)P.Nh * P.Kp / 4, 0, st);
            dim3h::dense_build_a(P.count_mode, rowR, colR, hid, hx2x, P.Mh, P.Kp, A, overflow, st);
            dim3h::dense_build_b(P.count_mode ? 1 : 0, dS, slotS, slotSz, P.z_ident ? nullptr : exZ, P.zlo, nS, exY,
                                 hid, zflag, hzpos, P.Kp, P.dwords, B, overflow, st);
            const int e = dim3h::gemm_i8((const int8_t*)A, (const int8_t*)B, P.Mh, P.Nh, P.Kp, dbits, P.dwords, cint,
                                         st);
            if (e != 0) return fail(c, DIM3_ERR_DEVICE, "dense product launch: %s", hipGetErrorString((hipError_t)e));
        } else {
            TRY(ws(c, WS_DB, (uint64_t)P.K * P.dwords, &B));
            dim3h::fill_u32(B, (uint64_t)P.K * P.dwords, 0, st);
            // fused: the tile layout (row tile t ORs words t W .. t W + W - 1);
            // product rows: plain bitmaps over the columns
            dim3h::dense_build_b(2, dS, slotS, slotSz, P.z_ident ? nullptr : exZ, P.zlo, nS, exY, hid, zflag, hzpos,
                                 0, P.dwords, B, overflow, st, P.fused ? P.tile_words : 0u,
                                 P.fused && !P.plain ? kTileBlockLog : 0u);
            if (!P.fused) {  // fused: the seeding reads B itself
                const int e = dim3h::bitor_rows(rowR, colR, hid, hx2x, P.Mh, B, P.dwords, dbits, st);
                if (e != 0)
                    return fail(c, DIM3_ERR_DEVICE, "dense bit-or launch: %s", hipGetErrorString((hipError_t)e));
            }
        }
        LAUNCHCHK(c, "dense");
    }
    HIPCHK(c, hipEventRecord(c->ev[6], st));
    return DIM3_OK;
}

// The three stages of one query, synchronising after each.
int prepare(dim3_ctx* c, bool count_mode, const int2* dR, uint64_t nR, const int2* dS, uint64_t nS,
            const dim3_opts& o, hipStream_t st, int force_wgs = 0, const ValueSpec& vals = ValueSpec()) {
    TRY(prepare_start(c, count_mode, dR, nR, dS, nS, o, st, vals));
    auto& P = c->bank().plan;
    if (P.empty) return DIM3_OK;
    HIPCHK(c, hipStreamSynchronize(st));
    TRY(prepare_map(c, o, st));
    HIPCHK(c, hipStreamSynchronize(st));
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    P.ms_map = ms;
    TRY(prepare_plan(c, o, st, force_wgs));
    if (P.gc && P.gc_dup) {
        dim3_opts o2 = o;
        o2.flags |= kOptNoGc;
        TRY(prepare_plan(c, o2, st, force_wgs));
    }
    return DIM3_OK;
}

// Expansion arguments of the current bank's prepared plan.
dim3h::ExpandArgs make_args(dim3_ctx* c, bool fp_mode, void* out, uint64_t cap) {
    auto& P = c->bank().plan;
    auto& W = c->bank().ws;
    dim3h::ExpandArgs a;
    memset(&a, 0, sizeof(a));
    a.sc = (Scalars*)W[WS_SCALARS].p;
    a.fp_mode = fp_mode ? 1u : 0u;
    if (P.empty) return a;  // nx = 0: no tasks
    a.rowptrR = (const uint32_t*)W[WS_ROWPTR_R].p;
    a.colR = (const int32_t*)W[WS_COL_R].p;
    a.rowptrS = (const uint32_t*)W[WS_ROWPTR_S].p;
    a.colS = (const int32_t*)W[WS_COL_S].p;
    a.id2x = (const int32_t*)W[WS_ID2X].p;
    a.id2z = P.z_ident ? nullptr : (const int32_t*)W[WS_ID2Z].p;
    a.zlo = P.zlo;
    a.x2z = (P.flags & DIM3_OPT_EXCLUDE_EQUAL) ? (const int32_t*)W[WS_X2Z].p : nullptr;
    a.nx = P.nx;
    a.ntiles = P.ntiles;
    a.task_lo = 0;
    a.task_hi = P.nx * P.ntiles;
    a.row_lo = 0;
    a.row_hi = P.nx;
    a.tile_size = P.tile_size;
    a.tile_words = P.tile_words;
    a.touched_cap = P.touched_cap;
    a.stage_cap = P.stage_cap;
    // plain bit layout only where the dense seed copies whole words (every z
    // ID a product column); elsewhere the swizzle keeps hub z off one word
    a.swz = P.plain ? 0u : kTileBlockLog;
    a.fp_zw = fp_mode && P.fp_zw ? (const uint32_t*)W[WS_FP_ZW].p : nullptr;
    a.out = out;
    a.cap = fp_mode ? 0 : cap;
    if (P.core) a.splitS = (const uint32_t*)W[WS_CNT_SD].p;
    if (P.agg) {
        a.valR = (const int32_t*)W[WS_VAL_R].p;
        a.valS = (const int32_t*)W[WS_VAL_S].p;
    }
    if (P.dense != DIM3_DENSE_OFF) {
        a.hx = (const int32_t*)W[WS_HX].p;
        if (P.count_mode) {
            a.cint = (const int32_t*)W[WS_CINT].p;
        } else if (P.fused) {
            a.bkeys = (const uint32_t*)W[WS_DB].p;
            a.hid = (const int32_t*)W[WS_HID].p;
            // an XCD's L2 is 4 MiB: larger key bitmaps are read tile by tile
            a.tile_major = (uint64_t)P.K * P.dwords * 4u > (4ull << 20) ? 1u : 0u;
        } else {
            a.dbits = (const uint32_t*)W[WS_DBITS].p;
        }
        a.hz2zid = (const uint32_t*)W[WS_HZ2ZID].p;
        a.hz_tile = (const uint32_t*)W[WS_HZTILE].p;
        a.dwords = P.dwords;
        a.Nh = P.Nh;
    }
    return a;
}

int expand_args(dim3_ctx* c, const dim3h::ExpandArgs& a, hipStream_t st);

// Task lists of a value expansion: rows with at most 256 join paths are
// expanded one per wavefront (k_value_rows); rows with at most
// min(H / 2, touched_cap) paths go through an LDS hash table of H slots
// (the largest power of two whose 16-byte slots fit the tile's 12-byte
// groups), one workgroup task per row; the others are tiled.
int value_task_lists(dim3_ctx* c, dim3h::ExpandArgs& a, hipStream_t st, bool hash_rows) {
    auto& P = c->bank().plan;
    if (P.nx == 0) return DIM3_OK;
    uint32_t bits = 0;
    while (4ull * (2ull << bits) <= 3ull * P.tile_size) bits++;
    // a table under 64 slots: no hash rows
    const uint32_t limit = bits < 6 || !hash_rows ? 0u : std::min<uint32_t>((1u << bits) / 2u, P.touched_cap);
    uint32_t *bound, *fw, *pw, *fh, *ph, *wx, *hx, *bx;
    TRY(ws(c, WS_ROW_BOUND, P.nx, &bound));
    TRY(ws(c, WS_VT_FW, P.nx, &fw));
    TRY(ws(c, WS_VT_PW, P.nx, &pw));
    TRY(ws(c, WS_VT_FH, P.nx, &fh));
    TRY(ws(c, WS_VT_PH, P.nx, &ph));
    TRY(ws(c, WS_VT_WAVE, P.nx, &wx));
    TRY(ws(c, WS_VT_HASH, P.nx, &hx));
    TRY(ws(c, WS_VT_BIG, P.nx, &bx));
    uint64_t* part;
    TRY(ws(c, WS_PARTIALS, dim3h::scan_partials_needed((uint64_t)P.nx + 1), &part));
    Scalars* sc = (Scalars*)c->bank().ws[WS_SCALARS].p;
    // untiled rows (and the row bounds) read S by y alone: a compact row
    // index (the tiled one is ntiles times larger and its y boundaries lie
    // ntiles words apart)
    if (P.ntiles > 1) {
        uint32_t* ry;
        TRY(ws(c, WS_ROWPTR_Y, (uint64_t)P.ny + 1, &ry));
        dim3h::rowptr_y(a.rowptrS, P.ny, P.ntiles, ry, st);
        a.rowptrY = ry;
    } else {
        a.rowptrY = a.rowptrS;
    }
    dim3h::value_tasks(a, P.nz, limit, bound, fw, pw, fh, ph, part, &sc->n_wave, &sc->n_hash, wx, hx, bx, st);
    uint32_t* rn;
    uint64_t* ro;
    TRY(ws(c, WS_VT_ROWN, P.nx, &rn));
    TRY(ws(c, WS_VT_ROWOFF, P.nx, &ro));
    a.row_n = rn;
    a.row_off = ro;
    a.scan_part = part;
    a.wave_x = wx;
    a.hash_x = hx;
    a.big_x = bx;
    a.n_wave = &sc->n_wave;
    a.n_hash = &sc->n_hash;
    a.hash_bits = bits;
    return DIM3_OK;
}

// Expansion of the prepared plan into `out` (cap rows) or the fingerprint.
int expand(dim3_ctx* c, bool fp_mode, void* out, uint64_t cap, hipStream_t st) {
    auto& P = c->bank().plan;
    Scalars* sc = (Scalars*)c->bank().ws[WS_SCALARS].p;
    if (P.empty) {
        // reset the counters the expansion owns (a re-run after a capacity miss)
        hipMemsetAsync(&sc->out_n, 0, sizeof(uint64_t) * 5, st);
        HIPCHK(c, hipMemcpyAsync(c->h_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        return DIM3_OK;
    }
    if (P.gc) {
        // the dense part's terms over the box, then the light groups'
        const auto& W = c->bank().ws;
        hipMemsetAsync(&sc->out_n, 0, sizeof(uint64_t) * 5, st);
        hipMemsetAsync(&sc->task, 0, 2 * sizeof(uint32_t), st);
        HIPCHK(c, hipEventRecord(c->ev[4], st));
        int e = dim3h::gc_dense((const int8_t*)W[WS_GC_A8].p, (const uint64_t*)W[WS_GC_XK].p, P.gc_Mp,
                                (const int8_t*)W[WS_GC_B8].p, W[WS_GC_G].p, dim3h::gc_cols_padded(P.gc_Nh), kGcCols,
                                sc, st);
        if (e != 0) return fail(c, DIM3_ERR_DEVICE, "group-count dense launch: %s", hipGetErrorString((hipError_t)e));
        HIPCHK(c, hipEventRecord(c->ev[7], st));
        e = dim3h::gc_light((const uint32_t*)W[WS_ROWPTR_R].p, (const int32_t*)W[WS_COL_R].p,
                            (const uint32_t*)W[WS_GC_ROWL].p, (const int32_t*)W[WS_GC_COLL].p, P.ny,
                            (const uint64_t*)W[WS_GC_HM].p, W[WS_GC_ZINFO].p, (const int32_t*)W[WS_ID2X].p, P.nx,
                            P.ntiles, P.tile_size, P.nz, c->grid_cus(), sc, st);
        if (e != 0) return fail(c, DIM3_ERR_DEVICE, "group-count light launch: %s", hipGetErrorString((hipError_t)e));
        HIPCHK(c, hipEventRecord(c->ev[5], st));
        HIPCHK(c, hipMemcpyAsync(c->h_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        return DIM3_OK;
    }
    dim3h::ExpandArgs a = make_args(c, fp_mode, out, cap);
    if (P.agg) TRY(value_task_lists(c, a, st, true));
    // an all-sparse COUNT or join-project: rows of at most 256 paths one per
    // wavefront, the others as (x, z-tile) tasks (a million short rows in z
    // tiles would be a million near-empty tasks per tile)
    else if (P.dense == DIM3_DENSE_OFF && !P.core) TRY(value_task_lists(c, a, st, false));
    return expand_args(c, a, st);
}

// One expansion launch with prepared arguments (all rows, or one chunk of
// rows of a streamed query); the scalars are reset first and read back after.
int expand_args(dim3_ctx* c, const dim3h::ExpandArgs& a, hipStream_t st) {
    auto& P = c->bank().plan;
    Scalars* sc = (Scalars*)c->bank().ws[WS_SCALARS].p;
    // reset the counters the expansion owns (a re-run after a capacity miss)
    hipMemsetAsync(&sc->out_n, 0, sizeof(uint64_t) * 5, st);
    hipMemsetAsync(&sc->task, 0, 2 * sizeof(uint32_t), st);  // task, task2
    HIPCHK(c, hipEventRecord(c->ev[4], st));
    if (!P.agg && a.wave_x != nullptr) {
        const int e = dim3h::expand_short_rows(a, P.count_mode, c->grid_cus(), st);
        if (e != 0) return fail(c, DIM3_ERR_DEVICE, "short-row launch failed: %s", hipGetErrorString((hipError_t)e));
    }
    int e = P.agg          ? dim3h::expand_value(a, P.agg, P.vf32, P.block, c->grid_cus(), st)
            : P.count_mode ? dim3h::expand_count(a, P.block, c->grid_cus(), st)
                           : dim3h::expand_project(a, P.block, c->grid_cus(), st);
    if (e != 0) return fail(c, DIM3_ERR_DEVICE, "expansion launch failed: %s", hipGetErrorString((hipError_t)e));
    HIPCHK(c, hipEventRecord(c->ev[5], st));
    HIPCHK(c, hipMemcpyAsync(c->h_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return DIM3_OK;
}

// COUNT with the int8 product: some (x, y) or (z, y) multiplicity passed 127
bool dense_overflowed(dim3_ctx* c, const Scalars& s) {
    const auto& P = c->bank().plan;
    return P.dense == DIM3_DENSE_MFMA && P.count_mode && !P.empty && s.spare[1] != 0;
}

// Result of the current bank's query from its scalars; timing = the event
// times of a single query (a batch fills its own).
void fill_result(dim3_ctx* c, dim3_result* res, const Scalars& s, bool timing) {
    auto& P = c->bank().plan;
    res->n = P.empty ? 0 : s.out_n;
    res->swapped = P.swapped ? 1u : 0u;
    res->fp.count = res->n;
    res->fp.sum = P.empty ? 0 : s.fp_sum;
    const bool path_sums = P.agg_fp || P.gc;  // sum2 and csum from the join paths
    res->fp.sum2 = P.empty ? 0 : path_sums ? s.agg_sum2 : s.fp_sum2;
    res->fp.csum = P.empty ? 0 : path_sums ? P.out_j : s.fp_csum;
    dim3_stats& st = res->stats;
    memset(&st, 0, sizeof(st));
    st.n_r = P.nR;
    st.n_s = P.nS;
    st.nx = P.nx;
    st.ny = P.ny;
    st.nz = P.nz;
    st.out_j = P.out_j;
    st.paths_sparse = P.empty ? 0 : s.paths;
    st.dense_mode = P.gc ? (uint32_t)DIM3_DENSE_MFMA : (uint32_t)P.dense;
    st.heavy_keys = P.gc ? P.gc_K : P.dense ? P.K : 0;
    st.nx_heavy = P.gc ? P.gc_Mh : P.dense ? P.Mh : 0;
    st.nz_heavy = P.gc ? P.gc_Nh : P.dense ? P.Nh : 0;
    st.tile_bits = P.empty ? 0 : P.tile_size;
    st.n_tiles = P.ntiles;
    st.block = (uint32_t)P.block;
    st.ms_map = P.ms_map;
    if (!P.empty && timing) {
        float ms = 0;
        hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
        st.ms_csr = ms;
        hipEventElapsedTime(&ms, c->ev[3], c->ev[6]);
        st.ms_dense = ms;
        if (P.gc) {  // the MFMA box kernel (part of the expansion)
            hipEventElapsedTime(&ms, c->ev[4], c->ev[7]);
            st.ms_dense = ms;
        }
        hipEventElapsedTime(&ms, c->ev[4], c->ev[5]);
        st.ms_expand = ms;
        hipEventElapsedTime(&ms, c->ev[0], c->ev[5]);
        st.ms_total = ms;
    }
}

int check_opts(dim3_ctx* c, const dim3_opts* opts, dim3_opts* o) {
    dim3_opts_default(o);
    if (opts) *o = *opts;
    if (o->output_mode > DIM3_OUT_FINGERPRINT) return fail(c, DIM3_ERR_INVALID, "bad output_mode %u", o->output_mode);
    if (o->dense_mode < -1 || o->dense_mode > 2) return fail(c, DIM3_ERR_INVALID, "bad dense_mode %d", o->dense_mode);
    if (o->tile_bits < 0 || o->tile_bits > (1 << 30)) return fail(c, DIM3_ERR_INVALID, "bad tile_bits %d", o->tile_bits);
    if (!(o->block == 0 || o->block == 256 || o->block == 512 || o->block == 1024))
        return fail(c, DIM3_ERR_INVALID, "bad block %d", o->block);
    return DIM3_OK;
}

// ---------------------------------------------------------------------------
// f1 (dim3.hpp:41-50): the classical path for low OUT_J
// ---------------------------------------------------------------------------
// The reference's f1 takes its classical radix hash join (dim3.hpp:273-391)
// when the estimated OUT_J is under about 2.2 n.  Here the classical path is
// a sort of one 64-bit (x, z) key per join path (dim3h::classical_join), a
// few launches and two host synchronisations; the general plan maps x, y and
// z to dense IDs and builds two CSRs first, work that only pays when the
// rows are long, and whose short rows (one per wavefront) cost per row.
// Measured (materialised join-project, ms general / classical): uniform x, y,
// z over 1e7 values, 1e7 tuples per side (about one path per x row) at 0.5,
// 2, 4, 8 paths per tuple 11.6 / 7.9, 14.7 / 10.1, 14.6 / 13.3, 14.7 / 19.2;
// the reference generator's data with a 1e7 key domain (x mod 1e6: ten paths
// per x row, OUT_J 1e7) 5.6 / 7.6; 1e6 tuples per side 1.55 / 2.3; 1e5
// 0.53 / 0.47.  COUNT fingerprints tie (8.9 / 8.7 at one path per row) and
// stay on the general plan.  So AUTO plans of at least kClassicalMinTuples
// tuples take it when EstimateJoinCardinality's estimate is at most
// kClassicalPathsPerTuple paths per input tuple and kClassicalPathsPerRow
// per x row, the rows bounded by nR and by R.x's value range (the estimate
// and a min / max pass, two small launches and synchronisations, run only
// for those sizes).
constexpr double kClassicalPathsPerTuple = 4.0;
constexpr double kClassicalPathsPerRow = 3.0;
constexpr uint64_t kClassicalMinTuples = 8000000;

// after the swap rule: nR >= nS
int want_classical(dim3_ctx* c, const dim3_opts& o, bool count_mode, const int2* dR, uint64_t nR, const int2* dS,
                   uint64_t nS, hipStream_t st, bool* yes) {
    *yes = false;
    if (o.flags & DIM3_OPT_NO_CLASSICAL) return DIM3_OK;
    if (nR == 0 || nS == 0) return DIM3_OK;  // empty: the general plan's early exit
    if (o.flags & DIM3_OPT_CLASSICAL) {
        *yes = true;
        return DIM3_OK;
    }
    if (o.dense_mode != DIM3_DENSE_AUTO || o.tile_bits || o.block || o.heavy_k || o.dense_core >= 0) return DIM3_OK;
    if (nR + nS < kClassicalMinTuples) return DIM3_OK;
    if (count_mode && o.output_mode == DIM3_OUT_FINGERPRINT) return DIM3_OK;
    int64_t est = 0;
    TRY(estimate_on(c, dR, nR, dS, nS, st, &est));
    if ((double)est > kClassicalPathsPerTuple * (double)(nR + nS) || est >= (1ll << 30)) return DIM3_OK;
    Scalars* sc;
    TRY(ws(c, WS_SCALARS, 1, &sc));
    dim3h::minmax(dR, nR, dS, 0, sc->mm, st);
    int32_t mm[2] = {0, 0};
    HIPCHK(c, dim3h::copy_wait(mm, sc->mm, sizeof(mm), hipMemcpyDeviceToHost, st));
    const double rows = std::min<double>((double)nR, (double)mm[1] - (double)mm[0] + 1.0);
    *yes = (double)est <= kClassicalPathsPerRow * rows;
    return DIM3_OK;
}

// The classical path of one project / COUNT query over device relations;
// out_for(rows, &buf, &cap) names the output buffer (materialise).
int run_classical(dim3_ctx* c, bool count_mode, const int2* dR, uint64_t nR, const int2* dS, uint64_t nS,
                  const dim3_opts& o, bool fp, const std::function<int(uint64_t, void**, uint64_t*)>& out_for,
                  hipStream_t st, dim3_result* res) {
    const bool swapped = !(o.flags & DIM3_OPT_NO_SWAP) && nR < nS;
    if (swapped) {
        std::swap(dR, dS);
        std::swap(nR, nS);
    }
    Scalars* sc;
    TRY(ws(c, WS_SCALARS, 1, &sc));
    HIPCHK(c, hipMemsetAsync(sc, 0, sizeof(Scalars), st));
    HIPCHK(c, hipEventRecord(c->ev[0], st));
    std::string e;
    const int rc = dim3h::classical_join(dR, nR, dS, nS, count_mode, (o.flags & DIM3_OPT_EXCLUDE_EQUAL) != 0, fp,
                                         out_for, sc, st, &e);
    if (rc != DIM3_OK) return fail(c, rc, "%s", e.c_str());
    HIPCHK(c, hipEventRecord(c->ev[5], st));
    HIPCHK(c, dim3h::copy_wait(c->h_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    const Scalars& s = *c->h_sc;
    res->n = s.out_n;
    res->swapped = swapped ? 1u : 0u;
    res->fp.count = s.out_n;
    res->fp.sum = s.fp_sum;
    res->fp.sum2 = s.fp_sum2;
    res->fp.csum = s.fp_csum;
    dim3_stats& st_ = res->stats;
    memset(&st_, 0, sizeof(st_));
    st_.n_r = nR;
    st_.n_s = nS;
    st_.out_j = s.paths;
    st_.paths_sparse = s.paths;
    st_.dense_mode = DIM3_DENSE_OFF;
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev[0], c->ev[5]);
    st_.ms_expand = ms;
    st_.ms_total = ms;
    return DIM3_OK;
}

int run_dev(dim3_ctx* c, bool count_mode, const int32_t* d_R, uint64_t nR, const int32_t* d_S, uint64_t nS,
            const dim3_opts* opts, void* d_out, uint64_t cap, void* stream, dim3_result* res) {
    if (!c || !res) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    memset(res, 0, sizeof(*res));
    if ((nR && !d_R) || (nS && !d_S)) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    dim3_opts o;
    TRY(check_opts(c, opts, &o));
    const bool fp = o.output_mode == DIM3_OUT_FINGERPRINT;
    if (!fp && cap && !d_out) return fail(c, DIM3_ERR_INVALID, "null output buffer");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    c->qstream = st;
    if (count_mode && o.dense_mode == DIM3_DENSE_BITOR)
        return fail(c, DIM3_ERR_UNSUPPORTED, "DIM3_DENSE_BITOR is join-project only");
    {
        const bool sw = !(o.flags & DIM3_OPT_NO_SWAP) && nR < nS;
        bool classical = false;
        TRY(want_classical(c, o, count_mode, (const int2*)(sw ? d_S : d_R), sw ? nS : nR,
                           (const int2*)(sw ? d_R : d_S), sw ? nR : nS, st, &classical));
        if (classical) {
            auto out_for = [&](uint64_t, void** buf, uint64_t* bcap) {
                *buf = d_out;
                *bcap = cap;
                return DIM3_OK;
            };
            TRY(run_classical(c, count_mode, (const int2*)d_R, nR, (const int2*)d_S, nS, o, fp, out_for, st, res));
            if (!fp && res->n > cap)
                return fail(c, DIM3_ERR_CAPACITY, "output needs %llu rows, buffer holds %llu",
                            (unsigned long long)res->n, (unsigned long long)cap);
            return DIM3_OK;
        }
    }
    TRY(prepare(c, count_mode, (const int2*)d_R, nR, (const int2*)d_S, nS, o, st));
    TRY(expand(c, fp, d_out, cap, st));
    if (dense_overflowed(c, *c->h_sc)) {  // a multiplicity above 127: redo without the int8 product
        o.dense_mode = DIM3_DENSE_OFF;
        TRY(prepare(c, count_mode, (const int2*)d_R, nR, (const int2*)d_S, nS, o, st));
        TRY(expand(c, fp, d_out, cap, st));
    }
    fill_result(c, res, *c->h_sc, true);
    if (!fp && res->n > cap)
        return fail(c, DIM3_ERR_CAPACITY, "output needs %llu rows, buffer holds %llu", (unsigned long long)res->n,
                    (unsigned long long)cap);
    return DIM3_OK;
}

int run_host(dim3_ctx* c, bool count_mode, const int32_t* R, uint64_t nR, const int32_t* S, uint64_t nS,
             const dim3_opts* opts, dim3_result* res) {
    if (!c || !res) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    memset(res, 0, sizeof(*res));
    if ((nR && !R) || (nS && !S)) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    dim3_opts o;
    TRY(check_opts(c, opts, &o));
    const bool fp = o.output_mode == DIM3_OUT_FINGERPRINT;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    c->qstream = st;
    int2 *dR, *dS;
    TRY(ws(c, WS_IN_R, std::max<uint64_t>(nR, 1), &dR));
    TRY(ws(c, WS_IN_S, std::max<uint64_t>(nS, 1), &dS));
    if (nR) HIPCHK(c, hipMemcpyAsync(dR, R, nR * 8, hipMemcpyHostToDevice, st));
    if (nS) HIPCHK(c, hipMemcpyAsync(dS, S, nS * 8, hipMemcpyHostToDevice, st));
    if (count_mode && o.dense_mode == DIM3_DENSE_BITOR)
        return fail(c, DIM3_ERR_UNSUPPORTED, "DIM3_DENSE_BITOR is join-project only");
    const size_t row_bytes = count_mode ? sizeof(dim3_count_row) : 8;
    {
        const bool sw = !(o.flags & DIM3_OPT_NO_SWAP) && nR < nS;
        bool classical = false;
        TRY(want_classical(c, o, count_mode, sw ? dS : dR, sw ? nS : nR, sw ? dR : dS, sw ? nR : nS, st,
                           &classical));
        if (classical) {
            void* dout = nullptr;
            auto out_for = [&](uint64_t rows, void** buf, uint64_t* bcap) {
                const int rc = ws_get(c, WS_OUT, std::max<uint64_t>(rows, 1) * row_bytes, &dout);
                *buf = dout;
                *bcap = rows;
                return rc;
            };
            TRY(run_classical(c, count_mode, dR, nR, dS, nS, o, fp, out_for, st, res));
            if (!fp && res->n) {
                void* h = malloc(res->n * row_bytes);
                if (!h) return fail(c, DIM3_ERR_OOM, "host malloc of %llu rows", (unsigned long long)res->n);
                hipError_t e = dim3h::copy_wait(h, dout, res->n * row_bytes, hipMemcpyDeviceToHost, st);
                if (e != hipSuccess) {
                    free(h);
                    return fail(c, DIM3_ERR_DEVICE, "D2H copy: %s", hipGetErrorString(e));
                }
                if (count_mode)
                    res->rows = (dim3_count_row*)h;
                else
                    res->pairs = (int32_t*)h;
            }
            return DIM3_OK;
        }
    }
    TRY(prepare(c, count_mode, dR, nR, dS, nS, o, st));
    void* dout = nullptr;
    uint64_t cap = 0;
    if (!fp && !c->bank().plan.empty) {
        // rows <= min(OUT_J, nx*nz); size for that bound when it is modest,
        // otherwise guess and re-run the expansion at the exact size
        uint64_t bound = std::min<uint64_t>(c->bank().plan.out_j, (uint64_t)c->bank().plan.nx * c->bank().plan.nz);
        const uint64_t room = c->budget > c->held ? (c->budget - c->held) / row_bytes / 2 : 0;
        cap = std::min<uint64_t>(bound, std::max<uint64_t>(room, 1u << 20));
        cap = std::max<uint64_t>(cap, 1);
        TRY(ws_get(c, WS_OUT, cap * row_bytes, &dout));
        cap = c->bank().ws[WS_OUT].bytes / row_bytes;
    }
    TRY(expand(c, fp, dout, cap, st));
    if (dense_overflowed(c, *c->h_sc)) {  // a multiplicity above 127: redo without the int8 product
        o.dense_mode = DIM3_DENSE_OFF;
        TRY(prepare(c, count_mode, dR, nR, dS, nS, o, st));
        TRY(expand(c, fp, dout, cap, st));
    }
    if (!fp && !c->bank().plan.empty && c->h_sc->out_n > cap) {
        cap = c->h_sc->out_n;
        TRY(ws_get(c, WS_OUT, cap * row_bytes, &dout));
        TRY(expand(c, fp, dout, cap, st));
    }
    fill_result(c, res, *c->h_sc, true);
    if (!fp && res->n) {
        void* h = malloc(res->n * row_bytes);
        if (!h) return fail(c, DIM3_ERR_OOM, "host malloc of %llu rows", (unsigned long long)res->n);
        hipError_t e = dim3h::copy_wait(h, dout, res->n * row_bytes, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) {
            free(h);
            return fail(c, DIM3_ERR_DEVICE, "D2H copy: %s", hipGetErrorString(e));
        }
        if (count_mode)
            res->rows = (dim3_count_row*)h;
        else
            res->pairs = (int32_t*)h;
    }
    return DIM3_OK;
}

// Value aggregates (DIM3_AGG_*): the COUNT pipeline with the tuple values
// carried into the CSRs and the value expansion; never a dense part.
int check_value_query(dim3_ctx* c, int agg, uint64_t nR, const int32_t* vR, uint64_t nS, const int32_t* vS) {
    if (agg < DIM3_AGG_SUM || agg > DIM3_AGG_AVG) return fail(c, DIM3_ERR_INVALID, "bad aggregate %d", agg);
    if ((nR && !vR) || (nS && !vS)) return fail(c, DIM3_ERR_INVALID, "null value pointer");
    return DIM3_OK;
}

int run_value_dev(dim3_ctx* c, int agg, bool f32, const int32_t* d_R, const int32_t* d_vR, uint64_t nR,
                  const int32_t* d_S, const int32_t* d_vS, uint64_t nS, const dim3_opts* opts, void* d_out,
                  uint64_t cap, void* stream, dim3_result* res) {
    if (!c || !res) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    memset(res, 0, sizeof(*res));
    if ((nR && !d_R) || (nS && !d_S)) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    TRY(check_value_query(c, agg, nR, d_vR, nS, d_vS));
    dim3_opts o;
    TRY(check_opts(c, opts, &o));
    o.dense_mode = DIM3_DENSE_OFF;
    const bool fp = o.output_mode == DIM3_OUT_FINGERPRINT;
    if (!fp && cap && !d_out) return fail(c, DIM3_ERR_INVALID, "null output buffer");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    c->qstream = st;
    ValueSpec vals;
    vals.agg = agg;
    vals.f32 = f32;
    vals.vR = d_vR;
    vals.vS = d_vS;
    TRY(prepare(c, true, (const int2*)d_R, nR, (const int2*)d_S, nS, o, st, 0, vals));
    TRY(expand(c, fp, d_out, cap, st));
    fill_result(c, res, *c->h_sc, true);
    if (!fp && res->n > cap)
        return fail(c, DIM3_ERR_CAPACITY, "output needs %llu rows, buffer holds %llu", (unsigned long long)res->n,
                    (unsigned long long)cap);
    return DIM3_OK;
}

// A value query (agg 0: COUNT) on device-resident inputs, rows copied to a
// malloc'd host buffer (res->rows, or res->rows_f32 for float values).
int value_query_to_host(dim3_ctx* c, int agg, bool f32, const int2* dR, const int32_t* dvR, uint64_t nR,
                        const int2* dS, const int32_t* dvS, uint64_t nS, dim3_opts o, hipStream_t st,
                        dim3_result* res) {
    const bool fp = o.output_mode == DIM3_OUT_FINGERPRINT;
    ValueSpec vals;
    vals.agg = agg;
    vals.f32 = f32;
    vals.vR = dvR;
    vals.vS = dvS;
    TRY(prepare(c, true, dR, nR, dS, nS, o, st, 0, vals));
    const size_t row_bytes = f32 ? sizeof(dim3_value_row_f32) : sizeof(dim3_count_row);
    const auto& P = c->bank().plan;
    void* dout = nullptr;
    uint64_t cap = 0;
    if (!fp && !P.empty) {
        // groups <= min(OUT_J, nx * nz): size for that bound when it is
        // modest, otherwise guess and re-run the expansion at the exact size
        const uint64_t bound = std::min<uint64_t>(P.out_j, (uint64_t)P.nx * P.nz);
        const uint64_t room = c->budget > c->held ? (c->budget - c->held) / row_bytes / 2 : 0;
        cap = std::max<uint64_t>(std::min<uint64_t>(bound, std::max<uint64_t>(room, 1u << 20)), 1);
        TRY(ws_get(c, WS_OUT, cap * row_bytes, &dout));
        cap = c->bank().ws[WS_OUT].bytes / row_bytes;
    }
    TRY(expand(c, fp, dout, cap, st));
    if (dense_overflowed(c, *c->h_sc)) {  // COUNT: a multiplicity above 127, redo without the int8 product
        o.dense_mode = DIM3_DENSE_OFF;
        TRY(prepare(c, true, dR, nR, dS, nS, o, st, 0, vals));
        TRY(expand(c, fp, dout, cap, st));
    }
    if (!fp && !P.empty && c->h_sc->out_n > cap) {
        cap = c->h_sc->out_n;
        TRY(ws_get(c, WS_OUT, cap * row_bytes, &dout));
        TRY(expand(c, fp, dout, cap, st));
    }
    fill_result(c, res, *c->h_sc, true);
    if (!fp && res->n) {
        void* h = malloc(res->n * row_bytes);
        if (!h) return fail(c, DIM3_ERR_OOM, "host malloc of %llu rows", (unsigned long long)res->n);
        hipError_t e = dim3h::copy_wait(h, dout, res->n * row_bytes, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) {
            free(h);
            return fail(c, DIM3_ERR_DEVICE, "D2H copy: %s", hipGetErrorString(e));
        }
        if (f32) res->rows_f32 = (dim3_value_row_f32*)h;
        else res->rows = (dim3_count_row*)h;
    }
    return DIM3_OK;
}

int run_value_host(dim3_ctx* c, int agg, bool f32, const int32_t* R, const int32_t* vR, uint64_t nR,
                   const int32_t* S, const int32_t* vS, uint64_t nS, const dim3_opts* opts, dim3_result* res) {
    if (!c || !res) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    memset(res, 0, sizeof(*res));
    if ((nR && !R) || (nS && !S)) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    TRY(check_value_query(c, agg, nR, vR, nS, vS));
    dim3_opts o;
    TRY(check_opts(c, opts, &o));
    o.dense_mode = DIM3_DENSE_OFF;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    c->qstream = st;
    int2 *dR, *dS;
    int32_t *dvR, *dvS;
    TRY(ws(c, WS_IN_R, std::max<uint64_t>(nR, 1), &dR));
    TRY(ws(c, WS_IN_S, std::max<uint64_t>(nS, 1), &dS));
    TRY(ws(c, WS_IN_VR, std::max<uint64_t>(nR, 1), &dvR));
    TRY(ws(c, WS_IN_VS, std::max<uint64_t>(nS, 1), &dvS));
    if (nR) HIPCHK(c, hipMemcpyAsync(dR, R, nR * 8, hipMemcpyHostToDevice, st));
    if (nS) HIPCHK(c, hipMemcpyAsync(dS, S, nS * 8, hipMemcpyHostToDevice, st));
    if (nR) HIPCHK(c, hipMemcpyAsync(dvR, vR, nR * 4, hipMemcpyHostToDevice, st));
    if (nS) HIPCHK(c, hipMemcpyAsync(dvS, vS, nS * 4, hipMemcpyHostToDevice, st));
    return value_query_to_host(c, agg, f32, dR, dvR, nR, dS, dvS, nS, o, st, res);
}

// Value aggregate (agg 0: COUNT) over the reference's TUPLE arrays: the
// tuples are copied as they are and split into (first, key) pairs and int32
// values (float bit patterns) on the device.
int run_tuples(dim3_ctx* c, int agg, const void* R, uint64_t nR, const void* S, uint64_t nS, uint32_t stride,
               uint32_t voff, uint32_t vtype, const dim3_opts* opts, dim3_result* res) {
    if (!c || !res) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    memset(res, 0, sizeof(*res));
    if ((nR && !R) || (nS && !S)) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    if (agg < 0 || agg > DIM3_AGG_AVG) return fail(c, DIM3_ERR_INVALID, "bad aggregate %d", agg);
    const uint32_t vbytes = vtype == DIM3_VALUE_I64 ? 8u : 4u;
    if (vtype > DIM3_VALUE_F32 || stride % 4u || voff % 4u || voff < 8u || voff + vbytes > stride)
        return fail(c, DIM3_ERR_INVALID, "bad tuple layout (stride %u, value at %u, type %u)", stride, voff, vtype);
    dim3_opts o;
    TRY(check_opts(c, opts, &o));
    if (agg) o.dense_mode = DIM3_DENSE_OFF;
    if (!agg && o.dense_mode == DIM3_DENSE_BITOR)
        return fail(c, DIM3_ERR_UNSUPPORTED, "DIM3_DENSE_BITOR is join-project only");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    c->qstream = st;
    int2 *dR, *dS;
    int32_t *dvR, *dvS;
    uint8_t* raw;
    Scalars* sc;
    TRY(ws(c, WS_IN_R, std::max<uint64_t>(nR, 1), &dR));
    TRY(ws(c, WS_IN_S, std::max<uint64_t>(nS, 1), &dS));
    TRY(ws(c, WS_IN_VR, std::max<uint64_t>(nR, 1), &dvR));
    TRY(ws(c, WS_IN_VS, std::max<uint64_t>(nS, 1), &dvS));
    TRY(ws(c, WS_SCALARS, 1, &sc));
    HIPCHK(c, hipMemsetAsync(sc, 0, sizeof(Scalars), st));
    uint32_t* bad = (uint32_t*)&sc->spare[2];
    const uint64_t nmax = std::max<uint64_t>(std::max(nR, nS), 1);
    TRY(ws(c, WS_OUT, nmax * stride, &raw));  // staging; the output reuses the slot afterwards
    if (nR) {
        HIPCHK(c, hipMemcpyAsync(raw, R, nR * stride, hipMemcpyHostToDevice, st));
        dim3h::split_tuples(raw, nR, stride, voff, vtype, dR, dvR, bad, st);
    }
    if (nS) {
        HIPCHK(c, hipMemcpyAsync(raw, S, nS * stride, hipMemcpyHostToDevice, st));
        dim3h::split_tuples(raw, nS, stride, voff, vtype, dS, dvS, bad, st);
    }
    LAUNCHCHK(c, "split tuples");
    uint32_t hbad = 0;
    HIPCHK(c, hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (hbad && agg) return fail(c, DIM3_ERR_INVALID, "a 64-bit tuple value does not fit int32");
    return value_query_to_host(c, agg, agg && vtype == DIM3_VALUE_F32, dR, agg ? dvR : nullptr, nR, dS,
                               agg ? dvS : nullptr, nS, o, st, res);
}

// Streamed output: the plan's rows are expanded in chunks of consecutive x
// rows whose output bound (row_bounds) fits a device buffer of
// max(chunk_rows, largest row bound) rows; each chunk is copied to a pinned
// host buffer and handed to the callback before the next one runs.
int run_stream(dim3_ctx* c, bool count_mode, const int32_t* R, uint64_t nR, const int32_t* S, uint64_t nS,
               const dim3_opts* opts, uint64_t chunk_rows, dim3_rows_cb cb, void* user, dim3_result* res, int agg = 0,
               bool f32 = false, const int32_t* vR = nullptr, const int32_t* vS = nullptr) {
    if (!c || !res) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    memset(res, 0, sizeof(*res));
    if (!cb) return fail(c, DIM3_ERR_INVALID, "null row callback");
    if ((nR && !R) || (nS && !S)) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    if (agg) TRY(check_value_query(c, agg, nR, vR, nS, vS));
    dim3_opts o;
    TRY(check_opts(c, opts, &o));
    o.output_mode = DIM3_OUT_MATERIALIZE;
    if (agg) o.dense_mode = DIM3_DENSE_OFF;
    if (count_mode && o.dense_mode == DIM3_DENSE_BITOR)
        return fail(c, DIM3_ERR_UNSUPPORTED, "DIM3_DENSE_BITOR is join-project only");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    c->qstream = st;
    int2 *dR, *dS;
    TRY(ws(c, WS_IN_R, std::max<uint64_t>(nR, 1), &dR));
    TRY(ws(c, WS_IN_S, std::max<uint64_t>(nS, 1), &dS));
    if (nR) HIPCHK(c, hipMemcpyAsync(dR, R, nR * 8, hipMemcpyHostToDevice, st));
    if (nS) HIPCHK(c, hipMemcpyAsync(dS, S, nS * 8, hipMemcpyHostToDevice, st));
    ValueSpec vals;
    if (agg) {
        int32_t *dvR, *dvS;
        TRY(ws(c, WS_IN_VR, std::max<uint64_t>(nR, 1), &dvR));
        TRY(ws(c, WS_IN_VS, std::max<uint64_t>(nS, 1), &dvS));
        if (nR) HIPCHK(c, hipMemcpyAsync(dvR, vR, nR * 4, hipMemcpyHostToDevice, st));
        if (nS) HIPCHK(c, hipMemcpyAsync(dvS, vS, nS * 4, hipMemcpyHostToDevice, st));
        vals.agg = agg;
        vals.f32 = f32;
        vals.vR = dvR;
        vals.vS = dvS;
    }
    TRY(prepare(c, count_mode, dR, nR, dS, nS, o, st, 0, vals));
    {
        // a multiplicity above 127 in the int8 product: plan again all-sparse
        Scalars* sc = (Scalars*)c->bank().ws[WS_SCALARS].p;
        HIPCHK(c, hipMemcpyAsync(c->h_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (dense_overflowed(c, *c->h_sc)) {
            o.dense_mode = DIM3_DENSE_OFF;
            TRY(prepare(c, count_mode, dR, nR, dS, nS, o, st, 0, vals));
        }
    }
    const auto& P = c->bank().plan;
    const size_t row_bytes = P.agg && P.vf32 ? sizeof(dim3_value_row_f32) : P.count_mode ? sizeof(dim3_count_row) : 8;
    uint64_t total = 0, paths = 0;
    double ms_expand = 0;
    if (!P.empty) {
        uint32_t* bound;
        TRY(ws(c, WS_ROW_BOUND, P.nx, &bound));
        dim3h::ExpandArgs a = make_args(c, false, nullptr, 0);
        dim3h::row_bounds(a, P.nz, bound, st);
        std::vector<uint32_t> hb(P.nx);
        HIPCHK(c, hipMemcpyAsync(hb.data(), bound, (size_t)P.nx * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        uint64_t cap = chunk_rows ? chunk_rows : (1ull << 24);
        for (uint32_t b : hb) cap = std::max<uint64_t>(cap, b);
        void* dout;
        TRY(ws_get(c, WS_OUT, cap * row_bytes, &dout));
        if (c->h_stream_bytes < cap * row_bytes) {
            if (c->h_stream) dim3h::host_free(c->h_stream);
            c->h_stream = nullptr;
            c->h_stream_bytes = 0;
            HIPCHK(c, dim3h::host_malloc(&c->h_stream, cap * row_bytes));
            c->h_stream_bytes = cap * row_bytes;
        }
        for (uint32_t x0 = 0; x0 < P.nx;) {
            uint32_t x1 = x0;
            uint64_t acc = 0;
            while (x1 < P.nx && (x1 == x0 || acc + hb[x1] <= cap)) acc += hb[x1++];
            a = make_args(c, false, dout, cap);
            a.task_lo = x0 * P.ntiles;
            a.task_hi = x1 * P.ntiles;
            a.row_lo = x0;
            a.row_hi = x1;
            // the chunk's rows by size, as a whole query's (value aggregates;
            // the short rows of an all-sparse plan)
            if (P.agg) TRY(value_task_lists(c, a, st, true));
            else if (P.dense == DIM3_DENSE_OFF && !P.core) TRY(value_task_lists(c, a, st, false));
            TRY(expand_args(c, a, st));
            const uint64_t n = c->h_sc->out_n;
            if (n > cap)
                return fail(c, DIM3_ERR_DEVICE, "rows %u..%u produced %llu rows above their bound %llu", x0, x1,
                            (unsigned long long)n, (unsigned long long)cap);
            float ms = 0;
            hipEventElapsedTime(&ms, c->ev[4], c->ev[5]);
            ms_expand += ms;
            paths += c->h_sc->paths;
            if (n) {
                HIPCHK(c, hipMemcpyAsync(c->h_stream, dout, n * row_bytes, hipMemcpyDeviceToHost, st));
                HIPCHK(c, hipStreamSynchronize(st));
                if (cb(user, c->h_stream, n) != 0) return fail(c, DIM3_ERR_CANCELLED, "stopped by the row callback");
            }
            total += n;
            x0 = x1;
        }
    }
    fill_result(c, res, *c->h_sc, false);
    res->n = total;
    res->fp.count = total;
    res->stats.paths_sparse = paths;
    res->stats.ms_expand = ms_expand;
    return DIM3_OK;
}

// Many instances: each is planned (mapping, CSR, dense part) into its own
// bank, then all of them expand in one persistent launch fed by one device
// work counter.  The launch (scalar reset, expansion, scalar gather and the
// copy to the host) is captured as a HIP graph and replayed once.
// A batch of at least kPipeMin instances is planned in two chunks, largest
// instances first: chunk A holds the largest instances up to half of the
// batch's tuples, chunk B the rest.  Chunk A's expansion (most of the
// batch's join paths) runs on a stream masked to kPipeShare of the compute
// units while the host plans chunk B on the call's stream, whose small
// kernels take the other units; chunk B's expansion then runs on the call's
// stream beside chunk A's tail.  (Only on the context's own stream: a batch
// on a caller's stream, a server worker's CU-masked one say, is one chunk.)
constexpr uint32_t kPipeMin = 16;
constexpr double kPipeShare = 0.875;

int pipe_stream(dim3_ctx* c, hipStream_t* out) {
    if (!c->pipe_stream) {
        const uint32_t cus = (uint32_t)std::max(2, c->cu_count);
        const uint32_t use = std::min<uint32_t>(cus - 1, std::max<uint32_t>(1, (uint32_t)(kPipeShare * cus + 0.5)));
        std::vector<uint32_t> mask((cus + 31) / 32, 0);
        for (uint32_t i = 0; i < use; i++) mask[i / 32] |= 1u << (i % 32);
        HIPCHK(c, hipExtStreamCreateWithCUMask(&c->pipe_stream, (uint32_t)mask.size(), mask.data()));
        c->pipe_cus = (int)use;
    }
    *out = c->pipe_stream;
    return DIM3_OK;
}

int run_batch(dim3_ctx* c, bool count_mode, const dim3_instance* inst, uint32_t n, const dim3_opts* opts,
              void* stream, dim3_result* results) {
    if (!c || (n && (!inst || !results))) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    if (n == 0) return DIM3_OK;
    memset(results, 0, sizeof(dim3_result) * n);
    dim3_opts o;
    TRY(check_opts(c, opts, &o));
    const bool fp = o.output_mode == DIM3_OUT_FINGERPRINT;
    if (count_mode && o.dense_mode == DIM3_DENSE_BITOR)
        return fail(c, DIM3_ERR_UNSUPPORTED, "DIM3_DENSE_BITOR is join-project only");
    for (uint32_t i = 0; i < n; i++) {
        if ((inst[i].n_r && !inst[i].R_xy) || (inst[i].n_s && !inst[i].S_zy))
            return fail(c, DIM3_ERR_INVALID, "instance %u: null relation pointer", i);
        if (!fp && inst[i].cap && !inst[i].out) return fail(c, DIM3_ERR_INVALID, "instance %u: null output", i);
    }
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    c->qstream = st;
    struct Reset {  // leave the context on bank 0 whatever happens
        dim3_ctx* c;
        ~Reset() { c->use_bank(0); }
    } reset{c};

    // plan every instance into bank i + 1 with one shared workgroup shape
    const int block = o.block ? o.block : 256;
    dim3_opts ob = o;
    ob.block = block;
    const int wgs = block == 256 ? 4 : block == 512 ? 2 : 1;
    std::vector<dim3h::ExpandArgs> args(n);
    std::vector<Scalars*> scl(n);
    // the instances largest first, and the chunk split
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return inst[a].n_r + inst[a].n_s > inst[b].n_r + inst[b].n_s;
    });
    uint64_t tuples = 0, acc = 0;
    for (uint32_t i = 0; i < n; i++) tuples += inst[i].n_r + inst[i].n_s;
    uint32_t na = n;
    if (n >= kPipeMin && stream == nullptr) {  // a caller's stream keeps one chunk (its CU mask is the caller's)
        na = 0;
        while (na < n && 2 * acc < tuples) {
            acc += inst[order[na]].n_r + inst[order[na]].n_s;
            na++;
        }
        na = std::max<uint32_t>(1, std::min<uint32_t>(na, n - 1));
    }
    // each planning stage over every instance of a chunk, then one synchronisation
    auto plan_chunk = [&](uint32_t first, uint32_t last) -> int {
        for (uint32_t k = first; k < last; k++) {
            const uint32_t i = order[k];
            c->use_bank(i + 1);
            TRY(prepare_start(c, count_mode, (const int2*)inst[i].R_xy, inst[i].n_r, (const int2*)inst[i].S_zy,
                              inst[i].n_s, ob, st));
        }
        HIPCHK(c, hipStreamSynchronize(st));
        for (uint32_t k = first; k < last; k++) {
            c->use_bank(order[k] + 1);
            TRY(prepare_map(c, ob, st));
        }
        HIPCHK(c, hipStreamSynchronize(st));
        for (uint32_t k = first; k < last; k++) {
            const uint32_t i = order[k];
            c->use_bank(i + 1);
            TRY(prepare_plan(c, ob, st, wgs));
            args[k] = make_args(c, fp, inst[i].out, inst[i].cap);
            scl[k] = args[k].sc;
        }
        c->use_bank(0);
        return DIM3_OK;
    };
    // the expansion kernel of every instance (a COUNT fingerprint from path
    // sums expands as a join-project)
    const bool kcount =
        count_mode && !(fp && (o.flags & DIM3_OPT_PATH_SUMS) && !(o.flags & DIM3_OPT_EXCLUDE_EQUAL));
    dim3h::ExpandArgs* d_args;
    uint32_t* d_start;
    Scalars** d_scl;
    Scalars* d_scout;
    TRY(ws(c, WS_BATCH_ARGS, n, &d_args));
    TRY(ws(c, WS_BATCH_START, (uint64_t)n + 4, &d_start));  // per chunk: its starts and a work counter
    TRY(ws(c, WS_BATCH_SCL, n, &d_scl));
    TRY(ws(c, WS_BATCH_SCOUT, n, &d_scout));
    if (c->h_batch_cap < n) {
        if (c->h_batch) dim3h::host_free(c->h_batch);
        c->h_batch = nullptr;
        c->h_batch_cap = 0;
        const uint32_t cap = std::max<uint32_t>(n, 512);
        HIPCHK(c, dim3h::host_malloc((void**)&c->h_batch, sizeof(Scalars) * cap));
        c->h_batch_cap = cap;
    }
    // Launch of one chunk (entries [first, last) of order) on stream sx with
    // cus compute units: a HIP graph (scalar reset, expansion, scalar
    // gather, copy to the host).  Its kernels read every instance's
    // arguments from device tables, so the instantiated graph (one per
    // chunk slot) is replayed by later batches whose shape (kernel,
    // workgroup, LDS, grid, instance count) and table and host buffers are
    // the same; otherwise it is captured again.
    auto launch_chunk = [&](uint32_t first, uint32_t last, hipStream_t sx, int cus, int slot) -> int {
        const uint32_t m = last - first;
        std::vector<uint32_t> start(m + 1, 0);
        size_t lds = 0;
        uint64_t total = 0;
        for (uint32_t k = first; k < last; k++) {
            const auto& P = c->banks[order[k] + 1]->plan;
            start[k - first] = (uint32_t)total;
            if (!P.empty) {
                total += (uint64_t)P.nx * P.ntiles;
                lds = std::max(lds, dim3h::expand_lds_bytes(P.tile_words, P.touched_cap, P.stage_cap, (uint32_t)block, fp));
            }
            if (total >= (1ull << 32)) return fail(c, DIM3_ERR_INVALID, "batch holds more than 2^32 row tasks");
        }
        start[m] = (uint32_t)total;
        dim3h::ExpandArgs* da = d_args + first;
        uint32_t* ds = d_start + first + 2u * (uint32_t)slot;  // m + 1 starts, then the counter
        Scalars** dl = d_scl + first;
        Scalars* dout = d_scout + first;
        Scalars* hout = c->h_batch + first;
        HIPCHK(c, hipMemcpyAsync(da, args.data() + first, sizeof(dim3h::ExpandArgs) * m, hipMemcpyHostToDevice, sx));
        HIPCHK(c, hipMemcpyAsync(ds, start.data(), sizeof(uint32_t) * (m + 1), hipMemcpyHostToDevice, sx));
        HIPCHK(c, hipMemcpyAsync(dl, scl.data() + first, sizeof(Scalars*) * m, hipMemcpyHostToDevice, sx));
        HIPCHK(c, hipStreamSynchronize(sx));  // host vectors are pageable
        if (lds == 0) lds = dim3h::expand_lds_bytes(1, 64, 3u * block + 1u, (uint32_t)block, fp);
        unsigned grid = 1;
        int e = dim3h::expand_batch_setup(kcount, block, lds, cus, std::max<uint64_t>(total, 1), &grid);
        if (e != 0) return fail(c, DIM3_ERR_DEVICE, "batch setup: %s", hipGetErrorString((hipError_t)e));
        const uint64_t key[12] = {(uint64_t)kcount, (uint64_t)fp, (uint64_t)block, (uint64_t)lds, (uint64_t)grid,
                                  (uint64_t)m, (uint64_t)(uintptr_t)da, (uint64_t)(uintptr_t)ds,
                                  (uint64_t)(uintptr_t)dl, (uint64_t)(uintptr_t)dout, (uint64_t)(uintptr_t)hout,
                                  (uint64_t)(uintptr_t)sx};
        hipGraphExec_t& ex = c->batch_exec[slot];
        if (ex && memcmp(key, c->batch_key[slot], sizeof(key)) != 0) {
            hipGraphExecDestroy(ex);
            ex = nullptr;
        }
        if (!ex) {
            hipGraph_t graph = nullptr;
            {
                // no allocation anywhere in the process while the capture is open
                std::unique_lock<std::shared_mutex> capture_lock(dim3h::capture_mutex());
                HIPCHK(c, hipStreamBeginCapture(sx, hipStreamCaptureModeThreadLocal));
                dim3h::expand_batch_launch(kcount, fp, block, lds, grid, da, ds, m, ds + m + 1, dl, dout, sx);
                hipMemcpyAsync(hout, dout, sizeof(Scalars) * m, hipMemcpyDeviceToHost, sx);
                HIPCHK(c, hipStreamEndCapture(sx, &graph));
            }
            const hipError_t ie = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                ex = nullptr;
                return fail(c, DIM3_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(ie));
            }
            memcpy(c->batch_key[slot], key, sizeof(key));
        }
        const hipError_t ge = hipGraphLaunch(ex, sx);
        if (ge != hipSuccess) {
            hipGraphExecDestroy(ex);
            ex = nullptr;
            return fail(c, DIM3_ERR_DEVICE, "batch launch: %s", hipGetErrorString(ge));
        }
        return DIM3_OK;
    };

    float ms_expand = 0;
    if (na == n) {
        TRY(plan_chunk(0, n));
        HIPCHK(c, hipEventRecord(c->ev[4], st));
        TRY(launch_chunk(0, n, st, c->grid_cus(), 0));
        HIPCHK(c, hipEventRecord(c->ev[5], st));
        HIPCHK(c, hipStreamSynchronize(st));
        hipEventElapsedTime(&ms_expand, c->ev[4], c->ev[5]);
    } else {
        hipStream_t sp;
        TRY(pipe_stream(c, &sp));
        TRY(plan_chunk(0, na));
        HIPCHK(c, hipEventRecord(c->ev[6], st));  // chunk A's CSRs are built on the call's stream
        HIPCHK(c, hipStreamWaitEvent(sp, c->ev[6], 0));
        HIPCHK(c, hipEventRecord(c->ev[4], sp));
        TRY(launch_chunk(0, na, sp, std::max(1, (int)(c->pipe_cus * c->cu_share)), 0));
        HIPCHK(c, hipEventRecord(c->ev[7], sp));
        int rc = plan_chunk(na, n);  // beside chunk A's expansion
        if (rc == DIM3_OK) rc = launch_chunk(na, n, st, c->grid_cus(), 1);
        if (rc == DIM3_OK) {
            hipError_t e = hipEventRecord(c->ev[5], st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) rc = fail(c, DIM3_ERR_DEVICE, "batch chunk B: %s", hipGetErrorString(e));
        }
        const hipError_t ea = hipStreamSynchronize(sp);  // chunk A always finishes before we return
        TRY(rc);
        if (ea != hipSuccess) return fail(c, DIM3_ERR_DEVICE, "batch chunk A: %s", hipGetErrorString(ea));
        float ma = 0, mb = 0;
        hipEventElapsedTime(&ma, c->ev[4], c->ev[7]);
        hipEventElapsedTime(&mb, c->ev[4], c->ev[5]);
        ms_expand = std::max(ma, mb);
    }

    int rc = DIM3_OK;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t i = order[k];
        c->use_bank(i + 1);
        Scalars s = c->h_batch[k];
        if (dense_overflowed(c, s)) {  // a multiplicity above 127: this instance again, all-sparse
            dim3_opts oi = ob;
            oi.dense_mode = DIM3_DENSE_OFF;
            TRY(prepare(c, count_mode, (const int2*)inst[i].R_xy, inst[i].n_r, (const int2*)inst[i].S_zy,
                        inst[i].n_s, oi, st, wgs));
            TRY(expand(c, fp, inst[i].out, inst[i].cap, st));
            s = *c->h_sc;
        }
        fill_result(c, &results[i], s, false);
        results[i].stats.ms_expand = ms_expand;
        if (!fp && results[i].n > inst[i].cap) rc = DIM3_ERR_CAPACITY;
    }
    if (rc == DIM3_ERR_CAPACITY) fail(c, rc, "at least one instance needs more rows than its buffer holds");
    return rc;
}

void model_to_tuning(const DenseModel& m, uint32_t source, dim3_tuning* t) {
    memset(t, 0, sizeof(*t));
    t->path_ns = m.path_s * 1e9;
    t->seed_bit_ns = m.seed_s * 1e9;
    t->seed_word_ns = m.seed_word_s * 1e9;
    t->mfma_tops = m.mfma_ops / 1e12;
    t->bitor_gbs = m.l2_bps / 1e9;
    t->hbm_gbs = m.hbm_bps / 1e9;
    t->source = source;
}

void tuning_to_model(const dim3_tuning& t, DenseModel* m) {
    m->path_s = t.path_ns * 1e-9;
    m->seed_s = t.seed_bit_ns * 1e-9;
    m->seed_word_s = t.seed_word_ns * 1e-9;
    m->mfma_ops = t.mfma_tops * 1e12;
    m->l2_bps = t.bitor_gbs * 1e9;
    m->hbm_bps = t.hbm_gbs * 1e9;
}

float elapsed(dim3_ctx* c, int a, int b) {
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev[a], c->ev[b]);
    return ms;
}

// One fingerprint query of R = X x Y (every row reaches every key) against
// S = Y keys of D scattered z values each (affine maps of 0 .. D - 1 into
// [0, 2^kTuneZBits), distinct per key): each row's output is the union of
// its keys' z sets, spread over the z range as the z IDs of real data are
// (consecutive z IDs would make the LDS updates of a path unrealistically
// cheap).  Returns the plan's timings and the output rows.
constexpr uint32_t kTuneZBits = 20;

int tune_query(dim3_ctx* c, uint32_t X, uint32_t Y, uint32_t D, int dense, float* ms_dense, float* ms_expand,
               QueryPlan* plan, uint64_t* rows) {
    hipStream_t st = c->stream;
    c->qstream = st;
    int2 *R, *S;
    TRY(ws(c, WS_IN_R, (uint64_t)X * Y, &R));
    TRY(ws(c, WS_IN_S, (uint64_t)D * Y, &S));
    dim3h::grid_relation(R, X, Y, st);
    dim3h::scattered_relation(S, D, Y, kTuneZBits, st);
    dim3_opts o;
    dim3_opts_default(&o);
    o.output_mode = DIM3_OUT_FINGERPRINT;
    o.flags = DIM3_OPT_NO_SWAP;
    o.dense_mode = dense;
    TRY(prepare(c, false, R, (uint64_t)X * Y, S, (uint64_t)D * Y, o, st));
    TRY(expand(c, true, nullptr, 0, st));
    *ms_dense = elapsed(c, 3, 6);
    *ms_expand = elapsed(c, 4, 5);
    *plan = c->bank().plan;
    *rows = c->h_sc->out_n;
    return DIM3_OK;
}

int autotune(dim3_ctx* c, const char* cache_path, dim3_tuning* out) {
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
    std::string dev = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
    for (auto& ch : dev)
        if (ch == ' ') ch = '_';
    if (cache_path) {
        FILE* f = fopen(cache_path, "r");
        if (f) {
            char name[256] = {0};
            dim3_tuning t;
            memset(&t, 0, sizeof(t));
            const int n = fscanf(f, "dim3-tuning-v2 %255s %lf %lf %lf %lf %lf %lf", name, &t.path_ns, &t.seed_bit_ns,
                                 &t.seed_word_ns, &t.mfma_tops, &t.bitor_gbs, &t.hbm_gbs);
            fclose(f);
            if (n == 7 && dev == name && t.path_ns > 0 && t.mfma_tops > 0 && t.hbm_gbs > 0) {
                tuning_to_model(t, &c->model);
                c->model_source = 2;
                if (out) model_to_tuning(c->model, 2, out);
                return DIM3_OK;
            }
        }
    }
    hipStream_t st = c->stream;
    c->qstream = st;
    DenseModel m = c->model;
    // streaming writes: 256 MiB fill
    {
        uint32_t* buf;
        const uint64_t n = 64ull << 20;
        TRY(ws(c, WS_OUT, n, &buf));
        dim3h::fill_u32(buf, n, 0, st);
        HIPCHK(c, hipEventRecord(c->ev[0], st));
        for (int r = 0; r < 4; r++) dim3h::fill_u32(buf, n, (uint32_t)r, st);
        HIPCHK(c, hipEventRecord(c->ev[1], st));
        HIPCHK(c, hipStreamSynchronize(st));
        m.hbm_bps = 4.0 * n * 4 / (elapsed(c, 0, 1) * 1e-3);
    }
    // int8 product on the MFMA cores: 8192 x 8192 x 4096
    {
        const uint32_t M = 8192, K = 4096;
        uint32_t *A, *B, *D;
        TRY(ws(c, WS_DA, (uint64_t)M * K / 4, &A));
        TRY(ws(c, WS_DB, (uint64_t)M * K / 4, &B));
        TRY(ws(c, WS_DBITS, (uint64_t)M * (M / 32), &D));
        dim3h::fill_u32(A, (uint64_t)M * K / 4, 0x01010101u, st);
        dim3h::fill_u32(B, (uint64_t)M * K / 4, 0x01000100u, st);
        dim3h::gemm_i8((const int8_t*)A, (const int8_t*)B, M, M, K, D, M / 32, nullptr, st);
        HIPCHK(c, hipEventRecord(c->ev[0], st));
        for (int r = 0; r < 3; r++) dim3h::gemm_i8((const int8_t*)A, (const int8_t*)B, M, M, K, D, M / 32, nullptr, st);
        HIPCHK(c, hipEventRecord(c->ev[1], st));
        HIPCHK(c, hipStreamSynchronize(st));
        m.mfma_ops = 3.0 * 2.0 * M * M * (double)K / (elapsed(c, 0, 1) * 1e-3);
    }
    // Sparse paths: 4096 rows, each reaching 32 keys of 32768 scattered z
    // values (4.3e9 paths, about 2 / 3 of the 2^20 z values per row), then
    // the same rows from one key whose z set has the same size: the emission
    // is the same, so the difference is the paths' cost.  Then the 32-key
    // query seeded from an MFMA product (word copies of the product rows).
    // The model's path and seeding rates are fitted on real workloads on
    // MI355X, whose rows mix key weights (partial pieces, per-task work):
    // these heavy-row micro-queries run about three times faster per path
    // than that, so they scale the fitted rates by this device's speed
    // relative to the same micro-queries on MI355X (kRef*) rather than
    // replacing them.  The bit-or rate keeps its fitted value.
    constexpr double kRefPathS = 0.297e-12, kRefSeedWordS = 0.0739e-9;
    auto scale = [](double measured, double ref) { return std::min(4.0, std::max(0.25, measured / ref)); };
    const DenseModel fitted;
    const uint32_t X = 4096, Y = 32, D = 32768;
    float d32, e32, d1, e1, dm, em;
    QueryPlan p32, p1, pm;
    uint64_t r32, r1, rm;
    TRY(tune_query(c, X, Y, D, DIM3_DENSE_OFF, &d32, &e32, &p32, &r32));
    const uint32_t D1 = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(r32 / X, 1), 1u << kTuneZBits);
    TRY(tune_query(c, X, 1, D1, DIM3_DENSE_OFF, &d1, &e1, &p1, &r1));
    const double dpaths = (double)p32.out_j - (double)p1.out_j;
    double path_s = 0;
    if (e32 > e1 && dpaths > 0) {
        path_s = (e32 - e1) * 1e-3 / dpaths;
        m.path_s = fitted.path_s * scale(path_s, kRefPathS);
        m.seed_s = fitted.seed_s * scale(path_s, kRefPathS);  // one z ID and one LDS op, as a path
    }
    // the emission of the 32-key query's rows (the 1-key query less its paths)
    const double emit_s = std::max(0.0, e1 * 1e-3 - (double)p1.out_j * path_s);
    TRY(tune_query(c, X, Y, D, DIM3_DENSE_MFMA, &dm, &em, &pm, &rm));
    const double words = (double)X * (double)(pm.dwords ? pm.dwords : (pm.nz + 31) / 32);
    if (pm.dense == DIM3_DENSE_MFMA && em * 1e-3 > emit_s && rm == r32)
        m.seed_word_s = fitted.seed_word_s * scale((em * 1e-3 - emit_s) / words, kRefSeedWordS);
    c->model = m;
    c->model_source = 1;
    if (out) model_to_tuning(m, 1, out);
    if (cache_path) {
        FILE* f = fopen(cache_path, "w");
        if (f) {
            dim3_tuning t;
            model_to_tuning(m, 1, &t);
            fprintf(f, "dim3-tuning-v2 %s %.6g %.6g %.6g %.6g %.6g %.6g\n", dev.c_str(), t.path_ns, t.seed_bit_ns,
                    t.seed_word_ns, t.mfma_tops, t.bitor_gbs, t.hbm_gbs);
            fclose(f);
        }
    }
    return DIM3_OK;
}

}  // namespace

extern "C" {

int dim3_autotune(dim3_ctx* c, const char* cache_path, dim3_tuning* out) {
    if (!c) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    HIPCHK(c, hipSetDevice(c->device));
    c->use_bank(0);
    return autotune(c, cache_path, out);
}

int dim3_get_tuning(const dim3_ctx* c, dim3_tuning* out) {
    if (!c || !out) return DIM3_ERR_INVALID;
    model_to_tuning(c->model, c->model_source, out);
    return DIM3_OK;
}

int dim3_set_tuning(dim3_ctx* c, const dim3_tuning* in) {
    if (!c || !in) return DIM3_ERR_INVALID;
    if (!(in->path_ns > 0 && in->seed_bit_ns > 0 && in->seed_word_ns > 0 && in->mfma_tops > 0 && in->bitor_gbs > 0 &&
          in->hbm_gbs > 0))
        return fail(c, DIM3_ERR_INVALID, "tuning rates must be positive");
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    tuning_to_model(*in, &c->model);
    c->model_source = in->source;
    return DIM3_OK;
}


const char* dim3_version(void) { return "dim3-mi355x 0.1 (gfx950)"; }

const char* dim3_status_string(int s) {
    switch (s) {
        case DIM3_OK: return "ok";
        case DIM3_ERR_INVALID: return "invalid argument";
        case DIM3_ERR_OOM: return "out of memory";
        case DIM3_ERR_CANCELLED: return "stopped by the row callback";
        case DIM3_ERR_DEVICE: return "device error";
        case DIM3_ERR_CAPACITY: return "output capacity too small";
        case DIM3_ERR_ADMISSION: return "request can never be admitted";
        case DIM3_ERR_BUSY: return "queue full";
        case DIM3_ERR_NOT_FOUND: return "not found";
        case DIM3_ERR_UNSUPPORTED: return "unsupported";
        default: return "unknown status";
    }
}

void dim3_opts_default(dim3_opts* o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->output_mode = DIM3_OUT_MATERIALIZE;
    o->dense_mode = DIM3_DENSE_AUTO;
}

int dim3_ctx_create(int device, uint64_t hbm_budget_bytes, dim3_ctx** out) {
    if (!out) return DIM3_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return DIM3_ERR_DEVICE;
    if (device < 0 || device >= ndev) return DIM3_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return DIM3_ERR_DEVICE;
    dim3_ctx* c = new (std::nothrow) dim3_ctx();
    if (!c) return DIM3_ERR_OOM;
    c->device = device;
    c->use_bank(0);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return DIM3_ERR_DEVICE;
    }
    c->cu_count = prop.multiProcessorCount;
    c->max_lds = prop.sharedMemPerBlock;
    if (c->max_lds > 160 * 1024) c->max_lds = 160 * 1024;
    if (c->max_lds < 64 * 1024) c->max_lds = 64 * 1024;
    size_t fr = 0, tot = 0;
    hipMemGetInfo(&fr, &tot);
    c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(fr * 0.9);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        dim3h::host_malloc((void**)&c->h_sc, sizeof(Scalars)) != hipSuccess ||
        dim3h::host_malloc((void**)&c->h_hist, 6 * dim3h::kDenseBins * sizeof(unsigned long long)) !=
            hipSuccess) {
        delete c;
        return DIM3_ERR_DEVICE;
    }
    memset(c->h_sc, 0, sizeof(Scalars));
    for (auto& e : c->ev) hipEventCreate(&e);
    c->qstream = c->stream;
    // the allocation-free entry point's bank (its pinned mirrors made here)
    // and this device's scratch array (device is current)
    c->nab.reset(new (std::nothrow) Bank());
    if (!c->nab || dim3h::host_malloc((void**)&c->nab->hsc, sizeof(Scalars)) != hipSuccess ||
        dim3h::host_malloc((void**)&c->nab->hhist, 6 * dim3h::kDenseBins * sizeof(unsigned long long)) !=
            hipSuccess) {
        dim3_ctx_destroy(c);
        return DIM3_ERR_DEVICE;
    }
    c->scratch = scratch_address(device);
    c->err.reserve(600);  // error texts (< 512 bytes) never allocate
    *out = c;
    return DIM3_OK;
}

int dim3_ctx_destroy(dim3_ctx* c) {
    if (!c) return DIM3_ERR_INVALID;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto& b : c->banks)
        for (auto& s : b->ws)
            if (s.p) dim3h::dev_free(s.p);
    for (auto& ex : c->batch_exec)
        if (ex) hipGraphExecDestroy(ex);
    if (c->pipe_stream) {
        hipStreamSynchronize(c->pipe_stream);
        hipStreamDestroy(c->pipe_stream);
    }
    if (c->h_batch) dim3h::host_free(c->h_batch);
    if (c->h_stream) dim3h::host_free(c->h_stream);
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->h_sc) dim3h::host_free(c->h_sc);
    if (c->h_hist) dim3h::host_free(c->h_hist);
    hipStreamDestroy(c->stream);
    delete c;
    return DIM3_OK;
}

const char* dim3_last_error(const dim3_ctx* c) { return c ? c->err.c_str() : "null context"; }

uint64_t dim3_ctx_device_bytes(const dim3_ctx* c) { return c ? c->held : 0; }

int dim3_ctx_set_cu_share(dim3_ctx* c, double share) {
    if (!c || !(share > 0.0) || share > 1.0) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->cu_share = share;
    return DIM3_OK;
}

int dim3_ctx_release(dim3_ctx* c) {
    if (!c) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto& b : c->banks)
        for (auto& s : b->ws)
            if (s.p) {
                dim3h::dev_free_async(s.p, c->stream);
                c->held -= s.bytes;
                s.p = nullptr;
                s.bytes = 0;
            }
    c->banks.resize(1);
    c->banks[0]->plan = QueryPlan();
    c->cur = 0;
    hipStreamSynchronize(c->stream);
    return DIM3_OK;
}

int dim3_join_project(dim3_ctx* c, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy, uint64_t n_s,
                      const dim3_opts* opts, dim3_result* res) {
    return run_host(c, false, R_xy, n_r, S_zy, n_s, opts, res);
}

int dim3_join_aggregate_count(dim3_ctx* c, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy, uint64_t n_s,
                              const dim3_opts* opts, dim3_result* res) {
    return run_host(c, true, R_xy, n_r, S_zy, n_s, opts, res);
}

int dim3_join_project_stream(dim3_ctx* c, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy, uint64_t n_s,
                             const dim3_opts* opts, uint64_t chunk_rows, dim3_rows_cb cb, void* user,
                             dim3_result* res) {
    return run_stream(c, false, R_xy, n_r, S_zy, n_s, opts, chunk_rows, cb, user, res);
}

int dim3_join_aggregate_count_stream(dim3_ctx* c, const int32_t* R_xy, uint64_t n_r, const int32_t* S_zy,
                                     uint64_t n_s, const dim3_opts* opts, uint64_t chunk_rows, dim3_rows_cb cb,
                                     void* user, dim3_result* res) {
    return run_stream(c, true, R_xy, n_r, S_zy, n_s, opts, chunk_rows, cb, user, res);
}

int dim3_join_aggregate(dim3_ctx* c, const int32_t* R_xy, const int32_t* R_v, uint64_t n_r, const int32_t* S_zy,
                        const int32_t* S_v, uint64_t n_s, int agg, const dim3_opts* opts, dim3_result* res) {
    return run_value_host(c, agg, false, R_xy, R_v, n_r, S_zy, S_v, n_s, opts, res);
}

int dim3_join_aggregate_stream(dim3_ctx* c, const int32_t* R_xy, const int32_t* R_v, uint64_t n_r, const int32_t* S_zy,
                               const int32_t* S_v, uint64_t n_s, int agg, const dim3_opts* opts, uint64_t chunk_rows,
                               dim3_rows_cb cb, void* user, dim3_result* res) {
    return run_stream(c, true, R_xy, n_r, S_zy, n_s, opts, chunk_rows, cb, user, res, agg ? agg : -1, false, R_v,
                      S_v);
}

int dim3_join_aggregate_f32_stream(dim3_ctx* c, const int32_t* R_xy, const float* R_v, uint64_t n_r,
                                   const int32_t* S_zy, const float* S_v, uint64_t n_s, int agg,
                                   const dim3_opts* opts, uint64_t chunk_rows, dim3_rows_cb cb, void* user,
                                   dim3_result* res) {
    return run_stream(c, true, R_xy, n_r, S_zy, n_s, opts, chunk_rows, cb, user, res, agg ? agg : -1, true,
                      (const int32_t*)R_v, (const int32_t*)S_v);
}

int dim3_join_aggregate_tuples(dim3_ctx* c, const void* R_tuples, uint64_t n_r, const void* S_tuples, uint64_t n_s,
                               uint32_t stride, uint32_t value_offset, uint32_t value_type, int agg,
                               const dim3_opts* opts, dim3_result* res) {
    return run_tuples(c, agg, R_tuples, n_r, S_tuples, n_s, stride, value_offset, value_type, opts, res);
}

int dim3_join_aggregate_f32(dim3_ctx* c, const int32_t* R_xy, const float* R_v, uint64_t n_r, const int32_t* S_zy,
                            const float* S_v, uint64_t n_s, int agg, const dim3_opts* opts, dim3_result* res) {
    return run_value_host(c, agg, true, R_xy, (const int32_t*)R_v, n_r, S_zy, (const int32_t*)S_v, n_s, opts, res);
}

int dim3_join_aggregate_f32_dev(dim3_ctx* c, const int32_t* d_R, const float* d_Rv, uint64_t n_r, const int32_t* d_S,
                                const float* d_Sv, uint64_t n_s, int agg, const dim3_opts* opts,
                                dim3_value_row_f32* d_rows, uint64_t cap, void* stream, dim3_result* res) {
    return run_value_dev(c, agg, true, d_R, (const int32_t*)d_Rv, n_r, d_S, (const int32_t*)d_Sv, n_s, opts, d_rows,
                         cap, stream, res);
}

int dim3_join_aggregate_dev(dim3_ctx* c, const int32_t* d_R, const int32_t* d_Rv, uint64_t n_r, const int32_t* d_S,
                            const int32_t* d_Sv, uint64_t n_s, int agg, const dim3_opts* opts, dim3_count_row* d_rows,
                            uint64_t cap, void* stream, dim3_result* res) {
    return run_value_dev(c, agg, false, d_R, d_Rv, n_r, d_S, d_Sv, n_s, opts, d_rows, cap, stream, res);
}

int dim3_join_paths_grouped(dim3_ctx* c, const void* R_tuples, uint64_t n_r, const void* S_tuples, uint64_t n_s,
                            uint32_t stride, uint32_t value_offset, uint32_t value_size, const dim3_opts* opts,
                            dim3_grouped* g) {
    if (!c || !g) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    memset(g, 0, sizeof(*g));
    if ((n_r && !R_tuples) || (n_s && !S_tuples)) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    if ((value_size != 4 && value_size != 8) || stride % value_size || value_offset % value_size || value_offset < 8 ||
        value_offset + value_size > stride)
        return fail(c, DIM3_ERR_INVALID, "bad tuple layout (stride %u, value of %u bytes at %u)", stride, value_size,
                    value_offset);
    if (n_r > dim3s::kMaxTuples || n_s > dim3s::kMaxTuples)
        return fail(c, DIM3_ERR_INVALID, "relations take at most %llu tuples", (unsigned long long)dim3s::kMaxTuples);
    dim3_opts o;
    TRY(check_opts(c, opts, &o));
    HIPCHK(c, hipSetDevice(c->device));
    std::string e;
    const int rc = dim3h::grouped_paths(R_tuples, n_r, S_tuples, n_s, stride, value_offset, value_size,
                                        (o.flags & DIM3_OPT_EXCLUDE_EQUAL) != 0, c->stream, g, &e);
    if (rc != DIM3_OK) {
        dim3_grouped_free(g);
        return fail(c, rc, "%s", e.c_str());
    }
    return DIM3_OK;
}

void dim3_grouped_free(dim3_grouped* g) {
    if (!g) return;
    free(g->x);
    free(g->z);
    free(g->zkey);
    free(g->path_start);
    free(g->v1);
    free(g->v2);
    memset(g, 0, sizeof(*g));
}

void dim3_result_free(dim3_result* res) {
    if (!res) return;
    free(res->pairs);
    free(res->rows);
    free(res->rows_f32);
    res->pairs = nullptr;
    res->rows = nullptr;
    res->rows_f32 = nullptr;
    res->n = 0;
}

int dim3_join_project_dev(dim3_ctx* c, const int32_t* d_R, uint64_t n_r, const int32_t* d_S, uint64_t n_s,
                          const dim3_opts* opts, int32_t* d_pairs, uint64_t cap, void* stream, dim3_result* res) {
    return run_dev(c, false, d_R, n_r, d_S, n_s, opts, d_pairs, cap, stream, res);
}

int dim3_join_aggregate_count_dev(dim3_ctx* c, const int32_t* d_R, uint64_t n_r, const int32_t* d_S, uint64_t n_s,
                                  const dim3_opts* opts, dim3_count_row* d_rows, uint64_t cap, void* stream,
                                  dim3_result* res) {
    return run_dev(c, true, d_R, n_r, d_S, n_s, opts, d_rows, cap, stream, res);
}

// dim3_join_project_dev with every workspace buffer carved from the
// device's scratch array: no device or host allocation.  The all-sparse
// plan (the same rows); at most kNoAllocMaxTuples tuples per relation.
int dim3_join_project_dev_noalloc(dim3_ctx* c, const int32_t* d_R, uint64_t n_r, const int32_t* d_S, uint64_t n_s,
                                  const dim3_opts* opts, int32_t* d_pairs, uint64_t cap, void* stream,
                                  dim3_result* res) {
    if (!c || !res) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    memset(res, 0, sizeof(*res));
    if (!c->scratch) return fail(c, DIM3_ERR_DEVICE, "no scratch array on device %d", c->device);
    if (n_r > kNoAllocMaxTuples || n_s > kNoAllocMaxTuples)
        return fail(c, DIM3_ERR_INVALID, "the allocation-free join takes at most %llu tuples per relation (got %llu, %llu)",
                    (unsigned long long)kNoAllocMaxTuples, (unsigned long long)n_r, (unsigned long long)n_s);
    if ((n_r && !d_R) || (n_s && !d_S)) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    dim3_opts o;
    TRY(check_opts(c, opts, &o));
    o.dense_mode = DIM3_DENSE_OFF;
    const bool fp = o.output_mode == DIM3_OUT_FINGERPRINT;
    if (!fp && cap && !d_pairs) return fail(c, DIM3_ERR_INVALID, "null output buffer");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    c->qstream = st;
    struct Arena {  // the scratch array for this call only; every slot starts empty
        dim3_ctx* c;
        explicit Arena(dim3_ctx* cc) : c(cc) {
            for (auto& w : c->nab->ws) w = Slot();
            c->arena = c->scratch;
            c->arena_used = 0;
            c->arena_size = kScratchBytes;
        }
        ~Arena() { c->arena = nullptr; }
    } arena(c);
    TRY(prepare(c, false, (const int2*)d_R, n_r, (const int2*)d_S, n_s, o, st));
    TRY(expand(c, fp, d_pairs, cap, st));
    fill_result(c, res, *c->h_sc, true);
    if (!fp && res->n > cap)
        return fail(c, DIM3_ERR_CAPACITY, "output needs %llu rows, buffer holds %llu", (unsigned long long)res->n,
                    (unsigned long long)cap);
    return DIM3_OK;
}

int dim3_batch_join_project_dev(dim3_ctx* c, const dim3_instance* inst, uint32_t n, const dim3_opts* opts,
                                void* stream, dim3_result* results) {
    return run_batch(c, false, inst, n, opts, stream, results);
}

int dim3_batch_join_aggregate_count_dev(dim3_ctx* c, const dim3_instance* inst, uint32_t n, const dim3_opts* opts,
                                        void* stream, dim3_result* results) {
    return run_batch(c, true, inst, n, opts, stream, results);
}

int dim3_count_join_paths(dim3_ctx* c, const int32_t* R_xy, uint64_t nR, const int32_t* S_zy, uint64_t nS,
                          uint64_t* out_j) {
    if (!c || !out_j) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    *out_j = 0;
    if (nR == 0 || nS == 0) return DIM3_OK;
    if (!R_xy || !S_zy) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    c->qstream = st;
    int2 *dR, *dS;
    TRY(ws(c, WS_IN_R, nR, &dR));
    TRY(ws(c, WS_IN_S, nS, &dS));
    HIPCHK(c, hipMemcpyAsync(dR, R_xy, nR * 8, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(dS, S_zy, nS * 8, hipMemcpyHostToDevice, st));
    if (nR > dim3s::kMaxTuples || nS > dim3s::kMaxTuples)
        return fail(c, DIM3_ERR_INVALID, "relations are limited to 2^30-1 tuples");
    const uint32_t cap = (uint32_t)dim3s::key_table_cap(nR);
    uint32_t *keys, *degR, *degS, *slotR;
    int32_t* slotS;
    Scalars* sc;
    TRY(ws(c, WS_KEYS_Y, (uint64_t)cap + 1, &keys));
    TRY(ws(c, WS_DEGR_Y, (uint64_t)cap + 1, &degR));
    TRY(ws(c, WS_DEGS_Y, (uint64_t)cap + 1, &degS));
    TRY(ws(c, WS_SLOT_R, nR, &slotR));
    TRY(ws(c, WS_SLOT_S, nS, &slotS));
    TRY(ws(c, WS_SCALARS, 1, &sc));
    HIPCHK(c, hipMemsetAsync(sc, 0, sizeof(Scalars), st));
    dim3h::KeyMap m{keys, cap - 1, 0, cap + 1, 0, 0};
    dim3h::map_init(m, st);
    dim3h::fill_u32(degR, (uint64_t)cap + 1, 0, st);
    dim3h::fill_u32(degS, (uint64_t)cap + 1, 0, st);
    dim3h::build_y((const int2*)dR, nR, m, degR, slotR, st);
    dim3h::probe_y((const int2*)dS, nS, m, degR, degS, slotS, &sc->s_surv, st);
    dim3h::outj_only(degR, degS, cap + 1, &sc->out_j, st);
    LAUNCHCHK(c, "count_join_paths");
    HIPCHK(c, hipMemcpyAsync(c->h_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    *out_j = c->h_sc->out_j;
    return DIM3_OK;
}

// The estimate over device relations (|R| >= |S| > 0), on stream st.
static int estimate_on(dim3_ctx* c, const int2* dR, uint64_t nR, const int2* dS, uint64_t nS, hipStream_t st,
                       int64_t* estimate) {
    // sample bits: halve |S| until it is at most 10000 (dim3.hpp:262-269)
    uint32_t bits = 0;
    for (uint64_t n = nS; n > 10000; n /= 2) bits++;
    const uint32_t smask = bits >= 32 ? 0xFFFFFFFFu : (1u << bits) - 1u;
    // the distinct sampled keys: at most |S|, and at most the int32 values
    // whose low `bits` bits are zero (the key -1 has its own slot)
    const uint32_t cap = (uint32_t)dim3s::estimate_table_cap(nS, bits);
    uint32_t *keys, *cnt;
    Scalars* sc;
    TRY(ws(c, WS_KEYS_Y, (uint64_t)cap + 1, &keys));
    TRY(ws(c, WS_DEGS_Y, (uint64_t)cap + 1, &cnt));
    TRY(ws(c, WS_SCALARS, 1, &sc));
    HIPCHK(c, hipMemsetAsync(sc, 0, sizeof(Scalars), st));
    dim3h::KeyMap m{keys, cap - 1, 0, cap + 1, 0, 0};
    dim3h::map_init(m, st);
    dim3h::fill_u32(cnt, (uint64_t)cap + 1, 0, st);
    dim3h::estimate_sample((const int2*)dR, nR, (const int2*)dS, nS, smask, keys, cap - 1, cnt,
                           (unsigned long long*)&sc->spare[0], st);
    dim3h::estimate_prefix((const int2*)dR, (uint32_t)std::min<uint64_t>(nR, 1000), (const int2*)dS,
                           (uint32_t)std::min<uint64_t>(nS, 1000), (unsigned long long*)&sc->spare[1], st);
    LAUNCHCHK(c, "estimate");
    HIPCHK(c, hipMemcpyAsync(c->h_sc, sc, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const int64_t sampled = (int64_t)c->h_sc->spare[0];
    if (sampled != 0) {
        *estimate = sampled << bits;
    } else {  // dim3.hpp:215-221
        const double r2 = (double)(int64_t)c->h_sc->spare[1];
        *estimate = (int64_t)(r2 * ((double)nR / 1e3) * ((double)nS / 1e3));
    }
    return DIM3_OK;
}

// Argument checks and the swap rule; the reference asserts |R| >= |S|
// (dim3.hpp:194).  Returns -1 when there is nothing to sample (estimate 0).
static int estimate_args(dim3_ctx* c, const int32_t*& R_xy, uint64_t& nR, const int32_t*& S_zy, uint64_t& nS,
                         int64_t* estimate) {
    *estimate = 0;
    if (nR < nS) {
        std::swap(R_xy, S_zy);
        std::swap(nR, nS);
    }
    if (nS == 0) return -1;
    if (!R_xy || !S_zy) return fail(c, DIM3_ERR_INVALID, "null relation pointer");
    if (nR > dim3s::kMaxTuples) return fail(c, DIM3_ERR_INVALID, "relations are limited to 2^30-1 tuples");
    return DIM3_OK;
}

int dim3_estimate_join_cardinality(dim3_ctx* c, const int32_t* R_xy, uint64_t nR, const int32_t* S_zy, uint64_t nS,
                                   int64_t* estimate) {
    if (!c || !estimate) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    const int a = estimate_args(c, R_xy, nR, S_zy, nS, estimate);
    if (a != DIM3_OK) return a < 0 ? DIM3_OK : a;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    c->qstream = st;
    int2 *dR, *dS;
    TRY(ws(c, WS_IN_R, nR, &dR));
    TRY(ws(c, WS_IN_S, nS, &dS));
    HIPCHK(c, hipMemcpyAsync(dR, R_xy, nR * 8, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(dS, S_zy, nS * 8, hipMemcpyHostToDevice, st));
    return estimate_on(c, dR, nR, dS, nS, st, estimate);
}

int dim3_estimate_join_cardinality_dev(dim3_ctx* c, const int32_t* d_R, uint64_t nR, const int32_t* d_S, uint64_t nS,
                                       void* stream, int64_t* estimate) {
    if (!c || !estimate) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    const int a = estimate_args(c, d_R, nR, d_S, nS, estimate);
    if (a != DIM3_OK) return a < 0 ? DIM3_OK : a;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    c->qstream = st;
    return estimate_on(c, (const int2*)d_R, nR, (const int2*)d_S, nS, st, estimate);
}

int dim3_estimate_join_cardinality_tuples(dim3_ctx* c, const void* R_tuples, uint64_t nR, const void* S_tuples,
                                          uint64_t nS, uint32_t stride, int64_t* estimate) {
    if (!c || !estimate) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    if (stride < 8u || stride % 4u) return fail(c, DIM3_ERR_INVALID, "bad tuple stride %u", stride);
    const int32_t* R = (const int32_t*)R_tuples;
    const int32_t* S = (const int32_t*)S_tuples;
    const int a = estimate_args(c, R, nR, S, nS, estimate);
    if (a != DIM3_OK) return a < 0 ? DIM3_OK : a;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    c->qstream = st;
    int2 *dR, *dS;
    int32_t* dv;
    uint8_t* raw;
    Scalars* sc;
    TRY(ws(c, WS_IN_R, nR, &dR));
    TRY(ws(c, WS_IN_S, nS, &dS));
    TRY(ws(c, WS_IN_VR, nR, &dv));  // the split's value output, unused
    TRY(ws(c, WS_SCALARS, 1, &sc));
    TRY(ws(c, WS_OUT, nR * stride, &raw));  // staging (nR >= nS after the swap)
    // (first, key) of each tuple, split on the device as dim3_join_aggregate_tuples does
    HIPCHK(c, hipMemcpyAsync(raw, R, nR * stride, hipMemcpyHostToDevice, st));
    dim3h::split_tuples(raw, nR, stride, 0, DIM3_VALUE_I32, dR, dv, (uint32_t*)&sc->spare[2], st);
    HIPCHK(c, hipMemcpyAsync(raw, S, nS * stride, hipMemcpyHostToDevice, st));
    dim3h::split_tuples(raw, nS, stride, 0, DIM3_VALUE_I32, dS, dv, (uint32_t*)&sc->spare[2], st);
    LAUNCHCHK(c, "split tuples");
    return estimate_on(c, dR, nR, dS, nS, st, estimate);
}

int dim3_dense_product(dim3_ctx* c, const int8_t* A, uint64_t M, const int8_t* Bt, uint64_t N, uint64_t K,
                       uint32_t* bits, int32_t* counts) {
    if (!c) return DIM3_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->err.clear();
    if (M == 0 || N == 0) return DIM3_OK;
    if ((K && (!A || !Bt)) || (!bits && !counts)) return fail(c, DIM3_ERR_INVALID, "null argument");
    if (K > 131072 || M >= 65535ull * 128ull || N >= (1ull << 31) || M * N >= (1ull << 40))
        return fail(c, DIM3_ERR_INVALID, "dense product shape %llu x %llu x %llu out of range",
                    (unsigned long long)M, (unsigned long long)N, (unsigned long long)K);
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    c->qstream = st;
    const uint64_t Kp = std::max<uint64_t>((K + 63) / 64 * 64, 64);
    const uint64_t dwords = (N + 31) / 32;
    uint32_t *dA, *dB, *dbits = nullptr;
    int32_t* dcnt = nullptr;
    TRY(ws(c, WS_DA, M * Kp / 4, &dA));
    TRY(ws(c, WS_DB, N * Kp / 4, &dB));
    if (bits) TRY(ws(c, WS_DBITS, M * dwords, &dbits));
    if (counts) TRY(ws(c, WS_CINT, M * N, &dcnt));
    dim3h::fill_u32(dA, M * Kp / 4, 0, st);
    dim3h::fill_u32(dB, N * Kp / 4, 0, st);
    if (K) {
        HIPCHK(c, hipMemcpy2DAsync(dA, Kp, A, K, K, M, hipMemcpyHostToDevice, st));
        HIPCHK(c, hipMemcpy2DAsync(dB, Kp, Bt, K, K, N, hipMemcpyHostToDevice, st));
    }
    const int e = dim3h::gemm_i8((const int8_t*)dA, (const int8_t*)dB, (uint32_t)M, (uint32_t)N, (uint32_t)Kp, dbits,
                                 (uint32_t)dwords, dcnt, st);
    if (e != 0) return fail(c, DIM3_ERR_DEVICE, "dense product launch: %s", hipGetErrorString((hipError_t)e));
    if (bits) HIPCHK(c, hipMemcpyAsync(bits, dbits, M * dwords * 4, hipMemcpyDeviceToHost, st));
    if (counts) HIPCHK(c, hipMemcpyAsync(counts, dcnt, M * N * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return DIM3_OK;
}

}  // extern "C"
