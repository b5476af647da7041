// libdim3 serving loop: resident relations, admission by HBM footprint,
// concurrent workers (one context and stream each) and continuous batching.
//
// Request life: dim3_submit checks the relation handles and the footprint
// estimate (refused with DIM3_ERR_ADMISSION when it could never fit next to
// the resident relations), queues the request and wakes a worker.  A worker
// takes the oldest request whose footprint fits the memory not held by the
// other workers, plus (for fingerprint requests) every later waiting request
// of the same kind and options up to max_batch, runs them as one batch (or
// one query), records the result and wakes the waiters and the other
// workers.  Workspaces stay with their worker for reuse; a worker whose
// request does not fit asks the idle workers to release theirs.
// With 2+ workers, worker 0 serves the large requests (>= large_tuples) on a
// stream whose CU mask holds large_share_pct of the compute units; while one
// runs, the other workers launch on streams masked to the remaining CUs, so
// small requests never share a CU with (or queue behind) a large one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/dim3.h"
#include "dim3_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Rel {
    int32_t* d = nullptr;
    uint64_t n = 0;
    uint64_t bytes = 0;
    int32_t* dv = nullptr;  // tuple values (value aggregates), DIM3_VALUE_I32 or DIM3_VALUE_F32
    uint32_t vtype = 0;
    uint64_t vbytes = 0;
    uint32_t users = 0;
    bool dropped = false;
};

void free_rel(Rel& r) {
    dim3h::dev_free(r.d);
    if (r.dv) dim3h::dev_free(r.dv);
}

struct Job {
    dim3_request req{};
    uint64_t est = 0;      // footprint: workspace, plus out_est
    uint64_t out_est = 0;  // materialising requests: the estimated output bytes
    uint64_t tuples = 0;
    Clock::time_point t_submit, t_start, t_end;
    int status = DIM3_OK;
    dim3_result res{};
    bool started = false, done = false;
};

// Workspace estimate of one query: hash tables, slot and degree arrays, the
// two CSRs and the per-tuple scratch stay below 96 bytes per input tuple
// (measured 40-70 on the benchmark workloads), plus 16 MiB of fixed buffers.
uint64_t footprint(uint64_t nr, uint64_t ns) { return 96ull * (nr + ns) + (16ull << 20); }

bool same_batch_class(const dim3_request& a, const dim3_request& b) {
    // value aggregates run one request at a time
    return a.kind == b.kind && a.kind != DIM3_REQ_VALUE && a.opts.output_mode == DIM3_OUT_FINGERPRINT &&
           b.opts.output_mode == DIM3_OUT_FINGERPRINT && a.opts.flags == b.opts.flags &&
           a.opts.dense_mode == b.opts.dense_mode && a.opts.heavy_k == b.opts.heavy_k &&
           a.opts.tile_bits == b.opts.tile_bits && a.opts.block == b.opts.block;
}

}  // namespace

struct dim3_server {
    int device = 0;
    uint64_t budget = 0;
    uint32_t nworkers = 4, max_batch = 32, queue_cap = 4096;
    uint64_t large_tuples = 20000000;  // lane 0 (with 2+ workers) serves these alone
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Job> jobs;
    std::unordered_map<uint64_t, Rel> rels;
    uint64_t next_ticket = 1, next_rel = 1;
    uint64_t resident = 0;
    std::vector<uint64_t> held;   // per worker: workspace bytes (or the running batch's estimate, if larger)
    std::vector<char> busy;       // per worker: running a request
    std::vector<char> release;    // per worker: asked to give its idle workspace back
    std::vector<dim3_ctx*> ctxs;
    std::vector<std::thread> threads;
    // submit-time join size estimates (materialising requests), one at a time
    std::mutex est_mu;
    dim3_ctx* est_ctx = nullptr;
    // Compute-unit partition (2+ workers): lane 0's large requests run on a
    // stream masked to the first large_cus CUs; while one runs, the other
    // workers launch on streams masked to the remaining CUs, so neither kind
    // shares a CU with the other.  Otherwise every worker uses all CUs.
    hipStream_t io_stream = nullptr;       // relation uploads (client threads)
    std::vector<hipStream_t> full_stream;  // per worker, all CUs
    std::vector<hipStream_t> part_stream;  // per worker: lane 0 the large part, others the rest
    uint32_t cus = 0, large_cus = 0;
    uint32_t large_running = 0;
    bool stop = false;
    dim3_server_stats st{};

    uint64_t others_held(uint32_t w) const {
        uint64_t s = 0;
        for (uint32_t v = 0; v < held.size(); v++)
            if (v != w) s += held[v];
        return s;
    }
    // a request's hold on its relations ends; a dropped relation is freed by its last user (mu held)
    void unpin(uint64_t hdl) {
        Rel& r = rels.at(hdl);
        if (--r.users == 0 && r.dropped) {
            hipSetDevice(device);
            free_rel(r);
            resident -= r.bytes + r.vbytes;
            rels.erase(hdl);
        }
    }
    int estimate_paths(const int32_t* dR, uint64_t nR, const int32_t* dS, uint64_t nS, int64_t* out);
    void note_peak() {
        uint64_t s = resident;
        for (auto h : held) s += h;
        if (s > st.peak_bytes) st.peak_bytes = s;
    }
    void worker(uint32_t w);
    void run(uint32_t w, std::vector<uint64_t>& batch, std::vector<Job*>& js, void** out_buf, uint64_t* out_bytes);
};

void dim3_server::run(uint32_t w, std::vector<uint64_t>& batch, std::vector<Job*>& js, void** out_buf,
                      uint64_t* out_bytes) {
    dim3_ctx* c = ctxs[w];
    const bool count = js[0]->req.kind == DIM3_REQ_COUNT;
    // the stream of this batch: the CU partition when it applies
    const bool large = nworkers > 1 && w == 0 && js[0]->tuples >= large_tuples;
    bool part = false;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (large) large_running++;
        part = large || (nworkers > 1 && w != 0 && large_running > 0);
    }
    hipStream_t st = part ? part_stream[w] : full_stream[w];
    dim3_ctx_set_cu_share(c, !part ? 1.0 : (double)(large ? large_cus : cus - large_cus) / (double)cus);
    struct Done {  // the large lane is free again when this batch ends
        dim3_server* s;
        bool large;
        ~Done() {
            if (!large) return;
            std::lock_guard<std::mutex> lk(s->mu);
            s->large_running--;
        }
    } done{this, large};
    std::vector<const Rel*> R(js.size()), S(js.size());
    {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < js.size(); i++) {
            R[i] = &rels.at(js[i]->req.r);
            S[i] = &rels.at(js[i]->req.s);
        }
    }
    if (js.size() > 1) {
        std::vector<dim3_instance> inst(js.size());
        for (size_t i = 0; i < js.size(); i++)
            inst[i] = dim3_instance{R[i]->d, R[i]->n, S[i]->d, S[i]->n, nullptr, 0};
        std::vector<dim3_result> res(js.size());
        const int rc = count ? dim3_batch_join_aggregate_count_dev(c, inst.data(), (uint32_t)js.size(),
                                                                  &js[0]->req.opts, st, res.data())
                             : dim3_batch_join_project_dev(c, inst.data(), (uint32_t)js.size(), &js[0]->req.opts,
                                                           st, res.data());
        for (size_t i = 0; i < js.size(); i++) {
            js[i]->status = rc;
            js[i]->res = res[i];
        }
        return;
    }
    Job* j = js[0];
    const dim3_opts& o = j->req.opts;
    const bool fp = o.output_mode == DIM3_OUT_FINGERPRINT;
    const bool value = j->req.kind == DIM3_REQ_VALUE;
    const bool f32 = value && R[0]->vtype == DIM3_VALUE_F32;
    const size_t row = f32 ? sizeof(dim3_value_row_f32)
                       : (count || value) ? sizeof(dim3_count_row) : 2 * sizeof(int32_t);
    if (!fp && j->out_est > *out_bytes) {  // the estimated output (admitted with the request)
        if (*out_buf) dim3h::dev_free(*out_buf);
        *out_buf = nullptr;
        *out_bytes = 0;
        if (dim3h::dev_malloc(out_buf, j->out_est) == hipSuccess) {
            *out_bytes = j->out_est;
        } else {
            (void)hipGetLastError();
            *out_buf = nullptr;
        }
    }
    uint64_t cap = fp ? 0 : *out_bytes / row;
    dim3_result res;
    for (int attempt = 0; attempt < 2; attempt++) {
        const int agg = (int)j->req.agg;
        int rc = f32     ? dim3_join_aggregate_f32_dev(c, R[0]->d, (const float*)R[0]->dv, R[0]->n, S[0]->d,
                                                       (const float*)S[0]->dv, S[0]->n, agg, &o,
                                                       (dim3_value_row_f32*)*out_buf, cap, st, &res)
                 : value ? dim3_join_aggregate_dev(c, R[0]->d, R[0]->dv, R[0]->n, S[0]->d, S[0]->dv, S[0]->n, agg, &o,
                                                   (dim3_count_row*)*out_buf, cap, st, &res)
                 : count ? dim3_join_aggregate_count_dev(c, R[0]->d, R[0]->n, S[0]->d, S[0]->n, &o,
                                                         (dim3_count_row*)*out_buf, cap, st, &res)
                         : dim3_join_project_dev(c, R[0]->d, R[0]->n, S[0]->d, S[0]->n, &o, (int32_t*)*out_buf, cap,
                                                 st, &res);
        if (rc == DIM3_ERR_CAPACITY && attempt == 0) {  // grow the output buffer to the size reported
            if (*out_buf) dim3h::dev_free(*out_buf);
            *out_buf = nullptr;
            *out_bytes = 0;
            uint64_t room = 0;
            {
                std::lock_guard<std::mutex> lk(mu);
                const uint64_t used = resident + others_held(w) + dim3_ctx_device_bytes(c);
                room = budget > used ? budget - used : 0;
            }
            // an output larger than the HBM budget leaves fails the request
            if (res.n * row > room || dim3h::dev_malloc(out_buf, res.n * row) != hipSuccess) {
                (void)hipGetLastError();
                *out_buf = nullptr;
                j->status = DIM3_ERR_OOM;
                return;
            }
            *out_bytes = res.n * row;
            cap = res.n;
            continue;
        }
        j->status = rc;
        break;
    }
    j->res = res;
    if (j->status == DIM3_OK && !fp && res.n) {
        void* h = malloc(res.n * row);
        if (!h || dim3h::copy_wait(h, *out_buf, res.n * row, hipMemcpyDeviceToHost, st) != hipSuccess) {
            free(h);
            j->status = DIM3_ERR_OOM;
            j->res.n = 0;
            return;
        }
        if (f32) j->res.rows_f32 = (dim3_value_row_f32*)h;
        else if (count || value) j->res.rows = (dim3_count_row*)h;
        else j->res.pairs = (int32_t*)h;
    }
}

void dim3_server::worker(uint32_t w) {
    hipSetDevice(device);
    void* out_buf = nullptr;
    uint64_t out_bytes = 0;
    for (;;) {
        std::vector<uint64_t> batch;
        std::vector<Job*> js;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (release[w]) {  // another worker needs the memory this idle context holds
                    release[w] = 0;
                    lk.unlock();
                    dim3_ctx_release(ctxs[w]);
                    if (out_buf) dim3h::dev_free(out_buf);
                    out_buf = nullptr;
                    out_bytes = 0;
                    lk.lock();
                    held[w] = 0;
                    cv_work.notify_all();
                    continue;
                }
                if (stop && pending.empty()) break;
                // requests that relations put since their admission have made impossible
                for (auto jt = pending.begin(); jt != pending.end();) {
                    Job& j = jobs.at(*jt);
                    if (resident + j.est > budget) {
                        j.status = DIM3_ERR_ADMISSION;
                        j.done = true;
                        j.t_start = j.t_end = Clock::now();
                        for (uint64_t hdl : {j.req.r, j.req.s}) unpin(hdl);
                        st.rejected++;
                        jt = pending.erase(jt);
                        cv_done.notify_all();
                    } else {
                        ++jt;
                    }
                }
                // the oldest request that fits next to the other workers' memory
                const uint64_t room = budget > resident + others_held(w) ? budget - resident - others_held(w) : 0;
                // with 2+ workers, worker 0 takes large requests first and the others
                // only small ones, so small requests never wait behind a large one
                const bool lane0 = nworkers > 1 && w == 0;
                auto it = pending.end();
                for (int pass = 0; pass < 2 && it == pending.end(); pass++) {
                    for (auto jt = pending.begin(); jt != pending.end(); ++jt) {
                        const Job& j = jobs.at(*jt);
                        const bool large = j.tuples >= large_tuples;
                        if (nworkers > 1 && !lane0 && large) continue;
                        if (lane0 && pass == 0 && !large) continue;
                        if (j.est <= room) {
                            it = jt;
                            break;
                        }
                    }
                    if (!lane0) break;
                }
                if (it == pending.end() && !pending.empty()) {
                    // nothing fits: idle workers give back their workspace
                    uint64_t idle = 0;
                    for (uint32_t v = 0; v < held.size(); v++)
                        if (v != w && !busy[v] && held[v]) {
                            release[v] = 1;
                            idle += held[v];
                        }
                    if (idle) cv_work.notify_all();
                }
                if (it != pending.end()) {
                    Job& j0 = jobs.at(*it);
                    uint64_t used = j0.est;
                    batch.push_back(*it);
                    js.push_back(&j0);
                    it = pending.erase(it);
                    // continuous batching: later waiting requests of the same class join
                    while (it != pending.end() && batch.size() < max_batch && max_batch > 1 &&
                           j0.req.opts.output_mode == DIM3_OUT_FINGERPRINT) {
                        Job& j = jobs.at(*it);
                        if (same_batch_class(j0.req, j.req) && used + j.est <= room) {
                            used += j.est;
                            batch.push_back(*it);
                            js.push_back(&j);
                            it = pending.erase(it);
                        } else {
                            ++it;
                        }
                    }
                    held[w] = std::max<uint64_t>(held[w], used);
                    note_peak();
                    break;
                }
                cv_work.wait(lk);
            }
            if (batch.empty()) break;  // stopping and drained
            busy[w] = 1;
            const auto now = Clock::now();
            for (Job* j : js) {
                j->t_start = now;
                j->started = true;
            }
            if (js.size() > 1) {
                st.batches++;
                st.batched_requests += js.size();
            }
        }
        run(w, batch, js, &out_buf, &out_bytes);
        uint64_t h = dim3_ctx_device_bytes(ctxs[w]) + out_bytes;
        if (h > budget / 2) {  // give very large workspaces back (others are released on demand)
            dim3_ctx_release(ctxs[w]);
            if (out_buf) dim3h::dev_free(out_buf);
            out_buf = nullptr;
            out_bytes = 0;
            h = 0;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            const auto now = Clock::now();
            for (Job* j : js) {
                j->t_end = now;
                j->done = true;
                st.completed++;
                for (uint64_t hdl : {j->req.r, j->req.s}) unpin(hdl);
            }
            held[w] = h;
            busy[w] = 0;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
    if (out_buf) dim3h::dev_free(out_buf);
}

extern "C" {

void dim3_server_opts_default(dim3_server_opts* o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->workers = 4;
    o->max_batch = 32;
    o->queue_cap = 4096;
}

int dim3_server_create(const dim3_server_opts* opts, dim3_server** out) {
    if (!out) return DIM3_ERR_INVALID;
    *out = nullptr;
    dim3_server_opts o;
    dim3_server_opts_default(&o);
    if (opts) o = *opts;
    if (o.workers == 0) o.workers = 4;
    if (o.max_batch == 0) o.max_batch = 32;
    if (o.queue_cap == 0) o.queue_cap = 4096;
    if (o.workers > 16) return DIM3_ERR_INVALID;
    if (hipSetDevice(o.device) != hipSuccess) return DIM3_ERR_DEVICE;
    size_t fr = 0, tot = 0;
    hipMemGetInfo(&fr, &tot);
    dim3_server* s = new (std::nothrow) dim3_server();
    if (!s) return DIM3_ERR_OOM;
    s->device = o.device;
    s->budget = o.hbm_budget_bytes ? o.hbm_budget_bytes : (uint64_t)(fr * 0.9);
    s->nworkers = o.workers;
    s->max_batch = o.max_batch;
    s->queue_cap = o.queue_cap;
    s->large_tuples = o.large_tuples ? o.large_tuples : 20000000ull;
    const double share = (o.large_share_pct ? std::min<uint32_t>(o.large_share_pct, 100) : 90) / 100.0;
    s->held.assign(o.workers, 0);
    s->busy.assign(o.workers, 0);
    s->release.assign(o.workers, 0);
    for (uint32_t w = 0; w < o.workers; w++) {
        dim3_ctx* c = nullptr;
        const int rc = dim3_ctx_create(o.device, s->budget, &c);
        if (rc != DIM3_OK) {
            for (auto* x : s->ctxs) dim3_ctx_destroy(x);
            delete s;
            return rc;
        }
        s->ctxs.push_back(c);
    }
    // streams: all CUs, and the two sides of the CU partition (masks of
    // consecutive CU indices; the first large_cus go to the large lane)
    {
        hipDeviceProp_t prop;
        hipGetDeviceProperties(&prop, o.device);
        s->cus = (uint32_t)std::max(2, prop.multiProcessorCount);
        s->large_cus = std::min<uint32_t>(s->cus - 1, std::max<uint32_t>(1, (uint32_t)(share * s->cus + 0.5)));
        std::vector<uint32_t> big((s->cus + 31) / 32, 0), rest((s->cus + 31) / 32, 0);
        for (uint32_t i = 0; i < s->cus; i++) (i < s->large_cus ? big : rest)[i / 32] |= 1u << (i % 32);
        bool ok = true;
        for (uint32_t w = 0; w < o.workers; w++) {
            hipStream_t f = nullptr, p = nullptr;
            ok = ok && hipStreamCreateWithFlags(&f, hipStreamNonBlocking) == hipSuccess;
            const std::vector<uint32_t>& m = w == 0 ? big : rest;
            ok = ok && hipExtStreamCreateWithCUMask(&p, (uint32_t)m.size(), m.data()) == hipSuccess;
            s->full_stream.push_back(f);
            s->part_stream.push_back(p);
        }
        ok = ok && hipStreamCreateWithFlags(&s->io_stream, hipStreamNonBlocking) == hipSuccess;
        if (!ok) {
            for (auto x : s->full_stream)
                if (x) hipStreamDestroy(x);
            for (auto x : s->part_stream)
                if (x) hipStreamDestroy(x);
            if (s->io_stream) hipStreamDestroy(s->io_stream);
            for (auto* x : s->ctxs) dim3_ctx_destroy(x);
            delete s;
            return DIM3_ERR_DEVICE;
        }
    }
    for (uint32_t w = 0; w < o.workers; w++) s->threads.emplace_back(&dim3_server::worker, s, w);
    *out = s;
    return DIM3_OK;
}

int dim3_server_destroy(dim3_server* s) {
    if (!s) return DIM3_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->threads) t.join();
    hipSetDevice(s->device);
    for (auto& kv : s->rels) free_rel(kv.second);
    for (auto& kv : s->jobs) dim3_result_free(&kv.second.res);
    for (auto* c : s->ctxs) dim3_ctx_destroy(c);
    if (s->est_ctx) dim3_ctx_destroy(s->est_ctx);
    for (auto x : s->full_stream) hipStreamDestroy(x);
    if (s->io_stream) hipStreamDestroy(s->io_stream);
    for (auto x : s->part_stream) hipStreamDestroy(x);
    delete s;
    return DIM3_OK;
}

static int put_relation(dim3_server* s, const int32_t* pairs, uint64_t n, uint64_t* handle, hipMemcpyKind kind) {
    if (!s || !handle || (n && !pairs)) return DIM3_ERR_INVALID;
    *handle = 0;
    const uint64_t bytes = std::max<uint64_t>(n, 1) * 8;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->resident + bytes > s->budget) return DIM3_ERR_ADMISSION;
        s->resident += bytes;
    }
    hipSetDevice(s->device);
    int32_t* d = nullptr;
    if (dim3h::dev_malloc((void**)&d, bytes) != hipSuccess) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lk(s->mu);
        s->resident -= bytes;
        return DIM3_ERR_OOM;
    }
    if (n && dim3h::copy_wait(d, pairs, n * 8, kind, s->io_stream) != hipSuccess) {
        dim3h::dev_free(d);
        std::lock_guard<std::mutex> lk(s->mu);
        s->resident -= bytes;
        return DIM3_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t h = s->next_rel++;
    Rel& r = s->rels[h];
    r.d = d;
    r.n = n;
    r.bytes = bytes;
    *handle = h;
    s->note_peak();
    return DIM3_OK;
}

int dim3_server_put_relation(dim3_server* s, const int32_t* pairs, uint64_t n, uint64_t* handle) {
    return put_relation(s, pairs, n, handle, hipMemcpyHostToDevice);
}

int dim3_server_put_relation_dev(dim3_server* s, const int32_t* d_pairs, uint64_t n, uint64_t* handle) {
    return put_relation(s, d_pairs, n, handle, hipMemcpyDeviceToDevice);
}

int dim3_server_put_values(dim3_server* s, uint64_t handle, const void* values, uint32_t value_type) {
    if (!s || (value_type != DIM3_VALUE_I32 && value_type != DIM3_VALUE_F32)) return DIM3_ERR_INVALID;
    uint64_t n, bytes;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->rels.find(handle);
        if (it == s->rels.end() || it->second.dropped) return DIM3_ERR_NOT_FOUND;
        if (it->second.dv || it->second.users) return DIM3_ERR_INVALID;  // once, before any request uses it
        n = it->second.n;
        bytes = std::max<uint64_t>(n, 1) * 4;
        if (n && !values) return DIM3_ERR_INVALID;
        if (s->resident + bytes > s->budget) return DIM3_ERR_ADMISSION;
        s->resident += bytes;
    }
    hipSetDevice(s->device);
    int32_t* d = nullptr;
    int rc = DIM3_OK;
    if (dim3h::dev_malloc((void**)&d, bytes) != hipSuccess) {
        (void)hipGetLastError();
        rc = DIM3_ERR_OOM;
    } else if (n && dim3h::copy_wait(d, values, n * 4, hipMemcpyDefault, s->io_stream) != hipSuccess) {
        dim3h::dev_free(d);
        rc = DIM3_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc != DIM3_OK) {
        s->resident -= bytes;
        return rc;
    }
    Rel& r = s->rels.at(handle);
    r.dv = d;
    r.vtype = value_type;
    r.vbytes = bytes;
    s->note_peak();
    return DIM3_OK;
}

int dim3_server_drop_relation(dim3_server* s, uint64_t handle) {
    if (!s) return DIM3_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->rels.find(handle);
    if (it == s->rels.end() || it->second.dropped) return DIM3_ERR_NOT_FOUND;
    if (it->second.users) {
        it->second.dropped = true;  // freed by the last request using it
        return DIM3_OK;
    }
    hipSetDevice(s->device);
    free_rel(it->second);
    s->resident -= it->second.bytes + it->second.vbytes;
    s->rels.erase(it);
    return DIM3_OK;
}

int dim3_server::estimate_paths(const int32_t* dR, uint64_t nR, const int32_t* dS, uint64_t nS, int64_t* out) {
    std::lock_guard<std::mutex> lk(est_mu);
    hipSetDevice(device);
    if (!est_ctx) {
        const int rc = dim3_ctx_create(device, budget, &est_ctx);
        if (rc != DIM3_OK) return rc;
    }
    const int rc = dim3_estimate_join_cardinality_dev(est_ctx, dR, nR, dS, nS, nullptr, out);
    // its hash table grows with |S|: give a large one back
    if (dim3_ctx_device_bytes(est_ctx) > (64ull << 20)) dim3_ctx_release(est_ctx);
    return rc;
}

int dim3_submit(dim3_server* s, const dim3_request* req, uint64_t* ticket) {
    if (!s || !req || !ticket) return DIM3_ERR_INVALID;
    *ticket = 0;
    if (req->kind > DIM3_REQ_VALUE) return DIM3_ERR_INVALID;
    if (req->kind == DIM3_REQ_COUNT && req->opts.dense_mode == DIM3_DENSE_BITOR) return DIM3_ERR_UNSUPPORTED;
    if (req->kind == DIM3_REQ_VALUE && (req->agg < DIM3_AGG_SUM || req->agg > DIM3_AGG_AVG)) return DIM3_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ir = s->rels.find(req->r), is = s->rels.find(req->s);
    if (ir == s->rels.end() || is == s->rels.end() || ir->second.dropped || is->second.dropped)
        return DIM3_ERR_NOT_FOUND;
    // a value aggregate needs values of one type on both relations
    if (req->kind == DIM3_REQ_VALUE &&
        (!ir->second.dv || !is->second.dv || ir->second.vtype != is->second.vtype))
        return DIM3_ERR_INVALID;
    if (s->pending.size() >= s->queue_cap) return DIM3_ERR_BUSY;
    const uint64_t nR = ir->second.n, nS = is->second.n;
    uint64_t est = footprint(nR, nS), out_est = 0;
    // the request holds its relations from here (the job takes these holds)
    ir->second.users++;
    is->second.users++;
    if (req->opts.output_mode != DIM3_OUT_FINGERPRINT) {
        // a materialising request also needs its output: at most one row per
        // join path, the count EstimateJoinCardinality samples (run on the
        // device, without the server lock)
        const size_t row = req->kind == DIM3_REQ_PROJECT ? 2 * sizeof(int32_t)
                           : (req->kind == DIM3_REQ_VALUE && ir->second.vtype == DIM3_VALUE_F32)
                               ? sizeof(dim3_value_row_f32)
                               : sizeof(dim3_count_row);
        const int32_t *dR = ir->second.d, *dS = is->second.d;
        lk.unlock();
        int64_t paths = 0;
        // advisory: when the estimate cannot run (no memory for its table),
        // the request is admitted on its workspace alone and its worker
        // sizes the output from the query itself
        if (s->estimate_paths(dR, nR, dS, nS, &paths) != DIM3_OK) paths = 0;
        lk.lock();
        if (s->rels.at(req->r).dropped || s->rels.at(req->s).dropped) {
            s->unpin(req->r);
            s->unpin(req->s);
            return DIM3_ERR_NOT_FOUND;
        }
        const double rows = std::min((double)std::max<int64_t>(paths, 0), (double)nR * (double)nS);
        out_est = rows * row > 1e18 ? (uint64_t)1e18 : (uint64_t)(rows * row);
        // The estimate counts join paths; the rows are distinct (x, z)
        // groups, often far fewer (a dense block: ~1e3 paths per row).  It
        // sizes a pre-allocation only: at most a quarter of the memory the
        // relations leave, so it never refuses a request by itself; a larger
        // output is grown to the exact size the query reports.
        const uint64_t room = s->budget > s->resident ? s->budget - s->resident : 0;
        out_est = std::min<uint64_t>(out_est, room / 4);
        est += out_est;
    }
    if (s->pending.size() >= s->queue_cap) {  // filled while the estimate ran
        s->unpin(req->r);
        s->unpin(req->s);
        return DIM3_ERR_BUSY;
    }
    s->st.submitted++;
    if (s->resident + est > s->budget) {
        s->unpin(req->r);
        s->unpin(req->s);
        s->st.rejected++;
        return DIM3_ERR_ADMISSION;
    }
    const uint64_t t = s->next_ticket++;
    Job& j = s->jobs[t];
    j.req = *req;
    j.est = est;
    j.out_est = out_est;
    j.tuples = nR + nS;
    j.t_submit = Clock::now();
    s->pending.push_back(t);
    *ticket = t;
    s->cv_work.notify_one();
    return DIM3_OK;
}

int dim3_poll(dim3_server* s, uint64_t ticket, int* done) {
    if (!s || !done) return DIM3_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return DIM3_ERR_NOT_FOUND;
    *done = it->second.done ? 1 : 0;
    return DIM3_OK;
}

int dim3_wait(dim3_server* s, uint64_t ticket, dim3_result* res, double* queued_s, double* total_s) {
    if (!s || !res) return DIM3_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return DIM3_ERR_NOT_FOUND;
    s->cv_done.wait(lk, [&] { return it->second.done; });
    Job& j = it->second;
    *res = j.res;
    if (queued_s) *queued_s = std::chrono::duration<double>(j.t_start - j.t_submit).count();
    if (total_s) *total_s = std::chrono::duration<double>(j.t_end - j.t_submit).count();
    const int rc = j.status;
    s->jobs.erase(it);
    return rc;
}

int dim3_server_get_stats(dim3_server* s, dim3_server_stats* out) {
    if (!s || !out) return DIM3_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->resident_bytes = s->resident;
    uint64_t w = 0;
    for (auto* c : s->ctxs) w += dim3_ctx_device_bytes(c);
    out->workspace_bytes = w;
    return DIM3_OK;
}

}  // extern "C"
