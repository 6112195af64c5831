// xm_server.hip -- serving loop: a resident catalog answering crossmatch requests
// with continuous batching (include/xmatch.h, xm_server_*).
//
// The reference has no serving mode: BoostXmatch.cpp:301-330 starts one thread
// per GPU over a fixed job list and exits.  Here one worker thread per server
// drains a request queue: every pass takes the waiting requests (up to the
// admission limits), uploads their records with one copy, prepares them as one
// many-segment catalog (one segment per request, the resident catalog's zone
// height), joins every segment with the resident catalog in one batched launch
// (xm_match_batch), copies the pairs back once and splits them per request by
// the job column.  Requests that arrive meanwhile wait for the next pass, so the
// batch size follows the load.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "xm_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    std::vector<xm_obj> recs;
    double radius_deg = 0;
    Clock::time_point t_submit, t_done;
    bool done = false;
    xm_status status = XM_OK;
    std::string error;
    std::vector<int64_t> id_q, id_c;
    std::vector<double> dist;
};

bool valid_record(const xm_obj& o) {
    return o.ra_deg >= 0.0 && o.ra_deg < 360.0 && o.dec_deg >= -90.0 && o.dec_deg <= 90.0;
}

// Device bytes a query row may need in a batch: its record, catalog build (key,
// sort scratch, row), chunk share and a few result slots.
constexpr uint64_t BYTES_PER_QUERY_ROW = 24 + 8 + 8 + 32 + 16 + 4 * 40;

}  // namespace

struct xm_server {
    xm_ctx* ctx = nullptr;
    const xm_catalog* resident = nullptr;
    double h_deg = 0;
    xm_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    xm_server_stats stats{};
    // host staging of a batch (pinned, grown on demand)
    xm_obj* h_recs = nullptr;
    size_t h_recs_cap = 0;

    void run();
    // joins one batch; returns (query rows, pairs)
    std::pair<uint64_t, uint64_t> run_batch(std::vector<std::shared_ptr<Request>>& batch);
};

namespace {

void fail_batch(std::vector<std::shared_ptr<Request>>& batch, xm_status st, const std::string& m) {
    for (auto& r : batch) {
        r->status = st;
        r->error = m;
    }
}

}  // namespace

std::pair<uint64_t, uint64_t> xm_server::run_batch(std::vector<std::shared_ptr<Request>>& batch) {
    uint64_t rows = 0;
    for (auto& r : batch) rows += r->recs.size();
    // one upload of all query records through pinned staging
    if (rows > h_recs_cap) {
        if (h_recs) (void)hipHostFree(h_recs);
        h_recs = nullptr;
        h_recs_cap = 0;
        if (hipHostMalloc(reinterpret_cast<void**>(&h_recs), rows * sizeof(xm_obj),
                          hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            h_recs = nullptr;
            fail_batch(batch, XM_ENOMEM, "pinned staging allocation failed");
            return {0, 0};
        }
        h_recs_cap = rows;
    }
    std::vector<xm_segment> segs(batch.size());
    std::vector<xm_job> jobs(batch.size());
    uint64_t off = 0;
    for (size_t i = 0; i < batch.size(); i++) {
        const auto& r = batch[i];
        if (!r->recs.empty())
            std::memcpy(h_recs + off, r->recs.data(), r->recs.size() * sizeof(xm_obj));
        off += r->recs.size();
    }
    void* d_recs = nullptr;
    xm_status st = XM_OK;
    if (rows) {
        st = xm_device_alloc(ctx, rows * sizeof(xm_obj), &d_recs);
        if (st == XM_OK) st = xm_memcpy_h2d(ctx, d_recs, h_recs, rows * sizeof(xm_obj));
    }
    off = 0;
    for (size_t i = 0; i < batch.size(); i++) {
        const uint64_t n = batch[i]->recs.size();
        segs[i].records = n ? static_cast<const xm_obj*>(d_recs) + off : nullptr;
        segs[i].n = n;
        segs[i].zone_height_deg = h_deg;
        jobs[i].seg_a = uint32_t(i);
        jobs[i].seg_b = 0;
        jobs[i].radius_deg = batch[i]->radius_deg;
        off += n;
    }
    xm_catalog* q = nullptr;
    xm_result* res = nullptr;
    if (st == XM_OK) st = xm_catalog_create_segments_device(ctx, segs.data(), segs.size(), &q);
    if (d_recs) (void)xm_device_free(ctx, d_recs);
    // a batch of near-catalog queries gives about one pair per row; the join grows
    // its output by itself when that is short
    const uint64_t hint = std::max<uint64_t>(4096, 4 * rows);
    if (st == XM_OK)
        st = xm_match_batch(ctx, q, resident, jobs.data(), jobs.size(), cfg.precision, hint, &res,
                            nullptr);
    if (st != XM_OK) {
        fail_batch(batch, st, xm_last_error_message(ctx));
        xm_result_destroy(res);
        xm_catalog_destroy(q);
        return {0, 0};
    }
    uint64_t n = 0;
    xm_result_count(res, &n);
    std::vector<int64_t> ia(n), ib(n);
    std::vector<double> d(n);
    std::vector<uint32_t> job(n);
    std::vector<uint64_t> counts(batch.size());
    st = xm_result_copy(ctx, res, ia.data(), ib.data(), d.data());
    if (st == XM_OK) st = xm_result_jobs(ctx, res, job.data(), counts.data());
    xm_result_destroy(res);
    xm_catalog_destroy(q);
    if (st != XM_OK) {
        fail_batch(batch, st, xm_last_error_message(ctx));
        return {0, 0};
    }
    for (size_t i = 0; i < batch.size(); i++) {
        batch[i]->id_q.resize(counts[i]);
        batch[i]->id_c.resize(counts[i]);
        batch[i]->dist.resize(counts[i]);
    }
    std::vector<uint64_t> fill(batch.size(), 0);
    for (uint64_t k = 0; k < n; k++) {
        Request& r = *batch[job[k]];
        const uint64_t at = fill[job[k]]++;
        r.id_q[at] = ia[k];
        r.id_c[at] = ib[k];
        r.dist[at] = d[k];
    }
    return {rows, n};
}

void xm_server::run() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !pending.empty(); });
            if (pending.empty()) return;  // stop requested and the queue is drained
            // admission: rows per batch, bounded by the configured limit and by
            // what the free device memory can hold; the first request always goes
            size_t free_b = 0, total_b = 0;
            uint64_t row_limit = cfg.max_batch_rows;
            if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
                row_limit = std::min<uint64_t>(row_limit, free_b / 2 / BYTES_PER_QUERY_ROW);
            else
                (void)hipGetLastError();
            uint64_t rows = 0;
            while (!pending.empty() && batch.size() < cfg.max_batch_requests) {
                const uint64_t n = pending.front()->recs.size();
                if (!batch.empty() && rows + n > row_limit) break;
                rows += n;
                batch.push_back(pending.front());
                pending.pop_front();
            }
        }
        const auto t0 = Clock::now();
        std::pair<uint64_t, uint64_t> done{0, 0};
        try {
            done = run_batch(batch);
        } catch (const std::exception& e) {
            fail_batch(batch, XM_ENOMEM, std::string("batch failed: ") + e.what());
        }
        const auto t1 = Clock::now();
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& r : batch) {
                r->recs.clear();
                r->recs.shrink_to_fit();
                r->t_done = t1;
                r->done = true;
            }
            stats.rows += done.first;
            stats.matches += done.second;
            stats.batches++;
            stats.requests_done += batch.size();
            stats.busy_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
        cv_done.notify_all();
    }
}

extern "C" {

xm_status xm_server_create(xm_ctx* ctx, const xm_catalog* resident, const xm_server_config* cfg,
                           xm_server** out) {
    if (!ctx || !resident || !out) return XM_EINVAL;
    *out = nullptr;
    uint64_t nseg = 0;
    if (xm_catalog_segment_count(resident, &nseg) != XM_OK || nseg != 1) {
        xm::set_error(ctx, "the resident catalog must have one segment");
        return XM_EINVAL;
    }
    xm_server* s = new (std::nothrow) xm_server();
    if (!s) return XM_ENOMEM;
    s->ctx = ctx;
    s->resident = resident;
    xm_catalog_info(resident, nullptr, &s->h_deg, nullptr, nullptr);
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.max_batch_rows) s->cfg.max_batch_rows = uint64_t(1) << 24;
    if (!s->cfg.max_batch_requests) s->cfg.max_batch_requests = 4096;
    if (s->cfg.precision != XM_FP64_PLAIN && s->cfg.precision != XM_FP32_FILTER_FP64_RECHECK) {
        delete s;
        xm::set_error(ctx, "unknown precision mode");
        return XM_EINVAL;
    }
    try {
        s->worker = std::thread([s] { s->run(); });
    } catch (...) {
        delete s;
        return XM_EINTERNAL;
    }
    *out = s;
    return XM_OK;
}

xm_status xm_server_submit(xm_server* srv, const xm_obj* records, uint64_t n, double radius_deg,
                           uint64_t* ticket) {
    if (!srv || !ticket || (n && !records)) return XM_EINVAL;
    if (!(radius_deg > 0.0) || !std::isfinite(radius_deg) || radius_deg >= 90.0 ||
        std::ceil(radius_deg / srv->h_deg) > 1e6)
        return XM_EINVAL;
    if (n >= 0xffffffffull) return XM_EINVAL;
    for (uint64_t i = 0; i < n; i++)
        if (!valid_record(records[i])) return XM_EINVAL;
    auto r = std::make_shared<Request>();
    try {
        r->recs.assign(records, records + n);
    } catch (...) {
        return XM_ENOMEM;
    }
    r->radius_deg = radius_deg;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) return XM_EINVAL;
        r->ticket = srv->next_ticket++;
        r->t_submit = Clock::now();
        srv->live[r->ticket] = r;
        srv->pending.push_back(r);
        *ticket = r->ticket;
    }
    srv->cv_work.notify_one();
    return XM_OK;
}

xm_status xm_server_poll(xm_server* srv, uint64_t ticket, int32_t timeout_ms, uint64_t* n_matches) {
    if (!srv) return XM_EINVAL;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->live.find(ticket);
    if (it == srv->live.end()) return XM_EINVAL;
    std::shared_ptr<Request> r = it->second;
    if (!r->done) {
        if (timeout_ms < 0)
            srv->cv_done.wait(lk, [&] { return r->done; });
        else if (timeout_ms > 0)
            srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return r->done; });
    }
    if (!r->done) return XM_EAGAIN;
    if (r->status != XM_OK) return r->status;
    if (n_matches) *n_matches = r->id_q.size();
    return XM_OK;
}

xm_status xm_server_fetch(xm_server* srv, uint64_t ticket, int64_t* id_query, int64_t* id_catalog,
                          double* dist_arcsec, double* latency_ms) {
    if (!srv) return XM_EINVAL;
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->live.find(ticket);
        if (it == srv->live.end() || !it->second->done) return XM_EINVAL;
        r = it->second;
        srv->live.erase(it);  // a failed request is forgotten too
        if (r->status != XM_OK) return r->status;
    }
    const size_t n = r->id_q.size();
    if (id_query && n) std::memcpy(id_query, r->id_q.data(), n * 8);
    if (id_catalog && n) std::memcpy(id_catalog, r->id_c.data(), n * 8);
    if (dist_arcsec && n) std::memcpy(dist_arcsec, r->dist.data(), n * 8);
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    return XM_OK;
}

const char* xm_server_error(xm_server* srv, uint64_t ticket) {
    if (!srv) return "null server";
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->live.find(ticket);
    if (it == srv->live.end()) return "unknown ticket";
    return it->second->error.empty() ? "no error" : it->second->error.c_str();
}

xm_status xm_server_get_stats(xm_server* srv, xm_server_stats* stats) {
    if (!srv || !stats) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(srv->mu);
    *stats = srv->stats;
    return XM_OK;
}

void xm_server_destroy(xm_server* srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    if (srv->h_recs) (void)hipHostFree(srv->h_recs);
    delete srv;
}

}  // extern "C"
