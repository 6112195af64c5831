// xm_server.hip -- serving loop: a resident catalog answering crossmatch requests
// with continuous batching (include/xmatch.h, xm_server_*).
//
// The reference has no serving mode: BoostXmatch.cpp:301-330 starts one thread
// per GPU over a fixed job list and exits.  Here a server runs a two-stage
// pipeline over two streams of one GPU:
//
//   preparing thread (a context of its own, so its own stream and memory):
//       takes the next pieces of the waiting requests, copies their records into
//       page-locked staging, uploads them with one copy and prepares them as one
//       many-segment catalog (one segment per piece, the resident zone height);
//   joining thread (the resident catalog's context):
//       joins every piece with the resident catalog in one batched launch
//       (xm_match_batch reads the query catalog of the other context without
//       changing it), copies the pairs back once and appends each piece's pairs
//       to its request.
//
// The preparing thread forms batch k+1 as soon as the joining thread takes batch
// k, so the upload and catalog build of one batch run while the other joins, and
// requests that arrive meanwhile join batch k+1.  Requests are cut into pieces of
// at most split_rows rows; a batch takes the next piece of each waiting request
// in turn (a request with rows left goes to the back of the queue), so a 1e5-row
// request shares batches with the small requests behind it.  A request is done
// when all its pieces are joined.
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
    uint64_t next_row = 0;    // first row not yet given to a batch (preparing thread)
    uint32_t pieces_out = 0;  // pieces given to batches and not yet joined
    bool done = false;
    xm_status status = XM_OK;
    std::string error;
    // pairs, appended piece by piece by the joining thread
    std::vector<int64_t> id_q, id_c;
    std::vector<double> dist;
};

// Rows [first, first + n) of a request.
struct Piece {
    std::shared_ptr<Request> req;
    uint64_t first, n;
};

struct Batch {
    std::vector<Piece> pieces;
    uint64_t rows = 0;
    xm_catalog* q = nullptr;  // prepared on the preparing context
    xm_status status = XM_OK;
    std::string error;
};

bool valid_record(const xm_obj& o) {
    return o.ra_deg >= 0.0 && o.ra_deg < 360.0 && o.dec_deg >= -90.0 && o.dec_deg <= 90.0;
}

// Device bytes a query row may need in a batch: its record, catalog build (key,
// sort scratch, row, FP32 row), chunk share and a few result slots.
constexpr uint64_t BYTES_PER_QUERY_ROW = 24 + 8 + 8 + 32 + 16 + 16 + 4 * 40;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct xm_server {
    xm_ctx* ctx = nullptr;       // the resident catalog's context: the joins
    xm_ctx* prep = nullptr;      // the preparing thread's context
    const xm_catalog* resident = nullptr;
    double h_deg = 0;
    xm_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_ready, cv_done;
    std::deque<std::shared_ptr<Request>> pending;   // requests with rows left to batch
    std::deque<Batch> ready;                        // prepared, waiting to join
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
    uint64_t next_ticket = 1;
    bool stop = false;
    bool prep_done = false;
    std::thread preparer, joiner;
    xm_server_stats stats{};
    // both-busy accounting: threads busy now, and since when both are
    int busy = 0;
    Clock::time_point both_since;
    // the preparing thread's page-locked staging (grown on demand)
    xm_obj* h_recs = nullptr;
    size_t h_recs_cap = 0;

    void prepare_loop();
    void join_loop();
    void prepare(Batch& b);
    // pairs of each piece of b, or the batch's failure; returns the pairs
    uint64_t join(Batch& b, std::vector<std::vector<int64_t>>& iq,
                  std::vector<std::vector<int64_t>>& ic, std::vector<std::vector<double>>& dd);
    void set_busy(bool on, Clock::time_point t);  // under mu
};

void xm_server::set_busy(bool on, Clock::time_point t) {
    if (on) {
        if (++busy == 2) both_since = t;
    } else {
        if (busy-- == 2) stats.overlap_ms += ms_between(both_since, t);
    }
}

void xm_server::prepare(Batch& b) {
    const uint64_t rows = b.rows;
    if (rows > h_recs_cap) {
        if (h_recs) (void)hipHostFree(h_recs);
        h_recs = nullptr;
        h_recs_cap = 0;
        if (hipHostMalloc(reinterpret_cast<void**>(&h_recs), rows * sizeof(xm_obj),
                          hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            h_recs = nullptr;
            b.status = XM_ENOMEM;
            b.error = "page-locked staging allocation failed";
            return;
        }
        h_recs_cap = rows;
    }
    uint64_t off = 0;
    for (const Piece& p : b.pieces) {
        // a request's records do not change until all its pieces are joined
        std::memcpy(h_recs + off, p.req->recs.data() + p.first, p.n * sizeof(xm_obj));
        off += p.n;
    }
    void* d = nullptr;
    xm_status st = xm_device_alloc(prep, rows * sizeof(xm_obj), &d);
    if (st == XM_OK) st = xm_memcpy_h2d(prep, d, h_recs, rows * sizeof(xm_obj));
    std::vector<xm_segment> segs(b.pieces.size());
    off = 0;
    for (size_t i = 0; i < b.pieces.size(); i++) {
        segs[i].records = static_cast<const xm_obj*>(d) + off;
        segs[i].n = b.pieces[i].n;
        segs[i].zone_height_deg = h_deg;
        off += b.pieces[i].n;
    }
    if (st == XM_OK) st = xm_catalog_create_segments_device(prep, segs.data(), segs.size(), &b.q);
    if (st != XM_OK) {
        b.status = st;
        b.error = xm_last_error_message(prep);
    }
    if (d) (void)xm_device_free(prep, d);
}

void xm_server::prepare_loop() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        Batch b;
        {
            std::unique_lock<std::mutex> lk(mu);
            // form the next batch once the previous prepared one has gone to join
            cv_work.wait(lk, [&] { return (stop && pending.empty()) || (!pending.empty() && ready.empty()); });
            if (pending.empty()) {
                prep_done = true;
                cv_ready.notify_all();
                return;
            }
            // admission: rows per batch, bounded by the configured limit and by
            // what the free device memory can hold; the first piece always goes
            size_t free_b = 0, total_b = 0;
            uint64_t row_limit = cfg.max_batch_rows;
            if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
                row_limit = std::min<uint64_t>(row_limit, free_b / 2 / BYTES_PER_QUERY_ROW);
            else
                (void)hipGetLastError();
            while (!pending.empty() && b.pieces.size() < cfg.max_batch_requests) {
                std::shared_ptr<Request> r = pending.front();
                const uint64_t left = r->recs.size() - r->next_row;
                if (left == 0) {  // nothing left to give (a failed request's rest)
                    pending.pop_front();
                    continue;
                }
                const uint64_t take = std::min<uint64_t>(left, cfg.split_rows);
                if (!b.pieces.empty() && b.rows + take > row_limit) break;
                pending.pop_front();
                b.pieces.push_back(Piece{r, r->next_row, take});
                r->next_row += take;
                r->pieces_out++;
                b.rows += take;
                if (r->next_row < r->recs.size()) pending.push_back(r);  // its turn again later
            }
            if (b.pieces.empty()) continue;
            set_busy(true, Clock::now());
        }
        const auto t0 = Clock::now();
        try {
            prepare(b);
        } catch (const std::exception& e) {
            b.status = XM_ENOMEM;
            b.error = std::string("batch preparation failed: ") + e.what();
        }
        const auto t1 = Clock::now();
        {
            std::lock_guard<std::mutex> lk(mu);
            set_busy(false, t1);
            stats.prep_ms += ms_between(t0, t1);
            ready.push_back(std::move(b));
        }
        cv_ready.notify_all();
    }
}

uint64_t xm_server::join(Batch& b, std::vector<std::vector<int64_t>>& iq,
                         std::vector<std::vector<int64_t>>& ic,
                         std::vector<std::vector<double>>& dd) {
    const size_t P = b.pieces.size();
    iq.assign(P, {});
    ic.assign(P, {});
    dd.assign(P, {});
    if (b.status != XM_OK) return 0;
    std::vector<xm_job> jobs(P);
    for (size_t i = 0; i < P; i++) jobs[i] = xm_job{uint32_t(i), 0u, b.pieces[i].req->radius_deg};
    // a batch of near-catalog queries gives about one pair per row; the join grows
    // its output by itself when that is short
    const uint64_t hint = std::max<uint64_t>(4096, 4 * b.rows);
    xm_result* res = nullptr;
    xm_status st = xm_match_batch(ctx, b.q, resident, jobs.data(), P, cfg.precision, hint, &res,
                                  nullptr);
    uint64_t n = 0;
    std::vector<int64_t> ia, ib;
    std::vector<double> d;
    std::vector<uint32_t> job;
    std::vector<uint64_t> counts(P);
    if (st == XM_OK) {
        xm_result_count(res, &n);
        ia.resize(n);
        ib.resize(n);
        d.resize(n);
        job.resize(n);
        st = xm_result_copy(ctx, res, ia.data(), ib.data(), d.data());
        if (st == XM_OK) st = xm_result_jobs(ctx, res, job.data(), counts.data());
    }
    if (st != XM_OK) {
        b.status = st;
        b.error = xm_last_error_message(ctx);
    }
    xm_result_destroy(res);
    if (st != XM_OK) return 0;
    for (size_t i = 0; i < P; i++) {
        iq[i].reserve(counts[i]);
        ic[i].reserve(counts[i]);
        dd[i].reserve(counts[i]);
    }
    for (uint64_t k = 0; k < n; k++) {
        iq[job[k]].push_back(ia[k]);
        ic[job[k]].push_back(ib[k]);
        dd[job[k]].push_back(d[k]);
    }
    return n;
}

void xm_server::join_loop() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        Batch b;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_ready.wait(lk, [&] { return !ready.empty() || prep_done; });
            if (ready.empty()) return;  // stopped, and every prepared batch joined
            b = std::move(ready.front());
            ready.pop_front();
            set_busy(true, Clock::now());
        }
        cv_work.notify_all();  // the preparing thread may form the next batch now
        const auto t0 = Clock::now();
        std::vector<std::vector<int64_t>> iq, ic;
        std::vector<std::vector<double>> dd;
        uint64_t pairs = 0;
        try {
            pairs = join(b, iq, ic, dd);
        } catch (const std::exception& e) {
            b.status = XM_ENOMEM;
            b.error = std::string("batch failed: ") + e.what();
        }
        xm_catalog_destroy(b.q);
        b.q = nullptr;
        // each piece's pairs to its request; only this thread appends to them
        for (size_t i = 0; i < b.pieces.size() && b.status == XM_OK; i++) {
            Request& r = *b.pieces[i].req;
            r.id_q.insert(r.id_q.end(), iq[i].begin(), iq[i].end());
            r.id_c.insert(r.id_c.end(), ic[i].begin(), ic[i].end());
            r.dist.insert(r.dist.end(), dd[i].begin(), dd[i].end());
        }
        const auto t1 = Clock::now();
        {
            std::lock_guard<std::mutex> lk(mu);
            for (const Piece& p : b.pieces) {
                Request& r = *p.req;
                r.pieces_out--;
                if (b.status != XM_OK && r.status == XM_OK) {
                    r.status = b.status;
                    r.error = b.error;
                    r.next_row = r.recs.size();  // no more pieces of a failed request
                }
                if (r.pieces_out == 0 && r.next_row == r.recs.size() && !r.done) {
                    r.recs.clear();
                    r.recs.shrink_to_fit();
                    r.t_done = t1;
                    r.done = true;
                    stats.requests_done++;
                }
            }
            if (b.status != XM_OK) {
                // failed requests with rows still queued leave the queue
                for (auto it = pending.begin(); it != pending.end();)
                    it = (*it)->status != XM_OK ? pending.erase(it) : it + 1;
            }
            set_busy(false, t1);
            stats.rows += b.rows;
            stats.matches += pairs;
            stats.batches++;
            stats.pieces += b.pieces.size();
            stats.busy_ms += ms_between(t0, t1);
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
    if (!s->cfg.max_batch_rows) s->cfg.max_batch_rows = uint64_t(1) << 18;
    if (!s->cfg.max_batch_requests) s->cfg.max_batch_requests = 4096;
    if (!s->cfg.split_rows) s->cfg.split_rows = 16384;
    if (s->cfg.precision != XM_FP64_PLAIN && s->cfg.precision != XM_FP32_FILTER_FP64_RECHECK) {
        delete s;
        xm::set_error(ctx, "unknown precision mode");
        return XM_EINVAL;
    }
    const xm_status st = xm_ctx_create(ctx->device, &s->prep);
    if (st != XM_OK) {
        xm::set_error(ctx, std::string("server context: ") + xm_last_error_message(nullptr));
        delete s;
        return st;
    }
    // FP32 form over a resident catalog stored with FP32 rows: the query catalogs
    // get theirs in the gather, so the filter reads FP32 rows on both sides (else it
    // rounds the FP64 rows in registers)
    if (s->cfg.precision == XM_FP32_FILTER_FP64_RECHECK && resident->row32)
        (void)xm_ctx_set_param(s->prep, "fp32_rows", 1);
    try {
        s->preparer = std::thread([s] { s->prepare_loop(); });
        s->joiner = std::thread([s] { s->join_loop(); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        if (s->preparer.joinable()) s->preparer.join();
        if (s->joiner.joinable()) s->joiner.join();
        xm_ctx_destroy(s->prep);
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
        if (n) {
            srv->pending.push_back(r);
        } else {
            r->t_done = r->t_submit;  // nothing to join
            r->done = true;
            srv->stats.requests_done++;
        }
        *ticket = r->ticket;
    }
    if (n) srv->cv_work.notify_all();
    else srv->cv_done.notify_all();
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
    if (latency_ms) *latency_ms = ms_between(r->t_submit, r->t_done);
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
    srv->cv_ready.notify_all();
    // the preparing thread drains the queue and stops; the joining thread joins
    // every prepared batch, then stops
    if (srv->preparer.joinable()) srv->preparer.join();
    if (srv->joiner.joinable()) srv->joiner.join();
    if (srv->h_recs) (void)hipHostFree(srv->h_recs);
    xm_ctx_destroy(srv->prep);
    delete srv;
}

}  // extern "C"
