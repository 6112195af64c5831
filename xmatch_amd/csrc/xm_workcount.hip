// xm_workcount.hip -- host-side work counts of a join, from the catalogs' sorted
// keys and zone tables copied to the host (never from device counters).
//
//   w_ref  = sum_z nA(z) * sum_{|z'-z|<=nz} nB(z')
//            the dense tests of the reference's zone loop (Worker.cu:330-359)
//   w_cand = #{(a,b): |za-zb| <= nz, cyclic |ra_a-ra_b| <= alpha(za)}
//            the pairs that reach the reference's chord test (Worker.cu:89-122),
//            alpha = calc_alpha (Common.cu:16-23) without widening, ra = atan2(y, x)
//            of the rows' unit vectors
//   w_tile = the tests the pair kernel executes under the job's shape (chunk rows
//            and dense rule, job_shape): per dense chunk, rows * B rows in its
//            windows; per sparse chunk, the chunk rows within each window B row's
//            search range (same integer window logic as the kernel: join_quanta,
//            window_intervals, chunk_is_dense)
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "xm_internal.h"

using namespace xm;

namespace {

double ref_alpha_rad(double theta_deg, double h_deg, int32_t zone) {
    const double d1 = std::fabs(zone * h_deg - 90.0);
    const double d2 = std::fabs((zone + 1) * h_deg - 90.0);
    const double d = std::max(d1, d2) + 0.01 * theta_deg;
    if (d + theta_deg > 89.99) return -1.0;  // whole circle
    return std::fabs(std::atan(std::sin(theta_deg / RAD2DEG)) /
                     std::sqrt(std::fabs(std::cos((d - theta_deg) / RAD2DEG) *
                                         std::cos((d + theta_deg) / RAD2DEG))));
}

uint64_t lower(const std::vector<uint32_t>& k, uint64_t lo, uint64_t hi, uint64_t t) {
    return uint64_t(std::lower_bound(k.begin() + lo, k.begin() + hi, t) - k.begin());
}

// One segment's rows on the host: keys (ra quanta), rows, and its zone table as
// local row indices (zoff[z - zmin] for z in [zmin, zmax + 1]).
struct HostSeg {
    SegDesc d;
    std::vector<uint32_t> key;
    std::vector<Row64> row;
    std::vector<uint32_t> zoff;
};

void fetch_seg(xm_ctx* ctx, const xm_catalog* c, uint32_t s, HostSeg& h) {
    h.d = c->seg[s];
    const uint64_t n = h.d.n;
    h.key.resize(n);
    h.row.resize(n);
    const uint64_t nt = n ? uint64_t(h.d.zmax - h.d.zmin) + 2 : 1;
    h.zoff.assign(nt, 0);
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) XM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    };
    d2h(h.key.data(), c->key + h.d.row_base, n * sizeof(uint32_t));
    d2h(h.row.data(), static_cast<const Row64*>(c->row) + h.d.row_base, n * sizeof(Row64));
    if (n) d2h(h.zoff.data(), c->zoff + h.d.zoff_base, nt * 4);
    XM_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& z : h.zoff) z -= h.d.row_base;
}

}  // namespace

extern "C" xm_status xm_work_counts_job(xm_ctx* ctx, const xm_catalog* a, const xm_catalog* b,
                                        const xm_job* job, uint64_t* w_ref, uint64_t* w_cand,
                                        uint64_t* w_tile) {
    if (!ctx || !a || !b || !job) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        XM_HIP(hipSetDevice(ctx->device));
        const double radius_deg = job->radius_deg;
        if (job->seg_a >= a->seg.size() || job->seg_b >= b->seg.size() || !(radius_deg > 0.0) ||
            a->seg[job->seg_a].h_deg != b->seg[job->seg_b].h_deg) {
            xm::set_error(ctx, "bad segment, bad radius or zone heights differ");
            return XM_EINVAL;
        }
        HostSeg A, B;
        fetch_seg(ctx, a, job->seg_a, A);
        fetch_seg(ctx, b, job->seg_b, B);
        const std::vector<uint32_t>& ka = A.key;
        const std::vector<uint32_t>& kb = B.key;
        const std::vector<uint32_t>& za = A.zoff;
        const std::vector<uint32_t>& zb = B.zoff;
        const std::vector<Row64>& ra_rows = A.row;
        const std::vector<Row64>& rb_rows = B.row;
        struct Info { uint64_t n; int32_t zmin, zmax, n_zones; double h_deg; int64_t ra_gran; };
        const Info ai{A.d.n, A.d.zmin, A.d.zmax, A.d.n_zones, A.d.h_deg, A.d.ra_gran};
        const Info bi{B.d.n, B.d.zmin, B.d.zmax, B.d.n_zones, B.d.h_deg, B.d.ra_gran};
        const Info* a_ = &ai;
        const Info* b_ = &bi;

        // exact ra (radians in [0, 2pi)) of every row, from its unit vector
        const double two_pi = 6.283185307179586476925286766559;
        auto ra_of = [&](const std::vector<Row64>& rows) {
            std::vector<double> r(rows.size());
            for (size_t i = 0; i < rows.size(); i++) {
                double v = std::atan2(rows[i].y, rows[i].x);
                r[i] = v < 0 ? v + two_pi : v;
            }
            return r;
        };
        const std::vector<double> raA = ra_of(ra_rows), raB = ra_of(rb_rows);

        const double h = a_->h_deg;
        const int32_t nz = int32_t(std::ceil(radius_deg / h));
        const int32_t n_zones = a_->n_zones;
        const double atan_sin_theta = std::atan(std::sin(radius_deg / RAD2DEG));
        const int64_t gran = std::max(a_->ra_gran, b_->ra_gran);
        const int64_t full = int64_t(1) << RA_BITS;
        // the shape the join of this job runs with
        const JobShape shape = job_shape(ctx, A.d, B.d, radius_deg);
        const uint64_t chunk_rows = shape.chunk_rows;

        std::atomic<int64_t> next{a_->n ? a_->zmin : 1};
        const int64_t zlast = a_->n ? a_->zmax : 0;
        unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
        std::vector<uint64_t> r_ref(nth, 0), r_cand(nth, 0), r_tile(nth, 0);
        auto work = [&](unsigned t) {
            uint64_t s_ref = 0, s_cand = 0, s_tile = 0;
            for (;;) {
                const int64_t z = next.fetch_add(1);
                if (z > zlast) break;
                const uint64_t a0 = za[z - a_->zmin], a1 = za[z - a_->zmin + 1];
                if (a1 <= a0) continue;
                const double al = ref_alpha_rad(radius_deg, h, int32_t(z));
                const int64_t aq = join_quanta(radius_deg, h, int32_t(z), atan_sin_theta, gran);
                for (int64_t z2 = std::max<int64_t>(0, z - nz);
                     z2 <= std::min<int64_t>(n_zones - 1, z + nz); z2++) {
                    if (z2 < b_->zmin || z2 > b_->zmax) continue;
                    const uint64_t b0 = zb[z2 - b_->zmin], b1 = zb[z2 - b_->zmin + 1];
                    if (b1 <= b0) continue;
                    s_ref += (a1 - a0) * (b1 - b0);
                    // w_cand: per A row, B rows within the unwidened alpha (cyclic, exact
                    // ra); the key window of the row bounds the rows to look at
                    for (uint64_t i = a0; i < a1; i++) {
                        if (al < 0 || aq == WHOLE_ZONE) { s_cand += b1 - b0; continue; }
                        const Window win = window_intervals(ka[i], ka[i], aq);
                        for (int t2 = 0; t2 < win.n; t2++) {
                            const Window wt = window_part(win, t2);
                            const uint64_t j0 = lower(kb, b0, b1, wt.lo0);
                            const uint64_t j1 = lower(kb, j0, b1, uint64_t(wt.hi0) + 1);
                            for (uint64_t j = j0; j < j1; j++) {
                                double d = std::fabs(raA[i] - raB[j]);
                                if (d > two_pi - d) d = two_pi - d;
                                if (d <= al) s_cand++;
                            }
                        }
                    }
                    // w_tile: the kernel's chunks, windows and per-chunk mode
                    for (uint64_t s = a0; s < a1; s += chunk_rows) {
                        const uint64_t cnt = std::min<uint64_t>(chunk_rows, a1 - s);
                        const uint64_t q_lo = ka[s], q_hi = ka[s + cnt - 1];
                        const bool dense =
                            chunk_is_dense(q_lo, q_hi, aq, shape.dense_min, uint32_t(cnt));
                        if (dense) {
                            // runs of 64 rows, each over its own window
                            for (uint64_t sc = s; sc < s + cnt; sc += 64) {
                                const uint64_t sub = std::min<uint64_t>(64, s + cnt - sc);
                                const Window win = window_intervals(ka[sc], ka[sc + sub - 1], aq);
                                for (int t2 = 0; t2 < win.n; t2++) {
                                    const Window wt = window_part(win, t2);
                                    const uint64_t j0 = lower(kb, b0, b1, wt.lo0);
                                    const uint64_t j1 = lower(kb, j0, b1, uint64_t(wt.hi0) + 1);
                                    s_tile += sub * (j1 - j0);
                                }
                            }
                            continue;
                        }
                        const Window win = window_intervals(q_lo, q_hi, aq);
                        for (int t2 = 0; t2 < win.n; t2++) {
                            const Window wt = window_part(win, t2);
                            const uint64_t j0 = lower(kb, b0, b1, wt.lo0);
                            const uint64_t j1 = lower(kb, j0, b1, uint64_t(wt.hi0) + 1);
                            // sparse: chunk rows within +-aq of each B row's quantum
                            int32_t qs[MAX_CHUNK_ROWS];
                            for (uint64_t k = 0; k < cnt; k++) qs[k] = int32_t(ka[s + k]);
                            for (uint64_t jj = j0; jj < j1; jj++) {
                                const int32_t qb = int32_t(kb[jj]) + wt.sh0;
                                const int32_t* e0 = std::lower_bound(qs, qs + cnt, qb - int32_t(aq));
                                const int32_t* e1 = std::upper_bound(qs, qs + cnt, qb + int32_t(aq));
                                s_tile += uint64_t(e1 - e0);
                            }
                        }
                    }
                }
            }
            r_ref[t] = s_ref;
            r_cand[t] = s_cand;
            r_tile[t] = s_tile;
        };
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; t++) th.emplace_back(work, t);
        for (auto& t : th) t.join();
        uint64_t s1 = 0, s2 = 0, s3 = 0;
        for (unsigned t = 0; t < nth; t++) { s1 += r_ref[t]; s2 += r_cand[t]; s3 += r_tile[t]; }
        if (w_ref) *w_ref = s1;
        if (w_cand) *w_cand = s2;
        if (w_tile) *w_tile = s3;
        (void)full;
        return XM_OK;
    } catch (const DeviceError& e) {
        xm::set_error(ctx, e.what);
        return XM_EDEVICE;
    } catch (const std::exception& e) {
        xm::set_error(ctx, e.what());
        return XM_EINTERNAL;
    }
}

extern "C" xm_status xm_work_counts(xm_ctx* ctx, const xm_catalog* a, const xm_catalog* b,
                                    double radius_deg, uint64_t* w_ref, uint64_t* w_cand,
                                    uint64_t* w_tile) {
    const xm_job job{0, 0, radius_deg};
    return xm_work_counts_job(ctx, a, b, &job, w_ref, w_cand, w_tile);
}
