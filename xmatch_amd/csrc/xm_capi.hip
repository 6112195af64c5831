// xm_capi.hip -- the C ABI (include/xmatch.h): context, memory, catalog
// preparation, join with overflow-and-retry, results, timing, work counts.
//
// Host-side counterparts in the reference (paths relative to its repository):
//   CudaManager / CudaContext  BoostXmatch/BoostXmatch/CudaManager.cpp:8-65,
//                              CudaContext.cpp:5-27          -> xm_ctx
//   Sorter::Sort               Sorter.cu:141-185             -> xm_catalog_create*
//   Worker::Match              Worker.cu:243-428             -> xm_match
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <thread>

#include "xm_internal.h"

using namespace xm;

namespace {

thread_local std::string g_tls_error = "no error";
// pinned host scratch: region counters and their prefix sums
constexpr size_t PINNED_BYTES = size_t(MAX_REGIONS) * (REGION_STRIDE + 1) * sizeof(uint64_t);

template <typename F>
xm_status guarded(xm_ctx* ctx, F&& f) {
    try {
        if (ctx) XM_HIP(hipSetDevice(ctx->device));
        f();
        return XM_OK;
    } catch (const InvalidArg& e) {
        if (ctx) ctx->last_error = e.what;
        g_tls_error = e.what;
        return XM_EINVAL;
    } catch (const OutOfMemory& e) {
        if (ctx) ctx->last_error = e.what;
        g_tls_error = e.what;
        return XM_ENOMEM;
    } catch (const DeviceError& e) {
        const std::string m = e.what;
        if (ctx) ctx->last_error = m;
        g_tls_error = m;
        return e.code == hipErrorOutOfMemory ? XM_ENOMEM : XM_EDEVICE;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_error = "host allocation failed";
        g_tls_error = "host allocation failed";
        return XM_ENOMEM;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        g_tls_error = e.what();
        return XM_EINTERNAL;
    }
}

void sync(xm_ctx* ctx) { XM_HIP(hipStreamSynchronize(ctx->stream)); }

int bit_length(uint32_t v) {
    int b = 0;
    while (v) { b++; v >>= 1; }
    return b;
}

// Build the A-side chunk table: zones in [zmin, zmax] split into runs of at most
// CHUNK_ROWS rows.  No synchronisation: the table is sized for the upper bound
// n / CHUNK_ROWS + zones and the kernels read the true count from the device.
void ensure_chunks(xm_ctx* ctx, xm_catalog* A) {
    if (A->chunks || A->n == 0) return;
    const int32_t nzone = A->zmax - A->zmin + 1;
    uint32_t* counts = static_cast<uint32_t*>(dalloc(ctx, (size_t(nzone) + 1) * sizeof(uint32_t)));
    launch_fill_u32(ctx, counts + nzone, 1, 0u);
    launch_chunk_counts(ctx, A->zoff, nzone, counts);
    scan_exclusive_u32(ctx, counts, uint64_t(nzone) + 1);
    const uint64_t bound = A->n / CHUNK_ROWS + uint64_t(nzone);
    Chunk* chunks = static_cast<Chunk*>(dalloc(ctx, std::max<uint64_t>(bound, 1) * sizeof(Chunk)));
    launch_chunk_fill(ctx, A->zoff, counts, A->zmin, nzone, chunks);
    A->chunks = chunks;
    A->max_chunks = bound;
    A->chunk_start = counts;
    A->n_chunks_dev = counts + nzone;
}

PairParams make_params(const xm_ctx* ctx, const xm_catalog* A, const xm_catalog* B,
                       double radius_deg) {
    PairParams pp;
    pp.theta_deg = radius_deg;
    pp.h_deg = A->h_deg;
    pp.atan_sin_theta = std::atan(std::sin(radius_deg / RAD2DEG));
    pp.chord2 = chord2_of_radius(radius_deg);
    pp.gran = std::max(A->ra_gran, B->ra_gran);
    // FP32 filter thresholds.  Unit-vector components rounded to FP32 move by at
    // most 2^-25; the FP32 differences, products and sums add a relative 2^-24 each.
    // So |sqrt(d2_32) - |u_a - u_b|| <= 2^-24 (sqrt(3) + 2.5 c) + O(2^-48) near the
    // threshold chord c.  The band is twice that bound; (c + band)^2 is rounded up
    // and (c - band)^2 down to FP32.
    const double c = std::sqrt(pp.chord2);
    const double band = 2.0 * std::ldexp(std::sqrt(3.0) + 2.5 * c + 1e-6, -24);
    const double hi = (c + band) * (c + band);
    const double lo = c > band ? (c - band) * (c - band) : 0.0;
    float fhi = float(hi), flo = float(lo);
    if (double(fhi) < hi) fhi = std::nextafter(fhi, INFINITY);
    if (double(flo) > lo) flo = std::nextafter(flo, 0.0f);
    pp.t_hi32 = fhi;
    pp.t_lo32 = flo;
    pp.nz = int32_t(std::ceil(radius_deg / A->h_deg));  // Worker.cu:313
    pp.n_zones = A->n_zones;
    pp.dense_min = ctx->dense_min;
    return pp;
}

void check_join_args(const xm_catalog* a, const xm_catalog* b, double radius_deg) {
    if (!a || !b) throw InvalidArg{"null catalog"};
    if (!(radius_deg > 0.0) || !std::isfinite(radius_deg) || radius_deg >= 90.0)
        throw InvalidArg{"radius must be in (0, 90) degrees"};
    if (a->h_deg != b->h_deg)
        throw InvalidArg{"catalogs were prepared with different zone heights"};
    if (std::ceil(radius_deg / a->h_deg) > 1e6)
        throw InvalidArg{"radius / zone height is too large"};
}

}  // namespace

namespace xm {

namespace {
size_t size_class(size_t bytes) {
    if (bytes <= 256) return 256;
    size_t p = 256;
    while (p * 2 < bytes) p *= 2;  // p < bytes <= 2p
    for (int q = 1; q <= 4; q++) {
        const size_t c = p + (p / 4) * size_t(q);
        if (c >= bytes) return c;
    }
    return 2 * p;
}
}  // namespace

void* dalloc(xm_ctx* ctx, size_t bytes) {
    const size_t cls = size_class(bytes ? bytes : 1);
    for (size_t i = ctx->free_blocks.size(); i-- > 0;) {
        if (ctx->free_blocks[i].first == cls) {
            void* p = ctx->free_blocks[i].second;
            ctx->free_blocks[i] = ctx->free_blocks.back();
            ctx->free_blocks.pop_back();
            ctx->live_blocks.push_back({p, cls});
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e == hipErrorOutOfMemory && !ctx->free_blocks.empty()) {
        // give the cache back (after the stream drains) and try once more
        (void)hipGetLastError();
        release_cache(ctx);
        e = hipMalloc(&p, cls);
    }
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        throw OutOfMemory{"device allocation of " + std::to_string(bytes) + " bytes failed"};
    }
    XM_HIP(e);
    ctx->cached_bytes += cls;
    ctx->live_blocks.push_back({p, cls});
    return p;
}

void dfree(xm_ctx* ctx, void* p) {
    if (!p) return;
    for (size_t i = ctx->live_blocks.size(); i-- > 0;) {
        if (ctx->live_blocks[i].first == p) {
            ctx->free_blocks.push_back({ctx->live_blocks[i].second, p});
            ctx->live_blocks[i] = ctx->live_blocks.back();
            ctx->live_blocks.pop_back();
            return;
        }
    }
    throw std::runtime_error("dfree of a pointer the context did not allocate");
}

void release_cache(xm_ctx* ctx) {
    if (ctx->free_blocks.empty()) return;
    XM_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& fb : ctx->free_blocks) {
        XM_HIP(hipFree(fb.second));
        ctx->cached_bytes -= fb.first;
    }
    ctx->free_blocks.clear();
}

void timing_begin(xm_ctx* ctx, const char* name, hipEvent_t* ev) {
    (void)name;
    hipEvent_t e;
    if (!ctx->event_pool.empty()) {
        e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
    } else {
        XM_HIP(hipEventCreate(&e));
    }
    XM_HIP(hipEventRecord(e, ctx->stream));
    *ev = e;
}

void timing_end(xm_ctx* ctx, const char* name, hipEvent_t start) {
    hipEvent_t e;
    if (!ctx->event_pool.empty()) {
        e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
    } else {
        XM_HIP(hipEventCreate(&e));
    }
    XM_HIP(hipEventRecord(e, ctx->stream));
    ctx->launches.push_back(TimedLaunch{name, start, e});
}

}  // namespace xm

namespace {

void resolve_timing(xm_ctx* ctx) {
    if (ctx->launches.empty()) return;
    sync(ctx);
    for (auto& l : ctx->launches) {
        float ms = 0.f;
        XM_HIP(hipEventElapsedTime(&ms, l.start, l.stop));
        bool found = false;
        for (auto& kt : ctx->kernel_totals) {
            if (kt.first == l.name) {
                kt.second.first += ms;
                kt.second.second += 1;
                found = true;
                break;
            }
        }
        if (!found) ctx->kernel_totals.push_back({l.name, {double(ms), 1}});
        ctx->event_pool.push_back(l.start);
        ctx->event_pool.push_back(l.stop);
    }
    ctx->launches.clear();
}

void catalog_free(xm_catalog* c) {
    if (!c || !c->ctx) return;
    xm_ctx* ctx = c->ctx;
    dfree(ctx, c->key);
    dfree(ctx, c->row);
    dfree(ctx, c->row32);
    dfree(ctx, c->zoff);
    dfree(ctx, c->chunks);
    dfree(ctx, c->chunk_start);
    c->key = nullptr; c->row = nullptr; c->row32 = nullptr; c->zoff = nullptr; c->chunks = nullptr;
    c->chunk_start = nullptr; c->n_chunks_dev = nullptr;
}

void catalog_build(xm_ctx* ctx, const xm_obj* d_recs, uint64_t n, double h_deg, xm_catalog* cat) {
    cat->ctx = ctx;
    cat->n = n;
    cat->h_deg = h_deg;
    cat->n_zones = n_zones_of(h_deg);
    cat->zmin = 0;
    cat->zmax = -1;
    if (n > 0) {
        PackLayout pl = pack_layout(n, cat->n_zones);
        if (pl.ra_bits > ctx->key_ra_bits) pl.ra_bits = ctx->key_ra_bits;
        cat->ra_gran = int64_t(1) << (RA_BITS - pl.ra_bits);
        uint64_t* words = static_cast<uint64_t*>(dalloc(ctx, n * sizeof(uint64_t)));
        uint64_t* words_alt = static_cast<uint64_t*>(dalloc(ctx, n * sizeof(uint64_t)));
        int32_t* small = static_cast<int32_t*>(dalloc(ctx, 4 * sizeof(int32_t)));
        launch_fill_i32(ctx, small, 1, INT_MAX);
        launch_fill_i32(ctx, small + 1, 1, -1);
        launch_fill_i32(ctx, small + 2, 1, 0);
        launch_make_keys(ctx, d_recs, n, h_deg, cat->n_zones, pl, words, small,
                         reinterpret_cast<uint32_t*>(small + 2));
        int32_t* pin = static_cast<int32_t*>(ctx->pinned);
        XM_HIP(hipMemcpyAsync(pin, small, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
        const int32_t zmin = pin[0], zmax = pin[1];
        const uint32_t nbad = uint32_t(pin[2]);
        dfree(ctx, small);
        if (nbad) {
            dfree(ctx, words);
            dfree(ctx, words_alt);
            throw InvalidArg{std::to_string(nbad) +
                             " record(s) with NaN or out-of-range coordinates "
                             "(need 0 <= ra < 360, -90 <= dec <= 90)"};
        }
        cat->zmin = zmin;
        cat->zmax = zmax;
        // Sort on the zone and ra bits (from the digit holding the lowest ra bit; the
        // row-index bits below it only order ties).  Zone digits whose bits agree
        // between zmin and zmax are constant over the segment and skipped.
        const int zone_lo = pl.ra_bits + pl.idx_bits;
        const int hi_bit = zone_lo + bit_length(uint32_t(zmax));
        uint64_t skip = 0;
        for (int d = 0; d < 8; d++) {
            if (d * 8 < zone_lo) continue;
            const int zs = d * 8 - zone_lo;
            if ((uint64_t(uint32_t(zmin)) >> zs) == (uint64_t(uint32_t(zmax)) >> zs))
                skip |= uint64_t(1) << d;
        }
        radix_sort_keys(ctx, &words, &words_alt, n, pl.idx_bits, hi_bit, skip);
        dfree(ctx, words_alt);
        cat->row = dalloc(ctx, n * sizeof(Row64));
        launch_gather(ctx, d_recs, words, n, pl, static_cast<Row64*>(cat->row));
        cat->key = words;
    }
    const int64_t ntab = int64_t(cat->zmax) - cat->zmin + 2;
    cat->zoff = static_cast<uint32_t*>(dalloc(ctx, size_t(ntab) * sizeof(uint32_t)));
    launch_zone_offsets(ctx, cat->key, n, cat->zmin, cat->zmax, cat->zoff);
}

}  // namespace

extern "C" {

int xm_abi_version(void) { return XM_ABI_VERSION; }

xm_status xm_ctx_create(int device, xm_ctx** out) {
    if (!out) return XM_EINVAL;
    *out = nullptr;
    xm_ctx* ctx = new (std::nothrow) xm_ctx();
    if (!ctx) return XM_ENOMEM;
    ctx->device = device;
    xm_status st = guarded(nullptr, [&] {
        int ndev = 0;
        XM_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev)
            throw InvalidArg{"device " + std::to_string(device) + " not present (" +
                             std::to_string(ndev) + " visible)"};
        XM_HIP(hipSetDevice(device));
        XM_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        XM_HIP(hipHostMalloc(&ctx->pinned, PINNED_BYTES, hipHostMallocDefault));
    });
    if (st != XM_OK) {
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return st;
    }
    *out = ctx;
    return XM_OK;
}

void xm_ctx_destroy(xm_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& l : ctx->launches) {
        (void)hipEventDestroy(l.start);
        (void)hipEventDestroy(l.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto& fb : ctx->free_blocks) (void)hipFree(fb.second);
    for (auto& lb : ctx->live_blocks) (void)hipFree(lb.first);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* xm_last_error_message(const xm_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : g_tls_error.c_str();
}

xm_status xm_ctx_set_timing(xm_ctx* ctx, int enable) {
    if (!ctx) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] { ctx->timing = enable != 0; });
}

xm_status xm_ctx_kernel_time(xm_ctx* ctx, const char* name, double* total_ms, uint64_t* count) {
    if (!ctx || !name) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        resolve_timing(ctx);
        double t = 0;
        uint64_t c = 0;
        for (auto& kt : ctx->kernel_totals)
            if (kt.first == name) { t = kt.second.first; c = kt.second.second; }
        if (total_ms) *total_ms = t;
        if (count) *count = c;
    });
}

xm_status xm_ctx_reset_timing(xm_ctx* ctx) {
    if (!ctx) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        resolve_timing(ctx);
        ctx->kernel_totals.clear();
    });
}

xm_status xm_ctx_set_param(xm_ctx* ctx, const char* name, int64_t value) {
    if (!ctx || !name) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        const std::string n = name;
        if (n == "dense_min") {
            if (value < 0 || value > 1000000) throw InvalidArg{"dense_min out of range"};
            ctx->dense_min = uint32_t(value);
        } else if (n == "key_ra_bits") {
            if (value < 1 || value > RA_BITS) throw InvalidArg{"key_ra_bits must be in [1, 24]"};
            ctx->key_ra_bits = int(value);
        } else {
            throw InvalidArg{"unknown parameter '" + n + "'"};
        }
    });
}

xm_status xm_ctx_get_param(xm_ctx* ctx, const char* name, int64_t* value) {
    if (!ctx || !name || !value) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        const std::string n = name;
        if (n == "dense_min") *value = ctx->dense_min;
        else if (n == "key_ra_bits") *value = ctx->key_ra_bits;
        else throw InvalidArg{"unknown parameter '" + n + "'"};
    });
}

xm_status xm_ctx_release_memory(xm_ctx* ctx, uint64_t* cached_bytes) {
    if (!ctx) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        release_cache(ctx);
        if (cached_bytes) *cached_bytes = ctx->cached_bytes;
    });
}

xm_status xm_ctx_synchronize(xm_ctx* ctx) {
    if (!ctx) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] { sync(ctx); });
}

xm_status xm_device_alloc(xm_ctx* ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        *dptr = dalloc(ctx, bytes);
        sync(ctx);
    });
}

xm_status xm_device_free(xm_ctx* ctx, void* dptr) {
    if (!ctx) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] { dfree(ctx, dptr); });
}

xm_status xm_memcpy_h2d(xm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        if (bytes) XM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        sync(ctx);
    });
}

xm_status xm_memcpy_d2h(xm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        if (bytes) XM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
    });
}

static xm_status catalog_create_impl(xm_ctx* ctx, const xm_obj* recs, bool on_device, uint64_t n,
                                     double zone_height_deg, xm_catalog** out) {
    if (!ctx || !out) return XM_EINVAL;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(ctx->mu);
    xm_catalog* cat = new (std::nothrow) xm_catalog();
    if (!cat) return XM_ENOMEM;
    xm_obj* staged = nullptr;
    xm_status st = guarded(ctx, [&] {
        if (!(zone_height_deg > 0.0) || !std::isfinite(zone_height_deg) || zone_height_deg > 180.0)
            throw InvalidArg{"zone height must be in (0, 180] degrees"};
        if (180.0 / zone_height_deg > double(1 << 24))
            throw InvalidArg{"zone height too small (more than 2^24 zones)"};
        if (n >= 0xffffffffull) throw InvalidArg{"a segment holds fewer than 2^32-1 records"};
        if (n && !recs) throw InvalidArg{"null records"};
        const xm_obj* d = recs;
        if (!on_device && n) {
            staged = static_cast<xm_obj*>(dalloc(ctx, n * sizeof(xm_obj)));
            XM_HIP(hipMemcpyAsync(staged, recs, n * sizeof(xm_obj), hipMemcpyHostToDevice, ctx->stream));
            d = staged;
        }
        catalog_build(ctx, d, n, zone_height_deg, cat);
        if (staged) { dfree(ctx, staged); staged = nullptr; }
        sync(ctx);
    });
    if (st != XM_OK) {
        (void)guarded(ctx, [&] {
            if (staged) dfree(ctx, staged);
            catalog_free(cat);
        });
        delete cat;
        return st;
    }
    *out = cat;
    return XM_OK;
}

xm_status xm_catalog_create(xm_ctx* ctx, const xm_obj* records, uint64_t n, double zone_height_deg,
                            xm_catalog** out) {
    return catalog_create_impl(ctx, records, false, n, zone_height_deg, out);
}

xm_status xm_catalog_create_device(xm_ctx* ctx, const xm_obj* d_records, uint64_t n,
                                   double zone_height_deg, xm_catalog** out) {
    return catalog_create_impl(ctx, d_records, true, n, zone_height_deg, out);
}

void xm_catalog_destroy(xm_catalog* cat) {
    if (!cat) return;
    if (cat->ctx) {
        std::lock_guard<std::mutex> lk(cat->ctx->mu);
        (void)guarded(cat->ctx, [&] { catalog_free(cat); });
    }
    delete cat;
}

xm_status xm_catalog_info(const xm_catalog* cat, uint64_t* n, double* zone_height_deg,
                          int32_t* zmin, int32_t* zmax) {
    if (!cat) return XM_EINVAL;
    if (n) *n = cat->n;
    if (zone_height_deg) *zone_height_deg = cat->h_deg;
    if (zmin) *zmin = cat->zmin;
    if (zmax) *zmax = cat->zmax;
    return XM_OK;
}

xm_status xm_catalog_export(xm_ctx* ctx, const xm_catalog* cat, int64_t* id, double* x, double* y,
                            double* z, int32_t* zone) {
    if (!ctx || !cat) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        const uint64_t n = cat->n;
        if (n == 0) return;
        std::vector<Row64> rows(n);
        XM_HIP(hipMemcpyAsync(rows.data(), cat->row, n * sizeof(Row64), hipMemcpyDeviceToHost,
                              ctx->stream));
        std::vector<uint64_t> keys;
        if (zone) {
            keys.resize(n);
            XM_HIP(hipMemcpyAsync(keys.data(), cat->key, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        }
        sync(ctx);
        for (uint64_t i = 0; i < n; i++) {
            if (id) id[i] = rows[i].id;
            if (x) x[i] = rows[i].x;
            if (y) y[i] = rows[i].y;
            if (z) z[i] = rows[i].z;
            if (zone) zone[i] = key_zone(keys[i]);
        }
    });
}

xm_status xm_match(xm_ctx* ctx, const xm_catalog* a, const xm_catalog* b, double radius_deg,
                   xm_precision precision, uint64_t capacity_hint, xm_result** out,
                   xm_match_stats* stats) {
    if (!ctx || !out) return XM_EINVAL;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(ctx->mu);
    xm_result* res = new (std::nothrow) xm_result();
    if (!res) return XM_ENOMEM;
    res->ctx = ctx;
    RegionOut ro{};
    uint64_t* scratch = nullptr;  // prefix[n_regions] + tail[2]
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    xm_match_stats st_local;
    std::memset(&st_local, 0, sizeof(st_local));
    xm_status st = guarded(ctx, [&] {
        check_join_args(a, b, radius_deg);
        if (precision != XM_FP64_PLAIN && precision != XM_FP32_FILTER_FP64_RECHECK)
            throw InvalidArg{"unknown precision mode"};
        const bool f32 = precision == XM_FP32_FILTER_FP64_RECHECK;
        xm_catalog* A = const_cast<xm_catalog*>(a);
        xm_catalog* Bm = const_cast<xm_catalog*>(b);
        ensure_chunks(ctx, A);
        if (f32) {
            for (xm_catalog* c : {A, Bm}) {
                if (c->row32 || c->n == 0) continue;
                c->row32 = dalloc(ctx, c->n * sizeof(Row32));
                launch_to_row32(ctx, static_cast<const Row64*>(c->row), c->n,
                                static_cast<Row32*>(c->row32));
            }
        }
        const PairParams pp = make_params(ctx, A, b, radius_deg);
        uint64_t cap = capacity_hint ? capacity_hint : 2 * std::max(a->n, b->n);
        cap = std::max<uint64_t>(cap, 1024);
        XM_HIP(hipEventCreate(&ev0));
        XM_HIP(hipEventCreate(&ev1));
        // Hits go to n_regions append regions (blocks spread over them round-robin),
        // a one-block scan sizes them and k_compact packs them into the result; the
        // only host round trip reads the total and the fullest region.
        ro.n_regions = std::max<uint32_t>(
            1, std::min<uint32_t>(MAX_REGIONS, pair_blocks(uint32_t(std::max<uint64_t>(A->max_chunks, 1)))));
        ro.region_cap = std::max<uint64_t>(1, (cap + ro.n_regions - 1) / ro.n_regions);
        ro.counters = static_cast<uint64_t*>(
            dalloc(ctx, size_t(ro.n_regions) * REGION_STRIDE * sizeof(uint64_t)));
        scratch = static_cast<uint64_t*>(dalloc(ctx, (size_t(ro.n_regions) + 4) * sizeof(uint64_t)));
        uint64_t* prefix = scratch;
        uint64_t* tail = scratch + ro.n_regions;
        uint64_t* pin = static_cast<uint64_t*>(ctx->pinned);
        for (int attempt = 0;; attempt++) {
            const uint64_t slots = uint64_t(ro.n_regions) * ro.region_cap;
            ro.id_a = static_cast<int64_t*>(dalloc(ctx, slots * 8));
            ro.id_b = static_cast<int64_t*>(dalloc(ctx, slots * 8));
            ro.d2 = static_cast<double*>(dalloc(ctx, slots * 8));
            res->id_a = static_cast<int64_t*>(dalloc(ctx, slots * 8));
            res->id_b = static_cast<int64_t*>(dalloc(ctx, slots * 8));
            res->dist = static_cast<double*>(dalloc(ctx, slots * 8));
            res->cap = slots;
            launch_fill_u64(ctx, ro.counters, size_t(ro.n_regions) * REGION_STRIDE, 0);
            XM_HIP(hipEventRecord(ev0, ctx->stream));
            launch_pair(ctx, A, b, pp, f32, ro);
            launch_region_scan(ctx, ro, prefix, tail);
            launch_compact(ctx, ro, prefix, res->id_a, res->id_b, res->dist);
            XM_HIP(hipEventRecord(ev1, ctx->stream));
            XM_HIP(hipMemcpyAsync(pin, tail, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  ctx->stream));
            if (A->n_chunks_dev)
                XM_HIP(hipMemcpyAsync(pin + 4, A->n_chunks_dev, sizeof(uint32_t),
                                      hipMemcpyDeviceToHost, ctx->stream));
            dfree(ctx, ro.id_a); dfree(ctx, ro.id_b); dfree(ctx, ro.d2);
            ro.id_a = nullptr; ro.id_b = nullptr; ro.d2 = nullptr;
            sync(ctx);
            float ms = 0.f;
            XM_HIP(hipEventElapsedTime(&ms, ev0, ev1));
            st_local.join_ms += ms;
            const uint64_t total = pin[0], most = pin[1];
            st_local.n_chunks = A->n_chunks_dev ? reinterpret_cast<uint32_t*>(pin + 4)[0] : 0;
            if (most <= ro.region_cap) {
                res->n = total;
                st_local.band_hits = pin[2];
                double err = 0.0;
                std::memcpy(&err, &pin[3], sizeof(double));
                st_local.max_dist_err_arcsec = err * RAD2DEG * 3600.0;
                break;
            }
            // overflow: every counter holds its region's exact total; re-run with
            // regions that fit the fullest one
            st_local.n_retries++;
            dfree(ctx, res->id_a); dfree(ctx, res->id_b); dfree(ctx, res->dist);
            res->id_a = nullptr; res->id_b = nullptr; res->dist = nullptr;
            ro.region_cap = most;
            if (attempt > 2) throw std::runtime_error("result size did not converge");
        }
        st_local.n_matches = res->n;
    });
    (void)guarded(ctx, [&] {
        dfree(ctx, ro.counters);
        dfree(ctx, scratch);
        dfree(ctx, ro.id_a); dfree(ctx, ro.id_b); dfree(ctx, ro.d2);
        if (ev0) XM_HIP(hipEventDestroy(ev0));
        if (ev1) XM_HIP(hipEventDestroy(ev1));
    });
    if (stats) *stats = st_local;
    if (st != XM_OK) {
        (void)guarded(ctx, [&] {
            dfree(ctx, res->id_a); dfree(ctx, res->id_b); dfree(ctx, res->dist);
        });
        delete res;
        return st;
    }
    *out = res;
    return XM_OK;
}

xm_status xm_result_count(const xm_result* res, uint64_t* n) {
    if (!res || !n) return XM_EINVAL;
    *n = res->n;
    return XM_OK;
}

xm_status xm_result_device_ptrs(const xm_result* res, const int64_t** id_a, const int64_t** id_b,
                                const double** dist_arcsec) {
    if (!res) return XM_EINVAL;
    if (id_a) *id_a = res->id_a;
    if (id_b) *id_b = res->id_b;
    if (dist_arcsec) *dist_arcsec = res->dist;
    return XM_OK;
}

xm_status xm_result_copy(xm_ctx* ctx, const xm_result* res, int64_t* id_a, int64_t* id_b,
                         double* dist_arcsec) {
    if (!ctx || !res) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        const size_t bytes = res->n * 8;
        if (!bytes) return;
        if (id_a) XM_HIP(hipMemcpyAsync(id_a, res->id_a, bytes, hipMemcpyDeviceToHost, ctx->stream));
        if (id_b) XM_HIP(hipMemcpyAsync(id_b, res->id_b, bytes, hipMemcpyDeviceToHost, ctx->stream));
        if (dist_arcsec)
            XM_HIP(hipMemcpyAsync(dist_arcsec, res->dist, bytes, hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
    });
}

void xm_result_destroy(xm_result* res) {
    if (!res) return;
    if (res->ctx) {
        std::lock_guard<std::mutex> lk(res->ctx->mu);
        (void)guarded(res->ctx, [&] {
            dfree(res->ctx, res->id_a);
            dfree(res->ctx, res->id_b);
            dfree(res->ctx, res->dist);
        });
    }
    delete res;
}

}  // extern "C"
