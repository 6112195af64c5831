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

// pinned host scratch: small device->host reads at the front, then a stage for
// the job table of small batches (one plain join included)
constexpr size_t PIN_JOB_BYTES = 64 * 1024;
constexpr size_t PINNED_BYTES = PIN_JOB_OFFSET + PIN_JOB_BYTES;

// The calling thread's last error message per context (null: no context).
thread_local std::vector<std::pair<const xm_ctx*, std::string>> t_errors;

std::string& error_slot(const xm_ctx* ctx) {
    for (auto& e : t_errors)
        if (e.first == ctx) return e.second;
    t_errors.push_back({ctx, "no error"});
    return t_errors.back().second;
}

void sync(xm_ctx* ctx) { XM_HIP(hipStreamSynchronize(ctx->stream)); }

int bit_length(uint32_t v) {
    int b = 0;
    while (v) { b++; v >>= 1; }
    return b;
}

// Build the A-side chunk table: every zone of segment s split into runs of at most
// rows[s] rows (a multiple of 64, at most MAX_CHUNK_ROWS).  A table built with
// other rows is replaced.  No synchronisation: the table is sized for the upper
// bound sum_s n_s / rows[s] + zones, and the kernels read the true counts from the
// device.  The segment table's device copy carries the rows per segment.
void ensure_chunks(xm_ctx* ctx, xm_catalog* A, const std::vector<uint32_t>& rows) {
    if (A->n == 0) return;
    bool same = A->chunks != nullptr;
    for (size_t s = 0; s < A->seg.size(); s++) same = same && A->seg[s].chunk_rows == rows[s];
    if (same) return;
    dfree(ctx, A->chunks);
    dfree(ctx, A->chunk_start);
    A->chunks = nullptr;
    A->chunk_start = nullptr;
    uint64_t bound = 0;
    for (size_t s = 0; s < A->seg.size(); s++) {
        SegDesc& d = A->seg[s];
        if (rows[s] == 0 || rows[s] % WAVE || rows[s] > uint32_t(MAX_CHUNK_ROWS))
            throw std::logic_error("chunk rows must be a multiple of 64 up to 512");
        d.chunk_rows = rows[s];
        if (d.n) bound += uint64_t(d.n) / d.chunk_rows + uint64_t(d.zmax - d.zmin + 1);
    }
    // the device copy of the segment table (the chunk kernels read chunk_rows), in
    // stream order; the host table changes only in a later call, after this call
    // has synchronised the stream
    XM_HIP(hipMemcpyAsync(A->seg_dev, A->seg.data(), A->seg.size() * sizeof(SegDesc),
                          hipMemcpyHostToDevice, ctx->stream));
    const uint64_t entries = A->n_zoff - 1;  // zone entries; the last zoff entry is n
    const uint32_t S = uint32_t(A->seg.size());
    ScopedBlocks tmp(ctx);
    uint32_t* counts = tmp.alloc<uint32_t>(A->n_zoff * sizeof(uint32_t));
    launch_fill_u32(ctx, counts + entries, 1, 0u);
    launch_chunk_counts(ctx, A->zoff, entries, A->seg_dev, S, counts);
    scan_exclusive_u32(ctx, counts, A->n_zoff);
    Chunk* chunks = tmp.alloc<Chunk>(std::max<uint64_t>(bound, 1) * sizeof(Chunk));
    launch_chunk_fill(ctx, A->zoff, counts, entries, A->seg_dev, S, chunks);
    tmp.keep(chunks);
    tmp.keep(counts);
    A->chunks = chunks;
    A->max_chunks = bound;
    A->chunk_start = counts;
}

// Upper bound on the chunks of segment s at rows per chunk (rows / chunk + zones).
uint64_t seg_max_chunks(const SegDesc& s, uint32_t rows) {
    return s.n ? uint64_t(s.n) / rows + uint64_t(s.zmax - s.zmin + 1) : 0;
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

PairParams make_params(const xm_ctx* ctx, const SegDesc& A, const SegDesc& B, double radius_deg) {
    PairParams pp;
    pp.theta_deg = radius_deg;
    pp.h_deg = A.h_deg;
    pp.atan_sin_theta = std::atan(std::sin(radius_deg / RAD2DEG));
    pp.chord2 = chord2_of_radius(radius_deg);
    pp.gran = std::max(A.ra_gran, B.ra_gran);
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
    pp.nz = int32_t(std::ceil(radius_deg / A.h_deg));  // Worker.cu:313
    pp.n_zones = A.n_zones;
    pp.dense_min = ctx->dense_min;
    return pp;
}

void check_radius(double radius_deg, double h_a, double h_b) {
    if (!(radius_deg > 0.0) || !std::isfinite(radius_deg) || radius_deg >= 90.0)
        throw InvalidArg{"radius must be in (0, 90) degrees"};
    if (h_a != h_b) throw InvalidArg{"segments were prepared with different zone heights"};
    if (std::ceil(radius_deg / h_a) > 1e6) throw InvalidArg{"radius / zone height is too large"};
}

void set_error(const xm_ctx* ctx, const std::string& msg) {
    error_slot(ctx) = msg;
    if (ctx) error_slot(nullptr) = msg;
}

// The pair kernel's shape when no tuned shape applies: the context's knobs.
JobShape default_shape(const xm_ctx* ctx) {
    return JobShape{ctx->chunk_rows ? ctx->chunk_rows : uint32_t(CHUNK_ROWS), ctx->dense_min};
}

uint32_t tune_key(const SegDesc& a, const SegDesc& b, double radius_deg) {
    auto density = [](const SegDesc& d) {
        const double zones = d.n ? double(d.zmax - d.zmin + 1) : 1.0;
        const double rows = std::max(1.0, double(d.n) / zones);
        return uint32_t(std::min(31L, std::lround(std::log2(rows))));
    };
    const double nz = std::ceil(radius_deg / a.h_deg);
    const uint32_t z = uint32_t(std::max(1.0, std::min(255.0, nz)));
    return density(a) | density(b) << 8 | z << 16;
}

JobShape job_shape(const xm_ctx* ctx, const SegDesc& a, const SegDesc& b, double radius_deg) {
    if (ctx->autotune) {
        const uint32_t k = tune_key(a, b, radius_deg);
        for (const auto& e : ctx->tune)
            if (e.first == k) return JobShape{e.second.chunk_rows, e.second.dense_min};
    }
    return default_shape(ctx);
}

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
    dfree(ctx, c->seg_dev);
    c->key = nullptr; c->row = nullptr; c->row32 = nullptr; c->zoff = nullptr; c->chunks = nullptr;
    c->chunk_start = nullptr; c->seg_dev = nullptr;
}

// Segment table checks and row bases; returns the total row count.
uint64_t seg_table(const std::vector<const xm_obj*>& recs, const std::vector<uint64_t>& ns,
                   const std::vector<double>& hs, xm_catalog* cat) {
    const size_t S = ns.size();
    if (S == 0) throw InvalidArg{"a catalog needs at least one segment"};
    if (S > MAX_SEGMENTS) throw InvalidArg{"more than 65536 segments in one catalog"};
    uint64_t total = 0;
    cat->seg.assign(S, SegDesc{});
    for (size_t s = 0; s < S; s++) {
        const double h = hs[s];
        if (!(h > 0.0) || !std::isfinite(h) || h > 180.0)
            throw InvalidArg{"zone height must be in (0, 180] degrees"};
        if (180.0 / h > double(1 << 24)) throw InvalidArg{"zone height too small (more than 2^24 zones)"};
        if (ns[s] && !recs[s]) throw InvalidArg{"null records"};
        SegDesc& d = cat->seg[s];
        d.recs = recs[s];
        d.row_base = uint32_t(total);
        d.n = uint32_t(ns[s]);
        d.zmin = 0;
        d.zmax = -1;
        d.zoff_base = 0;
        d.n_zones = n_zones_of(h);
        d.h_deg = h;
        d.ra_gran = 1;
        d.zone_base = 0;
        d.zone_bits = d.ra_bits = d.idx_bits = 0;
        total += ns[s];
        if (total > SORT_MAX_KEYS)
            throw InvalidArg{"a catalog holds at most 2^32-8193 records"};
    }
    return total;
}

// Zone tables of every segment (filled by the gathers), and the device copy
// of the segment table with the final zone ranges.
// Zone-offset table layout: segment s owns entries zoff_base .. zoff_base +
// (zmax - zmin) (none when empty) and one last entry holds n.  The gather fills
// it; an empty catalog holds only that last entry.
void zoff_layout(xm_ctx* ctx, xm_catalog* cat) {
    uint64_t entries = 0;
    for (auto& d : cat->seg) {
        d.zoff_base = uint32_t(entries);
        if (d.n) entries += uint64_t(d.zmax - d.zmin + 1);
    }
    cat->n_zoff = entries + 1;
    cat->zoff = static_cast<uint32_t*>(dalloc(ctx, cat->n_zoff * sizeof(uint32_t)));
    if (cat->n == 0) launch_fill_i32(ctx, reinterpret_cast<int32_t*>(cat->zoff), 1, 0);
}
void finish_tables(xm_ctx* ctx, xm_catalog* cat) {
    for (auto& d : cat->seg) d.recs = nullptr;
    if (!cat->seg_dev) cat->seg_dev = static_cast<SegDesc*>(dalloc(ctx, cat->seg.size() * sizeof(SegDesc)));
    XM_HIP(hipMemcpyAsync(cat->seg_dev, cat->seg.data(), cat->seg.size() * sizeof(SegDesc),
                          hipMemcpyHostToDevice, ctx->stream));
    // cat->seg stays as uploaded; the creating call synchronises before returning
}

// One segment: a single pass over the records packs zone | ra | row words with the
// absolute zone (the zone range comes back with the same pass).  A catalog of at
// least bucket_min_rows rows has that pass also move its records into zone buckets
// (BucketBuild), which the gather then reads from one XCD's L2 each; returns false,
// having built nothing, when the buckets' sampled plan failed or the bucket array
// does not fit the row field.
bool build_one_pass(xm_ctx* ctx, xm_catalog* cat, bool bucketed) {
    SegDesc& d = cat->seg[0];
    const uint64_t n = d.n;
    uint32_t pct = ctx->bucket_capacity_pct;
    BucketBuild bb = bucket_build(n, pct);
    // the row field holds the input row, or the record's place among the buckets
    PackLayout pl = pack_layout(n, d.n_zones);
    if (pl.ra_bits > ctx->key_ra_bits) pl.ra_bits = ctx->key_ra_bits;
    if (bucketed) {
        // The bucketed words keep the plain build's zone and ra bits, so the sort
        // orders the rows identically (ties by input row) and the catalog is the
        // same.  Where the wider row field the bucket array needs would cost an ra
        // bit, the array is cut to the positions the input's row field addresses:
        // the runs keep their capacity, or as much of it (down to 102% of the
        // estimate) as leaves 2% of the rows for the overflow run, which takes the
        // rest; with less room the build is plain.
        PackLayout pb = pack_layout(bb.c_total, d.n_zones);
        if (pb.ra_bits > ctx->key_ra_bits) pb.ra_bits = ctx->key_ra_bits;
        if (pb.ra_bits == pl.ra_bits) {
            pl = pb;
        } else {
            const uint64_t room = uint64_t(1) << pl.idx_bits;
            const uint64_t ovf_min = n / 50;
            if (room < bb.c_main_max + ovf_min) {
                const uint64_t slack = bb.c_main_max - n * pct / 100;  // the runs' fixed rows
                const uint64_t fit = room > ovf_min + slack ? room - ovf_min - slack : 0;
                const uint64_t pct_fit = fit * 100 / n;
                if (pct_fit < 102) return false;
                pct = uint32_t(std::min<uint64_t>(pct_fit, pct));
                bb = bucket_build(n, pct);
            }
            bb.c_total = room;
        }
    }
    d.idx_bits = pl.idx_bits;
    d.zone_bits = pl.zone_bits;
    d.zone_base = 0;
    ScopedBlocks tmp(ctx);
    uint64_t* words = tmp.alloc<uint64_t>(n * sizeof(uint64_t));
    uint64_t* words_alt = tmp.alloc<uint64_t>(n * sizeof(uint64_t));
    // the sort's digit counts, taken while the keys are written: digit positions
    // from the row bits up over the ra bits and every zone bit (the sort may need
    // fewer once the occupied zone range is known)
    // (8-bit digits take the most passes; 9-bit ones the most counters per pass)
    const int hist_pass = sort_passes(pl.ra_bits + pl.zone_bits, 8);
    const size_t digit_words = size_t(hist_pass) * 512;
    // scratch: zone range, bad count and the layout word; layout statistics; digit counts
    uint32_t* scratch = tmp.alloc<uint32_t>(
        (BUILD_SMALL_WORDS + LAYOUT_STATS_WORDS + digit_words) * sizeof(uint32_t));
    launch_build_init(ctx, scratch, digit_words);
    int32_t* small = reinterpret_cast<int32_t*>(scratch);
    uint32_t* stats = scratch + BUILD_SMALL_WORDS;
    uint32_t* digit_totals = stats + LAYOUT_STATS_WORDS;
    // the ra bits the sort and the key drop and the sort's digit width, decided on
    // the device from a sample
    launch_layout(ctx, d.recs, n, d.h_deg, d.n_zones, pl.ra_bits, pl.zone_bits, stats,
                  reinterpret_cast<uint32_t*>(small + 3));
    xm_obj* part = nullptr;
    uint32_t* plan = nullptr;
    const size_t runs = size_t(bb.n_buckets) * BUCKET_GROUPS;
    if (bucketed) {
        part = tmp.alloc<xm_obj>(bb.c_total * sizeof(xm_obj));
        plan = tmp.alloc<uint32_t>((2 * runs + BP_WORDS + bb.n_buckets) * sizeof(uint32_t));
        uint32_t* hist = plan + 2 * runs + BP_WORDS;
        launch_fill_u32(ctx, hist, bb.n_buckets, 0u);
        launch_bucket_plan(ctx, d.recs, n, d.h_deg, d.n_zones, stats, bb,
                           pct, hist, plan);
        launch_make_keys_bucketed(ctx, d.recs, n, d.h_deg, d.n_zones, pl, stats, bb, plan, part,
                                  words, small, reinterpret_cast<uint32_t*>(small + 2),
                                  pl.idx_bits, hist_pass, digit_totals,
                                  reinterpret_cast<uint32_t*>(small + 3));
    } else {
        launch_make_keys(ctx, d.recs, n, d.h_deg, d.n_zones, pl, words, small,
                         reinterpret_cast<uint32_t*>(small + 2), pl.idx_bits, hist_pass,
                         digit_totals, reinterpret_cast<uint32_t*>(small + 3));
    }
    int32_t* pin = static_cast<int32_t*>(ctx->pinned);
    XM_HIP(hipMemcpyAsync(pin, small, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    if (bucketed)
        XM_HIP(hipMemcpyAsync(pin + 4, plan + 2 * runs + BP_FAILED, sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    const int32_t zmin = pin[0], zmax = pin[1];
    const uint32_t nbad = uint32_t(pin[2]);
    const int ra_drop = layout_drop(uint32_t(pin[3]));
    const int digit_bits = layout_digit_bits(uint32_t(pin[3]));
    if (nbad) {
        throw InvalidArg{std::to_string(nbad) +
                         " record(s) with NaN or out-of-range coordinates "
                         "(need 0 <= ra < 360, -90 <= dec <= 90)"};
    }
    if (bucketed && pin[4]) return false;
    d.ra_bits = pl.ra_bits - ra_drop;
    d.ra_gran = int64_t(1) << (RA_BITS - d.ra_bits);
    d.zmin = zmin;
    d.zmax = zmax;
    // Sort on the ra bits and the zone bits that differ between zmin and zmax (the
    // zone bits above them are equal in every word; the row-index bits below only
    // order ties, which keep input order).
    const int zone_lo = pl.ra_bits + pl.idx_bits;
    const int hi_bit = zone_lo + bit_length(uint32_t(zmin) ^ uint32_t(zmax));
    radix_sort_keys(ctx, &words, &words_alt, n, pl.idx_bits + ra_drop, hi_bit, digit_totals,
                    digit_bits);
    tmp.free(scratch);
    tmp.free(words_alt);  // the sorted words are in words
    zoff_layout(ctx, cat);
    cat->row = dalloc(ctx, n * sizeof(Row64));
    if (ctx->fp32_rows) cat->row32 = dalloc(ctx, n * sizeof(Row32));
    cat->key = static_cast<uint32_t*>(dalloc(ctx, n * sizeof(uint32_t)));
    launch_gather(ctx, bucketed ? part : d.recs, words, n, pl, ra_drop, zmin,
                  static_cast<Row64*>(cat->row), static_cast<Row32*>(cat->row32), cat->zoff,
                  cat->n_zoff, cat->key, bucketed);
    return true;
}

void build_one(xm_ctx* ctx, xm_catalog* cat) {
    const uint64_t n = cat->seg[0].n;
    if (n == 0) return;
    const bool bucketed = ctx->bucket_min_rows && n >= ctx->bucket_min_rows &&
                          n <= BUCKET_MAX_ROWS;
    if (bucketed && build_one_pass(ctx, cat, true)) return;
    build_one_pass(ctx, cat, false);
}

// Many segments: pass 1 finds every segment's zone range, the host picks each
// segment's layout (segment-relative zone, so small segments keep more ra bits),
// pass 2 packs segment | zone | ra | row words, one radix sort orders them all.
void build_many(xm_ctx* ctx, xm_catalog* cat) {
    const uint32_t S = uint32_t(cat->seg.size());
    const uint64_t n = cat->n;
    if (n == 0) return;
    cat->seg_dev = static_cast<SegDesc*>(dalloc(ctx, S * sizeof(SegDesc)));
    XM_HIP(hipMemcpyAsync(cat->seg_dev, cat->seg.data(), S * sizeof(SegDesc), hipMemcpyHostToDevice,
                          ctx->stream));
    ScopedBlocks tmp(ctx);
    int32_t* range = tmp.alloc<int32_t>(size_t(S) * 3 * sizeof(int32_t));
    std::vector<int32_t> init(size_t(S) * 3);
    for (uint32_t s = 0; s < S; s++) {
        init[3 * s] = INT_MAX;
        init[3 * s + 1] = -1;
        init[3 * s + 2] = 0;
    }
    XM_HIP(hipMemcpyAsync(range, init.data(), init.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                          ctx->stream));
    launch_seg_range(ctx, cat->seg_dev, S, n, range);
    XM_HIP(hipMemcpyAsync(init.data(), range, init.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    sync(ctx);
    tmp.free(range);
    uint64_t nmax = 0;
    for (uint32_t s = 0; s < S; s++) {
        if (init[3 * s + 2])
            throw InvalidArg{"segment " + std::to_string(s) +
                             " holds record(s) with NaN or out-of-range coordinates "
                             "(need 0 <= ra < 360, -90 <= dec <= 90)"};
        nmax = std::max<uint64_t>(nmax, cat->seg[s].n);
    }
    const int seg_bits = bit_length_u64(uint64_t(S - 1));
    const int idx_bits = bit_length_u64(nmax - 1);
    int content = 0;
    for (uint32_t s = 0; s < S; s++) {
        SegDesc& d = cat->seg[s];
        d.idx_bits = idx_bits;
        if (d.n == 0) {
            d.zone_bits = 0;
            d.ra_bits = 1;
            continue;
        }
        d.zmin = init[3 * s];
        d.zmax = init[3 * s + 1];
        d.zone_base = d.zmin;
        d.zone_bits = bit_length_u64(uint64_t(d.zmax - d.zmin));
        const int room = 64 - seg_bits - idx_bits - d.zone_bits;
        d.ra_bits = std::min(std::min(room, RA_BITS), ctx->key_ra_bits);
        if (d.ra_bits < 1) throw InvalidArg{"segments too large to pack zone, ra and row in 64 bits"};
        d.ra_gran = int64_t(1) << (RA_BITS - d.ra_bits);
        content = std::max(content, idx_bits + d.ra_bits + d.zone_bits);
    }
    const int seg_shift = seg_bits ? content : 64;
    zoff_layout(ctx, cat);
    XM_HIP(hipMemcpyAsync(cat->seg_dev, cat->seg.data(), S * sizeof(SegDesc), hipMemcpyHostToDevice,
                          ctx->stream));
    uint64_t* words = tmp.alloc<uint64_t>(n * sizeof(uint64_t));
    uint64_t* words_alt = tmp.alloc<uint64_t>(n * sizeof(uint64_t));
    // the sort's digit counts, taken while the words are written; 9-bit digits
    // where they save a pass (the 10,000-segment batch: 44 bits in five passes
    // instead of six)
    const int sort_bits = content + seg_bits - idx_bits;
    const int db = sort_passes(sort_bits, 9) < sort_passes(sort_bits, 8) &&
                           sort_passes(sort_bits, 9) <= SEG9_MAX_PASSES
                       ? 9
                       : 8;
    const int n_pass = sort_passes(sort_bits, db);
    const size_t digit_words = size_t(n_pass) << db;
    uint32_t* digit_totals = tmp.alloc<uint32_t>(digit_words * 4);
    launch_fill_u32(ctx, digit_totals, digit_words, 0u);
    launch_seg_words(ctx, cat->seg_dev, S, n, seg_shift, words, idx_bits, n_pass, digit_totals,
                     db);
    radix_sort_keys(ctx, &words, &words_alt, n, idx_bits, content + seg_bits, digit_totals, db);
    tmp.free(digit_totals);
    tmp.free(words_alt);  // the sorted words are in words
    cat->row = dalloc(ctx, n * sizeof(Row64));
    if (ctx->fp32_rows) cat->row32 = dalloc(ctx, n * sizeof(Row32));
    cat->key = static_cast<uint32_t*>(dalloc(ctx, n * sizeof(uint32_t)));
    launch_gather_seg(ctx, cat->seg_dev, seg_shift, words, n, static_cast<Row64*>(cat->row),
                      static_cast<Row32*>(cat->row32), cat->zoff, cat->n_zoff, cat->key);
    sync(ctx);  // the host segment table is rewritten next; its uploads must be done
}

void catalog_build(xm_ctx* ctx, const std::vector<const xm_obj*>& recs,
                   const std::vector<uint64_t>& ns, const std::vector<double>& hs,
                   xm_catalog* cat) {
    cat->ctx = ctx;
    cat->n = seg_table(recs, ns, hs, cat);
    if (cat->n == 0) zoff_layout(ctx, cat);
    else if (cat->seg.size() == 1) build_one(ctx, cat);
    else build_many(ctx, cat);
    finish_tables(ctx, cat);
    sort_check(ctx);
}

// A batch of jobs on the device: JobDesc table, per-job wave counts scanned into
// job_prefix (n_jobs + 1 entries, the last = all waves), and the host-side bound
// on the waves that sizes the grid.
struct JobTable {
    JobDesc* jobs = nullptr;
    uint32_t* prefix = nullptr;
    uint32_t n_jobs = 0;
    uint64_t max_waves = 0;
    std::vector<JobDesc> host;  // upload source when the pinned stage is too small
};

// The table is uploaded asynchronously; t.host (or the context's pinned job stage)
// must stay untouched until the stream is synchronised.  shape[j] is job j's pair
// kernel shape; an A segment takes the chunk rows of the first job that uses it
// (segments no job uses keep the rows of their current chunk table).
void make_jobs(xm_ctx* ctx, xm_catalog* A, const xm_catalog* B, const xm_job* jobs,
               uint64_t n_jobs, const std::vector<JobShape>& shape, JobTable& t) {
    if (n_jobs == 0) throw InvalidArg{"no jobs"};
    if (n_jobs >= 0xffffffffull) throw InvalidArg{"too many jobs"};
    const bool pinned = n_jobs * sizeof(JobDesc) <= PIN_JOB_BYTES;
    if (!pinned) t.host.resize(n_jobs);
    JobDesc* h = pinned ? reinterpret_cast<JobDesc*>(static_cast<char*>(ctx->pinned) + PIN_JOB_OFFSET)
                        : t.host.data();
    std::vector<uint32_t> rows(A->seg.size(), 0);
    for (uint64_t j = 0; j < n_jobs; j++) {
        const xm_job& jb = jobs[j];
        if (jb.seg_a >= A->seg.size() || jb.seg_b >= B->seg.size())
            throw InvalidArg{"job " + std::to_string(j) + " names a segment the catalog lacks"};
        if (!rows[jb.seg_a]) rows[jb.seg_a] = shape[j].chunk_rows;
    }
    for (size_t s = 0; s < rows.size(); s++)
        if (!rows[s]) rows[s] = A->seg[s].chunk_rows ? A->seg[s].chunk_rows : default_shape(ctx).chunk_rows;
    uint64_t bound = 0;
    for (uint64_t j = 0; j < n_jobs; j++) {
        const xm_job& jb = jobs[j];
        const SegDesc& sa = A->seg[jb.seg_a];
        const SegDesc& sb = B->seg[jb.seg_b];
        check_radius(jb.radius_deg, sa.h_deg, sb.h_deg);
        JobDesc& d = h[j];
        d.pp = make_params(ctx, sa, sb, jb.radius_deg);
        d.pp.dense_min = shape[j].dense_min;
        d.a_zoff_lo = sa.zoff_base;
        d.a_zoff_hi = sa.zoff_base + (sa.n ? uint32_t(sa.zmax - sa.zmin + 1) : 0u);
        d.b_zoff_base = sb.zoff_base;
        d.b_zmin = sb.n ? sb.zmin : 0;
        d.b_zmax = sb.n ? sb.zmax : -1;
        d.a_chunk_base = 0;
        if (sb.n) bound += seg_max_chunks(sa, rows[jb.seg_a]);
    }
    if (bound >= 0xffffffffull) throw InvalidArg{"batch too large (more than 2^32 wavefronts)"};
    t.n_jobs = uint32_t(n_jobs);
    t.max_waves = bound;
    ensure_chunks(ctx, A, rows);
    t.jobs = static_cast<JobDesc*>(dalloc(ctx, n_jobs * sizeof(JobDesc)));
    t.prefix = static_cast<uint32_t*>(dalloc(ctx, (n_jobs + 1) * sizeof(uint32_t)));
    XM_HIP(hipMemcpyAsync(t.jobs, h, n_jobs * sizeof(JobDesc), hipMemcpyHostToDevice, ctx->stream));
    launch_fill_u32(ctx, t.prefix, n_jobs + 1, 0u);
    if (A->chunk_start) launch_job_prep(ctx, t.jobs, t.n_jobs, A->chunk_start, t.prefix);
    scan_exclusive_u32(ctx, t.prefix, n_jobs + 1);
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
        ctx->noalloc_scratch = noalloc_scratch(device);
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

const char* xm_last_error_message(const xm_ctx* ctx) { return error_slot(ctx).c_str(); }

xm_status xm_ctx_set_timing(xm_ctx* ctx, int enable) {
    if (!ctx) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] { ctx->timing = enable != 0; });
}

xm_status xm_ctx_set_timing_only(xm_ctx* ctx, const char* name) {
    if (!ctx) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] { ctx->timing_only = name ? name : ""; });
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
        } else if (n == "sort_wide_words") {
            if (value != 0 && value != 1) throw InvalidArg{"sort_wide_words must be 0 or 1"};
            ctx->sort_wide_words = value != 0;
        } else if (n == "bucket_min_rows") {
            if (value < 0) throw InvalidArg{"bucket_min_rows must be >= 0"};
            ctx->bucket_min_rows = uint64_t(value);
        } else if (n == "bucket_capacity_pct") {
            if (value < 1 || value > 1000)
                throw InvalidArg{"bucket_capacity_pct must be in [1, 1000]"};
            ctx->bucket_capacity_pct = uint32_t(value);
        } else if (n == "sort_fail_test") {
            if (value != 0 && value != 1) throw InvalidArg{"sort_fail_test must be 0 or 1"};
            ctx->sort_fail_test = value != 0;
        } else if (n == "fp32_rows") {
            if (value != 0 && value != 1) throw InvalidArg{"fp32_rows must be 0 or 1"};
            ctx->fp32_rows = value != 0;
        } else if (n == "autotune") {
            if (value != 0 && value != 1) throw InvalidArg{"autotune must be 0 or 1"};
            ctx->autotune = value != 0;
        } else if (n == "chunk_rows") {
            if (value < 0 || value > MAX_CHUNK_ROWS || value % WAVE)
                throw InvalidArg{"chunk_rows must be 0 or a multiple of 64 up to 512"};
            ctx->chunk_rows = uint32_t(value);
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
        else if (n == "sort_wide_words") *value = ctx->sort_wide_words ? 1 : 0;
        else if (n == "bucket_min_rows") *value = int64_t(ctx->bucket_min_rows);
        else if (n == "bucket_capacity_pct") *value = ctx->bucket_capacity_pct;
        else if (n == "sort_fail_test") *value = ctx->sort_fail_test ? 1 : 0;
        else if (n == "fp32_rows") *value = ctx->fp32_rows ? 1 : 0;
        else if (n == "chunk_rows") *value = ctx->chunk_rows ? ctx->chunk_rows : CHUNK_ROWS;
        else if (n == "autotune") *value = ctx->autotune ? 1 : 0;
        else throw InvalidArg{"unknown parameter '" + n + "'"};
    });
}

xm_status xm_ctx_tune_table(xm_ctx* ctx, xm_tune_entry* out, uint64_t cap, uint64_t* n) {
    if (!ctx || !n) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    *n = ctx->tune.size();
    for (uint64_t i = 0; out && i < cap && i < ctx->tune.size(); i++) {
        const auto& e = ctx->tune[i];
        out[i] = xm_tune_entry{e.first, e.second.chunk_rows, e.second.dense_min,
                               e.second.sample_jobs, e.second.tuned_ms, e.second.default_ms};
    }
    return XM_OK;
}

xm_status xm_ctx_release_memory(xm_ctx* ctx, uint64_t* cached_bytes) {
    if (!ctx) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        release_cache(ctx);
        if (cached_bytes) *cached_bytes = ctx->cached_bytes;
    });
}

xm_status xm_ctx_mem_info(xm_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!ctx) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        size_t f = 0, t = 0;
        XM_HIP(hipMemGetInfo(&f, &t));
        size_t cached_free = 0;
        for (auto& fb : ctx->free_blocks) cached_free += fb.first;
        if (free_bytes) *free_bytes = uint64_t(f) + cached_free;
        if (total_bytes) *total_bytes = uint64_t(t);
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

xm_status xm_host_alloc(xm_ctx* ctx, size_t bytes, void** hptr) {
    if (!ctx || !hptr) return XM_EINVAL;
    *hptr = nullptr;
    return guarded(ctx, [&] {
        const hipError_t e = hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            throw OutOfMemory{"pinned host allocation of " + std::to_string(bytes) + " bytes failed"};
        }
        XM_HIP(e);
    });
}

xm_status xm_host_free(xm_ctx* ctx, void* hptr) {
    if (!ctx) return XM_EINVAL;
    return guarded(ctx, [&] {
        if (hptr) XM_HIP(hipHostFree(hptr));
    });
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

static xm_status catalog_create_impl(xm_ctx* ctx, const xm_segment* segs, uint64_t n_segs,
                                     bool on_device, xm_catalog** out) {
    if (!ctx || !out || (n_segs && !segs)) return XM_EINVAL;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(ctx->mu);
    xm_catalog* cat = new (std::nothrow) xm_catalog();
    if (!cat) return XM_ENOMEM;
    xm_obj* staged = nullptr;
    xm_status st = guarded(ctx, [&] {
        if (n_segs == 0) throw InvalidArg{"a catalog needs at least one segment"};
        if (n_segs > MAX_SEGMENTS) throw InvalidArg{"more than 65536 segments in one catalog"};
        std::vector<const xm_obj*> recs(n_segs);
        std::vector<uint64_t> ns(n_segs);
        std::vector<double> hs(n_segs);
        uint64_t total = 0;
        for (uint64_t s = 0; s < n_segs; s++) {
            recs[s] = segs[s].records;
            ns[s] = segs[s].n;
            hs[s] = segs[s].zone_height_deg;
            if (ns[s] > SORT_MAX_KEYS) throw InvalidArg{"a catalog holds at most 2^32-8193 records"};
            total += ns[s];
            if (total > SORT_MAX_KEYS) throw InvalidArg{"a catalog holds at most 2^32-8193 records"};
        }
        if (!on_device && total) {
            // one staging buffer for all segments, one copy each
            staged = static_cast<xm_obj*>(dalloc(ctx, total * sizeof(xm_obj)));
            uint64_t off = 0;
            for (uint64_t s = 0; s < n_segs; s++) {
                if (ns[s] && !recs[s]) throw InvalidArg{"null records"};
                if (ns[s])
                    XM_HIP(hipMemcpyAsync(staged + off, recs[s], ns[s] * sizeof(xm_obj),
                                          hipMemcpyHostToDevice, ctx->stream));
                recs[s] = ns[s] ? staged + off : nullptr;
                off += ns[s];
            }
        }
        catalog_build(ctx, recs, ns, hs, cat);
        if (staged) { dfree(ctx, staged); staged = nullptr; }
        sync(ctx);
    });
    if (st != XM_OK) {
        (void)guarded(ctx, [&] {
            sync(ctx);
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
    const xm_segment sg{records, n, zone_height_deg};
    return catalog_create_impl(ctx, &sg, 1, false, out);
}

xm_status xm_catalog_create_device(xm_ctx* ctx, const xm_obj* d_records, uint64_t n,
                                   double zone_height_deg, xm_catalog** out) {
    const xm_segment sg{d_records, n, zone_height_deg};
    return catalog_create_impl(ctx, &sg, 1, true, out);
}

xm_status xm_catalog_create_segments(xm_ctx* ctx, const xm_segment* segs, uint64_t n_segs,
                                     xm_catalog** out) {
    return catalog_create_impl(ctx, segs, n_segs, false, out);
}

xm_status xm_catalog_create_segments_device(xm_ctx* ctx, const xm_segment* segs, uint64_t n_segs,
                                            xm_catalog** out) {
    return catalog_create_impl(ctx, segs, n_segs, true, out);
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
    if (!cat || cat->seg.empty()) return XM_EINVAL;
    int32_t lo = INT_MAX, hi = -1;
    for (const auto& d : cat->seg)
        if (d.n) { lo = std::min(lo, d.zmin); hi = std::max(hi, d.zmax); }
    if (hi < 0) lo = 0;
    if (n) *n = cat->n;
    if (zone_height_deg) *zone_height_deg = cat->seg[0].h_deg;
    if (zmin) *zmin = lo;
    if (zmax) *zmax = hi;
    return XM_OK;
}

xm_status xm_catalog_segment_count(const xm_catalog* cat, uint64_t* n_segs) {
    if (!cat || !n_segs) return XM_EINVAL;
    *n_segs = cat->seg.size();
    return XM_OK;
}

xm_status xm_catalog_segment_info(const xm_catalog* cat, uint64_t seg, uint64_t* first_row,
                                  uint64_t* n, double* zone_height_deg, int32_t* zmin,
                                  int32_t* zmax) {
    if (!cat || seg >= cat->seg.size()) return XM_EINVAL;
    const SegDesc& d = cat->seg[seg];
    if (first_row) *first_row = d.row_base;
    if (n) *n = d.n;
    if (zone_height_deg) *zone_height_deg = d.h_deg;
    if (zmin) *zmin = d.n ? d.zmin : 0;
    if (zmax) *zmax = d.n ? d.zmax : -1;
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
        std::vector<uint32_t> zoff;
        if (zone) {
            zoff.resize(cat->n_zoff);
            XM_HIP(hipMemcpyAsync(zoff.data(), cat->zoff, cat->n_zoff * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
        }
        sync(ctx);
        for (uint64_t i = 0; i < n; i++) {
            if (id) id[i] = rows[i].id;
            if (x) x[i] = rows[i].x;
            if (y) y[i] = rows[i].y;
            if (z) z[i] = rows[i].z;
        }
        // a segment's zone z holds rows [zoff[e], zoff[e + 1]), e = zoff_base + z - zmin
        if (zone)
            for (const SegDesc& d : cat->seg) {
                if (!d.n) continue;
                for (int32_t zz = d.zmin; zz <= d.zmax; zz++) {
                    const uint64_t e = d.zoff_base + uint64_t(zz - d.zmin);
                    for (uint64_t i = zoff[e]; i < zoff[e + 1]; i++) zone[i] = zz;
                }
            }
    });
}

}  // extern "C"

namespace {

// Above this much staging (24 bytes a slot) a join stages row indices (8 bytes a
// slot) and the compaction gathers the ids and distances from the rows.  Narrow
// slots cost two random 32-byte row reads per match in the compaction: on the
// 200M x 200M plane join at 1" (1.0e8 matches, 4e8 default slots = 9.6 GB wide)
// the compaction took 2.97 ms with narrow slots under an 8 GB limit.
constexpr uint64_t WIDE_STAGING_BYTES = uint64_t(64) << 30;

void free_slots(xm_ctx* ctx, RegionOut& ro) {
    dfree(ctx, ro.id_a); dfree(ctx, ro.id_b); dfree(ctx, ro.d2);
    dfree(ctx, ro.ra); dfree(ctx, ro.rb); dfree(ctx, ro.job);
    ro.id_a = nullptr; ro.id_b = nullptr; ro.d2 = nullptr;
    ro.ra = nullptr; ro.rb = nullptr; ro.job = nullptr;
}

// ---- autotuner: the pair kernel's shape per tuning key, by measurement.
// Candidate shapes: chunk rows x dense rule.  Each is timed on a sample of the
// batch's jobs of the key (up to TUNE_SAMPLE, the largest by nA * nB) as a
// count-only pair launch: regions and spill pools of zero slots, so every hit is
// counted and none is written.  The sample is repeated until the launch fills the
// GPU.  The fastest of three launches per shape counts,
// and a shape replaces the default only when it is faster by TUNE_MARGIN: on the
// 10,000-job batch (165 keys) the sample launches of small keys are mostly launch
// time, and choosing the fastest outright gave 249.6 ms against 247.0 ms with the
// default shape everywhere.
constexpr uint32_t TUNE_ROWS[] = {64, 128, 256, 512};
constexpr uint32_t TUNE_DENSE[] = {4, 16, 64};
constexpr size_t TUNE_SAMPLE = 16;
constexpr uint64_t TUNE_WAVES = 16384;
constexpr uint64_t TUNE_MAX_REPS = 256;
constexpr double TUNE_MARGIN = 0.97;
struct xm_tune_record {
    uint32_t rows, dense_min;
    double ms;
};

void tune_jobs(xm_ctx* ctx, xm_catalog* A, const xm_catalog* B, const xm_job* jobs,
               uint64_t n_jobs, bool f32) {
    // the jobs of every key the table lacks
    std::vector<std::pair<uint32_t, std::vector<uint64_t>>> todo;
    for (uint64_t j = 0; j < n_jobs; j++) {
        if (jobs[j].seg_a >= A->seg.size() || jobs[j].seg_b >= B->seg.size()) continue;
        const SegDesc& sa = A->seg[jobs[j].seg_a];
        const SegDesc& sb = B->seg[jobs[j].seg_b];
        if (!sa.n || !sb.n || sa.h_deg != sb.h_deg || !(jobs[j].radius_deg > 0.0)) continue;
        const uint32_t k = tune_key(sa, sb, jobs[j].radius_deg);
        bool known = false;
        for (const auto& e : ctx->tune) known = known || e.first == k;
        if (known) continue;
        bool found = false;
        for (auto& t : todo)
            if (t.first == k) { t.second.push_back(j); found = true; }
        if (!found) todo.push_back({k, {j}});
    }
    if (todo.empty()) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    XM_HIP(hipEventCreate(&e0));
    XM_HIP(hipEventCreate(&e1));
    ScopedBlocks tmp(ctx);
    const uint32_t P = SPILL_POOLS;
    uint64_t* counters = tmp.alloc<uint64_t>(size_t(MAX_REGIONS + P) * REGION_STRIDE * 8);
    uint64_t* zeros = tmp.alloc<uint64_t>(size_t(MAX_REGIONS + P) * 8);
    launch_fill_u64(ctx, zeros, MAX_REGIONS + P, 0);
    try {
        for (auto& t : todo) {
            std::vector<uint64_t>& js = t.second;
            auto work = [&](uint64_t j) {
                return double(A->seg[jobs[j].seg_a].n) * double(B->seg[jobs[j].seg_b].n);
            };
            std::stable_sort(js.begin(), js.end(),
                             [&](uint64_t x, uint64_t y) { return work(x) > work(y); });
            if (js.size() > TUNE_SAMPLE) js.resize(TUNE_SAMPLE);
            // the sample, repeated until one launch has TUNE_WAVES wavefronts (at the
            // default rows): the launch then fills the GPU as the batch does, and its
            // time measures a shape's throughput, not the latency of a few waves
            uint64_t waves = 0;
            for (uint64_t j : js)
                waves += seg_max_chunks(A->seg[jobs[j].seg_a], default_shape(ctx).chunk_rows);
            const uint64_t reps =
                std::max<uint64_t>(1, std::min<uint64_t>(TUNE_MAX_REPS, TUNE_WAVES / std::max<uint64_t>(waves, 1)));
            std::vector<xm_job> sample;
            for (uint64_t r = 0; r < reps; r++)
                for (uint64_t j : js) sample.push_back(jobs[j]);
            xm_tune_record best{};
            bool have = false;
            double default_ms = -1.0;
            for (uint32_t rows : TUNE_ROWS) {
                for (uint32_t dmin : TUNE_DENSE) {
                    JobTable jt;
                    make_jobs(ctx, A, B, sample.data(), sample.size(),
                              std::vector<JobShape>(sample.size(), JobShape{rows, dmin}), jt);
                    RegionOut ro{};
                    ro.n_regions = std::max<uint32_t>(
                        1, std::min<uint32_t>(MAX_REGIONS, pair_blocks(std::max<uint64_t>(jt.max_waves, 1))));
                    ro.counters = counters;
                    ro.base = zeros;
                    ro.cap = zeros;
                    ro.spill_cap = 0;
                    ro.wide = false;
                    float best_ms = 1e30f;
                    for (int rep = 0; rep < 3; rep++) {
                        launch_fill_u64(ctx, counters, size_t(ro.n_regions + P) * REGION_STRIDE, 0);
                        XM_HIP(hipEventRecord(e0, ctx->stream));
                        launch_pair(ctx, A, B, static_cast<const Chunk*>(A->chunks), jt.jobs, jt.prefix, jt.n_jobs, jt.max_waves, f32, ro);
                        XM_HIP(hipEventRecord(e1, ctx->stream));
                        XM_HIP(hipEventSynchronize(e1));
                        float ms = 0.f;
                        XM_HIP(hipEventElapsedTime(&ms, e0, e1));
                        best_ms = std::min(best_ms, ms);
                    }
                    dfree(ctx, jt.jobs);
                    dfree(ctx, jt.prefix);
                    const JobShape def = default_shape(ctx);
                    if (rows == def.chunk_rows && dmin == def.dense_min) default_ms = best_ms;
                    if (!have || best_ms < best.ms) {
                        best = xm_tune_record{rows, dmin, double(best_ms)};
                        have = true;
                    }
                }
            }
            const JobShape def = default_shape(ctx);
            if (default_ms >= 0.0 && best.ms > TUNE_MARGIN * default_ms)
                best = xm_tune_record{def.chunk_rows, def.dense_min, default_ms};
            xm_ctx::TuneEntry e{best.rows, best.dense_min, uint32_t(js.size()), best.ms,
                                default_ms};
            ctx->tune.push_back({t.first, e});
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    XM_HIP(hipEventDestroy(e0));
    XM_HIP(hipEventDestroy(e1));
}

// The join of a batch of jobs, with overflow-and-retry.  batch = keep the job
// column and the per-job counts in the result.
xm_status match_jobs(xm_ctx* ctx, const xm_catalog* a, const xm_catalog* b, const xm_job* jobs,
                     uint64_t n_jobs, xm_precision precision, uint64_t capacity_hint, bool batch,
                     xm_result** out, xm_match_stats* stats) {
    if (!ctx || !out || !a || !b || (n_jobs && !jobs)) return XM_EINVAL;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(ctx->mu);
    xm_result* res = new (std::nothrow) xm_result();
    if (!res) return XM_ENOMEM;
    res->ctx = ctx;
    RegionOut ro{};
    JobTable jt;
    uint64_t* scratch = nullptr;
    uint64_t* job_sums = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    xm_match_stats st_local;
    std::memset(&st_local, 0, sizeof(st_local));
    // views of catalogs of another context, and the blocks made for them
    xm_catalog a_view, b_view;
    std::vector<void*> local;
    xm_status st = guarded(ctx, [&] {
        if (a->ctx->device != ctx->device || b->ctx->device != ctx->device)
            throw InvalidArg{"catalog of another device"};
        if (precision != XM_FP64_PLAIN && precision != XM_FP32_FILTER_FP64_RECHECK)
            throw InvalidArg{"unknown precision mode"};
        const bool f32 = precision == XM_FP32_FILTER_FP64_RECHECK;
        xm_catalog* A = const_cast<xm_catalog*>(a);
        xm_catalog* Bm = const_cast<xm_catalog*>(b);
        // A catalog of another context on this device is read and never changed:
        // what the join caches in a catalog of its own context -- A's chunk table
        // (with its segment table's chunk rows) -- is made in this context's memory
        // for the join's duration.
        if (a->ctx != ctx) {
            a_view = *a;
            a_view.ctx = ctx;
            a_view.chunks = nullptr;
            a_view.chunk_start = nullptr;
            for (auto& d : a_view.seg) d.chunk_rows = 0;
            local.reserve(local.size() + 1);
            a_view.seg_dev = static_cast<SegDesc*>(dalloc(ctx, a_view.seg.size() * sizeof(SegDesc)));
            local.push_back(a_view.seg_dev);
            A = &a_view;
        }
        if (b->ctx != ctx) {
            b_view = *b;
            b_view.ctx = ctx;
            Bm = &b_view;
        }
        const xm_catalog* B = Bm;
        if (ctx->autotune) tune_jobs(ctx, A, B, jobs, n_jobs, f32);
        std::vector<JobShape> shape(n_jobs);
        for (uint64_t j = 0; j < n_jobs; j++) {
            if (jobs[j].seg_a >= A->seg.size() || jobs[j].seg_b >= B->seg.size())
                throw InvalidArg{"job " + std::to_string(j) + " names a segment the catalog lacks"};
            shape[j] = job_shape(ctx, A->seg[jobs[j].seg_a], B->seg[jobs[j].seg_b],
                                 jobs[j].radius_deg);
        }
        make_jobs(ctx, A, B, jobs, n_jobs, shape, jt);
        uint64_t cap = capacity_hint;
        if (!cap) {
            // the reference's default, 2 * max(nA, nB) per job (Worker.cu:295-300)
            for (uint64_t j = 0; j < n_jobs; j++)
                cap += 2 * std::max<uint64_t>(A->seg[jobs[j].seg_a].n, B->seg[jobs[j].seg_b].n);
        }
        // a default that free memory cannot hold is cut to a third of it (the
        // context's cached blocks count as free; slots narrow past the wide limit);
        // a join that needs more re-runs once at its exact size
        // free device memory, the context's cached blocks counted free
        uint64_t avail = ~uint64_t(0);
        {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
                avail = free_b;
                for (auto& fb : ctx->free_blocks) avail += fb.first;
            } else {
                (void)hipGetLastError();
            }
        }
        // wide slots while their staging stays under WIDE_STAGING_BYTES and a quarter
        // of the free memory: the staging and the result are held together while
        // the compaction runs (10,000 cap joins of 2.2e9 sources, FP32 form: 2.4e9
        // default slots would take 67 GB wide against 29 GB narrow)
        auto wide_fits = [&](uint64_t slots) {
            return slots * 24 <= WIDE_STAGING_BYTES && slots * 24 <= avail / 4;
        };
        if (!capacity_hint && avail != ~uint64_t(0)) {
            const uint64_t job_b = batch ? 4 : 0;
            const uint64_t slot_b = (wide_fits(cap) ? 24 : 8) + job_b;
            cap = std::min<uint64_t>(cap, avail / 3 / slot_b);
        }
        cap = std::max<uint64_t>(cap, 1024);
        for (auto& e : ev) XM_HIP(hipEventCreate(&e));
        // Hits go to n_regions append regions as (A row, B row[, job]); wavefronts
        // spread over them round-robin.  The first run gives the regions 3/4 of the slots
        // and keeps 1/4 as 64 spill pools (region r spills into pool r % 64, each
        // pool with its own counter line), so uneven regions do not overflow while
        // the total fits.  A one-block scan sizes the parts; one host round trip
        // reads the total; k_compact then gathers ids and distances into a result
        // of exactly the total.  If a pool overflowed, the re-run gives every region
        // exactly its count from the first run (the wave -> region assignment and
        // each wave's hits do not change), so it cannot overflow.
        ro.n_regions = std::max<uint32_t>(
            1, std::min<uint32_t>(MAX_REGIONS, pair_blocks(std::max<uint64_t>(jt.max_waves, 1))));
        const uint32_t R = ro.n_regions;
        const uint32_t P = SPILL_POOLS;
        ro.counters = static_cast<uint64_t*>(
            dalloc(ctx, size_t(R + P) * REGION_STRIDE * sizeof(uint64_t)));
        // scratch: prefix[R + P + 1], tail[4], base[R + P], cap[R], next_base[R + P], next_cap[R]
        scratch = static_cast<uint64_t*>(dalloc(ctx, (size_t(R) * 5 + 3 * P + 8) * sizeof(uint64_t)));
        uint64_t* prefix = scratch;
        uint64_t* tail = prefix + R + P + 1;
        uint64_t* base = tail + 4;
        uint64_t* capa = base + R + P;
        uint64_t* next_base = capa + R;
        uint64_t* next_cap = next_base + R + P;
        if (batch) {
            ro.n_jobs = uint32_t(n_jobs);
            ro.job_count = static_cast<unsigned long long*>(dalloc(ctx, size_t(JOB_SPREAD) * n_jobs * 8));
            job_sums = static_cast<uint64_t*>(dalloc(ctx, n_jobs * 8));
        }
        const uint64_t cap_each = std::max<uint64_t>(1, cap * 3 / 4 / R);
        ro.spill_cap = std::max<uint64_t>(1, (cap - std::min(cap, cap_each * R)) / P);
        launch_region_init(ctx, R, cap_each, ro.spill_cap, base, capa);
        uint64_t slots = cap_each * R + ro.spill_cap * P;
        uint64_t* pin = static_cast<uint64_t*>(ctx->pinned);
        uint64_t total = 0;
        for (int attempt = 0;; attempt++) {
            ro.base = base;
            ro.cap = capa;
            // wide slots unless the staging would pass WIDE_STAGING_BYTES
            ro.wide = wide_fits(slots);
            if (ro.wide) {
                ro.id_a = static_cast<int64_t*>(dalloc(ctx, slots * 8));
                ro.id_b = static_cast<int64_t*>(dalloc(ctx, slots * 8));
                ro.d2 = static_cast<double*>(dalloc(ctx, slots * 8));
            } else {
                ro.ra = static_cast<uint32_t*>(dalloc(ctx, slots * 4));
                ro.rb = static_cast<uint32_t*>(dalloc(ctx, slots * 4));
            }
            if (batch) {
                ro.job = static_cast<uint32_t*>(dalloc(ctx, slots * 4));
                launch_fill_u64(ctx, reinterpret_cast<uint64_t*>(ro.job_count),
                                size_t(JOB_SPREAD) * n_jobs, 0);
            }
            launch_fill_u64(ctx, ro.counters, size_t(R + P) * REGION_STRIDE, 0);
            XM_HIP(hipEventRecord(ev[0], ctx->stream));
            launch_pair(ctx, A, B, static_cast<const Chunk*>(A->chunks), jt.jobs, jt.prefix, jt.n_jobs, jt.max_waves, f32, ro);
            launch_region_scan(ctx, ro, prefix, tail, next_base, next_cap);
            XM_HIP(hipEventRecord(ev[1], ctx->stream));
            XM_HIP(hipMemcpyAsync(pin, tail, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  ctx->stream));
            XM_HIP(hipMemcpyAsync(pin + 4, jt.prefix + jt.n_jobs, sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
            sync(ctx);
            float ms = 0.f;
            XM_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            st_local.join_ms += ms;
            total = pin[0];
            const uint64_t spill = pin[1];
            st_local.n_chunks = reinterpret_cast<uint32_t*>(pin + 4)[0];
            if (spill <= ro.spill_cap) {
                st_local.band_hits = pin[2];
                double err = 0.0;
                std::memcpy(&err, &pin[3], sizeof(double));
                st_local.max_dist_err_arcsec = err * RAD2DEG * 3600.0;
                break;
            }
            // the pool overflowed: re-run with every region sized to its count
            st_local.n_retries++;
            if (attempt > 0) throw std::runtime_error("result size did not converge");
            free_slots(ctx, ro);
            std::swap(base, next_base);
            std::swap(capa, next_cap);
            slots = std::max<uint64_t>(total, 1);
            ro.spill_cap = 0;
        }
        const uint64_t slots_out = std::max<uint64_t>(total, 1);
        res->id_a = static_cast<int64_t*>(dalloc(ctx, slots_out * 8));
        res->id_b = static_cast<int64_t*>(dalloc(ctx, slots_out * 8));
        res->dist = static_cast<double*>(dalloc(ctx, slots_out * 8));
        if (batch) res->job = static_cast<uint32_t*>(dalloc(ctx, slots_out * 4));
        res->cap = slots_out;
        res->n = total;
        XM_HIP(hipEventRecord(ev[2], ctx->stream));
        launch_compact(ctx, ro, prefix, total, A, B, res->id_a, res->id_b, res->dist, res->job);
        XM_HIP(hipEventRecord(ev[3], ctx->stream));
        if (batch) {
            launch_job_sum(ctx, ro.job_count, uint32_t(n_jobs), job_sums);
            res->job_counts.resize(n_jobs);
            XM_HIP(hipMemcpyAsync(res->job_counts.data(), job_sums, n_jobs * 8,
                                  hipMemcpyDeviceToHost, ctx->stream));
        }
        free_slots(ctx, ro);
        sync(ctx);
        float ms = 0.f;
        XM_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
        st_local.join_ms += ms;
        st_local.n_matches = res->n;
    });
    (void)guarded(ctx, [&] {
        sync(ctx);
        for (void* p : local) dfree(ctx, p);
        dfree(ctx, a_view.chunks);
        dfree(ctx, a_view.chunk_start);
        dfree(ctx, ro.counters);
        dfree(ctx, scratch);
        free_slots(ctx, ro);
        dfree(ctx, reinterpret_cast<void*>(ro.job_count));
        dfree(ctx, job_sums);
        dfree(ctx, jt.jobs);
        dfree(ctx, jt.prefix);
        for (auto e : ev)
            if (e) XM_HIP(hipEventDestroy(e));
    });
    if (stats) *stats = st_local;
    if (st != XM_OK) {
        (void)guarded(ctx, [&] {
            dfree(ctx, res->id_a); dfree(ctx, res->id_b); dfree(ctx, res->dist);
            dfree(ctx, res->job);
        });
        delete res;
        return st;
    }
    *out = res;
    return XM_OK;
}

}  // namespace

extern "C" {

xm_status xm_match(xm_ctx* ctx, const xm_catalog* a, const xm_catalog* b, double radius_deg,
                   xm_precision precision, uint64_t capacity_hint, xm_result** out,
                   xm_match_stats* stats) {
    if (!ctx || !out) return XM_EINVAL;
    *out = nullptr;
    if (!a || !b) {
        set_error(ctx, "null catalog");
        return XM_EINVAL;
    }
    if (a->seg.size() != 1 || b->seg.size() != 1) {
        set_error(ctx, "xm_match joins one-segment catalogs; use xm_match_batch");
        return XM_EINVAL;
    }
    const xm_job job{0, 0, radius_deg};
    return match_jobs(ctx, a, b, &job, 1, precision, capacity_hint, false, out, stats);
}

xm_status xm_match_batch(xm_ctx* ctx, const xm_catalog* a, const xm_catalog* b, const xm_job* jobs,
                         uint64_t n_jobs, xm_precision precision, uint64_t capacity_hint,
                         xm_result** out, xm_match_stats* stats) {
    return match_jobs(ctx, a, b, jobs, n_jobs, precision, capacity_hint, true, out, stats);
}

xm_status xm_result_jobs(xm_ctx* ctx, const xm_result* res, uint32_t* job, uint64_t* job_counts) {
    if (!ctx || !res) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        if (!res->job && res->job_counts.empty())
            throw InvalidArg{"not the result of a batch join"};
        if (job && res->n) {
            XM_HIP(hipMemcpyAsync(job, res->job, res->n * 4, hipMemcpyDeviceToHost, ctx->stream));
            sync(ctx);
        }
        if (job_counts)
            std::memcpy(job_counts, res->job_counts.data(), res->job_counts.size() * 8);
    });
}

xm_status xm_result_job_device_ptr(const xm_result* res, const uint32_t** job) {
    if (!res || !job) return XM_EINVAL;
    *job = res->job;
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
        if (res->ctx != ctx) throw InvalidArg{"result belongs to another context"};
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
            dfree(res->ctx, res->job);
        });
    }
    delete res;
}

}  // extern "C"
