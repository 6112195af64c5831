// xm_noalloc.hip -- xm_match_noalloc: the one-segment join of xm_match written
// straight into caller-provided device arrays, with no device allocation.
//
// For callers that cannot allocate device memory: code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc.  The join's working memory
// -- the chunk table of A, the region counters and tables, the job table -- lives
// in one file-scope device array, `scratch`, of NOALLOC_SCRATCH_BYTES on every GPU
// the library runs on (held until the process ends).  Its device address on each
// GPU comes from hipGetSymbolAddress, called once per device when the first
// context of that device is created, with the device current; host code uses that
// address, never the array's name, for kernel arguments and fills.
//
// Two runs of the pair kernel.  The first counts every wavefront's hits into its
// region (regions of zero slots: nothing is written).  A one-block scan gives each
// region its exact count and offset.  The second run writes each region's hits
// from its offset on: a wavefront's region and its hits are the same in both runs,
// so every region is filled exactly and the output is contiguous -- (idA, idB,
// distance in arcsec) in the caller's arrays, in no particular order, as xm_match
// returns them.
//
// Every call on a GPU shares that GPU's array and leaves its contents behind; each
// run clears what must start at zero with fill kernels.  Calls that overlap on one
// GPU (two threads or two contexts of the same device) race on the array: the
// caller serialises them.  Launches of this entry point are not timed (timing
// events are not created here).
#include <cstring>

#include "xm_internal.h"

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte offsets of the join's arrays in the scratch, for A's zone and chunk counts.
struct Layout {
    size_t counters, zeros, prefix, tail, next_base, next_cap, base2, cap2, job, seg, jprefix,
        cstart, partial, chunks, total;
};

Layout layout(uint64_t n_zoff, uint64_t n_chunks) {
    using namespace xm;
    const size_t parts = MAX_REGIONS + SPILL_POOLS;
    Layout L;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at += align256(bytes);
        return o;
    };
    L.counters = take(parts * REGION_STRIDE * sizeof(uint64_t));
    L.zeros = take(parts * sizeof(uint64_t));
    L.prefix = take((parts + 1) * sizeof(uint64_t));
    L.tail = take(4 * sizeof(uint64_t));
    L.next_base = take(parts * sizeof(uint64_t));
    L.next_cap = take(MAX_REGIONS * sizeof(uint64_t));
    L.base2 = take(parts * sizeof(uint64_t));
    L.cap2 = take(MAX_REGIONS * sizeof(uint64_t));
    L.job = take(256);  // the job and segment descriptors are uploaded in one copy:
    L.seg = take(256);  // job at +0, segment at +256
    L.jprefix = take(2 * sizeof(uint32_t));
    L.cstart = take(n_zoff * sizeof(uint32_t));
    L.partial = take(scan_partial_words(n_zoff) * sizeof(uint32_t));
    L.chunks = take(n_chunks * sizeof(Chunk));
    L.total = at;
    return L;
}

// Timing off for the duration of a call (creating timing events is left to the
// allocating entry points).
struct NoTiming {
    xm_ctx* ctx;
    bool was;
    explicit NoTiming(xm_ctx* c) : ctx(c), was(c->timing) { c->timing = false; }
    ~NoTiming() { ctx->timing = was; }
};

}  // namespace

namespace xm {

// The device address of `scratch` on device `device` (which must be current),
// from hipGetSymbolAddress on the first call for that device.
void* noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        throw InvalidArg{"device index beyond the allocation-free scratch table"};
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        XM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw DeviceError{hipErrorInvalidSymbol, "hipGetSymbolAddress gave no address"};
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

}  // namespace xm

using namespace xm;

extern "C" xm_status xm_match_noalloc(xm_ctx* ctx, const xm_catalog* a, const xm_catalog* b,
                                      double radius_deg, xm_precision precision,
                                      int64_t* d_id_a, int64_t* d_id_b, double* d_dist_arcsec,
                                      uint64_t capacity, uint64_t* n_matches,
                                      xm_match_stats* stats) {
    if (!ctx || !n_matches) return XM_EINVAL;
    *n_matches = 0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        if (!a || !b) throw InvalidArg{"null catalog"};
        if (a->ctx != ctx || b->ctx != ctx) throw InvalidArg{"catalog belongs to another context"};
        if (a->seg.size() != 1 || b->seg.size() != 1)
            throw InvalidArg{"xm_match_noalloc joins one-segment catalogs"};
        if (precision != XM_FP64_PLAIN && precision != XM_FP32_FILTER_FP64_RECHECK)
            throw InvalidArg{"unknown precision mode"};
        const bool f32 = precision == XM_FP32_FILTER_FP64_RECHECK;
        if (capacity && (!d_id_a || !d_id_b || !d_dist_arcsec))
            throw InvalidArg{"null output arrays"};
        if (!ctx->noalloc_scratch) throw InvalidArg{"the context has no allocation-free scratch"};
        const SegDesc& sa = a->seg[0];
        const SegDesc& sb = b->seg[0];
        check_radius(radius_deg, sa.h_deg, sb.h_deg);
        if (!sa.n || !sb.n) return;
        const JobShape shape = job_shape(ctx, sa, sb, radius_deg);
        const uint64_t entries = a->n_zoff - 1;
        const uint64_t n_chunks =
            uint64_t(sa.n) / shape.chunk_rows + uint64_t(sa.zmax - sa.zmin + 1);
        const Layout L = layout(a->n_zoff, n_chunks);
        if (L.total > NOALLOC_SCRATCH_BYTES)
            throw InvalidArg{"the join needs " + std::to_string(L.total) +
                             " bytes of working memory (A: " + std::to_string(sa.n) + " rows in " +
                             std::to_string(entries) + " zones); xm_match_noalloc holds " +
                             std::to_string(NOALLOC_SCRATCH_BYTES)};
        NoTiming no_timing(ctx);
        char* base = static_cast<char*>(ctx->noalloc_scratch);
        auto at = [&](size_t off) { return static_cast<void*>(base + off); };
        uint64_t* counters = static_cast<uint64_t*>(at(L.counters));
        uint64_t* zeros = static_cast<uint64_t*>(at(L.zeros));
        uint64_t* prefix = static_cast<uint64_t*>(at(L.prefix));
        uint64_t* tail = static_cast<uint64_t*>(at(L.tail));
        uint64_t* next_base = static_cast<uint64_t*>(at(L.next_base));
        uint64_t* next_cap = static_cast<uint64_t*>(at(L.next_cap));
        uint64_t* base2 = static_cast<uint64_t*>(at(L.base2));
        uint64_t* cap2 = static_cast<uint64_t*>(at(L.cap2));
        JobDesc* job = static_cast<JobDesc*>(at(L.job));
        SegDesc* seg = static_cast<SegDesc*>(at(L.seg));
        uint32_t* jprefix = static_cast<uint32_t*>(at(L.jprefix));
        uint32_t* cstart = static_cast<uint32_t*>(at(L.cstart));
        uint32_t* partial = static_cast<uint32_t*>(at(L.partial));
        Chunk* chunks = static_cast<Chunk*>(at(L.chunks));

        // the job and the A segment (with this join's chunk rows), through the
        // context's pinned stage, in one copy
        static_assert(sizeof(JobDesc) <= 256 && sizeof(SegDesc) <= 256, "descriptor slots");
        JobDesc d;
        std::memset(&d, 0, sizeof(d));
        d.pp = make_params(ctx, sa, sb, radius_deg);
        d.pp.dense_min = shape.dense_min;
        d.a_zoff_lo = sa.zoff_base;
        d.a_zoff_hi = sa.zoff_base + uint32_t(sa.zmax - sa.zmin + 1);
        d.b_zoff_base = sb.zoff_base;
        d.b_zmin = sb.zmin;
        d.b_zmax = sb.zmax;
        d.a_chunk_base = 0;
        SegDesc sd = sa;
        sd.chunk_rows = shape.chunk_rows;
        char* stage = static_cast<char*>(ctx->pinned) + PIN_JOB_OFFSET;
        std::memcpy(stage, &d, sizeof(d));
        std::memcpy(stage + 256, &sd, sizeof(sd));
        XM_HIP(hipMemcpyAsync(job, stage, 512, hipMemcpyHostToDevice, ctx->stream));

        // A's chunk table and the job's wave count
        launch_fill_u32(ctx, cstart + entries, 1, 0u);
        launch_chunk_counts(ctx, a->zoff, entries, seg, 1, cstart);
        scan_exclusive_u32_with(ctx, cstart, a->n_zoff, partial);
        launch_chunk_fill(ctx, a->zoff, cstart, entries, seg, 1, chunks);
        launch_fill_u32(ctx, jprefix, 2, 0u);
        launch_job_prep(ctx, job, 1, cstart, jprefix);
        scan_exclusive_u32_with(ctx, jprefix, 2, nullptr);

        // run 1: count every region's hits
        RegionOut ro{};
        ro.n_regions = std::max<uint32_t>(1, std::min<uint32_t>(MAX_REGIONS, pair_blocks(n_chunks)));
        const uint32_t parts = ro.n_regions + SPILL_POOLS;
        ro.counters = counters;
        ro.base = zeros;
        ro.cap = zeros;
        ro.spill_cap = 0;
        ro.wide = true;
        launch_fill_u64(ctx, zeros, parts, 0);
        launch_fill_u64(ctx, counters, uint64_t(parts) * REGION_STRIDE, 0);
        launch_pair(ctx, a, b, chunks, job, jprefix, 1, n_chunks, f32, ro);
        launch_region_scan(ctx, ro, prefix, tail, next_base, next_cap);
        uint64_t* pin = static_cast<uint64_t*>(ctx->pinned);
        XM_HIP(hipMemcpyAsync(pin, tail, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
        XM_HIP(hipMemcpyAsync(pin + 4, jprefix + 1, sizeof(uint32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        XM_HIP(hipStreamSynchronize(ctx->stream));
        const uint64_t total = pin[0];
        const uint32_t waves = reinterpret_cast<uint32_t*>(pin + 4)[0];
        *n_matches = total;
        if (stats) {
            stats->n_matches = total;
            stats->n_chunks = waves;
        }
        if (total > capacity)
            throw OutOfMemory{"the join has " + std::to_string(total) +
                              " pairs; the output arrays hold " + std::to_string(capacity)};
        if (total == 0) return;

        // run 2: each region's hits from its offset on, straight into the outputs
        // (chord^2 in the distance array, turned into arcsec after the run)
        ro.base = next_base;
        ro.cap = next_cap;
        ro.id_a = d_id_a;
        ro.id_b = d_id_b;
        ro.d2 = d_dist_arcsec;
        launch_fill_u64(ctx, counters, uint64_t(parts) * REGION_STRIDE, 0);
        launch_pair(ctx, a, b, chunks, job, jprefix, 1, n_chunks, f32, ro);
        launch_region_scan(ctx, ro, prefix, tail, base2, cap2);
        launch_chord_to_arcsec(ctx, d_dist_arcsec, total);
        XM_HIP(hipMemcpyAsync(pin, tail, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
        XM_HIP(hipStreamSynchronize(ctx->stream));
        if (pin[0] != total || pin[1] != 0)
            throw std::runtime_error("the writing run found other counts than the counting run");
        if (stats) {
            stats->band_hits = pin[2];
            double err = 0.0;
            std::memcpy(&err, &pin[3], sizeof(double));
            stats->max_dist_err_arcsec = err * RAD2DEG * 3600.0;
        }
    });
}
