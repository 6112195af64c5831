// xm_internal.h -- library-internal types shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xmatch.h"
#include "xm_geom.h"

namespace xm {

// Result of a failed HIP call, carried up to the C ABI as XM_EDEVICE.
struct DeviceError {
    hipError_t code;
    std::string what;
};
struct InvalidArg {
    std::string what;
};
struct OutOfMemory {
    std::string what;
};

#define XM_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::xm::DeviceError{e_, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// Per-launch timing records (HIP events on the context stream).
struct TimedLaunch {
    const char* name;
    hipEvent_t start, stop;
};

// The device address of the allocation-free join's scratch array on device
// `device`, which must be current (hipGetSymbolAddress, once per device).
void* noalloc_scratch(int device);

// Records msg as the calling thread's last error on ctx (null: with no context);
// xm_last_error_message returns it.  Per thread, so that threads sharing a
// context never read or overwrite each other's messages.
void set_error(const xm_ctx* ctx, const std::string& msg);

}  // namespace xm

struct xm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    bool timing = false;
    std::string timing_only;                 // time only this kernel (empty: all)
    std::vector<xm::TimedLaunch> launches;   // pending, resolved on query
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<std::string, std::pair<double, uint64_t>>> kernel_totals;
    // pinned host scratch for small device->host reads
    void* pinned = nullptr;
    // device address of the allocation-free join's scratch array on this device
    void* noalloc_scratch = nullptr;
    // radix sort error word (device), set by a look-back that hit its polling limit
    uint32_t* sort_err = nullptr;
    // device memory cache: blocks keyed by size class; reused in stream order
    std::vector<std::pair<size_t, void*>> free_blocks;          // (class bytes, ptr)
    std::vector<std::pair<void*, size_t>> live_blocks;          // ptr -> class bytes
    size_t cached_bytes = 0;
    // tuning knobs (xm_ctx_set_param)
    uint32_t dense_min = 16;
    int key_ra_bits = 24;  // upper bound on the ra bits of a catalog's sort key
    bool sort_wide_words = false;  // 64-bit look-back words at every sort size
    bool sort_fail_test = false;   // sorts fail by their look-back limit (error-path tests)
    bool fp32_rows = false;        // catalog builds also write the FP32 rows (gather pass)
    uint32_t chunk_rows = 0;       // A rows per pair-kernel chunk (0: CHUNK_ROWS)
    uint64_t bucket_min_rows = uint64_t(1) << 24;  // one-segment builds of at least this
                                   // many rows write their records into zone buckets
    uint32_t bucket_capacity_pct = 115;  // bucket runs hold this % of the sampled estimate
    bool autotune = false;         // joins pick their shapes by measurement (tune table)
    // tune table: key (tune_key) -> measured shape, its time and the default's
    struct TuneEntry {
        uint32_t chunk_rows, dense_min, sample_jobs;
        double tuned_ms, default_ms;
    };
    std::vector<std::pair<uint32_t, TuneEntry>> tune;
};

namespace xm {

// One segment of a catalog: a contiguous run of rows, zone-sorted on its own and
// indexed by its own zone table.  A catalog made by xm_catalog_create has one
// segment; xm_catalog_create_segments puts many in one set of arrays, built by one
// sort, so that a batch of joins runs as one launch.
struct SegDesc {
    const xm_obj* recs;   // records during the build (device), null afterwards
    uint32_t row_base;    // first row in the catalog arrays
    uint32_t n;           // rows
    int32_t zmin, zmax;   // occupied zone range (zmin > zmax when empty)
    uint32_t zoff_base;   // zoff index of zone zmin; entries zoff_base .. + (zmax-zmin+1)
    int32_t n_zones;      // ceil(180 / h)
    double h_deg;         // zone height, degrees
    int64_t ra_gran;      // ra granularity of the keys, in 2^-24 turn quanta
    // build word layout: zone - zone_base in zone_bits above ra_bits above idx_bits
    int32_t zone_base;
    int32_t zone_bits, ra_bits, idx_bits;
    uint32_t chunk_rows;  // A side: rows per chunk of the current chunk table (0: none yet)
    uint32_t pad_;
};

}  // namespace xm

struct xm_catalog {
    xm_ctx* ctx = nullptr;
    uint64_t n = 0;                  // rows of all segments
    std::vector<xm::SegDesc> seg;    // host copy of the segment table
    xm::SegDesc* seg_dev = nullptr;  // device copy
    uint32_t* key = nullptr;         // ra quantum of each row (2^-24 turn), sorted order; the
                                     // row's segment and zone come from the zone table
    void* row = nullptr;             // xm::Row64[n]: FP64 unit vector + id, sorted order
    void* row32 = nullptr;           // xm::Row32[n]: FP32 unit vectors, written by the gather
                                     // when the context's fp32_rows is set (FP32 joins of two
                                     // such catalogs read them; otherwise they round the FP64
                                     // rows in registers)
    uint32_t* zoff = nullptr;        // [n_zoff]: first row of each segment zone, then n
    uint64_t n_zoff = 0;             // sum over segments of (zmax - zmin + 1), plus 1
    // A-side row chunks (built on first use as the A catalog of a join): the
    // table is sized for max_chunks, the true count stays on the device
    void* chunks = nullptr;
    uint64_t max_chunks = 0;
    uint32_t* chunk_start = nullptr;   // [n_zoff]: exclusive scan of chunks per zoff entry
};

struct xm_result {
    xm_ctx* ctx = nullptr;
    uint64_t n = 0;
    uint64_t cap = 0;
    int64_t* id_a = nullptr;
    int64_t* id_b = nullptr;
    double* dist = nullptr;
    uint32_t* job = nullptr;            // batch joins: job index of every pair
    std::vector<uint64_t> job_counts;   // batch joins: pairs per job
};

namespace xm {

constexpr int WAVE = 64;
// A rows per wavefront in the pair kernel: a chunk is a run of at most chunk_rows
// consecutive rows of one zone, chunk_rows a multiple of 64 up to MAX_CHUNK_ROWS,
// chosen per A segment (default CHUNK_ROWS).  A zone of k chunks scans its
// neighbour B zones k times: the 200M x 200M plane join at 1" (309 rows per zone on
// average, up to ~505) took 7.09 ms in the pair kernel with 256-row chunks and
// 5.93 ms with 512 (FP64, same box).
constexpr int CHUNK_ROWS = 512;
// the pair kernel stages a chunk's ra quanta in LDS (4 B a row per wavefront):
// 1024-row chunks tied with 512 on every workload measured, and their 14 KB of
// LDS per block held the kernel to 6 wavefronts per SIMD against 7
constexpr int MAX_CHUNK_ROWS = 512;

// One catalog row on the device: FP64 unit vector and id in one 32-byte line, so
// a candidate B row is a single aligned fetch.
struct Row64 {
    double x, y, z;
    int64_t id;
};

// FP32 unit vector of a row, for the FP32 filter (16 bytes, one aligned fetch).
struct Row32 {
    float x, y, z;
    uint32_t pad;
};

// A run of at most chunk_rows consecutive A rows of one zone of one segment.
struct Chunk {
    uint32_t start;
    uint32_t count;
    int32_t zone;
    int32_t pad;
};

// ---- memory: the context's cache of device blocks.  Every kernel of a context
// runs on its one stream, so a block freed after its last use is enqueued can be
// handed to later work on that stream at once.  Sizes are rounded up to a class
// (4 classes per power of two, at most 25% slack); hipMalloc is called only when
// no cached block of the class is free.  xm_ctx_release_memory returns the cache.
void* dalloc(xm_ctx* ctx, size_t bytes);
void dfree(xm_ctx* ctx, void* p);
void release_cache(xm_ctx* ctx);

// Device blocks a call takes for its own use.  They go back to the context's cache
// when the object goes out of scope, also when an exception ends the call early
// (kernels already enqueued on the stream still run first, as for any freed
// block).  keep() hands a block on to an owner that outlives the call.
class ScopedBlocks {
  public:
    explicit ScopedBlocks(xm_ctx* ctx) : ctx_(ctx) {}
    ScopedBlocks(const ScopedBlocks&) = delete;
    ScopedBlocks& operator=(const ScopedBlocks&) = delete;
    ~ScopedBlocks() {
        for (void* p : held_) {
            try {
                dfree(ctx_, p);
            } catch (...) {
            }
        }
    }
    template <typename T>
    T* alloc(size_t bytes) {
        held_.reserve(held_.size() + 1);  // the push below cannot throw once dalloc has
        void* p = dalloc(ctx_, bytes);
        held_.push_back(p);
        return static_cast<T*>(p);
    }
    void free(void* p) {
        forget(p);
        dfree(ctx_, p);
    }
    void keep(void* p) { forget(p); }

  private:
    void forget(void* p) {
        for (size_t i = 0; i < held_.size(); i++) {
            if (held_[i] == p) {
                held_[i] = held_.back();
                held_.pop_back();
                return;
            }
        }
        throw std::logic_error("a device block the call does not hold");
    }
    xm_ctx* ctx_;
    std::vector<void*> held_;
};

// Runs f with the context's device current and turns the exceptions of the
// library into a status (its message recorded by set_error).
template <typename F>
xm_status guarded(xm_ctx* ctx, F&& f) {
    try {
        if (ctx) XM_HIP(hipSetDevice(ctx->device));
        f();
        return XM_OK;
    } catch (const InvalidArg& e) {
        set_error(ctx, e.what);
        return XM_EINVAL;
    } catch (const OutOfMemory& e) {
        set_error(ctx, e.what);
        return XM_ENOMEM;
    } catch (const DeviceError& e) {
        set_error(ctx, e.what);
        return e.code == hipErrorOutOfMemory ? XM_ENOMEM : XM_EDEVICE;
    } catch (const std::bad_alloc&) {
        set_error(ctx, "host allocation failed");
        return XM_ENOMEM;
    } catch (const std::exception& e) {
        set_error(ctx, e.what());
        return XM_EINTERNAL;
    }
}

// ---- launch bookkeeping: events around a kernel when timing is on
void timing_begin(xm_ctx* ctx, const char* name, hipEvent_t* ev);
void timing_end(xm_ctx* ctx, const char* name, hipEvent_t ev);

template <typename F>
inline void launch_timed(xm_ctx* ctx, const char* name, F&& f) {
    hipEvent_t ev = nullptr;
    const bool timed = ctx->timing && (ctx->timing_only.empty() || ctx->timing_only == name);
    if (timed) timing_begin(ctx, name, &ev);
    f();
    XM_HIP(hipGetLastError());
    if (timed) timing_end(ctx, name, ev);
}

// ---- scan / sort (xm_sort.hip)
// Exclusive prefix sum of n uint32 values in place.
void scan_exclusive_u32(xm_ctx* ctx, uint32_t* data, uint64_t n);
// The same with caller scratch for the tile sums: scan_partial_words(n) words.
constexpr uint64_t SCAN_TILE = 4096;
inline uint64_t scan_partial_words(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }
void scan_exclusive_u32_with(xm_ctx* ctx, uint32_t* data, uint64_t n, uint32_t* partial);
// Stable LSD radix sort of 64-bit keys on bits [lo_bit, hi_bit), in digits of
// digit_bits (8 or 9) bits from lo_bit up; bits from hi_bit up must be equal in all
// keys.  Buffers are ping-ponged; on return *keys points at the sorted keys and
// *keys_alt at scratch.  digit_totals, when given, holds the counts of every digit
// value at digit positions lo_bit + digit_bits p (2^digit_bits per position, at
// least the sort's passes), as k_make_keys counts them while it writes the keys;
// else the sort reads the keys once to count them.
void radix_sort_keys(xm_ctx* ctx, uint64_t** keys, uint64_t** keys_alt, uint64_t n, int lo_bit,
                     int hi_bit, const uint32_t* digit_totals = nullptr, int digit_bits = 8);
// digit positions of a 64-bit key in 8-bit digits
constexpr int SORT_MAX_PASSES = 8;
// 9-bit digits serve keys of at most this many passes: they save a pass over 8-bit
// ones for keys of 33..36 sorted bits, and 4 x 512 digit counters take no more
// LDS than 8 x 256 (k_make_keys keeps four blocks a CU)
constexpr int SORT9_MAX_PASSES = 4;
// Many-segment sorts take 9-bit digits up to this many passes (63 bits) where they
// save a pass: k_seg_words keeps one histogram copy, whose 7 x 512 counters fit its
// LDS at full occupancy.
constexpr int SEG9_MAX_PASSES = 7;
// digit counters of one histogram copy for every digit layout a sort takes
constexpr int SORT_HIST_BINS = SORT_MAX_PASSES * 256 > SEG9_MAX_PASSES * 512
                                   ? SORT_MAX_PASSES * 256
                                   : SEG9_MAX_PASSES * 512;
// Sort passes of a key of `bits` sorted bits in digits of digit_bits.
XM_HD int sort_passes(int bits, int digit_bits) { return (bits + digit_bits - 1) / digit_bits; }
// Largest key count the radix sort takes (32-bit tile indices: the last 8192-key
// tile must end below 2^32), and so the largest catalog.
constexpr uint64_t SORT_MAX_KEYS = 0xffffffffull - 8192;

// Adds one key's digits of DB bits at positions lo_bit + DB p (p < n_pass) to the
// LDS histogram h[p][2^DB]; the active lanes of a wavefront call it together.  A
// digit that all 64 lanes hold (segment and zone bits of ordered input, the
// constant top digits) is added once, by lane 0: 64 atomics on one LDS word would
// run one after another.
template <int DB>
__device__ inline void count_digits(uint32_t* h, uint64_t key, bool valid, int lo_bit,
                                    int n_pass) {
    constexpr uint32_t R = 1u << DB;
    const bool full = __ballot(valid) == ~uint64_t(0);
    for (int p = 0; p < n_pass; p++) {
        const uint32_t d = uint32_t(key >> (lo_bit + DB * p)) & (R - 1);
        if (full && __all(d == uint32_t(__builtin_amdgcn_readfirstlane(int(d))))) {
            if ((threadIdx.x & (WAVE - 1)) == 0) atomicAdd(&h[p * R + d], uint32_t(WAVE));
        } else if (valid) {
            atomicAdd(&h[p * R + d], 1u);
        }
    }
}

// The context's sort error word (device, allocated and cleared on first use): set
// by a look-back that reached its polling limit; later passes and the gathers do
// nothing while it is set.
uint32_t* sort_error_word(xm_ctx* ctx);
// Reads the sort error word (synchronises the stream); throws if a sort since the
// last check failed.
void sort_check(xm_ctx* ctx);
// byte offset of the sort error word's host copy in the context's pinned scratch
constexpr size_t PIN_SORT_ERR_OFFSET = 1024;
// byte offset of the stage for small job tables (and other small uploads)
constexpr size_t PIN_JOB_OFFSET = 4096;

// ---- catalog kernels (xm_catalog.hip)
// One-segment build: packed sort words (pack_word) of the records, their zone
// range and bad count, in one pass.
// With n_pass > 0 it also adds the counts of every digit value at digit positions
// lo_bit + Dp (p < n_pass) of the valid records' words into digit_totals[p][2^D],
// D = 8.  With layout (the device word launch_layout writes) the digit width D and
// the dropped ra bits come from it: the digit positions start that many bits
// higher and cover only the passes the sort then needs (at most n_pass).
void launch_make_keys(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg, int32_t n_zones,
                      PackLayout pl, uint64_t* words, int32_t* zminmax, uint32_t* bad,
                      int lo_bit = 0, int n_pass = 0, uint32_t* digit_totals = nullptr,
                      const uint32_t* layout = nullptr);
// One-segment key layout: from a strided sample of the records, the number of low
// ra bits the sort and the catalog key drop and the sort's digit width, written to
// *layout_out as drop | digit_bits << 8 (device; stats is scratch of
// LAYOUT_STATS_WORDS words, set up by launch_build_init).  zone_bits: the sort word's
// zone field.  TPB must be 256 (one thread per bin).
constexpr int LAYOUT_STATS_WORDS = 516;
void launch_layout(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg, int32_t n_zones,
                   int ra_bits, int zone_bits, uint32_t* stats, uint32_t* layout_out);
XM_HD int layout_drop(uint32_t layout) { return int(layout & 255u); }
XM_HD int layout_digit_bits(uint32_t layout) { return (layout >> 8) == 9u ? 9 : 8; }
// The one-segment build's scratch, BUILD_SMALL_WORDS + LAYOUT_STATS_WORDS +
// n_digit_words words, set up by one launch: small words {zone min = INT_MAX, zone
// max = -1, bad records = 0, layout = 0}, the layout statistics, the sort's digit
// counters (zero).
constexpr int BUILD_SMALL_WORDS = 4;
void launch_build_init(xm_ctx* ctx, uint32_t* scratch, uint64_t n_digit_words);
// Many-segment build, pass 1: zone range and bad count of every segment
// (range[3 s] = zmin, range[3 s + 1] = zmax, range[3 s + 2] = bad).
void launch_seg_range(xm_ctx* ctx, const SegDesc* seg, uint32_t n_seg, uint64_t n, int32_t* range);
// Pass 2: words = s << seg_shift | (zone - zone_base) | ra | row within the segment,
// in the segment's layout; with n_pass > 0 it also adds the counts of every digit
// value (digit_bits 8 or 9) at positions lo_bit + digit_bits p (p < n_pass) into
// digit_totals[p][2^digit_bits].
void launch_seg_words(xm_ctx* ctx, const SegDesc* seg, uint32_t n_seg, uint64_t n, int seg_shift,
                      uint64_t* words, int lo_bit = 0, int n_pass = 0,
                      uint32_t* digit_totals = nullptr, int digit_bits = 8);
// From the sorted words: rows[i] (unit vector + id of the input row the word
// carries), rows32[i] when rows32 is not null (the FP32 unit vector, for the FP32
// filter), keys[i] (the row's ra quantum) and every segment's zone offsets
// zoff[zoff_base + z - zmin] (first row of zone z; the entry after a segment's
// last zone is the next segment's first row, zoff[n_zoff - 1] = n).  The
// one-segment form takes the layout by value; the many-segment form reads it
// from the segment table.
// With xcd_runs the blocks are dealt so that each XCD takes one contiguous run of
// rows (for records in zone buckets, whose reads then stay in that XCD's L2).
void launch_gather(xm_ctx* ctx, const xm_obj* recs, const uint64_t* words, uint64_t n,
                   PackLayout pl, int ra_drop, int32_t zmin, Row64* rows, Row32* rows32,
                   uint32_t* zoff, uint64_t n_zoff, uint32_t* keys, bool xcd_runs = false);

// ---- one-segment build with the records in zone buckets
// The key pass also moves every record into a bucket of consecutive zones (~4 MB of
// records each), one run per bucket and XCD group, so that the gather, whose blocks
// take one contiguous run of rows per XCD, reads each bucket from one L2 instead of
// whole random 128-B lines.  The runs' sizes come from a sample of the records; a
// record whose run is full goes to an overflow run; if that fills too, the plan has
// failed (plan word BP_FAILED) and the caller rebuilds without buckets.
constexpr uint32_t BUCKET_GROUPS = 8;       // bucket runs per bucket: one per XCD
constexpr uint32_t MAX_BUCKETS = 2048;
constexpr uint32_t BUCKET_SAMPLE = 1u << 20;
constexpr uint32_t BUCKET_SLACK = 64;       // rows added to every run's estimate
// largest bucketed catalog: run cursors (32-bit) stay below 2^32
constexpr uint64_t BUCKET_MAX_ROWS = uint64_t(1) << 30;
// plan words after the 2 x BUCKET_GROUPS x n_buckets run cursors and ends
constexpr int BP_OVF_CURSOR = 0, BP_FAILED = 1, BP_WORDS = 4;
struct BucketBuild {
    uint32_t n_buckets;
    uint64_t c_main_max;  // bound on the bucket runs' rows
    uint64_t c_total;     // rows of the bucketed record array (runs + overflow)
};
BucketBuild bucket_build(uint64_t n, uint32_t capacity_pct);
// plan: BUCKET_GROUPS x n_buckets cursors, as many run ends, BP_WORDS words; hist:
// n_buckets counters (zero).  Samples the records' zones (bucket edges from the
// layout statistics' zone range), sizes the runs and sets the cursors.
void launch_bucket_plan(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg,
                        int32_t n_zones, const uint32_t* stats, const BucketBuild& bb,
                        uint32_t capacity_pct, uint32_t* hist, uint32_t* plan);
// k_make_keys for bucketed records: the same words (input order, the row field =
// the record's place in part), digit counts, zone range and bad count.
void launch_make_keys_bucketed(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg,
                               int32_t n_zones, PackLayout pl, const uint32_t* stats,
                               const BucketBuild& bb, uint32_t* plan, xm_obj* part,
                               uint64_t* words, int32_t* zminmax, uint32_t* bad, int lo_bit,
                               int n_pass, uint32_t* digit_totals, const uint32_t* layout);
void launch_gather_seg(xm_ctx* ctx, const SegDesc* seg, int seg_shift, const uint64_t* words,
                       uint64_t n, Row64* rows, Row32* rows32, uint32_t* zoff, uint64_t n_zoff,
                       uint32_t* keys);
// counts[k] = chunks of zoff entry k (k < n_entries) at its segment's chunk_rows;
// chunks from the scanned counts.
void launch_chunk_counts(xm_ctx* ctx, const uint32_t* zoff, uint64_t n_entries,
                         const SegDesc* seg, uint32_t n_seg, uint32_t* counts);
void launch_chunk_fill(xm_ctx* ctx, const uint32_t* zoff, const uint32_t* cstart,
                       uint64_t n_entries, const SegDesc* seg, uint32_t n_seg, Chunk* chunks);
void launch_fill_u32(xm_ctx* ctx, uint32_t* p, uint64_t n, uint32_t v);
void launch_fill_u64(xm_ctx* ctx, uint64_t* p, uint64_t n, uint64_t v);
void launch_fill_i32(xm_ctx* ctx, int32_t* p, uint64_t n, int32_t v);

// ---- pair kernel (xm_pair.hip)
struct PairParams {
    double theta_deg;       // search radius, degrees
    double h_deg;           // zone height, degrees
    double atan_sin_theta;  // atan(sin(theta)), host-computed once
    double chord2;          // (2 sin(r/2))^2
    int64_t gran;           // coarser ra granularity of the two catalogs (quanta)
    float t_lo32, t_hi32;   // FP32 filter: (c - band)^2 rounded down, (c + band)^2 up
    int32_t nz;             // ceil(r / h)
    int32_t n_zones;
    uint32_t dense_min;     // chunk mode rule, see chunk_is_dense()
};
// Output of a join.  Hits are appended to n_regions regions; wavefront g appends
// to region g % n_regions through counters[r * REGION_STRIDE] (one 128-byte line
// per counter, so the waves of a join never queue on one address).  Region r owns
// slots [base[r], base[r] + cap[r]); a flush that runs past its region's end puts
// the rest in spill pool p = r % SPILL_POOLS, slots [base[n_regions + p], +
// spill_cap), through that pool's counter counters[(n_regions + p) * REGION_STRIDE].
// Counters count every hit, so when a pool overflows the caller re-runs with each
// region sized to its count.
constexpr uint32_t MAX_REGIONS = 1024;
constexpr uint32_t SPILL_POOLS = 64;
constexpr uint32_t REGION_STRIDE = 16;
constexpr uint32_t JOB_SPREAD = 64;  // partial hit counters per job, on separate lines
struct RegionOut {
    uint64_t* counters;      // [(n_regions + SPILL_POOLS) * REGION_STRIDE]
    uint32_t n_regions;
    const uint64_t* base;    // [n_regions + SPILL_POOLS]: first slot of each region, then pool
    const uint64_t* cap;     // [n_regions]: slots of each region
    uint64_t spill_cap;      // slots of each pool
    // wide slots hold (idA, idB, chord^2), written while the rows are in cache;
    // narrow slots hold (A row, B row), 8 bytes instead of 24, for very large outputs
    bool wide;
    int64_t* id_a;           // wide: [slots]
    int64_t* id_b;
    double* d2;
    uint32_t* ra;            // narrow: [slots]: A row of each hit
    uint32_t* rb;            // B row of each hit
    uint32_t* job;           // batch joins: job of each hit (else null)
    unsigned long long* job_count;  // batch joins: [JOB_SPREAD * n_jobs] partial counts
    uint32_t n_jobs;
};

// One join of a batch: A segment x B segment at one radius.  The host fills
// everything but a_chunk_base; launch_job_prep fills that and the wave counts.
struct JobDesc {
    PairParams pp;
    uint32_t a_zoff_lo;        // A segment's zoff entries [a_zoff_lo, a_zoff_hi)
    uint32_t a_zoff_hi;
    uint32_t b_zoff_base;      // zoff index of B zone b_zmin
    int32_t b_zmin, b_zmax;
    uint32_t a_chunk_base;     // first chunk of the A segment (device-filled)
};
// The pair kernel's shape for one job: A rows per chunk (a multiple of 64 up to
// MAX_CHUNK_ROWS) and the dense rule (PairParams::dense_min).
struct JobShape {
    uint32_t chunk_rows;
    uint32_t dense_min;
};
JobShape default_shape(const xm_ctx* ctx);
// The shape a join of segment a with segment b at radius_deg runs with: the tuned
// shape of its key when the context tunes and has measured that key, else the
// default shape.
JobShape job_shape(const xm_ctx* ctx, const SegDesc& a, const SegDesc& b, double radius_deg);
// Tuning key of a job: log2 rows per occupied zone of each segment (rounded,
// 0..31) and nz = ceil(r / h) (1..255): da | db << 8 | nz << 16.
uint32_t tune_key(const SegDesc& a, const SegDesc& b, double radius_deg);

// Join parameters of segment A with segment B at radius_deg (FP32 filter bounds
// included); check_radius rejects radii outside (0, 90) degrees, unequal zone
// heights and more than 1e6 zones of search.
PairParams make_params(const xm_ctx* ctx, const SegDesc& A, const SegDesc& B, double radius_deg);
void check_radius(double radius_deg, double h_a, double h_b);

// Per job: a_chunk_base = cstart[a_zoff_lo], waves[j] = its chunk count.
void launch_job_prep(xm_ctx* ctx, JobDesc* jobs, uint32_t n_jobs, const uint32_t* cstart,
                     uint32_t* waves);
uint32_t pair_blocks(uint64_t n_waves);
// One launch over every (job, A chunk): wave g belongs to the job j with
// job_prefix[j] <= g < job_prefix[j+1] (job_prefix[n_jobs] = total, on the
// device); the grid covers max_waves.  f32 = FP32 filter + FP64 recheck, on the
// FP32 rows when both catalogs carry them, else on the FP64 rows rounded to FP32 in
// registers (the same values); otherwise FP64 plain.
void launch_pair(xm_ctx* ctx, const xm_catalog* A, const xm_catalog* B, const Chunk* chunks,
                 const JobDesc* jobs, const uint32_t* job_prefix, uint32_t n_jobs,
                 uint64_t max_waves, bool f32, const RegionOut& ro);
// d[i] = 2 asin(sqrt(d[i]) / 2) in arcsec for i < n (chord^2 to distance, in place).
void launch_chord_to_arcsec(xm_ctx* ctx, double* d, uint64_t n);
// Regions of cap_each slots back to back, then SPILL_POOLS pools of pool_cap.
void launch_region_init(xm_ctx* ctx, uint32_t n_regions, uint64_t cap_each, uint64_t pool_cap,
                        uint64_t* base, uint64_t* cap);
// Parts = the regions, then the pools.  prefix[k] (k <= parts) = kept hits of the
// parts before k; tail[0] = all hits, tail[1] = the largest pool demand, tail[2] =
// FP32 band hits, tail[3] = largest |chord_32 - chord_64| (double bits);
// next_base / next_cap = regions sized to their counts and empty pools, for a
// re-run.
void launch_region_scan(xm_ctx* ctx, const RegionOut& ro, uint64_t* prefix, uint64_t* tail,
                        uint64_t* next_base, uint64_t* next_cap);
// job_counts[j] = sum of the JOB_SPREAD partial counts of job j.
void launch_job_sum(xm_ctx* ctx, const unsigned long long* partial, uint32_t n_jobs,
                    uint64_t* job_counts);
// Packs the regions' total hits into contiguous arrays (region r's hits at
// prefix[r]): ids from the rows, FP64 distance in arcsec; job (may be null) gets
// the job column of batch joins.
void launch_compact(xm_ctx* ctx, const RegionOut& ro, const uint64_t* prefix, uint64_t total,
                    const xm_catalog* A, const xm_catalog* B, int64_t* id_a, int64_t* id_b,
                    double* dist, uint32_t* job);

}  // namespace xm
