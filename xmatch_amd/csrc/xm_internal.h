// xm_internal.h -- library-internal types shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
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

}  // namespace xm

struct xm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::string last_error = "no error";
    bool timing = false;
    std::vector<xm::TimedLaunch> launches;   // pending, resolved on query
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<std::string, std::pair<double, uint64_t>>> kernel_totals;
    // pinned host scratch for small device->host reads
    void* pinned = nullptr;
    // device memory cache: blocks keyed by size class; reused in stream order
    std::vector<std::pair<size_t, void*>> free_blocks;          // (class bytes, ptr)
    std::vector<std::pair<void*, size_t>> live_blocks;          // ptr -> class bytes
    size_t cached_bytes = 0;
    // tuning knobs (xm_ctx_set_param)
    uint32_t dense_min = 16;
    int key_ra_bits = 24;  // upper bound on the ra bits of a catalog's sort key
};

struct xm_catalog {
    xm_ctx* ctx = nullptr;
    uint64_t n = 0;
    double h_deg = 0;
    int32_t n_zones = 0;
    int32_t zmin = 0, zmax = -1;
    int64_t ra_gran = 1;         // ra granularity of the keys, in 2^-24 turn quanta
    uint64_t* key = nullptr;     // sorted (zone << RA_BITS | ra quantum)
    void* row = nullptr;         // xm::Row64[n]: FP64 unit vector + id, sorted order
    void* row32 = nullptr;       // xm::Row32[n]: FP32 unit vectors, made on first FP32 join
    uint32_t* zoff = nullptr;    // [zmax-zmin+2]: first row of zones zmin..zmax+1
    // A-side row chunks (built on first use as the A catalog of a join): the
    // table is sized for max_chunks, the true count stays on the device
    void* chunks = nullptr;
    uint64_t max_chunks = 0;
    uint32_t* chunk_start = nullptr;   // [zmax-zmin+2]: exclusive scan of chunks per zone
    const uint32_t* n_chunks_dev = nullptr;  // = chunk_start + (zmax-zmin+1)
};

struct xm_result {
    xm_ctx* ctx = nullptr;
    uint64_t n = 0;
    uint64_t cap = 0;
    int64_t* id_a = nullptr;
    int64_t* id_b = nullptr;
    double* dist = nullptr;
};

namespace xm {

constexpr int WAVE = 64;
constexpr int CHUNK_ROWS = 64;   // A rows per wavefront in the pair kernel

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

// A run of at most CHUNK_ROWS consecutive A rows of one zone.  aq is the ra
// half-width of the zone's window in quanta (alpha_quanta), filled per join.
struct Chunk {
    uint32_t start;
    uint32_t count;
    int32_t zone;
    int32_t aq;
};

// ---- memory: the context's cache of device blocks.  Every kernel of a context
// runs on its one stream, so a block freed after its last use is enqueued can be
// handed to later work on that stream at once.  Sizes are rounded up to a class
// (4 classes per power of two, at most 25% slack); hipMalloc is called only when
// no cached block of the class is free.  xm_ctx_release_memory returns the cache.
void* dalloc(xm_ctx* ctx, size_t bytes);
void dfree(xm_ctx* ctx, void* p);
void release_cache(xm_ctx* ctx);

// ---- launch bookkeeping: events around a kernel when timing is on
void timing_begin(xm_ctx* ctx, const char* name, hipEvent_t* ev);
void timing_end(xm_ctx* ctx, const char* name, hipEvent_t ev);

template <typename F>
inline void launch_timed(xm_ctx* ctx, const char* name, F&& f) {
    hipEvent_t ev = nullptr;
    if (ctx->timing) timing_begin(ctx, name, &ev);
    f();
    XM_HIP(hipGetLastError());
    if (ctx->timing) timing_end(ctx, name, ev);
}

// ---- scan / sort (xm_sort.hip)
// Exclusive prefix sum of n uint32 values in place.
void scan_exclusive_u32(xm_ctx* ctx, uint32_t* data, uint64_t n);
// Stable LSD radix sort of 64-bit keys on bits [lo_bit, hi_bit), 8-bit digits,
// skipping the digits flagged in skip_mask_digits.  Buffers are ping-ponged; on
// return *keys points at the sorted keys and *keys_alt at scratch.
void radix_sort_keys(xm_ctx* ctx, uint64_t** keys, uint64_t** keys_alt, uint64_t n, int lo_bit,
                     int hi_bit, uint64_t skip_mask_digits);

// ---- catalog kernels (xm_catalog.hip)
// Packed sort words (pack_word) of the records, their zone range and bad count.
void launch_make_keys(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg, int32_t n_zones,
                      PackLayout pl, uint64_t* words, int32_t* zminmax, uint32_t* bad);
// From the sorted words: rows[i] (unit vector + id of the input row the word
// carries) and, in place, words[i] -> catalog key zone << 24 | ra << (24 - ra_bits).
void launch_gather(xm_ctx* ctx, const xm_obj* recs, uint64_t* words, uint64_t n, PackLayout pl,
                   Row64* rows);
void launch_zone_offsets(xm_ctx* ctx, const uint64_t* keys, uint64_t n, int32_t zmin,
                         int32_t zmax, uint32_t* zoff);
void launch_chunk_counts(xm_ctx* ctx, const uint32_t* zoff, int32_t nzone, uint32_t* counts);
void launch_chunk_fill(xm_ctx* ctx, const uint32_t* zoff, const uint32_t* cstart, int32_t zmin,
                       int32_t nzone, Chunk* chunks);
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
// Hits are appended to n_regions regions of region_cap slots; block b appends to
// region b % n_regions through counters[region * REGION_STRIDE] (one 128-byte line
// per counter, so the waves of a join never queue on one address).  Counters count
// every hit; slots past region_cap are dropped and the caller re-runs.
constexpr uint32_t MAX_REGIONS = 1024;
constexpr uint32_t REGION_STRIDE = 16;
struct RegionOut {
    uint64_t* counters;
    uint32_t n_regions;
    uint64_t region_cap;
    int64_t* id_a;   // [n_regions * region_cap]
    int64_t* id_b;
    double* d2;      // chord^2 of each hit
};
uint32_t pair_blocks(uint32_t n_chunks);
// f32 = FP32 filter + FP64 recheck (needs A->row32 and B->row32), else FP64 plain.
void launch_pair(xm_ctx* ctx, const xm_catalog* A, const xm_catalog* B, const PairParams& pp,
                 bool f32, const RegionOut& ro);
void launch_to_row32(xm_ctx* ctx, const Row64* in, uint64_t n, Row32* out);
// prefix[r] = hits of the regions before r (capped at region_cap each);
// tail[0] = total hits, tail[1] = largest region count, tail[2] = FP32 band hits,
// tail[3] = largest |chord_32 - chord_64| of a kept pair (double bits).
void launch_region_scan(xm_ctx* ctx, const RegionOut& ro, uint64_t* prefix, uint64_t* tail);
// Packs the regions into contiguous arrays (region r's hits at prefix[r]) and
// converts chord^2 to the FP64 distance in arcsec.
void launch_compact(xm_ctx* ctx, const RegionOut& ro, const uint64_t* prefix, int64_t* id_a,
                    int64_t* id_b, double* dist);

}  // namespace xm
