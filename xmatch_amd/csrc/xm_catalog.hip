// xm_catalog.hip -- segment preparation kernels: keys, gather to SoA unit
// vectors, zone offsets, A-row chunks.
//
// Replaces Sorter::Sort (BoostXmatch/BoostXmatch/Sorter.cu:141-185):
//   GetZoneId + less_zonera (Sorter.cu:20-43)   -> k_make_keys (packed sort words)
//   get_id / get_radec_radian (Sorter.cu:45-62) -> k_gather (one 32-byte row per source:
//                                                  FP64 unit vector, radec2xyz of
//                                                  Worker.cu:31-40 done once per source
//                                                  instead of per pair, and the id)
//   lower_bound/upper_bound (Sorter.cu:161-164) -> also k_gather: a row whose zone
//                                                  differs from its left neighbour's
//                                                  starts that zone (and every empty
//                                                  zone between them) in zoff
//
// Many segments (the reference sorts each Segment on its own, one Sorter::Sort
// per segment, BoostXmatch.cpp:250-263) are prepared together: k_seg_range finds
// every segment's zone range, k_seg_words packs segment | zone | ra | row words in
// a per-segment layout, one radix sort orders all of them, and k_gather_seg
// splits them back into per-segment zone tables.
#include <algorithm>
#include <climits>

#include "xm_internal.h"

namespace xm {

namespace {

constexpr int TPB = 256;
// k_make_keys grid cap (grid-stride; also the number of digit-histogram flushes).
// Measured on 50M records: 1024 blocks 0.300 ms, 768 0.318, 512 0.391, 2048 0.320,
// 4096 0.328.
constexpr uint32_t MAKE_KEYS_BLOCKS = 1024;
// digit counters per histogram copy: every pass of 8-bit digits, or up to
// SORT9_MAX_PASSES of 9-bit digits
constexpr int MAKE_KEYS_HIST_BINS =
    SORT_MAX_PASSES * 256 > SORT9_MAX_PASSES * 512 ? SORT_MAX_PASSES * 256 : SORT9_MAX_PASSES * 512;

__device__ inline bool valid_record(double ra, double dec) {
    return ra >= 0.0 && ra < 360.0 && dec >= -90.0 && dec <= 90.0;  // false for NaN
}

// Grid-stride over the records with a capped grid, so the zone range and the
// bad-record count cost one atomic per block rather than one per 256 records.
// ---- one segment: how many low ra bits the sort key drops, and the digit width
// The sort runs over the zone and ra bits in 8- or 9-bit digits.  When they are a
// few bits over a whole number of digits, dropping those ra bits saves a pass; the
// coarser ra widens each B row's search by the granularity.  9-bit digits are
// taken where they need fewer passes than 8-bit ones (the 200M-row plane catalog
// at 1": 36 bits, 4 passes instead of 5).  k_layout_sample
// takes LAYOUT_SAMPLE records at an even stride: histograms of their zone (256
// bins over all zones) and ra (256 bins over the circle), their zone range and
// count.  k_layout_decide (one thread) estimates the rows a search meets per
// granule of ra: the mean rows per zone, n / n_zones, scaled by how concentrated
// the sample is in declination and in ra (256 sum(h^2) / m^2 for each; 1 for a
// uniform spread, ~1.23 in declination for a uniform sphere).  It drops the bits
// only if that adds at most LAYOUT_MAX_EXTRA rows to each search.  Measured on the
// 50M x 50M full-sky join at 2": dropping 6 of 19 ra bits (0.047 rows per search)
// took the sort from 5 to 4 passes, 8.43 -> 8.09 ms per step; 7 bits (0.093) gave
// 8.22 ms.
constexpr uint32_t LAYOUT_SAMPLE = 16384;
constexpr double LAYOUT_MAX_EXTRA = 0.1;
// stats: [0, 256) zone bins, [256, 512) ra bins, [512] zmin, [513] zmax,
// [514] sampled records, [515] ra bits to drop
static_assert(TPB == 256, "the layout kernels take one thread per histogram bin");

__global__ __launch_bounds__(TPB) void k_layout_sample(const xm_obj* __restrict__ recs, uint64_t n,
                                                       uint32_t m, double h_deg, int32_t n_zones,
                                                       uint32_t* __restrict__ stats) {
    __shared__ uint32_t hz[256], hr[256];
    __shared__ int32_t s_mn, s_mx;
    __shared__ uint32_t s_c;
    hz[threadIdx.x] = 0;
    hr[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        s_mn = INT_MAX;
        s_mx = -1;
        s_c = 0;
    }
    __syncthreads();
    for (uint32_t t = blockIdx.x * TPB + threadIdx.x; t < m; t += gridDim.x * TPB) {
        const uint64_t i = uint64_t(t) * n / m;
        const double ra = recs[i].ra_deg;
        const double dec = recs[i].dec_deg;
        if (!valid_record(ra, dec)) continue;
        const int32_t z = zone_of(dec, h_deg, n_zones);
        atomicAdd(&hz[uint32_t(uint64_t(z) * 256 / uint64_t(n_zones))], 1u);
        const uint32_t rb = uint32_t(ra * (256.0 / 360.0));
        atomicAdd(&hr[rb < 255u ? rb : 255u], 1u);
        atomicMin(&s_mn, z);
        atomicMax(&s_mx, z);
        atomicAdd(&s_c, 1u);
    }
    __syncthreads();
    if (hz[threadIdx.x]) atomicAdd(&stats[threadIdx.x], hz[threadIdx.x]);
    if (hr[threadIdx.x]) atomicAdd(&stats[256 + threadIdx.x], hr[threadIdx.x]);
    if (threadIdx.x == 0 && s_c) {
        atomicMin(reinterpret_cast<int32_t*>(&stats[512]), s_mn);
        atomicMax(reinterpret_cast<int32_t*>(&stats[513]), s_mx);
        atomicAdd(&stats[514], s_c);
    }
}

__global__ __launch_bounds__(TPB) void k_layout_decide(uint32_t* __restrict__ stats, uint64_t n,
                                                       int32_t n_zones, int ra_bits, int zone_bits,
                                                       uint32_t* __restrict__ layout_out) {
    __shared__ double s_z[TPB / WAVE], s_r[TPB / WAVE];
    const double hz = double(stats[threadIdx.x]), hr = double(stats[256 + threadIdx.x]);
    double sz = hz * hz, sr = hr * hr;
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) {
        sz += __shfl_xor(sz, d, WAVE);
        sr += __shfl_xor(sr, d, WAVE);
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        s_z[threadIdx.x / WAVE] = sz;
        s_r[threadIdx.x / WAVE] = sr;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const uint32_t m = stats[514];
    uint32_t drop = 0, digit_bits = 8;
    if (m > 0) {
        sz = s_z[0] + s_z[1] + s_z[2] + s_z[3];
        sr = s_r[0] + s_r[1] + s_r[2] + s_r[3];
        const double mm = double(m) * double(m);
        const double conc = (256.0 * sz / mm) * (256.0 * sr / mm);
        const int32_t zmn = int32_t(stats[512]), zmx = int32_t(stats[513]);
        const int bits = ra_bits + bit_length_u64(uint64_t(uint32_t(zmn) ^ uint32_t(zmx)));
        // for each digit width, the ra bits to drop to reach a whole number of
        // digits (if the rows each search meets grow little) and the passes left;
        // 9-bit digits only where they save a pass
        int best = SORT_MAX_PASSES + 1;
        for (int db = 8; db <= 9; db++) {
            int d = bits % db;
            if (d > 0) {
                bool ok = d < ra_bits;
                if (ok) {
                    const double granule = ldexp(1.0, d - ra_bits);  // of the circle
                    const double extra = double(n) / double(n_zones) * conc * 2.0 * granule;
                    ok = extra <= LAYOUT_MAX_EXTRA;
                }
                if (!ok) d = 0;
            }
            // the make_keys histograms cover the whole zone field
            if (db == 9 && sort_passes(ra_bits - d + zone_bits, 9) > SORT9_MAX_PASSES) continue;
            const int passes = sort_passes(bits - d, db);
            if (passes < best) {
                best = passes;
                drop = uint32_t(d);
                digit_bits = uint32_t(db);
            }
        }
    }
    stats[515] = drop;
    *layout_out = drop | digit_bits << 8;
}

// The block also counts the digits of its valid words for the radix sort: four
// LDS histograms, one per lane mod 4, so that spatially ordered input (whole
// wavefronts of one zone) meets 16-way rather than 64-way atomic conflicts
// (50M records sorted by dec: 0.31 ms against 0.42 ms with one histogram; random
// order 0.30 ms either way); their sums go to digit_totals at the end, one atomic
// per non-zero bin.
__global__ __launch_bounds__(TPB) void k_make_keys(const xm_obj* __restrict__ recs, uint32_t n,
                                                   double h_deg, int32_t n_zones, PackLayout pl,
                                                   uint64_t* __restrict__ words,
                                                   int32_t* __restrict__ zminmax,
                                                   uint32_t* __restrict__ bad, int lo_bit,
                                                   int n_pass, uint32_t* __restrict__ digit_totals,
                                                   const uint32_t* __restrict__ layout) {
    // the sort starts above the ra bits the key drops, in digits of the width the
    // layout chose (both decided on the device); the digit counts follow
    int db = 8;
    if (layout) {
        const uint32_t lw = *layout;
        const int drop = layout_drop(lw);
        db = layout_digit_bits(lw);
        lo_bit += drop;
        const int need = sort_passes(pl.ra_bits - drop + pl.zone_bits, db);
        n_pass = n_pass < need ? n_pass : need;
    }
    const uint32_t radix = 1u << db;
    const uint32_t dmask = radix - 1u;
    __shared__ int32_t s_min[TPB / WAVE], s_max[TPB / WAVE];
    __shared__ uint32_t s_bad[TPB / WAVE];
    __shared__ uint32_t s_h[4][MAKE_KEYS_HIST_BINS];
    const int bins = n_pass * int(radix);
    for (int i = threadIdx.x; i < bins; i += TPB)
#pragma unroll
        for (int c = 0; c < 4; c++) s_h[c][i] = 0;
    __syncthreads();
    uint32_t* const h = s_h[threadIdx.x & 3];
    int32_t zmn = 0x7fffffff, zmx = -1;
    uint32_t nbad = 0;
    const uint64_t stride = uint64_t(gridDim.x) * TPB;
    for (uint64_t i = uint64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += stride) {
        const double ra = recs[i].ra_deg;
        const double dec = recs[i].dec_deg;
        uint64_t word = ~uint64_t(0);
        if (valid_record(ra, dec)) {
            const int32_t z = zone_of(dec, h_deg, n_zones);
            word = pack_word(pl, z, ra, i);
            zmn = min(zmn, z);
            zmx = max(zmx, z);
            for (int p = 0; p < n_pass; p++)
                atomicAdd(&h[uint32_t(p) * radix + (uint32_t(word >> (lo_bit + db * p)) & dmask)],
                          1u);
        } else {
            nbad++;
        }
        words[i] = word;
    }
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) {
        zmn = min(zmn, __shfl_xor(zmn, d, WAVE));
        zmx = max(zmx, __shfl_xor(zmx, d, WAVE));
        nbad += __shfl_xor(nbad, d, WAVE);
    }
    const int w = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        s_min[w] = zmn;
        s_max[w] = zmx;
        s_bad[w] = nbad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t a = s_min[0], b = s_max[0];
        uint32_t nb = s_bad[0];
        for (int q = 1; q < TPB / WAVE; q++) {
            a = min(a, s_min[q]);
            b = max(b, s_max[q]);
            nb += s_bad[q];
        }
        if (b >= 0) {
            atomicMin(&zminmax[0], a);
            atomicMax(&zminmax[1], b);
        }
        if (nb) atomicAdd(bad, nb);
    }
    for (int i = threadIdx.x; i < bins; i += TPB)
        if (const uint32_t t = s_h[0][i] + s_h[1][i] + s_h[2][i] + s_h[3][i])
            atomicAdd(&digit_totals[i], t);
}

// ---- many segments: a block takes a run of consecutive rows of the concatenation
// (k_seg_range: a tile of SEG_TILE rows) and walks the segments that overlap it, so
// every segment-wide quantity costs one atomic per (block, segment).
constexpr uint32_t SEG_TILE = TPB * 16;

// Last segment whose row_base <= row (segments are in row order; empty segments
// share their base with the next one, so the search lands on a non-empty one).
__device__ inline uint32_t seg_of_row(const SegDesc* __restrict__ seg, uint32_t n_seg, uint64_t row) {
    uint32_t lo = 0, hi = n_seg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (seg[mid].row_base <= row) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline int32_t block_min(int32_t v, int32_t* s) {
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0) s[threadIdx.x / WAVE] = v;
    __syncthreads();
    v = s[0];
    for (int q = 1; q < TPB / WAVE; q++) v = min(v, s[q]);
    __syncthreads();
    return v;
}

__device__ inline int32_t block_max(int32_t v, int32_t* s) {
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0) s[threadIdx.x / WAVE] = v;
    __syncthreads();
    v = s[0];
    for (int q = 1; q < TPB / WAVE; q++) v = max(v, s[q]);
    __syncthreads();
    return v;
}

__global__ __launch_bounds__(TPB) void k_seg_range(const SegDesc* __restrict__ seg, uint32_t n_seg,
                                                   uint64_t n, int32_t* __restrict__ range) {
    __shared__ int32_t s_red[TPB / WAVE];
    const uint64_t t0 = uint64_t(blockIdx.x) * SEG_TILE;
    const uint64_t t1 = t0 + SEG_TILE < n ? t0 + SEG_TILE : n;
    for (uint32_t s = seg_of_row(seg, n_seg, t0); s < n_seg; s++) {
        const uint64_t base = seg[s].row_base, end = base + seg[s].n;
        if (base >= t1) break;
        if (end <= t0) continue;
        const uint64_t lo = base > t0 ? base : t0, hi = end < t1 ? end : t1;
        const xm_obj* __restrict__ recs = seg[s].recs;
        const double h = seg[s].h_deg;
        const int32_t nzs = seg[s].n_zones;
        int32_t zmn = 0x7fffffff, zmx = -1, nbad = 0;
        for (uint64_t i = lo + threadIdx.x; i < hi; i += TPB) {
            const double ra = recs[i - base].ra_deg;
            const double dec = recs[i - base].dec_deg;
            if (valid_record(ra, dec)) {
                const int32_t z = zone_of(dec, h, nzs);
                zmn = min(zmn, z);
                zmx = max(zmx, z);
            } else {
                nbad++;
            }
        }
        zmn = block_min(zmn, s_red);
        zmx = block_max(zmx, s_red);
        nbad = block_max(nbad > 0 ? 1 : 0, s_red);
        if (threadIdx.x == 0) {
            if (zmx >= 0) {
                atomicMin(&range[3 * s], zmn);
                atomicMax(&range[3 * s + 1], zmx);
            }
            if (nbad) atomicAdd(&range[3 * s + 2], 1);
        }
    }
}

// Each block takes one contiguous run of rows (finding its first segment once) and
// walks the segments in it.  Like k_make_keys, it also counts the digits of its
// words for the radix sort (segment bits are uniform over long runs, so
// count_digits adds a wavefront's shared digit once) and adds its counts to
// digit_totals at the end.
template <int DB>
__global__ __launch_bounds__(TPB) void k_seg_words(const SegDesc* __restrict__ seg, uint32_t n_seg,
                                                   uint64_t n, int seg_shift,
                                                   uint64_t* __restrict__ words, int lo_bit,
                                                   int n_pass, uint32_t* __restrict__ digit_totals) {
    constexpr int R = 1 << DB;
    __shared__ uint32_t s_h[SORT_HIST_BINS];
    for (int i = threadIdx.x; i < n_pass * R; i += TPB) s_h[i] = 0;
    __syncthreads();
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t t0 = uint64_t(blockIdx.x) * per;
    const uint64_t t1 = t0 + per < n ? t0 + per : n;
    if (t0 < t1) {
        for (uint32_t s = seg_of_row(seg, n_seg, t0); s < n_seg; s++) {
            const uint64_t base = seg[s].row_base, end = base + seg[s].n;
            if (base >= t1) break;
            if (end <= t0) continue;
            const uint64_t lo = base > t0 ? base : t0, hi = end < t1 ? end : t1;
            const xm_obj* __restrict__ recs = seg[s].recs;
            const double h = seg[s].h_deg;
            const int32_t nzs = seg[s].n_zones, zb = seg[s].zone_base;
            const int ib = seg[s].idx_bits, rb = seg[s].ra_bits;
            const uint64_t top = seg_shift < 64 ? uint64_t(s) << seg_shift : 0;
            for (uint64_t i = lo + threadIdx.x; i < hi; i += TPB) {
                const double ra = recs[i - base].ra_deg;
                const double dec = recs[i - base].dec_deg;
                const int32_t z = zone_of(dec, h, nzs) - zb;
                const uint64_t q = ra_quantum(ra) >> (RA_BITS - rb);
                const uint64_t word = top | (uint64_t(uint32_t(z)) << (rb + ib)) | (q << ib) | (i - base);
                words[i] = word;
                count_digits<DB>(s_h, word, true, lo_bit, n_pass);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_pass * R; i += TPB)
        if (s_h[i]) atomicAdd(&digit_totals[i], s_h[i]);
}

// The gathers write their rows and keys with streaming (non-temporal) stores, so
// the 2 GB they write per 50M rows does not push the records, which they read at
// random, out of the caches: 50M records in random order, 1.74 -> 1.61 ms.
// Streaming loads of the records made it slower (1.94 ms).
__device__ inline void store_row_nt(Row64* p, const Row64& r) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    d2 a, b;
    a.x = r.x;
    a.y = r.y;
    b.x = r.z;
    b.y = __builtin_bit_cast(double, r.id);
    __builtin_nontemporal_store(a, reinterpret_cast<d2*>(p));
    __builtin_nontemporal_store(b, reinterpret_cast<d2*>(p) + 1);
}

// The FP32 row of a unit vector (each component rounded to nearest), for catalogs
// built for the FP32 filter: written by the gather next to the FP64 row, so the
// FP32 form costs 16 more bytes of writes per row rather than a pass of its own
// (a separate pass reads the 32-byte rows back and writes the 16-byte ones).
__device__ inline void store_row32_nt(Row32* p, const Row64& r) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 v;
    v.x = float(r.x);
    v.y = float(r.y);
    v.z = float(r.z);
    v.w = 0.0f;
    __builtin_nontemporal_store(v, reinterpret_cast<f4*>(p));
}

// Blocks are dealt to the 8 XCDs round-robin.  The block that launch order b
// stands for is remapped so that XCD x takes the x-th contiguous run of rows (the
// blocks past the last whole round keep their place): the rows an XCD gathers at
// once draw their records from a narrow range of the input (one zone bucket of a
// bucketed build, a few segments of a many-segment one) held in its own L2.
__device__ inline uint32_t xcd_run_block(uint32_t b, uint32_t nb) {
    const uint32_t per = nb / 8;
    return b < per * 8 ? (b % 8) * per + b / 8 : b;
}

// k_gather takes one row per thread in blocks of GATHER_TPB: 50M random-order
// records, 256 threads 1.61 ms, 512 1.60 ms, 1024 1.59 ms.  k_gather_seg keeps
// blocks of TPB: the 10,000-segment batch's gather took 49.5 ms with 256 threads and
// 50.1 ms with 1024; dealt in runs per XCD, 46.0 ms (46.1 with 512 threads, 46.7
// with 1024).
constexpr int GATHER_TPB = 1024;

__global__ __launch_bounds__(TPB) void k_gather_seg(const SegDesc* __restrict__ seg, int seg_shift,
                                                    const uint64_t* __restrict__ words, uint32_t n,
                                                    Row64* __restrict__ rows,
                                                    Row32* __restrict__ rows32,
                                                    uint32_t* __restrict__ zoff, uint32_t n_zoff,
                                                    uint32_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ sort_err) {
    const uint32_t i = xcd_run_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
    if (i >= n || *sort_err) return;
    const uint64_t wd = words[i];
    const uint32_t s = seg_shift < 64 ? uint32_t(wd >> seg_shift) : 0u;
    const int ib = seg[s].idx_bits, rb = seg[s].ra_bits;
    const uint64_t zmask = (uint64_t(1) << seg[s].zone_bits) - 1;
    const uint64_t loc = wd & ((uint64_t(1) << ib) - 1);
    const uint64_t ra_q = (wd >> ib) & ((uint64_t(1) << rb) - 1);
    const int32_t zrel = int32_t((wd >> (ib + rb)) & zmask);
    // Zone offsets: rows of segment s hold zone-relative entries
    // zoff[zoff_base + zrel]; the first row of a segment opens its zone 0.
    int32_t zprev = -1;
    if (i > 0) {
        const uint64_t wp = words[i - 1];
        const uint32_t sp = seg_shift < 64 ? uint32_t(wp >> seg_shift) : 0u;
        if (sp == s) zprev = int32_t((wp >> (ib + rb)) & zmask);
    }
    const uint32_t base = seg[s].zoff_base;
    for (int32_t z = zprev + 1; z <= zrel; z++) zoff[base + uint32_t(z)] = i;
    if (i == n - 1) zoff[n_zoff - 1] = n;
    const xm_obj rec = seg[s].recs[loc];
    Row64 r;
    unit_vector(rec.ra_deg, rec.dec_deg, &r.x, &r.y, &r.z);
    r.id = rec.id;
    store_row_nt(&rows[i], r);
    if (rows32) store_row32_nt(&rows32[i], r);
    __builtin_nontemporal_store(uint32_t(ra_q << (RA_BITS - rb)), &keys[i]);
}

// A sort whose look-back failed (sort_err set) leaves words that may index past
// the records and the zone table: the gathers then write nothing (the catalog call
// reports the error).
__global__ __launch_bounds__(GATHER_TPB) void k_gather(const xm_obj* __restrict__ recs,
                                                const uint64_t* __restrict__ words, uint32_t n,
                                                PackLayout pl, int ra_drop, int32_t zmin,
                                                Row64* __restrict__ rows,
                                                Row32* __restrict__ rows32,
                                                uint32_t* __restrict__ zoff, uint32_t n_zoff,
                                                uint32_t* __restrict__ keys,
                                                const uint32_t* __restrict__ sort_err,
                                                uint32_t xcd_runs) {
    const uint32_t blk = xcd_runs ? xcd_run_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t i = blk * GATHER_TPB + threadIdx.x;
    if (i >= n || *sort_err) return;
    const uint64_t wd = words[i];
    const int zs = pl.idx_bits + pl.ra_bits;
    const uint64_t s = wd & ((uint64_t(1) << pl.idx_bits) - 1);
    const uint64_t ra_q = (wd >> pl.idx_bits) & ((uint64_t(1) << pl.ra_bits) - 1);
    const int32_t zone = int32_t(wd >> zs);
    const int32_t zprev = i > 0 ? int32_t(words[i - 1] >> zs) : zmin - 1;
    for (int32_t z = zprev + 1; z <= zone; z++) zoff[uint32_t(z - zmin)] = i;
    if (i == n - 1) zoff[n_zoff - 1] = n;
    const int64_t oid = recs[s].id;
    const double ra = recs[s].ra_deg;
    const double dec = recs[s].dec_deg;
    Row64 r;
    unit_vector(ra, dec, &r.x, &r.y, &r.z);
    r.id = oid;
    store_row_nt(&rows[i], r);
    if (rows32) store_row32_nt(&rows32[i], r);
    __builtin_nontemporal_store(
        uint32_t((ra_q >> ra_drop) << (RA_BITS - pl.ra_bits + ra_drop)), &keys[i]);
}

// zoff[zoff_base(s) + z - zmin(s)] = first row of segment s whose zone is >= z,
// for z in [zmin(s), zmax(s)]; the entry after a segment's last is the next
// segment's first row, and zoff[n_zoff - 1] = n.  Row i writes the entries of the
// zones in (zone(i-1), zone(i)] of its segment (from zmin(s) at a segment start).
//
// Chunks: zoff entry k (zone zmin(s) + k - zoff_base(s) of segment s) holds
// ceil(rows / chunk_rows(s)) chunks.  The entry's segment is the last one whose
// zoff_base <= k: an empty segment shares its base with the next, so the search
// lands on the segment that owns the entry.
__device__ inline uint32_t seg_of_entry(const SegDesc* __restrict__ seg, uint32_t n_seg,
                                        uint32_t k) {
    uint32_t lo = 0, hi = n_seg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (seg[mid].zoff_base <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(TPB) void k_chunk_counts(const uint32_t* __restrict__ zoff,
                                                      uint32_t n_entries,
                                                      const SegDesc* __restrict__ seg,
                                                      uint32_t n_seg,
                                                      uint32_t* __restrict__ counts) {
    const uint32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= n_entries) return;
    const uint32_t rows = zoff[k + 1] - zoff[k];
    const uint32_t cr = seg[seg_of_entry(seg, n_seg, k)].chunk_rows;
    counts[k] = (rows + cr - 1) / cr;
}

__global__ __launch_bounds__(TPB) void k_chunk_fill(const uint32_t* __restrict__ zoff,
                                                    const uint32_t* __restrict__ cstart,
                                                    uint32_t n_entries,
                                                    const SegDesc* __restrict__ seg,
                                                    uint32_t n_seg,
                                                    Chunk* __restrict__ chunks) {
    const uint32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= n_entries) return;
    const uint32_t r0 = zoff[k], r1 = zoff[k + 1];
    if (r1 <= r0) return;
    const SegDesc& sd = seg[seg_of_entry(seg, n_seg, k)];
    const uint32_t cr = sd.chunk_rows;
    const int32_t zone = sd.zmin + int32_t(k - sd.zoff_base);
    uint32_t c = cstart[k];
    for (uint32_t r = r0; r < r1; r += cr, c++) {
        Chunk ch;
        ch.start = r;
        ch.count = (r1 - r) < cr ? (r1 - r) : cr;
        ch.zone = zone;
        ch.pad = 0;
        chunks[c] = ch;
    }
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_fill(T* __restrict__ p, uint64_t n, T v) {
    const uint64_t i = uint64_t(blockIdx.x) * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}

inline dim3 grid_for(uint64_t n) { return dim3(uint32_t((n + TPB - 1) / TPB)); }

// The one-segment build's scratch words, set in one launch rather than one fill
// each: small = {INT_MAX, -1, 0, 0} (zone min, zone max, bad records, layout),
// then the layout statistics (zero but zone min INT_MAX and zone max -1),
// then the digit counters (zero).
// ---- one segment, records in zone buckets
// Bucket of zone z: the sampled zone range [zlo, zhi] (layout statistics) cut into
// n_buckets equal parts; zones outside it go to the first or last bucket.  The
// scale nb / span is one FP64 product per record (an int64 division took ~80
// instructions); the sample and the key pass use the same function.
struct BucketMap {
    int32_t zlo;
    uint32_t nb;
    double scale;
};

__device__ inline BucketMap bucket_map(const uint32_t* __restrict__ stats, int32_t n_zones,
                                       uint32_t nb) {
    int32_t a = int32_t(stats[512]), b = int32_t(stats[513]);
    if (b < a) {  // nothing valid sampled
        a = 0;
        b = n_zones - 1;
    }
    BucketMap m;
    m.zlo = a;
    m.nb = nb;
    m.scale = double(nb) / (double(b) - double(a) + 1.0);
    return m;
}

__device__ inline uint32_t bucket_of(int32_t z, const BucketMap& m) {
    const double k = (double(z) - double(m.zlo)) * m.scale;
    return k < 0.0 ? 0u : (k >= double(m.nb) ? m.nb - 1 : uint32_t(k));
}

// hist[b] += sampled records of bucket b (m records at an even stride)
__global__ __launch_bounds__(TPB) void k_bucket_sample(const xm_obj* __restrict__ recs, uint64_t n,
                                                       uint32_t m, double h_deg, int32_t n_zones,
                                                       const uint32_t* __restrict__ stats,
                                                       uint32_t nb, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[MAX_BUCKETS];
    for (uint32_t b = threadIdx.x; b < nb; b += TPB) h[b] = 0;
    __syncthreads();
    const BucketMap bm = bucket_map(stats, n_zones, nb);
    for (uint32_t t = blockIdx.x * TPB + threadIdx.x; t < m; t += gridDim.x * TPB) {
        const uint64_t i = uint64_t(t) * n / m;
        const double ra = recs[i].ra_deg;
        const double dec = recs[i].dec_deg;
        if (!valid_record(ra, dec)) continue;
        atomicAdd(&h[bucket_of(zone_of(dec, h_deg, n_zones), bm)], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += TPB)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

// One block.  Run (b, g) -- bucket b, group g -- holds est(b) pct / 100 / groups +
// BUCKET_SLACK rows, est(b) = n hist[b] / sampled; runs are laid out bucket by bucket
// (a bucket's groups side by side); cursor[g][b] = the run's start, end[g][b] its end,
// the overflow cursor starts after the last run.
constexpr uint32_t PLAN_TPB = 1024;
__global__ __launch_bounds__(PLAN_TPB) void k_bucket_plan(const uint32_t* __restrict__ hist,
                                                          uint32_t nb, uint64_t n, uint32_t pct,
                                                          uint32_t* __restrict__ plan) {
    __shared__ uint64_t s_part[PLAN_TPB];
    __shared__ uint64_t s_tot;
    const uint32_t t = threadIdx.x;
    uint64_t sum = 0;
    for (uint32_t b = t; b < nb; b += PLAN_TPB) sum += hist[b];
    s_part[t] = sum;
    __syncthreads();
    if (t == 0) {
        uint64_t a = 0;
        for (uint32_t k = 0; k < PLAN_TPB; k++) a += s_part[k];
        s_tot = a;
    }
    __syncthreads();
    const uint64_t sampled = s_tot;
    const uint32_t runs = nb * BUCKET_GROUPS;
    auto run_rows = [&](uint32_t e) -> uint64_t {
        const uint32_t b = e / BUCKET_GROUPS;
        const uint64_t est = sampled ? uint64_t(hist[b]) * n / sampled : n / nb;
        return est * pct / (100 * BUCKET_GROUPS) + BUCKET_SLACK;
    };
    // each thread a contiguous range of runs; a serial scan of the thread sums
    const uint32_t per = (runs + PLAN_TPB - 1) / PLAN_TPB;
    const uint32_t e0 = t * per, e1 = e0 + per < runs ? e0 + per : runs;
    uint64_t mine = 0;
    for (uint32_t e = e0; e < e1; e++) mine += run_rows(e);
    __syncthreads();
    s_part[t] = mine;
    __syncthreads();
    if (t == 0) {
        uint64_t a = 0;
        for (uint32_t k = 0; k < PLAN_TPB; k++) {
            const uint64_t v = s_part[k];
            s_part[k] = a;
            a += v;
        }
        s_tot = a;
    }
    __syncthreads();
    uint64_t run = s_part[t];
    uint32_t* const cursor = plan;
    uint32_t* const end = plan + runs;
    for (uint32_t e = e0; e < e1; e++) {
        const uint32_t b = e / BUCKET_GROUPS, g = e % BUCKET_GROUPS;
        cursor[g * nb + b] = uint32_t(run);
        run += run_rows(e);
        end[g * nb + b] = uint32_t(run);
    }
    if (t == 0) {
        plan[2 * runs + BP_OVF_CURSOR] = uint32_t(s_tot);
        plan[2 * runs + BP_FAILED] = 0;
    }
}

// The key pass of a bucketed build.  Block k takes rows [k per_blk, (k + 1) per_blk)
// in chunks of BK_CHUNK and appends to the runs of group k % BUCKET_GROUPS: the
// chunk's records are counted per bucket in LDS, each bucket reserves its count in
// its run with one atomic (the rest, past the run's end, in the overflow run), and
// the records are regrouped by bucket in LDS and written as contiguous runs.  The
// words go out in input order with the row field = the record's place in part, so
// the sort orders ties by input row as the plain build does, and the catalog is the
// same.  A block holds one chunk's records in LDS, so a CU runs one block: the
// next chunk's records are loaded into registers while this chunk reserves and
// writes, and the staged records sit in one LDS array per field (8-byte lanes
// read back without bank conflicts).  The 200M-row plane catalog (2048 buckets):
// 3.47 ms per pass with records loaded where they are ranked, one 24-byte LDS
// array and an int64 division per bucket; 2.97 ms with the prefetch, the LDS
// arrays per field and the FP64 bucket map; 2.74 ms with each lane counting its
// own digits.  Deeper pipelines measured slower: counting and reserving chunk k+1
// while chunk k is written out (3.45 ms at 1024 threads, which spills registers;
// 3.57 ms at 512 threads).
constexpr uint32_t BK_TPB = 1024;
constexpr uint32_t BK_CHUNK = 4096;
constexpr uint32_t BK_ITEMS = BK_CHUNK / BK_TPB;
constexpr uint32_t BK_NONE = 0xffffffffu;
static_assert(MAX_BUCKETS <= 2 * BK_TPB, "a thread owns at most two buckets");
__global__ __launch_bounds__(BK_TPB) void k_make_keys_bucketed(
    const xm_obj* __restrict__ recs, uint32_t n, uint32_t per_blk, double h_deg,
    int32_t n_zones, PackLayout pl, const uint32_t* __restrict__ stats, uint32_t nb,
    uint32_t* __restrict__ plan, uint32_t ovf_end, xm_obj* __restrict__ part,
    uint64_t* __restrict__ words, int32_t* __restrict__ zminmax, uint32_t* __restrict__ bad,
    int lo_bit, int n_pass, uint32_t* __restrict__ digit_totals,
    const uint32_t* __restrict__ layout) {
    __shared__ int64_t s_id[BK_CHUNK];
    __shared__ double s_ra[BK_CHUNK], s_dec[BK_CHUNK];
    __shared__ uint32_t s_pos[BK_CHUNK];
    __shared__ uint32_t s_lb[MAX_BUCKETS + 1];  // counts, then chunk-local starts
    __shared__ uint32_t s_base[MAX_BUCKETS];    // the run start each bucket reserved
    __shared__ uint32_t s_fit[MAX_BUCKETS];     // how many of them fit the run
    __shared__ uint32_t s_obase[MAX_BUCKETS];   // overflow start of the rest
    __shared__ uint32_t s_h[MAKE_KEYS_HIST_BINS];
    __shared__ uint32_t s_part[BK_TPB / WAVE];
    __shared__ int32_t s_min[BK_TPB / WAVE], s_max[BK_TPB / WAVE];
    __shared__ uint32_t s_bad[BK_TPB / WAVE];

    int db = 8;
    if (layout) {
        const uint32_t lw = *layout;
        const int drop = layout_drop(lw);
        db = layout_digit_bits(lw);
        lo_bit += drop;
        const int need = sort_passes(pl.ra_bits - drop + pl.zone_bits, db);
        n_pass = n_pass < need ? n_pass : need;
    }
    const int bins = n_pass << db;
    for (int i = threadIdx.x; i < bins; i += BK_TPB) s_h[i] = 0;
    for (uint32_t b = threadIdx.x; b <= nb; b += BK_TPB) s_lb[b] = 0;
    const BucketMap bm = bucket_map(stats, n_zones, nb);
    const uint32_t runs = nb * BUCKET_GROUPS;
    const uint32_t g = blockIdx.x % BUCKET_GROUPS;
    uint32_t* const cursor = plan + g * nb;
    const uint32_t* const run_end = plan + runs + g * nb;
    uint32_t* const ovf = plan + 2 * runs;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    int32_t zmn = 0x7fffffff, zmx = -1;
    uint32_t nbad = 0;
    const uint64_t lo = uint64_t(blockIdx.x) * per_blk;
    const uint64_t hi = lo + per_blk < n ? lo + per_blk : n;
    xm_obj x[BK_ITEMS];
#pragma unroll
    for (uint32_t j = 0; j < BK_ITEMS; j++) {
        const uint64_t i = lo + j * BK_TPB + threadIdx.x;
        if (i < hi) x[j] = recs[i];
    }
    __syncthreads();
    for (uint64_t c0 = lo; c0 < hi; c0 += BK_CHUNK) {
        const uint32_t m = uint32_t(hi - c0 < BK_CHUNK ? hi - c0 : BK_CHUNK);
        uint64_t wd[BK_ITEMS];
        uint32_t bk[BK_ITEMS], rk[BK_ITEMS];
#pragma unroll
        for (uint32_t j = 0; j < BK_ITEMS; j++) {
            const uint32_t sl = j * BK_TPB + threadIdx.x;
            bk[j] = BK_NONE;
            wd[j] = ~uint64_t(0);
            bool valid = false;
            if (sl < m) {
                valid = valid_record(x[j].ra_deg, x[j].dec_deg);
                if (valid) {
                    const int32_t z = zone_of(x[j].dec_deg, h_deg, n_zones);
                    wd[j] = pack_word(pl, z, x[j].ra_deg, 0);
                    zmn = min(zmn, z);
                    zmx = max(zmx, z);
                    bk[j] = bucket_of(z, bm);
                    rk[j] = atomicAdd(&s_lb[bk[j]], 1u);
                } else {
                    nbad++;
                }
            }
            // records in random order seldom share a digit across a wavefront, so
            // each counts its own (no uniform-digit test as in count_digits).  Testing
            // the top digit first and taking count_digits where a wavefront shares it
            // made dec-ordered input faster (50M full sky: 0.54 -> 0.43 ms) and random
            // input slower (the 200M plane: 2.48 -> 2.54 ms); the workloads' records
            // are in random order.
            if (valid) {
                const uint32_t dm = (1u << db) - 1u;
                for (int p = 0; p < n_pass; p++)
                    atomicAdd(&s_h[(uint32_t(p) << db) + (uint32_t(wd[j] >> (lo_bit + db * p)) & dm)],
                              1u);
            }
        }
        // the next chunk's records, in flight while this one reserves and writes
        const uint64_t c1 = c0 + BK_CHUNK;
        xm_obj xn[BK_ITEMS];
#pragma unroll
        for (uint32_t j = 0; j < BK_ITEMS; j++) {
            const uint64_t i = c1 + j * BK_TPB + threadIdx.x;
            if (i < hi) xn[j] = recs[i];
        }
        __syncthreads();
        // exclusive scan of the bucket counts (two per thread), s_lb[nb] = total
        {
            const uint32_t b0 = 2 * threadIdx.x;
            const uint32_t v0 = b0 < nb ? s_lb[b0] : 0u;
            const uint32_t v1 = b0 + 1 < nb ? s_lb[b0 + 1] : 0u;
            uint32_t inc = v0 + v1;
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                const uint32_t o = __shfl_up(inc, d, WAVE);
                if (lane >= d) inc += o;
            }
            if (lane == WAVE - 1) s_part[w] = inc;
            __syncthreads();
            uint32_t wbase = 0;
            for (int q = 0; q < w; q++) wbase += s_part[q];
            const uint32_t ex = wbase + inc - v0 - v1;
            __syncthreads();
            if (b0 < nb) s_lb[b0] = ex;
            if (b0 + 1 < nb) s_lb[b0 + 1] = ex + v0;
            if (threadIdx.x == BK_TPB - 1) {
                uint32_t tot = 0;
                for (int q = 0; q < BK_TPB / WAVE; q++) tot += s_part[q];
                s_lb[nb] = tot;
            }
        }
        __syncthreads();
        // each bucket reserves its rows in its run, the rest in the overflow run;
        // a thread's two buckets issue their atomics together
        {
            uint32_t cnt[2], base[2], e[2];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const uint32_t b = threadIdx.x + q * BK_TPB;
                cnt[q] = b < nb ? s_lb[b + 1] - s_lb[b] : 0u;
                base[q] = cnt[q] ? atomicAdd(&cursor[b], cnt[q]) : 0u;
                e[q] = cnt[q] ? run_end[b] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const uint32_t b = threadIdx.x + q * BK_TPB;
                if (!cnt[q]) continue;
                const uint32_t fit =
                    base[q] >= e[q] ? 0u : (e[q] - base[q] < cnt[q] ? e[q] - base[q] : cnt[q]);
                s_base[b] = base[q];
                s_fit[b] = fit;
                if (fit < cnt[q]) {
                    const uint32_t ob = atomicAdd(&ovf[BP_OVF_CURSOR], cnt[q] - fit);
                    if (uint64_t(ob) + (cnt[q] - fit) > ovf_end) {
                        atomicOr(&ovf[BP_FAILED], 1u);
                        s_obase[b] = BK_NONE;
                    } else {
                        s_obase[b] = ob;
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < BK_ITEMS; j++) {
            const uint32_t sl = j * BK_TPB + threadIdx.x;
            if (sl >= m) continue;
            uint32_t pos = BK_NONE;
            if (bk[j] != BK_NONE) {
                const uint32_t b = bk[j];
                if (rk[j] < s_fit[b]) pos = s_base[b] + rk[j];
                else if (s_obase[b] != BK_NONE) pos = s_obase[b] + (rk[j] - s_fit[b]);
                const uint32_t slot = s_lb[b] + rk[j];
                s_id[slot] = x[j].id;
                s_ra[slot] = x[j].ra_deg;
                s_dec[slot] = x[j].dec_deg;
                s_pos[slot] = pos;
            }
            words[c0 + sl] = pos != BK_NONE ? (wd[j] | pos) : ~uint64_t(0);
        }
        const uint32_t staged = s_lb[nb];
        __syncthreads();
        for (uint32_t sl = threadIdx.x; sl < staged; sl += BK_TPB) {
            const uint32_t pos = s_pos[sl];
            if (pos != BK_NONE) {
                xm_obj r;
                r.id = s_id[sl];
                r.ra_deg = s_ra[sl];
                r.dec_deg = s_dec[sl];
                part[pos] = r;
            }
        }
        // the counts start from zero for the next chunk (nothing reads them now)
        for (uint32_t b = threadIdx.x; b <= nb; b += BK_TPB) s_lb[b] = 0;
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < BK_ITEMS; j++) x[j] = xn[j];
    }
    // zone range and bad count: one atomic each per block
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) {
        zmn = min(zmn, __shfl_xor(zmn, d, WAVE));
        zmx = max(zmx, __shfl_xor(zmx, d, WAVE));
        nbad += __shfl_xor(nbad, d, WAVE);
    }
    if (lane == 0) {
        s_min[w] = zmn;
        s_max[w] = zmx;
        s_bad[w] = nbad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t a = s_min[0], b = s_max[0];
        uint32_t nbb = s_bad[0];
        for (int q = 1; q < BK_TPB / WAVE; q++) {
            a = min(a, s_min[q]);
            b = max(b, s_max[q]);
            nbb += s_bad[q];
        }
        if (b >= 0) {
            atomicMin(&zminmax[0], a);
            atomicMax(&zminmax[1], b);
        }
        if (nbb) atomicAdd(bad, nbb);
    }
    for (int i = threadIdx.x; i < bins; i += BK_TPB)
        if (s_h[i]) atomicAdd(&digit_totals[i], s_h[i]);
}

__global__ __launch_bounds__(TPB) void k_build_init(uint32_t* __restrict__ w, uint32_t n_words) {
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n_words) return;
    constexpr uint32_t zmin_at = BUILD_SMALL_WORDS + 512, zmax_at = BUILD_SMALL_WORDS + 513;
    uint32_t v = 0;
    if (i == 0 || i == zmin_at) v = 0x7fffffffu;
    else if (i == 1 || i == zmax_at) v = 0xffffffffu;
    w[i] = v;
}

}  // namespace

void launch_make_keys(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg, int32_t n_zones,
                      PackLayout pl, uint64_t* words, int32_t* zminmax, uint32_t* bad, int lo_bit,
                      int n_pass, uint32_t* digit_totals, const uint32_t* layout) {
    if (n == 0) return;
    if (n_pass < 0 || n_pass > SORT_MAX_PASSES || (n_pass > 0 && !digit_totals))
        throw InvalidArg{"make_keys: bad digit histogram request"};
    const uint32_t nblk = std::min<uint32_t>(grid_for(n).x, MAKE_KEYS_BLOCKS);
    launch_timed(ctx, "make_keys", [&] {
        hipLaunchKernelGGL(k_make_keys, dim3(nblk), dim3(TPB), 0, ctx->stream, recs, uint32_t(n),
                           h_deg, n_zones, pl, words, zminmax, bad, lo_bit, n_pass, digit_totals,
                           layout);
    });
}

void launch_build_init(xm_ctx* ctx, uint32_t* scratch, uint64_t n_digit_words) {
    const uint64_t n = BUILD_SMALL_WORDS + LAYOUT_STATS_WORDS + n_digit_words;
    launch_timed(ctx, "fill", [&] {
        hipLaunchKernelGGL(k_build_init, grid_for(n), dim3(TPB), 0, ctx->stream, scratch,
                           uint32_t(n));
    });
}

void launch_layout(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg, int32_t n_zones,
                   int ra_bits, int zone_bits, uint32_t* stats, uint32_t* layout_out) {
    if (n == 0) return;
    const uint32_t m = uint32_t(std::min<uint64_t>(n, LAYOUT_SAMPLE));
    launch_timed(ctx, "layout", [&] {
        hipLaunchKernelGGL(k_layout_sample, dim3(64), dim3(TPB), 0, ctx->stream, recs, n, m,
                           h_deg, n_zones, stats);
        hipLaunchKernelGGL(k_layout_decide, dim3(1), dim3(TPB), 0, ctx->stream, stats, n, n_zones,
                           ra_bits, zone_bits, layout_out);
    });
}

void launch_gather(xm_ctx* ctx, const xm_obj* recs, const uint64_t* words, uint64_t n,
                   PackLayout pl, int ra_drop, int32_t zmin, Row64* rows, Row32* rows32,
                   uint32_t* zoff, uint64_t n_zoff, uint32_t* keys, bool xcd_runs) {
    if (n == 0) return;
    const uint32_t* err = sort_error_word(ctx);
    launch_timed(ctx, "gather", [&] {
        hipLaunchKernelGGL(k_gather, dim3(uint32_t((n + GATHER_TPB - 1) / GATHER_TPB)),
                           dim3(GATHER_TPB), 0, ctx->stream, recs, words,
                           uint32_t(n), pl, ra_drop, zmin, rows, rows32, zoff, uint32_t(n_zoff),
                           keys, err, uint32_t(xcd_runs ? 1 : 0));
    });
}

void launch_seg_range(xm_ctx* ctx, const SegDesc* seg, uint32_t n_seg, uint64_t n, int32_t* range) {
    if (n == 0 || n_seg == 0) return;
    launch_timed(ctx, "seg_range", [&] {
        hipLaunchKernelGGL(k_seg_range, dim3(uint32_t((n + SEG_TILE - 1) / SEG_TILE)), dim3(TPB), 0,
                           ctx->stream, seg, n_seg, n, range);
    });
}

void launch_seg_words(xm_ctx* ctx, const SegDesc* seg, uint32_t n_seg, uint64_t n, int seg_shift,
                      uint64_t* words, int lo_bit, int n_pass, uint32_t* digit_totals,
                      int digit_bits) {
    if (n == 0 || n_seg == 0) return;
    if ((digit_bits != 8 && digit_bits != 9) || n_pass < 0 ||
        n_pass > (digit_bits == 8 ? SORT_MAX_PASSES : SEG9_MAX_PASSES) ||
        (n_pass > 0 && !digit_totals))
        throw InvalidArg{"seg_words: bad digit histogram request"};
    const uint32_t nblk = uint32_t(std::min<uint64_t>((n + TPB - 1) / TPB, MAKE_KEYS_BLOCKS));
    launch_timed(ctx, "make_keys", [&] {
        if (digit_bits == 9)
            hipLaunchKernelGGL(k_seg_words<9>, dim3(nblk), dim3(TPB), 0, ctx->stream, seg, n_seg,
                               n, seg_shift, words, lo_bit, n_pass, digit_totals);
        else
            hipLaunchKernelGGL(k_seg_words<8>, dim3(nblk), dim3(TPB), 0, ctx->stream, seg, n_seg,
                               n, seg_shift, words, lo_bit, n_pass, digit_totals);
    });
}

void launch_gather_seg(xm_ctx* ctx, const SegDesc* seg, int seg_shift, const uint64_t* words,
                       uint64_t n, Row64* rows, Row32* rows32, uint32_t* zoff, uint64_t n_zoff,
                       uint32_t* keys) {
    if (n == 0) return;
    const uint32_t* err = sort_error_word(ctx);
    launch_timed(ctx, "gather", [&] {
        hipLaunchKernelGGL(k_gather_seg, grid_for(n), dim3(TPB), 0, ctx->stream, seg, seg_shift,
                           words, uint32_t(n), rows, rows32, zoff, uint32_t(n_zoff), keys, err);
    });
}

void launch_chunk_counts(xm_ctx* ctx, const uint32_t* zoff, uint64_t n_entries,
                         const SegDesc* seg, uint32_t n_seg, uint32_t* counts) {
    if (n_entries == 0) return;
    launch_timed(ctx, "chunk_counts", [&] {
        hipLaunchKernelGGL(k_chunk_counts, grid_for(n_entries), dim3(TPB), 0, ctx->stream, zoff,
                           uint32_t(n_entries), seg, n_seg, counts);
    });
}

void launch_chunk_fill(xm_ctx* ctx, const uint32_t* zoff, const uint32_t* cstart,
                       uint64_t n_entries, const SegDesc* seg, uint32_t n_seg, Chunk* chunks) {
    if (n_entries == 0) return;
    launch_timed(ctx, "chunk_fill", [&] {
        hipLaunchKernelGGL(k_chunk_fill, grid_for(n_entries), dim3(TPB), 0, ctx->stream, zoff,
                           cstart, uint32_t(n_entries), seg, n_seg, chunks);
    });
}

void launch_fill_u32(xm_ctx* ctx, uint32_t* p, uint64_t n, uint32_t v) {
    if (n == 0) return;
    launch_timed(ctx, "fill", [&] {
        hipLaunchKernelGGL(k_fill<uint32_t>, grid_for(n), dim3(TPB), 0, ctx->stream, p, n, v);
    });
}

void launch_fill_u64(xm_ctx* ctx, uint64_t* p, uint64_t n, uint64_t v) {
    if (n == 0) return;
    launch_timed(ctx, "fill", [&] {
        hipLaunchKernelGGL(k_fill<uint64_t>, grid_for(n), dim3(TPB), 0, ctx->stream, p, n, v);
    });
}

void launch_fill_i32(xm_ctx* ctx, int32_t* p, uint64_t n, int32_t v) {
    if (n == 0) return;
    launch_timed(ctx, "fill", [&] {
        hipLaunchKernelGGL(k_fill<int32_t>, grid_for(n), dim3(TPB), 0, ctx->stream, p, n, v);
    });
}

BucketBuild bucket_build(uint64_t n, uint32_t capacity_pct) {
    BucketBuild bb;
    // ~4 MB of records a bucket, in [64, MAX_BUCKETS].  Fewer buckets make the key
    // pass cheaper (fewer reservations, longer runs) and the gather dearer (a bucket
    // outgrows the XCD's 4 MB L2).  The 200M-row plane catalog, key pass + gather per
    // catalog: 2 MB (2048 buckets) 2.74 + 3.88 ms, 3 MB 2.60 + 3.98, 4 MB 2.49 +
    // 4.07, 6 MB 2.39 + 4.20, 8 MB 2.36 + 4.34.
    const uint64_t want = n * sizeof(xm_obj) / (uint64_t(4) << 20);
    bb.n_buckets = uint32_t(std::min<uint64_t>(MAX_BUCKETS, std::max<uint64_t>(64, want)));
    const uint64_t runs = uint64_t(bb.n_buckets) * BUCKET_GROUPS;
    bb.c_main_max = n * capacity_pct / 100 + runs * BUCKET_SLACK;
    // the overflow run: 15% of the rows (and at least 64Ki)
    bb.c_total = bb.c_main_max + n * 15 / 100 + 65536;
    return bb;
}

void launch_bucket_plan(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg,
                        int32_t n_zones, const uint32_t* stats, const BucketBuild& bb,
                        uint32_t capacity_pct, uint32_t* hist, uint32_t* plan) {
    const uint32_t m = uint32_t(std::min<uint64_t>(n, BUCKET_SAMPLE));
    launch_timed(ctx, "bucket_plan", [&] {
        hipLaunchKernelGGL(k_bucket_sample, dim3(256), dim3(TPB), 0, ctx->stream, recs, n, m,
                           h_deg, n_zones, stats, bb.n_buckets, hist);
        hipLaunchKernelGGL(k_bucket_plan, dim3(1), dim3(PLAN_TPB), 0, ctx->stream,
                           (const uint32_t*)hist, bb.n_buckets, n, capacity_pct, plan);
    });
}

void launch_make_keys_bucketed(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg,
                               int32_t n_zones, PackLayout pl, const uint32_t* stats,
                               const BucketBuild& bb, uint32_t* plan, xm_obj* part,
                               uint64_t* words, int32_t* zminmax, uint32_t* bad, int lo_bit,
                               int n_pass, uint32_t* digit_totals, const uint32_t* layout) {
    if (n == 0) return;
    if (n_pass < 0 || n_pass > SORT_MAX_PASSES || (n_pass > 0 && !digit_totals))
        throw InvalidArg{"make_keys: bad digit histogram request"};
    if (bb.c_total >= (uint64_t(1) << 32) || bb.n_buckets > MAX_BUCKETS)
        throw InvalidArg{"bucketed build: too many rows"};
    // 512 blocks (one per CU at a time: 152 KB of LDS each), rows in whole chunks
    constexpr uint32_t BLOCKS = 512;
    static_assert(BLOCKS % BUCKET_GROUPS == 0, "blocks cycle through the groups");
    const uint64_t per = ((n + BLOCKS - 1) / BLOCKS + BK_CHUNK - 1) / BK_CHUNK * BK_CHUNK;
    launch_timed(ctx, "make_keys", [&] {
        hipLaunchKernelGGL(k_make_keys_bucketed, dim3(BLOCKS), dim3(BK_TPB), 0, ctx->stream, recs,
                           uint32_t(n), uint32_t(per), h_deg, n_zones, pl, stats, bb.n_buckets,
                           plan, uint32_t(bb.c_total), part, words, zminmax, bad, lo_bit, n_pass,
                           digit_totals, layout);
    });
}

}  // namespace xm
