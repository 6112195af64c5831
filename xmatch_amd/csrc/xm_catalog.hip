// xm_catalog.hip -- segment preparation kernels: keys, gather to SoA unit
// vectors, zone offsets, A-row chunks.
//
// Replaces Sorter::Sort (BoostXmatch/BoostXmatch/Sorter.cu:141-185):
//   GetZoneId + less_zonera (Sorter.cu:20-43)   -> k_make_keys (packed sort words)
//   lower_bound/upper_bound (Sorter.cu:161-164) -> k_zone_offsets
//   get_id / get_radec_radian (Sorter.cu:45-62) -> k_gather (one 32-byte row per source:
//                                                  FP64 unit vector, radec2xyz of
//                                                  Worker.cu:31-40 done once per source
//                                                  instead of per pair, and the id)
#include <algorithm>

#include "xm_internal.h"

namespace xm {

namespace {

constexpr int TPB = 256;

__device__ inline bool valid_record(double ra, double dec) {
    return ra >= 0.0 && ra < 360.0 && dec >= -90.0 && dec <= 90.0;  // false for NaN
}

// Grid-stride over the records with a capped grid, so the zone range and the
// bad-record count cost one atomic per block rather than one per 256 records.
__global__ __launch_bounds__(TPB) void k_make_keys(const xm_obj* __restrict__ recs, uint32_t n,
                                                   double h_deg, int32_t n_zones, PackLayout pl,
                                                   uint64_t* __restrict__ words,
                                                   int32_t* __restrict__ zminmax,
                                                   uint32_t* __restrict__ bad) {
    __shared__ int32_t s_min[TPB / WAVE], s_max[TPB / WAVE];
    __shared__ uint32_t s_bad[TPB / WAVE];
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
}

__global__ __launch_bounds__(TPB) void k_gather(const xm_obj* __restrict__ recs,
                                                uint64_t* __restrict__ words, uint32_t n,
                                                PackLayout pl, Row64* __restrict__ rows) {
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const uint64_t wd = words[i];
    const uint64_t s = wd & ((uint64_t(1) << pl.idx_bits) - 1);
    const uint64_t ra_q = (wd >> pl.idx_bits) & ((uint64_t(1) << pl.ra_bits) - 1);
    const uint64_t zone = wd >> (pl.idx_bits + pl.ra_bits);
    const int64_t oid = recs[s].id;
    const double ra = recs[s].ra_deg;
    const double dec = recs[s].dec_deg;
    Row64 r;
    unit_vector(ra, dec, &r.x, &r.y, &r.z);
    r.id = oid;
    rows[i] = r;
    words[i] = (zone << RA_BITS) | (ra_q << (RA_BITS - pl.ra_bits));
}

// zoff[z - zmin] = first sorted row whose zone is >= z, for z in [zmin, zmax+1].
// Row i writes the entries of the zones in (zone(i-1), zone(i)].
__global__ __launch_bounds__(TPB) void k_zone_offsets(const uint64_t* __restrict__ keys, uint32_t n,
                                                      int32_t zmin, int32_t zmax,
                                                      uint32_t* __restrict__ zoff) {
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i > n) return;
    const int32_t zp = i == 0 ? zmin - 1 : key_zone(keys[i - 1]);
    const int32_t zc = i == n ? zmax + 1 : key_zone(keys[i]);
    for (int32_t z = zp + 1; z <= zc; z++) zoff[z - zmin] = i;
}

__global__ __launch_bounds__(TPB) void k_chunk_counts(const uint32_t* __restrict__ zoff,
                                                      int32_t nzone,
                                                      uint32_t* __restrict__ counts) {
    const int32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= nzone) return;
    const uint32_t rows = zoff[k + 1] - zoff[k];
    counts[k] = (rows + CHUNK_ROWS - 1) / CHUNK_ROWS;
}

__global__ __launch_bounds__(TPB) void k_chunk_fill(const uint32_t* __restrict__ zoff,
                                                    const uint32_t* __restrict__ cstart,
                                                    int32_t zmin, int32_t nzone,
                                                    Chunk* __restrict__ chunks) {
    const int32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= nzone) return;
    const uint32_t r0 = zoff[k], r1 = zoff[k + 1];
    uint32_t c = cstart[k];
    for (uint32_t r = r0; r < r1; r += CHUNK_ROWS, c++) {
        Chunk ch;
        ch.start = r;
        ch.count = (r1 - r) < uint32_t(CHUNK_ROWS) ? (r1 - r) : uint32_t(CHUNK_ROWS);
        ch.zone = zmin + k;
        ch.aq = 0;
        chunks[c] = ch;
    }
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_fill(T* __restrict__ p, uint64_t n, T v) {
    const uint64_t i = uint64_t(blockIdx.x) * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}

inline dim3 grid_for(uint64_t n) { return dim3(uint32_t((n + TPB - 1) / TPB)); }

}  // namespace

void launch_make_keys(xm_ctx* ctx, const xm_obj* recs, uint64_t n, double h_deg, int32_t n_zones,
                      PackLayout pl, uint64_t* words, int32_t* zminmax, uint32_t* bad) {
    if (n == 0) return;
    const uint32_t nblk = std::min<uint32_t>(grid_for(n).x, 4096u);
    launch_timed(ctx, "make_keys", [&] {
        hipLaunchKernelGGL(k_make_keys, dim3(nblk), dim3(TPB), 0, ctx->stream, recs, uint32_t(n),
                           h_deg, n_zones, pl, words, zminmax, bad);
    });
}

void launch_gather(xm_ctx* ctx, const xm_obj* recs, uint64_t* words, uint64_t n, PackLayout pl,
                   Row64* rows) {
    if (n == 0) return;
    launch_timed(ctx, "gather", [&] {
        hipLaunchKernelGGL(k_gather, grid_for(n), dim3(TPB), 0, ctx->stream, recs, words,
                           uint32_t(n), pl, rows);
    });
}

void launch_zone_offsets(xm_ctx* ctx, const uint64_t* keys, uint64_t n, int32_t zmin, int32_t zmax,
                         uint32_t* zoff) {
    launch_timed(ctx, "zone_offsets", [&] {
        hipLaunchKernelGGL(k_zone_offsets, grid_for(n + 1), dim3(TPB), 0, ctx->stream, keys,
                           uint32_t(n), zmin, zmax, zoff);
    });
}

void launch_chunk_counts(xm_ctx* ctx, const uint32_t* zoff, int32_t nzone, uint32_t* counts) {
    if (nzone <= 0) return;
    launch_timed(ctx, "chunk_counts", [&] {
        hipLaunchKernelGGL(k_chunk_counts, grid_for(uint64_t(nzone)), dim3(TPB), 0, ctx->stream,
                           zoff, nzone, counts);
    });
}

void launch_chunk_fill(xm_ctx* ctx, const uint32_t* zoff, const uint32_t* cstart, int32_t zmin,
                       int32_t nzone, Chunk* chunks) {
    if (nzone <= 0) return;
    launch_timed(ctx, "chunk_fill", [&] {
        hipLaunchKernelGGL(k_chunk_fill, grid_for(uint64_t(nzone)), dim3(TPB), 0, ctx->stream,
                           zoff, cstart, zmin, nzone, chunks);
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

}  // namespace xm
