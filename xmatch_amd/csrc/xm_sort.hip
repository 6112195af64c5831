// xm_sort.hip -- device-wide exclusive scan and LSD radix sort of (key, index) pairs.
//
// Replaces thrust::sort(less_zonera) of BoostXmatch/BoostXmatch/Sorter.cu:153.  The
// keys are packed 64-bit words zone | ra | input row, so one stable LSD radix sort
// on the zone and ra bits orders the segment by (zone, ra) without comparisons and
// carries each row's input position in the low bits.
//
// Radix pass (8-bit digit), reduce-then-scan:
//   k_radix_count   per 4096-key tile: digit histogram -> counts[digit][tile]
//   scan            exclusive prefix over counts (digit-major) -> global offsets
//   k_radix_scatter per tile: wave64 ballot match of equal digits gives each key
//                   its stable rank; keys are regrouped by digit in LDS, then
//                   written out so equal digits land in contiguous runs.
// Algorithmic HBM traffic per pass and key: 8 (count) + 8 (scatter read) + 8
// (scatter write) bytes.
#include "xm_internal.h"

namespace xm {

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;   // 4096 keys per tile
constexpr int RS_WAVE_TILE = RS_TILE / 4;        // 1024 keys per wavefront

constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;   // 4096 values per tile
constexpr int TOP_THREADS = 1024;

__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t t = __shfl_up(v, d, WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}

// Exclusive scan of one value per thread over a block of NT threads.
// s_w must hold NT/64 + 1 words; s_w[NT/64] receives the block total.
template <int NT>
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* s_w) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const uint32_t inc = wave_incl_scan(v, lane);
    if (lane == WAVE - 1) s_w[w] = inc;
    __syncthreads();
    if (threadIdx.x < WAVE) {
        const int nw = NT / WAVE;
        const uint32_t ws = threadIdx.x < nw ? s_w[threadIdx.x] : 0u;
        const uint32_t wi = wave_incl_scan(ws, threadIdx.x);
        if (threadIdx.x < nw) s_w[threadIdx.x] = wi - ws;
        if (threadIdx.x == nw - 1) s_w[nw] = wi;
    }
    __syncthreads();
    const uint32_t r = inc - v + s_w[w];
    return r;
}

// ----------------------------------------------------------------- scan ----
__global__ __launch_bounds__(SC_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                            uint64_t n,
                                                            uint32_t* __restrict__ partial) {
    __shared__ uint32_t s_w[SC_THREADS / WAVE];
    const uint64_t base = uint64_t(blockIdx.x) * SC_TILE;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint64_t i = base + uint64_t(k) * SC_THREADS + threadIdx.x;
        if (i < n) sum += in[i];
    }
    sum = wave_sum(sum);
    if ((threadIdx.x & (WAVE - 1)) == 0) s_w[threadIdx.x / WAVE] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < SC_THREADS / WAVE; w++) t += s_w[w];
        partial[blockIdx.x] = t;
    }
}

// One block: exclusive scan of the per-tile partials, looping with a carry.
__global__ __launch_bounds__(TOP_THREADS) void k_scan_top(uint32_t* __restrict__ p, uint32_t np) {
    __shared__ uint32_t s_w[TOP_THREADS / WAVE + 1];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < np; base += TOP_THREADS * 4) {
        uint32_t v[4];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = base + threadIdx.x * 4 + k;
            v[k] = i < np ? p[i] : 0u;
            s += v[k];
        }
        const uint32_t ex = block_excl_scan<TOP_THREADS>(s, s_w);
        uint32_t run = carry + ex;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = base + threadIdx.x * 4 + k;
            if (i < np) p[i] = run;
            run += v[k];
        }
        carry += s_w[TOP_THREADS / WAVE];
        __syncthreads();
    }
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_apply(uint32_t* __restrict__ data, uint64_t n,
                                                           const uint32_t* __restrict__ partial) {
    // padded so that a thread's 16 consecutive values sit in distinct banks
    __shared__ uint32_t s[SC_TILE + SC_TILE / SC_ITEMS];
    __shared__ uint32_t s_w[SC_THREADS / WAVE + 1];
    const uint64_t base = uint64_t(blockIdx.x) * SC_TILE;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint32_t e = k * SC_THREADS + threadIdx.x;
        const uint64_t i = base + e;
        s[e + e / SC_ITEMS] = i < n ? data[i] : 0u;
    }
    __syncthreads();
    uint32_t v[SC_ITEMS];
    uint32_t sum = 0;
    const uint32_t t0 = threadIdx.x * SC_ITEMS;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint32_t e = t0 + k;
        v[k] = s[e + e / SC_ITEMS];
        sum += v[k];
    }
    uint32_t run = block_excl_scan<SC_THREADS>(sum, s_w) + (partial ? partial[blockIdx.x] : 0u);
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint32_t e = t0 + k;
        s[e + e / SC_ITEMS] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        const uint32_t e = k * SC_THREADS + threadIdx.x;
        const uint64_t i = base + e;
        if (i < n) data[i] = s[e + e / SC_ITEMS];
    }
}

// ----------------------------------------------------------- radix sort ----
__global__ __launch_bounds__(RS_THREADS) void k_radix_count(const uint64_t* __restrict__ keys,
                                                            uint32_t n, int shift,
                                                            uint32_t* __restrict__ counts,
                                                            uint32_t nblocks) {
    __shared__ uint32_t h[4][256];
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int q = 0; q < 4; q++) h[q][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const uint32_t i = base + k * RS_THREADS + threadIdx.x;
        if (i < n) {
            const uint32_t d = uint32_t(keys[i] >> shift) & 255u;
            atomicAdd(&h[w][d], 1u);
        }
    }
    __syncthreads();
    const uint32_t t = threadIdx.x;
    counts[size_t(t) * nblocks + blockIdx.x] = h[0][t] + h[1][t] + h[2][t] + h[3][t];
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(
    const uint64_t* __restrict__ kin, uint64_t* __restrict__ kout, uint32_t n, int shift,
    const uint32_t* __restrict__ offs, uint32_t nblocks) {
    __shared__ uint64_t s_k[RS_TILE];
    __shared__ uint32_t s_wc[4][256];   // per-wave running digit counts, then wave prefixes
    __shared__ uint32_t s_ts[256];      // tile-local start of each digit
    __shared__ uint32_t s_go[256];      // global start of each digit for this tile
    __shared__ uint32_t s_w[RS_THREADS / WAVE + 1];

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w = tid / WAVE;
#pragma unroll
    for (int q = 0; q < 4; q++) s_wc[q][tid] = 0;
    s_go[tid] = offs[size_t(tid) * nblocks + blockIdx.x];
    __syncthreads();

    const uint64_t lt = (uint64_t(1) << lane) - 1u;
    const uint32_t base = blockIdx.x * RS_TILE + w * RS_WAVE_TILE;
    uint64_t key[RS_ITEMS];
    uint32_t rk[RS_ITEMS];
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const uint32_t i = base + k * WAVE + lane;
        key[k] = i < n ? kin[i] : ~uint64_t(0);
    }
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const uint32_t i = base + k * WAVE + lane;
        const bool valid = i < n;
        const uint32_t d = uint32_t(key[k] >> shift) & 255u;
        // lanes holding the same digit: intersect the 8 per-bit ballots
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        uint32_t pre = 0;
        if (valid) pre = s_wc[w][d];
        rk[k] = pre + uint32_t(__popcll(m & lt));
        // the highest lane of each digit group advances that digit's count; the
        // wavefront's LDS read above is ordered before this write
        if (valid && (m >> lane) == 1u) s_wc[w][d] = pre + uint32_t(__popcll(m));
    }
    __syncthreads();
    {
        const uint32_t c0 = s_wc[0][tid], c1 = s_wc[1][tid], c2 = s_wc[2][tid], c3 = s_wc[3][tid];
        s_wc[0][tid] = 0;
        s_wc[1][tid] = c0;
        s_wc[2][tid] = c0 + c1;
        s_wc[3][tid] = c0 + c1 + c2;
        const uint32_t ex = block_excl_scan<RS_THREADS>(c0 + c1 + c2 + c3, s_w);
        s_ts[tid] = ex;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const uint32_t i = base + k * WAVE + lane;
        if (i < n) {
            const uint32_t d = uint32_t(key[k] >> shift) & 255u;
            s_k[s_ts[d] + s_wc[w][d] + rk[k]] = key[k];
        }
    }
    __syncthreads();
    const uint32_t tile0 = blockIdx.x * RS_TILE;
    const uint32_t tile_n = (n - tile0) < uint32_t(RS_TILE) ? (n - tile0) : uint32_t(RS_TILE);
    for (uint32_t p = tid; p < tile_n; p += RS_THREADS) {
        const uint64_t kk = s_k[p];
        const uint32_t d = uint32_t(kk >> shift) & 255u;
        kout[s_go[d] + (p - s_ts[d])] = kk;
    }
}

}  // namespace

void scan_exclusive_u32(xm_ctx* ctx, uint32_t* data, uint64_t n) {
    if (n == 0) return;
    const uint64_t nt = (n + SC_TILE - 1) / SC_TILE;
    if (nt == 1) {
        launch_timed(ctx, "scan_apply", [&] {
            hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SC_THREADS), 0, ctx->stream, data, n,
                               (const uint32_t*)nullptr);
        });
        return;
    }
    if (nt > 0xffffffffull) throw InvalidArg{"scan too large"};
    uint32_t* partial = static_cast<uint32_t*>(dalloc(ctx, nt * sizeof(uint32_t)));
    launch_timed(ctx, "scan_reduce", [&] {
        hipLaunchKernelGGL(k_scan_reduce, dim3(uint32_t(nt)), dim3(SC_THREADS), 0, ctx->stream,
                           data, n, partial);
    });
    launch_timed(ctx, "scan_top", [&] {
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(TOP_THREADS), 0, ctx->stream, partial,
                           uint32_t(nt));
    });
    launch_timed(ctx, "scan_apply", [&] {
        hipLaunchKernelGGL(k_scan_apply, dim3(uint32_t(nt)), dim3(SC_THREADS), 0, ctx->stream,
                           data, n, (const uint32_t*)partial);
    });
    dfree(ctx, partial);
}

void radix_sort_keys(xm_ctx* ctx, uint64_t** keys, uint64_t** keys_alt, uint64_t n, int lo_bit,
                     int hi_bit) {
    if (n <= 1) return;
    if (n > 0xffffffffull) throw InvalidArg{"radix sort supports fewer than 2^32 keys"};
    if (lo_bit < 0 || hi_bit > 64) throw InvalidArg{"radix sort bit range outside the key"};
    const uint32_t nblocks = uint32_t((n + RS_TILE - 1) / RS_TILE);
    uint32_t* counts = static_cast<uint32_t*>(dalloc(ctx, size_t(256) * nblocks * sizeof(uint32_t)));
    // 8-bit digits starting at lo_bit (the last one may reach past hi_bit into bits
    // that are equal in every key)
    for (int shift = lo_bit; shift < hi_bit; shift += 8) {
        launch_timed(ctx, "radix_count", [&] {
            hipLaunchKernelGGL(k_radix_count, dim3(nblocks), dim3(RS_THREADS), 0, ctx->stream,
                               (const uint64_t*)*keys, uint32_t(n), shift, counts, nblocks);
        });
        scan_exclusive_u32(ctx, counts, uint64_t(256) * nblocks);
        launch_timed(ctx, "radix_scatter", [&] {
            hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(RS_THREADS), 0, ctx->stream,
                               (const uint64_t*)*keys, *keys_alt, uint32_t(n), shift,
                               (const uint32_t*)counts, nblocks);
        });
        std::swap(*keys, *keys_alt);
    }
    dfree(ctx, counts);
}

}  // namespace xm
