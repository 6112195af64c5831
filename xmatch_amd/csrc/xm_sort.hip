// xm_sort.hip -- device-wide exclusive scan and LSD radix sort of (key, index) pairs.
//
// Replaces thrust::sort(less_zonera) of BoostXmatch/BoostXmatch/Sorter.cu:153.  The
// keys are packed 64-bit words zone | ra | input row, so one stable LSD radix sort
// on the zone and ra bits orders the segment by (zone, ra) without comparisons and
// carries each row's input position in the low bits.
//
// Radix sort, one histogram read and then one kernel per 8-bit digit:
//   k_radix_hist    one read of the keys: counts of every digit value at every
//                   digit position (block histograms in LDS, added up atomically)
//   k_digit_offsets per digit position: global start of each digit value
//   k_radix_scatter per 8192-key tile: each key's stable rank among the equal
//                   digits of its wavefront comes from a 64-bit lane mask per
//                   digit built in LDS; the tile's offset for each digit is
//                   the count of that digit in all earlier tiles, found by
//                   decoupled look-back over the tiles' published counts; keys
//                   are regrouped by digit in LDS, then written out so equal
//                   digits land in contiguous runs.
// Algorithmic HBM traffic per key: 8 bytes once (histogram), then 8 (read) + 8
// (write) bytes per digit.
#include <algorithm>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "xm_internal.h"

namespace xm {

namespace {

// 8192-key tiles of 8 wavefronts: measured on the 50M-key sort against 4096-key
// (256 x 16, 256 x 12), 8192-key (256 x 32), 6144-, 12288- and 16384-key tiles.
// Larger tiles mean fewer look-backs and longer runs of equal digits in the
// write-out; at 512 x 16 two blocks still fit a CU (LDS) at 4 waves per SIMD.
constexpr int RS_THREADS = 512;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;   // keys per tile
constexpr int RS_WAVES = RS_THREADS / WAVE;
constexpr int RS_WAVE_TILE = RS_TILE / RS_WAVES;  // keys per wavefront

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
constexpr int HIST_THREADS = 256;
constexpr int HIST_ITEMS = 8;          // keys per thread per round of the histogram
constexpr uint32_t HIST_BLOCKS = 1024; // histogram grid cap

// Look-back words, one per (tile, digit value): top bit = pass parity, next bit
// = inclusive prefix (count over this and all earlier tiles; else the tile's own
// count), low bits = count.  Every tile writes all 256 of its words in every
// pass, so after a pass all words carry that pass's parity and the next pass
// tells its fresh words from stale ones by that bit; the buffer is cleared (parity
// 0, which is not the first pass's) once per sort.  32-bit words hold counts below
// 2^30, 64-bit words any count; the smaller words halve the look-back's traffic.
template <typename W>
struct LbWord {
    static constexpr int TOP = int(sizeof(W) * 8) - 1;
    static constexpr W COUNT_MASK = sizeof(W) == 4 ? W(0x3fffffffu) : W(0xffffffffu);
    __device__ static W make(uint32_t parity, bool prefix, uint32_t count) {
        return (W(parity & 1u) << TOP) | (W(prefix ? 1u : 0u) << (TOP - 1)) | W(count);
    }
    __device__ static bool fresh(W v, uint32_t parity) { return uint32_t(v >> TOP) == (parity & 1u); }
    __device__ static bool prefix(W v) { return ((v >> (TOP - 1)) & 1u) != 0; }
    __device__ static uint32_t count(W v) { return uint32_t(v & COUNT_MASK); }
};
// A look-back waits only on tiles that started before it and publish their own
// counts before they wait on anything, so each wait ends.  The limit (seconds of
// polling) turns a broken invariant into an error instead of a stalled kernel.
constexpr uint32_t LB_SPIN_LIMIT = 1u << 22;
// earlier tiles' words read per round trip of the look-back (measured on the
// 50M-key sort: 4 beats 1, 2, 8 and 16; 3, 5 and 6 are within 1%; wider windows
// cost more traffic than the round trips they save)
constexpr int LB_WINDOW = 4;
// Tiles are dealt in stripes of XS_RUN consecutive tiles per XCD (blocks go to the 8
// XCDs round-robin), so that the runs neighbouring tiles write for one digit meet in
// one L2: at 200M keys (9-bit digits) 0.955 ms a pass with one tile counter, 0.903 /
// 0.882 / 0.879 / 0.888 ms with stripes of 4 / 8 / 16 / 32 tiles.
constexpr uint32_t XS_RUN = 16;
constexpr uint32_t XS_XCDS = 8;

// digit counters of a histogram: every pass of 8-bit digits, or up to
// SEG9_MAX_PASSES of 9-bit digits
constexpr int HIST_BINS = SORT_HIST_BINS;

// total[pass][digit] += counts of every DB-bit digit value over the block's keys
// (total starts at zero).
template <int DB>
__global__ __launch_bounds__(HIST_THREADS) void k_radix_hist(const uint64_t* __restrict__ keys,
                                                           uint32_t n, int lo_bit, int n_pass,
                                                           uint32_t* __restrict__ total) {
    constexpr int R = 1 << DB;
    __shared__ uint32_t h[HIST_BINS];
    for (int i = threadIdx.x; i < n_pass * R; i += HIST_THREADS) h[i] = 0;
    __syncthreads();
    const uint64_t stride = uint64_t(gridDim.x) * HIST_THREADS * HIST_ITEMS;
    for (uint64_t base = uint64_t(blockIdx.x) * HIST_THREADS * HIST_ITEMS; base < n; base += stride) {
        uint64_t k[HIST_ITEMS];
#pragma unroll
        for (int j = 0; j < HIST_ITEMS; j++) {
            const uint64_t i = base + uint64_t(j) * HIST_THREADS + threadIdx.x;
            k[j] = i < n ? keys[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < HIST_ITEMS; j++) {
            const uint64_t i = base + uint64_t(j) * HIST_THREADS + threadIdx.x;
            count_digits<DB>(h, k[j], i < n, lo_bit, n_pass);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_pass * R; i += HIST_THREADS)
        if (h[i]) atomicAdd(&total[i], h[i]);
}

// One block per digit position: start[pass][d] = keys whose digit is below d.
template <int DB>
__global__ __launch_bounds__(1 << DB) void k_digit_offsets(const uint32_t* __restrict__ total,
                                                          uint32_t* __restrict__ start) {
    constexpr int R = 1 << DB;
    __shared__ uint32_t s_w[R / WAVE + 1];
    const uint32_t i = blockIdx.x * R + threadIdx.x;
    start[i] = block_excl_scan<R>(total[i], s_w);
}

// One pass over the DB-bit digit at `shift`.  A block takes a ticket from its XCD's
// counter (tile_ctr[blockIdx % 8]) when it starts; the XCD's i-th ticket is tile
// (i / XS_RUN) 8 XS_RUN + xcd XS_RUN + i % XS_RUN while whole stripes last, then the
// tail goes round-robin (a bijection onto the tiles for any tile count).  A tile
// waits only on tiles of smaller index; with blocks dispatched in order, those are
// held by blocks at most 8 XS_RUN - 1 = 127 places later than the waiter, far inside
// the 512 the GPU holds at once, so they run while it waits (and the polling limit
// turns a broken invariant into an error, never a hang).  A tile publishes its digit counts (an LDS histogram taken as its keys
// arrive) before it ranks them, and looks back only after ranking and regrouping,
// when the earlier tiles have mostly published their prefixes.  With 9-bit digits
// the per-wave counts are 16-bit (a wave holds 1024 keys, a tile 8192), so two
// blocks still fit a CU's LDS.
template <typename W, int DB>
__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(
    const uint64_t* __restrict__ kin, uint64_t* __restrict__ kout, uint32_t n, int shift,
    uint32_t parity, const uint32_t* __restrict__ digit_start, W* status, uint32_t* tile_ctr,
    uint32_t* err, uint32_t spin_limit, uint32_t publish) {
    using LB = LbWord<W>;
    constexpr uint32_t R = 1u << DB;
    using WC = typename std::conditional<DB == 8, uint32_t, uint16_t>::type;
    static_assert(R <= uint32_t(RS_THREADS), "one thread per digit value");
    __shared__ uint64_t s_k[RS_TILE];
    __shared__ WC s_wc[RS_WAVES][R];   // per-wave running digit counts, then prefixes
    __shared__ uint32_t s_hist[R];     // the tile's digit counts
    __shared__ uint32_t s_go[R];       // global start minus tile-local start of each digit
    __shared__ uint32_t s_w[RS_THREADS / WAVE + 1];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_abort;

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w = tid / WAVE;
    // threads below R own one digit value each in the per-digit steps
    const bool dig = uint32_t(tid) < R;
    if (tid == 0) {
        // a failed look-back in an earlier pass left holes of stale words in this
        // pass's input; their digits could place keys past n, so the pass stops
        s_abort = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s_abort) {
            s_tile = 0u;
        } else {
            const uint32_t nt = (n + RS_TILE - 1) / RS_TILE;
            const uint32_t x = blockIdx.x % XS_XCDS;
            const uint32_t i = atomicAdd(tile_ctr + x, 1u);
            const uint32_t full = nt / (XS_XCDS * XS_RUN);  // whole rounds of stripes
            s_tile = i < full * XS_RUN
                         ? (i / XS_RUN) * XS_XCDS * XS_RUN + x * XS_RUN + i % XS_RUN
                         : full * XS_XCDS * XS_RUN + (i - full * XS_RUN) * XS_XCDS + x;
        }
    }
    // per-wave digit masks for the ranking live in s_k until the regroup
    uint64_t* const s_mask = s_k;
    static_assert(RS_TILE >= RS_WAVES * R, "s_k holds the ranking masks");
    for (uint32_t i = tid; i < RS_WAVES * R; i += RS_THREADS) {
        (&s_wc[0][0])[i] = 0;
        s_mask[i] = 0;
    }
    if (dig) s_hist[tid] = 0;
    __syncthreads();
    if (s_abort) return;  // the whole block: s_abort is one LDS word
    const uint32_t tile = s_tile;

    const uint32_t base = tile * RS_TILE + w * RS_WAVE_TILE;
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
        if (i < n) atomicAdd(&s_hist[uint32_t(key[k] >> shift) & (R - 1)], 1u);
    }
    __syncthreads();
    const uint32_t cnt = dig ? s_hist[tid] : 0u;
    W* const mine = status + size_t(tile) * R + tid;
    if (dig && publish)
        __hip_atomic_store(mine, LB::make(parity, tile == 0, cnt), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);

    const uint64_t lt = (uint64_t(1) << lane) - 1u;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const uint32_t i = base + k * WAVE + lane;
        const bool valid = i < n;
        const uint32_t d = uint32_t(key[k] >> shift) & (R - 1);
        // Valid lanes are a prefix of the wavefront (keys run from the tile start),
        // so lane 0 holds a valid digit whenever any lane does.
        const uint32_t d0 = uint32_t(__builtin_amdgcn_readfirstlane(int(d)));
        if (__all(!valid || d == d0)) {
            // one digit across the wavefront: the valid lanes are its group
            const uint64_t m = __ballot(valid);
            const uint32_t pre = s_wc[w][d0];
            rk[k] = pre + uint32_t(__popcll(m & lt));
            if (valid && (m >> lane) == 1u) s_wc[w][d0] = WC(pre + uint32_t(__popcll(m)));
            continue;
        }
        // lanes holding the same digit: each sets its bit in the digit's mask, then
        // reads the whole mask back (a wavefront's LDS operations run in order)
        uint64_t m = 0;
        uint32_t pre = 0;
        if (valid) {
            atomicOr(reinterpret_cast<unsigned long long*>(&s_mask[w * R + d]),
                     (unsigned long long)(uint64_t(1) << lane));
            m = s_mask[w * R + d];
            pre = s_wc[w][d];
        }
        rk[k] = pre + uint32_t(__popcll(m & lt));
        // the highest lane of each digit group advances that digit's count and
        // clears its mask, after the wavefront's reads above
        if (valid && (m >> lane) == 1u) {
            s_wc[w][d] = WC(pre + uint32_t(__popcll(m)));
            s_mask[w * R + d] = 0;
        }
    }
    // tile-local start of each digit; then s_wc[q][d] becomes the tile-local
    // position of wavefront q's first key of digit d, so the regroup reads one word
    // per key
    const uint32_t ts = block_excl_scan<RS_THREADS>(cnt, s_w);
    if (dig) {
        uint32_t run = ts;
#pragma unroll
        for (int q = 0; q < RS_WAVES; q++) {
            const uint32_t c = s_wc[q][tid];
            s_wc[q][tid] = WC(run);
            run += c;
        }
    }
    __syncthreads();
    // regroup by digit in LDS (tile-local positions only)
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const uint32_t i = base + k * WAVE + lane;
        if (i < n) {
            const uint32_t d = uint32_t(key[k] >> shift) & (R - 1);
            s_k[s_wc[w][d] + rk[k]] = key[k];
        }
    }
    // look-back: sum the counts published by the earlier tiles, nearest first, up
    // to and including the first inclusive prefix
    uint32_t excl = 0;
    if (dig && tile > 0) {
        int32_t j = int32_t(tile) - 1;
        uint32_t spins = 0;
        for (;;) {
            W v[LB_WINDOW];
#pragma unroll
            for (int k = 0; k < LB_WINDOW; k++) {
                // tile 0 always holds a prefix, so the walk never passes it
                v[k] = j - k >= 0 ? __hip_atomic_load(status + size_t(j - k) * R + tid,
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : LB::make(parity, true, 0);
            }
            int used = 0;
            bool done = false;
#pragma unroll
            for (int k = 0; k < LB_WINDOW; k++) {
                if (!done && used == k && LB::fresh(v[k], parity)) {
                    excl += LB::count(v[k]);
                    used = k + 1;
                    done = LB::prefix(v[k]);
                }
            }
            if (done) break;
            j -= used;
            if (used == 0) {
                if (++spins >= spin_limit) {
                    atomicOr(err, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (publish)
            __hip_atomic_store(mine, LB::make(parity, true, excl + cnt), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    if (dig) s_go[tid] = digit_start[tid] + excl - ts;
    __syncthreads();
    const uint32_t tile0 = tile * RS_TILE;
    const uint32_t tile_n = (n - tile0) < uint32_t(RS_TILE) ? (n - tile0) : uint32_t(RS_TILE);
    for (uint32_t p = tid; p < tile_n; p += RS_THREADS) {
        const uint64_t kk = s_k[p];
        kout[s_go[uint32_t(kk >> shift) & (R - 1)] + p] = kk;
    }
}

template <typename W, int DB>
void scatter_passes(xm_ctx* ctx, uint64_t** keys, uint64_t** keys_alt, uint64_t n, int lo_bit,
                    int n_pass, uint32_t nblocks, const uint32_t* start, void* status,
                    uint32_t* ctr) {
    // a look-back that reaches its polling limit sets the context's sort error
    // word (reported by sort_check)
    uint32_t* const err = sort_error_word(ctx);
    const auto kernel = k_radix_scatter<W, DB>;
    for (int p = 0; p < n_pass; p++) {
        launch_timed(ctx, "radix_scatter", [&] {
            hipLaunchKernelGGL(kernel, dim3(nblocks), dim3(RS_THREADS), 0,
                               ctx->stream, (const uint64_t*)*keys, *keys_alt, uint32_t(n),
                               lo_bit + DB * p, uint32_t(p + 1), start + p * (1 << DB),
                               static_cast<W*>(status),
                               ctr + XS_XCDS * p, err,
                               // a context set to fail its sorts (tests of the error
                               // path): tiles publish nothing and poll 64 times
                               ctx->sort_fail_test ? 64u : LB_SPIN_LIMIT,
                               ctx->sort_fail_test ? 0u : 1u);
        });
        std::swap(*keys, *keys_alt);
    }
}

}  // namespace

static_assert(SC_TILE == SCAN_TILE, "the scan's tile");

void scan_exclusive_u32(xm_ctx* ctx, uint32_t* data, uint64_t n) {
    if (n == 0) return;
    const uint64_t nt = (n + SC_TILE - 1) / SC_TILE;
    if (nt == 1) {
        scan_exclusive_u32_with(ctx, data, n, nullptr);
        return;
    }
    ScopedBlocks tmp(ctx);
    scan_exclusive_u32_with(ctx, data, n, tmp.alloc<uint32_t>(nt * sizeof(uint32_t)));
}

void scan_exclusive_u32_with(xm_ctx* ctx, uint32_t* data, uint64_t n, uint32_t* partial) {
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
}

uint32_t* sort_error_word(xm_ctx* ctx) {
    if (!ctx->sort_err) {
        ctx->sort_err = static_cast<uint32_t*>(dalloc(ctx, sizeof(uint32_t)));
        launch_fill_u32(ctx, ctx->sort_err, 1, 0u);
    }
    return ctx->sort_err;
}

namespace {

template <int DB>
void radix_sort_db(xm_ctx* ctx, uint64_t** keys, uint64_t** keys_alt, uint64_t n, int lo_bit,
                   int hi_bit, const uint32_t* digit_totals) {
    constexpr int R = 1 << DB;
    // digits starting at lo_bit (the last one may reach past hi_bit into bits that
    // are equal in every key)
    const int n_pass = sort_passes(hi_bit - lo_bit, DB);
    if (n_pass * R > HIST_BINS) throw InvalidArg{"radix sort: too many digit passes"};
    const uint32_t nblocks = uint32_t((n + RS_TILE - 1) / RS_TILE);
    const uint32_t hist_blocks = std::min<uint32_t>(
        HIST_BLOCKS, uint32_t((n + HIST_THREADS * HIST_ITEMS - 1) / (HIST_THREADS * HIST_ITEMS)));
    // one block: look-back words | tile counters | digit totals | digit starts
    // 32-bit look-back words suffice below 2^30 keys
    const bool narrow = n < (uint64_t(1) << 30) && !ctx->sort_wide_words;
    const size_t status_bytes = size_t(nblocks) * R * (narrow ? 4 : 8);
    const size_t ctr_words = 64;   // per pass, one tile counter per XCD
    static_assert(SORT_MAX_PASSES * XS_XCDS <= 64, "tile counters");
    const size_t digit_words = size_t(n_pass) * R;
    ScopedBlocks tmp(ctx);
    char* blk = tmp.alloc<char>(status_bytes + (ctr_words + 2 * digit_words) * 4);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(blk + status_bytes);
    uint32_t* total = ctr + ctr_words;
    uint32_t* start = total + digit_words;
    launch_fill_u32(ctx, reinterpret_cast<uint32_t*>(blk),
                    status_bytes / 4 + ctr_words + (digit_totals ? 0 : digit_words), 0u);
    if (!digit_totals) {
        launch_timed(ctx, "radix_hist", [&] {
            hipLaunchKernelGGL(k_radix_hist<DB>, dim3(hist_blocks), dim3(HIST_THREADS), 0,
                               ctx->stream, (const uint64_t*)*keys, uint32_t(n), lo_bit, n_pass,
                               total);
        });
    }
    launch_timed(ctx, "radix_offsets", [&] {
        hipLaunchKernelGGL(k_digit_offsets<DB>, dim3(n_pass), dim3(R), 0, ctx->stream,
                           digit_totals ? digit_totals : (const uint32_t*)total, start);
    });
    if (narrow)
        scatter_passes<uint32_t, DB>(ctx, keys, keys_alt, n, lo_bit, n_pass, nblocks, start, blk,
                                     ctr);
    else
        scatter_passes<uint64_t, DB>(ctx, keys, keys_alt, n, lo_bit, n_pass, nblocks, start, blk,
                                     ctr);
}

}  // namespace

void radix_sort_keys(xm_ctx* ctx, uint64_t** keys, uint64_t** keys_alt, uint64_t n, int lo_bit,
                     int hi_bit, const uint32_t* digit_totals, int digit_bits) {
    if (n <= 1 || hi_bit <= lo_bit) return;
    static_assert(SORT_MAX_KEYS == 0xffffffffull - RS_TILE, "the key limit follows the tile size");
    if (n > SORT_MAX_KEYS)
        throw InvalidArg{"radix sort supports up to 2^32 - " + std::to_string(RS_TILE + 1) + " keys"};
    if (lo_bit < 0 || hi_bit > 64) throw InvalidArg{"radix sort bit range outside the key"};
    if (digit_bits == 8)
        radix_sort_db<8>(ctx, keys, keys_alt, n, lo_bit, hi_bit, digit_totals);
    else if (digit_bits == 9)
        radix_sort_db<9>(ctx, keys, keys_alt, n, lo_bit, hi_bit, digit_totals);
    else
        throw InvalidArg{"radix sort digits are 8 or 9 bits"};
}

void sort_check(xm_ctx* ctx) {
    if (!ctx->sort_err) return;
    uint32_t* pin = reinterpret_cast<uint32_t*>(static_cast<char*>(ctx->pinned) + PIN_SORT_ERR_OFFSET);
    XM_HIP(hipMemcpyAsync(pin, ctx->sort_err, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    XM_HIP(hipStreamSynchronize(ctx->stream));
    if (*pin) {
        launch_fill_u32(ctx, ctx->sort_err, 1, 0u);
        throw std::runtime_error("radix sort: a tile's look-back did not complete");
    }
}

}  // namespace xm
