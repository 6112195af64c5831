// xm_pair.hip -- the zone-join pair kernel (FP64 plain form).
//
// Replaces the per-zone-pair launch loop of Worker::Match and xmatch_kernel
// (BoostXmatch/BoostXmatch/Worker.cu:330-359 and :65-241).  One launch covers the
// whole join.  Each wavefront owns a chunk of up to 64 consecutive A rows of one
// zone.  For each B zone z-nz..z+nz it finds, on the sorted packed keys, the rows
// whose ra lies in the chunk's ra span widened by alpha(zone) (Common.cu:16-23),
// the 0/360 seam split in two intervals, and tests them in one of two modes:
//
//  dense   one A row per lane (unit vector in registers); the window's B rows are
//          walked in wave-uniform order, their unit vectors arriving as wide scalar
//          (SGPR) loads, every lane testing its A row against each of them.
//  sparse  the chunk's A unit vectors and ra quanta are staged in LDS; each lane
//          takes one B row of the window, binary-searches the chunk for the A rows
//          within +-alpha of its ra, and tests only those.  At zone height = radius
//          a zone holds a few hundred sources over 360 deg, so a 64-row chunk spans
//          tens of degrees while alpha is arcseconds: almost every dense test would
//          be wasted.
//
// Predicate (Worker.cu:88-97 with the seam bug D2 corrected), on FP64 unit vectors:
//   d2 = fma(dz, dz, fma(dy, dy, dx*dx))  <  (2 sin(r/2))^2
// Hits are compacted with ballot + mbcnt into a per-wave LDS stage; a full stage
// (or the end of the chunk) costs one global atomic on the 64-bit append counter.
// Slots beyond the buffer capacity are counted but not written; the host then
// grows the buffer and re-runs, so nothing is ever truncated (Worker.cu:147-148,
// 384-387 dropped them).
#include <climits>

#include "xm_internal.h"

namespace xm {

namespace {

constexpr int PK_WAVES = 4;
constexpr int PK_THREADS = PK_WAVES * WAVE;
constexpr int STAGE = 256;
constexpr int UNROLL = 4;

__device__ inline uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ inline uint32_t mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

__device__ inline uint32_t uni(uint32_t v) {
    return uint32_t(__builtin_amdgcn_readfirstlane(int(v)));
}

__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// First index in [lo, hi) whose key is >= target (hi if none), by 64-ary search:
// each step probes 64 evenly spaced keys and keeps the one gap that brackets the
// answer.  Called by all 64 lanes with uniform arguments.
__device__ inline uint32_t wave_lower_bound(const uint64_t* __restrict__ keys, uint32_t lo,
                                            uint32_t hi, uint64_t target, uint32_t lane) {
    while (hi - lo > uint32_t(WAVE)) {
        const uint64_t step = (uint64_t(hi - lo) + WAVE - 1) / WAVE;
        const uint64_t p = uint64_t(lo) + (lane + 1) * step - 1;
        const bool pred = p < hi && keys[p] < target;
        const uint64_t c = uint64_t(__popcll(__ballot(pred)));
        const uint64_t nlo = uint64_t(lo) + c * step;
        const uint64_t cand = uint64_t(lo) + (c + 1) * step - 1;
        hi = uni(uint32_t(cand < hi ? cand : hi));
        lo = uni(uint32_t(nlo));
    }
    const uint32_t p = lo + lane;
    const bool pred = p < hi && keys[p] < target;
    return uni(lo + uint32_t(__popcll(__ballot(pred))));
}

__device__ inline double dist_arcsec(double d2) {
    return 2.0 * asin(sqrt(d2) * 0.5) * RAD2DEG * 3600.0;
}

// The chord of the reference's exact test (Worker.cu:93-97) on FP64 unit vectors:
// d2 = fma(dz, dz, fma(dy, dy, dx*dx)), identical on host and device.
__device__ inline double chord2(double ax, double ay, double az, double bx, double by, double bz) {
    const double dx = ax - bx;
    const double dy = ay - by;
    const double dz = az - bz;
    return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}

struct Stage {
    uint32_t* ra;
    uint32_t* rb;
    double* d2;
    uint32_t n;  // wave-uniform fill
};

struct Out {
    const int64_t* __restrict__ aid;
    const int64_t* __restrict__ bid;
    unsigned long long* __restrict__ counter;
    uint64_t cap;
    int64_t* __restrict__ id_a;
    int64_t* __restrict__ id_b;
    double* __restrict__ dist;
};

// Write the wave's staged hits to the result buffer: one atomic per call.
__device__ inline void flush_stage(Stage& st, uint32_t lane, const Out& o) {
    wave_lds_sync();
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(o.counter, (unsigned long long)st.n);
    base = __shfl(base, 0, WAVE);
    for (uint32_t e = lane; e < st.n; e += WAVE) {
        const uint64_t slot = uint64_t(base) + e;
        if (slot < o.cap) {
            o.id_a[slot] = o.aid[st.ra[e]];
            o.id_b[slot] = o.bid[st.rb[e]];
            o.dist[slot] = dist_arcsec(st.d2[e]);
        }
    }
    wave_lds_sync();
    st.n = 0;
}

// Append the lanes of hit mask m (A row ra, B row rb, per lane) to the stage;
// flush when it cannot take another full wave of hits.
__device__ inline void stage_hits(uint64_t m, double d2, uint32_t ra, uint32_t rb, uint32_t lane,
                                  Stage& st, const Out& o) {
    if ((m >> lane) & 1u) {
        const uint32_t pos = st.n + mbcnt64(m);
        st.ra[pos] = ra;
        st.rb[pos] = rb;
        st.d2[pos] = d2;
    }
    st.n += uint32_t(__popcll(m));
    if (st.n > uint32_t(STAGE - WAVE)) flush_stage(st, lane, o);
}

// Blocks are dealt to the 8 XCDs round-robin; give each XCD a contiguous run of
// chunks so that neighbouring chunks, which read the same B zones, share an L2.
__device__ inline uint32_t xcd_swizzle(uint32_t b, uint32_t nb) {
    const uint32_t per = nb / 8, rem = nb % 8, xcd = b % 8, idx = b / 8;
    return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Dense mode: B rows [j0, j1) against the lane's A row (registers).
__device__ inline void dense_range(uint32_t j0, uint32_t j1, double ax, double ay, double az,
                                   uint32_t row, double thr, const double* __restrict__ bx_,
                                   const double* __restrict__ by_, const double* __restrict__ bz_,
                                   uint32_t lane, Stage& st, const Out& o) {
    uint32_t j = j0;
    // UNROLL B rows per step: each coordinate arrives as one wide scalar load from a
    // uniform base, and one branch covers the UNROLL hit masks
    for (; j + UNROLL <= j1; j += UNROLL) {
        const double* __restrict__ px = bx_ + size_t(j);
        const double* __restrict__ py = by_ + size_t(j);
        const double* __restrict__ pz = bz_ + size_t(j);
        double d2[UNROLL];
        uint64_t m[UNROLL];
        uint64_t any = 0;
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            d2[u] = chord2(ax, ay, az, px[u], py[u], pz[u]);
            m[u] = __ballot(d2[u] < thr);
            any |= m[u];
        }
        if (any) {
#pragma unroll
            for (int u = 0; u < UNROLL; u++)
                if (m[u]) stage_hits(m[u], d2[u], row, j + u, lane, st, o);
        }
    }
    for (; j < j1; j++) {
        const double d2 = chord2(ax, ay, az, bx_[j], by_[j], bz_[j]);
        const uint64_t m = __ballot(d2 < thr);
        if (m) stage_hits(m, d2, row, j, lane, st, o);
    }
}

// Sparse mode: one B row of [j0, j1) per lane, tested against the chunk rows whose
// ra quantum lies within +-aq of its own (shifted into the chunk's frame).
// s_aq holds the chunk's ra quanta ascending, padded with INT_MAX to 64 entries.
__device__ inline void sparse_range(uint32_t j0, uint32_t j1, int32_t shift, int32_t aq,
                                    uint32_t start, uint32_t cnt, const int32_t* s_aq,
                                    const double* s_ax, const double* s_ay, const double* s_az,
                                    double thr, const uint64_t* __restrict__ bkey,
                                    const double* __restrict__ bx_, const double* __restrict__ by_,
                                    const double* __restrict__ bz_, uint32_t lane, Stage& st,
                                    const Out& o) {
    for (uint32_t jb = j0; jb < j1; jb += WAVE) {
        const uint32_t j = jb + lane;
        const bool valid = j < j1;
        const int32_t qb = valid ? int32_t(bkey[j] & RA_MASK) + shift : 0;
        const int32_t lo_q = qb - aq, hi_q = qb + aq;
        // c = number of chunk rows with quantum < lo_q (branchless search of 64)
        uint32_t c = 0;
#pragma unroll
        for (uint32_t step = WAVE / 2; step >= 1; step >>= 1)
            if (s_aq[c + step - 1] < lo_q) c += step;
        if (s_aq[c] < lo_q) c++;
        bool have_b = false;
        double bx = 0.0, by = 0.0, bz = 0.0;
        for (uint32_t k = 0;; k++) {
            const uint32_t cc = c + k;
            const bool act = valid && cc < cnt && s_aq[cc < uint32_t(WAVE) ? cc : 0] <= hi_q;
            if (!__ballot(act)) break;
            double d2 = 0.0;
            bool hit = false;
            if (act) {
                if (!have_b) {
                    bx = bx_[j];
                    by = by_[j];
                    bz = bz_[j];
                    have_b = true;
                }
                d2 = chord2(s_ax[cc], s_ay[cc], s_az[cc], bx, by, bz);
                hit = d2 < thr;
            }
            const uint64_t m = __ballot(hit);
            if (m) stage_hits(m, d2, start + cc, j, lane, st, o);
        }
    }
}

__global__ __launch_bounds__(PK_THREADS) void k_pair_fp64(
    const Chunk* __restrict__ chunks, uint32_t n_chunks, const uint64_t* __restrict__ akey,
    const double* __restrict__ ax_, const double* __restrict__ ay_, const double* __restrict__ az_,
    const int64_t* __restrict__ aid, const uint64_t* __restrict__ bkey,
    const double* __restrict__ bx_, const double* __restrict__ by_, const double* __restrict__ bz_,
    const int64_t* __restrict__ bid, const uint32_t* __restrict__ bzoff, int32_t bzmin,
    int32_t bzmax, PairParams pp, unsigned long long* __restrict__ counter, uint64_t cap,
    int64_t* __restrict__ out_a, int64_t* __restrict__ out_b, double* __restrict__ out_d) {
    __shared__ uint32_t s_ra[PK_WAVES][STAGE];
    __shared__ uint32_t s_rb[PK_WAVES][STAGE];
    __shared__ double s_d2[PK_WAVES][STAGE];
    __shared__ int32_t s_aq[PK_WAVES][WAVE];
    __shared__ double s_ax[PK_WAVES][WAVE];
    __shared__ double s_ay[PK_WAVES][WAVE];
    __shared__ double s_az[PK_WAVES][WAVE];

    const uint32_t lane = lane_id();
    const uint32_t w = uni(threadIdx.x / WAVE);
    const uint32_t c = xcd_swizzle(blockIdx.x, gridDim.x) * PK_WAVES + w;
    if (c >= n_chunks) return;  // whole wavefront; the kernel has no block barrier

    const uint32_t start = uni(chunks[c].start);
    const uint32_t cnt = uni(chunks[c].count);
    const int32_t za = int32_t(uni(uint32_t(chunks[c].zone)));

    const bool act = lane < cnt;
    const uint32_t row = start + lane;
    double ax = 10.0, ay = 10.0, az = 10.0;  // inactive lanes sit far off the sphere
    int32_t qa = INT_MAX;
    if (act) {
        ax = ax_[row];
        ay = ay_[row];
        az = az_[row];
        qa = int32_t(akey[row] & RA_MASK);
    }
    s_aq[w][lane] = qa;
    s_ax[w][lane] = ax;
    s_ay[w][lane] = ay;
    s_az[w][lane] = az;
    wave_lds_sync();

    const uint64_t q_lo = akey[start] & RA_MASK;
    const uint64_t q_hi = akey[start + cnt - 1] & RA_MASK;
    const int64_t aq = alpha_quanta(pp.theta_deg, pp.h_deg, za, pp.atan_sin_theta);
    const bool dense = chunk_is_dense(q_lo, q_hi, aq, pp.dense_min);

    Stage st{s_ra[w], s_rb[w], s_d2[w], 0};
    const Out o{aid, bid, counter, cap, out_a, out_b, out_d};

    for (int32_t dz = -pp.nz; dz <= pp.nz; dz++) {
        const int32_t zb = za + dz;
        if (zb < 0 || zb >= pp.n_zones || zb < bzmin || zb > bzmax) continue;
        const uint32_t b0 = uni(bzoff[zb - bzmin]);
        const uint32_t b1 = uni(bzoff[zb - bzmin + 1]);
        if (b0 >= b1) continue;
        uint64_t lo[2], hi[2];
        int32_t shift[2];
        const int ni = window_intervals(zb, q_lo, q_hi, aq, lo, hi, shift);
        for (int t = 0; t < ni; t++) {
            const uint32_t j0 = wave_lower_bound(bkey, b0, b1, lo[t], lane);
            const uint32_t j1 = wave_lower_bound(bkey, j0, b1, hi[t] + 1u, lane);
            if (dense)
                dense_range(j0, j1, ax, ay, az, row, pp.chord2, bx_, by_, bz_, lane, st, o);
            else
                sparse_range(j0, j1, shift[t], int32_t(aq), start, cnt, s_aq[w], s_ax[w], s_ay[w],
                             s_az[w], pp.chord2, bkey, bx_, by_, bz_, lane, st, o);
        }
    }
    if (st.n) flush_stage(st, lane, o);
}

}  // namespace

void launch_pair_fp64(xm_ctx* ctx, const xm_catalog* A, const xm_catalog* B, const PairParams& pp,
                      uint64_t* counter, uint64_t cap, int64_t* out_a, int64_t* out_b,
                      double* out_d) {
    if (A->n_chunks == 0 || B->n == 0) return;
    const uint64_t nblk = (A->n_chunks + PK_WAVES - 1) / PK_WAVES;
    launch_timed(ctx, "pair_fp64", [&] {
        hipLaunchKernelGGL(k_pair_fp64, dim3(uint32_t(nblk)), dim3(PK_THREADS), 0, ctx->stream,
                           (const Chunk*)A->chunks, uint32_t(A->n_chunks), A->key, A->x, A->y, A->z,
                           A->id, B->key, B->x, B->y, B->z, B->id, B->zoff, B->zmin, B->zmax, pp,
                           (unsigned long long*)counter, cap, out_a, out_b, out_d);
    });
}

}  // namespace xm
