// xm_pair.hip -- the zone-join pair kernel (FP64 plain form).
//
// Replaces the per-zone-pair launch loop of Worker::Match and xmatch_kernel
// (BoostXmatch/BoostXmatch/Worker.cu:330-359 and :65-241).  One launch covers the
// whole join: each wavefront owns a chunk of up to 64 consecutive A rows of one
// zone (one row per lane, unit vector in registers).  For each B zone z-nz..z+nz
// it finds, on the sorted packed keys, the rows whose ra lies in the chunk's ra
// span widened by alpha(zone) (Common.cu:16-23), with the 0/360 seam split in two
// intervals, and walks them in wave-uniform order: the B unit vector is a scalar
// (SGPR) operand and every lane tests its A row against it.
//
// Predicate (Worker.cu:88-97 with the seam bug D2 corrected):
//   d2 = fma(dz, dz, fma(dy, dy, dx*dx))  <  (2 sin(r/2))^2
// Hits are compacted with ballot + mbcnt into a per-wave LDS stage; a full stage
// (or the end of the chunk) costs one global atomic on the 64-bit append counter.
// Slots beyond the buffer capacity are counted but not written; the host then
// grows the buffer and re-runs, so nothing is ever truncated (Worker.cu:147-148,
// 384-387 dropped them).
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

// Write the wave's staged hits to the result buffer: one atomic per call.
__device__ inline void flush_stage(uint32_t nst, uint32_t lane, const uint32_t* s_ra,
                                   const uint32_t* s_rb, const double* s_d2,
                                   const int64_t* __restrict__ aid,
                                   const int64_t* __restrict__ bid,
                                   unsigned long long* __restrict__ counter, uint64_t cap,
                                   int64_t* __restrict__ out_a, int64_t* __restrict__ out_b,
                                   double* __restrict__ out_d) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(counter, (unsigned long long)nst);
    base = __shfl(base, 0, WAVE);
    for (uint32_t e = lane; e < nst; e += WAVE) {
        const uint64_t o = uint64_t(base) + e;
        if (o < cap) {
            out_a[o] = aid[s_ra[e]];
            out_b[o] = bid[s_rb[e]];
            out_d[o] = dist_arcsec(s_d2[e]);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The chord of the reference's exact test (Worker.cu:93-97) on FP64 unit vectors:
// d2 = fma(dz, dz, fma(dy, dy, dx*dx)), identical on host and device.
__device__ inline double chord2(double ax, double ay, double az, double bx, double by, double bz) {
    const double dx = ax - bx;
    const double dy = ay - by;
    const double dz = az - bz;
    return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}

// Append the lanes of hit mask m (A row per lane, B row j) to the wave's stage;
// flush it to the result buffer when it cannot take another full wave of hits.
__device__ inline void stage_hits(uint64_t m, double d2, uint32_t row, uint32_t j, uint32_t lane,
                                  uint32_t& nst, uint32_t* st_ra, uint32_t* st_rb, double* st_d2,
                                  const int64_t* __restrict__ aid, const int64_t* __restrict__ bid,
                                  unsigned long long* __restrict__ counter, uint64_t cap,
                                  int64_t* __restrict__ out_a, int64_t* __restrict__ out_b,
                                  double* __restrict__ out_d) {
    if ((m >> lane) & 1u) {
        const uint32_t pos = nst + mbcnt64(m);
        st_ra[pos] = row;
        st_rb[pos] = j;
        st_d2[pos] = d2;
    }
    nst += uint32_t(__popcll(m));
    if (nst > uint32_t(STAGE - WAVE)) {
        flush_stage(nst, lane, st_ra, st_rb, st_d2, aid, bid, counter, cap, out_a, out_b, out_d);
        nst = 0;
    }
}

// Blocks are dealt to the 8 XCDs round-robin; give each XCD a contiguous run of
// chunks so that neighbouring chunks, which read the same B zones, share an L2.
__device__ inline uint32_t xcd_swizzle(uint32_t b, uint32_t nb) {
    const uint32_t per = nb / 8, rem = nb % 8, xcd = b % 8, idx = b / 8;
    return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
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
    if (act) {
        ax = ax_[row];
        ay = ay_[row];
        az = az_[row];
    }
    const uint64_t q_lo = akey[start] & RA_MASK;
    const uint64_t q_hi = akey[start + cnt - 1] & RA_MASK;
    const int64_t aq = alpha_quanta(pp.theta_deg, pp.h_deg, za, pp.atan_sin_theta);

    uint32_t* st_ra = s_ra[w];
    uint32_t* st_rb = s_rb[w];
    double* st_d2 = s_d2[w];
    uint32_t nst = 0;

    for (int32_t dz = -pp.nz; dz <= pp.nz; dz++) {
        const int32_t zb = za + dz;
        if (zb < 0 || zb >= pp.n_zones || zb < bzmin || zb > bzmax) continue;
        const uint32_t b0 = uni(bzoff[zb - bzmin]);
        const uint32_t b1 = uni(bzoff[zb - bzmin + 1]);
        if (b0 >= b1) continue;
        uint64_t lo[2], hi[2];
        const int ni = window_intervals(zb, q_lo, q_hi, aq, lo, hi);
        for (int t = 0; t < ni; t++) {
            const uint32_t j0 = wave_lower_bound(bkey, b0, b1, lo[t], lane);
            const uint32_t j1 = wave_lower_bound(bkey, j0, b1, hi[t] + 1u, lane);
            uint32_t j = j0;
            // UNROLL B rows per step: each coordinate arrives as one wide scalar load
            // from a uniform base, and one branch covers the UNROLL hit masks
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
                    m[u] = __ballot(d2[u] < pp.chord2);
                    any |= m[u];
                }
                if (any) {
#pragma unroll
                    for (int u = 0; u < UNROLL; u++)
                        if (m[u]) stage_hits(m[u], d2[u], row, j + u, lane, nst, st_ra, st_rb,
                                             st_d2, aid, bid, counter, cap, out_a, out_b, out_d);
                }
            }
            for (; j < j1; j++) {
                const double d2 = chord2(ax, ay, az, bx_[j], by_[j], bz_[j]);
                const uint64_t m = __ballot(d2 < pp.chord2);
                if (m) stage_hits(m, d2, row, j, lane, nst, st_ra, st_rb, st_d2, aid, bid, counter,
                                  cap, out_a, out_b, out_d);
            }
        }
    }
    if (nst) flush_stage(nst, lane, st_ra, st_rb, st_d2, aid, bid, counter, cap, out_a, out_b, out_d);
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
