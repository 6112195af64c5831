// xm_pair.hip -- the zone-join pair kernel, FP64 plain and FP32 filter + FP64 recheck.
//
// Replaces the per-zone-pair launch loop of Worker::Match and xmatch_kernel
// (BoostXmatch/BoostXmatch/Worker.cu:330-359 and :65-241).  The A catalog is cut
// into chunks of up to chunk_rows (64..512, per A segment; 512 by default)
// consecutive rows of one zone; each wavefront owns one
// chunk and joins it with B zones z-nz..z+nz over the chunk's ra span widened by
// alpha(zone) (Common.cu:16-23), the 0/360 seam split in two intervals.  Each
// chunk is joined in one of two modes, chosen per chunk (chunk_is_dense):
//
//  sparse  the chunk's ra quanta are staged in LDS; each lane takes one B row,
//          checks it against the window, binary-searches the chunk for the A rows
//          within +-alpha of its ra and tests only those.  At zone height = radius
//          a zone holds a few hundred sources over 360 deg, so a chunk spans most
//          of the circle while alpha is arcseconds: a dense test of every (A, B)
//          pair of the window would almost always be wasted.  When a zone is one
//          chunk, each B row of its neighbour zones is examined once per A zone;
//          a zone of k chunks examines them k times.
//  dense   (pole rule, windows covering the circle, clustered data) 64 rows at a
//          time, one A row per lane in registers; the window's B rows are walked
//          in wave-uniform order as wide scalar (SGPR) loads, every lane testing
//          its A row against each.
//
// Predicate (Worker.cu:88-97 with the seam bug D2 corrected), on FP64 unit vectors:
//   d2 = fma(dz, dz, fma(dy, dy, dx*dx))  <  (2 sin(r/2))^2
// FP32 filter form: the same chord on FP32 unit vectors keeps a pair when
// d2_32 < (c + band)^2, c = sqrt(chord threshold), band >= twice the proven FP32
// chord error; every kept pair is re-evaluated in FP64 exactly as above when it is
// written, so the set and the distances are those of the FP64 form.
//
// A launch joins a batch of jobs (A segment x B segment x radius, JobDesc): the
// waves are numbered over the concatenation of every job's A chunks, and a wave
// finds its job by binary search of the jobs' wave prefix on the device.  A plain
// join is a batch of one job.  The reference ran one job at a time per GPU thread
// (JobManager.cpp:5-41, Worker.cu:430-485), each as (2nz+1) launches per A zone.
//
// Hits are compacted with ballot + mbcnt into a per-wave LDS stage; a full stage
// (or the end of the chunk) costs one global atomic on the 64-bit append counter
// of the block's output region.  A join's ~10^6 wavefronts spread over up to 1024
// region counters, each on its own 128-byte line: same-address device atomics
// serialise at ~10 ns each, so one shared counter would cost milliseconds.  Slots
// beyond a region's capacity are counted but not written; the host then grows the
// regions and re-runs, so nothing is ever truncated (Worker.cu:147-148, 384-387
// dropped them).  k_compact packs the regions into the contiguous result.
#include <climits>

#include "xm_internal.h"

namespace xm {

namespace {

// wavefronts per block (each works alone); 50M x 50M at 2": 1 -> 1.39 ms, 2 -> 1.39,
// 4 -> 1.43, 8 -> 1.60, 16 -> 1.91
constexpr int PK_WAVES = 2;
constexpr int PK_THREADS = PK_WAVES * WAVE;
constexpr int STAGE = 128;  // two wavefronts of hits (50M x 50M: 64 -> 1.45 ms, 128 -> 1.38, 256 -> 1.44)
constexpr int UNROLL = 4;
constexpr uint32_t ONE_PASS_ROWS = 8 * WAVE;  // B zones up to this size are read in one pass

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
__device__ inline uint32_t wave_lower_bound(const uint32_t* __restrict__ keys, uint32_t lo,
                                            uint32_t hi, uint32_t target, uint32_t lane) {
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

// The chord of the reference's exact test (Worker.cu:93-97) on unit vectors:
// d2 = fma(dz, dz, fma(dy, dy, dx*dx)), in the precision T.  For T = double it is
// the identical computation the host oracle performs.
template <typename T>
__device__ inline T chord2(T ax, T ay, T az, T bx, T by, T bz) {
    const T dx = ax - bx;
    const T dy = ay - by;
    const T dz = az - bz;
    return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}

// The three forms of a join: the test's arithmetic (T), the rows the test reads
// (Row) and whether it is the FP32 filter (kept pairs rechecked in FP64 when
// written).  Fp32Cvt rounds the FP64 rows to FP32 in registers, exactly as the
// FP32 rows are made, so its filter decisions are those of Fp32Rows bit for bit;
// it reads the FP64 rows the recheck needs anyway, where Fp32Rows reads 16-byte
// rows for the test and the FP64 rows of every kept pair on top.
// In Fp32Cvt's sparse path both FP64 rows are in registers at the test, so the
// recheck runs there (REG) and only final pairs are staged.
struct Fp64 {
    using T = double;
    using Row = Row64;
    static constexpr bool F32 = false;
    static constexpr bool REG = false;
};
struct Fp32Rows {
    using T = float;
    using Row = Row32;
    static constexpr bool F32 = true;
    static constexpr bool REG = false;
};
struct Fp32Cvt {
    using T = float;
    using Row = Row64;
    static constexpr bool F32 = true;
    static constexpr bool REG = true;
};

struct Stage {
    uint32_t* ra;
    uint32_t* rb;
    double* d2;  // chord^2 of the test (FP32 form: the filter's, for the band statistics)
    uint32_t n;  // wave-uniform fill
};

struct Out {
    const Row64* __restrict__ arow;
    const Row64* __restrict__ brow;
    unsigned long long* __restrict__ counter;  // this block's region counter line
    unsigned long long* __restrict__ spill;    // the spill pool's counter
    uint64_t base, cap;                        // this block's region slots
    uint64_t spill_base, spill_cap;
    bool wide;
    int64_t* __restrict__ id_a;                // wide slot arrays
    int64_t* __restrict__ id_b;
    double* __restrict__ d2;
    uint32_t* __restrict__ ra;                 // narrow slot arrays
    uint32_t* __restrict__ rb;
    uint32_t* __restrict__ job;                // batch joins: job column
    unsigned long long* __restrict__ job_count;  // batch joins: a partial counter of this job
    double thr;                                // FP64 chord threshold
    double t_lo;                               // FP32 form: below it the filter is sure
    uint32_t jid;
};

// Wave g appends to region g % n_regions: consecutive chunks of one job, which
// carry similar numbers of hits, are dealt to all regions in turn.
__device__ inline Out make_out(const Row64* arow, const Row64* brow, const RegionOut& ro,
                               const PairParams& pp, uint32_t jid, uint32_t g) {
    const uint32_t r = g % ro.n_regions;
    Out o;
    o.arow = arow;
    o.brow = brow;
    o.counter = (unsigned long long*)(ro.counters + size_t(r) * REGION_STRIDE);
    const uint32_t pool = ro.n_regions + r % SPILL_POOLS;
    o.spill = (unsigned long long*)(ro.counters + size_t(pool) * REGION_STRIDE);
    o.base = ro.base[r];
    o.cap = ro.cap[r];
    o.spill_base = ro.base[pool];
    o.spill_cap = ro.spill_cap;
    o.wide = ro.wide;
    o.id_a = ro.id_a;
    o.id_b = ro.id_b;
    o.d2 = ro.d2;
    o.ra = ro.ra;
    o.rb = ro.rb;
    o.job = ro.job;
    o.job_count = ro.job_count
                      ? ro.job_count + size_t(g % JOB_SPREAD) * ro.n_jobs + jid
                      : nullptr;
    o.thr = pp.chord2;
    o.t_lo = double(pp.t_lo32);
    o.jid = jid;
    return o;
}

// Write the wave's staged hits to its region: one atomic per call.  Wide slots get
// (idA, idB, chord^2) -- the rows were just read by the test, so the id loads hit
// in cache -- narrow slots get (A row, B row).  In the FP64 form every staged pair
// is a match.  In the FP32 form each staged pair is re-evaluated in FP64 here and
// dropped if it fails; the band-hit count and the largest |chord_32 - chord_64| go
// to the region's line.
template <bool F32>
__device__ inline void flush_stage(Stage& st, uint32_t lane, const Out& o) {
    wave_lds_sync();
    uint64_t mk[STAGE / WAVE];
    uint32_t total = 0, band = 0;
    double err = 0.0;
#pragma unroll
    for (int k = 0; k < STAGE / WAVE; k++) {
        const uint32_t e = lane + k * WAVE;
        bool keep = e < st.n;
        if (F32 && keep) {
            // the FP64 chord replaces the filter's in the stage (each lane owns its e)
            const Row64 a = o.arow[st.ra[e]];
            const Row64 b = o.brow[st.rb[e]];
            const double d32 = st.d2[e];
            const double d64 = chord2<double>(a.x, a.y, a.z, b.x, b.y, b.z);
            st.d2[e] = d64;
            keep = d64 < o.thr;
            band += d32 >= o.t_lo ? 1u : 0u;
            if (keep) {
                const double e1 = __builtin_fabs(__builtin_sqrt(d32) - __builtin_sqrt(d64));
                err = e1 > err ? e1 : err;
            }
        }
        mk[k] = __ballot(keep);
        total += uint32_t(__popcll(mk[k]));
    }
    // positions [pos, pos + total) of the region's hit stream; those at or past the
    // region's capacity go to the spill pool, reserved with a second atomic
    unsigned long long pos = 0, sp = 0;
    if (lane == 0 && total) {
        pos = atomicAdd(o.counter, (unsigned long long)total);
        if (o.job_count) atomicAdd(o.job_count, (unsigned long long)total);
        const uint64_t first_over = pos > o.cap ? pos : o.cap;
        if (pos + total > o.cap)
            sp = atomicAdd(o.spill, (unsigned long long)(pos + total - first_over));
    }
    pos = __shfl(pos, 0, WAVE);
    sp = __shfl(sp, 0, WAVE);
    const uint64_t first_over = pos > o.cap ? uint64_t(pos) : o.cap;
    uint32_t before = 0;
#pragma unroll
    for (int k = 0; k < STAGE / WAVE; k++) {
        if ((mk[k] >> lane) & 1u) {
            const uint32_t e = lane + k * WAVE;
            const uint64_t p = uint64_t(pos) + before + mbcnt64(mk[k]);
            uint64_t slot = ~uint64_t(0);
            if (p < o.cap) {
                slot = o.base + p;
            } else {
                const uint64_t q = uint64_t(sp) + (p - first_over);
                if (q < o.spill_cap) slot = o.spill_base + q;
            }
            if (slot != ~uint64_t(0)) {
                if (o.wide) {
                    o.id_a[slot] = o.arow[st.ra[e]].id;
                    o.id_b[slot] = o.brow[st.rb[e]].id;
                    o.d2[slot] = st.d2[e];
                } else {
                    o.ra[slot] = st.ra[e];
                    o.rb[slot] = st.rb[e];
                }
                if (o.job) o.job[slot] = o.jid;
            }
        }
        before += uint32_t(__popcll(mk[k]));
    }
    if (F32) {
#pragma unroll
        for (int d = WAVE / 2; d >= 1; d >>= 1) {
            band += __shfl_xor(band, d, WAVE);
            const double e2 = __shfl_xor(err, d, WAVE);
            err = e2 > err ? e2 : err;
        }
        if (lane == 0) {
            if (band) atomicAdd(o.counter + 1, (unsigned long long)band);
            if (err > 0.0)
                atomicMax(o.counter + 2, (unsigned long long)__double_as_longlong(err));
        }
    }
    wave_lds_sync();
    st.n = 0;
}

// Append the lanes of keep mask m (A row ra, B row rb, per lane) to the stage;
// flush when it cannot take another full wave of hits.
template <bool F32>
__device__ inline void stage_hits(uint64_t m, double d2, uint32_t ra, uint32_t rb, uint32_t lane,
                                  Stage& st, const Out& o) {
    if ((m >> lane) & 1u) {
        const uint32_t pos = st.n + mbcnt64(m);
        st.ra[pos] = ra;
        st.rb[pos] = rb;
        st.d2[pos] = d2;
    }
    st.n += uint32_t(__popcll(m));
    if (st.n > uint32_t(STAGE - WAVE)) flush_stage<F32>(st, lane, o);
}

// Blocks are dealt to the 8 XCDs round-robin.  Give each XCD runs of XCD_RUN
// consecutive blocks, the runs dealt in turn: neighbouring chunks, which read the
// same B zones, share an L2, and every XCD gets runs from the whole range (chunks
// are in zone order, and a zone's cost follows its population, which varies with
// declination: contiguous eighths of the range would leave the polar XCDs idle).
constexpr uint32_t XCD_RUN = 16;
__device__ inline uint32_t xcd_swizzle(uint32_t b, uint32_t nb) {
    constexpr uint32_t X = 8, G = XCD_RUN;
    const uint32_t full = nb / (X * G) * (X * G);
    if (b >= full) return b;  // the last partial round keeps its order
    const uint32_t xcd = b % X, k = b / X;
    return (k / G) * (X * G) + xcd * G + k % G;
}

// Output row i of the packed result comes from part r (a region, or a spill pool
// as parts n_regions ..; prefix[r] <= i < prefix[r+1]) at slot base[r] + i -
// prefix[r].  Wide slots (wa != null) hold the ids and chord^2; for narrow slots
// the pair's rows give the ids and the FP64 chord -- the same chord2<double> of the
// same two unit vectors the pair kernel tested, so the bits agree.  chord^2 turns
// into 2 asin(sqrt(d2)/2) in arcsec.  The part prefix is searched in LDS.
constexpr uint32_t COMPACT_ITEMS = 4;  // measured on 30M matches: 1 -> 0.329 ms, 4 -> 0.267, 16 -> 0.275
constexpr uint32_t COMPACT_TPB = 1024; // 30M matches: 256 threads 0.267 ms, 1024 0.257
__global__ __launch_bounds__(COMPACT_TPB) void k_compact(uint32_t n_parts,
                                                 const uint64_t* __restrict__ prefix,
                                                 const uint64_t* __restrict__ base,
                                                 uint64_t total,
                                                 const int64_t* __restrict__ wa,
                                                 const int64_t* __restrict__ wb,
                                                 const double* __restrict__ wd,
                                                 const uint32_t* __restrict__ ta,
                                                 const uint32_t* __restrict__ tb,
                                                 const uint32_t* __restrict__ tj,
                                                 const Row64* __restrict__ arow,
                                                 const Row64* __restrict__ brow,
                                                 int64_t* __restrict__ id_a,
                                                 int64_t* __restrict__ id_b,
                                                 double* __restrict__ dist,
                                                 uint32_t* __restrict__ job) {
    __shared__ uint64_t s_pre[MAX_REGIONS + SPILL_POOLS];
    __shared__ uint64_t s_base[MAX_REGIONS + SPILL_POOLS];
    for (uint32_t r = threadIdx.x; r < n_parts; r += COMPACT_TPB) {
        s_pre[r] = prefix[r];
        s_base[r] = base[r];
    }
    __syncthreads();
    // COMPACT_ITEMS outputs per thread, so each block's load of the part tables
    // serves COMPACT_TPB * COMPACT_ITEMS outputs
    for (uint32_t k = 0; k < COMPACT_ITEMS; k++) {
        const uint64_t i = (uint64_t(blockIdx.x) * COMPACT_ITEMS + k) * COMPACT_TPB + threadIdx.x;
        if (i >= total) return;
        uint32_t lo = 0, hi = n_parts - 1;  // last part with prefix <= i
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) / 2;
            if (s_pre[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        const uint64_t src = s_base[lo] + (i - s_pre[lo]);
        if (wa) {
            id_a[i] = wa[src];
            id_b[i] = wb[src];
            dist[i] = dist_arcsec(wd[src]);
        } else {
            const Row64 a = arow[ta[src]];
            const Row64 b = brow[tb[src]];
            id_a[i] = a.id;
            id_b[i] = b.id;
            dist[i] = dist_arcsec(chord2<double>(a.x, a.y, a.z, b.x, b.y, b.z));
        }
        if (job) job[i] = tj[src];
    }
}

// --------------------------------------------------------------- sparse ----
struct SparseCtx {
    uint32_t start, cnt;
    uint32_t half;            // search width / 2: the power of two >= cnt, halved
    int32_t aq;
    const int32_t* s_aq;      // ra quanta of the chunk rows, padded with INT_MAX
};

// Candidate pairs of the sparse path wait in a per-wavefront LDS ring until a
// full round of 64 can be tested, one pair per lane: the B-row scan finds about
// one candidate per row, so testing in the scan loop left most lanes idle.
constexpr uint32_t CQ = 128;  // ring entries (a round is drained once 64 wait)
struct CandQ {
    uint32_t* a;              // A row of each queued pair
    uint32_t* b;              // B row
    uint32_t head, tail;      // wave-uniform
};

// The FP32 filter's statistics of the pairs a lane rechecked in registers: pairs
// kept with the FP32 chord at or above the band's lower edge, and the largest
// |chord_32 - chord_64| of a match -- as flush_stage counts them for staged pairs.
struct FilterStats {
    uint32_t band = 0;
    double err = 0.0;
};

template <typename P>
__device__ inline void flush_filter_stats(const FilterStats& fs, uint32_t lane, const Out& o) {
    if constexpr (P::REG) {
        uint32_t band = fs.band;
        double err = fs.err;
#pragma unroll
        for (int d = WAVE / 2; d >= 1; d >>= 1) {
            band += __shfl_xor(band, d, WAVE);
            const double e2 = __shfl_xor(err, d, WAVE);
            err = e2 > err ? e2 : err;
        }
        if (lane == 0) {
            if (band) atomicAdd(o.counter + 1, (unsigned long long)band);
            if (err > 0.0)
                atomicMax(o.counter + 2, (unsigned long long)__double_as_longlong(err));
        }
    }
}

template <typename P>
__device__ inline void cand_round(CandQ& q, uint32_t k,
                                  const typename P::Row* __restrict__ arow,
                                  const typename P::Row* __restrict__ brow,
                                  typename P::T thr, uint32_t lane, Stage& st,
                                  const Out& o, FilterStats& fs) {
    using T = typename P::T;
    wave_lds_sync();  // the ring entries written by other lanes
    if constexpr (P::REG) {
        // the FP32 filter on the rows rounded to FP32; the pairs it keeps are
        // decided by the FP64 chord of the same rows, and only matches are staged
        double d64 = 0.0;
        bool hit = false;
        uint32_t ra = 0, rb = 0;
        if (lane < k) {
            const uint32_t s = (q.head + lane) & (CQ - 1);
            ra = q.a[s];
            rb = q.b[s];
            const Row64 a = arow[ra];
            const Row64 b = brow[rb];
            const float d32 = chord2<float>(a.x, a.y, a.z, b.x, b.y, b.z);
            if (d32 < thr) {
                const double dd = double(d32);
                fs.band += dd >= o.t_lo ? 1u : 0u;
                d64 = chord2<double>(a.x, a.y, a.z, b.x, b.y, b.z);
                hit = d64 < o.thr;
                if (hit) {
                    const double e1 = __builtin_fabs(__builtin_sqrt(dd) - __builtin_sqrt(d64));
                    fs.err = e1 > fs.err ? e1 : fs.err;
                }
            }
        }
        q.head += k;
        const uint64_t m = __ballot(hit);
        if (m) stage_hits<false>(m, d64, ra, rb, lane, st, o);
        return;
    }
    T d2 = 0;
    bool hit = false;
    uint32_t ra = 0, rb = 0;
    if (lane < k) {
        const uint32_t s = (q.head + lane) & (CQ - 1);
        ra = q.a[s];
        rb = q.b[s];
        const auto a = arow[ra];
        const auto b = brow[rb];
        d2 = chord2<T>(a.x, a.y, a.z, b.x, b.y, b.z);
        hit = d2 < thr;
    }
    q.head += k;
    const uint64_t m = __ballot(hit);
    if (m) stage_hits<P::F32>(m, double(d2), ra, rb, lane, st, o);
}

// One B row per lane (row j, key), against the chunk rows within +-aq of its ra if
// it lies in one of the window's intervals (ra quanta).  The chunk's
// ra quanta are in LDS: a binary search finds the first, a scan the rest, and each
// candidate (A row, B row) goes to the ring; full rounds are tested at once.
template <typename P>
__device__ inline void sparse_row(const SparseCtx& x, uint32_t j, bool valid, uint32_t key,
                                  const Window& w, CandQ& q,
                                  const typename P::Row* __restrict__ arow,
                                  const typename P::Row* __restrict__ brow,
                                  typename P::T thr, uint32_t lane, Stage& st,
                                  const Out& o, FilterStats& fs) {
    const int32_t qv = int32_t(key);
    bool in = false;
    int32_t qb = 0;
    if (valid) {
        if (qv >= int32_t(w.lo0) && qv <= int32_t(w.hi0)) {
            in = true;
            qb = qv + w.sh0;
        } else if (w.n > 1 && qv >= int32_t(w.lo1) && qv <= int32_t(w.hi1)) {
            in = true;
            qb = qv + w.sh1;
        }
    }
    if (!__ballot(in)) return;
    const int32_t lo_q = qb - x.aq, hi_q = qb + x.aq;
    uint32_t c = 0;
    for (uint32_t step = x.half; step >= 1; step >>= 1)
        if (x.s_aq[c + step - 1] < lo_q) c += step;
    if (x.s_aq[c] < lo_q) c++;
    for (uint32_t k = 0;; k++) {
        const uint32_t cc = c + k;
        const bool act = in && cc < x.cnt && x.s_aq[cc < uint32_t(MAX_CHUNK_ROWS) ? cc : 0] <= hi_q;
        const uint64_t m = __ballot(act);
        if (!m) break;
        if (act) {
            const uint32_t s = (q.tail + mbcnt64(m)) & (CQ - 1);
            q.a[s] = x.start + cc;
            q.b[s] = j;
        }
        q.tail += uint32_t(__popcll(m));
        if (q.tail - q.head >= uint32_t(WAVE))
            cand_round<P>(q, WAVE, arow, brow, thr, lane, st, o, fs);
    }
}

// Order of a wave's B zones.  In phase p (0 .. 2nz) the wave of A zone za joins B
// zone za + dz with (za + dz) = p (mod 2nz + 1): in every phase each B zone that
// is joined at all is joined by all 2nz + 1 of its A zones at once, and the waves
// of neighbouring A zones start together, so they read that zone's rows while the
// first reader's lines are still in the L2.  In zone order (dz = -nz first) the
// three readers of a B zone met it in three different phases.
__device__ inline int32_t phase_dz(int32_t p, int32_t za, int32_t nz) {
    const int32_t m = 2 * nz + 1;
    int32_t r = (p - za) % m;
    if (r < 0) r += m;
    return r <= nz ? r : r - m;
}

// ---------------------------------------------------------------- dense ----
// B rows [j0, j1) against the lane's A row (registers), the rows arriving as wide
// scalar loads from a uniform base, one branch per UNROLL masks.
template <typename P>
__device__ inline void dense_range(uint32_t j0, uint32_t j1, typename P::T ax,
                                   typename P::T ay, typename P::T az,
                                   uint32_t row, typename P::T thr,
                                   const typename P::Row* __restrict__ brow,
                                   uint32_t lane, Stage& st, const Out& o) {
    using T = typename P::T;
    using Row = typename P::Row;
    uint32_t j = j0;
    for (; j + UNROLL <= j1; j += UNROLL) {
        const Row* __restrict__ pb = brow + size_t(j);
        T d2[UNROLL];
        uint64_t m[UNROLL];
        uint64_t any = 0;
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            d2[u] = chord2<T>(ax, ay, az, pb[u].x, pb[u].y, pb[u].z);
            m[u] = __ballot(d2[u] < thr);
            any |= m[u];
        }
        if (any) {
#pragma unroll
            for (int u = 0; u < UNROLL; u++)
                if (m[u]) stage_hits<P::F32>(m[u], double(d2[u]), row, j + u, lane, st, o);
        }
    }
    for (; j < j1; j++) {
        const T d2 = chord2<T>(ax, ay, az, brow[j].x, brow[j].y, brow[j].z);
        const uint64_t m = __ballot(d2 < thr);
        if (m) stage_hits<P::F32>(m, double(d2), row, j, lane, st, o);
    }
}

// ----------------------------------------------------------------- join ----
// One wavefront per (job, A chunk).  The wave total lives on the device
// (job_prefix[n_jobs]); the grid covers an upper bound and the waves past the
// total return at once.  P is the form: the tests read arow_t/brow_t (the FP64
// rows, or for Fp32Rows the catalogs' FP32 rows) and, in the FP32 forms, keep
// pairs below t_hi32; the FP64 rows serve ids and the recheck.
template <typename P>
__global__ __launch_bounds__(PK_THREADS) void k_pair(
    const JobDesc* __restrict__ jobs, const uint32_t* __restrict__ job_prefix, uint32_t n_jobs,
    const Chunk* __restrict__ chunks, const uint32_t* __restrict__ akey,
    const Row64* __restrict__ arow, const typename P::Row* __restrict__ arow_t,
    const uint32_t* __restrict__ bkey, const Row64* __restrict__ brow,
    const typename P::Row* __restrict__ brow_t, const uint32_t* __restrict__ zoff_b,
    RegionOut ro) {
    using T = typename P::T;
    __shared__ uint32_t s_ra[PK_WAVES][STAGE];
    __shared__ uint32_t s_rb[PK_WAVES][STAGE];
    __shared__ double s_d2[PK_WAVES][STAGE];
    __shared__ int32_t s_aq[PK_WAVES][MAX_CHUNK_ROWS];
    __shared__ uint32_t s_qa[PK_WAVES][CQ];
    __shared__ uint32_t s_qb[PK_WAVES][CQ];

    const uint32_t lane = lane_id();
    const uint32_t w = uni(threadIdx.x / WAVE);
    const uint32_t g = xcd_swizzle(blockIdx.x, gridDim.x) * PK_WAVES + w;
    if (g >= job_prefix[n_jobs]) return;  // whole wavefront; the kernel has no block barrier
    // the job: last j with job_prefix[j] <= g (jobs without chunks are skipped over)
    uint32_t jlo = 0, jhi = n_jobs - 1;
    while (jlo < jhi) {
        const uint32_t mid = (jlo + jhi + 1) / 2;
        if (job_prefix[mid] <= g) jlo = mid;
        else jhi = mid - 1;
    }
    const uint32_t jid = uni(jlo);
    const PairParams pp = jobs[jid].pp;
    const int32_t bzmin = jobs[jid].b_zmin, bzmax = jobs[jid].b_zmax;
    const uint32_t* __restrict__ bzoff = zoff_b + jobs[jid].b_zoff_base;
    const uint32_t c = jobs[jid].a_chunk_base + (g - job_prefix[jid]);

    const uint32_t start = uni(chunks[c].start);
    const uint32_t cnt = uni(chunks[c].count);  // 1 .. MAX_CHUNK_ROWS
    const int32_t za = int32_t(uni(uint32_t(chunks[c].zone)));
    // the sparse search probes indices below 2 * half, half the largest power of
    // two below cnt (1 for cnt <= 2): 2 * half <= MAX_CHUNK_ROWS
    uint32_t half = 1;
    while (2 * half < cnt) half *= 2;

    // the chunk's ra quanta in LDS, padded with INT_MAX to 2 * half (>= cnt); the
    // loads of four rows per lane are in flight together
    for (uint32_t r0 = 0; r0 < 2 * half; r0 += 4 * WAVE) {
        int32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t r = r0 + u * WAVE + lane;
            v[u] = r < cnt ? int32_t(akey[start + r]) : INT_MAX;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t r = r0 + u * WAVE + lane;
            if (r < 2 * half) s_aq[w][r] = v[u];
        }
    }
    wave_lds_sync();
    const uint64_t q_lo = uint32_t(s_aq[w][0]);
    const uint64_t q_hi = uint32_t(s_aq[w][cnt - 1]);
    const int64_t aq = join_quanta(pp.theta_deg, pp.h_deg, za, pp.atan_sin_theta, pp.gran);
    const bool dense = chunk_is_dense(q_lo, q_hi, aq, pp.dense_min, cnt);
    const T thr = P::F32 ? T(pp.t_hi32) : T(pp.chord2);
    Stage st{s_ra[w], s_rb[w], s_d2[w], 0};
    const Out o = make_out(arow, brow, ro, pp, jid, g);

    if (dense) {
        // 64 rows at a time, one A row per lane in registers; each run of 64 joins
        // the B rows of its own window
        for (uint32_t sc = 0; sc < cnt; sc += WAVE) {
            const uint32_t sub = cnt - sc < uint32_t(WAVE) ? cnt - sc : uint32_t(WAVE);
            const uint32_t row = start + sc + lane;
            T ax = T(10), ay = T(10), az = T(10);  // inactive lanes sit far off the sphere
            if (lane < sub) {
                const auto a = arow_t[row];
                ax = a.x;
                ay = a.y;
                az = a.z;
            }
            const uint64_t s_lo = uint32_t(s_aq[w][sc]);
            const uint64_t s_hi = uint32_t(s_aq[w][sc + sub - 1]);
            for (int32_t ph = 0; ph <= 2 * pp.nz; ph++) {
                const int32_t zb = za + phase_dz(ph, za, pp.nz);
                if (zb < 0 || zb >= pp.n_zones || zb < bzmin || zb > bzmax) continue;
                const uint32_t b0 = uni(bzoff[zb - bzmin]);
                const uint32_t b1 = uni(bzoff[zb - bzmin + 1]);
                if (b0 >= b1) continue;
                const Window win = window_intervals(s_lo, s_hi, aq);
                for (int t = 0; t < win.n; t++) {
                    const Window wt = window_part(win, t);
                    const uint32_t j0 = wave_lower_bound(bkey, b0, b1, wt.lo0, lane);
                    const uint32_t j1 = wave_lower_bound(bkey, j0, b1, wt.hi0 + 1u, lane);
                    dense_range<P>(j0, j1, ax, ay, az, row, thr, brow_t, lane, st, o);
                }
            }
        }
        if (st.n) flush_stage<P::F32>(st, lane, o);
        return;
    }

    const SparseCtx x{start, cnt, half, int32_t(aq), s_aq[w]};
    CandQ q{s_qa[w], s_qb[w], 0, 0};
    FilterStats fs;
    for (int32_t ph = 0; ph <= 2 * pp.nz; ph++) {
        const int32_t zb = za + phase_dz(ph, za, pp.nz);
        if (zb < 0 || zb >= pp.n_zones || zb < bzmin || zb > bzmax) continue;
        const uint32_t b0 = uni(bzoff[zb - bzmin]);
        const uint32_t b1 = uni(bzoff[zb - bzmin + 1]);
        if (b0 >= b1) continue;
        const Window win = window_intervals(q_lo, q_hi, aq);
        if (b1 - b0 <= ONE_PASS_ROWS) {
            // the whole zone in one round trip; rows outside the window drop out
            uint32_t kk[ONE_PASS_ROWS / WAVE];
#pragma unroll
            for (uint32_t k = 0; k < ONE_PASS_ROWS / WAVE; k++) {
                const uint32_t j = b0 + k * WAVE + lane;
                kk[k] = j < b1 ? bkey[j] : 0;
            }
#pragma unroll 1
            for (uint32_t k = 0; k < ONE_PASS_ROWS / WAVE; k++) {
                if (b0 + k * WAVE >= b1) break;
                const uint32_t j = b0 + k * WAVE + lane;
                sparse_row<P>(x, j, j < b1, kk[k], win, q, arow_t, brow_t, thr, lane, st, o, fs);
            }
        } else {
            for (int t = 0; t < win.n; t++) {
                const Window wt = window_part(win, t);
                const uint32_t j0 = wave_lower_bound(bkey, b0, b1, wt.lo0, lane);
                const uint32_t j1 = wave_lower_bound(bkey, j0, b1, wt.hi0 + 1u, lane);
                for (uint32_t jb = j0; jb < j1; jb += WAVE) {
                    const uint32_t j = jb + lane;
                    const bool valid = j < j1;
                    const uint32_t key = valid ? bkey[j] : 0;
                    sparse_row<P>(x, j, valid, key, wt, q, arow_t, brow_t, thr, lane, st, o, fs);
                }
            }
        }
    }
    while (q.tail != q.head) {
        const uint32_t k = q.tail - q.head < uint32_t(WAVE) ? q.tail - q.head : uint32_t(WAVE);
        cand_round<P>(q, k, arow_t, brow_t, thr, lane, st, o, fs);
    }
    // staged pairs are final where the recheck ran in registers
    if (st.n) flush_stage<P::F32 && !P::REG>(st, lane, o);
    flush_filter_stats<P>(fs, lane, o);
}

// One block.  Parts: the n_regions regions (kept = min(count, cap)), then the
// SPILL_POOLS pools (kept = min(demand, spill_cap)); prefix = exclusive scan of
// kept over the parts; tail[0] = sum of the region counts (every hit), tail[1] =
// the largest pool demand, tail[2] = band hits, tail[3] = the largest chord error
// (double bits); next_base = exclusive scan of the region counts, then the pools
// (empty, at the total), next_cap = the region counts.
__global__ __launch_bounds__(1024) void k_region_scan(const uint64_t* __restrict__ counters,
                                                      uint32_t n_regions,
                                                      const uint64_t* __restrict__ cap,
                                                      uint64_t spill_cap,
                                                      uint64_t* __restrict__ prefix,
                                                      uint64_t* __restrict__ tail,
                                                      uint64_t* __restrict__ next_base,
                                                      uint64_t* __restrict__ next_cap) {
    __shared__ uint64_t s_kept[1024];
    __shared__ uint64_t s_cnt[1024];
    __shared__ uint64_t s_band[1024];
    __shared__ uint64_t s_err[1024];
    __shared__ uint64_t s_pool[SPILL_POOLS];
    __shared__ uint64_t s_pmax[SPILL_POOLS];
    const uint32_t t = threadIdx.x;
    const bool v = t < n_regions;
    const uint64_t c = v ? counters[size_t(t) * REGION_STRIDE] : 0;
    const uint64_t kept = v ? (c < cap[t] ? c : cap[t]) : 0;
    s_kept[t] = kept;
    s_cnt[t] = c;
    s_band[t] = v ? counters[size_t(t) * REGION_STRIDE + 1] : 0;
    s_err[t] = v ? counters[size_t(t) * REGION_STRIDE + 2] : 0;
    if (t < SPILL_POOLS) {
        const uint64_t d = counters[size_t(n_regions + t) * REGION_STRIDE];
        s_pool[t] = d < spill_cap ? d : spill_cap;
        s_pmax[t] = d;
    }
    __syncthreads();
    // Hillis-Steele inclusive scans of kept and of count; tree reductions of the rest
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint64_t x = t >= d ? s_kept[t - d] : 0;
        const uint64_t y = t >= d ? s_cnt[t - d] : 0;
        __syncthreads();
        s_kept[t] += x;
        s_cnt[t] += y;
        __syncthreads();
    }
    if (v) {
        prefix[t] = s_kept[t] - kept;
        next_base[t] = s_cnt[t] - c;
        next_cap[t] = c;
    }
    __syncthreads();
    for (uint32_t d = 512; d >= 1; d >>= 1) {
        if (t < d) {
            if (s_err[t + d] > s_err[t]) s_err[t] = s_err[t + d];
            s_band[t] += s_band[t + d];
            if (t < SPILL_POOLS && t + d < SPILL_POOLS && s_pmax[t + d] > s_pmax[t])
                s_pmax[t] = s_pmax[t + d];
        }
        __syncthreads();
    }
    if (t == 0) {
        uint64_t run = s_kept[1023];
        for (uint32_t p = 0; p < SPILL_POOLS; p++) {
            prefix[n_regions + p] = run;
            run += s_pool[p];
            next_base[n_regions + p] = s_cnt[1023];
        }
        prefix[n_regions + SPILL_POOLS] = run;
        tail[0] = s_cnt[1023];
        tail[1] = s_pmax[0];
        tail[2] = s_band[0];
        tail[3] = s_err[0];
    }
}

__global__ __launch_bounds__(256) void k_region_init(uint32_t n_regions, uint64_t cap_each,
                                                     uint64_t pool_cap,
                                                     uint64_t* __restrict__ base,
                                                     uint64_t* __restrict__ cap) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n_regions + SPILL_POOLS) return;
    if (r < n_regions) {
        base[r] = uint64_t(r) * cap_each;
        cap[r] = cap_each;
    } else {
        base[r] = uint64_t(n_regions) * cap_each + uint64_t(r - n_regions) * pool_cap;
    }
}

__global__ __launch_bounds__(256) void k_job_sum(const unsigned long long* __restrict__ partial,
                                                 uint32_t n_jobs, uint64_t* __restrict__ out) {
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n_jobs) return;
    uint64_t s = 0;
    for (uint32_t p = 0; p < JOB_SPREAD; p++) s += partial[size_t(p) * n_jobs + j];
    out[j] = s;
}

__global__ __launch_bounds__(256) void k_job_prep(JobDesc* __restrict__ jobs, uint32_t n_jobs,
                                                  const uint32_t* __restrict__ cstart,
                                                  uint32_t* __restrict__ waves) {
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n_jobs) return;
    const uint32_t c0 = cstart[jobs[j].a_zoff_lo], c1 = cstart[jobs[j].a_zoff_hi];
    jobs[j].a_chunk_base = c0;
    waves[j] = jobs[j].b_zmin <= jobs[j].b_zmax ? c1 - c0 : 0u;  // empty B: no work
}

// chord^2 -> distance in arcsec, in place: the same conversion as k_compact's.
__global__ __launch_bounds__(256) void k_chord_to_arcsec(double* __restrict__ d, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) d[i] = dist_arcsec(d[i]);
}

}  // namespace

void launch_compact(xm_ctx* ctx, const RegionOut& ro, const uint64_t* prefix, uint64_t total,
                    const xm_catalog* A, const xm_catalog* B, int64_t* id_a, int64_t* id_b,
                    double* dist, uint32_t* job) {
    if (total == 0) return;
    launch_timed(ctx, "compact", [&] {
        hipLaunchKernelGGL(k_compact,
                           dim3(uint32_t((total + COMPACT_TPB * COMPACT_ITEMS - 1) /
                                         (COMPACT_TPB * COMPACT_ITEMS))),
                           dim3(COMPACT_TPB), 0,
                           ctx->stream, ro.n_regions + SPILL_POOLS, prefix, ro.base, total,
                           (const int64_t*)(ro.wide ? ro.id_a : nullptr),
                           (const int64_t*)ro.id_b, (const double*)ro.d2,
                           (const uint32_t*)ro.ra, (const uint32_t*)ro.rb,
                           (const uint32_t*)ro.job, static_cast<const Row64*>(A->row),
                           static_cast<const Row64*>(B->row), id_a, id_b, dist,
                           ro.job ? job : nullptr);
    });
}

void launch_region_init(xm_ctx* ctx, uint32_t n_regions, uint64_t cap_each, uint64_t pool_cap,
                        uint64_t* base, uint64_t* cap) {
    launch_timed(ctx, "fill", [&] {
        hipLaunchKernelGGL(k_region_init, dim3((n_regions + SPILL_POOLS + 255) / 256), dim3(256),
                           0, ctx->stream, n_regions, cap_each, pool_cap, base, cap);
    });
}

void launch_job_sum(xm_ctx* ctx, const unsigned long long* partial, uint32_t n_jobs,
                    uint64_t* job_counts) {
    if (n_jobs == 0) return;
    launch_timed(ctx, "job_sum", [&] {
        hipLaunchKernelGGL(k_job_sum, dim3((n_jobs + 255) / 256), dim3(256), 0, ctx->stream,
                           partial, n_jobs, job_counts);
    });
}

uint32_t pair_blocks(uint64_t n_waves) { return uint32_t((n_waves + PK_WAVES - 1) / PK_WAVES); }

void launch_job_prep(xm_ctx* ctx, JobDesc* jobs, uint32_t n_jobs, const uint32_t* cstart,
                     uint32_t* waves) {
    if (n_jobs == 0) return;
    launch_timed(ctx, "job_prep", [&] {
        hipLaunchKernelGGL(k_job_prep, dim3((n_jobs + 255) / 256), dim3(256), 0, ctx->stream, jobs,
                           n_jobs, cstart, waves);
    });
}

void launch_region_scan(xm_ctx* ctx, const RegionOut& ro, uint64_t* prefix, uint64_t* tail,
                        uint64_t* next_base, uint64_t* next_cap) {
    launch_timed(ctx, "region_scan", [&] {
        hipLaunchKernelGGL(k_region_scan, dim3(1), dim3(1024), 0, ctx->stream,
                           (const uint64_t*)ro.counters, ro.n_regions, ro.cap, ro.spill_cap,
                           prefix, tail, next_base, next_cap);
    });
}

void launch_chord_to_arcsec(xm_ctx* ctx, double* d, uint64_t n) {
    if (n == 0) return;
    launch_timed(ctx, "chord_to_arcsec", [&] {
        hipLaunchKernelGGL(k_chord_to_arcsec, dim3(uint32_t((n + 255) / 256)), dim3(256), 0,
                           ctx->stream, d, n);
    });
}


void launch_pair(xm_ctx* ctx, const xm_catalog* A, const xm_catalog* B, const Chunk* chunks,
                 const JobDesc* jobs, const uint32_t* job_prefix, uint32_t n_jobs,
                 uint64_t max_waves, bool f32, const RegionOut& ro) {
    if (B->n == 0 || A->n == 0 || n_jobs == 0 || max_waves == 0) return;
    const dim3 grid(pair_blocks(max_waves));
    const Row64* arow = static_cast<const Row64*>(A->row);
    const Row64* brow = static_cast<const Row64*>(B->row);
    if (f32 && A->row32 && B->row32) {
        launch_timed(ctx, "pair_fp32", [&] {
            hipLaunchKernelGGL(k_pair<Fp32Rows>, grid, dim3(PK_THREADS), 0, ctx->stream, jobs,
                               job_prefix, n_jobs, chunks, A->key, arow,
                               static_cast<const Row32*>(A->row32), B->key, brow,
                               static_cast<const Row32*>(B->row32), B->zoff, ro);
        });
    } else if (f32) {
        launch_timed(ctx, "pair_fp32", [&] {
            hipLaunchKernelGGL(k_pair<Fp32Cvt>, grid, dim3(PK_THREADS), 0, ctx->stream, jobs,
                               job_prefix, n_jobs, chunks, A->key, arow, arow, B->key, brow, brow,
                               B->zoff, ro);
        });
    } else {
        launch_timed(ctx, "pair_fp64", [&] {
            hipLaunchKernelGGL(k_pair<Fp64>, grid, dim3(PK_THREADS), 0, ctx->stream, jobs,
                               job_prefix, n_jobs, chunks, A->key, arow, arow, B->key, brow, brow,
                               B->zoff, ro);
        });
    }
}

}  // namespace xm
