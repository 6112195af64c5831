/*
 * xmatch.h -- C ABI of the MI355X zone-algorithm crossmatch (libxmatch_amd.so).
 *
 * Catalog segments of (id, ra, dec) go in; (idA, idB, distance) pairs within a
 * search radius come out.  The method is the zone algorithm of BoostXmatch
 * (matthiaslee/Xmatch): declination zones of height h, sources sorted by
 * (zone, ra), each A zone joined with B zones z-nz..z+nz over an ra window
 * widened by alpha(dec), then an exact chord test.
 *
 * Every entry point returns xm_status.  Nothing is truncated: when the result
 * buffer overflows the library grows it and re-runs the join.  Plain pointers
 * and sizes only; device pointers are HIP device addresses on the context's GPU.
 *
 * Threading: calls on one xm_ctx are serialised by an internal mutex; all device
 * work of a context runs on its one HIP stream.  Different contexts are
 * independent.
 */
#ifndef XMATCH_H
#define XMATCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XM_ABI_VERSION 1

typedef enum {
    XM_OK = 0,
    XM_EINVAL = 1,    /* bad argument or bad input record (NaN, dec outside [-90,90], ra outside [0,360)) */
    XM_ENOMEM = 2,    /* device or host allocation failed */
    XM_EDEVICE = 3,   /* no usable GPU, or a HIP call failed */
    XM_EINTERNAL = 4,
    XM_EAGAIN = 5     /* xm_server_poll: the request is not done yet */
} xm_status;

typedef enum {
    XM_FP64_PLAIN = 0,              /* FP64 unit vectors, FP64 chord test on every candidate */
    XM_FP32_FILTER_FP64_RECHECK = 1 /* FP32 chord test with a guard band; every pair it keeps
                                       is decided and measured again in FP64 */
} xm_precision;

/* One input record, 24 bytes, packed little-endian.
 * Replaces xmatch::Obj (BoostXmatch/BoostXmatch/Obj.h:21-25). */
typedef struct {
    int64_t id;
    double ra_deg;
    double dec_deg;
} xm_obj;

typedef struct xm_ctx xm_ctx;
typedef struct xm_catalog xm_catalog;
typedef struct xm_result xm_result;

/* One segment of records for a many-segment catalog: the reference's Segment
 * (BoostXmatch/BoostXmatch/Segment.h:22-41) before Sorter::Sort, i.e. a run of
 * input records and the zone height it is indexed with. */
typedef struct {
    const xm_obj *records;   /* host or device pointer, per the creating call */
    uint64_t n;
    double zone_height_deg;
} xm_segment;

/* One join of a batch: segment seg_a of the A catalog against segment seg_b of
 * the B catalog within radius_deg.  Replaces xmatch::Job (Job.h:16-29), which
 * paired two Segment pointers with the search radius. */
typedef struct {
    uint32_t seg_a;
    uint32_t seg_b;
    double radius_deg;
} xm_job;

/* Work and timing figures of one xm_match call. */
typedef struct {
    uint64_t n_matches;      /* pairs in the result */
    uint64_t n_chunks;       /* A row chunks (one wavefront each) */
    uint64_t n_retries;      /* result-buffer overflows that forced a re-run */
    uint64_t band_hits;      /* FP32 filter only: pairs inside the guard band (FP64 rechecked) */
    double join_ms;          /* device time of the join (HIP events), including re-runs */
    double max_dist_err_arcsec; /* FP32 filter only: largest |chord_32 - chord_64| of a
                                   match, in arcsec (the result holds the FP64 distance) */
} xm_match_stats;

/* ---- context: replaces CudaManager/CudaContext (CudaManager.cpp:8-65,
 *      CudaContext.cpp:5-27) with one explicit per-device handle. */
xm_status xm_ctx_create(int device, xm_ctx **out);
void xm_ctx_destroy(xm_ctx *ctx);
/* Message of the last failed call on this context (never NULL). */
const char *xm_last_error_message(const xm_ctx *ctx);
/* Per-kernel device timing with HIP events on the context stream (0 = off). */
xm_status xm_ctx_set_timing(xm_ctx *ctx, int enable);
/* While timing is on, record only the launches of kernel `name` (NULL or "": all).
 * Each timed launch adds two event records to the stream: the 50M x 50M join step,
 * when it made 45 launches, took 7.94 ms with all of them timed and 7.56 ms with
 * none on MI355X.  A caller that times one kernel inside a measured loop names it. */
xm_status xm_ctx_set_timing_only(xm_ctx *ctx, const char *name);
/* Sum and count of recorded durations for kernel `name` since the last reset. */
xm_status xm_ctx_kernel_time(xm_ctx *ctx, const char *name, double *total_ms, uint64_t *count);
xm_status xm_ctx_reset_timing(xm_ctx *ctx);
/* Tuning knobs of the context (no counterpart in the reference, whose kernel
 * shape was fixed at compile time, Worker.cu:22,351-353):
 *   "dense_min"  an A chunk (up to chunk_rows rows of one zone, 512 by default) is
 *                joined densely (every B row of its window against every chunk row)
 *                when a B row
 *                is expected to meet at least this many chunk rows, else sparsely
 *                (each B row against the chunk rows within the ra window).
 *                0 = always dense.  Default 16.
 *   "key_ra_bits" upper bound (1..24) on the ra bits kept in a catalog's sort key
 *                (applied at catalog creation; large catalogs keep fewer, so that
 *                zone | ra | row index fits 64 bits).  Default 24.
 *   "sort_wide_words" 1 = the catalog sort's tile look-back uses 64-bit words at
 *                every size (by default only from 2^30 rows up; 32-bit words
 *                below).  Default 0.
 *   "chunk_rows", "autotune": the pair kernel's shape, see xm_ctx_tune_table.
 *   "fp32_rows"  1 = catalog creation also writes each row's FP32 unit vector (16
 *                bytes a row, in the same pass as the FP64 rows).  The FP32 filter
 *                (XM_FP32_FILTER_FP64_RECHECK) of two catalogs that carry them reads
 *                them; otherwise it rounds the FP64 rows to FP32 in registers, the
 *                same values, and reads no more than the FP64 form.  Default 0.
 *   "bucket_min_rows" one-segment catalogs of at least this many rows are built
 *                with their records moved into zone buckets by the key pass, so that
 *                the gather reads them from one XCD's L2.  0 = never.  Default 2^24.
 *   "bucket_capacity_pct" a bucket's runs hold this % (1..1000) of its sampled
 *                estimate; records past a full run go to an overflow run, and a
 *                build whose overflow run fills is redone without buckets.
 *                Default 115.
 *   "sort_fail_test" 1 = every catalog sort of more than one tile (8192 rows)
 *                fails through its look-back's bounded wait, and catalog creation
 *                returns XM_EINTERNAL: for tests of that error path.  Default 0.
 * Results never depend on the knobs (but "sort_fail_test"); only the work done
 * does. */
xm_status xm_ctx_set_param(xm_ctx *ctx, const char *name, int64_t value);
xm_status xm_ctx_get_param(xm_ctx *ctx, const char *name, int64_t *value);
/* The pair kernel's shape per job: "chunk_rows" (A rows per wavefront, a multiple
 * of 64 up to 512; 0 = 512) and "dense_min" above are the shape every join uses,
 * unless "autotune" is 1.  Then a join first looks up each job's tuning key --
 * log2 rows per occupied zone of its A and B segments (rounded) and nz =
 * ceil(r / h) -- in the context's tune table.  For a key the table lacks, it times
 * count-only pair launches of up to 16 of the batch's jobs with that key (the
 * largest, repeated until the launch fills the GPU) for chunk rows {64, 128,
 * 256, 512} x dense_min {4, 16, 64} and
 * keeps the fastest shape if it beats the default shape by more than 3%.  The reference's kernel shape was fixed at compile time
 * (Worker.cu:22,351-353).  Results never depend on the shape. */
typedef struct {
    uint32_t key;          /* log2 rows/zone of A | log2 rows/zone of B << 8 | nz << 16 */
    uint32_t chunk_rows;   /* chosen A rows per chunk */
    uint32_t dense_min;    /* chosen dense rule */
    uint32_t sample_jobs;  /* jobs the shapes were timed on */
    double tuned_ms;       /* pair kernel time of the chosen shape on the sample */
    double default_ms;     /* the same for the default shape (the context's knobs) */
} xm_tune_entry;
/* The tune table: up to cap entries into out (may be NULL), their count in *n. */
xm_status xm_ctx_tune_table(xm_ctx *ctx, xm_tune_entry *out, uint64_t cap, uint64_t *n);
/* Free and total device memory; blocks the context caches for reuse count as free. */
xm_status xm_ctx_mem_info(xm_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);
/* Wait for all device work of the context. */
xm_status xm_ctx_synchronize(xm_ctx *ctx);
/* The context keeps the device memory of finished catalogs, results and scratch
 * for reuse.  This returns every cached block that is not in use to the device;
 * cached_bytes (may be NULL) receives what the context still holds. */
xm_status xm_ctx_release_memory(xm_ctx *ctx, uint64_t *cached_bytes);
/* Device malloc/free through the context's memory cache (for callers that stage
 * records on the device). */
xm_status xm_device_alloc(xm_ctx *ctx, size_t bytes, void **dptr);
xm_status xm_device_free(xm_ctx *ctx, void *dptr);
/* Page-locked host memory (hipHostMalloc): copies from it reach the device at
 * full link speed.  For callers that stream records from files. */
xm_status xm_host_alloc(xm_ctx *ctx, size_t bytes, void **hptr);
xm_status xm_host_free(xm_ctx *ctx, void *hptr);
xm_status xm_memcpy_h2d(xm_ctx *ctx, void *dst, const void *src, size_t bytes);
xm_status xm_memcpy_d2h(xm_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- segment preparation: replaces Sorter::Sort (BoostXmatch/BoostXmatch/
 *      Sorter.cu:141-185) and Segment::Load's in-memory result (Segment.cpp:14-22).
 * Builds a device-resident, zone-sorted, zone-indexed catalog.  zone_height_deg
 * is the reference's zh_deg (BoostXmatch.cpp:158,271: default = radius).
 * The host variant copies the records to the device first.  n <= 2^32-8193
 * (XM_EINVAL above). */
xm_status xm_catalog_create(xm_ctx *ctx, const xm_obj *records, uint64_t n,
                            double zone_height_deg, xm_catalog **out);
xm_status xm_catalog_create_device(xm_ctx *ctx, const xm_obj *d_records, uint64_t n,
                                   double zone_height_deg, xm_catalog **out);
/* Many segments in one catalog, prepared together by one sort (the reference ran
 * one Sorter::Sort per segment, BoostXmatch.cpp:250-263,284-297).  Each segment
 * keeps its own zone height, zone table and rows; rows of segment s follow those
 * of segment s-1.  At most 65536 segments and 2^32-8193 records in all.  The
 * _device variant takes device record pointers; the other copies host records. */
xm_status xm_catalog_create_segments(xm_ctx *ctx, const xm_segment *segs, uint64_t n_segs,
                                     xm_catalog **out);
xm_status xm_catalog_create_segments_device(xm_ctx *ctx, const xm_segment *segs,
                                            uint64_t n_segs, xm_catalog **out);
void xm_catalog_destroy(xm_catalog *cat);
/* n (all segments), zone height (of segment 0), occupied zone range [zmin, zmax]
 * over all segments (zmin > zmax when empty). */
xm_status xm_catalog_info(const xm_catalog *cat, uint64_t *n, double *zone_height_deg,
                          int32_t *zmin, int32_t *zmax);
xm_status xm_catalog_segment_count(const xm_catalog *cat, uint64_t *n_segs);
/* Segment seg: its first row in the catalog's sorted order, rows, zone height and
 * occupied zone range.  Any output pointer may be NULL. */
xm_status xm_catalog_segment_info(const xm_catalog *cat, uint64_t seg, uint64_t *first_row,
                                  uint64_t *n, double *zone_height_deg, int32_t *zmin,
                                  int32_t *zmax);
/* Sorted-order export to host arrays of length n (any pointer may be NULL):
 * ids, FP64 unit vectors, zone ids.  For tests and tooling. */
xm_status xm_catalog_export(xm_ctx *ctx, const xm_catalog *cat, int64_t *id, double *x,
                            double *y, double *z, int32_t *zone);

/* ---- join: replaces Worker::Match (BoostXmatch/BoostXmatch/Worker.cu:243-428)
 *      and xmatch_kernel (Worker.cu:65-241).
 * radius_deg is Job::sr_deg (Job.h:21).  capacity_hint sizes the first result
 * buffer (0 = 2*max(nA,nB), as the reference's maxout default, Worker.cu:295-300);
 * overflow never truncates.  stats may be NULL. */
xm_status xm_match(xm_ctx *ctx, const xm_catalog *a, const xm_catalog *b, double radius_deg,
                   xm_precision precision, uint64_t capacity_hint, xm_result **out,
                   xm_match_stats *stats);
/* A batch of joins in one launch: job j joins segment jobs[j].seg_a of a with
 * segment jobs[j].seg_b of b (both of the same zone height) within
 * jobs[j].radius_deg.  Replaces JobManager's queue of Jobs, each run by
 * Worker::Match (JobManager.cpp:5-41, Worker.cu:430-485).  The result holds the
 * pairs of every job (unordered, like xm_match); xm_result_jobs gives the job of
 * each pair and the count per job.  capacity_hint 0 = sum over jobs of
 * 2 * max(nA, nB).  a may equal b. */
xm_status xm_match_batch(xm_ctx *ctx, const xm_catalog *a, const xm_catalog *b,
                         const xm_job *jobs, uint64_t n_jobs, xm_precision precision,
                         uint64_t capacity_hint, xm_result **out, xm_match_stats *stats);
/* xm_match without device allocation: the pairs of a join of two one-segment
 * catalogs go straight into the caller's device arrays d_id_a, d_id_b and
 * d_dist_arcsec, each of `capacity` elements, in no particular order; *n_matches
 * receives the number of pairs.  No counterpart in the reference: for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.
 * Working memory comes from a file-scope device array of 64 MiB per GPU, held
 * until the process ends; A's chunk table must fit in it (about 17 MB for 2e8 rows
 * at 1"): a join that needs more returns XM_EINVAL with the sizes.  A join of more
 * pairs than capacity returns XM_ENOMEM with the count in *n_matches and the
 * arrays unwritten.  Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (two threads, two contexts of one device) race, and the caller
 * serialises them.  Not timed by xm_ctx_set_timing; stats (may be NULL) carries
 * n_matches, n_chunks, band_hits and max_dist_err_arcsec, join_ms is 0. */
xm_status xm_match_noalloc(xm_ctx *ctx, const xm_catalog *a, const xm_catalog *b,
                           double radius_deg, xm_precision precision, int64_t *d_id_a,
                           int64_t *d_id_b, double *d_dist_arcsec, uint64_t capacity,
                           uint64_t *n_matches, xm_match_stats *stats);
xm_status xm_result_count(const xm_result *res, uint64_t *n);
/* Device pointers to the n result triples (valid until xm_result_destroy). */
xm_status xm_result_device_ptrs(const xm_result *res, const int64_t **id_a,
                                const int64_t **id_b, const double **dist_arcsec);
/* Copy the result to host arrays of length >= count (pointers may be NULL). */
xm_status xm_result_copy(xm_ctx *ctx, const xm_result *res, int64_t *id_a, int64_t *id_b,
                         double *dist_arcsec);
/* Batch results: job index of each pair (host array of length count) and pairs
 * per job (host array of length n_jobs); either may be NULL. */
xm_status xm_result_jobs(xm_ctx *ctx, const xm_result *res, uint32_t *job, uint64_t *job_counts);
xm_status xm_result_job_device_ptr(const xm_result *res, const uint32_t **job);
void xm_result_destroy(xm_result *res);

/* ---- serving: a resident catalog answering crossmatch requests.  No counterpart
 * in the reference, which ran one batch of jobs and exited (BoostXmatch.cpp:
 * 301-330); built on xm_catalog_create_segments + xm_match_batch.
 * A server runs two threads, each on its own stream of the resident catalog's
 * GPU (both of the default priority).  The preparing thread, on a context of its
 * own, takes the waiting requests up to the admission limits (query rows and
 * requests per batch, free device memory), uploads them and prepares them as one
 * many-segment catalog.  The joining thread, on the resident catalog's context,
 * joins every segment with the resident catalog in one batched launch and hands
 * each request its pairs.  Batch k+1 is prepared while batch k joins.  A request
 * of more than split_rows sources is cut into pieces of split_rows; a batch takes
 * the next piece of each waiting request in turn, so a large request shares its
 * batches with the requests behind it instead of holding them up.  A request's
 * pairs do not depend on what it was batched with or how it was cut.  The
 * resident catalog must outlive the server; requests use its zone height.  Other
 * calls on the resident catalog's context wait for the server's joins (the
 * context's mutex). */
typedef struct xm_server xm_server;

typedef struct {
    uint64_t max_batch_rows;      /* query rows per batch (0 = 1 << 18) */
    uint32_t max_batch_requests;  /* request pieces per batch (0 = 4096) */
    xm_precision precision;
    uint64_t split_rows;          /* rows per piece of a large request (0 = 16384) */
} xm_server_config;

typedef struct {
    uint64_t requests_done;
    uint64_t batches;
    uint64_t rows;               /* query rows joined */
    uint64_t matches;
    double busy_ms;              /* wall time the joining thread spent on batches */
    double prep_ms;              /* wall time the preparing thread spent on batches */
    double overlap_ms;           /* wall time both were busy at once */
    uint64_t pieces;             /* request pieces joined */
} xm_server_stats;

xm_status xm_server_create(xm_ctx *ctx, const xm_catalog *resident, const xm_server_config *cfg,
                           xm_server **out);
/* Queue n host records to be matched with the resident catalog within radius_deg;
 * the records are copied before the call returns.  *ticket names the request. */
xm_status xm_server_submit(xm_server *srv, const xm_obj *records, uint64_t n, double radius_deg,
                           uint64_t *ticket);
/* Wait up to timeout_ms (0 = do not wait, < 0 = wait until done) for the request.
 * XM_OK: done, *n_matches set; XM_EAGAIN: not done yet; another status: the
 * request failed (xm_server_error gives the message) or the ticket is unknown. */
xm_status xm_server_poll(xm_server *srv, uint64_t ticket, int32_t timeout_ms, uint64_t *n_matches);
/* Copy a done request's pairs (query id, resident id, distance in arcsec) into
 * host arrays of length n_matches (any may be NULL) and its submit-to-done time,
 * then forget the request.  A failed request is forgotten as well and its status
 * returned: read its message with xm_server_error first. */
xm_status xm_server_fetch(xm_server *srv, uint64_t ticket, int64_t *id_query, int64_t *id_catalog,
                          double *dist_arcsec, double *latency_ms);
/* Message of a failed request (valid until the request is fetched). */
const char *xm_server_error(xm_server *srv, uint64_t ticket);
xm_status xm_server_get_stats(xm_server *srv, xm_server_stats *stats);
/* Finishes the queued requests, stops the worker and frees the server. */
void xm_server_destroy(xm_server *srv);

/* ---- work counts, computed on the host from the catalogs' sorted keys and
 * zone tables (never from device counters):
 *   w_ref  = sum_z nA(z) * sum_{|z'-z|<=nz} nB(z')        (reference dense tests)
 *   w_cand = #{(a,b): |za-zb|<=nz, cyclic |dra| <= alpha(za)} (pairs reaching the chord)
 *   w_tile = tests the pair kernel executes under its chunking */
xm_status xm_work_counts(xm_ctx *ctx, const xm_catalog *a, const xm_catalog *b,
                         double radius_deg, uint64_t *w_ref, uint64_t *w_cand,
                         uint64_t *w_tile);
/* The same for one job of a batch (its two segments). */
xm_status xm_work_counts_job(xm_ctx *ctx, const xm_catalog *a, const xm_catalog *b,
                             const xm_job *job, uint64_t *w_ref, uint64_t *w_cand,
                             uint64_t *w_tile);

/* ---- version */
int xm_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* XMATCH_H */
