/*
 * zone_oracle.c -- CPU oracle (test infrastructure, never shipped in the product
 * path).  FP64 restatement of the BoostXmatch zone join; see xmo.h for the map of
 * reference functions and the corrected deviations.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off).  The chord is evaluated as
 *   d2 = fma(dz, dz, fma(dy, dy, dx*dx))
 * with the exactly rounded C99 fma(), the same operations the device kernel issues
 * (and the contraction nvcc applies by default to the reference's
 * dist2 = x*x + y*y + z*z, Worker.cu:96).  Nothing else may be contracted.
 */
#define _GNU_SOURCE
#include "xmo.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- sincos ---- */
/* pi/2 split for Cody-Waite reduction.  P1 and P2 carry 33 significant bits so
 * that k*P1 and k*P2 are exact for |k| < 2^20; P3 is the rounded remainder.
 * Derived from Machin's formula pi/4 = 4 atan(1/5) - atan(1/239) in 400-bit
 * integer arithmetic. */
static const double XMO_P1 = 0x1.921fb54400000p+0;
static const double XMO_P2 = 0x1.0b4611a600000p-34;
static const double XMO_P3 = 0x1.3198a2e037073p-69;
static const double XMO_2_OVER_PI = 0x1.45f306dc9c883p-1;

/* Taylor series of sin and cos on |r| <= pi/4 with coefficients 1/n! (each the
 * correctly rounded double of the exact rational).  Truncation error of the
 * degree-21 sine and degree-22 cosine is below 2e-24. */
static double xmo_sin_poly(double r)
{
    const double z = r * r;
    double p = 1.0 / 51090942171709440000.0;         /* 1/21! */
    p = -1.0 / 121645100408832000.0 + z * p;         /* 1/19! */
    p = 1.0 / 355687428096000.0 + z * p;             /* 1/17! */
    p = -1.0 / 1307674368000.0 + z * p;              /* 1/15! */
    p = 1.0 / 6227020800.0 + z * p;                  /* 1/13! */
    p = -1.0 / 39916800.0 + z * p;                   /* 1/11! */
    p = 1.0 / 362880.0 + z * p;                      /* 1/9!  */
    p = -1.0 / 5040.0 + z * p;                       /* 1/7!  */
    p = 1.0 / 120.0 + z * p;                         /* 1/5!  */
    p = -1.0 / 6.0 + z * p;                          /* 1/3!  */
    return r + (r * z) * p;
}

static double xmo_cos_poly(double r)
{
    const double z = r * r;
    double p = 1.0 / 1124000727777607680000.0;       /* 1/22! */
    p = -1.0 / 2432902008176640000.0 + z * p;        /* 1/20! */
    p = 1.0 / 6402373705728000.0 + z * p;            /* 1/18! */
    p = -1.0 / 20922789888000.0 + z * p;             /* 1/16! */
    p = 1.0 / 87178291200.0 + z * p;                 /* 1/14! */
    p = -1.0 / 479001600.0 + z * p;                  /* 1/12! */
    p = 1.0 / 3628800.0 + z * p;                     /* 1/10! */
    p = -1.0 / 40320.0 + z * p;                      /* 1/8!  */
    p = 1.0 / 720.0 + z * p;                         /* 1/6!  */
    p = -1.0 / 24.0 + z * p;                         /* 1/4!  */
    p = 0.5 + z * p;                                 /* 1/2!: p = (1 - cos r) / r^2 */
    return 1.0 - z * p;
}

void xmo_sincos(double x, double *s, double *c)
{
    const double kd = rint(x * XMO_2_OVER_PI);
    const double r = ((x - kd * XMO_P1) - kd * XMO_P2) - kd * XMO_P3;
    const double sr = xmo_sin_poly(r);
    const double cr = xmo_cos_poly(r);
    const int q = ((int)kd) & 3;
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

/* radec2xyz, Worker.cu:31-40, on radians obtained as in Sorter.cu:58-59. */
void xmo_unitvec(double ra_deg, double dec_deg, double *x, double *y, double *z)
{
    const double ra = ra_deg / XMO_RAD2DEG;
    const double dec = dec_deg / XMO_RAD2DEG;
    double sr, cr, sd, cd;
    xmo_sincos(ra, &sr, &cr);
    xmo_sincos(dec, &sd, &cd);
    *x = cd * cr;
    *y = cd * sr;
    *z = sd;
}

/* Sorter.cu:156 */
int32_t xmo_n_zones(double h_deg) { return (int32_t)ceil(180.0 / h_deg); }

/* Sorter.cu:20-24, clamped into the table (D3). */
int32_t xmo_zone(double dec_deg, double h_deg, int32_t n_zones)
{
    int32_t z = (int32_t)floor((dec_deg + 90.0) / h_deg);
    if (z < 0) z = 0;
    if (z > n_zones - 1) z = n_zones - 1;
    return z;
}

/* Common.cu:5-14.  Returns radians; 180 is the reference's whole-circle sentinel. */
double xmo_alpha(double theta_deg, double abs_dec_deg)
{
    if (abs_dec_deg + theta_deg > 89.99) return 180.0;
    return fabs(atan(sin(theta_deg / XMO_RAD2DEG)) /
                sqrt(fabs(cos((abs_dec_deg - theta_deg) / XMO_RAD2DEG) *
                          cos((abs_dec_deg + theta_deg) / XMO_RAD2DEG))));
}

/* Common.cu:16-23: the zone's larger |dec| edge plus a 0.01*theta overshoot. */
double xmo_alpha_zone(double theta_deg, double h_deg, int32_t zone)
{
    double d1 = fabs(zone * h_deg - 90.0);
    double d2 = fabs((zone + 1) * h_deg - 90.0);
    double d = d1 > d2 ? d1 : d2;
    return xmo_alpha(theta_deg, d + 0.01 * theta_deg);
}

/* Worker.cu:315-317, with the sine of this file. */
double xmo_sr_dist2(double sr_deg)
{
    const double sr_rad = sr_deg / XMO_RAD2DEG;
    double s, c;
    xmo_sincos(sr_rad / 2.0, &s, &c);
    const double d = 2.0 * s;
    return d * d;
}

double xmo_dist_arcsec(double d2)
{
    return 2.0 * asin(sqrt(d2) / 2.0) * XMO_RAD2DEG * 3600.0;
}

/* ------------------------------------------------------------- segments ---- */
typedef struct {
    int32_t zone;
    double ra;
    int64_t idx;
} xmo_sortrec;

/* less_zonera (Sorter.cu:27-43); ties broken by input position so the oracle
 * is deterministic (the reference leaves them unordered). */
static int xmo_cmp_zonera(const void *pa, const void *pb)
{
    const xmo_sortrec *a = (const xmo_sortrec *)pa, *b = (const xmo_sortrec *)pb;
    if (a->zone != b->zone) return a->zone < b->zone ? -1 : 1;
    if (a->ra != b->ra) return a->ra < b->ra ? -1 : 1;
    return (a->idx > b->idx) - (a->idx < b->idx);
}

int xmo_segment_prepare(const xmo_obj *recs, int64_t n, double h_deg, xmo_segment *s)
{
    memset(s, 0, sizeof(*s));
    if (n < 0 || !(h_deg > 0)) return -1;
    s->n = n;
    s->h_deg = h_deg;
    s->n_zones = xmo_n_zones(h_deg);
    size_t nn = (size_t)(n > 0 ? n : 1);
    xmo_sortrec *tmp = (xmo_sortrec *)malloc(nn * sizeof(xmo_sortrec));
    s->id = (int64_t *)malloc(nn * sizeof(int64_t));
    s->ra_rad = (double *)malloc(nn * sizeof(double));
    s->dec_rad = (double *)malloc(nn * sizeof(double));
    s->x = (double *)malloc(nn * sizeof(double));
    s->y = (double *)malloc(nn * sizeof(double));
    s->z = (double *)malloc(nn * sizeof(double));
    s->zone = (int32_t *)malloc(nn * sizeof(int32_t));
    if (!tmp || !s->id || !s->ra_rad || !s->dec_rad || !s->x || !s->y || !s->z || !s->zone) {
        free(tmp);
        xmo_segment_free(s);
        return -2;
    }
    for (int64_t i = 0; i < n; i++) {
        tmp[i].zone = xmo_zone(recs[i].dec, h_deg, s->n_zones);
        tmp[i].ra = recs[i].ra;
        tmp[i].idx = i;
    }
    qsort(tmp, (size_t)n, sizeof(xmo_sortrec), xmo_cmp_zonera);
    for (int64_t i = 0; i < n; i++) {
        const xmo_obj *o = &recs[tmp[i].idx];
        s->id[i] = o->id;                       /* get_id, Sorter.cu:45-49 */
        s->ra_rad[i] = o->ra / XMO_RAD2DEG;     /* get_radec_radian, Sorter.cu:52-62 */
        s->dec_rad[i] = o->dec / XMO_RAD2DEG;
        s->zone[i] = tmp[i].zone;
        xmo_unitvec(o->ra, o->dec, &s->x[i], &s->y[i], &s->z[i]);
    }
    free(tmp);
    if (n > 0) {
        s->zmin = s->zone[0];
        s->zmax = s->zone[n - 1];
    } else {
        s->zmin = 0;
        s->zmax = -1;
    }
    /* lower_bound over zones (Sorter.cu:161-164), for the occupied range only. */
    int64_t nz_tab = (int64_t)s->zmax - s->zmin + 2;
    s->zbegin = (int64_t *)malloc((size_t)(nz_tab > 0 ? nz_tab : 1) * sizeof(int64_t));
    if (!s->zbegin) { xmo_segment_free(s); return -2; }
    int64_t i = 0;
    for (int64_t k = 0; k < nz_tab; k++) {
        int32_t zz = s->zmin + (int32_t)k;
        while (i < n && s->zone[i] < zz) i++;
        s->zbegin[k] = i;
    }
    return 0;
}

void xmo_segment_free(xmo_segment *s)
{
    free(s->zbegin); free(s->id); free(s->ra_rad); free(s->dec_rad);
    free(s->x); free(s->y); free(s->z); free(s->zone);
    memset(s, 0, sizeof(*s));
}

/* ---------------------------------------------------------------- pairs ---- */
static int xmo_pairs_push(xmo_pairs *p, int64_t ida, int64_t idb, double dist)
{
    if (p->n == p->cap) {
        int64_t nc = p->cap ? p->cap * 2 : 1024;
        int64_t *a = (int64_t *)realloc(p->ida, (size_t)nc * sizeof(int64_t));
        if (!a) return -2;
        p->ida = a;
        int64_t *b = (int64_t *)realloc(p->idb, (size_t)nc * sizeof(int64_t));
        if (!b) return -2;
        p->idb = b;
        double *d = (double *)realloc(p->dist_arcsec, (size_t)nc * sizeof(double));
        if (!d) return -2;
        p->dist_arcsec = d;
        p->cap = nc;
    }
    p->ida[p->n] = ida;
    p->idb[p->n] = idb;
    p->dist_arcsec[p->n] = dist;
    p->n++;
    return 0;
}

void xmo_pairs_free(xmo_pairs *p)
{
    free(p->ida); free(p->idb); free(p->dist_arcsec);
    memset(p, 0, sizeof(*p));
}

typedef struct { int64_t a, b; double d; } xmo_trip;

static int xmo_cmp_trip(const void *pa, const void *pb)
{
    const xmo_trip *a = (const xmo_trip *)pa, *b = (const xmo_trip *)pb;
    if (a->a != b->a) return a->a < b->a ? -1 : 1;
    return (a->b > b->b) - (a->b < b->b);
}

static int xmo_pairs_sort(xmo_pairs *p)
{
    if (p->n < 2) return 0;
    xmo_trip *t = (xmo_trip *)malloc((size_t)p->n * sizeof(xmo_trip));
    if (!t) return -2;
    for (int64_t i = 0; i < p->n; i++) { t[i].a = p->ida[i]; t[i].b = p->idb[i]; t[i].d = p->dist_arcsec[i]; }
    qsort(t, (size_t)p->n, sizeof(xmo_trip), xmo_cmp_trip);
    for (int64_t i = 0; i < p->n; i++) { p->ida[i] = t[i].a; p->idb[i] = t[i].b; p->dist_arcsec[i] = t[i].d; }
    free(t);
    return 0;
}

/* ------------------------------------------------------------ zone join ---- */
typedef struct {
    const xmo_segment *A, *B;
    double sr_deg, sr_rad, sr_dist2, dec_box;
    int32_t nz;
    int mode;
    int32_t z_lo, z_hi;          /* A zones handled */
    volatile int32_t next_zone;  /* dynamic schedule, taken with __atomic */
} xmo_job;

typedef struct {
    xmo_job *job;
    xmo_pairs out;
    uint64_t w_ref, w_cand;
    int err;
} xmo_worker;

static double xmo_cyc_dra(double ra1, double ra2)
{
    const double two_pi = 2.0 * 3.14159265358979323846;
    double d = fabs(ra1 - ra2);
    if (d > two_pi - d) d = two_pi - d;
    return d;
}

/* first index in [lo,hi) with ra_rad >= v (ra sorted inside a zone) */
static int64_t xmo_lower_ra(const double *ra, int64_t lo, int64_t hi, double v)
{
    while (lo < hi) {
        int64_t m = lo + (hi - lo) / 2;
        if (ra[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
}

static int xmo_test_pair(xmo_worker *w, int64_t ia, int64_t ib)
{
    const xmo_segment *A = w->job->A, *B = w->job->B;
    const double dx = A->x[ia] - B->x[ib];
    const double dy = A->y[ia] - B->y[ib];
    const double dz = A->z[ia] - B->z[ib];
    const double d2 = fma(dz, dz, fma(dy, dy, dx * dx));
    if (d2 < w->job->sr_dist2)
        return xmo_pairs_push(&w->out, A->id[ia], B->id[ib], xmo_dist_arcsec(d2));
    return 0;
}

static void xmo_join_zone(xmo_worker *w, int32_t z1)
{
    const xmo_job *j = w->job;
    const xmo_segment *A = j->A, *B = j->B;
    if (z1 < A->zmin || z1 > A->zmax) return;
    const int64_t a0 = A->zbegin[z1 - A->zmin], a1 = A->zbegin[z1 - A->zmin + 1];
    if (a1 <= a0) return;
    const double two_pi = 2.0 * 3.14159265358979323846;
    const double alpha = xmo_alpha_zone(j->sr_deg, A->h_deg, z1);  /* Worker.cu:349 */
    const int whole = alpha >= 180.0;
    const double alpha_w = alpha * (1.0 + 1e-6) + 1e-12;            /* pruning margin */
    int32_t zs = z1 - j->nz, ze = z1 + j->nz;                        /* Worker.cu:338-339 */
    if (zs < 0) zs = 0;
    if (ze > A->n_zones - 1) ze = A->n_zones - 1;
    for (int32_t z2 = zs; z2 <= ze; z2++) {
        if (z2 < B->zmin || z2 > B->zmax) continue;
        const int64_t b0 = B->zbegin[z2 - B->zmin], b1 = B->zbegin[z2 - B->zmin + 1];
        if (b1 <= b0) continue;
        w->w_ref += (uint64_t)(a1 - a0) * (uint64_t)(b1 - b0);
        for (int64_t ia = a0; ia < a1; ia++) {
            if (j->mode == 0) {
                /* dense tile as xmatch_kernel, Worker.cu:82-122 */
                for (int64_t ib = b0; ib < b1; ib++) {
                    const double dra = xmo_cyc_dra(A->ra_rad[ia], B->ra_rad[ib]);
                    if (whole || dra <= alpha) w->w_cand++;
                    if (fabs(A->dec_rad[ia] - B->dec_rad[ib]) >= j->dec_box) continue;
                    if (!whole && dra >= alpha_w) continue;
                    if ((w->err = xmo_test_pair(w, ia, ib))) return;
                }
            } else {
                /* per-source ra window by binary search, split at the seam */
                if (whole || alpha_w >= 3.14159265358979323846) {
                    for (int64_t ib = b0; ib < b1; ib++)
                        if ((w->err = xmo_test_pair(w, ia, ib))) return;
                    continue;
                }
                const double lo = A->ra_rad[ia] - alpha_w, hi = A->ra_rad[ia] + alpha_w;
                int64_t r[4];
                int nr = 0;
                if (lo < 0) {
                    r[nr++] = xmo_lower_ra(B->ra_rad, b0, b1, lo + two_pi); r[nr++] = b1;
                    r[nr++] = b0; r[nr++] = xmo_lower_ra(B->ra_rad, b0, b1, hi);
                } else if (hi >= two_pi) {
                    r[nr++] = xmo_lower_ra(B->ra_rad, b0, b1, lo); r[nr++] = b1;
                    r[nr++] = b0; r[nr++] = xmo_lower_ra(B->ra_rad, b0, b1, hi - two_pi);
                } else {
                    r[nr++] = xmo_lower_ra(B->ra_rad, b0, b1, lo);
                    r[nr++] = xmo_lower_ra(B->ra_rad, b0, b1, hi);
                }
                /* the two seam pieces never overlap because alpha_w < pi */
                for (int k = 0; k < nr; k += 2)
                    for (int64_t ib = r[k]; ib < r[k + 1]; ib++)
                        if ((w->err = xmo_test_pair(w, ia, ib))) return;
            }
        }
    }
}

static void *xmo_worker_main(void *arg)
{
    xmo_worker *w = (xmo_worker *)arg;
    xmo_job *j = w->job;
    for (;;) {
        int32_t z = __atomic_fetch_add(&j->next_zone, 1, __ATOMIC_RELAXED);
        if (z > j->z_hi || w->err) break;
        xmo_join_zone(w, z);
    }
    return NULL;
}

int xmo_match_zones(const xmo_segment *A, const xmo_segment *B, double sr_deg, int mode,
                    int n_threads, int32_t z_lo, int32_t z_hi, xmo_pairs *out,
                    xmo_counts *counts)
{
    memset(out, 0, sizeof(*out));
    if (counts) memset(counts, 0, sizeof(*counts));
    if (A->h_deg != B->h_deg || !(sr_deg > 0)) return -1;
    if (n_threads < 1) n_threads = 1;
    xmo_job job;
    memset(&job, 0, sizeof(job));
    job.A = A; job.B = B; job.mode = mode;
    job.sr_deg = sr_deg;
    job.sr_rad = sr_deg / XMO_RAD2DEG;                       /* Worker.cu:315 */
    job.sr_dist2 = xmo_sr_dist2(sr_deg);                     /* Worker.cu:316-317 */
    job.dec_box = job.sr_rad * (1.0 + 1e-6) + 1e-12;         /* Worker.cu:88 + margin */
    job.nz = (int32_t)ceil(sr_deg / A->h_deg);              /* Worker.cu:313 */
    if (z_lo < A->zmin) z_lo = A->zmin;
    if (z_hi > A->zmax) z_hi = A->zmax;
    job.z_lo = z_lo; job.z_hi = z_hi; job.next_zone = z_lo;
    xmo_worker *ws = (xmo_worker *)calloc((size_t)n_threads, sizeof(xmo_worker));
    pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    if (!ws || !th) { free(ws); free(th); return -2; }
    for (int t = 0; t < n_threads; t++) ws[t].job = &job;
    if (n_threads == 1 || A->n == 0) {
        for (int t = 0; t < n_threads; t++) xmo_worker_main(&ws[t]);
    } else {
        for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, xmo_worker_main, &ws[t]);
        for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    }
    int err = 0;
    for (int t = 0; t < n_threads; t++) {
        xmo_worker *w = &ws[t];
        if (w->err) err = w->err;
        if (counts) { counts->w_ref += w->w_ref; counts->w_cand += w->w_cand; }
        for (int64_t i = 0; i < w->out.n && !err; i++)
            err = xmo_pairs_push(out, w->out.ida[i], w->out.idb[i], w->out.dist_arcsec[i]);
        xmo_pairs_free(&w->out);
    }
    free(ws); free(th);
    if (!err) err = xmo_pairs_sort(out);
    if (counts) counts->n_match = (uint64_t)out->n;
    return err;
}

int xmo_match(const xmo_segment *A, const xmo_segment *B, double sr_deg, int mode,
              int n_threads, xmo_pairs *out, xmo_counts *counts)
{
    return xmo_match_zones(A, B, sr_deg, mode, n_threads, A->zmin, A->zmax, out, counts);
}

int xmo_brute(const xmo_obj *a, int64_t na, const xmo_obj *b, int64_t nb, double sr_deg,
              double h_deg, xmo_pairs *out)
{
    memset(out, 0, sizeof(*out));
    const int32_t n_zones = xmo_n_zones(h_deg);
    const int32_t nz = (int32_t)ceil(sr_deg / h_deg);
    const double sr_dist2 = xmo_sr_dist2(sr_deg);
    double *bx = (double *)malloc((size_t)(nb > 0 ? nb : 1) * 3 * sizeof(double));
    int32_t *bz = (int32_t *)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(int32_t));
    if (!bx || !bz) { free(bx); free(bz); return -2; }
    for (int64_t j = 0; j < nb; j++) {
        xmo_unitvec(b[j].ra, b[j].dec, &bx[3 * j], &bx[3 * j + 1], &bx[3 * j + 2]);
        bz[j] = xmo_zone(b[j].dec, h_deg, n_zones);
    }
    int err = 0;
    for (int64_t i = 0; i < na && !err; i++) {
        double x, y, z;
        xmo_unitvec(a[i].ra, a[i].dec, &x, &y, &z);
        const int32_t za = xmo_zone(a[i].dec, h_deg, n_zones);
        for (int64_t j = 0; j < nb; j++) {
            const int32_t dzn = za - bz[j];
            if (dzn > nz || -dzn > nz) continue;
            const double dx = x - bx[3 * j], dy = y - bx[3 * j + 1], dz = z - bx[3 * j + 2];
            const double d2 = fma(dz, dz, fma(dy, dy, dx * dx));
            if (d2 < sr_dist2) {
                err = xmo_pairs_push(out, a[i].id, b[j].id, xmo_dist_arcsec(d2));
                if (err) break;
            }
        }
    }
    free(bx); free(bz);
    if (!err) err = xmo_pairs_sort(out);
    return err;
}

/* --------------------------------------------------------- array helpers ---- */
void xmo_sincos_many(const double *x, int64_t n, double *s, double *c)
{
    for (int64_t i = 0; i < n; i++) xmo_sincos(x[i], &s[i], &c[i]);
}

void xmo_unitvec_many(const double *ra_deg, const double *dec_deg, int64_t n, double *x,
                      double *y, double *z)
{
    for (int64_t i = 0; i < n; i++) xmo_unitvec(ra_deg[i], dec_deg[i], &x[i], &y[i], &z[i]);
}

int xmo_match_objs(const xmo_obj *a, int64_t na, const xmo_obj *b, int64_t nb, double sr_deg,
                   double h_deg, int mode, int n_threads, int32_t z_lo, int32_t z_hi,
                   xmo_pairs *out, xmo_counts *counts)
{
    xmo_segment A, B;
    int err = xmo_segment_prepare(a, na, h_deg, &A);
    if (err) return err;
    err = xmo_segment_prepare(b, nb, h_deg, &B);
    if (err) { xmo_segment_free(&A); return err; }
    err = xmo_match_zones(&A, &B, sr_deg, mode, n_threads, z_lo, z_hi, out, counts);
    xmo_segment_free(&A);
    xmo_segment_free(&B);
    return err;
}
