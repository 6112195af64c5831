/*
 * xmo.h -- CPU oracle for the zone-algorithm crossmatch (test infrastructure only).
 *
 * This is a plain-C, FP64 restatement of the reference zone join of BoostXmatch
 * (matthiaslee/Xmatch).  It exists to CHECK the HIP implementation: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
 * library never links or calls it.
 *
 * Reference functions restated here (paths relative to the reference repository):
 *   record layout            BoostXmatch/BoostXmatch/Obj.h:21-25
 *   zone id                  BoostXmatch/BoostXmatch/Sorter.cu:20-24
 *   (zone, ra) ordering      BoostXmatch/BoostXmatch/Sorter.cu:27-43,153
 *   zone table size          BoostXmatch/BoostXmatch/Sorter.cu:156
 *   deg -> rad by division   BoostXmatch/BoostXmatch/Sorter.cu:52-62, Common.h:9
 *   calc_alpha               BoostXmatch/BoostXmatch/Common.cu:5-23
 *   radec2xyz                BoostXmatch/BoostXmatch/Worker.cu:31-40
 *   zone loop, nz, sr_dist2  BoostXmatch/BoostXmatch/Worker.cu:310-359
 *   pair predicate           BoostXmatch/BoostXmatch/Worker.cu:82-122
 *   id mapping of matches    BoostXmatch/BoostXmatch/Worker.cu:408-419
 *
 * Deviations from the reference (its bugs, corrected):
 *   D1 (per-job H2D byte count) does not arise: there is no device here.
 *   D2 (y -= |y2| in the +2pi branch): every pair uses the plain chord of the
 *      unit vectors; the ra wrap only widens the window.
 *   D3 (dec = +90 lands in zone n_zones): the zone is clamped to n_zones-1.
 *   D4 (output overflow drops blocks): output grows; nothing is truncated.
 *   D6: an FP64 distance is produced for every pair.
 * Boundary rule: the dec box and the alpha window are pruning only.  They are
 * widened by a margin (relative 1e-6 plus 1e-12 rad) that exceeds every rounding
 * error of the chord, so the match set is exactly
 *   { (a,b) : |zone_a - zone_b| <= nz  and  d2(u_a,u_b) < sr_dist2 },
 * d2 = fma(dz, dz, fma(dy, dy, dx*dx)) on the FP64 unit-vector differences.
 */
#ifndef XMO_H
#define XMO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XMO_RAD2DEG 57.295779513082323 /* Common.h:9 */

typedef struct {
    int64_t id;
    double ra;  /* degrees, [0,360) */
    double dec; /* degrees, [-90,90] */
} xmo_obj;       /* 24 bytes, packed little-endian: Obj.h:21-25 */

/* Sine and cosine of x (radians, |x| < 2^20*pi/2): Cody-Waite reduction by pi/2,
 * then Taylor series with 1/n! coefficients to degree 21/22. */
void xmo_sincos(double x, double *s, double *c);
/* radec2xyz (Worker.cu:31-40) on ra/dec in degrees converted as Sorter.cu:58-59. */
void xmo_unitvec(double ra_deg, double dec_deg, double *x, double *y, double *z);
int32_t xmo_n_zones(double h_deg);                                   /* Sorter.cu:156 */
int32_t xmo_zone(double dec_deg, double h_deg, int32_t n_zones);     /* Sorter.cu:20-24 + D3 */
double xmo_alpha(double theta_deg, double abs_dec_deg);             /* Common.cu:5-14 */
double xmo_alpha_zone(double theta_deg, double h_deg, int32_t zone); /* Common.cu:16-23 */
double xmo_sr_dist2(double sr_deg);                                  /* Worker.cu:315-317 */
double xmo_dist_arcsec(double d2);

/* A prepared (zone-sorted) segment: the result of Sorter::Sort (Sorter.cu:141-185). */
typedef struct {
    int64_t n;
    int32_t n_zones;
    double h_deg;
    int32_t zmin, zmax;   /* occupied zone range (zmin > zmax when n == 0) */
    int64_t *zbegin;      /* [zmax-zmin+2] start offsets of zones zmin..zmax+1 */
    int64_t *id;
    double *ra_rad, *dec_rad;
    double *x, *y, *z;
    int32_t *zone;
} xmo_segment;

int xmo_segment_prepare(const xmo_obj *recs, int64_t n, double h_deg, xmo_segment *out);
void xmo_segment_free(xmo_segment *s);

typedef struct {
    int64_t n, cap;
    int64_t *ida, *idb;
    double *dist_arcsec;
} xmo_pairs;

typedef struct {
    uint64_t w_ref;   /* sum over zone pairs of n_A(z) * n_B(z') */
    uint64_t w_cand;  /* pairs with |dzone| <= nz and cyclic |dra| <= alpha(z_a) */
    uint64_t n_match;
} xmo_counts;

/* Worker::Match (Worker.cu:243-428).  mode 0: dense zone-pair tests as the
 * reference kernel does (W_ref tests); mode 1: per-source ra window by binary
 * search (same set, fewer tests).  Output sorted by (idA, idB). */
int xmo_match(const xmo_segment *A, const xmo_segment *B, double sr_deg, int mode,
              int n_threads, xmo_pairs *out, xmo_counts *counts);
/* Same as xmo_match but only for A zones in [z_lo, z_hi] (a deterministic subset). */
int xmo_match_zones(const xmo_segment *A, const xmo_segment *B, double sr_deg, int mode,
                    int n_threads, int32_t z_lo, int32_t z_hi, xmo_pairs *out,
                    xmo_counts *counts);
/* All-pairs check of the predicate (no zones sorted, no windows).  O(na*nb). */
int xmo_brute(const xmo_obj *a, int64_t na, const xmo_obj *b, int64_t nb, double sr_deg,
              double h_deg, xmo_pairs *out);
void xmo_pairs_free(xmo_pairs *p);

/* Convenience for ctypes callers: array forms and prepare-then-match in one call. */
void xmo_sincos_many(const double *x, int64_t n, double *s, double *c);
void xmo_unitvec_many(const double *ra_deg, const double *dec_deg, int64_t n, double *x,
                      double *y, double *z);
int xmo_match_objs(const xmo_obj *a, int64_t na, const xmo_obj *b, int64_t nb, double sr_deg,
                   double h_deg, int mode, int n_threads, int32_t z_lo, int32_t z_hi,
                   xmo_pairs *out, xmo_counts *counts);

#ifdef __cplusplus
}
#endif
#endif
