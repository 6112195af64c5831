/*
 * xmatch_synth.h -- device-side synthetic catalogs for benchmarks and tests.
 *
 * Not part of the crossmatch path (the reference has no generator); these entry
 * points fill device buffers of xm_obj records so that 10^8..10^9-row workloads
 * are made in milliseconds instead of minutes.  Every record is a pure function of
 * (seed, stream, row): random draws come from Philox-4x32-10 (Salmon, Moraes, Dror
 * and Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011) keyed by the
 * seed, with the counter (row, stream, draw index).
 */
#ifndef XMATCH_SYNTH_H
#define XMATCH_SYNTH_H

#include "xmatch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    XM_SHAPE_SPHERE = 0, /* uniform on the sphere */
    XM_SHAPE_PATCH = 1,  /* p = {ra0, ra1, dec0, dec1}: uniform in ra, uniform in sin(dec) */
    XM_SHAPE_PLANE = 2,  /* p = {pole_ra, pole_dec, contrast, width}: density proportional to
                            1 + contrast * exp(-b^2 / (2 width^2)), b = latitude relative to
                            the great circle with that pole (degrees); rejection sampling */
    XM_SHAPE_CAP = 3     /* p = {ra_c, dec_c, radius}: uniform inside a spherical cap */
} xm_shape_kind;

typedef struct {
    int32_t kind;
    double p[4];
} xm_shape;

/* n fresh records of the shape, ids id0 .. id0+n-1, into d_out (device). */
xm_status xm_synth_fresh(xm_ctx *ctx, const xm_shape *shape, uint64_t n, uint64_t seed,
                         uint64_t stream, int64_t id0, xm_obj *d_out);

/* n records: rows [0, round(frac_copy*n)) copy a uniformly chosen record of d_a
 * (na rows) moved by a Gaussian of sigma_arcsec per tangent-plane axis (gnomonic
 * inverse); the other rows are fresh records of `fresh`.  ids id0 .. id0+n-1. */
xm_status xm_synth_copies(xm_ctx *ctx, const xm_obj *d_a, uint64_t na, double frac_copy,
                          double sigma_arcsec, const xm_shape *fresh, uint64_t n, uint64_t seed,
                          uint64_t stream, int64_t id0, xm_obj *d_out);

#ifdef __cplusplus
}
#endif
#endif /* XMATCH_SYNTH_H */
