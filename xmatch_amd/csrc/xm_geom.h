// xm_geom.h -- single-source FP64 geometry of the zone crossmatch.
//
// Everything here is __host__ __device__ and compiled with -ffp-contract=off, so
// the host (x86-64 SSE2) and the device (gfx950) evaluate the same IEEE-754
// operations in the same order and round identically.  That is what makes the
// device unit vectors bit-identical to a host evaluation, and the match set
// identical to an FP64 host join.
//
// Reference semantics (paths relative to the reference repository):
//   zone id          BoostXmatch/BoostXmatch/Sorter.cu:20-24 (clamped: dec=+90 lands
//                    in zone n_zones there, outside its table Sorter.cu:156)
//   deg -> rad       Sorter.cu:52-62 (division by RAD2DEG, Common.h:9)
//   unit vector      Worker.cu:31-40 (radec2xyz)
//   ra half-width    Common.cu:5-23 (calc_alpha, Gray-Nieto-Szalay 2007)
//   chord threshold  Worker.cu:315-317
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define XM_HD __host__ __device__ inline

namespace xm {

constexpr double RAD2DEG = 57.295779513082323;  // Common.h:9
// Rows are sorted by (zone, ra quantum), ra quantised to 2^-24 turn (0.077 arcsec);
// rows inside one quantum keep their input order (the reference leaves (zone, ra) ties
// unordered).  A catalog keeps each row's quantum as its 32-bit key: within a zone
// the keys ascend, and the zone table gives every row's zone.
constexpr int RA_BITS = 24;
constexpr uint64_t RA_MASK = (uint64_t(1) << RA_BITS) - 1;
constexpr double RA_SCALE = double(uint64_t(1) << RA_BITS) / 360.0;  // quanta per degree
constexpr int SORT_LO_BIT = 0;
static_assert(RA_BITS % 8 == 0 && RA_BITS <= 32, "zone digits start on a byte boundary");
constexpr int64_t WHOLE_ZONE = -1;

// ---------------------------------------------------------------------------
// sin/cos: Cody-Waite reduction by pi/2 (P1, P2 with 33 significant bits, so
// k*P1 and k*P2 are exact for |k| < 2^20), then Taylor series with the rounded
// 1/n! coefficients to degree 21 (sin) and 22 (cos) on |r| <= pi/4.  The split of
// pi/2 was computed from Machin's formula in 400-bit integer arithmetic.
XM_HD double sin_poly(double r) {
    const double z = r * r;
    double p = 1.0 / 51090942171709440000.0;
    p = -1.0 / 121645100408832000.0 + z * p;
    p = 1.0 / 355687428096000.0 + z * p;
    p = -1.0 / 1307674368000.0 + z * p;
    p = 1.0 / 6227020800.0 + z * p;
    p = -1.0 / 39916800.0 + z * p;
    p = 1.0 / 362880.0 + z * p;
    p = -1.0 / 5040.0 + z * p;
    p = 1.0 / 120.0 + z * p;
    p = -1.0 / 6.0 + z * p;
    return r + (r * z) * p;
}

XM_HD double cos_poly(double r) {
    const double z = r * r;
    double p = 1.0 / 1124000727777607680000.0;
    p = -1.0 / 2432902008176640000.0 + z * p;
    p = 1.0 / 6402373705728000.0 + z * p;
    p = -1.0 / 20922789888000.0 + z * p;
    p = 1.0 / 87178291200.0 + z * p;
    p = -1.0 / 479001600.0 + z * p;
    p = 1.0 / 3628800.0 + z * p;
    p = -1.0 / 40320.0 + z * p;
    p = 1.0 / 720.0 + z * p;
    p = -1.0 / 24.0 + z * p;
    p = 0.5 + z * p;  // (1 - cos r) / r^2
    return 1.0 - z * p;
}

XM_HD void sincos_cw(double x, double* s, double* c) {
    const double P1 = 0x1.921fb54400000p+0;
    const double P2 = 0x1.0b4611a600000p-34;
    const double P3 = 0x1.3198a2e037073p-69;
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double kd = __builtin_rint(x * TWO_OVER_PI);
    const double r = ((x - kd * P1) - kd * P2) - kd * P3;
    const double sr = sin_poly(r);
    const double cr = cos_poly(r);
    const int q = int(kd) & 3;
    if (q == 0) { *s = sr; *c = cr; }
    else if (q == 1) { *s = cr; *c = -sr; }
    else if (q == 2) { *s = -sr; *c = -cr; }
    else { *s = -cr; *c = sr; }
}

// radec2xyz (Worker.cu:31-40) of degrees converted as get_radec_radian (Sorter.cu:58-59).
XM_HD void unit_vector(double ra_deg, double dec_deg, double* x, double* y, double* z) {
    const double ra = ra_deg / RAD2DEG;
    const double dec = dec_deg / RAD2DEG;
    double sr, cr, sd, cd;
    sincos_cw(ra, &sr, &cr);
    sincos_cw(dec, &sd, &cd);
    *x = cd * cr;
    *y = cd * sr;
    *z = sd;
}

XM_HD int32_t n_zones_of(double h_deg) { return int32_t(__builtin_ceil(180.0 / h_deg)); }

// GetZoneId (Sorter.cu:20-24), clamped into [0, n_zones-1].
XM_HD int32_t zone_of(double dec_deg, double h_deg, int32_t n_zones) {
    int32_t z = int32_t(__builtin_floor((dec_deg + 90.0) / h_deg));
    z = z < 0 ? 0 : z;
    return z > n_zones - 1 ? n_zones - 1 : z;
}

XM_HD uint64_t ra_quantum(double ra_deg) {
    double q = ra_deg * RA_SCALE;
    uint64_t u = q <= 0.0 ? 0 : uint64_t(q);
    return u > RA_MASK ? RA_MASK : u;
}

// segments one catalog holds
constexpr uint64_t MAX_SEGMENTS = uint64_t(1) << 16;

// Chord threshold (2 sin(r/2))^2 of Worker.cu:315-317 with the sine above.
XM_HD double chord2_of_radius(double sr_deg) {
    const double sr_rad = sr_deg / RAD2DEG;
    double s, c;
    sincos_cw(sr_rad / 2.0, &s, &c);
    const double d = 2.0 * s;
    return d * d;
}

// Half-width, in ra quanta, of the window for A zone `zone` (calc_alpha of
// Common.cu:16-23 then Common.cu:5-14, i.e. atan(sin theta)/sqrt(|cos(d-t)cos(d+t)|)
// at the zone's larger |dec| edge plus 0.01 theta), widened by a relative 1e-6
// plus 1e-10 deg and two quanta, so the window is a strict superset of every pair
// the chord test can accept.  atan_sin_theta = atan(sin(theta)) is passed in
// (computed once per join on the host) so host and device use the same value.
// Returns WHOLE_ZONE when the reference's pole rule (|dec| + theta > 89.99)
// applies or the window covers the circle.
XM_HD int64_t alpha_quanta(double theta_deg, double h_deg, int32_t zone, double atan_sin_theta) {
    const double d1 = __builtin_fabs(double(zone) * h_deg - 90.0);
    const double d2 = __builtin_fabs(double(zone + 1) * h_deg - 90.0);
    const double d = (d1 > d2 ? d1 : d2) + 0.01 * theta_deg;
    if (d + theta_deg > 89.99) return WHOLE_ZONE;
    double s1, c1, s2, c2;
    sincos_cw((d - theta_deg) / RAD2DEG, &s1, &c1);
    sincos_cw((d + theta_deg) / RAD2DEG, &s2, &c2);
    const double alpha_rad = __builtin_fabs(atan_sin_theta / __builtin_sqrt(__builtin_fabs(c1 * c2)));
    const double alpha_deg = alpha_rad * RAD2DEG * (1.0 + 1e-6) + 1e-10;
    const double q = __builtin_ceil(alpha_deg * RA_SCALE) + 2.0;
    if (q >= double(uint64_t(1) << (RA_BITS - 1))) return WHOLE_ZONE;
    return int64_t(q);
}

// Half-width of a join's window in ra quanta for A zone `zone`: alpha_quanta plus
// the coarser of the two catalogs' ra granularities (gran quanta), since each
// catalog's keys floor ra to its own granularity.  WHOLE_ZONE as alpha_quanta.
XM_HD int64_t join_quanta(double theta_deg, double h_deg, int32_t zone, double atan_sin_theta,
                          int64_t gran) {
    int64_t aq = alpha_quanta(theta_deg, h_deg, zone, atan_sin_theta);
    if (aq == WHOLE_ZONE) return WHOLE_ZONE;
    aq += gran;
    if (aq >= int64_t(uint64_t(1) << (RA_BITS - 1))) return WHOLE_ZONE;
    return aq;
}

// Packed sort word of a catalog build: zone | ra (ra_bits) | input row (idx_bits),
// 64 bits in all.  idx_bits = bits to hold the row index, zone_bits = bits to hold
// the zone id, ra_bits = min(24, 64 - zone_bits - idx_bits); sorting the word on
// its zone and ra bits orders the rows by (zone, ra) and carries each row's input
// position along.  The catalog key keeps ra << (24 - ra_bits).
struct PackLayout {
    int idx_bits, ra_bits, zone_bits;
};

XM_HD int bit_length_u64(uint64_t v) {
    int b = 0;
    while (v) { b++; v >>= 1; }
    return b;
}

XM_HD PackLayout pack_layout(uint64_t n, int32_t n_zones) {
    PackLayout p;
    p.idx_bits = bit_length_u64(n > 0 ? n - 1 : 0);
    p.zone_bits = bit_length_u64(uint64_t(n_zones > 0 ? n_zones - 1 : 0));
    const int room = 64 - p.idx_bits - p.zone_bits;
    p.ra_bits = room < RA_BITS ? room : RA_BITS;
    return p;
}

XM_HD uint64_t pack_word(const PackLayout& p, int32_t zone, double ra_deg, uint64_t idx) {
    const uint64_t q = ra_quantum(ra_deg) >> (RA_BITS - p.ra_bits);
    return (uint64_t(uint32_t(zone)) << (p.ra_bits + p.idx_bits)) | (q << p.idx_bits) | idx;
}

// The (at most two) ra quantum intervals of a B zone that a chunk whose A rows span
// ra quanta [q_lo, q_hi] must test, for half-width aq (or WHOLE_ZONE): inclusive
// bounds [lo0, hi0] and, when n = 2, [lo1, hi1] on the B rows' keys (their ra
// quanta); sh0 and sh1 map the quanta of each interval into the chunk's frame
// across the 0/360 seam (-2^24, 0 or +2^24).  Plain fields rather than arrays, so
// that a kernel keeps them in scalar registers.
struct Window {
    int n;
    uint32_t lo0, hi0, lo1, hi1;
    int32_t sh0, sh1;
};

XM_HD Window window_intervals(uint64_t q_lo, uint64_t q_hi, int64_t aq) {
    const int64_t full = int64_t(uint64_t(1) << RA_BITS);
    Window w{1, 0, uint32_t(RA_MASK), 0, 0, 0, 0};
    if (aq == WHOLE_ZONE) return w;
    const int64_t l = int64_t(q_lo) - aq;
    const int64_t h = int64_t(q_hi) + aq;
    if (h - l + 1 >= full) return w;
    if (l < 0) {
        w.n = 2;
        w.lo0 = uint32_t(l + full); w.hi0 = uint32_t(RA_MASK); w.sh0 = -int32_t(full);
        w.lo1 = 0; w.hi1 = uint32_t(h);
        return w;
    }
    if (h >= full) {
        w.n = 2;
        w.lo0 = uint32_t(l); w.hi0 = uint32_t(RA_MASK);
        w.lo1 = 0; w.hi1 = uint32_t(h - full); w.sh1 = int32_t(full);
        return w;
    }
    w.lo0 = uint32_t(l); w.hi0 = uint32_t(h);
    return w;
}

// Interval t (0 or 1) of a window as a window of its own.
XM_HD Window window_part(const Window& w, int t) {
    return t == 0 ? Window{1, w.lo0, w.hi0, 0, 0, w.sh0, 0}
                  : Window{1, w.lo1, w.hi1, 0, 0, w.sh1, 0};
}

// Join mode of one A chunk of cnt rows spanning ra quanta [q_lo, q_hi].  Dense:
// every B row of the window is tested against every row of the chunk (one wave
// instruction per B row and 64 chunk rows).  Sparse: each B row is tested only
// against the chunk rows within +-aq of its ra (binary search on the chunk's ra
// quanta).  Dense is used for the pole rule, for windows that cover the circle,
// and when a B row is expected to meet at least dense_min chunk rows:
// cnt (2aq+1) / (span + 2aq + 1) >= dense_min.
XM_HD bool chunk_is_dense(uint64_t q_lo, uint64_t q_hi, int64_t aq, uint32_t dense_min,
                          uint32_t cnt) {
    if (aq == WHOLE_ZONE) return true;
    const int64_t full = int64_t(uint64_t(1) << RA_BITS);
    const int64_t w = 2 * aq + 1;
    const int64_t span = int64_t(q_hi) - int64_t(q_lo);
    if (span + w >= full) return true;
    return int64_t(cnt) * w >= int64_t(dense_min) * (span + w);
}

}  // namespace xm
