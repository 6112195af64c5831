// xm_synth.hip -- device-side synthetic catalogs (include/xmatch_synth.h).
//
// Random numbers: Philox-4x32-10 (Salmon et al., SC 2011): a 4x32-bit counter and
// a 2x32-bit key, ten rounds of two 32x32->64 multiplications with the published
// multipliers and Weyl key increments.  Counter = (row lo, row hi, stream, draw),
// key = seed; each call yields four 32-bit words = two 53-bit uniform doubles.
#include <cmath>

#include "../../include/xmatch_synth.h"
#include "xm_internal.h"

namespace xm {

namespace {

struct U4 {
    uint32_t v[4];
};

__device__ inline U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = uint64_t(M0) * ctr.v[0];
        const uint64_t p1 = uint64_t(M1) * ctr.v[2];
        const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        U4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

struct Rng {
    uint64_t row;
    uint32_t stream, k0, k1, draw;
    __device__ void uniform2(double& a, double& b) {
        U4 c;
        c.v[0] = uint32_t(row);
        c.v[1] = uint32_t(row >> 32);
        c.v[2] = stream;
        c.v[3] = draw++;
        const U4 o = philox4x32_10(c, k0, k1);
        const uint64_t x = (uint64_t(o.v[0]) << 32) | o.v[1];
        const uint64_t y = (uint64_t(o.v[2]) << 32) | o.v[3];
        a = double(x >> 11) * 0x1.0p-53;
        b = double(y >> 11) * 0x1.0p-53;
    }
};

__device__ inline Rng make_rng(uint64_t seed, uint64_t stream, uint64_t row) {
    Rng r;
    r.row = row;
    r.stream = uint32_t(stream);
    r.k0 = uint32_t(seed);
    r.k1 = uint32_t(seed >> 32) ^ uint32_t(stream >> 32);
    r.draw = 0;
    return r;
}

constexpr double PI = 3.14159265358979323846;
constexpr double D2R = PI / 180.0;

__device__ inline double wrap_ra(double ra) {
    ra = fmod(ra, 360.0);
    if (ra < 0) ra += 360.0;
    if (ra >= 360.0) ra -= 360.0;
    return ra;
}

__device__ inline void unit(double ra, double dec, double& x, double& y, double& z) {
    double sr, cr, sd, cd;
    sincos(ra * D2R, &sr, &cr);
    sincos(dec * D2R, &sd, &cd);
    x = cd * cr;
    y = cd * sr;
    z = sd;
}

__device__ inline void radec(double x, double y, double z, double& ra, double& dec) {
    const double n = sqrt(x * x + y * y + z * z);
    ra = wrap_ra(atan2(y, x) / D2R);
    double s = z / n;
    s = s > 1.0 ? 1.0 : (s < -1.0 ? -1.0 : s);
    dec = asin(s) / D2R;
}

__device__ inline void sphere_point(Rng& g, double& ra, double& dec) {
    double u1, u2;
    g.uniform2(u1, u2);
    ra = wrap_ra(360.0 * u1);
    dec = asin(2.0 * u2 - 1.0) / D2R;
}

__device__ void shape_point(const xm_shape& s, Rng& g, double& ra, double& dec) {
    if (s.kind == XM_SHAPE_PATCH) {
        double u1, u2;
        g.uniform2(u1, u2);
        ra = wrap_ra(s.p[0] + (s.p[1] - s.p[0]) * u1);
        const double s0 = sin(s.p[2] * D2R), s1 = sin(s.p[3] * D2R);
        dec = asin(s0 + (s1 - s0) * u2) / D2R;
        dec = dec < s.p[2] ? s.p[2] : (dec > s.p[3] ? s.p[3] : dec);
        return;
    }
    if (s.kind == XM_SHAPE_PLANE) {
        double px, py, pz;
        unit(s.p[0], s.p[1], px, py, pz);
        const double contrast = s.p[2], w = s.p[3];
        for (int attempt = 0; attempt < 512; attempt++) {
            sphere_point(g, ra, dec);
            double x, y, z;
            unit(ra, dec, x, y, z);
            double d = x * px + y * py + z * pz;
            d = d > 1.0 ? 1.0 : (d < -1.0 ? -1.0 : d);
            const double b = asin(d) / D2R;
            const double acc = (1.0 + contrast * exp(-b * b / (2.0 * w * w))) / (1.0 + contrast);
            double u, unused;
            g.uniform2(u, unused);
            if (u < acc) return;
        }
        return;  // (1 - 1/(1+contrast))^512 chance: keep the last point
    }
    if (s.kind == XM_SHAPE_CAP) {
        // uniform in the cap around (ra_c, dec_c) of radius p[2]: cos(theta) uniform
        double u1, u2;
        g.uniform2(u1, u2);
        const double ct = 1.0 - u1 * (1.0 - cos(s.p[2] * D2R));
        const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
        const double ph = 2.0 * PI * u2;
        // point around the north pole, then rotate the pole onto the centre
        const double lx = st * cos(ph), ly = st * sin(ph), lz = ct;
        const double a = s.p[0] * D2R, d = s.p[1] * D2R;
        const double sa = sin(a), ca = cos(a), sd = sin(d), cd = cos(d);
        // columns: east = (-sa, ca, 0), north = (-sd ca, -sd sa, cd), centre = (cd ca, cd sa, sd)
        const double x = -sa * lx - sd * ca * ly + cd * ca * lz;
        const double y = ca * lx - sd * sa * ly + cd * sa * lz;
        const double z = cd * ly + sd * lz;
        radec(x, y, z, ra, dec);
        return;
    }
    sphere_point(g, ra, dec);
}

__global__ __launch_bounds__(256) void k_synth_fresh(xm_shape s, uint64_t n, uint64_t seed,
                                                     uint64_t stream, int64_t id0,
                                                     xm_obj* __restrict__ out) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    Rng g = make_rng(seed, stream, i);
    double ra, dec;
    shape_point(s, g, ra, dec);
    xm_obj o;
    o.id = id0 + int64_t(i);
    o.ra_deg = ra;
    o.dec_deg = dec;
    out[i] = o;
}

__global__ __launch_bounds__(256) void k_synth_copies(const xm_obj* __restrict__ a, uint64_t na,
                                                      uint64_t n_copy, double sigma_rad,
                                                      xm_shape fresh, uint64_t n, uint64_t seed,
                                                      uint64_t stream, int64_t id0,
                                                      xm_obj* __restrict__ out) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    Rng g = make_rng(seed, stream, i);
    double ra, dec;
    if (i < n_copy && na > 0) {
        double u1, u2, u3, u4;
        g.uniform2(u1, u2);
        g.uniform2(u3, u4);
        uint64_t k = uint64_t(u1 * double(na));
        k = k < na ? k : na - 1;
        const xm_obj src = a[k];
        // Box-Muller pair of N(0, 1), scaled to radians on the tangent plane
        const double rr = sqrt(-2.0 * log(1.0 - u3)) * sigma_rad;
        const double xi = rr * cos(2.0 * PI * u4), eta = rr * sin(2.0 * PI * u4);
        const double ra0 = src.ra_deg * D2R, de0 = src.dec_deg * D2R;
        const double sr = sin(ra0), cr = cos(ra0), sd = sin(de0), cd = cos(de0);
        // gnomonic inverse: normalise u + xi e_ra + eta e_dec
        const double px = cd * cr - xi * sr - eta * sd * cr;
        const double py = cd * sr + xi * cr - eta * sd * sr;
        const double pz = sd + eta * cd;
        radec(px, py, pz, ra, dec);
        (void)u2;
    } else {
        shape_point(fresh, g, ra, dec);
    }
    xm_obj o;
    o.id = id0 + int64_t(i);
    o.ra_deg = ra;
    o.dec_deg = dec < -90.0 ? -90.0 : (dec > 90.0 ? 90.0 : dec);
    out[i] = o;
}

}  // namespace

}  // namespace xm

using namespace xm;

extern "C" xm_status xm_synth_fresh(xm_ctx* ctx, const xm_shape* shape, uint64_t n, uint64_t seed,
                                    uint64_t stream, int64_t id0, xm_obj* d_out) {
    if (!ctx || !shape || (n && !d_out)) return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        XM_HIP(hipSetDevice(ctx->device));
        if (n) {
            hipLaunchKernelGGL(k_synth_fresh, dim3(uint32_t((n + 255) / 256)), dim3(256), 0,
                               ctx->stream, *shape, n, seed, stream, id0, d_out);
            XM_HIP(hipGetLastError());
        }
        XM_HIP(hipStreamSynchronize(ctx->stream));
        return XM_OK;
    } catch (const DeviceError& e) {
        xm::set_error(ctx, e.what);
        return XM_EDEVICE;
    }
}

extern "C" xm_status xm_synth_copies(xm_ctx* ctx, const xm_obj* d_a, uint64_t na, double frac_copy,
                                     double sigma_arcsec, const xm_shape* fresh, uint64_t n,
                                     uint64_t seed, uint64_t stream, int64_t id0, xm_obj* d_out) {
    if (!ctx || !fresh || (n && !d_out) || (na && !d_a) || !(frac_copy >= 0.0 && frac_copy <= 1.0))
        return XM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        XM_HIP(hipSetDevice(ctx->device));
        const uint64_t n_copy = uint64_t(std::llround(frac_copy * double(n)));
        const double sigma_rad = sigma_arcsec / 3600.0 * (3.14159265358979323846 / 180.0);
        if (n) {
            hipLaunchKernelGGL(k_synth_copies, dim3(uint32_t((n + 255) / 256)), dim3(256), 0,
                               ctx->stream, d_a, na, n_copy, sigma_rad, *fresh, n, seed, stream,
                               id0, d_out);
            XM_HIP(hipGetLastError());
        }
        XM_HIP(hipStreamSynchronize(ctx->stream));
        return XM_OK;
    } catch (const DeviceError& e) {
        xm::set_error(ctx, e.what);
        return XM_EDEVICE;
    }
}
