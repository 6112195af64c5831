"""Synthetic catalogs for the crossmatch workloads (numpy PCG64, seeded).

Every catalog is a record array of the reference's 24-byte object layout
{int64 id; float64 ra_deg; float64 dec_deg} (BoostXmatch/BoostXmatch/Obj.h:21-25).

Draws:
  uniform doubles        numpy Generator(PCG64(seed)).random()  (53-bit)
  uniform sphere         ra = 360 U,  dec = asin(2U - 1)
  patch                  ra = U[ra0, ra1),  dec = asin(U[sin dec0, sin dec1])
  perturbed copies       (xi, eta) ~ N(0, sigma^2) on the tangent plane at the
                         source, mapped back to the sphere by the gnomonic inverse
                         (normalise u + xi e_ra + eta e_dec), ra wrapped to [0, 360)
"""
import numpy as np

OBJ_DTYPE = np.dtype([("id", "<i8"), ("ra", "<f8"), ("dec", "<f8")])
ARCSEC = 1.0 / 3600.0
B_ID_OFFSET = 1_000_000_000


def make_objs(ids, ra, dec):
    out = np.empty(len(ids), dtype=OBJ_DTYPE)
    out["id"] = ids
    out["ra"] = ra
    out["dec"] = dec
    return out


def _wrap_ra(ra):
    ra = np.mod(ra, 360.0)
    return np.where(ra >= 360.0, ra - 360.0, ra)


def uniform_sphere(rng, n):
    ra = 360.0 * rng.random(n)
    dec = np.degrees(np.arcsin(2.0 * rng.random(n) - 1.0))
    return _wrap_ra(ra), np.clip(dec, -90.0, 90.0)


def uniform_patch(rng, n, ra0, ra1, dec0, dec1):
    ra = ra0 + (ra1 - ra0) * rng.random(n)
    s0, s1 = np.sin(np.radians(dec0)), np.sin(np.radians(dec1))
    dec = np.degrees(np.arcsin(s0 + (s1 - s0) * rng.random(n)))
    return _wrap_ra(ra), np.clip(dec, dec0, dec1)


def perturb(rng, ra_deg, dec_deg, sigma_arcsec):
    """Gaussian offsets of sigma per tangent-plane axis, gnomonic inverse."""
    n = len(ra_deg)
    sig = np.radians(sigma_arcsec * ARCSEC)
    xi = rng.standard_normal(n) * sig
    eta = rng.standard_normal(n) * sig
    ra = np.radians(ra_deg)
    dec = np.radians(dec_deg)
    sr, cr, sd, cd = np.sin(ra), np.cos(ra), np.sin(dec), np.cos(dec)
    px = cd * cr - xi * sr - eta * sd * cr
    py = cd * sr + xi * cr - eta * sd * sr
    pz = sd + eta * cd
    norm = np.sqrt(px * px + py * py + pz * pz)
    ra2 = np.degrees(np.arctan2(py, px))
    dec2 = np.degrees(np.arcsin(np.clip(pz / norm, -1.0, 1.0)))
    return _wrap_ra(ra2), np.clip(dec2, -90.0, 90.0)


def copies_plus_fresh(rng, a, n_b, frac_copy, sigma_arcsec, fresh_fn, id_offset=B_ID_OFFSET):
    """B = frac_copy perturbed copies of random A sources (with replacement) plus
    fresh sources from fresh_fn(rng, n); shuffled; ids id_offset + 0..n_b-1."""
    n_copy = int(round(frac_copy * n_b))
    pick = rng.integers(0, len(a), n_copy)
    ra_c, dec_c = perturb(rng, a["ra"][pick], a["dec"][pick], sigma_arcsec)
    ra_f, dec_f = fresh_fn(rng, n_b - n_copy)
    ra = np.concatenate([ra_c, ra_f])
    dec = np.concatenate([dec_c, dec_f])
    perm = rng.permutation(n_b)
    return make_objs(id_offset + np.arange(n_b, dtype=np.int64), ra[perm], dec[perm])


def config_small_patch(seed=1, n_a=100_000, n_b=100_000):
    """Plain zone crossmatch at small size: ra U[0,10), dec in [-5,5] deg (uniform
    on the sphere), B 70% copies (sigma 0.3") + 30% fresh, radius 1"."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ra, dec = uniform_patch(rng, n_a, 0.0, 10.0, -5.0, 5.0)
    a = make_objs(np.arange(n_a, dtype=np.int64), ra, dec)
    b = copies_plus_fresh(rng, a, n_b, 0.7, 0.3,
                          lambda r, n: uniform_patch(r, n, 0.0, 10.0, -5.0, 5.0))
    return a, b, 1.0 * ARCSEC


def config_full_sky(seed=2, n_a=50_000_000, n_b=50_000_000):
    """Full-sky plain FP64 join: A uniform on the sphere; B 60% copies
    (sigma 0.5") + 40% fresh uniform; radius 2"."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ra, dec = uniform_sphere(rng, n_a)
    a = make_objs(np.arange(n_a, dtype=np.int64), ra, dec)
    b = copies_plus_fresh(rng, a, n_b, 0.6, 0.5, uniform_sphere)
    return a, b, 2.0 * ARCSEC


# Pole of the dense great circle: 60 deg from the celestial pole, at ra = 0.
PLANE_POLE_RA, PLANE_POLE_DEC = 0.0, 30.0


def plane_density_sample(rng, n, chunk=1 << 22):
    """Sources with density proportional to 1 + 20 exp(-b^2 / (2 (10 deg)^2)),
    b = latitude relative to the great circle whose pole is (PLANE_POLE_RA,
    PLANE_POLE_DEC); rejection sampling from the uniform sphere."""
    pr, pd = np.radians(PLANE_POLE_RA), np.radians(PLANE_POLE_DEC)
    pole = np.array([np.cos(pd) * np.cos(pr), np.cos(pd) * np.sin(pr), np.sin(pd)])
    ra_out, dec_out = [], []
    have = 0
    while have < n:
        ra, dec = uniform_sphere(rng, chunk)
        r, d = np.radians(ra), np.radians(dec)
        dot = np.cos(d) * np.cos(r) * pole[0] + np.cos(d) * np.sin(r) * pole[1] + np.sin(d) * pole[2]
        b = np.degrees(np.arcsin(np.clip(dot, -1.0, 1.0)))
        w = (1.0 + 20.0 * np.exp(-b * b / (2.0 * 10.0 ** 2))) / 21.0
        keep = rng.random(chunk) < w
        ra_out.append(ra[keep])
        dec_out.append(dec[keep])
        have += int(keep.sum())
    return np.concatenate(ra_out)[:n], np.concatenate(dec_out)[:n]


def config_plane(seed=3, n_a=200_000_000, n_b=200_000_000):
    """Galactic-plane-like density, B 50% copies (sigma 0.3") + 50% fresh, radius 1"."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ra, dec = plane_density_sample(rng, n_a)
    a = make_objs(np.arange(n_a, dtype=np.int64), ra, dec)
    b = copies_plus_fresh(rng, a, n_b, 0.5, 0.3, plane_density_sample)
    return a, b, 1.0 * ARCSEC


def uniform_cap(rng, n, ra_c, dec_c, radius_deg):
    """Uniform inside the spherical cap of radius_deg about (ra_c, dec_c): cos of the
    angle from the centre uniform in [cos r, 1], position angle uniform."""
    ct = 1.0 - rng.random(n) * (1.0 - np.cos(np.radians(radius_deg)))
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = 2.0 * np.pi * rng.random(n)
    lx, ly, lz = st * np.cos(ph), st * np.sin(ph), ct
    a, d = np.radians(ra_c), np.radians(dec_c)
    sa, ca, sd, cd = np.sin(a), np.cos(a), np.sin(d), np.cos(d)
    x = -sa * lx - sd * ca * ly + cd * ca * lz
    y = ca * lx - sd * sa * ly + cd * sa * lz
    z = cd * ly + sd * lz
    ra = _wrap_ra(np.degrees(np.arctan2(y, x)))
    dec = np.degrees(np.arcsin(np.clip(z, -1.0, 1.0)))
    return ra, np.clip(dec, -90.0, 90.0)


def batch_job_specs(seed=4, n_jobs=10_000, n_lo=1e2, n_hi=1e6, cap_lo=0.1, cap_hi=2.0,
                    r_lo_arcsec=0.5, r_hi_arcsec=10.0):
    """Parameters of a batch of independent joins: N_A, N_B ~ logU[n_lo, n_hi],
    cap centre uniform on the sphere, cap radius U[cap_lo, cap_hi] deg, radius
    ~ logU[r_lo, r_hi] arcsec, zone height = radius.  Returns a structured array."""
    rng = np.random.Generator(np.random.PCG64(seed))
    spec = np.empty(n_jobs, dtype=[("n_a", "<i8"), ("n_b", "<i8"), ("ra_c", "<f8"),
                                   ("dec_c", "<f8"), ("cap_deg", "<f8"), ("r_deg", "<f8")])
    lg = np.log10
    spec["n_a"] = np.round(10.0 ** rng.uniform(lg(n_lo), lg(n_hi), n_jobs)).astype(np.int64)
    spec["n_b"] = np.round(10.0 ** rng.uniform(lg(n_lo), lg(n_hi), n_jobs)).astype(np.int64)
    ra, dec = uniform_sphere(rng, n_jobs)
    spec["ra_c"], spec["dec_c"] = ra, dec
    spec["cap_deg"] = rng.uniform(cap_lo, cap_hi, n_jobs)
    spec["r_deg"] = 10.0 ** rng.uniform(lg(r_lo_arcsec), lg(r_hi_arcsec), n_jobs) * ARCSEC
    return spec


def batch_job_catalogs(spec, seed=4):
    """Host catalogs of a batch: per job, A uniform in its cap; B 50% copies of A
    (sigma 0.3") + 50% fresh in the cap.  Ids are unique across the batch."""
    rng = np.random.Generator(np.random.PCG64([seed, 1]))
    out = []
    id0 = 0
    for s in spec:
        na, nb = int(s["n_a"]), int(s["n_b"])
        ra, dec = uniform_cap(rng, na, s["ra_c"], s["dec_c"], s["cap_deg"])
        a = make_objs(id0 + np.arange(na, dtype=np.int64), ra, dec)
        b = copies_plus_fresh(rng, a, nb, 0.5, 0.3,
                              lambda r, n, s=s: uniform_cap(r, n, s["ra_c"], s["dec_c"],
                                                            s["cap_deg"]),
                              id_offset=B_ID_OFFSET + id0) if na else make_objs(
                                  B_ID_OFFSET + id0 + np.arange(nb, dtype=np.int64),
                                  *uniform_cap(rng, nb, s["ra_c"], s["dec_c"], s["cap_deg"]))
        id0 += max(na, nb)
        out.append((a, b, float(s["r_deg"])))
    return out


def read_idradec(path, count=None, offset=0):
    """Raw little-endian 24-byte records, as Segment::Load reads them
    (BoostXmatch/BoostXmatch/Segment.cpp:14-22); count = file bytes / 24."""
    return np.fromfile(path, dtype=OBJ_DTYPE, count=-1 if count is None else count,
                       offset=offset * OBJ_DTYPE.itemsize)
