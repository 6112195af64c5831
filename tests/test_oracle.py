"""Pins the CPU oracle (oracle/zone_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the oracle (a dropped Taylor
term, a wrong quadrant, a transposed coordinate, an off-by-one zone, a window that
is too narrow, a missing seam piece) fails one of them:
  - sin/cos against glibc (numpy) to 1 ulp over many arguments, and closed forms;
  - unit vectors at the axes and poles;
  - zone ids at exact zone edges and the dec = +90 clamp;
  - calc_alpha against the exact Gray-Nieto-Szalay half-width it bounds;
  - the join against an all-pairs evaluation (no zones, no windows) on random,
    seam and polar data, and against exact mathematics for pairs placed at
    r(1 +- 1e-9) along meridians and parallels;
  - the test.idradec self-join counts (10,004 / 10,070 / 10,980 pairs at
    1/2/5 arcsec, 10,000 of them identities) and the golden fixtures.
"""
import hashlib
import json
import os
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from xmatch_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _ulps(a, b):
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(b), np.finfo(float).tiny))


def test_sincos_vs_libm():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-2 * np.pi, 2 * np.pi, 400_000),
                        rng.uniform(-1e-3, 1e-3, 50_000),
                        np.arange(-8, 9) * (np.pi / 4)])
    s, c = O.sincos(x)
    assert _ulps(s, np.sin(x)).max() <= 1.0
    assert _ulps(c, np.cos(x)).max() <= 1.0
    assert np.abs(s - np.sin(x)).max() <= 2.3e-16
    assert np.abs(c - np.cos(x)).max() <= 2.3e-16


def test_sincos_closed_forms():
    s, c = O.sincos(np.array([0.0, np.pi / 2, np.pi, 3 * np.pi / 2, np.pi / 6, np.pi / 3]))
    assert s[0] == 0.0 and c[0] == 1.0
    assert s[1] == 1.0 and abs(c[1] - 6.123233995736766e-17) < 1e-30
    assert c[2] == -1.0 and abs(s[2] - 1.2246467991473532e-16) < 1e-30
    assert s[3] == -1.0
    assert abs(s[4] - 0.5) <= 1.2e-16 and abs(c[5] - 0.5) <= 1.2e-16


def test_unit_vectors_axes_and_poles():
    x, y, z = O.unitvec([0, 90, 180, 270, 123.4, 17.0], [0, 0, 0, 0, 90, -90])
    assert (x[0], y[0], z[0]) == (1.0, 0.0, 0.0)
    assert abs(x[1]) < 1e-16 and y[1] == 1.0 and z[1] == 0.0
    assert x[2] == -1.0 and abs(y[2]) < 1e-15
    assert y[3] == -1.0
    assert z[4] == 1.0 and abs(x[4]) < 1e-16 and abs(y[4]) < 1e-16
    assert z[5] == -1.0
    rng = np.random.default_rng(3)
    ra, dec = rng.uniform(0, 360, 10000), rng.uniform(-90, 90, 10000)
    x, y, z = O.unitvec(ra, dec)
    assert np.abs(x * x + y * y + z * z - 1).max() < 1e-15
    # coordinates are not transposed: atan2(y, x) gives ra back, asin(z) gives dec
    np.testing.assert_allclose(np.degrees(np.arctan2(y, x)) % 360, ra, atol=1e-9)
    np.testing.assert_allclose(np.degrees(np.arcsin(z)), dec, atol=1e-6)


def test_zone_edges():
    h = 0.5  # exact binary fraction: zone edges are exact doubles
    nzn = O.n_zones(h)
    assert nzn == 360
    decs = np.array([-90.0, -89.5, -89.5 - 1e-13, 0.0, -1e-13, 89.5, 90.0])
    z = O.zone(decs, h)
    assert list(z) == [0, 1, 0, 180, 179, 359, 359]  # dec = +90 clamped (D3)
    h = 1.0 / 3600
    assert O.n_zones(h) == 648000
    assert O.zone(np.array([90.0]), h)[0] == 647999


def test_alpha_bounds_exact_halfwidth():
    # calc_alpha must never be narrower than the exact ra half-width
    # asin(sin t / cos d) for any |dec| within the zone.
    for theta in (0.1 / 3600, 1 / 3600, 5 / 3600, 0.05):
        for d in np.linspace(0, 89.9, 200):
            a = O.alpha(theta, d)
            if a >= 180:
                assert d + theta > 89.99
                continue
            exact = np.arcsin(np.sin(np.radians(theta)) / np.cos(np.radians(d)))
            assert a >= exact * (1 - 1e-14)
            if d < 80:  # away from the pole the bound is tight
                assert a <= exact * 1.01
    # at dec 0 the formula reduces to atan(sin t) / cos t
    t = np.radians(2 / 3600)
    assert abs(O.alpha(2 / 3600, 0.0) - np.arctan(np.sin(t)) / np.cos(t)) < 1e-20
    assert O.alpha(1 / 3600, 89.995) == 180.0


def _rand_objs(rng, n, ra0, ra1, dec0, dec1, id0=0):
    ra, dec = synth.uniform_patch(rng, n, ra0, ra1, dec0, dec1)
    return synth.make_objs(id0 + np.arange(n), ra, dec)


@pytest.mark.parametrize("r_deg,ra_c,dec_c,n", [(3.0, 0.0, 85.0, 1000), (1.0, 180.0, -60.0, 2000),
                                                (0.2, 359.5, 0.0, 4000)])
def test_join_degree_radii_equals_allpairs(r_deg, ra_c, dec_c, n):
    # Degree radii: tall zones, ra windows reaching the whole circle near the pole
    # (the alpha sentinel) and windows across the 0/360 seam; both oracle modes
    # must equal the all-pairs set.
    rng = np.random.default_rng(int(r_deg * 10))
    ra, dec = synth.uniform_cap(rng, n, ra_c, dec_c, 10.0)
    a = synth.make_objs(np.arange(n), ra, dec)
    bra, bdec = synth.uniform_cap(rng, n, ra_c, dec_c, 10.0)
    b = synth.make_objs(10**9 + np.arange(n), bra, bdec)
    ref = O.brute(a, b, r_deg)
    assert len(ref[0]) > 5000
    for mode in (0, 1):
        (ia, ib, d), _ = O.match(a, b, r_deg, mode=mode, threads=4)
        assert np.array_equal(ia, ref[0]) and np.array_equal(ib, ref[1])
        assert np.abs(d - ref[2]).max() <= 1e-9


@pytest.mark.parametrize("where", ["equator", "seam", "north_pole", "south_pole", "mid"])
def test_join_equals_allpairs(where):
    rng = np.random.default_rng(zlib.crc32(where.encode()))
    box = {"equator": (10, 10.05, -0.03, 0.03), "seam": (0, 360, -0.02, 0.02),
           "north_pole": (0, 360, 89.97, 90.0), "south_pole": (0, 360, -90.0, -89.97),
           "mid": (200, 200.04, 45, 45.03)}[where]
    a = _rand_objs(rng, 3000, *box)
    if where == "seam":  # crowd both sides of ra = 0
        a["ra"] = np.where(rng.random(3000) < 0.5, rng.uniform(0, 0.01, 3000),
                           rng.uniform(359.99, 360, 3000) % 360)
    b = a.copy()
    b["id"] += 10**9
    bra, bdec = synth.perturb(rng, b["ra"], b["dec"], 2.0)
    b["ra"], b["dec"] = bra, bdec
    for r in (1.5 / 3600, 4 / 3600):
        ref = O.brute(a, b, r)
        for mode in (0, 1):
            (ia, ib, d), _ = O.match(a, b, r, mode=mode, threads=4)
            assert np.array_equal(ia, ref[0]) and np.array_equal(ib, ref[1])
            assert np.abs(d - ref[2]).max() <= 1e-9
        assert len(ref[0]) > 100


def test_join_zone_height_below_radius():
    # zone height r/3 -> nz = 3 neighbour zones each side
    rng = np.random.default_rng(11)
    a = _rand_objs(rng, 2000, 30, 30.02, 10, 10.02)
    b = _rand_objs(rng, 2000, 30, 30.02, 10, 10.02, id0=10**9)
    r = 6 / 3600
    ref = O.brute(a, b, r, zone_height_deg=r / 3)
    (ia, ib, _), _ = O.match(a, b, r, zone_height_deg=r / 3, mode=0)
    assert np.array_equal(ia, ref[0]) and np.array_equal(ib, ref[1]) and len(ia) > 10


@pytest.mark.parametrize("dec0", [0.0, 37.0, -62.0, 88.0])
def test_boundary_pairs_follow_mathematics(dec0):
    # B placed at distance r(1+eps) from A along a meridian and along a parallel:
    # for |eps| = 1e-9 rounding cannot decide, so the set is fixed by mathematics.
    r = 1.0 / 3600
    recs_a, recs_b, expect = [], [], []
    k = 0
    for ra0 in (0.0, 45.0, 180.0, 359.9999999):
        for eps in (-1e-9, 1e-9, -1e-6, 1e-6):
            dist = r * (1 + eps)
            # meridian: same ra, dec shifted by dist
            recs_a.append((k, ra0, dec0)); recs_b.append((10**9 + k, ra0, dec0 + dist))
            expect.append(eps < 0); k += 1
            # parallel: same dec, ra offset giving great-circle distance dist
            c = np.cos(np.radians(dec0))
            dra = np.degrees(2 * np.arcsin(np.sin(np.radians(dist) / 2) / c))
            recs_a.append((k, ra0, dec0)); recs_b.append((10**9 + k, (ra0 + dra) % 360, dec0))
            expect.append(eps < 0); k += 1
    a = np.array(recs_a, dtype=synth.OBJ_DTYPE)
    b = np.array(recs_b, dtype=synth.OBJ_DTYPE)
    got = set()
    for mode in (0, 1):
        (ia, ib, _), _ = O.match(a, b, r, mode=mode)
        pairs = {(x, y) for x, y in zip(ia, ib) if y - 10**9 == x}
        if got:
            assert pairs == got
        got = pairs
    want = {(i, 10**9 + i) for i, e in enumerate(expect) if e}
    assert got == want


@pytest.mark.parametrize("r_arcsec,count,w_ref", [(1, 10004, 19786), (2, 10070, 29540),
                                                  (5, 10980, 58838)])
def test_selfjoin_sample_data(r_arcsec, count, w_ref):
    a = synth.read_idradec(os.path.join(GOLD, "test.idradec"))
    (ia, ib, d), cnt = O.match(a, a, r_arcsec / 3600.0, mode=0)
    assert len(ia) == count and cnt["w_ref"] == w_ref
    ident = ia == ib
    assert ident.sum() == 10000 and np.all(d[ident] == 0.0)
    # symmetric: (a, b) in the set iff (b, a) is
    fwd = set(zip(ia.tolist(), ib.tolist()))
    assert all((y, x) in fwd for x, y in fwd)
    g = np.load(os.path.join(GOLD, "selfjoin_r%d.npz" % r_arcsec))
    assert np.array_equal(g["ida"], ia) and np.array_equal(g["idb"], ib)
    assert np.abs(g["dist_arcsec"] - d).max() <= 1e-9


def test_small_patch_golden():
    with open(os.path.join(GOLD, "small_patch_seed1.json")) as f:
        g = json.load(f)
    A, B, r = synth.config_small_patch(seed=1)
    (ia, ib, d), cnt = O.match(A, B, r, mode=1)
    h = hashlib.sha256()
    h.update(ia.astype("<i8").tobytes())
    h.update(ib.astype("<i8").tobytes())
    assert len(ia) == g["count"] and h.hexdigest() == g["sha256_pairs"]
    assert [list(map(float, x)) for x in zip(ia[:5], ib[:5], d[:5])] == \
        [[float(v) for v in row] for row in g["first"]]


def test_segmentation_invariance():
    # union over segment pairs == unsegmented join (JobManager.cpp:5-13 partitions A x B)
    A, B, r = synth.config_small_patch(seed=5, n_a=20000, n_b=20000)
    (ia, ib, _), _ = O.match(A, B, r)
    got_a, got_b = [], []
    for sa in np.array_split(A, 3):
        for sb in np.array_split(B, 4):
            (x, y, _), _ = O.match(sa, sb, r)
            got_a.append(x); got_b.append(y)
    x, y = np.concatenate(got_a), np.concatenate(got_b)
    o = np.lexsort((y, x))
    assert np.array_equal(x[o], ia) and np.array_equal(y[o], ib)
