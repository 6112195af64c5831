"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the
golden fixtures.  The bar: identical (idA, idB) sets (bit-exact), distances within
1e-6 arcsec, FP64 unit vectors bit-identical to the oracle's."""
import hashlib
import json
import os
import zlib

import numpy as np
import pytest

import xmatch_amd as xm
from oracle import oracle as O
from xmatch_amd import synth

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DIST_TOL_ARCSEC = 1e-6


def gpu_join(ctx, a, b, r, h=None, capacity_hint=0, precision=xm.FP64_PLAIN):
    h = r if h is None else h
    ca = ctx.catalog(a, h)
    cb = ctx.catalog(b, h)
    res = ctx.match(ca, cb, r, precision=precision, capacity_hint=capacity_hint)
    out = res.to_host()
    stats = res.stats.as_dict()
    res.close(); ca.close(); cb.close()
    return out, stats


def assert_same(got, ref):
    ia, ib, d = got
    ra, rb, rd = ref
    assert len(ia) == len(ra), (len(ia), len(ra))
    assert np.array_equal(ia, ra) and np.array_equal(ib, rb)
    if len(d):
        assert np.abs(d - rd).max() <= DIST_TOL_ARCSEC


def test_unit_vectors_and_zones_bit_exact(gpu_ctx):
    A, _, r = synth.config_small_patch(seed=1)
    special = synth.make_objs(np.arange(10**7, 10**7 + 8),
                              [0.0, 90.0, 180.0, 270.0, 359.999999999, 12.5, 0.0, 45.0],
                              [0.0, 0.0, 0.0, 0.0, 0.0, 90.0, -90.0, -45.0])
    recs = np.concatenate([A, special])
    cat = gpu_ctx.catalog(recs, r)
    e = cat.export()
    cat.close()
    o = np.argsort(recs["id"])
    p = np.argsort(e["id"])
    assert np.array_equal(e["id"][p], recs["id"][o])
    x, y, z = O.unitvec(recs["ra"][o], recs["dec"][o])
    assert np.array_equal(e["x"][p].view(np.int64), x.view(np.int64))
    assert np.array_equal(e["y"][p].view(np.int64), y.view(np.int64))
    assert np.array_equal(e["z"][p].view(np.int64), z.view(np.int64))
    assert np.array_equal(e["zone"][p], O.zone(recs["dec"][o], r))


def test_sorted_by_zone_then_ra(gpu_ctx):
    A, _, r = synth.config_small_patch(seed=3)
    cat = gpu_ctx.catalog(A, r)
    e = cat.export()
    info = cat.info()
    cat.close()
    assert np.all(np.diff(e["zone"]) >= 0)
    assert info["zmin"] == e["zone"][0] and info["zmax"] == e["zone"][-1]
    ra = dict(zip(A["id"].tolist(), A["ra"].tolist()))
    ras = np.array([ra[i] for i in e["id"]])
    same = np.diff(e["zone"]) == 0
    # ra order within a zone, to the sort key's granularity: this catalog's key
    # has 24 ra bits, of which the layout drops fewer than one digit (at most 8),
    # rows of one quantum keeping input order
    assert np.all(np.diff(ras)[same] >= -360.0 / 2**16)
    assert np.mean(np.diff(ras)[same] >= 0) > 0.99


def test_sort_is_stable_and_word_width_invariant(gpu_ctx):
    # Records that share (ra, dec) share their sort key: the radix sort must keep
    # them in input order (ids rise with input position here).  ~70k rows make
    # 18 sort tiles, so the tiles' look-back chains are exercised; 64-bit
    # look-back words must give the same catalog as the default 32-bit words.
    A, _, r = synth.config_small_patch(seed=5, n_a=60_000, n_b=10)
    rng = np.random.default_rng(5)
    dup = np.repeat(A[rng.choice(len(A), 2500, replace=False)], 4)
    recs = np.concatenate([A, dup])
    recs = recs[rng.permutation(len(recs))]
    recs["id"] = np.arange(len(recs), dtype=np.int64) * 7 + 3
    cat = gpu_ctx.catalog(recs, r)
    e = cat.export()
    cat.close()
    pos = dict(zip(recs["id"].tolist(), range(len(recs))))
    assert sorted(pos[i] for i in e["id"].tolist()) == list(range(len(recs)))
    groups = {}
    for i in e["id"].tolist():
        groups.setdefault((recs["ra"][pos[i]], recs["dec"][pos[i]]), []).append(i)
    assert sum(len(v) > 1 for v in groups.values()) == 2500
    for k, ids in groups.items():
        assert ids == sorted(ids), k
    old = gpu_ctx.get_param("sort_wide_words")
    try:
        gpu_ctx.set_param("sort_wide_words", 1)
        cat = gpu_ctx.catalog(recs, r)
        e64 = cat.export()
        cat.close()
    finally:
        gpu_ctx.set_param("sort_wide_words", old)
    for f in e:
        assert np.array_equal(e[f], e64[f]), f


def test_sort_lookback_failure_is_reported(gpu_ctx):
    # With "sort_fail_test" the sort's tiles publish no counts and poll only 64
    # times, so every look-back ends at its limit (the kernels still exit): the
    # catalog call must fail with XM_EINTERNAL, and the context must then sort
    # correctly again.
    A, _, r = synth.config_small_patch(seed=7, n_a=60_000, n_b=10)   # 8 tiles
    assert gpu_ctx.get_param("sort_fail_test") == 0
    try:
        gpu_ctx.set_param("sort_fail_test", 1)
        with pytest.raises(xm.XmatchError) as ei:
            gpu_ctx.catalog(A, r)
        assert ei.value.status == xm.XM_EINTERNAL
        assert "look-back" in str(ei.value)
    finally:
        gpu_ctx.set_param("sort_fail_test", 0)
    cat = gpu_ctx.catalog(A, r)
    e = cat.export()
    cat.close()
    assert np.array_equal(np.sort(e["id"]), np.sort(A["id"]))
    assert np.all(np.diff(e["zone"]) >= 0)
    ref, _ = O.match(A, A, r)
    got, _ = gpu_join(gpu_ctx, A, A, r)
    assert_same(got, ref)


def test_timing_only_one_kernel(gpu_ctx):
    # Timing with a kernel named records that kernel's launches and no other; with
    # none named, every kernel's.
    A, _, r = synth.config_small_patch(seed=3, n_a=40_000, n_b=10)
    try:
        gpu_ctx.reset_timing()
        gpu_ctx.set_timing(True, only="gather")
        gpu_ctx.catalog(A, r).close()
        gpu_ctx.catalog(A, r).close()
        gpu_ctx.set_timing(False)
        t, c = gpu_ctx.kernel_time("gather")
        assert c == 2 and t > 0
        assert gpu_ctx.kernel_time("make_keys")[1] == 0
        assert gpu_ctx.kernel_time("radix_scatter")[1] == 0
        gpu_ctx.reset_timing()
        gpu_ctx.set_timing(True)
        gpu_ctx.catalog(A, r).close()
        gpu_ctx.set_timing(False)
        assert gpu_ctx.kernel_time("gather")[1] == 1
        assert gpu_ctx.kernel_time("make_keys")[1] == 1
        assert gpu_ctx.kernel_time("radix_scatter")[1] >= 1
    finally:
        gpu_ctx.set_timing(False)
        gpu_ctx.reset_timing()


@pytest.mark.parametrize("order", ["dec", "dec_ra", "reversed"])
def test_ordered_input_same_set(gpu_ctx, order):
    # Spatially ordered input (as survey catalogs are often stored) gives whole
    # wavefronts of equal zone digits in the sort; the join must not change.
    A, B, r = synth.config_small_patch(seed=6, n_a=80_000, n_b=80_000)
    if order == "dec":
        A = A[np.argsort(A["dec"], kind="stable")]
        B = B[np.argsort(B["dec"], kind="stable")]
    elif order == "dec_ra":
        A = A[np.lexsort((A["ra"], A["dec"]))]
        B = B[np.lexsort((B["ra"], B["dec"]))]
    else:
        A = A[np.lexsort((A["ra"], A["dec"]))][::-1].copy()
        B = B[np.lexsort((B["ra"], B["dec"]))][::-1].copy()
    ref, _ = O.match(A, B, r)
    got, _ = gpu_join(gpu_ctx, A, B, r)
    assert_same(got, ref)
    cat = gpu_ctx.catalog(A, r)
    e = cat.export()
    cat.close()
    assert np.all(np.diff(e["zone"]) >= 0)


@pytest.mark.parametrize("r_arcsec", [1, 2, 5])
def test_selfjoin_sample_data_golden(gpu_ctx, r_arcsec):
    a = synth.read_idradec(os.path.join(GOLD, "test.idradec"))
    g = np.load(os.path.join(GOLD, "selfjoin_r%d.npz" % r_arcsec))
    got, _ = gpu_join(gpu_ctx, a, a, r_arcsec / 3600.0)
    assert_same(got, (g["ida"], g["idb"], g["dist_arcsec"]))
    ident = got[0] == got[1]
    assert ident.sum() == 10000 and np.all(got[2][ident] == 0.0)


def test_small_patch_vs_oracle_and_golden(gpu_ctx):
    A, B, r = synth.config_small_patch(seed=1)
    got, stats = gpu_join(gpu_ctx, A, B, r)
    ref, _ = O.match(A, B, r)
    assert_same(got, ref)
    with open(os.path.join(GOLD, "small_patch_seed1.json")) as f:
        g = json.load(f)
    h = hashlib.sha256()
    h.update(got[0].astype("<i8").tobytes())
    h.update(got[1].astype("<i8").tobytes())
    assert h.hexdigest() == g["sha256_pairs"] and stats["n_matches"] == g["count"]
    assert stats["n_retries"] == 0


def test_overflow_retry_never_truncates(gpu_ctx):
    A, B, r = synth.config_small_patch(seed=1)
    got, stats = gpu_join(gpu_ctx, A, B, r, capacity_hint=1000)
    assert stats["n_retries"] >= 1
    ref, _ = O.match(A, B, r)
    assert_same(got, ref)


def test_spill_pool_takes_region_overflow(gpu_ctx):
    # ~70k matches over 1024 regions of ~73 slots each: many regions overflow into
    # the shared spill pool, which holds the rest, so no re-run is needed
    A, B, r = synth.config_small_patch(seed=1)
    got, stats = gpu_join(gpu_ctx, A, B, r, capacity_hint=100_000)
    assert stats["n_retries"] == 0
    assert_same(got, O.match(A, B, r)[0])


@pytest.mark.parametrize("case", ["r3deg_pole_cap", "r1deg_south", "r0.2deg_seam",
                                  "r0.05arcsec_dense"])
def test_large_and_small_radii_vs_oracle(gpu_ctx, case):
    # Radii far from the arcsecond scale: degree radii give few, tall zones, ra
    # windows that reach the whole circle near the pole (the reference's alpha
    # sentinel) and cross the 0/360 seam; a 0.05" radius gives 1.3e7 zones (24 zone
    # bits in the sort key).  The set must equal the oracle's.
    rng = np.random.default_rng(sum(map(ord, case)))
    if case == "r3deg_pole_cap":
        r, n = 3.0, 1000
        ra, dec = synth.uniform_cap(rng, n, 0.0, 85.0, 10.0)
    elif case == "r1deg_south":
        r, n = 1.0, 3000
        ra, dec = synth.uniform_cap(rng, n, 180.0, -60.0, 10.0)
    elif case == "r0.2deg_seam":
        r, n = 0.2, 15000
        ra, dec = synth.uniform_cap(rng, n, 359.5, 0.0, 10.0)
    else:
        r, n = 0.05 / 3600, 50000
        ra, dec = synth.uniform_patch(rng, n, 30.0, 30.2, 10.0, 10.2)
    a = synth.make_objs(np.arange(n), ra, dec)
    if case == "r0.05arcsec_dense":
        bra, bdec = synth.perturb(rng, ra, dec, 0.02)
    else:
        if case == "r3deg_pole_cap":
            bra, bdec = synth.uniform_cap(rng, n, 0.0, 85.0, 10.0)
        elif case == "r1deg_south":
            bra, bdec = synth.uniform_cap(rng, n, 180.0, -60.0, 10.0)
        else:
            bra, bdec = synth.uniform_cap(rng, n, 359.5, 0.0, 10.0)
    b = synth.make_objs(10**9 + np.arange(n), bra, bdec)
    ref, _ = O.match(a, b, r)
    assert len(ref[0]) > 10_000
    got, _ = gpu_join(gpu_ctx, a, b, r)
    assert_same(got, ref)
    # the FP32 filter with FP64 recheck returns the same set at these radii too
    got32, _ = gpu_join(gpu_ctx, a, b, r, precision=xm.FP32_FILTER_FP64_RECHECK)
    assert_same(got32, ref)


@pytest.mark.parametrize("where", ["seam", "north_pole", "south_pole", "dense_cluster", "mid"])
def test_edge_geometry_vs_allpairs(gpu_ctx, where):
    rng = np.random.default_rng(zlib.crc32(where.encode()))
    n = 4000
    if where == "seam":
        ra = np.where(rng.random(n) < 0.5, rng.uniform(0, 0.01, n),
                      rng.uniform(359.99, 360, n) % 360)
        dec = rng.uniform(-0.02, 0.02, n)
    elif where == "north_pole":
        ra, dec = synth.uniform_patch(rng, n, 0, 360, 89.97, 90.0)
    elif where == "south_pole":
        ra, dec = synth.uniform_patch(rng, n, 0, 360, -90.0, -89.97)
    elif where == "dense_cluster":  # everything in two zones
        ra, dec = synth.uniform_patch(rng, n, 150.0, 150.003, 20.0, 20.0003)
    else:
        ra, dec = synth.uniform_patch(rng, n, 200, 200.04, 45, 45.03)
    a = synth.make_objs(np.arange(n), ra, dec)
    bra, bdec = synth.perturb(rng, ra, dec, 2.0)
    b = synth.make_objs(10**9 + np.arange(n), bra, bdec)
    if where == "north_pole":
        a["dec"][:3] = 90.0
        b["dec"][:3] = 90.0
    for r in (1.5 / 3600, 4 / 3600):
        got, _ = gpu_join(gpu_ctx, a, b, r)
        assert_same(got, O.brute(a, b, r))


def test_zone_height_below_radius(gpu_ctx):
    rng = np.random.default_rng(11)
    ra, dec = synth.uniform_patch(rng, 5000, 30, 30.02, 10, 10.02)
    a = synth.make_objs(np.arange(5000), ra, dec)
    ra, dec = synth.uniform_patch(rng, 5000, 30, 30.02, 10, 10.02)
    b = synth.make_objs(10**9 + np.arange(5000), ra, dec)
    r = 6 / 3600
    got, _ = gpu_join(gpu_ctx, a, b, r, h=r / 3)
    assert_same(got, O.brute(a, b, r, zone_height_deg=r / 3))


@pytest.mark.parametrize("h_over_r", [1.7, 4.0, 30.0])
def test_zone_height_above_radius(gpu_ctx, h_over_r):
    # Zones taller than the radius (the reference's -z option above -r): one
    # neighbour zone each side, and each zone holds many more rows per search.
    rng = np.random.default_rng(int(h_over_r * 10))
    ra, dec = synth.uniform_patch(rng, 20000, 30, 30.1, 10, 10.1)
    a = synth.make_objs(np.arange(20000), ra, dec)
    bra, bdec = synth.perturb(rng, ra, dec, 2.0)
    b = synth.make_objs(10**9 + np.arange(20000), bra, bdec)
    r = 3 / 3600
    h = r * h_over_r
    ref, _ = O.match(a, b, r, zone_height_deg=h)
    assert len(ref[0]) > 10_000
    got, _ = gpu_join(gpu_ctx, a, b, r, h=h)
    assert_same(got, ref)
    assert_same(got, O.brute(a, b, r, zone_height_deg=h))


def test_boundary_pairs_match_oracle_exactly(gpu_ctx):
    # pairs at distance r(1 + eps) with eps down to 2^-50: rounding decides these,
    # so only an identical FP64 evaluation gives the oracle's set
    rng = np.random.default_rng(5)
    r = 1.0 / 3600
    recs_a, recs_b = [], []
    k = 0
    for dec0 in (0.0, 0.3, 37.0, -62.0, 88.0):
        for ra0 in (0.0, 45.0, 180.0, 359.9999999):
            for e in (-2.0**-30, -2.0**-40, -2.0**-45, -2.0**-50, 0.0, 2.0**-50, 2.0**-45,
                      2.0**-40, 2.0**-30):
                dist = r * (1 + e)
                c = np.cos(np.radians(dec0))
                dra = np.degrees(2 * np.arcsin(np.sin(np.radians(dist) / 2) / c))
                for (pa, pb) in (((ra0, dec0), (ra0, dec0 + dist)),
                                 ((ra0, dec0), ((ra0 + dra) % 360, dec0)),
                                 ((ra0, dec0), ((ra0 - dra) % 360, dec0))):
                    recs_a.append((k, pa[0], pa[1]))
                    recs_b.append((10**9 + k, pb[0], pb[1]))
                    k += 1
    a = np.array(recs_a, dtype=synth.OBJ_DTYPE)
    b = np.array(recs_b, dtype=synth.OBJ_DTYPE)
    b["dec"] = np.clip(b["dec"], -90, 90)
    got, _ = gpu_join(gpu_ctx, a, b, r)
    assert_same(got, O.match(a, b, r)[0])
    assert_same(got, O.brute(a, b, r))


def test_empty_and_single(gpu_ctx):
    one = synth.make_objs([7], [10.0], [10.0])
    other = synth.make_objs([8], [10.0], [10.0 + 0.5 / 3600])
    empty = one[:0]
    r = 1 / 3600
    got, _ = gpu_join(gpu_ctx, empty, one, r)
    assert len(got[0]) == 0
    got, _ = gpu_join(gpu_ctx, one, empty, r)
    assert len(got[0]) == 0
    got, _ = gpu_join(gpu_ctx, empty, empty, r)
    assert len(got[0]) == 0
    got, _ = gpu_join(gpu_ctx, one, other, r)
    assert got[0].tolist() == [7] and got[1].tolist() == [8]
    assert abs(got[2][0] - 0.5) < 1e-6


def test_ids_are_carried_bit_for_bit(gpu_ctx):
    # Ids are opaque int64 payload (Obj.h:21-25): the extremes, negatives and
    # duplicates come back exactly as given, paired as the positions say.
    ids_a = np.array([-2**63, -1, 0, 2**63 - 1, 42, 42], dtype=np.int64)
    ids_b = np.array([2**63 - 1, -2**63, 123456789012345678, -5, 42, 7], dtype=np.int64)
    ra = np.array([0.0, 10.0, 20.0, 359.9999999, 100.0, 200.0])
    dec = np.array([-90.0, 0.0, 45.0, 0.0, 89.9999, -30.0])
    a = synth.make_objs(ids_a, ra, dec)
    b = synth.make_objs(ids_b, ra, dec + 0.2 / 3600 * np.where(dec > 89, -1, 1))
    r = 1 / 3600
    got, _ = gpu_join(gpu_ctx, a, b, r)
    ref, _ = O.match(a, b, r)
    assert_same(got, ref)
    pairs = sorted(zip(got[0].tolist(), got[1].tolist()))
    assert pairs == sorted(zip(ids_a.tolist(), ids_b.tolist()))


@pytest.mark.parametrize("bad", ["nan_ra", "nan_dec", "dec_high", "dec_low", "ra_360", "ra_neg"])
def test_invalid_records_rejected(gpu_ctx, bad):
    a = synth.make_objs([1, 2, 3], [10.0, 20.0, 30.0], [0.0, 1.0, 2.0])
    field, val = {"nan_ra": ("ra", np.nan), "nan_dec": ("dec", np.nan),
                  "dec_high": ("dec", 90.0001), "dec_low": ("dec", -90.5),
                  "ra_360": ("ra", 360.0), "ra_neg": ("ra", -1e-9)}[bad]
    a[field][1] = val
    with pytest.raises(xm.XmatchError) as ei:
        gpu_ctx.catalog(a, 1 / 3600)
    assert ei.value.status == xm.XM_EINVAL
    # the context stays usable after the error
    got, _ = gpu_join(gpu_ctx, a[:1], a[:1], 1 / 3600)
    assert len(got[0]) == 1


@pytest.mark.parametrize("bucketed", [False, True])
def test_failed_builds_return_their_memory(gpu_ctx, bucketed):
    # A build that fails (bad records; a sort whose look-back fails) returns every
    # device block it took: free memory, with the context's cached blocks counted
    # as free, is the same before and after.  Also with the records moved into
    # zone buckets (forced at this size), whose build takes the bucket array too.
    A, _, r = synth.config_small_patch(seed=5, n_a=2_000_000, n_b=10)
    default_min = gpu_ctx.get_param("bucket_min_rows")
    try:
        gpu_ctx.set_param("bucket_min_rows", 1 if bucketed else 0)
        gpu_ctx.catalog(A, r).close()   # the context's blocks and error word exist
        gpu_ctx.synchronize()
        free0 = gpu_ctx.mem_info()[0]
        bad = A.copy()
        bad["dec"][1_000_000] = np.nan
        with pytest.raises(xm.XmatchError) as ei:
            gpu_ctx.catalog(bad, r)
        assert ei.value.status == xm.XM_EINVAL
        try:
            gpu_ctx.set_param("sort_fail_test", 1)
            with pytest.raises(xm.XmatchError):
                gpu_ctx.catalog(A, r)
        finally:
            gpu_ctx.set_param("sort_fail_test", 0)
        gpu_ctx.synchronize()
        free1 = gpu_ctx.mem_info()[0]
        # the catalog builds again, as before the failures
        ca = gpu_ctx.catalog(A, r)
        assert ca.info()["n"] == len(A)
        ca.close()
    finally:
        gpu_ctx.set_param("bucket_min_rows", default_min)
    # a leak would hold at least the 16-MB key buffers of the 2M-row sort
    assert abs(int(free1) - int(free0)) < (4 << 20)


def test_bad_join_arguments(gpu_ctx):
    a = synth.make_objs([1], [10.0], [0.0])
    ca = gpu_ctx.catalog(a, 1 / 3600)
    cb = gpu_ctx.catalog(a, 2 / 3600)
    with pytest.raises(xm.XmatchError) as ei:
        gpu_ctx.match(ca, cb, 1 / 3600)
    assert ei.value.status == xm.XM_EINVAL
    with pytest.raises(xm.XmatchError):
        gpu_ctx.match(ca, ca, -1.0)
    ca.close(); cb.close()


def test_segmentation_invariance(gpu_ctx):
    A, B, r = synth.config_small_patch(seed=5, n_a=30000, n_b=30000)
    whole, _ = gpu_join(gpu_ctx, A, B, r)
    parts = []
    for sa in np.array_split(A, 3):
        for sb in np.array_split(B, 2):
            parts.append(gpu_join(gpu_ctx, sa, sb, r)[0])
    x = np.concatenate([p[0] for p in parts])
    y = np.concatenate([p[1] for p in parts])
    d = np.concatenate([p[2] for p in parts])
    o = np.lexsort((y, x))
    assert_same((x[o], y[o], d[o]), whole)


def test_full_sky_2m_vs_oracle(gpu_ctx):
    A, B, r = synth.config_full_sky(seed=2, n_a=2_000_000, n_b=2_000_000)
    got, stats = gpu_join(gpu_ctx, A, B, r)
    ref, cnt = O.match(A, B, r, mode=0)
    assert_same(got, ref)
    ca, cb = gpu_ctx.catalog(A, r), gpu_ctx.catalog(B, r)
    w = gpu_ctx.work_counts(ca, cb, r)
    ca.close(); cb.close()
    assert w["w_ref"] == cnt["w_ref"]
    # the host count resolves ra to the key quantum (2^-24 turn), the oracle exactly
    assert abs(w["w_cand"] - cnt["w_cand"]) <= 1e-3 * cnt["w_cand"] + 10
    assert w["w_tile"] >= w["w_cand"] * 0.99


@pytest.mark.parametrize("dense_min", [0, 4, 16, 1000000])
def test_join_mode_does_not_change_the_set(gpu_ctx, dense_min):
    # dense_min 0 forces the dense tile loop on every chunk, 1e6 the sparse window
    # search on every chunk the pole rule allows; the set must not move
    A, B, r = synth.config_full_sky(seed=9, n_a=400_000, n_b=400_000)
    clus = synth.make_objs(np.arange(5_000_000, 5_003_000),
                           *synth.uniform_patch(np.random.default_rng(1), 3000, 150, 150.01, 20,
                                                20.001))
    A = np.concatenate([A, clus])
    B = np.concatenate([B, clus])
    B["id"][-3000:] += 10**9
    ref, _ = O.match(A, B, r)
    old = gpu_ctx.get_param("dense_min")
    try:
        gpu_ctx.set_param("dense_min", dense_min)
        assert gpu_ctx.get_param("dense_min") == dense_min
        got, _ = gpu_join(gpu_ctx, A, B, r)
    finally:
        gpu_ctx.set_param("dense_min", old)
    assert_same(got, ref)


def dense_band(seed, n=300_000):
    """A 10 deg x 0.05 deg band at 1": 180 zones of ~1700 rows each, so a zone is
    several chunks at every chunk size, B 70% copies (0.3") + 30% fresh."""
    rng = np.random.default_rng(seed)
    ra, dec = synth.uniform_patch(rng, n, 100.0, 110.0, 10.0, 10.05)
    a = synth.make_objs(np.arange(n), ra, dec)
    k = int(0.7 * n)
    pick = rng.integers(0, n, k)
    bra, bdec = synth.perturb(rng, ra[pick], dec[pick], 0.3)
    fra, fdec = synth.uniform_patch(rng, n - k, 100.0, 110.0, 10.0, 10.05)
    b = synth.make_objs(10**9 + np.arange(n), np.concatenate([bra, fra]),
                        np.concatenate([bdec, fdec]))
    return a, b, 1.0 / 3600


@pytest.mark.parametrize("chunk_rows", [64, 128, 256, 512])
def test_chunk_rows_do_not_change_the_set(gpu_ctx, chunk_rows):
    # rows per wavefront chunk, from one chunk per zone to many; both precisions
    a, b, r = dense_band(31)
    ref, _ = O.match(a, b, r)
    assert len(ref[0]) > 150_000
    old = gpu_ctx.get_param("chunk_rows")
    try:
        gpu_ctx.set_param("chunk_rows", chunk_rows)
        assert gpu_ctx.get_param("chunk_rows") == chunk_rows
        got, _ = gpu_join(gpu_ctx, a, b, r)
        got32, _ = gpu_join(gpu_ctx, a, b, r, precision=xm.FP32_FILTER_FP64_RECHECK)
        ca, cb = gpu_ctx.catalog(a, r), gpu_ctx.catalog(b, r)
        w = gpu_ctx.work_counts(ca, cb, r)
        ca.close(); cb.close()
    finally:
        gpu_ctx.set_param("chunk_rows", old)
    assert_same(got, ref)
    assert_same(got32, ref)
    assert w["w_tile"] >= w["w_cand"]


def test_chunk_rows_rejects_bad_values(gpu_ctx):
    for v in (-64, 1, 100, 1024):
        with pytest.raises(xm.XmatchError):
            gpu_ctx.set_param("chunk_rows", v)


def test_fp32_rows_and_fp64_rows_rounded_in_registers_agree(gpu_ctx):
    # the FP32 filter on the catalogs' stored FP32 rows (fp32_rows = 1, written by
    # the gather) and on the FP64 rows rounded to FP32 in registers: the same
    # pairs and distances, and -- the FP32 values being the same bits -- the same
    # band-hit count and largest filter error
    a, b, r = dense_band(32, n=100_000)
    ref, _ = O.match(a, b, r)
    cvt, st_cvt = gpu_join(gpu_ctx, a, b, r, precision=xm.FP32_FILTER_FP64_RECHECK)
    gpu_ctx.set_param("fp32_rows", 1)
    try:
        assert gpu_ctx.get_param("fp32_rows") == 1
        rows, st_rows = gpu_join(gpu_ctx, a, b, r, precision=xm.FP32_FILTER_FP64_RECHECK)
        seg = synth.make_objs(np.arange(5000), *synth.uniform_patch(
            np.random.default_rng(3), 5000, 100.0, 100.5, 10.0, 10.05))
        cs = gpu_ctx.catalog_segments([a, seg], r)  # the many-segment gather writes them too
        cb = gpu_ctx.catalog(b, r)
        res = gpu_ctx.match_batch(cs, cb, [(0, 0, r)], precision=xm.FP32_FILTER_FP64_RECHECK)
        seg_pairs = res.by_job()[0]
        res.close(); cs.close(); cb.close()
        ca = gpu_ctx.catalog(a, r)
    finally:
        gpu_ctx.set_param("fp32_rows", 0)
    cb = gpu_ctx.catalog(b, r)   # without FP32 rows: the join rounds in registers
    res = gpu_ctx.match(ca, cb, r, precision=xm.FP32_FILTER_FP64_RECHECK)
    mixed, st_mixed = res.to_host(), res.stats.as_dict()
    res.close(); ca.close(); cb.close()
    for got in (cvt, rows, seg_pairs, mixed):
        assert_same(got, ref)
    assert st_cvt["band_hits"] == st_rows["band_hits"] == st_mixed["band_hits"]
    assert st_cvt["max_dist_err_arcsec"] == st_rows["max_dist_err_arcsec"] > 0


@pytest.mark.parametrize("pct,n", [(115, 300_000), (70, 300_000), (10, 300_000),
                                   (115, 5_000_123), (70, 5_000_123)])
def test_bucketed_build_equals_plain(gpu_ctx, pct, n):
    # the key pass that moves the records into zone buckets, forced at this size:
    # with runs of 115% of the sampled estimate (the default), 70% (records spill
    # into the overflow run) and 10% (the overflow run fills too and the build falls
    # back to the plain one) the catalog is the plain build's row for row, and so
    # are the joins.  At 300,000 rows each of the pass's 512 blocks takes at most
    # one 4096-record chunk; at 5,000,123 rows a block takes three, the last block
    # a partial one, so the pipelined chunks are all exercised.
    A, B, r = synth.config_small_patch(seed=71, n_a=n, n_b=n)
    default_min = gpu_ctx.get_param("bucket_min_rows")
    try:
        gpu_ctx.set_param("bucket_min_rows", 0)   # the plain build
        cp, cbp = gpu_ctx.catalog(A, r), gpu_ctx.catalog(B, r)
        ep = cp.export()
        res = gpu_ctx.match(cp, cbp, r)
        ref = res.to_host()
        res.close(); cp.close(); cbp.close()
        gpu_ctx.set_param("bucket_min_rows", 1)
        gpu_ctx.set_param("bucket_capacity_pct", pct)
        assert gpu_ctx.get_param("bucket_capacity_pct") == pct
        ca, cb = gpu_ctx.catalog(A, r), gpu_ctx.catalog(B, r)
        eb = ca.export()
        res = gpu_ctx.match(ca, cb, r, precision=xm.FP32_FILTER_FP64_RECHECK)
        got = res.to_host()
        res.close(); ca.close(); cb.close()
    finally:
        gpu_ctx.set_param("bucket_min_rows", default_min)
        gpu_ctx.set_param("bucket_capacity_pct", 115)
    for f in ep:
        assert np.array_equal(ep[f], eb[f]), f
    assert_same(got, ref)


@pytest.mark.parametrize("n,r_arcsec,cap_deg", [(1_700_000, 0.5, 0.1), (1_950_000, 0.5, 0.1),
                                                (2_000_000, 4.0, 1.0)])
def test_bucketed_build_clustered_equals_plain(gpu_ctx, n, r_arcsec, cap_deg):
    # zone buckets come from a sample of the records; here 90% of them sit in a
    # small cap (nearly every record lands in one bucket) and 10% are spread over
    # the sky (zones the layout sample may miss fall into the first or last
    # bucket).  The catalog, forced bucketed, equals the plain build row for row,
    # and so does the self-join.  At 1.7M rows and 0.5" the bucket array would
    # need a wider row field than the input, costing the key an ra bit, so it is
    # cut to the input's field (its overflow run shrinks); at 1.95M rows the runs
    # also hold less than 115% of their estimates (103%); at 2M rows and 4" the
    # key keeps all 24 ra bits either way.
    rng = np.random.default_rng(zlib.crc32(b"clustered %d" % n))
    n_c = n * 9 // 10
    ra_c, dec_c = synth.uniform_cap(rng, n_c, 123.4, -45.6, cap_deg)
    ra_s, dec_s = synth.uniform_sphere(rng, n - n_c)
    A = synth.make_objs(np.arange(n), np.concatenate([ra_c, ra_s]),
                        np.concatenate([dec_c, dec_s]))
    A = A[rng.permutation(n)]
    r = r_arcsec / 3600
    default_min = gpu_ctx.get_param("bucket_min_rows")
    try:
        gpu_ctx.set_param("bucket_min_rows", 0)
        cp = gpu_ctx.catalog(A, r)
        ep = cp.export()
        res = gpu_ctx.match(cp, cp, r)
        ref = res.to_host()
        res.close(); cp.close()
        gpu_ctx.set_param("bucket_min_rows", 1)
        gpu_ctx.reset_timing()
        gpu_ctx.set_timing(True, only="bucket_plan")
        ca = gpu_ctx.catalog(A, r)
        gpu_ctx.set_timing(False)
        _, planned = gpu_ctx.kernel_time("bucket_plan")
        gpu_ctx.reset_timing()
        eb = ca.export()
        res = gpu_ctx.match(ca, ca, r)
        got = res.to_host()
        res.close(); ca.close()
    finally:
        gpu_ctx.set_param("bucket_min_rows", default_min)
    assert planned == 1   # the build took the bucketed path (its plan was made)
    for f in ep:
        assert np.array_equal(ep[f], eb[f]), f
    assert_same(got, ref)
    assert len(ref[0]) >= n   # every source matches itself


def test_bucketed_build_duplicate_positions(gpu_ctx):
    # 2000 distinct positions on a 10" grid, 100 records at each, in random order:
    # every sort key ties 100 ways.  The forced bucketed catalog orders the ties by
    # input row as the plain build does, and the self-join at 1" pairs exactly the
    # records of one position: 2000 x 100^2 pairs, all at distance 0.
    rng = np.random.default_rng(zlib.crc32(b"duplicates"))
    g = np.arange(2000)
    ra0 = 10.0 + (g % 50) * (10.0 / 3600)
    dec0 = 5.0 + (g // 50) * (10.0 / 3600)
    pos = np.repeat(g, 100)
    perm = rng.permutation(len(pos))
    A = synth.make_objs(np.arange(len(pos)), ra0[pos[perm]], dec0[pos[perm]])
    r = 1.0 / 3600
    default_min = gpu_ctx.get_param("bucket_min_rows")
    counts = []
    try:
        exports = []
        for bmr in (0, 1):
            gpu_ctx.set_param("bucket_min_rows", bmr)
            c = gpu_ctx.catalog(A, r)
            exports.append(c.export())
            res = gpu_ctx.match(c, c, r)
            counts.append(res.count)
            res.close(); c.close()
    finally:
        gpu_ctx.set_param("bucket_min_rows", default_min)
    for f in exports[0]:
        assert np.array_equal(exports[0][f], exports[1][f]), f
    # ties in input order: within one position the ids ascend
    ids = exports[0]["id"]
    same = (exports[0]["x"][1:] == exports[0]["x"][:-1]) & (exports[0]["y"][1:] == exports[0]["y"][:-1])
    assert np.all(ids[1:][same] > ids[:-1][same])
    assert counts == [2000 * 100 * 100] * 2


def test_bucket_knobs_reject_bad_values(gpu_ctx):
    for name, v in (("bucket_min_rows", -1), ("bucket_capacity_pct", 0),
                    ("bucket_capacity_pct", 1001)):
        with pytest.raises(xm.XmatchError):
            gpu_ctx.set_param(name, v)


def test_autotune_keeps_the_sets(gpu_ctx):
    # a batch of jobs of different densities and radii, with and without the
    # tuner: the same pairs per job; the tune table names the measured keys
    rng = np.random.default_rng(41)
    segs_a, segs_b, rs = [], [], []
    for j in range(12):
        n = int(10 ** rng.uniform(2.5, 5.0))
        cap = float(rng.uniform(0.05, 0.5))
        ra_c, dec_c = float(rng.uniform(0, 360)), float(rng.uniform(-80, 80))
        ra, dec = synth.uniform_cap(rng, n, ra_c, dec_c, cap)
        segs_a.append(synth.make_objs(10**6 * j + np.arange(n), ra, dec))
        pb = rng.integers(0, n, n // 2)
        bra, bdec = synth.perturb(rng, ra[pb], dec[pb], 0.3)
        fra, fdec = synth.uniform_cap(rng, n - n // 2, ra_c, dec_c, cap)
        segs_b.append(synth.make_objs(10**9 + 10**6 * j + np.arange(n),
                                      np.concatenate([bra, fra]), np.concatenate([bdec, fdec])))
        rs.append(float(10 ** rng.uniform(np.log10(0.5), 1.0)) / 3600)
    jobs = [(j, j, rs[j]) for j in range(12)]
    out = {}
    for tune in (0, 1):
        gpu_ctx.set_param("autotune", tune)
        try:
            ca = gpu_ctx.catalog_segments(segs_a, rs)
            cb = gpu_ctx.catalog_segments(segs_b, rs)
            res = gpu_ctx.match_batch(ca, cb, jobs)
            out[tune] = res.by_job()
            res.close(); ca.close(); cb.close()
            table = gpu_ctx.tune_table()
        finally:
            gpu_ctx.set_param("autotune", 0)
    assert len(table) >= 1
    for e in table:
        assert e["chunk_rows"] in (64, 128, 256, 512) and e["dense_min"] in (4, 16, 64)
        assert 1 <= e["sample_jobs"] <= 16 and e["tuned_ms"] <= e["default_ms"] + 1e-9
    for j in range(12):
        ref, _ = O.match(segs_a[j], segs_b[j], rs[j])
        assert_same(out[0][j], ref)
        assert_same(out[1][j], ref)


@pytest.mark.parametrize("ra_bits", [6, 12, 18])
def test_coarse_ra_keys_do_not_change_the_set(gpu_ctx, ra_bits):
    # keys that floor ra to 2^-ra_bits turn (as 10^8-row catalogs do) only widen
    # the windows; A and B may even use different granularities
    A, B, r = synth.config_small_patch(seed=4, n_a=60_000, n_b=60_000)
    ref, _ = O.match(A, B, r)
    old = gpu_ctx.get_param("key_ra_bits")
    try:
        gpu_ctx.set_param("key_ra_bits", ra_bits)
        ca = gpu_ctx.catalog(A, r)
        gpu_ctx.set_param("key_ra_bits", 24 if ra_bits != 24 else 20)
        cb = gpu_ctx.catalog(B, r)
    finally:
        gpu_ctx.set_param("key_ra_bits", old)
    res = gpu_ctx.match(ca, cb, r)
    got = res.to_host()
    res.close()
    res = gpu_ctx.match(cb, ca, r)  # the other way round: A coarse on the B side
    got2 = res.to_host()
    res.close(); ca.close(); cb.close()
    assert_same(got, ref)
    o = np.lexsort((got2[0], got2[1]))
    assert np.array_equal(got2[1][o], ref[0]) and np.array_equal(got2[0][o], ref[1])


# ---------------------------------------------------------------- FP32 filter
F32 = xm.FP32_FILTER_FP64_RECHECK


def test_fp32_filter_small_patch_identical(gpu_ctx):
    A, B, r = synth.config_small_patch(seed=1)
    got, stats = gpu_join(gpu_ctx, A, B, r, precision=F32)
    ref, _ = O.match(A, B, r)
    assert_same(got, ref)  # distances are the FP64 ones
    assert stats["band_hits"] >= 0 and stats["max_dist_err_arcsec"] < 0.1


@pytest.mark.parametrize("where", ["seam", "north_pole", "south_pole", "dense_cluster", "mid"])
def test_fp32_filter_edge_geometry(gpu_ctx, where):
    rng = np.random.default_rng(zlib.crc32(where.encode()) + 1)
    n = 3000
    if where == "seam":
        ra = np.where(rng.random(n) < 0.5, rng.uniform(0, 0.01, n),
                      rng.uniform(359.99, 360, n) % 360)
        dec = rng.uniform(-0.02, 0.02, n)
    elif where == "north_pole":
        ra, dec = synth.uniform_patch(rng, n, 0, 360, 89.97, 90.0)
    elif where == "south_pole":
        ra, dec = synth.uniform_patch(rng, n, 0, 360, -90.0, -89.97)
    elif where == "dense_cluster":
        ra, dec = synth.uniform_patch(rng, n, 150.0, 150.003, 20.0, 20.0003)
    else:
        ra, dec = synth.uniform_patch(rng, n, 200, 200.04, 45, 45.03)
    a = synth.make_objs(np.arange(n), ra, dec)
    bra, bdec = synth.perturb(rng, ra, dec, 2.0)
    b = synth.make_objs(10**9 + np.arange(n), bra, bdec)
    for r in (1.5 / 3600, 4 / 3600, 60 / 3600):
        got, _ = gpu_join(gpu_ctx, a, b, r, precision=F32)
        assert_same(got, O.brute(a, b, r))


def test_fp32_filter_boundary_pairs(gpu_ctx):
    # pairs at r(1 + eps), eps down to 2^-50, sit inside the guard band: only the FP64
    # recheck can decide them, and it must decide them as the oracle does
    r = 1.0 / 3600
    recs_a, recs_b = [], []
    k = 0
    for dec0 in (0.0, 0.3, 37.0, -62.0, 88.0):
        for ra0 in (0.0, 45.0, 180.0, 359.9999999):
            for e in (-2.0**-20, -2.0**-30, -2.0**-45, -2.0**-50, 0.0, 2.0**-50, 2.0**-45,
                      2.0**-30, 2.0**-20):
                dist = r * (1 + e)
                c = np.cos(np.radians(dec0))
                dra = np.degrees(2 * np.arcsin(np.sin(np.radians(dist) / 2) / c))
                for (pa, pb) in (((ra0, dec0), (ra0, dec0 + dist)),
                                 ((ra0, dec0), ((ra0 + dra) % 360, dec0)),
                                 ((ra0, dec0), ((ra0 - dra) % 360, dec0))):
                    recs_a.append((k, pa[0], pa[1]))
                    recs_b.append((10**9 + k, pb[0], pb[1]))
                    k += 1
    a = np.array(recs_a, dtype=synth.OBJ_DTYPE)
    b = np.array(recs_b, dtype=synth.OBJ_DTYPE)
    b["dec"] = np.clip(b["dec"], -90, 90)
    got, stats = gpu_join(gpu_ctx, a, b, r, precision=F32)
    assert_same(got, O.match(a, b, r)[0])
    assert stats["band_hits"] > 0


def test_fp32_filter_full_sky_2m_same_as_fp64(gpu_ctx):
    A, B, r = synth.config_full_sky(seed=7, n_a=2_000_000, n_b=2_000_000)
    ca, cb = gpu_ctx.catalog(A, r), gpu_ctx.catalog(B, r)
    r64 = gpu_ctx.match(ca, cb, r)
    r32 = gpu_ctx.match(ca, cb, r, precision=F32)
    g64, g32 = r64.to_host(), r32.to_host()
    s32 = r32.stats.as_dict()
    r64.close(); r32.close(); ca.close(); cb.close()
    assert np.array_equal(g64[0], g32[0]) and np.array_equal(g64[1], g32[1])
    assert np.array_equal(g64[2], g32[2])  # bit-identical FP64 distances
    assert_same(g64, O.match(A, B, r)[0])
    assert s32["max_dist_err_arcsec"] < 0.1
