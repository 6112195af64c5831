"""GPU parity of many-segment catalogs and batched joins (xm_catalog_create_segments,
xm_match_batch) against the CPU oracle, per job, and against one-job joins.

The reference runs a queue of jobs, each a (segment A, segment B, radius) triple
(JobManager.cpp:5-41, Job.h:16-29); a batch must give every job exactly the
pairs that job alone gives.  Bar: identical (idA, idB) sets per job, distances
within 1e-6 arcsec."""
import numpy as np
import pytest

import xmatch_amd as xm
from oracle import oracle as O
from xmatch_amd import synth

pytestmark = pytest.mark.gpu

DIST_TOL_ARCSEC = 1e-6


def assert_same(got, ref, what=""):
    ia, ib, d = got
    ra, rb, rd = ref
    assert len(ia) == len(ra), (what, len(ia), len(ra))
    assert np.array_equal(ia, ra) and np.array_equal(ib, rb), what
    if len(d):
        assert np.abs(d - rd).max() <= DIST_TOL_ARCSEC, what


def batch(ctx, a_segs, b_segs, hs_a, hs_b, jobs, precision=xm.FP64_PLAIN, capacity_hint=0):
    A = ctx.catalog_segments(a_segs, hs_a)
    B = ctx.catalog_segments(b_segs, hs_b)
    res = ctx.match_batch(A, B, jobs, precision=precision, capacity_hint=capacity_hint)
    per_job = res.by_job()
    stats = res.stats.as_dict()
    res.close(); A.close(); B.close()
    return per_job, stats


def small_caps(seed, n_jobs, n_hi=3000):
    spec = synth.batch_job_specs(seed=seed, n_jobs=n_jobs, n_lo=1, n_hi=n_hi, cap_lo=0.01,
                                 cap_hi=0.05)
    return synth.batch_job_catalogs(spec, seed=seed)


def test_batch_equals_oracle_per_job(gpu_ctx):
    jobs_data = small_caps(seed=7, n_jobs=40)
    a_segs = [a for a, _, _ in jobs_data]
    b_segs = [b for _, b, _ in jobs_data]
    rs = [r for _, _, r in jobs_data]
    jobs = [(j, j, rs[j]) for j in range(len(rs))]
    per_job, stats = batch(gpu_ctx, a_segs, b_segs, rs, rs, jobs)
    total = 0
    for j, (a, b, r) in enumerate(jobs_data):
        ref, _ = O.match(a, b, r)
        assert_same(per_job[j], ref, "job %d" % j)
        total += len(ref[0])
    assert stats["n_matches"] == total and total > 1000


def test_batch_equals_one_job_joins(gpu_ctx):
    jobs_data = small_caps(seed=8, n_jobs=12, n_hi=20000)
    rs = [r for _, _, r in jobs_data]
    per_job, _ = batch(gpu_ctx, [a for a, _, _ in jobs_data], [b for _, b, _ in jobs_data],
                       rs, rs, [(j, j, rs[j]) for j in range(len(rs))])
    for j, (a, b, r) in enumerate(jobs_data):
        ca, cb = gpu_ctx.catalog(a, r), gpu_ctx.catalog(b, r)
        res = gpu_ctx.match(ca, cb, r)
        assert_same(per_job[j], res.to_host(), "job %d" % j)
        res.close(); ca.close(); cb.close()


def test_cross_product_of_segments_is_the_whole_join(gpu_ctx):
    A, B, r = synth.config_small_patch(seed=9, n_a=30000, n_b=30000)
    sa = np.array_split(A, 3)
    sb = np.array_split(B, 2)
    jobs = xm.cross_product_jobs(3, 2, r)
    per_job, _ = batch(gpu_ctx, sa, sb, r, r, jobs)
    for j, (ia, ib, _) in enumerate(jobs):
        assert_same(per_job[j], O.match(sa[ia], sb[ib], r)[0], "job %d" % j)
    x = np.concatenate([p[0] for p in per_job])
    y = np.concatenate([p[1] for p in per_job])
    d = np.concatenate([p[2] for p in per_job])
    o = np.lexsort((y, x))
    assert_same((x[o], y[o], d[o]), O.match(A, B, r)[0])


def test_self_join_of_a_segment_catalog(gpu_ctx):
    A, _, r = synth.config_small_patch(seed=10, n_a=20000, n_b=10)
    segs = np.array_split(A, 4)
    cat = gpu_ctx.catalog_segments(segs, r)
    jobs = [(i, k, r) for i in range(4) for k in range(4)]
    res = gpu_ctx.match_batch(cat, cat, jobs)
    per_job = res.by_job()
    for j, (ia, ib, _) in enumerate(jobs):
        assert_same(per_job[j], O.match(segs[ia], segs[ib], r)[0], "job %d" % j)
    res.close(); cat.close()


def test_empty_and_single_segments(gpu_ctx):
    r = 2.0 / 3600
    one = synth.make_objs([5], [100.0], [-30.0])
    near = synth.make_objs([6], [100.0], [-30.0 + 1.0 / 3600])
    empty = one[:0]
    rng = np.random.default_rng(3)
    ra, dec = synth.uniform_patch(rng, 500, 100.0, 100.01, -30.0, -29.99)
    a3 = synth.make_objs(np.arange(1000, 1500), ra, dec)
    ra, dec = synth.perturb(rng, ra, dec, 0.5)
    b3 = synth.make_objs(np.arange(5000, 5500), ra, dec)
    a_segs = [empty, one, a3, empty]
    b_segs = [near, empty, b3, empty]
    jobs = [(0, 0, r), (1, 0, r), (1, 1, r), (2, 2, r), (2, 0, r), (3, 3, r), (2, 3, r),
            (1, 2, r)]
    per_job, _ = batch(gpu_ctx, a_segs, b_segs, r, r, jobs)
    for j, (ia, ib, _) in enumerate(jobs):
        assert_same(per_job[j], O.brute(a_segs[ia], b_segs[ib], r), "job %d" % j)
    assert len(per_job[1][0]) == 1 and len(per_job[3][0]) > 400


def test_all_segments_empty(gpu_ctx):
    e = synth.make_objs([], [], [])
    per_job, stats = batch(gpu_ctx, [e, e], [e], 1 / 3600, 1 / 3600, [(0, 0, 1 / 3600),
                                                                    (1, 0, 1 / 3600)])
    assert stats["n_matches"] == 0 and all(len(p[0]) == 0 for p in per_job)


def test_many_small_segments(gpu_ctx):
    # 3000 segments of 0..20 sources: blocks of the segment passes span many
    # segments, and the wave -> job search crosses many jobs per block
    rng = np.random.default_rng(12)
    r = 3.0 / 3600
    a_segs, b_segs = [], []
    k = 0
    for s in range(3000):
        n = int(rng.integers(0, 21))
        ra, dec = synth.uniform_patch(rng, n, 50.0 + 0.01 * (s % 7), 50.0 + 0.01 * (s % 7) + 0.002,
                                      10.0, 10.002)
        a_segs.append(synth.make_objs(np.arange(k, k + n), ra, dec))
        rb, db = synth.perturb(rng, ra, dec, 1.0)
        b_segs.append(synth.make_objs(10**9 + np.arange(k, k + n), rb, db))
        k += n
    jobs = [(s, s, r) for s in range(3000)] + [(s, (s + 7) % 3000, r) for s in range(0, 3000, 5)]
    per_job, _ = batch(gpu_ctx, a_segs, b_segs, r, r, jobs)
    for j, (ia, ib, _) in enumerate(jobs):
        assert_same(per_job[j], O.brute(a_segs[ia], b_segs[ib], r), "job %d" % j)


def test_per_segment_zone_heights(gpu_ctx):
    # each job at its own radius with zone height = radius (and one at h = r / 2)
    jobs_data = small_caps(seed=13, n_jobs=6)
    hs = [r for _, _, r in jobs_data]
    hs[2] = hs[2] / 2
    jobs = [(j, j, jobs_data[j][2]) for j in range(6)]
    per_job, _ = batch(gpu_ctx, [a for a, _, _ in jobs_data], [b for _, b, _ in jobs_data], hs, hs,
                       jobs)
    for j, (a, b, r) in enumerate(jobs_data):
        assert_same(per_job[j], O.match(a, b, r, zone_height_deg=hs[j])[0], "job %d" % j)


def test_batch_overflow_retry(gpu_ctx):
    jobs_data = small_caps(seed=14, n_jobs=10)
    rs = [r for _, _, r in jobs_data]
    jobs = [(j, j, rs[j]) for j in range(10)]
    segs_a = [a for a, _, _ in jobs_data]
    segs_b = [b for _, b, _ in jobs_data]
    per_job, stats = batch(gpu_ctx, segs_a, segs_b, rs, rs, jobs, capacity_hint=64)
    assert stats["n_retries"] >= 1
    for j, (a, b, r) in enumerate(jobs_data):
        assert_same(per_job[j], O.match(a, b, r)[0], "job %d" % j)


def test_batch_fp32_filter_same_as_fp64(gpu_ctx):
    jobs_data = small_caps(seed=15, n_jobs=20)
    rs = [r for _, _, r in jobs_data]
    jobs = [(j, j, rs[j]) for j in range(20)]
    sa = [a for a, _, _ in jobs_data]
    sb = [b for _, b, _ in jobs_data]
    p64, _ = batch(gpu_ctx, sa, sb, rs, rs, jobs)
    p32, st = batch(gpu_ctx, sa, sb, rs, rs, jobs, precision=xm.FP32_FILTER_FP64_RECHECK)
    for j in range(20):
        assert_same(p32[j], p64[j], "job %d" % j)


def test_device_segments_and_layout(gpu_ctx):
    jobs_data = small_caps(seed=16, n_jobs=5)
    segs = [a for a, _, _ in jobs_data]
    hs = [r for _, _, r in jobs_data]
    bufs = [gpu_ctx.to_device(s) for s in segs]
    cat = gpu_ctx.catalog_segments(bufs, hs)
    assert cat.segment_count == 5
    e = cat.export()
    first = 0
    for s, recs in enumerate(segs):
        info = cat.segment_info(s)
        assert info["first_row"] == first and info["n"] == len(recs)
        sl = slice(first, first + len(recs))
        first += len(recs)
        zones = O.zone(recs["dec"], hs[s])
        assert info["zone_height_deg"] == hs[s]
        if len(recs):
            assert info["zmin"] == zones.min() and info["zmax"] == zones.max()
        assert np.all(np.diff(e["zone"][sl]) >= 0)
        o = np.argsort(recs["id"])
        p = np.argsort(e["id"][sl])
        assert np.array_equal(e["id"][sl][p], recs["id"][o])
        x, y, z = O.unitvec(recs["ra"][o], recs["dec"][o])
        assert np.array_equal(e["x"][sl][p].view(np.int64), x.view(np.int64))
        assert np.array_equal(e["z"][sl][p].view(np.int64), z.view(np.int64))
        assert np.array_equal(e["zone"][sl][p], zones[o])
    cat.close()
    for b in bufs:
        b.free()


def test_export_zones_with_empty_segments(gpu_ctx):
    # zones come from the zone table, per segment: empty segments first, between and
    # last own no entries, and every row of the others gets its own zone
    rng = np.random.default_rng(31)
    h = 2.0 / 3600
    empty = synth.make_objs([], [], [])
    segs = [empty]
    for k, (d0, d1) in enumerate([(-89.999, -89.99), (10.0, 10.02), (89.995, 90.0)]):
        ra, dec = synth.uniform_patch(rng, 300 + 50 * k, 0.0, 360.0, d0, d1)
        segs += [synth.make_objs(np.arange(300 + 50 * k) + 1000 * k, ra, dec), empty]
    cat = gpu_ctx.catalog_segments(segs, [h] * len(segs))
    e = cat.export()
    first = 0
    for s, recs in enumerate(segs):
        sl = slice(first, first + len(recs))
        first += len(recs)
        o = np.argsort(recs["id"])
        p = np.argsort(e["id"][sl])
        assert np.array_equal(e["zone"][sl][p], O.zone(recs["dec"], h)[o]), "segment %d" % s
    assert first == len(e["zone"])
    cat.close()


def test_batch_work_counts_equal_one_job_counts(gpu_ctx):
    jobs_data = small_caps(seed=17, n_jobs=4)
    rs = [r for _, _, r in jobs_data]
    A = gpu_ctx.catalog_segments([a for a, _, _ in jobs_data], rs)
    B = gpu_ctx.catalog_segments([b for _, b, _ in jobs_data], rs)
    for j, (a, b, r) in enumerate(jobs_data):
        wb = gpu_ctx.work_counts_job(A, B, (j, j, r))
        ca, cb = gpu_ctx.catalog(a, r), gpu_ctx.catalog(b, r)
        w1 = gpu_ctx.work_counts(ca, cb, r)
        assert wb["w_ref"] == w1["w_ref"] and wb["w_cand"] == w1["w_cand"]
        ca.close(); cb.close()
    A.close(); B.close()


def test_batch_bad_arguments(gpu_ctx):
    a = synth.make_objs([1, 2], [10.0, 11.0], [0.0, 0.0])
    A = gpu_ctx.catalog_segments([a, a], [1 / 3600, 2 / 3600])
    for jobs in ([(2, 0, 1 / 3600)], [(0, 5, 1 / 3600)], [(0, 1, 1 / 3600)], [(0, 0, -1.0)], []):
        with pytest.raises(xm.XmatchError) as ei:
            gpu_ctx.match_batch(A, A, jobs)
        assert ei.value.status == xm.XM_EINVAL
    with pytest.raises(xm.XmatchError) as ei:
        gpu_ctx.match(A, A, 1 / 3600)
    assert ei.value.status == xm.XM_EINVAL
    bad = a.copy()
    bad["dec"][1] = np.nan
    with pytest.raises(xm.XmatchError) as ei:
        gpu_ctx.catalog_segments([a, bad], 1 / 3600)
    assert ei.value.status == xm.XM_EINVAL and "segment 1" in str(ei.value)
    # the context keeps working
    res = gpu_ctx.match_batch(A, A, [(1, 1, 1 / 3600)])
    assert res.count == 2
    res.close(); A.close()
