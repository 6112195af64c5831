"""The serving form at its own shape: a resident catalog of 5e8 sources with the
galactic-plane-like density (1 + 20 exp(-b^2 / (2 (10 deg)^2)), pole (0, 30) deg),
zone-sorted at 1 arcsec and built with FP32 rows beside the FP64 ones, answering
requests shaped as the serving workload draws them: N ~ logU[1, 1e5] sources in a
cap of radius U[1, 30] arcmin around a resident source, 70% of them 0.5" copies of
resident sources inside the cap and 30% fresh in the cap, radius U[1, 5] arcsec.

Both server forms (FP64 plain, FP32 filter + FP64 recheck) answer a burst of
requests; every request's pairs equal the oracle's join of the request with the
resident rows of the zones around it (selected on the host from the catalog's
zone-sorted export), and a request answered alone equals the same request answered
in the middle of the burst."""
import numpy as np
import pytest

import xmatch_amd as xm
from oracle import oracle as O
from xmatch_amd import synth

pytestmark = pytest.mark.gpu

N_RES = 500_000_000
H = 1.0 / 3600.0
N_REQ = 24
DIST_TOL_ARCSEC = 1e-6


@pytest.fixture(scope="module")
def resident(gpu_ctx):
    shp = xm.shape(xm.SHAPE_PLANE, synth.PLANE_POLE_RA, synth.PLANE_POLE_DEC, 20.0, 10.0)
    d = gpu_ctx.synth_fresh(shp, N_RES, 5005, stream=0, id0=0)
    gpu_ctx.set_param("fp32_rows", 1)
    try:
        cat = gpu_ctx.catalog(d, H)
    finally:
        gpu_ctx.set_param("fp32_rows", 0)
    recs = gpu_ctx.download(d)
    d.free()
    e = cat.export(fields=("id", "zone"))
    rows = recs[e["id"]]        # ids are 0..N_RES-1: the records in catalog order
    zones = e["zone"]
    del recs, e
    yield cat, rows, zones
    cat.close()
    gpu_ctx.release_memory()


def zone_rows(rows, zones, dec_lo, dec_hi, pad):
    """The resident rows of the zones covering [dec_lo, dec_hi], pad zones more on
    either side (the zone formula of Sorter.cu:20-24)."""
    z0 = int(np.floor((max(dec_lo, -90.0) + 90.0) / H)) - pad
    z1 = int(np.floor((min(dec_hi, 90.0) + 90.0) / H)) + pad
    i0, i1 = np.searchsorted(zones, [z0, z1 + 1])
    return rows[i0:i1]


def ang_deg(ra1, dec1, ra2, dec2):
    r1, d1, r2, d2 = map(np.radians, (ra1, dec1, ra2, dec2))
    c = np.sin(d1) * np.sin(d2) + np.cos(d1) * np.cos(d2) * np.cos(r1 - r2)
    return np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))


def make_requests(rows, zones, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    reqs = []
    id0 = 10**12
    for i in range(N_REQ):
        c = rows[rng.integers(0, len(rows))]
        cap = float(rng.uniform(1.0, 30.0)) / 60.0
        n = max(1, int(round(10.0 ** rng.uniform(0.0, 5.0))))
        r = float(rng.uniform(1.0, 5.0)) / 3600.0
        near = zone_rows(rows, zones, c["dec"] - cap, c["dec"] + cap, 1)
        members = near[ang_deg(near["ra"], near["dec"], c["ra"], c["dec"]) < cap]
        ra, dec = synth.uniform_cap(rng, n, c["ra"], c["dec"], cap)
        k = int(round(0.7 * n)) if len(members) else 0
        if k:
            pick = members[rng.integers(0, len(members), k)]
            ra[:k], dec[:k] = synth.perturb(rng, pick["ra"], pick["dec"], 0.5)
        reqs.append((synth.make_objs(id0 + np.arange(n), ra, dec), r))
        id0 += n
    return reqs


def oracle_pairs(q, r, rows, zones):
    nz = int(np.ceil(r / H))
    b = zone_rows(rows, zones, q["dec"].min(), q["dec"].max(), nz + 1)
    return O.match(q, b, r, zone_height_deg=H, threads=16)[0]


def sorted_result(res):
    ia, ib, d, _ = res
    o = np.lexsort((ib, ia))
    return ia[o], ib[o], d[o]


def same(got, ref):
    assert len(got[0]) == len(ref[0]), (len(got[0]), len(ref[0]))
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    if len(got[2]):
        assert np.abs(got[2] - ref[2]).max() <= DIST_TOL_ARCSEC


@pytest.fixture(scope="module")
def requests(resident):
    _, rows, zones = resident
    reqs = make_requests(rows, zones, 55)
    sizes = np.array([len(q) for q, _ in reqs])
    assert sizes.max() > 10_000 and sizes.min() < 100   # the logU[1, 1e5] spread
    refs = [oracle_pairs(q, r, rows, zones) for q, r in reqs]
    # most requests find pairs: 70% of their sources copy resident ones
    assert sum(len(x[0]) > 0 for x in refs) >= N_REQ // 2
    return reqs, refs


@pytest.mark.parametrize("precision", [xm.FP64_PLAIN, xm.FP32_FILTER_FP64_RECHECK],
                         ids=["fp64", "fp32_filter"])
def test_c5_requests_match_oracle_alone_and_in_mix(gpu_ctx, resident, requests, precision):
    cat, _, _ = resident
    reqs, refs = requests
    k = N_REQ // 2
    with xm.Server(gpu_ctx, cat, precision=precision) as srv:
        alone = sorted_result(srv.result(srv.submit(*reqs[k])))
        s0 = srv.stats()
        tickets = [srv.submit(q, r) for q, r in reqs]   # a burst: batched together
        got = [sorted_result(srv.result(t)) for t in tickets]
        s1 = srv.stats()
    assert s1["requests_done"] - s0["requests_done"] == N_REQ
    # the burst went through in fewer batches than requests
    assert s1["batches"] - s0["batches"] < N_REQ
    for g, ref in zip(got, refs):
        same(g, ref)
    same(alone, refs[k])
    for a, m in zip(alone, got[k]):
        assert np.array_equal(a, m)   # bit for bit, distances included
