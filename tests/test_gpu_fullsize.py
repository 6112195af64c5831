"""The two full-size workloads, checked through properties that do not depend on
size and against the oracle on parts of them, since the oracle cannot join them
whole in a test's time.

Full-sky 50M x 50M at 2 arcsec:
- the catalog is ordered by (zone, ra key) and its unit vectors equal the oracle's
  bit for bit on a sample of rows;
- joining B against A gives exactly the swapped pairs of A against B;
- every pair is within the radius by the oracle's chord test, and its distance
  agrees with one recomputed on the host;
- the pairs of A rows in three regions equal the oracle's join of those rows: a
  band of 1000 zones at the equator (0.49% of A), both polar caps above |dec| =
  89.5 deg (ra windows of whole zones), and a strip across the 0/360 seam.

Galactic-plane density 200M x 200M at 1 arcsec (device-generated): the FP32 filter
with FP64 recheck returns the FP64 set with the same distances, and in the 1000-zone
band of the highest zone populations (zones of several chunks, B zones of hundreds
of rows) both forms equal the oracle's join of the band."""
import numpy as np
import pytest

from oracle import oracle as O
from xmatch_amd import synth

pytestmark = pytest.mark.gpu

N = 50_000_000
DIST_TOL_ARCSEC = 1e-6


@pytest.fixture(scope="module")
def full(gpu_ctx):
    A, B, r = synth.config_full_sky(seed=2, n_a=N, n_b=N)
    dA = gpu_ctx.to_device(A)
    dB = gpu_ctx.to_device(B)
    ca = gpu_ctx.catalog(dA, r)
    cb = gpu_ctx.catalog(dB, r)
    res = gpu_ctx.match(ca, cb, r)
    ab = res.to_host()
    res.close()
    res = gpu_ctx.match(cb, ca, r)
    ba = res.to_host()
    res.close()
    ea = ca.export()
    info = ca.info()
    ca.close()
    cb.close()
    del dA, dB
    gpu_ctx.release_memory()
    return A, B, r, ab, ba, ea, info


def _pair_key(ia, ib):
    # A ids are 0..N-1 (< 2^26), B ids 1e9 + 0..N-1 (< 2^31): one sortable int64
    return ia.astype(np.int64) * (1 << 31) + ib.astype(np.int64)


def test_full_catalog_order_and_unit_vectors(full):
    A, _, r, _, _, e, info = full
    assert np.array_equal(np.sort(e["id"]), np.arange(N, dtype=np.int64))
    assert np.all(np.diff(e["zone"]) >= 0)
    assert info["zmin"] == e["zone"][0] and info["zmax"] == e["zone"][-1]
    # within a zone, ra ascends up to the key's ra granularity (at least 13 ra bits are
    # kept at 50M rows: 6 of 19 are dropped when that saves a sort pass)
    ra = A["ra"][e["id"]]
    same = np.diff(e["zone"]) == 0
    assert np.all(np.diff(ra)[same] >= -360.0 / 2**13)
    rng = np.random.default_rng(11)
    s = rng.choice(N, 1_000_000, replace=False)
    x, y, z = O.unitvec(A["ra"][e["id"][s]], A["dec"][e["id"][s]])
    assert np.array_equal(e["x"][s].view(np.int64), x.view(np.int64))
    assert np.array_equal(e["y"][s].view(np.int64), y.view(np.int64))
    assert np.array_equal(e["z"][s].view(np.int64), z.view(np.int64))


def test_full_swap_symmetry(full):
    _, _, _, ab, ba, _, _ = full
    assert len(ab[0]) == len(ba[0]) > 29_000_000
    k1 = _pair_key(ab[0], ab[1])
    k2 = _pair_key(ba[1], ba[0])
    o1 = np.argsort(k1, kind="stable")
    o2 = np.argsort(k2, kind="stable")
    assert np.array_equal(k1[o1], k2[o2])
    assert np.abs(ab[2][o1] - ba[2][o2]).max() <= DIST_TOL_ARCSEC


def test_full_pairs_within_radius(full):
    A, B, r, ab, _, _, _ = full
    ia, ib, d = ab
    xa, ya, za = O.unitvec(A["ra"][ia], A["dec"][ia])
    jb = ib - synth.B_ID_OFFSET
    xb, yb, zb = O.unitvec(B["ra"][jb], B["dec"][jb])
    dx, dy, dz = xa - xb, ya - yb, za - zb
    d2 = dx * dx + dy * dy + dz * dz
    assert np.all(d2 < O.sr_dist2(r))
    d_host = np.degrees(2.0 * np.arcsin(np.sqrt(d2) / 2.0)) * 3600.0
    assert np.abs(d_host - d).max() <= DIST_TOL_ARCSEC


def test_full_equator_band_vs_oracle(full):
    A, B, r, ab, _, _, _ = full
    nzn = int(np.ceil(180.0 / r))
    z_lo, z_hi = nzn // 2 - 500, nzn // 2 + 499
    dec_lo, dec_hi = z_lo * r - 90.0, (z_hi + 1) * r - 90.0
    cand = A[(A["dec"] >= dec_lo - r) & (A["dec"] <= dec_hi + r)]
    zc = O.zone(cand["dec"], r)
    a_band = cand[(zc >= z_lo) & (zc <= z_hi)]
    b_band = B[(B["dec"] >= dec_lo - 3 * r) & (B["dec"] <= dec_hi + 3 * r)]
    assert 200_000 < len(a_band) < 300_000   # 0.49% of the sphere
    ref, _ = O.match(a_band, b_band, r, threads=16)
    sel = np.isin(ab[0], a_band["id"])
    k_gpu = _pair_key(ab[0][sel], ab[1][sel])
    o = np.argsort(k_gpu, kind="stable")
    k_ref = _pair_key(ref[0], ref[1])
    assert len(k_ref) > 100_000
    assert np.array_equal(k_gpu[o], k_ref)
    assert np.abs(ab[2][sel][o] - ref[2]).max() <= DIST_TOL_ARCSEC


def _region_vs_oracle(A, B, r, ab, a_sel, b_sel, min_pairs):
    """Pairs of the A rows a_sel (boolean over A) against the oracle's join of
    those rows with the B rows b_sel (a superset of every B row within r)."""
    a_reg, b_reg = A[a_sel], B[b_sel]
    ref, _ = O.match(a_reg, b_reg, r, threads=16)
    mask = np.zeros(N, bool)
    mask[a_reg["id"]] = True   # A ids are 0..N-1
    sel = mask[ab[0]]
    k_gpu = _pair_key(ab[0][sel], ab[1][sel])
    o = np.argsort(k_gpu, kind="stable")
    k_ref = _pair_key(ref[0], ref[1])
    assert len(k_ref) >= min_pairs
    assert np.array_equal(k_gpu[o], k_ref)
    assert np.abs(ab[2][sel][o] - ref[2]).max() <= DIST_TOL_ARCSEC


def test_full_polar_caps_vs_oracle(full):
    # above |dec| = 89.5 deg the ra window of a zone is the whole circle (the pole
    # rule, Common.cu:9): every B row of the neighbour zones is a candidate
    A, B, r, ab, _, _, _ = full
    a_sel = np.abs(A["dec"]) > 89.5
    b_sel = np.abs(B["dec"]) > 89.5 - 3 * r
    assert 1_500 < a_sel.sum() < 3_000
    _region_vs_oracle(A, B, r, ab, a_sel, b_sel, 500)


def test_full_seam_strip_vs_oracle(full):
    # A rows within 0.1 deg of ra = 0 at |dec| < 85 deg: their windows cross the
    # 0/360 seam (Worker.cu:98-122); at 85 deg alpha is 23", far inside the B strip
    A, B, r, ab, _, _, _ = full
    near = lambda ra, w: (ra < w) | (ra >= 360.0 - w)  # noqa: E731
    a_sel = near(A["ra"], 0.1) & (np.abs(A["dec"]) < 85.0)
    b_sel = near(B["ra"], 0.2) & (np.abs(B["dec"]) < 85.0 + 3 * r)
    assert 20_000 < a_sel.sum() < 40_000
    crossing = np.sum(near(A["ra"][a_sel], 0.0003))
    assert crossing > 10
    _region_vs_oracle(A, B, r, ab, a_sel, b_sel, 10_000)


def _digest(ida, idb, dist):
    # order-independent: sums of 64-bit mixes of the pairs and of the distance bits
    h = (ida.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ \
        (idb.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F))
    h ^= h >> np.uint64(29)
    return (len(ida), int(np.sum(h, dtype=np.uint64)),
            int(np.sum(dist.view(np.uint64), dtype=np.uint64)))


@pytest.fixture(scope="module")
def plane(gpu_ctx):
    """The 200M x 200M galactic-plane join at 1 arcsec, records generated on the
    device: the pairs of FP64 plain, of the FP32 filter on the FP64 rows rounded in
    registers and of the FP32 filter on FP32 rows the builds store, and the records
    on the host."""
    import xmatch_amd as xm
    n, r = 200_000_000, 1.0 / 3600.0
    shp = xm.shape(xm.SHAPE_PLANE, synth.PLANE_POLE_RA, synth.PLANE_POLE_DEC, 20.0, 10.0)
    dA = gpu_ctx.synth_fresh(shp, n, 2003, stream=0, id0=0)
    dB = gpu_ctx.synth_copies(dA, 0.5, 0.3, shp, n, 2003, stream=1)
    ca = gpu_ctx.catalog(dA, r)
    cb = gpu_ctx.catalog(dB, r)
    out = {}
    for name, prec in (("fp64", xm.FP64_PLAIN), ("fp32", xm.FP32_FILTER_FP64_RECHECK)):
        res = gpu_ctx.match(ca, cb, r, precision=prec)
        out[name] = res.to_host(sort=False)
        res.close()
    ca.close(); cb.close()
    try:
        gpu_ctx.set_param("fp32_rows", 1)
        ca, cb = gpu_ctx.catalog(dA, r), gpu_ctx.catalog(dB, r)
    finally:
        gpu_ctx.set_param("fp32_rows", 0)
    res = gpu_ctx.match(ca, cb, r, precision=xm.FP32_FILTER_FP64_RECHECK)
    out["fp32rows"] = res.to_host(sort=False)
    res.close(); ca.close(); cb.close()
    A, B = gpu_ctx.download(dA), gpu_ctx.download(dB)
    dA.free(); dB.free()
    gpu_ctx.release_memory()
    return n, r, A, B, out


def test_full_plane_fp32_filter_equals_fp64(plane):
    # the FP32 filter with FP64 recheck returns the FP64 set, distances bit for bit,
    # whether it rounds the FP64 rows in registers or reads stored FP32 rows
    _, _, _, _, out = plane
    dig = {k: _digest(*v) for k, v in out.items()}
    assert 99_500_000 < dig["fp64"][0] < 101_000_000
    assert dig["fp32"] == dig["fp64"]
    assert dig["fp32rows"] == dig["fp64"]


def test_full_plane_densest_band_vs_oracle(plane):
    # the 1000 consecutive zones holding the most rows (where the plane runs along
    # the zones), joined by the oracle, against both forms' pairs of those A rows
    n, r, A, B, out = plane
    nzn = O.n_zones(r)
    za = np.minimum(np.floor((A["dec"] + 90.0) / r).astype(np.int64), nzn - 1)
    counts = np.bincount(za, minlength=nzn)
    win = np.convolve(counts, np.ones(1000, np.int64), mode="valid")
    z_lo = int(np.argmax(win))
    z_hi = z_lo + 999
    assert counts[z_lo:z_hi + 1].max() > 512      # zones of more than one 512-row chunk
    dec_lo, dec_hi = z_lo * r - 90.0, (z_hi + 1) * r - 90.0
    cand = (A["dec"] >= dec_lo - r) & (A["dec"] <= dec_hi + r)
    idx = np.nonzero(cand)[0]
    zc = O.zone(A["dec"][idx], r)
    a_band = A[idx[(zc >= z_lo) & (zc <= z_hi)]]
    b_band = B[(B["dec"] >= dec_lo - 3 * r) & (B["dec"] <= dec_hi + 3 * r)]
    zb = np.floor((b_band["dec"] + 90.0) / r).astype(np.int64)
    assert np.bincount(zb - zb.min()).max() > 256   # B zones read in several rounds
    ref, _ = O.match(a_band, b_band, r, threads=16)
    assert len(ref[0]) > 200_000
    mask = np.zeros(n, bool)
    mask[a_band["id"]] = True
    k_ref = ref[0].astype(np.int64) * (1 << 31) + (ref[1] - synth.B_ID_OFFSET)
    for name, (ia, ib, d) in out.items():
        sel = mask[ia]
        k = ia[sel].astype(np.int64) * (1 << 31) + (ib[sel] - synth.B_ID_OFFSET)
        o = np.argsort(k, kind="stable")
        assert np.array_equal(k[o], k_ref), name
        assert np.abs(d[sel][o] - ref[2]).max() <= DIST_TOL_ARCSEC, name
