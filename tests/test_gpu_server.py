"""GPU tests of the serving loop (xm_server_*): every request's pairs equal the
oracle's join of that request with the resident catalog, whatever it was batched
with (alone = in a mix)."""
import threading

import numpy as np
import pytest

import xmatch_amd as xm
from oracle import oracle as O
from xmatch_amd import synth

pytestmark = pytest.mark.gpu

H = 1.0 / 3600


@pytest.fixture(scope="module")
def resident(gpu_ctx):
    A, _, _ = synth.config_small_patch(seed=21, n_a=200_000, n_b=10)
    cat = gpu_ctx.catalog(A, H)
    yield A, cat
    cat.close()


def make_query(rng, A, n, r_cap_deg, near=0.7, id0=0):
    """n sources in a cap around a resident source: near copies (0.5") + fresh."""
    c = A[rng.integers(0, len(A))]
    ra, dec = synth.uniform_cap(rng, n, c["ra"], c["dec"], r_cap_deg)
    k = int(round(near * n))
    if k:
        # copies of resident sources nearest to the cap: perturb resident rows
        pick = rng.integers(0, len(A), k)
        d = np.abs(A["ra"] - c["ra"]) + np.abs(A["dec"] - c["dec"])
        close = np.argsort(d)[:max(k, 1)]
        pick = close[rng.integers(0, len(close), k)]
        ra[:k], dec[:k] = synth.perturb(rng, A["ra"][pick], A["dec"][pick], 0.5)
    return synth.make_objs(id0 + np.arange(n), ra, dec)


def check(res, q, A, r):
    ia, ib, d, lat = res
    o = np.lexsort((ib, ia))
    ref = O.match(q, A, r, zone_height_deg=H)[0]
    assert np.array_equal(ia[o], ref[0]) and np.array_equal(ib[o], ref[1])
    if len(d):
        assert np.abs(d[o] - ref[2]).max() <= 1e-6
    assert lat >= 0.0


def test_requests_match_oracle(gpu_ctx, resident):
    A, cat = resident
    rng = np.random.default_rng(1)
    with xm.Server(gpu_ctx, cat) as srv:
        reqs = []
        for i in range(40):
            n = int(rng.integers(1, 3000))
            r = float(rng.uniform(1, 5)) / 3600
            q = make_query(rng, A, n, 0.05, id0=10**9 + 10**5 * i)
            reqs.append((srv.submit(q, r), q, r))
        for t, q, r in reqs:
            check(srv.result(t), q, A, r)
        st = srv.stats()
        assert st["requests_done"] == 40 and st["batches"] >= 1


def test_alone_equals_in_mix(gpu_ctx, resident):
    A, cat = resident
    rng = np.random.default_rng(2)
    q = make_query(rng, A, 2000, 0.03, id0=5 * 10**9)
    r = 3.0 / 3600
    with xm.Server(gpu_ctx, cat) as srv:
        alone = srv.result(srv.submit(q, r))
        others = [make_query(rng, A, int(rng.integers(1, 2000)), 0.05, id0=6 * 10**9 + 10**5 * i)
                  for i in range(60)]
        tickets = [srv.submit(o, 2.0 / 3600) for o in others[:30]]
        t = srv.submit(q, r)
        tickets += [srv.submit(o, 4.0 / 3600) for o in others[30:]]
        mixed = srv.result(t)
        for tk in tickets:
            srv.result(tk)
        assert srv.stats()["batches"] >= 2
    oa = np.lexsort((alone[1], alone[0]))
    om = np.lexsort((mixed[1], mixed[0]))
    assert np.array_equal(alone[0][oa], mixed[0][om])
    assert np.array_equal(alone[1][oa], mixed[1][om])
    assert np.array_equal(alone[2][oa], mixed[2][om])
    check(mixed, q, A, r)


def test_concurrent_clients(gpu_ctx, resident):
    A, cat = resident
    out = {}
    errors = []

    def client(k):
        try:
            rng = np.random.default_rng(100 + k)
            for i in range(10):
                q = make_query(rng, A, int(rng.integers(1, 500)), 0.02,
                               id0=7 * 10**9 + 10**6 * k + 10**4 * i)
                r = float(rng.uniform(1, 5)) / 3600
                t = srv.submit(q, r)
                out[(k, i)] = (srv.result(t), q, r)
        except Exception as e:  # reported below
            errors.append(e)

    with xm.Server(gpu_ctx, cat) as srv:
        th = [threading.Thread(target=client, args=(k,)) for k in range(8)]
        for x in th:
            x.start()
        for x in th:
            x.join()
    assert not errors
    assert len(out) == 80
    for res, q, r in out.values():
        check(res, q, A, r)


def test_empty_request_and_errors(gpu_ctx, resident):
    A, cat = resident
    with xm.Server(gpu_ctx, cat) as srv:
        e = synth.make_objs([], [], [])
        ia, ib, d, _ = srv.result(srv.submit(e, 2.0 / 3600))
        assert len(ia) == 0
        bad = synth.make_objs([1], [400.0], [0.0])
        with pytest.raises(xm.XmatchError):
            srv.submit(bad, 1.0 / 3600)
        with pytest.raises(xm.XmatchError):
            srv.submit(A[:5], -1.0)
        with pytest.raises(xm.XmatchError):
            srv.poll(123456789)
        # far from the resident patch: no pairs
        far = synth.make_objs([1, 2], [200.0, 200.1], [60.0, 60.0])
        assert len(srv.result(srv.submit(far, 5.0 / 3600))[0]) == 0


def test_destroy_with_queued_requests(gpu_ctx, resident):
    A, cat = resident
    rng = np.random.default_rng(3)
    srv = xm.Server(gpu_ctx, cat)
    for i in range(20):
        srv.submit(make_query(rng, A, 1000, 0.05, id0=10**10 + 10**4 * i), 2.0 / 3600)
    srv.close()  # drains the queue, then stops
