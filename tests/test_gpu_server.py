"""GPU tests of the serving loop (xm_server_*): every request's pairs equal the
oracle's join of that request with the resident catalog, whatever it was batched
with (alone = in a mix)."""
import threading
import time

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


def test_large_request_does_not_hold_up_small_ones(gpu_ctx, resident):
    # a 1e6-row request cut into 8192-row pieces, batches of at most 32768 rows: a
    # small request submitted just after it shares the next batches and finishes
    # long before it, where a stop-and-wait server would finish the large request
    # first; both equal the oracle
    A, cat = resident
    rng = np.random.default_rng(7)
    big = make_query(rng, A, 1_000_000, 0.3, id0=2 * 10**10)
    small = make_query(rng, A, 50, 0.02, id0=3 * 10**10)
    r = 2.0 / 3600
    with xm.Server(gpu_ctx, cat, max_batch_rows=32768, split_rows=8192) as srv:
        t0 = time.perf_counter()
        tb = srv.submit(big, r)
        ts = srv.submit(small, r)
        t1 = time.perf_counter()
        rs = srv.result(ts)
        rb = srv.result(tb)
        st = srv.stats()
    # the small one's completion (at most t1 + its latency) precedes the large
    # one's (at least t0 + its latency) by a wide margin
    assert t1 + rs[3] / 1e3 < t0 + rb[3] / 1e3
    assert st["pieces"] >= 1_000_000 // 8192 + 1 and st["batches"] >= 30
    check(rs, small, A, r)
    check(rb, big, A, r)


def test_alone_equals_pieces_in_mix(gpu_ctx, resident):
    # a request cut into pieces and mixed with others gives the pairs it gets alone
    A, cat = resident
    rng = np.random.default_rng(8)
    q = make_query(rng, A, 30_000, 0.1, id0=4 * 10**10)
    r = 3.0 / 3600
    with xm.Server(gpu_ctx, cat) as srv:
        alone = srv.result(srv.submit(q, r))
    with xm.Server(gpu_ctx, cat, max_batch_rows=20_000, split_rows=4096) as srv:
        others = [srv.submit(make_query(rng, A, int(rng.integers(1, 5000)), 0.05,
                                        id0=5 * 10**10 + 10**5 * i), 2.0 / 3600)
                  for i in range(20)]
        t = srv.submit(q, r)
        mixed = srv.result(t)
        for tk in others:
            srv.result(tk)
        st = srv.stats()
    assert st["pieces"] > 20 + 30_000 // 4096
    oa = np.lexsort((alone[1], alone[0]))
    om = np.lexsort((mixed[1], mixed[0]))
    for x, y in zip(alone[:3], mixed[:3]):
        assert np.array_equal(x[oa], y[om])
    check(mixed, q, A, r)


def test_preparation_overlaps_joins(gpu_ctx, resident):
    # a stream of requests keeps both threads busy: the next batch is prepared
    # while the current one joins
    A, cat = resident
    rng = np.random.default_rng(9)
    qs = [make_query(rng, A, int(rng.integers(2000, 20000)), 0.1, id0=6 * 10**10 + 10**5 * i)
          for i in range(60)]
    with xm.Server(gpu_ctx, cat, max_batch_rows=40_000) as srv:
        tickets = [srv.submit(q, 2.0 / 3600) for q in qs]
        for t in tickets:
            srv.result(t)
        st = srv.stats()
    assert st["batches"] >= 10
    assert st["prep_ms"] > 0 and st["busy_ms"] > 0 and st["overlap_ms"] > 0
