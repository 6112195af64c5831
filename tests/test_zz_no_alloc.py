"""xm_match_noalloc, the join that allocates no device memory, against xm_match
(the allocating entry point it is a variant of): the same pairs and distances, on
two calls in a row with the second one smaller, in both precision forms.  The
scratch array keeps the last call's contents, so the second call also shows that
each call starts from cleared counters."""
import numpy as np
import pytest

import xmatch_amd as xm
from xmatch_amd import synth

pytestmark = pytest.mark.gpu

SENTINEL = -7


def out_buffers(ctx, cap):
    fill = np.full(cap, SENTINEL, np.int64)
    return (ctx.to_device(fill), ctx.to_device(fill),
            ctx.to_device(np.full(cap, -1.0)))


def host(ctx, bufs, n):
    ia, ib, d = (ctx.download(b)[:n] for b in bufs)
    o = np.lexsort((ib, ia))
    return ia[o], ib[o], d[o]


def reference(ctx, ca, cb, r, precision):
    res = ctx.match(ca, cb, r, precision=precision)
    out = res.to_host()
    res.close()
    return out


@pytest.mark.parametrize("precision,fp32_rows",
                         [(xm.FP64_PLAIN, 0), (xm.FP32_FILTER_FP64_RECHECK, 0),
                          (xm.FP32_FILTER_FP64_RECHECK, 1)],
                         ids=["fp64", "fp32_filter", "fp32_filter_stored_rows"])
def test_noalloc_equals_match_on_two_calls(gpu_ctx, precision, fp32_rows):
    ctx = gpu_ctx
    ctx.set_param("fp32_rows", fp32_rows)
    try:
        big = synth.config_small_patch(seed=61, n_a=300_000, n_b=300_000)
        small = synth.config_small_patch(seed=62, n_a=40_000, n_b=50_000)
        cats = [(ctx.catalog(a, r), ctx.catalog(b, r), r) for a, b, r in (big, small)]
    finally:
        ctx.set_param("fp32_rows", 0)
    cap = 400_000
    bufs = out_buffers(ctx, cap)
    counts = []
    for ca, cb, r in cats:   # the second call joins the smaller pair
        ref = reference(ctx, ca, cb, r, precision)
        n, st = ctx.match_noalloc(ca, cb, r, *bufs, cap, precision=precision)
        counts.append(n)
        assert n == len(ref[0]) == st.n_matches
        got = host(ctx, bufs, n)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        assert np.abs(got[2] - ref[2]).max() <= 1e-6
        assert np.array_equal(got[2], ref[2])   # the same FP64 chord, the same arcsec
        # nothing written past the pairs
        tail = ctx.download(bufs[0])[n:]
        assert np.all(tail == SENTINEL) or counts[0] > n
    assert counts[0] > 200_000 and 0 < counts[1] < counts[0]
    for ca, cb, _ in cats:
        ca.close(); cb.close()


def test_noalloc_rejects_what_it_cannot_hold(gpu_ctx):
    ctx = gpu_ctx
    a, b, r = synth.config_small_patch(seed=63, n_a=20_000, n_b=20_000)
    ca, cb = ctx.catalog(a, r), ctx.catalog(b, r)
    ref = reference(ctx, ca, cb, r, xm.FP64_PLAIN)
    small = out_buffers(ctx, 100)
    with pytest.raises(xm.XmatchError) as e:
        ctx.match_noalloc(ca, cb, r, *small, 100)
    assert e.value.status == xm.XM_ENOMEM and e.value.n_matches == len(ref[0])
    assert np.all(ctx.download(small[0]) == SENTINEL)   # nothing written
    # the FP32 form (FP64 rows rounded in registers) counts the same pairs
    with pytest.raises(xm.XmatchError) as e:
        ctx.match_noalloc(ca, cb, r, *small, 100, precision=xm.FP32_FILTER_FP64_RECHECK)
    assert e.value.status == xm.XM_ENOMEM and e.value.n_matches == len(ref[0])
    assert np.all(ctx.download(small[0]) == SENTINEL)
    # a join after the rejections still equals xm_match
    bufs = out_buffers(ctx, 50_000)
    n, _ = ctx.match_noalloc(ca, cb, r, *bufs, 50_000)
    got = host(ctx, bufs, n)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    ca.close(); cb.close()
