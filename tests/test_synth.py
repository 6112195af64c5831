"""The device catalog generator: Philox-4x32-10 pinned by its published
known-answer vectors, and the device records checked against the same draws."""
import numpy as np
import pytest

from tests.philox_ref import philox4x32_10, uniforms

KAT = [  # (counter, key, output) from the Random123 distribution's kat_vectors
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff, 0xffffffff], [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert philox4x32_10(ctr, key) == out


@pytest.mark.gpu
def test_device_sphere_records_follow_philox(gpu_ctx):
    import xmatch_amd as xm
    seed, stream = 0x1234_5678_9abc, 5
    buf = gpu_ctx.synth_fresh(xm.shape(xm.SHAPE_SPHERE), 2000, seed, stream=stream, id0=77)
    recs = gpu_ctx.download(buf)
    buf.free()
    assert np.array_equal(recs["id"], 77 + np.arange(2000))
    for i in (0, 1, 2, 999, 1999):
        u1, u2 = uniforms(seed, stream, i, 0)
        assert abs(recs["ra"][i] - 360.0 * u1) < 1e-9
        assert abs(recs["dec"][i] - np.degrees(np.arcsin(2.0 * u2 - 1.0))) < 1e-9


@pytest.mark.gpu
def test_device_shapes_statistics(gpu_ctx):
    import xmatch_amd as xm
    n = 400_000
    sph = gpu_ctx.download(gpu_ctx.synth_fresh(xm.shape(xm.SHAPE_SPHERE), n, 3))
    # uniform on the sphere: sin(dec) uniform on [-1, 1], ra uniform on [0, 360)
    assert abs(np.mean(np.sin(np.radians(sph["dec"])))) < 0.01
    assert abs(np.mean(sph["ra"]) - 180.0) < 1.5
    assert sph["ra"].min() >= 0 and sph["ra"].max() < 360
    cap = gpu_ctx.download(gpu_ctx.synth_fresh(xm.shape(xm.SHAPE_CAP, 120.0, -30.0, 2.0), n, 4))
    c = np.array([np.cos(np.radians(-30)) * np.cos(np.radians(120)),
                  np.cos(np.radians(-30)) * np.sin(np.radians(120)), np.sin(np.radians(-30))])
    r, d = np.radians(cap["ra"]), np.radians(cap["dec"])
    ang = np.degrees(np.arccos(np.clip(np.cos(d) * np.cos(r) * c[0] + np.cos(d) * np.sin(r) * c[1]
                                       + np.sin(d) * c[2], -1, 1)))
    assert ang.max() <= 2.0 + 1e-9 and np.median(ang) > 1.3  # area-uniform: median at r/sqrt(2)
    pl = gpu_ctx.download(gpu_ctx.synth_fresh(xm.shape(xm.SHAPE_PLANE, 0.0, 30.0, 20.0, 10.0),
                                              n, 5))
    p = np.array([np.cos(np.radians(30)), 0.0, np.sin(np.radians(30))])
    r, d = np.radians(pl["ra"]), np.radians(pl["dec"])
    b = np.degrees(np.arcsin(np.cos(d) * np.cos(r) * p[0] + np.cos(d) * np.sin(r) * p[1]
                             + np.sin(d) * p[2]))
    # the density contrast concentrates sources within a few widths of the circle
    assert np.mean(np.abs(b) < 10.0) > 0.45
    copies = gpu_ctx.synth_copies(gpu_ctx.synth_fresh(xm.shape(xm.SHAPE_SPHERE), 1000, 6), 1.0,
                                  0.3, xm.shape(xm.SHAPE_SPHERE), 1000, 7)
    assert len(gpu_ctx.download(copies)) == 1000
