"""ctypes wrapper of the CPU oracle (oracle/zone_oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker.  The product library never
routes through it.  See xmo.h for the reference file:line map.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libxmo.so")

OBJ_DTYPE = np.dtype([("id", "<i8"), ("ra", "<f8"), ("dec", "<f8")])  # Obj.h:21-25


class _Pairs(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("cap", ctypes.c_int64),
                ("ida", ctypes.POINTER(ctypes.c_int64)), ("idb", ctypes.POINTER(ctypes.c_int64)),
                ("dist", ctypes.POINTER(ctypes.c_double))]


class _Counts(ctypes.Structure):
    _fields_ = [("w_ref", ctypes.c_uint64), ("w_cand", ctypes.c_uint64),
                ("n_match", ctypes.c_uint64)]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.xmo_sincos_many.argtypes = [dp, ctypes.c_int64, dp, dp]
        L.xmo_unitvec_many.argtypes = [dp, dp, ctypes.c_int64, dp, dp, dp]
        L.xmo_zone.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int32]
        L.xmo_zone.restype = ctypes.c_int32
        L.xmo_n_zones.argtypes = [ctypes.c_double]
        L.xmo_n_zones.restype = ctypes.c_int32
        L.xmo_alpha.argtypes = [ctypes.c_double, ctypes.c_double]
        L.xmo_alpha.restype = ctypes.c_double
        L.xmo_alpha_zone.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int32]
        L.xmo_alpha_zone.restype = ctypes.c_double
        L.xmo_sr_dist2.argtypes = [ctypes.c_double]
        L.xmo_sr_dist2.restype = ctypes.c_double
        L.xmo_match_objs.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.POINTER(_Pairs), ctypes.POINTER(_Counts)]
        L.xmo_brute.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                ctypes.c_double, ctypes.c_double, ctypes.POINTER(_Pairs)]
        L.xmo_pairs_free.argtypes = [ctypes.POINTER(_Pairs)]
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def sincos(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().xmo_sincos_many(_dptr(x), x.size, _dptr(s), _dptr(c))
    return s, c


def unitvec(ra_deg, dec_deg):
    ra = np.ascontiguousarray(ra_deg, dtype=np.float64)
    dec = np.ascontiguousarray(dec_deg, dtype=np.float64)
    x, y, z = (np.empty_like(ra) for _ in range(3))
    lib().xmo_unitvec_many(_dptr(ra), _dptr(dec), ra.size, _dptr(x), _dptr(y), _dptr(z))
    return x, y, z


def zone(dec_deg, h_deg):
    L = lib()
    nzn = L.xmo_n_zones(h_deg)
    return np.array([L.xmo_zone(float(d), h_deg, nzn) for d in np.atleast_1d(dec_deg)],
                    dtype=np.int32)


def n_zones(h_deg):
    return lib().xmo_n_zones(h_deg)


def alpha(theta_deg, abs_dec_deg):
    return lib().xmo_alpha(theta_deg, abs_dec_deg)


def alpha_zone(theta_deg, h_deg, z):
    return lib().xmo_alpha_zone(theta_deg, h_deg, z)


def sr_dist2(sr_deg):
    return lib().xmo_sr_dist2(sr_deg)


def _take(p):
    n = p.n
    if n:
        ida = np.ctypeslib.as_array(p.ida, shape=(n,)).copy()
        idb = np.ctypeslib.as_array(p.idb, shape=(n,)).copy()
        dist = np.ctypeslib.as_array(p.dist, shape=(n,)).copy()
    else:
        ida = np.empty(0, np.int64)
        idb = np.empty(0, np.int64)
        dist = np.empty(0, np.float64)
    lib().xmo_pairs_free(ctypes.byref(p))
    return ida, idb, dist


def _objs(a):
    a = np.ascontiguousarray(a)
    if a.dtype != OBJ_DTYPE:
        raise TypeError("records must have dtype %r" % (OBJ_DTYPE,))
    return a


def match(a, b, radius_deg, zone_height_deg=None, mode=1, threads=None, zone_range=None):
    """Zone join of record arrays a and b (Worker::Match semantics).

    Returns (idA, idB, dist_arcsec) sorted by (idA, idB) and a dict of work counts.
    mode 0 is the reference's dense zone-pair loop (W_ref tests), mode 1 the
    per-source ra window; both give the same set.
    """
    a = _objs(a)
    b = _objs(b)
    h = radius_deg if zone_height_deg is None else zone_height_deg
    if threads is None:
        threads = os.cpu_count() or 1
    z_lo, z_hi = (-(2 ** 31), 2 ** 31 - 1) if zone_range is None else zone_range
    p = _Pairs()
    c = _Counts()
    err = lib().xmo_match_objs(a.ctypes.data, a.size, b.ctypes.data, b.size, radius_deg, h,
                               mode, threads, z_lo, z_hi, ctypes.byref(p), ctypes.byref(c))
    if err:
        lib().xmo_pairs_free(ctypes.byref(p))
        raise RuntimeError("oracle match failed: %d" % err)
    res = _take(p)
    return res, {"w_ref": c.w_ref, "w_cand": c.w_cand, "n_match": c.n_match}


def brute(a, b, radius_deg, zone_height_deg=None):
    """All-pairs evaluation of the predicate; O(len(a)*len(b))."""
    a = _objs(a)
    b = _objs(b)
    h = radius_deg if zone_height_deg is None else zone_height_deg
    p = _Pairs()
    err = lib().xmo_brute(a.ctypes.data, a.size, b.ctypes.data, b.size, radius_deg, h,
                          ctypes.byref(p))
    if err:
        lib().xmo_pairs_free(ctypes.byref(p))
        raise RuntimeError("oracle brute failed: %d" % err)
    return _take(p)
