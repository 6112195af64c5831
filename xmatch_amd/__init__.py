"""xmatch_amd -- MI355X-native zone-algorithm crossmatch (ctypes binding of the C ABI).

The product is libxmatch_amd.so (HIP kernels for gfx950 behind include/xmatch.h).
This module is a thin ctypes layer over that ABI for tests, the bench and Python
callers.  It has no CPU fallback: if the library is missing or no GPU is usable,
calls raise.

Reference interface mirrored (paths relative to the reference repository):
  Segment + Sorter::Sort   BoostXmatch/BoostXmatch/Segment.h:22-41, Sorter.cu:141-185
                           -> Catalog
  Job + Worker::Match      Job.h:16-29, Worker.cu:243-428 -> Context.match
  JobManager cross product JobManager.cpp:5-13           -> cross_product_jobs +
                                                            Context.match_batch
"""
import ctypes
import os
import weakref

import numpy as np

from .synth import OBJ_DTYPE, ARCSEC, B_ID_OFFSET  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxmatch_amd.so")

XM_OK, XM_EINVAL, XM_ENOMEM, XM_EDEVICE, XM_EINTERNAL, XM_EAGAIN = range(6)
FP64_PLAIN = 0
FP32_FILTER_FP64_RECHECK = 1
_STATUS = {0: "XM_OK", 1: "XM_EINVAL", 2: "XM_ENOMEM", 3: "XM_EDEVICE", 4: "XM_EINTERNAL",
           5: "XM_EAGAIN"}

# every symbol declared in include/xmatch.h
EXPORTED = [
    "xm_abi_version", "xm_ctx_create", "xm_ctx_destroy", "xm_last_error_message",
    "xm_ctx_set_timing", "xm_ctx_set_timing_only", "xm_ctx_kernel_time", "xm_ctx_reset_timing", "xm_ctx_synchronize",
    "xm_ctx_set_param", "xm_ctx_get_param", "xm_ctx_release_memory",
    "xm_device_alloc", "xm_device_free", "xm_memcpy_h2d", "xm_memcpy_d2h",
    "xm_host_alloc", "xm_host_free", "xm_ctx_mem_info",
    "xm_catalog_create", "xm_catalog_create_device", "xm_catalog_destroy", "xm_catalog_info",
    "xm_catalog_export", "xm_match", "xm_result_count", "xm_result_device_ptrs",
    "xm_result_copy", "xm_result_destroy", "xm_work_counts",
    "xm_catalog_create_segments", "xm_catalog_create_segments_device",
    "xm_catalog_segment_count", "xm_catalog_segment_info", "xm_match_batch",
    "xm_result_jobs", "xm_result_job_device_ptr", "xm_work_counts_job",
    "xm_server_create", "xm_server_submit", "xm_server_poll", "xm_server_fetch",
    "xm_server_error", "xm_server_get_stats", "xm_server_destroy",
    "xm_synth_fresh", "xm_synth_copies", "xm_ctx_tune_table", "xm_match_noalloc",
]


class XmatchError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (_STATUS.get(status, status), message))
        self.status = status


class MatchStats(ctypes.Structure):
    _fields_ = [("n_matches", ctypes.c_uint64), ("n_chunks", ctypes.c_uint64),
                ("n_retries", ctypes.c_uint64), ("band_hits", ctypes.c_uint64),
                ("join_ms", ctypes.c_double), ("max_dist_err_arcsec", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Segment(ctypes.Structure):
    """xm_segment: records (host or device pointer), count, zone height."""
    _fields_ = [("records", ctypes.c_void_p), ("n", ctypes.c_uint64),
                ("zone_height_deg", ctypes.c_double)]


class Job(ctypes.Structure):
    """xm_job: segment of A x segment of B within radius_deg."""
    _fields_ = [("seg_a", ctypes.c_uint32), ("seg_b", ctypes.c_uint32),
                ("radius_deg", ctypes.c_double)]


def job_table(jobs):
    """A prepared xm_job array for Context.match_batch (at least one job)."""
    jobs = list(jobs)
    if not jobs:
        raise ValueError("no jobs")
    return _job_array(jobs)


def cross_product_jobs(n_seg_a, n_seg_b, radius_deg):
    """Every A segment against every B segment, A-major, as the reference's
    JobManager builds its queue (JobManager.cpp:5-13)."""
    return [(a, b, radius_deg) for a in range(n_seg_a) for b in range(n_seg_b)]


def _job_array(jobs):
    if isinstance(jobs, ctypes.Array):
        return jobs
    arr = (Job * max(len(jobs), 1))()
    for i, (a, b, r) in enumerate(jobs):
        arr[i].seg_a = int(a)
        arr[i].seg_b = int(b)
        arr[i].radius_deg = float(r)
    return arr


class TuneEntry(ctypes.Structure):
    """xm_tune_entry: one measured key of the context's tune table."""
    _fields_ = [("key", ctypes.c_uint32), ("chunk_rows", ctypes.c_uint32),
                ("dense_min", ctypes.c_uint32), ("sample_jobs", ctypes.c_uint32),
                ("tuned_ms", ctypes.c_double), ("default_ms", ctypes.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["log2_rows_per_zone_a"] = self.key & 0xff
        d["log2_rows_per_zone_b"] = (self.key >> 8) & 0xff
        d["nz"] = self.key >> 16
        return d


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch_rows", ctypes.c_uint64), ("max_batch_requests", ctypes.c_uint32),
                ("precision", ctypes.c_int), ("split_rows", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests_done", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("rows", ctypes.c_uint64), ("matches", ctypes.c_uint64),
                ("busy_ms", ctypes.c_double), ("prep_ms", ctypes.c_double),
                ("overlap_ms", ctypes.c_double), ("pieces", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Shape(ctypes.Structure):
    """Synthetic-catalog shape (include/xmatch_synth.h)."""
    _fields_ = [("kind", ctypes.c_int32), ("p", ctypes.c_double * 4)]


SHAPE_SPHERE, SHAPE_PATCH, SHAPE_PLANE, SHAPE_CAP = range(4)


def shape(kind, *p):
    s = Shape()
    s.kind = kind
    for i, v in enumerate(p):
        s.p[i] = v
    return s


_lib = None


def lib():
    """Load libxmatch_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("%s is not built; run python xmatch_amd/build.py "
                              "(hipcc --offload-arch=gfx950)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i32, dbl = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double
        pvp = ctypes.POINTER(ctypes.c_void_p)
        sigs = {
            "xm_abi_version": ([], ctypes.c_int),
            "xm_ctx_create": ([ctypes.c_int, pvp], ctypes.c_int),
            "xm_ctx_destroy": ([vp], None),
            "xm_last_error_message": ([vp], ctypes.c_char_p),
            "xm_ctx_set_timing": ([vp, ctypes.c_int], ctypes.c_int),
            "xm_ctx_set_timing_only": ([vp, ctypes.c_char_p], ctypes.c_int),
            "xm_ctx_kernel_time": ([vp, ctypes.c_char_p, ctypes.POINTER(dbl),
                                    ctypes.POINTER(u64)], ctypes.c_int),
            "xm_ctx_reset_timing": ([vp], ctypes.c_int),
            "xm_ctx_synchronize": ([vp], ctypes.c_int),
            "xm_ctx_release_memory": ([vp, ctypes.POINTER(u64)], ctypes.c_int),
            "xm_ctx_set_param": ([vp, ctypes.c_char_p, ctypes.c_int64], ctypes.c_int),
            "xm_ctx_get_param": ([vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)],
                                 ctypes.c_int),
            "xm_device_alloc": ([vp, ctypes.c_size_t, pvp], ctypes.c_int),
            "xm_device_free": ([vp, vp], ctypes.c_int),
            "xm_memcpy_h2d": ([vp, vp, vp, ctypes.c_size_t], ctypes.c_int),
            "xm_host_alloc": ([vp, ctypes.c_size_t, pvp], ctypes.c_int),
            "xm_ctx_mem_info": ([vp, ctypes.POINTER(u64), ctypes.POINTER(u64)], ctypes.c_int),
            "xm_host_free": ([vp, vp], ctypes.c_int),
            "xm_memcpy_d2h": ([vp, vp, vp, ctypes.c_size_t], ctypes.c_int),
            "xm_catalog_create": ([vp, vp, u64, dbl, pvp], ctypes.c_int),
            "xm_catalog_create_device": ([vp, vp, u64, dbl, pvp], ctypes.c_int),
            "xm_catalog_destroy": ([vp], None),
            "xm_catalog_info": ([vp, ctypes.POINTER(u64), ctypes.POINTER(dbl),
                                 ctypes.POINTER(i32), ctypes.POINTER(i32)], ctypes.c_int),
            "xm_catalog_export": ([vp, vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "xm_match": ([vp, vp, vp, dbl, ctypes.c_int, u64, pvp, ctypes.POINTER(MatchStats)],
                         ctypes.c_int),
            "xm_match_noalloc": ([vp, vp, vp, dbl, ctypes.c_int, vp, vp, vp, u64,
                                  ctypes.POINTER(u64), ctypes.POINTER(MatchStats)], ctypes.c_int),
            "xm_result_count": ([vp, ctypes.POINTER(u64)], ctypes.c_int),
            "xm_result_device_ptrs": ([vp, pvp, pvp, pvp], ctypes.c_int),
            "xm_result_copy": ([vp, vp, vp, vp, vp], ctypes.c_int),
            "xm_result_destroy": ([vp], None),
            "xm_work_counts": ([vp, vp, vp, dbl, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                ctypes.POINTER(u64)], ctypes.c_int),
            "xm_catalog_create_segments": ([vp, ctypes.POINTER(Segment), u64, pvp], ctypes.c_int),
            "xm_catalog_create_segments_device": ([vp, ctypes.POINTER(Segment), u64, pvp],
                                                  ctypes.c_int),
            "xm_catalog_segment_count": ([vp, ctypes.POINTER(u64)], ctypes.c_int),
            "xm_catalog_segment_info": ([vp, u64, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                         ctypes.POINTER(dbl), ctypes.POINTER(i32),
                                         ctypes.POINTER(i32)], ctypes.c_int),
            "xm_match_batch": ([vp, vp, vp, ctypes.POINTER(Job), u64, ctypes.c_int, u64, pvp,
                                ctypes.POINTER(MatchStats)], ctypes.c_int),
            "xm_result_jobs": ([vp, vp, vp, vp], ctypes.c_int),
            "xm_result_job_device_ptr": ([vp, pvp], ctypes.c_int),
            "xm_work_counts_job": ([vp, vp, vp, ctypes.POINTER(Job), ctypes.POINTER(u64),
                                    ctypes.POINTER(u64), ctypes.POINTER(u64)], ctypes.c_int),
            "xm_server_create": ([vp, vp, ctypes.POINTER(ServerConfig), pvp], ctypes.c_int),
            "xm_server_submit": ([vp, vp, u64, dbl, ctypes.POINTER(u64)], ctypes.c_int),
            "xm_server_poll": ([vp, u64, ctypes.c_int32, ctypes.POINTER(u64)], ctypes.c_int),
            "xm_server_fetch": ([vp, u64, vp, vp, vp, ctypes.POINTER(dbl)], ctypes.c_int),
            "xm_server_error": ([vp, u64], ctypes.c_char_p),
            "xm_server_get_stats": ([vp, ctypes.POINTER(ServerStats)], ctypes.c_int),
            "xm_server_destroy": ([vp], None),
            "xm_ctx_tune_table": ([vp, ctypes.POINTER(TuneEntry), u64, ctypes.POINTER(u64)],
                                  ctypes.c_int),
            "xm_synth_fresh": ([vp, ctypes.POINTER(Shape), u64, u64, u64, ctypes.c_int64, vp],
                               ctypes.c_int),
            "xm_synth_copies": ([vp, vp, u64, dbl, dbl, ctypes.POINTER(Shape), u64, u64, u64,
                                 ctypes.c_int64, vp], ctypes.c_int),
        }
        for name, (args, res) in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Context:
    """One GPU: replaces the reference's CudaManager/CudaContext pair."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        self._children = weakref.WeakSet()
        st = lib().xm_ctx_create(device, ctypes.byref(self._h))
        if st != XM_OK:
            raise XmatchError(st, lib().xm_last_error_message(None).decode())

    def _check(self, st):
        if st != XM_OK:
            raise XmatchError(st, lib().xm_last_error_message(self._h).decode())

    @property
    def handle(self):
        return self._h

    def close(self):
        """Destroy the context after every catalog and result made from it."""
        if self._h:
            # servers first: their workers use the catalogs
            for c in sorted(list(self._children), key=lambda c: not isinstance(c, Server)):
                c.close()
            lib().xm_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_timing(self, on=True, only=None):
        """Per-launch event timing; `only` names the one kernel to time (None: all)."""
        self._check(lib().xm_ctx_set_timing_only(self._h, only.encode() if only else None))
        self._check(lib().xm_ctx_set_timing(self._h, 1 if on else 0))

    def kernel_time(self, name):
        t = ctypes.c_double()
        c = ctypes.c_uint64()
        self._check(lib().xm_ctx_kernel_time(self._h, name.encode(), ctypes.byref(t), ctypes.byref(c)))
        return t.value, c.value

    def reset_timing(self):
        self._check(lib().xm_ctx_reset_timing(self._h))

    def synchronize(self):
        self._check(lib().xm_ctx_synchronize(self._h))

    def release_memory(self):
        """Return cached device blocks; returns the bytes the context still holds."""
        b = ctypes.c_uint64()
        self._check(lib().xm_ctx_release_memory(self._h, ctypes.byref(b)))
        return b.value

    def mem_info(self):
        """(free, total) device bytes; blocks the context caches count as free."""
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(lib().xm_ctx_mem_info(self._h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def set_param(self, name, value):
        self._check(lib().xm_ctx_set_param(self._h, name.encode(), int(value)))

    def get_param(self, name):
        v = ctypes.c_int64()
        self._check(lib().xm_ctx_get_param(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def tune_table(self):
        """The measured pair-kernel shapes (param "autotune"), one dict per key."""
        n = ctypes.c_uint64()
        self._check(lib().xm_ctx_tune_table(self._h, None, 0, ctypes.byref(n)))
        arr = (TuneEntry * max(n.value, 1))()
        self._check(lib().xm_ctx_tune_table(self._h, arr, n.value, ctypes.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    def to_device(self, arr):
        """Copy a host array to a device buffer owned by the context pool."""
        arr = np.ascontiguousarray(arr)
        d = ctypes.c_void_p()
        self._check(lib().xm_device_alloc(self._h, max(arr.nbytes, 1), ctypes.byref(d)))
        self._check(lib().xm_memcpy_h2d(self._h, d, _ptr(arr), arr.nbytes))
        return DeviceBuffer(self, d, arr.nbytes, arr.dtype, arr.shape)

    def alloc_records(self, n):
        """Uninitialised device buffer of n 24-byte records."""
        d = ctypes.c_void_p()
        nbytes = int(n) * OBJ_DTYPE.itemsize
        self._check(lib().xm_device_alloc(self._h, max(nbytes, 1), ctypes.byref(d)))
        return DeviceBuffer(self, d, nbytes, OBJ_DTYPE, (int(n),))

    def synth_fresh(self, shp, n, seed, stream=0, id0=0, out=None):
        """n device records of shape shp (Philox-4x32-10 draws keyed by seed), into
        out (a record DeviceBuffer or view of n rows) or a new buffer."""
        buf = self.alloc_records(n) if out is None else out
        self._check(lib().xm_synth_fresh(self._h, ctypes.byref(shp), int(n), seed, stream,
                                         int(id0), buf.ptr))
        return buf

    def synth_copies(self, a, frac_copy, sigma_arcsec, fresh, n, seed, stream=1,
                     id0=B_ID_OFFSET, out=None):
        """n device records: perturbed copies of records of DeviceBuffer a, then fresh."""
        buf = self.alloc_records(n) if out is None else out
        na = a.nbytes // OBJ_DTYPE.itemsize
        self._check(lib().xm_synth_copies(self._h, a.ptr, na, frac_copy, sigma_arcsec,
                                          ctypes.byref(fresh), int(n), seed, stream, int(id0),
                                          buf.ptr))
        return buf

    def download(self, buf):
        """Copy a DeviceBuffer back to a host numpy array."""
        out = np.empty(buf.shape, dtype=buf.dtype)
        self._check(lib().xm_memcpy_d2h(self._h, _ptr(out), buf.ptr, buf.nbytes))
        return out

    def catalog(self, records, zone_height_deg):
        """Zone-sorted device catalog from host records (dtype OBJ_DTYPE) or a DeviceBuffer."""
        return Catalog(self, records, zone_height_deg)

    def catalog_segments(self, segments, zone_height_deg=None):
        """One catalog of many segments, prepared by one sort.  segments: host record
        arrays (dtype OBJ_DTYPE) or DeviceBuffers (not mixed), or a SegmentTable;
        zone_height_deg: one value or one per segment."""
        if isinstance(segments, SegmentTable):
            return Catalog(self, segments, None)
        return Catalog(self, list(segments), zone_height_deg)

    def match(self, a, b, radius_deg, precision=FP64_PLAIN, capacity_hint=0):
        r = ctypes.c_void_p()
        stats = MatchStats()
        self._check(lib().xm_match(self._h, a.handle, b.handle, radius_deg, precision,
                                   capacity_hint, ctypes.byref(r), ctypes.byref(stats)))
        return Result(self, r, stats)

    def match_noalloc(self, a, b, radius_deg, id_a, id_b, dist, capacity,
                      precision=FP64_PLAIN):
        """xm_match_noalloc: the join's pairs into the caller's device buffers id_a,
        id_b (int64) and dist (float64, arcsec) of `capacity` elements each; returns
        (pair count, MatchStats).  Raises XmatchError (XM_ENOMEM) when the pairs
        outnumber capacity; the count is then in the exception's `n_matches`."""
        n = ctypes.c_uint64()
        stats = MatchStats()
        st = lib().xm_match_noalloc(self._h, a.handle, b.handle, radius_deg, precision,
                                    id_a.ptr, id_b.ptr, dist.ptr, int(capacity),
                                    ctypes.byref(n), ctypes.byref(stats))
        if st != XM_OK:
            err = XmatchError(st, lib().xm_last_error_message(self._h).decode())
            err.n_matches = n.value
            raise err
        return n.value, stats

    def match_batch(self, a, b, jobs, precision=FP64_PLAIN, capacity_hint=0):
        """One launch over jobs [(seg_a, seg_b, radius_deg), ...]; the result carries
        the job of every pair (Result.jobs)."""
        arr = _job_array(jobs)
        n_jobs = getattr(jobs, "n_jobs", len(jobs))
        r = ctypes.c_void_p()
        stats = MatchStats()
        self._check(lib().xm_match_batch(self._h, a.handle, b.handle, arr, n_jobs, precision,
                                         capacity_hint, ctypes.byref(r), ctypes.byref(stats)))
        return Result(self, r, stats, n_jobs=n_jobs)

    def work_counts(self, a, b, radius_deg):
        w = [ctypes.c_uint64() for _ in range(3)]
        self._check(lib().xm_work_counts(self._h, a.handle, b.handle, radius_deg,
                                         *(ctypes.byref(x) for x in w)))
        return {"w_ref": w[0].value, "w_cand": w[1].value, "w_tile": w[2].value}

    def work_counts_job(self, a, b, job):
        arr = _job_array([job])
        w = [ctypes.c_uint64() for _ in range(3)]
        self._check(lib().xm_work_counts_job(self._h, a.handle, b.handle, arr,
                                             *(ctypes.byref(x) for x in w)))
        return {"w_ref": w[0].value, "w_cand": w[1].value, "w_tile": w[2].value}


class DeviceBuffer:
    """Device memory of a context.  An owning buffer returns its memory to the
    context's cache on free() or when it is garbage-collected while the context is
    open; a view keeps its parent alive."""

    def __init__(self, ctx, ptr, nbytes, dtype, shape, owner=True, parent=None):
        self.ctx, self.ptr, self.nbytes, self.dtype, self.shape = ctx, ptr, nbytes, dtype, shape
        self.owner = owner
        self._parent = parent

    def view(self, first, count):
        """Rows [first, first + count) of a 1-D buffer, without ownership."""
        item = np.dtype(self.dtype).itemsize
        if first < 0 or count < 0 or (first + count) * item > self.nbytes:
            raise IndexError("view outside the buffer")
        return DeviceBuffer(self.ctx, ctypes.c_void_p((self.ptr.value or 0) + first * item),
                            count * item, self.dtype, (int(count),), owner=False, parent=self)

    def free(self):
        if self.ptr and self.owner and self.ctx.handle:
            self.ctx._check(lib().xm_device_free(self.ctx.handle, self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SegmentTable:
    """A prepared xm_segment array (e.g. built once, outside a timed loop)."""

    def __init__(self, segments, zone_height_deg):
        self.keep = []
        n = len(segments)
        hs = (list(zone_height_deg) if np.ndim(zone_height_deg) else
              [float(zone_height_deg)] * n)
        self.segs = (Segment * max(n, 1))()
        self.n = n
        self.on_device = n > 0 and isinstance(segments[0], DeviceBuffer)
        for i, r in enumerate(segments):
            if isinstance(r, DeviceBuffer) != self.on_device:
                raise TypeError("segments must be all host arrays or all DeviceBuffers")
            if self.on_device:
                self.keep.append(r)   # the table points into it
                self.segs[i].records = r.ptr
                self.segs[i].n = r.nbytes // OBJ_DTYPE.itemsize
            else:
                rec = np.ascontiguousarray(r)
                if rec.dtype != OBJ_DTYPE:
                    raise TypeError("records must have dtype %r" % (OBJ_DTYPE,))
                self.keep.append(rec)
                self.segs[i].records = rec.ctypes.data if rec.size else None
                self.segs[i].n = rec.size
            self.segs[i].zone_height_deg = hs[i]


class Catalog:
    def __init__(self, ctx, records, zone_height_deg):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        if isinstance(records, SegmentTable):
            fn = (lib().xm_catalog_create_segments_device if records.on_device
                  else lib().xm_catalog_create_segments)
            st = fn(ctx.handle, records.segs, records.n, ctypes.byref(self._h))
        elif isinstance(records, list):
            t = SegmentTable(records, zone_height_deg)
            fn = (lib().xm_catalog_create_segments_device if t.on_device
                  else lib().xm_catalog_create_segments)
            st = fn(ctx.handle, t.segs, t.n, ctypes.byref(self._h))
        elif isinstance(records, DeviceBuffer):
            n = records.nbytes // OBJ_DTYPE.itemsize
            st = lib().xm_catalog_create_device(ctx.handle, records.ptr, n, zone_height_deg,
                                                ctypes.byref(self._h))
        else:
            rec = np.ascontiguousarray(records)
            if rec.dtype != OBJ_DTYPE:
                raise TypeError("records must have dtype %r" % (OBJ_DTYPE,))
            st = lib().xm_catalog_create(ctx.handle, _ptr(rec), rec.size, zone_height_deg,
                                         ctypes.byref(self._h))
        ctx._check(st)
        ctx._children.add(self)

    @property
    def handle(self):
        return self._h

    def info(self):
        n = ctypes.c_uint64()
        h = ctypes.c_double()
        z0 = ctypes.c_int32()
        z1 = ctypes.c_int32()
        self.ctx._check(lib().xm_catalog_info(self._h, ctypes.byref(n), ctypes.byref(h),
                                              ctypes.byref(z0), ctypes.byref(z1)))
        return {"n": n.value, "zone_height_deg": h.value, "zmin": z0.value, "zmax": z1.value}

    @property
    def segment_count(self):
        n = ctypes.c_uint64()
        self.ctx._check(lib().xm_catalog_segment_count(self._h, ctypes.byref(n)))
        return n.value

    def segment_info(self, s):
        f, n, h = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        z0, z1 = ctypes.c_int32(), ctypes.c_int32()
        self.ctx._check(lib().xm_catalog_segment_info(self._h, s, ctypes.byref(f), ctypes.byref(n),
                                                      ctypes.byref(h), ctypes.byref(z0),
                                                      ctypes.byref(z1)))
        return {"first_row": f.value, "n": n.value, "zone_height_deg": h.value,
                "zmin": z0.value, "zmax": z1.value}

    def export(self, fields=("id", "x", "y", "z", "zone")):
        """Sorted-order host arrays of the catalog: ids, FP64 unit vectors, zones
        (only the fields named)."""
        n = self.info()["n"]
        types = {"id": np.int64, "x": np.float64, "y": np.float64, "z": np.float64,
                 "zone": np.int32}
        out = {f: np.empty(n, types[f]) for f in fields}
        self.ctx._check(lib().xm_catalog_export(self.ctx.handle, self._h, _ptr(out.get("id")),
                                                _ptr(out.get("x")), _ptr(out.get("y")),
                                                _ptr(out.get("z")), _ptr(out.get("zone"))))
        return out

    def close(self):
        if self._h:
            lib().xm_catalog_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Result:
    def __init__(self, ctx, h, stats, n_jobs=None):
        self.ctx = ctx
        self._h = h
        self.stats = stats
        self.n_jobs = n_jobs
        ctx._children.add(self)

    def jobs(self, column=True):
        """Batch results: (job index of each pair, in to_host(sort=False) order, or
        None without column; pairs per job)."""
        job = np.empty(self.count, np.uint32) if column else None
        counts = np.empty(self.n_jobs or 0, np.uint64)
        self.ctx._check(lib().xm_result_jobs(self.ctx.handle, self._h, _ptr(job), _ptr(counts)))
        return job, counts

    def by_job(self):
        """Batch results split per job: a list of (id_a, id_b, dist) sorted by (id_a, id_b)."""
        ida, idb, dist = self.to_host(sort=False)
        job, counts = self.jobs()
        o = np.lexsort((idb, ida, job))
        ida, idb, dist, job = ida[o], idb[o], dist[o], job[o]
        edges = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
        return [(ida[edges[j]:edges[j + 1]], idb[edges[j]:edges[j + 1]],
                 dist[edges[j]:edges[j + 1]]) for j in range(len(counts))]

    @property
    def count(self):
        n = ctypes.c_uint64()
        self.ctx._check(lib().xm_result_count(self._h, ctypes.byref(n)))
        return n.value

    def to_host(self, sort=True):
        n = self.count
        ida = np.empty(n, np.int64)
        idb = np.empty(n, np.int64)
        dist = np.empty(n, np.float64)
        self.ctx._check(lib().xm_result_copy(self.ctx.handle, self._h, _ptr(ida), _ptr(idb),
                                             _ptr(dist)))
        if sort:
            o = np.lexsort((idb, ida))
            ida, idb, dist = ida[o], idb[o], dist[o]
        return ida, idb, dist

    def close(self):
        if self._h:
            lib().xm_result_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """A resident catalog answering crossmatch requests with continuous batching
    (xm_server_*).  submit() returns a ticket; result(ticket) waits and returns
    (id_query, id_catalog, dist_arcsec, latency_ms)."""

    def __init__(self, ctx, resident, max_batch_rows=0, max_batch_requests=0,
                 precision=FP64_PLAIN, split_rows=0):
        self.ctx = ctx
        self.resident = resident  # kept alive for the server's lifetime
        cfg = ServerConfig(max_batch_rows, max_batch_requests, precision, split_rows)
        self._h = ctypes.c_void_p()
        ctx._check(lib().xm_server_create(ctx.handle, resident.handle, ctypes.byref(cfg),
                                          ctypes.byref(self._h)))
        ctx._children.add(self)

    def submit(self, records, radius_deg):
        rec = np.ascontiguousarray(records)
        if rec.dtype != OBJ_DTYPE:
            raise TypeError("records must have dtype %r" % (OBJ_DTYPE,))
        t = ctypes.c_uint64()
        st = lib().xm_server_submit(self._h, _ptr(rec) if rec.size else None, rec.size,
                                    radius_deg, ctypes.byref(t))
        if st != XM_OK:
            raise XmatchError(st, "request rejected (bad radius or records)")
        return t.value

    def poll(self, ticket, timeout_ms=0):
        """Pair count when done, None while pending."""
        n = ctypes.c_uint64()
        st = lib().xm_server_poll(self._h, ticket, int(timeout_ms), ctypes.byref(n))
        if st == XM_EAGAIN:
            return None
        if st != XM_OK:
            msg = lib().xm_server_error(self._h, ticket).decode()
            # a failed request is released by fetching it
            lib().xm_server_fetch(self._h, ticket, None, None, None, None)
            raise XmatchError(st, msg)
        return n.value

    def result(self, ticket, timeout_ms=-1):
        n = self.poll(ticket, timeout_ms)
        if n is None:
            return None
        ida = np.empty(n, np.int64)
        idb = np.empty(n, np.int64)
        dist = np.empty(n, np.float64)
        lat = ctypes.c_double()
        st = lib().xm_server_fetch(self._h, ticket, _ptr(ida), _ptr(idb), _ptr(dist),
                                   ctypes.byref(lat))
        if st != XM_OK:
            raise XmatchError(st, "fetch failed")
        return ida, idb, dist, lat.value

    def stats(self):
        s = ServerStats()
        self.ctx._check(lib().xm_server_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def close(self):
        if self._h:
            lib().xm_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

