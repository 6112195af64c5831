"""Benchmark of the zone crossmatch on one MI355X.

Headline workload (default): the precision form of the join, 200M x 200M all-sky
with a galactic-plane-like density, radius = zone height = 1 arcsec.  Source
density is proportional to 1 + 20 exp(-b^2 / (2 (10 deg)^2)), b = latitude
relative to the great circle whose pole is at (ra, dec) = (0, 30) deg, 60 deg from
the celestial pole.  B: 50% copies of A sources moved by a 0.3" Gaussian per
tangent-plane axis, 50% fresh sources of the same density.  Both catalogs are
generated on the device (Philox-4x32-10).

One step = the hot path over one batch, records resident in HBM: prepare catalog
A (keys, radix sort, zone index, unit vectors), prepare catalog B, join A x B into
device-resident (idA, idB, dist) triples.  The step runs in three forms on the
same records:
  fp32_filter_fp64_recheck      the FP32 chord filters with a guard band on the FP64
                                rows rounded to FP32 in registers; every kept pair
                                is rechecked in FP64 (the library's FP32 form);
  fp32_filter_stored_fp32_rows  the same filter on FP32 rows that the catalog
                                builds write in the gather pass beside the FP64
                                rows (fp32_rows = 1);
  fp64_plain                    FP64 rows, the FP64 chord on every candidate.
Both catalogs are rebuilt in every step of every form: nothing carries over.

Metric: pair-tests/s = W_cand / step time.  W_cand = pairs reaching the exact chord
test under the reference's zone and alpha windows, computed on the host from the
sorted zone tables.  value = the FP32-filter form's rate; the other forms' rates,
the ratios, the match sets' digests, the band hits and the largest FP32 chord error
go beside it, with matches/s and the join-only rates.  Per-kernel durations are
measured with HIP events on the stream the kernels run on.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import xmatch_amd as xm  # noqa: E402
from xmatch_amd import synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6
FP32_VECTOR_PEAK_TFLOPS = 157.3
# a pair test in difference form: 3 sub, 1 mul, 2 FMA (= 4 FLOP) -> 8 FLOP
FLOP_PER_PAIR_TEST = 8.0
# untimed steps with events around every launch, for the per-kernel breakdown
BREAKDOWN_STEPS = 3

# algorithmic HBM bytes per source for the memory-bound kernels of a catalog build
BYTES_PER_SOURCE = {
    "radix_scatter": 16.0,   # read packed word 8, write packed word 8
    "radix_hist": 8.0,       # read packed word (once per sort, every digit)
    "make_keys": 32.0,       # read 24-B record, write packed word 8
    # catalogs of 2^24 rows and more: the key pass also writes each record into its
    # zone bucket: read 24, write 24 + 8
    "make_keys_bucketed": 56.0,
    "gather": 68.0,          # read word 8 + 24-B record, write 32-B row + 4-B key
    "gather_fp32": 84.0,     # the same + the 16-B FP32 row
}
KERNELS = ["layout", "bucket_plan", "make_keys", "radix_hist", "radix_offsets",
           "radix_scatter", "scan_reduce", "scan_top", "scan_apply", "gather", "chunk_counts",
           "chunk_fill", "fill", "pair_fp64", "pair_fp32", "region_scan", "compact", "job_prep"]

# rocprofv3 names of kernels whose timing name differs
PROFILE_NAMES = {"pair_fp32": "pair<Fp32Cvt>", "pair_fp64": "pair<Fp64>"}
# the library's default bucket_min_rows: one-segment catalogs from this size up are
# built with their records in zone buckets
BUCKET_MIN_ROWS = 1 << 24

PLANE_SEED = 2003  # device generator key of the headline catalogs
# the forms of the headline join: (name, precision, fp32_rows, --mode key).  The FP32
# filter runs on the FP64 rows rounded in registers (the library's default) or on
# FP32 rows stored by the catalog builds (fp32_rows = 1): the same filter values.
MODES = (("fp32_filter_fp64_recheck", xm.FP32_FILTER_FP64_RECHECK, 0, "fp32"),
         ("fp32_filter_stored_fp32_rows", xm.FP32_FILTER_FP64_RECHECK, 1, "fp32rows"),
         ("fp64_plain", xm.FP64_PLAIN, 0, "fp64"))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel, profile="full_sky_50m_fp64.md"):
    """HBM bytes per launch of `kernel` (a rocprofv3 kernel name without its
    template arguments, less the leading k_) from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (already scaled by the copy calibration), or
    None."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    in_table = False
    for line in open(path):
        if line.startswith("## "):
            in_table = "HBM bytes per dispatch" in line
            continue
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if in_table and len(cells) == 5 and cells[0] == "k_" + kernel:
            try:
                return (float(cells[3]) + float(cells[4])) * 1e6
            except ValueError:
                return None
    return None


def host_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on, capped
    by OMP_NUM_THREADS when it is set (the GPU box allots 16 host threads per GPU)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        return min(avail, int(cap))
    return avail


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_band(A, B, r, threads, target_s):
    """The CPU oracle (FP64 restatement of the reference zone join: its dense
    zone-pair loop with the box and alpha prefilters, sort + zone index + join) on
    a declination band |dec| < w of both catalogs (B widened by 2r), w sized from a
    short calibration run so that the timed run takes about target_s seconds."""
    from oracle import oracle as O

    def run(w):
        sa = A[np.abs(A["dec"]) < w]
        sb = B[np.abs(B["dec"]) < w + 2 * r]
        t0 = time.perf_counter()
        (ia, _, _), cnt = O.match(sa, sb, r, mode=0, threads=threads)
        return time.perf_counter() - t0, sa, sb, ia, cnt

    w = 0.25
    t, *_ = run(w)
    w = float(min(60.0, max(0.05, w * target_s / max(t, 1e-3))))
    t, sa, sb, ia, cnt = run(w)
    return {
        "value": cnt["w_cand"] / t, "unit": "pair-tests/s", "cores": threads, "kind": "port",
        "nproc": os.cpu_count(), "cpu_model": cpu_model(),
        "sample": "A and B restricted to |dec| < %.3f deg (B widened by 2r): %d x %d sources, "
                  "%.2f%% of A; sort + zone index + the reference's dense zone join, %.1f s on "
                  "%d threads" % (w, len(sa), len(sb), 100.0 * len(sa) / len(A), t, threads),
        "seconds": t, "w_cand": cnt["w_cand"], "w_ref": cnt["w_ref"], "matches": len(ia),
    }


def pair_digest(ida, idb, dist):
    """Order-independent digest of a match set: sums of 64-bit mixes of the pairs
    and of the distance bits."""
    h = (ida.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ \
        (idb.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F))
    h ^= h >> np.uint64(29)
    return {"n": int(len(ida)), "pairs": int(np.sum(h, dtype=np.uint64)),
            "dist_bits": int(np.sum(dist.view(np.uint64), dtype=np.uint64))}


def kernel_table(ctx, steps):
    out = {}
    for k in KERNELS:
        tot, cnt = ctx.kernel_time(k)
        if cnt:
            out[k] = {"total_ms": tot / steps, "launches": cnt // steps, "avg_ms": tot / cnt}
    return out


def timed_steps(ctx, step, warmup, steps, only=None):
    """warmup untimed steps, BREAKDOWN_STEPS with events around every launch, then
    `steps` timed steps (events around kernel `only` alone, None: none) bracketed by
    device synchronisation.  Returns (ms per step, breakdown table, (total ms,
    launches) of `only` over the timed steps, the last step's stats)."""
    for _ in range(warmup):
        step()
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    for _ in range(BREAKDOWN_STEPS):
        step()
    ctx.set_timing(False)
    table = kernel_table(ctx, BREAKDOWN_STEPS)
    if only is None:
        only = max(table.items(), key=lambda kv: kv[1]["total_ms"])[0]
    ctx.reset_timing()
    ctx.set_timing(True, only=only)
    ctx.synchronize()
    t0 = time.perf_counter()
    last = None
    for _ in range(steps):
        last = step()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_timing(False)
    dom = ctx.kernel_time(only)
    ctx.reset_timing()
    return 1e3 * elapsed / steps, table, only, dom, last


def make_plane(ctx, n, seed):
    plane = xm.shape(xm.SHAPE_PLANE, synth.PLANE_POLE_RA, synth.PLANE_POLE_DEC, 20.0, 10.0)
    dA = ctx.synth_fresh(plane, n, seed, stream=0, id0=0)
    dB = ctx.synth_copies(dA, 0.5, 0.3, plane, n, seed, stream=1)
    return dA, dB


def join_step(ctx, dA, dB, r, prec):
    """One step: both catalogs built from the resident records, then the join."""
    def step():
        ca = ctx.catalog(dA, r)
        cb = ctx.catalog(dB, r)
        res = ctx.match(ca, cb, r, precision=prec)
        out = (res.count, res.stats.as_dict())
        res.close(); ca.close(); cb.close()
        return out
    return step


def pair_roofline(mode, avg_ms, n_a, n_b, matches, w_tile):
    """Bytes roofline of the pair kernel: each catalog row read once (its key and
    the rows the form reads) and each match written once as a 24-B (idA, idB,
    chord^2) slot, over the kernel's average launch; and the VALU fraction, W_tile
    tests of FLOP_PER_PAIR_TEST against the vector peak of the form's precision."""
    if mode in ("fp64_plain", "fp32_filter_fp64_recheck"):
        # FP64 rows (the FP32 filter rounds them in registers)
        per_row, per_match = 4 + 32, 24
        peak_tf = FP64_VECTOR_PEAK_TFLOPS if mode == "fp64_plain" else FP32_VECTOR_PEAK_TFLOPS
        a_bytes, b_bytes = n_a * per_row, n_b * per_row
    else:
        # FP32 rows for the filter, plus the FP64 row of each matched B row for the
        # recheck and the distance (every A row of a chunk with a kept pair: all of them)
        per_match, peak_tf = 24 + 32, FP32_VECTOR_PEAK_TFLOPS
        a_bytes, b_bytes = n_a * (4 + 16 + 32), n_b * (4 + 16)
    alg = a_bytes + b_bytes + matches * per_match
    ach = alg / (avg_ms / 1e3) / 1e9
    valu = w_tile * FLOP_PER_PAIR_TEST / (avg_ms / 1e3) / 1e12
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": alg,
            "avg_launch_ms": avg_ms,
            "valu": {"achieved_tflops": valu, "peak_tflops": peak_tf, "frac": valu / peak_tf,
                     "flop_per_test": FLOP_PER_PAIR_TEST, "w_tile": w_tile}}


def run_full_sky_device(ctx, n, warmup, steps):
    """The full-sky plain FP64 join, 50M x 50M, r = h = 2": A uniform on the
    sphere, B 60% copies (0.5" per axis) + 40% fresh, generated on the device; one
    step = prepare A + prepare B + join."""
    r = 2.0 / 3600.0
    sphere = xm.shape(xm.SHAPE_SPHERE)
    dA = ctx.synth_fresh(sphere, n, 2002, stream=0, id0=0)
    dB = ctx.synth_copies(dA, 0.6, 0.5, sphere, n, 2002, stream=1)
    ms, table, dom, (dtot, dcnt), (nm, st) = timed_steps(
        ctx, join_step(ctx, dA, dB, r, xm.FP64_PLAIN), warmup, steps)
    ca, cb = ctx.catalog(dA, r), ctx.catalog(dB, r)
    w = ctx.work_counts(ca, cb, r)
    ca.close(); cb.close(); dA.free(); dB.free()
    ctx.release_memory()
    return {"workload": "full-sky plain FP64 zone join, %d x %d, r = h = 2 arcsec (device-"
                        "generated: A uniform on the sphere, B 60%% copies at 0.5 arcsec + 40%% "
                        "fresh)" % (n, n),
            "value": w["w_cand"] / (ms / 1e3), "unit": "pair-tests/s", "ms_per_step": ms,
            "steps": steps, "matches": nm, "matches_per_s": nm / (ms / 1e3),
            "join_ms": st["join_ms"], "w_cand": w["w_cand"], "w_tile": w["w_tile"],
            "w_ref": w["w_ref"], "dominant_kernel": dom, "dominant_avg_ms": dtot / dcnt,
            "kernels": table}


def run_headline(args):
    """The precision form (default): 200M x 200M galactic-plane density at 1",
    each step building both catalogs and joining, in the FP32 filter + FP64
    recheck forms and FP64 plain on the same records."""
    n = args.n or 200_000_000
    r = 1.0 / 3600.0
    ctx = xm.Context(0)
    if args.chunk_rows:
        ctx.set_param("chunk_rows", args.chunk_rows)
    if args.autotune:
        ctx.set_param("autotune", 1)
    if args.key_ra_bits:
        ctx.set_param("key_ra_bits", args.key_ra_bits)
    t0 = time.time()
    dA, dB = make_plane(ctx, n, PLANE_SEED)
    log("generated %d x %d plane-density sources on the device in %.2f s"
        % (n, n, time.time() - t0))
    # work counts from the sorted zone tables (host), one build; with the tuner on,
    # one join first, so that the counts follow the tuned shape
    ca, cb = ctx.catalog(dA, r), ctx.catalog(dB, r)
    if args.autotune:
        ctx.match(ca, cb, r).close()
    w = ctx.work_counts(ca, cb, r)
    ca.close(); cb.close()
    log("work counts: %s" % w)
    modes = {}
    digests = {}
    dom = None
    # the FP32 form first: its dominant kernel is the one whose launches the timed
    # steps of every form carry events around
    for name, prec, rows32, key in MODES:
        if args.mode not in ("all", key):
            continue
        ctx.set_param("fp32_rows", rows32)
        step = join_step(ctx, dA, dB, r, prec)
        ms, table, only, (dtot, dcnt), (nm, st) = timed_steps(ctx, step, args.warmup,
                                                              args.steps, only=dom)
        dom = only
        pk = "pair_fp32" if prec == xm.FP32_FILTER_FP64_RECHECK else "pair_fp64"
        # one more step, untimed, whose pairs are kept for the digest
        ca, cb = ctx.catalog(dA, r), ctx.catalog(dB, r)
        res = ctx.match(ca, cb, r, precision=prec)
        digests[name] = pair_digest(*res.to_host(sort=False))
        res.close(); ca.close(); cb.close()
        ctx.release_memory()
        modes[name] = {
            "ms_per_step": ms, "pair_tests_per_s": w["w_cand"] / (ms / 1e3),
            "matches": nm, "matches_per_s": nm / (ms / 1e3),
            "join_ms": st["join_ms"], "join_pair_tests_per_s": w["w_cand"] / (st["join_ms"] / 1e3),
            "band_hits": st["band_hits"], "max_dist_err_arcsec": st["max_dist_err_arcsec"],
            "n_retries": st["n_retries"],
            "dominant_kernel_avg_ms": dtot / dcnt, "dominant_kernel_launches": dcnt,
            "kernels": table,
            "pair_roofline": pair_roofline(name, table[pk]["avg_ms"], n, n, nm, w["w_tile"]),
        }
        log("%s: %.2f ms per step (join %.2f ms), %d matches"
            % (name, ms, st["join_ms"], nm))
    ctx.set_param("fp32_rows", 0)
    if args.mode != "all":
        # one form alone (for counter runs that must not mix the forms' kernels)
        print(json.dumps({"mode": args.mode, "modes": modes, "w_cand": w["w_cand"],
                          "w_tile": w["w_tile"], "digests": digests}), flush=True)
        ctx.close()
        return
    f32, f64 = modes["fp32_filter_fp64_recheck"], modes["fp64_plain"]
    f32s = modes["fp32_filter_stored_fp32_rows"]
    # the dominant kernel of the FP32 form: its algorithmic bytes per launch over
    # its average launch in the timed steps
    if dom == "make_keys" and n >= BUCKET_MIN_ROWS:
        alg = BYTES_PER_SOURCE["make_keys_bucketed"] * n
    elif dom in BYTES_PER_SOURCE:
        alg = BYTES_PER_SOURCE[dom] * n
    elif dom == "pair_fp32":
        alg = f32["pair_roofline"]["algorithmic_bytes_per_launch"]
    else:
        alg = None
    dom_avg = f32["dominant_kernel_avg_ms"]
    roof = {"kernel": dom, "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "avg_launch_ms": dom_avg, "launches_timed": f32["dominant_kernel_launches"],
            "traffic": measured_traffic(PROFILE_NAMES.get(dom, dom), "plane_200m_precision.md"),
            "algorithmic_bytes_per_launch": alg, "achieved": None, "frac": None}
    if alg:
        roof["achieved"] = alg / (dom_avg / 1e3) / 1e9
        roof["frac"] = roof["achieved"] / HBM_PEAK_GBS
    line = {
        "metric": "pair-tests/s", "value": f32["pair_tests_per_s"], "unit": "pair-tests/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": f32["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 filter + f64 recheck",
        "data": "synthetic (generated on the device, Philox-4x32-10)",
        "config": {"workload": "galactic-plane density zone join, %d x %d, r = h = 1 arcsec, "
                               "FP32 filter + FP64 recheck against FP64 plain; a step builds "
                               "both catalogs and joins" % (n, n),
                   "n_a": n, "n_b": n, "radius_arcsec": 1.0, "seed": PLANE_SEED,
                   "density": "1 + 20 exp(-b^2 / (2 (10 deg)^2)), pole (ra, dec) = (0, 30) deg",
                   "b": "50% copies (0.3 arcsec per axis) + 50% fresh"},
        "matches": f32["matches"], "matches_per_s": f32["matches_per_s"],
        "fp64_plain_value": f64["pair_tests_per_s"],
        "fp32_over_fp64": f32["pair_tests_per_s"] / f64["pair_tests_per_s"],
        "fp32_over_fp64_join_only": f64["join_ms"] / f32["join_ms"],
        "identical_match_set": (digests["fp64_plain"] == digests["fp32_filter_fp64_recheck"]
                                == digests["fp32_filter_stored_fp32_rows"]),
        "fp32_stored_rows_value": f32s["pair_tests_per_s"],
        "fp32_stored_rows_over_fp64": f32s["pair_tests_per_s"] / f64["pair_tests_per_s"],
        "digest": digests["fp64_plain"],
        "band_hits": f32["band_hits"], "max_dist_err_arcsec": f32["max_dist_err_arcsec"],
        "w_ref": w["w_ref"], "w_cand": w["w_cand"], "w_tile": w["w_tile"],
        "modes": modes, "roofline": roof,
        "pair_shape": {"chunk_rows": ctx.get_param("chunk_rows"),
                       "dense_min": ctx.get_param("dense_min"),
                       "autotune": bool(args.autotune), "tune_table": ctx.tune_table()},
        "kernels_from": "%d untimed steps with events around every launch; the timed steps "
                        "carry events around the dominant kernel only" % BREAKDOWN_STEPS,
    }
    if not args.no_full_sky:
        log("full-sky 50M x 50M FP64 ...")
        line["full_sky_50m"] = run_full_sky_device(ctx, 50_000_000, 2, 10)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_threads()
        log("cpu baseline on %d threads ..." % threads)
        A, B = ctx.download(dA), ctx.download(dB)
        line["cpu_baseline"] = cpu_baseline_band(A, B, r, threads, args.cpu_seconds)
        del A, B
    else:
        line["cpu_baseline"] = None
    dA.free(); dB.free()
    ctx.close()
    print(json.dumps(line), flush=True)


def run_full_sky(args):
    """The full-sky plain FP64 join, 50M x 50M, r = h = 2" (host-generated, seeded
    PCG64), one step = prepare A + prepare B + join; the CPU oracle beside it on
    |dec| < cpu_dec_band."""
    from oracle import oracle as O  # noqa: F401  (the cpu_baseline leg)
    n = args.n or 50_000_000
    t0 = time.time()
    A, B, r = synth.config_full_sky(seed=args.seed, n_a=n, n_b=n)
    log("generated %d x %d sources in %.1f s" % (len(A), len(B), time.time() - t0))
    if args.input_order == "dec":
        # spatially ordered input, as survey catalogs are often stored (a diagnostic
        # form; the workload keeps the random order)
        A = A[np.argsort(A["dec"], kind="stable")]
        B = B[np.argsort(B["dec"], kind="stable")]
    ctx = xm.Context(0)
    if args.chunk_rows:
        ctx.set_param("chunk_rows", args.chunk_rows)
    dA = ctx.to_device(A)
    dB = ctx.to_device(B)
    ms, table, dom, (dtot, dcnt), (nm, st) = timed_steps(
        ctx, join_step(ctx, dA, dB, r, xm.FP64_PLAIN), args.warmup, args.steps)
    ca = ctx.catalog(dA, r)
    cb = ctx.catalog(dB, r)
    w = ctx.work_counts(ca, cb, r)
    ca.close(); cb.close()
    dom_avg = dtot / dcnt
    per_launch = BYTES_PER_SOURCE.get(dom, 0.0) * n
    ach = per_launch / (dom_avg / 1e3) / 1e9
    roof = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(dom),
            "algorithmic_bytes_per_launch": per_launch, "avg_launch_ms": dom_avg,
            "launches_timed": dcnt}
    sort_ms = sum(table.get(k, {}).get("total_ms", 0.0)
                  for k in ("radix_hist", "radix_offsets", "radix_scatter"))
    passes = table.get("radix_scatter", {}).get("launches", 0) / 2
    sort_bytes = 2 * n * (8 + passes * 16)
    line = {
        "metric": "pair-tests/s", "value": w["w_cand"] / (ms / 1e3), "unit": "pair-tests/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "full-sky plain FP64 zone join, %d x %d, r = h = 2 arcsec" % (n, n),
                   "n_a": n, "n_b": n, "radius_arcsec": 2.0, "seed": args.seed,
                   "input_order": args.input_order},
        "matches": nm, "matches_per_s": nm / (ms / 1e3),
        "join_ms": st["join_ms"], "join_pair_tests_per_s": w["w_cand"] / (st["join_ms"] / 1e3),
        "w_ref": w["w_ref"], "w_cand": w["w_cand"], "w_tile": w["w_tile"],
        "sort_ms_per_step": sort_ms, "sort_passes_per_catalog": passes,
        "sort_gb_per_s": sort_bytes / (sort_ms / 1e3) / 1e9 if sort_ms else None,
        "kernels": table, "kernels_from": "%d untimed steps with events around every launch"
                                          % BREAKDOWN_STEPS,
        "roofline": roof,
        "pair_roofline": pair_roofline("fp64_plain", table["pair_fp64"]["avg_ms"], n, n, nm,
                                       w["w_tile"]),
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_threads()
        log("cpu baseline on %d threads ..." % threads)
        line["cpu_baseline"] = cpu_baseline_band(A, B, r, threads, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(line), flush=True)


def run_batch(args):
    """Batched form: n_jobs independent joins (cap-shaped segment pairs of logU sizes,
    each at its own radius, zone height = radius) as one many-segment catalog per
    side and one batched join, against a loop of one-job calls (catalog A, catalog
    B, join) over the same device-resident records."""
    n_jobs = args.jobs
    spec = synth.batch_job_specs(seed=4, n_jobs=n_jobs)
    na, nb = spec["n_a"], spec["n_b"]
    rs = [float(r) for r in spec["r_deg"]]
    ctx = xm.Context(0)
    t0 = time.time()
    dA = ctx.alloc_records(int(na.sum()))
    dB = ctx.alloc_records(int(nb.sum()))
    a_views, b_views = [], []
    oa = ob = 0
    for j in range(n_jobs):
        cap = xm.shape(xm.SHAPE_CAP, float(spec["ra_c"][j]), float(spec["dec_c"][j]),
                       float(spec["cap_deg"][j]))
        va = dA.view(oa, int(na[j]))
        vb = dB.view(ob, int(nb[j]))
        ctx.synth_fresh(cap, int(na[j]), 4000 + args.seed, stream=2 * j, id0=oa, out=va)
        ctx.synth_copies(va, 0.5, 0.3, cap, int(nb[j]), 4000 + args.seed, stream=2 * j + 1,
                         id0=10**11 + ob, out=vb)
        a_views.append(va)
        b_views.append(vb)
        oa += int(na[j])
        ob += int(nb[j])
    log("generated %d jobs, %d + %d sources on the device in %.1f s"
        % (n_jobs, oa, ob, time.time() - t0))
    seg_a = xm.SegmentTable(a_views, rs)
    seg_b = xm.SegmentTable(b_views, rs)
    jobs = xm.job_table([(j, j, rs[j]) for j in range(n_jobs)])

    prec = xm.FP32_FILTER_FP64_RECHECK if args.precision == "fp32" else xm.FP64_PLAIN
    if args.chunk_rows:
        ctx.set_param("chunk_rows", args.chunk_rows)

    def batched():
        A = ctx.catalog_segments(seg_a)
        B = ctx.catalog_segments(seg_b)
        res = ctx.match_batch(A, B, jobs, precision=prec)
        return A, B, res

    fixed_ms = None
    if args.autotune:
        # the fixed default shape first, then the tuned shapes (the warmup calls
        # measure every tuning key the batch holds)
        for _ in range(args.warmup):
            A, B, res = batched()
            res.close(); A.close(); B.close()
        ctx.synchronize()
        fm = []
        for i in range(args.steps):
            t = time.perf_counter()
            A, B, res = batched()
            fm.append(1e3 * (time.perf_counter() - t))
            res.close(); A.close(); B.close()
        fixed_ms = float(np.median(fm))
        log("fixed default shape: %.1f ms per step" % fixed_ms)
        ctx.set_param("autotune", 1)
        t0 = time.time()
        A, B, res = batched()
        res.close(); A.close(); B.close()
        log("tuning: %.1f s, %d keys" % (time.time() - t0, len(ctx.tune_table())))

    ctx.set_timing(True)
    for _ in range(args.warmup):
        A, B, res = batched()
        res.close(); A.close(); B.close()
    ctx.synchronize()
    ctx.reset_timing()
    ms = []
    for i in range(args.steps):
        t = time.perf_counter()
        A, B, res = batched()
        ms.append(1e3 * (time.perf_counter() - t))
        if i < args.steps - 1:
            res.close(); A.close(); B.close()
    kernels = {}
    for k in KERNELS + ["seg_range", "job_prep", "pair_fp32", "to_row32"]:
        tot, cnt = ctx.kernel_time(k)
        if cnt:
            kernels[k] = {"total_ms": tot / args.steps, "launches": cnt // args.steps}
    ctx.set_timing(False)
    t_batch = float(np.median(ms))
    stats = res.stats.as_dict()
    _, counts_batch = res.jobs(column=False)
    log("batched: %.1f ms per step, %d matches" % (t_batch, stats["n_matches"]))
    # work counts of every job, from the sorted segment tables (host)
    t0 = time.time()
    w = {"w_ref": 0, "w_cand": 0, "w_tile": 0}
    if not args.no_work_counts:
        for j in range(n_jobs):
            wj = ctx.work_counts_job(A, B, (j, j, rs[j]))
            for k in w:
                w[k] += wj[k]
        log("work counts in %.1f s" % (time.time() - t0))
    res.close(); A.close(); B.close()
    ctx.release_memory()

    def loop():
        counts = np.zeros(n_jobs, np.uint64)
        for j in range(n_jobs):
            ca = ctx.catalog(a_views[j], rs[j])
            cb = ctx.catalog(b_views[j], rs[j])
            r = ctx.match(ca, cb, rs[j], precision=prec)
            counts[j] = r.count
            r.close(); cb.close(); ca.close()
        return counts

    loop()
    ctx.synchronize()
    t = time.perf_counter()
    counts_loop = loop()
    t_loop = 1e3 * (time.perf_counter() - t)
    log("loop of one-job calls: %.1f ms" % t_loop)
    # a sample of jobs compared pair by pair (batch of the sample vs one-job joins)
    rng = np.random.default_rng(0)
    sample = sorted(set(rng.choice(n_jobs, size=min(24, n_jobs), replace=False).tolist()))
    sa = ctx.catalog_segments([a_views[j] for j in sample], [rs[j] for j in sample])
    sb = ctx.catalog_segments([b_views[j] for j in sample], [rs[j] for j in sample])
    sres = ctx.match_batch(sa, sb, [(i, i, rs[j]) for i, j in enumerate(sample)],
                           precision=prec)
    per_job = sres.by_job()
    sample_ok = True
    for i, j in enumerate(sample):
        ca = ctx.catalog(a_views[j], rs[j])
        cb = ctx.catalog(b_views[j], rs[j])
        r1 = ctx.match(ca, cb, rs[j])
        x, y, d = r1.to_host()
        g = per_job[i]
        sample_ok &= bool(np.array_equal(x, g[0]) and np.array_equal(y, g[1]) and
                          (len(d) == 0 or np.abs(d - g[2]).max() <= 1e-6))
        r1.close(); cb.close(); ca.close()
    sres.close(); sa.close(); sb.close()
    n_src = int(na.sum() + nb.sum())
    out = {
        "metric": "pair-tests/s", "unit": "pair-tests/s", "n_gpus": 1, "data": "synthetic",
        "config": {"workload": "batched joins: %d cap-shaped segment pairs, N_A, N_B ~ logU[1e2, "
                               "1e6], cap radius U[0.1, 2] deg, r ~ logU[0.5, 10] arcsec, h = r"
                               % n_jobs,
                   "n_jobs": n_jobs, "sources": n_src, "b": "50% copies (0.3 arcsec) + 50% fresh",
                   "precision": ("fp32_filter_fp64_recheck" if args.precision == "fp32"
                                 else "fp64_plain")},
        "steps": args.steps, "warmup": args.warmup,
        "batched_ms": t_batch, "loop_ms": t_loop, "batched_over_loop": t_loop / t_batch,
        "value": (w["w_cand"] / (t_batch / 1e3)) if w["w_cand"] else None,
        "loop_pair_tests_per_s": (w["w_cand"] / (t_loop / 1e3)) if w["w_cand"] else None,
        "matches": stats["n_matches"], "matches_per_s": stats["n_matches"] / (t_batch / 1e3),
        "sources_per_s": n_src / (t_batch / 1e3), "n_retries": stats["n_retries"],
        "device_join_ms": stats["join_ms"],
        "per_job_counts_equal": bool(np.array_equal(counts_batch, counts_loop)),
        "sample_jobs_identical": sample_ok, "sample_jobs": len(sample),
        "kernels": kernels,
        "shapes": {"autotune": bool(args.autotune), "fixed_default_ms": fixed_ms,
                   "tuned_over_fixed": (fixed_ms / t_batch) if fixed_ms else None,
                   "default": {"chunk_rows": ctx.get_param("chunk_rows"),
                               "dense_min": ctx.get_param("dense_min")},
                   "tune_table": ctx.tune_table()},
    }
    out.update(w)
    ctx.set_param("autotune", 0)
    ctx.set_param("fp32_rows", 0)
    dA.free(); dB.free()
    ctx.close()
    print(json.dumps(out), flush=True)


def make_serving_queries(ctx, res_cat, res_host, n_queries, seed):
    """Requests of the serving form: cap radius U[1, 30] arcmin around a random
    resident source, N ~ logU[1, 1e5] sources, 70% of them 0.5" copies of resident
    sources inside the cap, 30% fresh in the cap; radius r ~ U[1, 5] arcsec.  The
    resident sources inside each cap come from one batched join of the cap centres
    with the resident catalog."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n_res = len(res_host)
    centre = rng.integers(0, n_res, n_queries)
    cap_deg = rng.uniform(1.0, 30.0, n_queries) / 60.0
    nq = np.maximum(1, np.round(10.0 ** rng.uniform(0.0, 5.0, n_queries))).astype(np.int64)
    r_deg = rng.uniform(1.0, 5.0, n_queries) / 3600.0
    h = res_cat.info()["zone_height_deg"]
    segs = [synth.make_objs([i], [res_host["ra"][c]], [res_host["dec"][c]])
            for i, c in enumerate(centre)]
    cc = ctx.catalog_segments(segs, h)
    res = ctx.match_batch(cc, res_cat, [(i, 0, float(cap_deg[i])) for i in range(n_queries)],
                          capacity_hint=1 << 26)
    _, in_cap, _ = res.to_host(sort=False)
    job, counts = res.jobs()
    res.close(); cc.close()
    o = np.argsort(job, kind="stable")
    in_cap = in_cap[o]
    edges = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    queries = []
    id0 = 10**12
    for i in range(n_queries):
        n = int(nq[i])
        k = int(round(0.7 * n))
        members = in_cap[edges[i]:edges[i + 1]]
        if len(members) == 0:
            k = 0
        ra, dec = synth.uniform_cap(rng, n, res_host["ra"][centre[i]], res_host["dec"][centre[i]],
                                    cap_deg[i])
        if k:
            pick = members[rng.integers(0, len(members), k)]
            ra[:k], dec[:k] = synth.perturb(rng, res_host["ra"][pick], res_host["dec"][pick], 0.5)
        queries.append((synth.make_objs(id0 + np.arange(n), ra, dec), float(r_deg[i])))
        id0 += n
    return queries


def run_serve(args):
    """Serving form: a resident 5e8-source catalog (galactic-plane density, h = 1")
    answers requests arriving as a Poisson process; latency = submit -> done,
    measured by the server; p50 / p99 per arrival rate.  --precision fp32: the
    resident catalog stores FP32 unit vectors beside its FP64 rows (fp32_rows) and
    the server runs the FP32 filter with the FP64 recheck."""
    import threading
    n_res = args.n or 500_000_000
    ctx = xm.Context(0)
    fp32 = args.precision == "fp32"
    prec = xm.FP32_FILTER_FP64_RECHECK if fp32 else xm.FP64_PLAIN
    plane = xm.shape(xm.SHAPE_PLANE, synth.PLANE_POLE_RA, synth.PLANE_POLE_DEC, 20.0, 10.0)
    t0 = time.time()
    d_res = ctx.synth_fresh(plane, n_res, 5000 + args.seed, stream=0, id0=0)
    if fp32:
        ctx.set_param("fp32_rows", 1)
    res_cat = ctx.catalog(d_res, 1.0 / 3600)
    res_host = ctx.download(d_res)
    d_res.free()
    log("resident catalog: %d sources, built in %.1f s" % (n_res, time.time() - t0))
    rates = [float(x) for x in args.rates.split(",")]
    per_rate = []
    rng = np.random.Generator(np.random.PCG64(args.seed + 77))
    for rate in rates:
        n_q = max(1, int(round(rate * args.duration)))
        t0 = time.time()
        queries = make_serving_queries(ctx, res_cat, res_host, n_q, seed=int(rate * 1000) + args.seed)
        log("rate %.0f q/s: %d requests generated in %.1f s" % (rate, n_q, time.time() - t0))
        gaps = rng.exponential(1.0 / rate, n_q)
        with xm.Server(ctx, res_cat, precision=prec) as srv:
            # warm the server's buffers with a few requests
            for q, r in queries[:4]:
                srv.result(srv.submit(q, r))
            s0 = srv.stats()
            tickets = [None] * n_q
            lat = np.zeros(n_q)
            nm = np.zeros(n_q, np.int64)
            done = threading.Event()

            def collect():
                for i in range(n_q):
                    while tickets[i] is None:
                        time.sleep(0.0002)
                    ia, _, _, l = srv.result(tickets[i])
                    lat[i] = l
                    nm[i] = len(ia)
                done.set()

            th = threading.Thread(target=collect)
            th.start()
            start = time.perf_counter()
            t_next = start
            for i, (q, r) in enumerate(queries):
                t_next += gaps[i]
                wait = t_next - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
                tickets[i] = srv.submit(q, r)
            done.wait()
            th.join()
            wall = time.perf_counter() - start
            s1 = srv.stats()
        sizes = np.array([len(q) for q, _ in queries])
        small = sizes <= 10_000
        batches = s1["batches"] - s0["batches"]
        per_rate.append({
            "rate_qps": rate, "requests": n_q, "wall_s": wall,
            "achieved_qps": n_q / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "p99_ms_n_le_1e4": float(np.percentile(lat[small], 99)) if small.any() else None,
            "max_ms": float(lat.max()),
            "mean_sources_per_request": float(sizes.mean()),
            "matches": int(nm.sum()), "matches_per_s": float(nm.sum() / wall),
            "batches": batches, "mean_requests_per_batch": n_q / max(batches, 1),
            "worker_busy_frac": (s1["busy_ms"] - s0["busy_ms"]) / 1e3 / wall,
            "prep_busy_frac": (s1["prep_ms"] - s0["prep_ms"]) / 1e3 / wall,
            "overlap_frac": (s1["overlap_ms"] - s0["overlap_ms"]) / 1e3 / wall,
            "pieces": s1["pieces"] - s0["pieces"],
        })
        log("rate %.0f q/s: p50 %.2f ms, p99 %.2f ms (N <= 1e4: %.2f ms), %d batches"
            % (rate, per_rate[-1]["p50_ms"], per_rate[-1]["p99_ms"],
               per_rate[-1]["p99_ms_n_le_1e4"] or -1, batches))
    out = {"metric": "request latency", "unit": "ms", "n_gpus": 1, "data": "synthetic",
           "higher_is_better": False,
           "config": {"workload": "serving: resident %d-source catalog (galactic-plane density, "
                                  "h = 1 arcsec), Poisson arrivals, requests N ~ logU[1, 1e5] in "
                                  "caps of U[1, 30] arcmin, 70%% 0.5-arcsec copies of resident "
                                  "sources, r ~ U[1, 5] arcsec" % n_res,
                      "resident_sources": n_res, "duration_s_per_rate": args.duration,
                      "precision": "fp32_filter_fp64_recheck, FP32 rows stored" if fp32
                      else "fp64_plain"},
           "rates": per_rate}
    res_cat.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="plane",
                    choices=["plane", "full_sky", "batch", "serve"],
                    help="plane: the 200M x 200M precision-form join, FP32 filter against "
                         "FP64 plain (default, the headline); full_sky: the 50M x 50M FP64 "
                         "join; batch: many independent joins, batched against a loop of "
                         "one-job calls; serve: a resident catalog answering Poisson request "
                         "streams")
    ap.add_argument("--rates", default="50,200,800", help="serve config: arrival rates, q/s")
    ap.add_argument("--duration", type=float, default=10.0,
                    help="serve config: seconds of arrivals per rate")
    ap.add_argument("--jobs", type=int, default=10_000, help="jobs of the batch config")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="batch and serve configs: FP64 plain or the FP32 filter with FP64 "
                         "recheck")
    ap.add_argument("--no-work-counts", action="store_true",
                    help="batch config: skip the host work counts")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=0,
                    help="sources per catalog (0: the config's own: plane 2e8, full_sky 5e7, "
                         "serve 5e8 resident)")
    ap.add_argument("--seed", type=int, default=2, help="full_sky and serve configs: seed")
    ap.add_argument("--input-order", choices=["random", "dec"], default="random",
                    help="full_sky config: record order of both catalogs (dec = sorted by "
                         "declination; a diagnostic form)")
    ap.add_argument("--mode", choices=["all", "fp32", "fp32rows", "fp64"], default="all",
                    help="plane config: every form (the headline line) or one alone: the FP32 "
                         "filter on FP64 rows rounded in registers, on stored FP32 rows, or "
                         "FP64 plain")
    ap.add_argument("--no-full-sky", action="store_true",
                    help="plane config: skip the 50M x 50M FP64 figures reported beside it")
    ap.add_argument("--chunk-rows", type=int, default=0,
                    help="pair kernel: A rows per chunk (a multiple of 64 up to 512; 0: the "
                         "library default)")
    ap.add_argument("--key-ra-bits", type=int, default=0,
                    help="catalog sort key: at most this many ra bits (0: the library's choice)")
    ap.add_argument("--autotune", action="store_true",
                    help="pair kernel: shapes picked by the library's tuner")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU baseline: target length of the timed oracle run")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of 1 GiB (for scaling PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library drives one GPU and has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": 1 << 30,
                          "bytes_written": 1 << 30, "launches": 3}), flush=True)
        return
    {"plane": run_headline, "full_sky": run_full_sky, "batch": run_batch,
     "serve": run_serve}[args.config](args)


if __name__ == "__main__":
    main()
