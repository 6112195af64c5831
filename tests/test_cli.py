"""The command-line tool (xmatch_amd/xmatch, built from csrc/xm_cli.cpp): the
reference's flags, defaults and exit codes (BoostXmatch.cpp:95-193), and on the GPU
its output files against the golden self-join sets of the sample data."""
import os
import subprocess

import numpy as np
import pytest

import xmatch_amd.build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
SAMPLE = os.path.join(GOLD, "test.idradec")


def run(args, stdin=""):
    if not os.path.exists(B.CLI):
        pytest.skip("command-line tool not built")
    return subprocess.run([B.CLI] + args, input=stdin, capture_output=True, text=True, timeout=120)


def test_help_exits_1():
    r = run(["-h"])
    assert r.returncode == 1 and "Usage:" in r.stdout and "--radius" in r.stdout


def test_needs_two_inputs():
    r = run([SAMPLE])
    assert r.returncode == 2 and "Input 2 files!" in r.stdout


def test_missing_file_exits_3(tmp_path):
    r = run([SAMPLE, str(tmp_path / "nope.idradec")])
    assert r.returncode == 3 and "File not found!" in r.stdout


@pytest.mark.parametrize("bad", [["--bogus"], ["-f", "12"], ["-r", "abc"], ["-r", "-1"],
                                 ["-p", "fp16"], ["-n"]])
def test_bad_options_exit_1(bad):
    r = run(bad + [SAMPLE, SAMPLE])
    assert r.returncode == 1 and "Usage:" in r.stdout


def test_existing_outdir_without_overwrite(tmp_path):
    out = tmp_path / "out"
    out.mkdir()
    (out / "0").write_bytes(b"old")
    r = run(["-o", str(out), SAMPLE, SAMPLE], stdin="no\n")
    assert r.returncode == 255
    assert "use option -x to overwrite" in r.stdout
    assert (out / "0").read_bytes() == b"old"


def read_pairs(outdir, fmt=24):
    rec = np.dtype([("a", "<i8"), ("b", "<i8")] + ([("d", "<f8")] if fmt == 24 else []))
    parts = [np.fromfile(os.path.join(outdir, f), dtype=rec)
             for f in sorted((f for f in os.listdir(outdir) if f.isdigit()), key=int)]
    x = np.concatenate(parts) if parts else np.empty(0, rec)
    o = np.lexsort((x["b"], x["a"]))
    return x[o], len(parts)


@pytest.mark.gpu
@pytest.mark.parametrize("r_arcsec", [1, 2, 5])
def test_selfjoin_output_matches_golden(tmp_path, r_arcsec):
    out = tmp_path / "out"
    r = run(["-o", str(out), "-r", str(r_arcsec), SAMPLE, SAMPLE])
    assert r.returncode == 0, r.stdout + r.stderr
    g = np.load(os.path.join(GOLD, "selfjoin_r%d.npz" % r_arcsec))
    x, nfiles = read_pairs(out)
    assert nfiles == 1
    assert np.array_equal(x["a"], g["ida"]) and np.array_equal(x["b"], g["idb"])
    assert np.abs(x["d"] - g["dist_arcsec"]).max() <= 1e-6
    assert r.stdout.startswith("%d matches" % len(g["ida"]))


def manifest_plan(outdir):
    """(bands, round rows) of the plan in <outdir>/manifest."""
    lines = open(os.path.join(outdir, "manifest")).read().splitlines()
    plan = [ln for ln in lines if ln.startswith("plan ")][0].split()
    kv = dict(x.split("=") for x in plan[1:])
    bands = [tuple(int(v) for v in ln.split()[1:]) for ln in lines if ln.startswith("band ")]
    assert len(bands) == int(kv["bands"])
    return bands, int(kv["round_rows"])


@pytest.mark.gpu
def test_segments_rounds_and_reference_format(tmp_path):
    # -n 2500: A cut into declination bands of at most 2500 records, each one
    # segment; B streamed in rounds of 5000 records; 16-byte {idA, idB} records
    out = tmp_path / "out"
    out.mkdir()
    (out / "stale").write_bytes(b"x")
    r = run(["-o", str(out), "-x", "-r", "2", "-n", "2500", "-b", "5000", "-f", "16", "-v",
             SAMPLE, SAMPLE])
    assert r.returncode == 0, r.stdout + r.stderr
    assert not (out / "stale").exists()
    g = np.load(os.path.join(GOLD, "selfjoin_r2.npz"))
    x, nfiles = read_pairs(out, fmt=16)
    bands, rows = manifest_plan(out)
    assert len(bands) >= 4 and rows == 5000
    assert all(b[2] <= 2500 for b in bands) and sum(b[2] for b in bands) == 10_000
    assert nfiles >= len(bands)
    assert np.array_equal(x["a"], g["ida"]) and np.array_equal(x["b"], g["idb"])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_declination_bands_and_rounds_equal_the_golden(tmp_path, precision):
    # A out of core: bands of at most 3000 records (at least 3), B streamed past
    # each band in rounds of at most 1500 records (at least 2 in a band); every
    # A row in one band, so the union of the round files is the self-join once
    out = tmp_path / "out"
    r = run(["-o", str(out), "-r", "2", "-a", "3000", "-b", "1500", "-p", precision, "-v",
             SAMPLE, SAMPLE])
    assert r.returncode == 0, r.stdout + r.stderr
    bands, rows = manifest_plan(out)
    assert len(bands) >= 3 and rows == 1500
    assert all(b[2] <= 3000 for b in bands) and sum(b[2] for b in bands) == 10_000
    # declination bands: disjoint zone ranges in order
    assert all(bands[i][1] < bands[i + 1][0] for i in range(len(bands) - 1))
    x, nfiles = read_pairs(out)
    assert nfiles >= len(bands) + 1    # some band took two rounds or more
    g = np.load(os.path.join(GOLD, "selfjoin_r2.npz"))
    assert np.array_equal(x["a"], g["ida"]) and np.array_equal(x["b"], g["idb"])
    assert np.abs(x["d"] - g["dist_arcsec"]).max() <= 1e-6


@pytest.mark.gpu
def test_count_only_and_fp32(tmp_path):
    g = np.load(os.path.join(GOLD, "selfjoin_r5.npz"))
    r = run(["-r", "5", "-p", "fp32", SAMPLE, SAMPLE])
    assert r.returncode == 0 and r.stdout.startswith("%d matches" % len(g["ida"]))


@pytest.mark.gpu
def test_resume_skips_finished_rounds(tmp_path):
    out = tmp_path / "out"
    args = ["-o", str(out), "-r", "2", "-n", "2500", "-b", "2500", SAMPLE, SAMPLE]
    r = run(args)
    assert r.returncode == 0, r.stdout + r.stderr
    full, nfiles = read_pairs(out)
    assert nfiles >= 4
    os.remove(out / "2")                   # as if the run had stopped in round 2
    (out / "2.part").write_bytes(b"partial")
    r = run(args[:2] + ["-c"] + args[2:])
    assert r.returncode == 0 and "%d round(s) done before" % (nfiles - 1) in r.stdout
    again, n2 = read_pairs(out)
    assert n2 == nfiles and np.array_equal(again, full)


def write_manifest_of(tmp_path, args):
    """An output directory whose manifest another run wrote: the first lines of a
    manifest name the inputs and the options."""
    out = tmp_path / "out"
    out.mkdir()
    ident = ["xmatch-manifest 1"]
    for tag, path in (("a", SAMPLE), ("b", SAMPLE)):
        st = os.stat(path)
        ident.append("%s %d %d.%d %s" % (tag, st.st_size, st.st_mtime_ns // 10**9,
                                         st.st_mtime_ns % 10**9, path))
    ident.append(args)
    (out / "manifest").write_text("\n".join(ident) + "\nplan round_rows=5000 bands=1\n"
                                  "band 0 648000 10000\n")
    return out


def test_resume_refuses_other_options(tmp_path):
    out = write_manifest_of(tmp_path, "options r=2 z=2 n=0 f=24 p=fp64 a=0 b=0")
    r = run(["-o", str(out), "-c", "-r", "5", SAMPLE, SAMPLE])
    assert r.returncode == 1 and "other inputs or options" in r.stdout + r.stderr
    r = run(["-o", str(out), "-c", "-r", "2", "-p", "fp32", SAMPLE, SAMPLE])
    assert r.returncode == 1 and "other inputs or options" in r.stdout + r.stderr


def test_resume_refuses_other_inputs(tmp_path):
    out = write_manifest_of(tmp_path, "options r=2 z=2 n=0 f=24 p=fp64 a=0 b=0")
    other = tmp_path / "other.idradec"
    other.write_bytes(open(SAMPLE, "rb").read())
    r = run(["-o", str(out), "-c", "-r", "2", SAMPLE, str(other)])
    assert r.returncode == 1 and "other inputs or options" in r.stdout + r.stderr


def test_resume_without_manifest_is_refused(tmp_path):
    out = tmp_path / "out"
    out.mkdir()
    (out / "0").write_bytes(b"")
    r = run(["-o", str(out), "-c", "-r", "2", SAMPLE, SAMPLE])
    assert r.returncode == 1 and "no manifest" in r.stdout + r.stderr


def test_resume_with_the_same_options_passes_the_check(tmp_path):
    # the same inputs and options: the run goes on to the GPU (on a machine
    # without one it stops there, after the check)
    out = write_manifest_of(tmp_path, "options r=2 z=2 n=0 f=24 p=fp64 a=0 b=0")
    r = run(["-o", str(out), "-c", "-r", "2", SAMPLE, SAMPLE])
    assert r.returncode == 0 or (r.returncode == 4 and "no usable GPU" in r.stderr), \
        r.stdout + r.stderr
