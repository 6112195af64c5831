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


@pytest.mark.gpu
def test_segments_rounds_and_reference_format(tmp_path):
    # 2500-object segments, B in rounds of 5000 records: 2 rounds of 4 x 2 jobs,
    # written as 16-byte {idA, idB} records to <outdir>/0 and <outdir>/1
    out = tmp_path / "out"
    out.mkdir()
    (out / "stale").write_bytes(b"x")
    r = run(["-o", str(out), "-x", "-r", "2", "-n", "2500", "-b", "5000", "-f", "16", "-v",
             SAMPLE, SAMPLE])
    assert r.returncode == 0, r.stdout + r.stderr
    assert not (out / "stale").exists()
    g = np.load(os.path.join(GOLD, "selfjoin_r2.npz"))
    x, nfiles = read_pairs(out, fmt=16)
    assert nfiles == 2
    assert np.array_equal(x["a"], g["ida"]) and np.array_equal(x["b"], g["idb"])


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
    assert nfiles == 4
    os.remove(out / "2")                   # as if the run had stopped in round 2
    (out / "2.part").write_bytes(b"partial")
    r = run(args[:2] + ["-c"] + args[2:])
    assert r.returncode == 0 and "3 round(s) done before" in r.stdout
    again, nfiles = read_pairs(out)
    assert nfiles == 4 and np.array_equal(again, full)
