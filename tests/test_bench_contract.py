"""The parts of bench.py's output contract that need no GPU: it refuses more than
one GPU with a message, and the roofline's traffic figure is read from the
committed rocprofv3 summary of the headline workload."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_traffic_is_read_from_the_committed_profile():
    import bench
    plane = "plane_200m_precision.md"
    t = bench.measured_traffic("gather", plane)
    # FETCH_SIZE x 2.0 + WRITE_SIZE of the 200M-row gather: tens of GB per launch,
    # above its algorithmic bytes (its record reads, from zone buckets, fetch some
    # lines more than once)
    assert t is not None and t > bench.BYTES_PER_SOURCE["gather"] * 200_000_000
    # the sort's scatter pass, which may be the dominant kernel instead
    s = bench.measured_traffic("radix_scatter", plane)
    assert s is not None and s > bench.BYTES_PER_SOURCE["radix_scatter"] * 200_000_000
    # the 50M-row full-sky summary
    assert bench.measured_traffic("gather") > bench.BYTES_PER_SOURCE["gather"] * 50_000_000
    assert bench.measured_traffic("no_such_kernel") is None
