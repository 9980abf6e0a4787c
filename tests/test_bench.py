"""bench.py's contract checks that need no GPU: one GPU only, and the HBM
traffic figure it reports comes from the committed rocprofv3 summary."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_comes_from_the_committed_profile():
    sys.path.insert(0, ROOT)
    import bench

    t = bench.read_traffic()
    text = open(bench.TRAFFIC_PROFILE).read()
    assert t is not None and t > 0
    # the figure is FETCH + WRITE per launch, scaled by the copy calibration
    fetch = float(text.split("fetch_bytes_per_launch = ")[1].split()[0])
    write = float(text.split("write_bytes_per_launch = ")[1].split()[0])
    assert abs(t - (fetch + write)) <= 1e-3 * t
