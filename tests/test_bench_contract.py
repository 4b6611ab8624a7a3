"""The host side of bench.py that needs no GPU: the refusal of more than one GPU,
the calibrated traffic read from the committed profile and the CPU description of
the baseline."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_measured_traffic_reads_the_committed_profile():
    path = os.path.join(ROOT, "profiles", "collide_traffic_256cubed.txt")
    text = open(path).read()
    m = re.search(r"traffic_bytes_per_launch k_collide_march_fp32 (\d+)", text)
    assert m
    t = bench.measured_traffic("k_collide_march_fp32")
    assert t == int(m.group(1))
    # the calibrated traffic of the FP32 collide lies between its algorithmic
    # population bytes (152 B per node at 256^3) and twice that
    assert 152 * 256 ** 3 < t < 2 * 152 * 256 ** 3
    assert bench.measured_traffic("no_such_kernel") is None


def test_cpu_model_names_the_host():
    s = bench.cpu_model()
    assert "logical CPUs" in s
    assert s.split(" (")[0]
