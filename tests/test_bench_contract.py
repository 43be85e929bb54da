"""bench.py's command-line contract, on the CPU: one GPU only (a request for more exits
with a JSON message and status 2), and the host description the CPU baseline reports."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_host_cpu_description():
    sys.path.insert(0, ROOT)
    import bench

    h = bench.host_cpu()
    assert h["physical_cores"] >= 1 and h["logical_cpus"] >= h["physical_cores"]
