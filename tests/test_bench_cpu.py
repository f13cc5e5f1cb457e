"""bench.py's host-side pieces: the CPU-baseline child process and the
one-GPU guard (no GPU needed)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")
sys.path.insert(0, os.path.join(ROOT, "tests"))
import _oracle as O  # noqa: E402


@pytest.mark.skipif(not os.path.exists(O.REF_OMP_SO), reason="oracle/_ref not built")
def test_cpu_child_runs_reference_to_each_time():
    """The child runs the reference's OpenMP build from the initial state to
    each requested time; the later time takes more steps and more time."""
    env = dict(os.environ, OMP_NUM_THREADS="2", OMP_PLACES="cores", OMP_PROC_BIND="close")
    out = subprocess.run([sys.executable, BENCH, "--cpu-child", "64", "0.002", "0.02"], env=env,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    assert res["kind"] == "reference"
    (s0, t0), (s1, t1) = res["runs"]
    assert 0 < s0 < s1 and 0.0 < t0 and 0.0 < t1


def test_host_topology_reports_cores():
    sys.path.insert(0, ROOT)
    import bench
    topo = bench.host_topology()
    assert topo["physical_cores_allowed"] >= 1
    assert topo["logical_cpus_allowed"] >= topo["physical_cores_allowed"]


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr
