"""bench.py's command line on the CPU: more than one GPU is refused with a
message (the build measures one GPU's shard), and the workload names and
algorithmic byte counts are the documented ones."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_per_cell_update():
    # reads N,F (4G), src (G), n_H, x, eint; writes N,F (4G), x, eint, T
    assert bench.alg_bytes_per_cell(1, "fp64") == 120
    assert bench.alg_bytes_per_cell(3, "fp64") == 264
    assert bench.alg_bytes_per_cell(1, "mixed") == 84
    assert bench.alg_bytes_per_cell(3, "mixed") == 156


def test_workload_names():
    assert set(bench.WORKLOAD_NAMES) == {"c2", "c3", "c4", "c5"}
    assert "128^3" in bench.WORKLOAD_NAMES["c3"]
