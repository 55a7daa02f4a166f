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


def test_tuning_table_written_elsewhere_keeps_the_source_entries(tmp_path):
    """`python -m emma_amd.tune --out NEW`: the new table starts from the
    table the per-level choices were read from, so a schedule-only run does
    not drop them; an existing output table is merged into"""
    import json

    from emma_amd import tune

    src = tmp_path / "src.json"
    src.write_text(json.dumps({"G3-complete-64": {"variant": "listed", "chem_budget": 8, "chem_keep": 0.0,
                                                   "us": 1.0, "default_us": 2.0, "workload": "w", "level": 5}}))
    sched = {"key": "tree-G3-64", "sched": "persistent", "ms": {"graph": 0.7, "persistent": 0.69}}
    out = tmp_path / "out.json"
    t = tune.save({}, "w2", str(out), sched, base=str(src))
    assert set(t) == {"G3-complete-64", "tree-G3-64"}
    assert json.loads(out.read_text()) == t
    sched2 = {"key": "tree-G3-128", "sched": "graph", "ms": {"graph": 1.0, "persistent": 1.1}}
    t2 = tune.save({}, "w3", str(out), sched2, base=str(src))
    assert set(t2) == {"G3-complete-64", "tree-G3-64", "tree-G3-128"}
    assert tune.save({}, "w4", str(tmp_path / "fresh.json"), sched) == {"tree-G3-64": t["tree-G3-64"] | {"workload": "w4"}}


def test_tuner_times_without_counters_and_restores_them():
    """the tuner's timing helpers run with the device counters off, as the
    bench's timed steps do, and give the context back as it was, also when
    the timing raises"""
    import pytest

    from emma_amd import tune

    class Ctx:
        def __init__(self, on):
            self.counters = on
            self.seen = []

        def enable_counters(self, on):
            self.counters = bool(on)

    @tune._uncounted
    def timed(g, fail):
        g.seen.append(g.counters)
        if fail:
            raise RuntimeError("timing failed")
        return 1.0

    for on in (True, False):
        g = Ctx(on)
        assert timed(g, False) == 1.0 and g.seen == [False] and g.counters is on
        with pytest.raises(RuntimeError):
            timed(g, True)
        assert g.seen == [False, False] and g.counters is on
