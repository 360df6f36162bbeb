"""bench.py's contract on the CPU: the one-GPU guard, the committed profile
figures it reads for the roofline, and the host description."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_exits_with_a_message(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_profile_figures_for_the_roofline():
    import bench
    for k in ("density", "force", "reorder", "sort"):
        assert bench.load_valu(k) > 0, k
        assert bench.load_traffic(k) > 0, k
    # The staged forms of density and force have figures of their own, picked
    # by the tuning the bench ran with.
    for k in ("density", "force"):
        staged = {f"{k}_staged": 1}
        assert bench.profile_key(k, staged) == f"{k}_staged" and bench.profile_key(k, {}) == k
        assert bench.load_valu(k, staged) > 0 and bench.load_valu(k, staged) != bench.load_valu(k)
        assert bench.load_traffic(k, staged) > 0
    t = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))
    # HBM bytes per launch of the force pass: at least its state writes
    # (new position and velocity, 32 B per particle of the 1M workload).
    assert t["force"]["bytes_per_launch"] >= 32 * 1048576
    assert t["copy_fetch_factor"] > 0 and t["copy_write_factor"] > 0


def test_algorithmic_bytes_per_step():
    import bench
    # sort 16 + reorder 56 + density 16 + force and integrate 100; the hash
    # pass moves no bytes per step.
    assert sum(v for k, v in bench.B_ALG.items() if k != "hash") == 188.0


def test_host_info():
    import bench
    h = bench.host_info()
    assert h["logical_cpus"] >= 1 and h["usable_cpus"] >= 1 and h["numa_nodes"] >= 0
