"""bench.py's command line on a machine without a GPU: more than one GPU is
refused with a message before anything touches a device, and every
workload name and variant is known to the parser."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in r.stderr + r.stdout
    assert r.stdout.strip() == ""


def test_unknown_workload_is_rejected():
    r = run("--workload", "no-such-workload")
    assert r.returncode != 0


def test_workloads_listed():
    sys.path.insert(0, ROOT)
    import bench
    assert {"ring", "cube", "lattice", "clouds", "sheet", "relax", "influence", "remesh",
            "small"} <= set(bench.WORKLOADS)
    # the flop counts the roofline uses are the frozen per-pair constants
    assert bench.FLOPS_PER_PAIR[("p3d", "vel+dvort", "gaussian")] == 79
    assert bench.FLOPS_PER_PAIR[("p3d", "vel+dvort", "winckelmans")] == 61
    assert bench.FLOPS_PER_PAIR[("f3d", "vel+dvort", "singular")] == 100
