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


def test_json_line_carries_the_contract_fields():
    sys.path.insert(0, ROOT)
    import argparse
    import bench
    line = bench.base_line(argparse.Namespace(steps=5, warmup=2), 1.6e12, "interactions/s",
                           660.0, {"workload": "ring"})
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config"):
        assert key in line
    assert line["n_gpus"] == 1 and line["vs_baseline"] is None and line["higher_is_better"]
    rf = bench.roofline(1.0995e12, 79, 660.0)
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in rf
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12
    assert abs(rf["achieved"] - 1.0995e12 * 79 / 0.660 / 1e12) < 1e-9
    # the counter-measured HBM bytes of the ring and the cube are on record
    assert bench.traffic_for("p3d_ring_vel_dvort_gaussian", 1 << 20) > 1e9
    assert bench.traffic_for("p3d_cube_vel_winckelmans", 8192) > 1e6
    assert bench.traffic_for("p3d_cube_vel_winckelmans", 4096) is None
