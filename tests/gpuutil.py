"""Helpers for the GPU parity tests."""
import sys

import numpy as np

import cvortex_amd as cv
import oracle


def have_gpu():
    try:
        cv.initialise()
        return cv.num_accelerators() > 0
    except Exception:
        return False


def case(msg):
    """Name the case on stdout before a GPU step (survives a lost box)."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def report(name, got, ref, ref64=None):
    e = oracle.rel_l2(got, ref)
    m = oracle.ref_test_metric(got, ref)
    line = f"  {name}: relL2 vs ref {e:.2e}, max |a-b|/|a+b| {m.max() if m.size else 0:.2e}"
    if ref64 is not None:
        r = oracle.per_target_rel(got, ref64, floor=1e-30)
        line += f", vs FP64 max {r.max() if r.size else 0:.2e} rms {np.sqrt(np.mean(r * r)) if r.size else 0:.2e}"
    print(line, flush=True)
    return e


def assert_parity(name, got, ref, tol=1e-5):
    got = np.asarray(got)
    assert np.all(np.isfinite(got)), f"{name}: non-finite GPU output"
    e = report(name, got, ref)
    assert e <= tol, f"{name}: relative L2 {e:.3e} > {tol:g}"
