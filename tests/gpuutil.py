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


def assert_parity_limited(name, got, ref32, exact, tol=1e-5):
    """Parity where the reference's own FP32 evaluation is the limit.

    Two pair terms of the reference lose accuracy in FP32 by themselves:
    the Gaussian g(rho), evaluated with the Abramowitz & Stegun erf
    (src/VortFunc.cpp:155-176), and the filament stretching, whose |r1 x r0|
    cancels near a filament (src/F3D.cpp:68-73).  The GPU evaluates both
    more accurately, so its distance to the reference is the reference's
    own error.  Require: the GPU matches the FP64 oracle to `tol`, and is no
    farther from the reference than max(tol, 1.5 x the reference's distance
    from the FP64 oracle).
    """
    got = np.asarray(got)
    assert np.all(np.isfinite(got)), f"{name}: non-finite GPU output"
    e_ex = oracle.rel_l2(got, exact)
    e_ref = oracle.rel_l2(got, ref32)
    d_ref = oracle.rel_l2(ref32, exact)
    print(f"  {name}: relL2 vs FP64 {e_ex:.2e}, vs ref {e_ref:.2e} (ref vs FP64 {d_ref:.2e})",
          flush=True)
    assert e_ex <= tol, f"{name}: relative L2 vs FP64 oracle {e_ex:.3e} > {tol:g}"
    assert e_ref <= max(tol, 1.5 * d_ref), (
        f"{name}: relative L2 vs reference {e_ref:.3e} > max({tol:g}, 1.5 x {d_ref:.3e})")


assert_parity_gaussian = assert_parity_limited
