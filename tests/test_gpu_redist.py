"""Redistribution on a grid (remeshing) on the GPU against the oracle's
restatement of the reference CPU path (reference src/P3D.cpp:509-665,
src/P2D.cpp:283-436).  The node set, its Morton order and the node
positions must be identical (the weights are the reference's FP32
operations); strengths within relative L2 1e-5 (the reference sums them in
FP32, the library in FP64).  With an output limit the histogram threshold
search can move by a rounding step between the two, so that case is checked
through its invariants (at most max_out particles, all of them nodes of the
unlimited result, total strength conserved) and a count close to the
oracle's."""
import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import workloads as W
from gpuutil import assert_parity, case, have_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-5
RD = ("lambda0", "lambda1", "lambda2", "lambda3", "m4p")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")


def cloud3(n, seed):
    p = W.random_p3d(n, seed, box=1.0, vort=1.0, vol=0.001)
    p["vorticity"] -= 0.5
    return p


def cloud2(n, seed):
    q = W.random_p2d(n, seed, box=1.0, vort=1.0)
    q["vorticity"] -= 0.5
    return q


def same_nodes(name, got, ref):
    assert len(got) == len(ref), f"{name}: {len(got)} particles, reference {len(ref)}"
    assert np.array_equal(got["coord"], ref["coord"]), f"{name}: node positions or order differ"


@pytest.mark.parametrize("r", RD)
@pytest.mark.parametrize("neg", [0.0, 0.2])
def test_p3d_redistribution_matches_oracle(orc, r, neg):
    case(f"P3D redistribute {r} negligible={neg}")
    p = cloud3(500, 1001)
    got = cv.P3D_redistribute_on_grid(p, r, 0.07, negligible=neg)
    ref = orc.P3D_redistribute_on_grid(p, r, 0.07, negligible=neg)
    same_nodes(f"P3D {r}", got, ref)
    assert_parity(f"P3D {r} strengths", got["vorticity"], ref["vorticity"], TOL)
    assert np.array_equal(got["volume"], ref["volume"])


@pytest.mark.parametrize("r", RD)
@pytest.mark.parametrize("neg", [0.0, 0.2])
def test_p2d_redistribution_matches_oracle(orc, r, neg):
    case(f"P2D redistribute {r} negligible={neg}")
    q = cloud2(700, 1002)
    got = cv.P2D_redistribute_on_grid(q, r, 0.05, negligible=neg)
    ref = orc.P2D_redistribute_on_grid(q, r, 0.05, negligible=neg)
    same_nodes(f"P2D {r}", got, ref)
    assert_parity(f"P2D {r} strengths", got["vorticity"], ref["vorticity"], TOL)
    assert np.array_equal(got["area"], ref["area"])


@pytest.mark.parametrize("r", ("lambda1", "m4p"))
def test_output_limit(orc, r):
    case(f"P3D redistribute {r} with an output limit")
    p = cloud3(400, 1003)
    full = cv.P3D_redistribute_on_grid(p, r, 0.07)
    lim = len(full) // 3
    got = cv.P3D_redistribute_on_grid(p, r, 0.07, max_out=lim)
    ref = orc.P3D_redistribute_on_grid(p, r, 0.07, max_out=lim)
    assert 0 < len(got) <= lim
    assert abs(len(got) - len(ref)) <= max(2, len(ref) // 50)
    # every survivor is a node of the unlimited result, in the same order
    pos = {tuple(c) for c in full["coord"]}
    assert all(tuple(c) in pos for c in got["coord"])
    a0 = p["vorticity"].astype(np.float64).sum(0)
    assert np.abs(got["vorticity"].astype(np.float64).sum(0) - a0).max() < 1e-5 * np.abs(p["vorticity"]).sum()


def test_larger_cloud_conserves_moments(orc):
    case("P3D redistribute m4p n=20000")
    p = W.random_p3d(20000, 1004, box=1.0, vort=1.0, vol=5e-5)
    p["vorticity"] -= 0.5
    got = cv.P3D_redistribute_on_grid(p, "m4p", 0.02)
    ref = orc.P3D_redistribute_on_grid(p, "m4p", 0.02)
    same_nodes("P3D m4p 20000", got, ref)
    assert_parity("P3D m4p 20000 strengths", got["vorticity"], ref["vorticity"], TOL)
    x = p["coord"].astype(np.float64)
    a = p["vorticity"].astype(np.float64)
    y = got["coord"].astype(np.float64)
    b = got["vorticity"].astype(np.float64)
    assert np.abs(b.sum(0) - a.sum(0)).max() < 1e-6 * np.abs(a).sum()
    assert np.abs(y.T @ b - x.T @ a).max() < 1e-5 * np.abs(x.T @ a).max()


def test_count_only_and_empty():
    case("redistribute count-only and empty")
    p = cloud3(100, 1005)
    n = cv.lib().cvtx_P3D_redistribute_on_grid
    from cvortex_amd.types import pointer_array
    ptrs, keep = pointer_array(p)
    import ctypes
    r = cv.redistfunc("m4p")
    cnt = n(ctypes.c_void_p(ptrs.ctypes.data), len(p), None, 0, ctypes.byref(r), 0.1, 0.0)
    assert cnt == len(cv.P3D_redistribute_on_grid(p, "m4p", 0.1))
    assert len(cv.P3D_redistribute_on_grid(p[:0], "m4p", 0.1)) == 0
