"""Vorticity field (cvtx_P3D_M2M_vort) and Pedrizzetti relaxation on the GPU
against the oracle's restatement of the reference CPU path (reference
src/P3D.cpp:298-322, 458-498, 669-707).  Tolerance: relative L2 <= 1e-5
over all targets, written per assertion.  Sizes large enough to take the
sorted, tile-skipping path (n_src >= 4096, n_src * n_tgt >= 2^26) are
checked on sampled targets against the oracle and against the unsorted
path."""
import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import device as D
from cvortex_amd import workloads as W
from cvortex_amd.types import make_p3d
from gpuutil import assert_parity, case, have_gpu
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5
REGS = ("singular", "winckelmans", "planetary", "gaussian")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")
    cv.set_accumulation(cv.ACC_COMP)


@pytest.mark.parametrize("k", REGS)
@pytest.mark.parametrize("ns,nt,sigma", [(1, 1, 0.3), (257, 513, 0.08), (2000, 700, 0.05),
                                         (3000, 3000, 0.3)])
def test_vort_matches_oracle(orc, k, ns, nt, sigma):
    case(f"vort {k} ns={ns} nt={nt} sigma={sigma}")
    p = W.random_p3d(ns, 100 + ns, box=1.0, vort=1.0)
    t = W.random_p3d(nt, 200 + nt, box=1.0, vort=1.0)["coord"]
    got = cv.P3D_M2M_vort(p, t, k, sigma)
    ref = orc.P3D_M2M_vort(p, t, k, sigma)
    if k == "singular":
        assert np.all(got == 0) and np.all(ref == 0)
    else:
        assert_parity(f"vort {k}", got, ref, TOL)


def test_vort_cutoff_box_and_self_pair(orc):
    case("vort cut-off box")
    s = 0.2
    alpha = np.array([0.3, -0.5, 0.7], np.float32)
    x0 = np.array([0.1, -0.2, 0.3], np.float32)
    offs = np.array([[0, 0, 0], [4.5, 0, 0], [4, 4, 4], [5.1, 0, 0], [0, -5.2, 0]], np.float32)
    src = make_p3d(x0 + s * offs, np.tile(alpha, (5, 1)), [0.01] * 5)
    for k in ("winckelmans", "gaussian", "planetary"):
        got = cv.P3D_M2M_vort(src, x0.reshape(1, 3), k, s)
        assert_parity(f"vort box {k}", got, orc.P3D_M2M_vort(src, x0.reshape(1, 3), k, s), TOL)
        # only the first three sources are inside the box; each alone:
        for i in range(5):
            one = cv.P3D_M2M_vort(src[i:i + 1], x0.reshape(1, 3), k, s)
            assert (np.any(one != 0) or k == "planetary" and i > 0) if i < 3 else np.all(one == 0)


def test_vort_negative_radius_is_zero():
    case("vort negative sigma")
    p = W.random_p3d(300, 7, box=1.0, vort=1.0)
    assert np.all(cv.P3D_M2M_vort(p, p["coord"], "winckelmans", -0.1) == 0)


@pytest.mark.parametrize("k", ("winckelmans", "gaussian", "planetary"))
def test_vort_sorted_path_large(orc, k):
    # 24000 sources x 12000 targets takes the Morton-sorted tile-skipping path
    case(f"vort sorted {k}")
    p = W.random_p3d(24000, 301, box=1.0, vort=1.0)
    t = W.random_p3d(12000, 302, box=1.2, vort=1.0)["coord"]
    sigma = 0.02
    got = cv.P3D_M2M_vort(p, t, k, sigma)
    idx = np.random.default_rng(5).choice(len(t), 600, replace=False)
    ref = orc.P3D_M2M_vort(p, t[idx], k, sigma)
    assert_parity(f"vort sorted {k} (sampled)", got[idx], ref, TOL)
    L = cv.lib()
    prev = L.cvtx_hip_set_target_order(0)
    try:
        brute = cv.P3D_M2M_vort(p, t, k, sigma)
    finally:
        L.cvtx_hip_set_target_order(prev)
    # same pairs, different summation order
    assert_parity(f"vort sorted vs unsorted {k}", got, brute, 1e-6)


@pytest.mark.parametrize("k", ("winckelmans", "gaussian"))
def test_relaxation_matches_oracle(orc, k):
    case(f"relaxation {k}")
    p = W.random_p3d(2500, 401, box=1.0, vort=1.0, vol=0.001)
    for fdt in (0.0, 0.25, 1.0):
        got = cv.P3D_pedrizzetti_relaxation(p, fdt, k, 0.1)
        ref = orc.P3D_pedrizzetti_relaxation(p, fdt, k, 0.1)
        assert_parity(f"relax {k} fdt={fdt}", got["vorticity"], ref["vorticity"], TOL)
        assert np.array_equal(got["coord"], p["coord"]) and np.array_equal(got["volume"], p["volume"])


def test_relaxation_isolated_and_singular():
    case("relaxation invariants")
    iso = make_p3d([[0, 0, 0], [1, 0, 0], [0, 2, 0]], [[0.3, -0.2, 0.5], [1, 1, 0], [0, 0, -2]],
                   [0.01] * 3)
    r = cv.P3D_pedrizzetti_relaxation(iso, 0.3, "winckelmans", 0.1)
    assert np.abs(r["vorticity"] - iso["vorticity"]).max() < 1e-6
    r = cv.P3D_pedrizzetti_relaxation(iso, 0.3, "singular", 0.1)
    assert np.all(r["vorticity"] == 0)


def test_relaxation_sorted_path(orc):
    case("relaxation sorted path n=12000")
    p = W.random_p3d(12000, 402, box=1.0, vort=1.0, vol=1e-4)
    got = cv.P3D_pedrizzetti_relaxation(p, 0.3, "gaussian", 0.02)
    ref = orc.P3D_pedrizzetti_relaxation(p, 0.3, "gaussian", 0.02)
    assert_parity("relax sorted", got["vorticity"], ref["vorticity"], TOL)


def test_device_api_vort_and_relax():
    case("vort/relax device API")
    p = W.random_p3d(5000, 501, box=1.0, vort=1.0, vol=0.001)
    pv, a = D.p3d_soa(p)
    dpv, da = D.DeviceArray.from_host(pv), D.DeviceArray.from_host(a)
    out = D.DeviceArray((5000, 3))
    D.vort_eval(D.VortProblem(dpv.addr, da.addr, 5000, dpv.addr, 5000, out.addr,
                              oracle_reg("winckelmans"), 0.05, cv.ACC_COMP))
    host = cv.P3D_M2M_vort(p, p["coord"], "winckelmans", 0.05)
    assert_parity("vort device vs host", out.download(), host, 1e-6)
    D.relax_eval(dpv, da, 5000, 0.25, "winckelmans", 0.05)
    relaxed = da.download()[:, :3]
    host_r = cv.P3D_pedrizzetti_relaxation(p, 0.25, "winckelmans", 0.05)["vorticity"]
    assert_parity("relax device vs host", relaxed, host_r, 1e-6)


def oracle_reg(name):
    from cvortex_amd.types import REG_ID
    return REG_ID[name]
