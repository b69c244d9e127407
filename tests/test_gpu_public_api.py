"""The public void cvtx_* functions on the GPU, called as CVortex.jl calls
them (reference include/cvortex/libcvtx.h:213-370): in the automatic
dispatch mode a problem that passes the reference's size guards runs on the
device, one below them (the N < 256 forcing hook, reference
src/P3D.cpp:353-355) or with a kernel name the device does not know runs on
the CPU leg.  Every result is checked against the oracle's restatement of
the reference CPU path at relative L2 <= 1e-5 (the Gaussian stretching
against FP64, the reference's own FP32 erf being the limit)."""
import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import capi
from cvortex_amd import workloads as W
from gpuutil import assert_parity, assert_parity_limited, case, have_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-5
REGS = ("singular", "winckelmans", "planetary", "gaussian")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")
    cv.set_accumulation(cv.ACC_COMP)
    prev = cv.set_dispatch(cv.DISPATCH_AUTO)
    yield
    cv.set_dispatch(prev)


def leg(want):
    assert cv.last_leg() == want, f"leg {cv.last_leg()}, expected {want}"


@pytest.mark.parametrize("k", REGS)
def test_p3d_public_on_device(orc, k):
    p = W.random_p3d(2000, 1301, box=1.0, vort=1.0)
    t = W.random_p3d(700, 1302, box=1.0, vort=1.0)
    s = 0.1
    case(f"public P3D M2M {k}: 2000 sources x 700 targets")
    assert_parity(f"vel {k}", capi.P3D_M2M_vel(p, t["coord"], k, s),
                  orc.P3D_M2M_vel(p, t["coord"], k, s), TOL)
    leg(cv.LEG_DEVICE)
    d = capi.P3D_M2M_dvort(p, t, k, s)
    leg(cv.LEG_DEVICE)
    r = orc.P3D_M2M_dvort(p, t, k, s)
    if k == "gaussian":
        _, d64, _ = orc.f64_P3D(p, t, k, s, gauss_mode=1)
        assert_parity_limited(f"dvort {k}", d, r, d64, TOL)
    else:
        assert_parity(f"dvort {k}", d, r, TOL)
    if k in ("winckelmans", "gaussian"):
        assert_parity(f"visc {k}", capi.P3D_M2M_visc_dvort(p, t, k, s, 0.01),
                      orc.P3D_M2M_visc_dvort(p, t, k, s, 0.01), TOL)
        leg(cv.LEG_DEVICE)
    assert_parity(f"vort {k}", capi.P3D_M2M_vort(p, t["coord"], k, s),
                  orc.P3D_M2M_vort(p, t["coord"], k, s), TOL)
    leg(cv.LEG_DEVICE)


def test_relaxation_public_on_device(orc):
    p = W.random_p3d(1500, 1303, box=1.0, vort=1.0)
    case("public relaxation: 1500 particles")
    got = capi.P3D_pedrizzetti_relaxation(p, 0.2, "gaussian", 0.1)
    leg(cv.LEG_DEVICE)
    assert_parity("relaxation", got["vorticity"],
                  orc.P3D_pedrizzetti_relaxation(p, 0.2, "gaussian", 0.1)["vorticity"], TOL)


def test_f3d_public_on_device(orc):
    f = W.random_f3d(600, 1304, box=1.0, strength=1.0)
    t = W.random_p3d(900, 1305, box=1.0, vort=1.0)
    case("public F3D M2M: 600 filaments x 900 targets")
    assert_parity("F3D vel", capi.F3D_M2M_vel(f, t["coord"]), orc.F3D_M2M_vel(f, t["coord"]), TOL)
    leg(cv.LEG_DEVICE)
    _, d64 = orc.f64_F3D(f, t, vel=False)
    assert_parity_limited("F3D dvort", capi.F3D_M2M_dvort(f, t), orc.F3D_M2M_dvort(f, t), d64, TOL)
    leg(cv.LEG_DEVICE)
    dirs = np.random.default_rng(9).normal(size=(900, 3)).astype(np.float32)
    assert_parity("F3D influence matrix", capi.F3D_inf_mtrx(f, t["coord"], dirs),
                  orc.F3D_inf_mtrx(f, t["coord"], dirs), TOL)
    leg(cv.LEG_DEVICE)


@pytest.mark.parametrize("k", REGS)
def test_p2d_public_on_device(orc, k):
    q = W.random_p2d(1800, 1306, box=1.0, vort=1.0)
    r = W.random_p2d(600, 1307, box=1.0, vort=1.0)
    case(f"public P2D M2M {k}: 1800 x 600")
    assert_parity(f"P2D vel {k}", capi.P2D_M2M_vel(q, r["coord"], k, 0.05),
                  orc.P2D_M2M_vel(q, r["coord"], k, 0.05), TOL)
    leg(cv.LEG_DEVICE)
    if k in ("winckelmans", "gaussian"):
        assert_parity(f"P2D visc {k}", capi.P2D_M2M_visc_dvort(q, r, k, 0.05, 0.01),
                      orc.P2D_M2M_visc_dvort(q, r, k, 0.05, 0.01), TOL)
        leg(cv.LEG_DEVICE)


def test_size_guard_sends_small_problems_to_the_host(orc):
    # fewer than 256 targets (or sources): the reference's CPU path
    p = W.random_p3d(2000, 1308, box=1.0, vort=1.0)
    t = W.random_p3d(255, 1309, box=1.0, vort=1.0)
    case("public P3D vel with 255 targets: host leg")
    assert_parity("vel 255 targets", capi.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.1),
                  orc.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.1), TOL)
    leg(cv.LEG_HOST)
    assert_parity("dvort 255 sources", capi.P3D_M2M_dvort(t, p, "winckelmans", 0.1),
                  orc.P3D_M2M_dvort(t, p, "winckelmans", 0.1), TOL)
    leg(cv.LEG_HOST)
    # at 256 the device takes it
    t = W.random_p3d(256, 1310, box=1.0, vort=1.0)
    assert_parity("vel 256 targets", capi.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.1),
                  orc.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.1), TOL)
    leg(cv.LEG_DEVICE)
    # DEVICE mode lifts the guard
    prev = cv.set_dispatch(cv.DISPATCH_DEVICE)
    try:
        t = W.random_p3d(17, 1311, box=1.0, vort=1.0)
        assert_parity("vel 17 targets on device",
                      capi.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.1),
                      orc.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.1), TOL)
        leg(cv.LEG_DEVICE)
    finally:
        cv.set_dispatch(prev)


def test_custom_name_runs_on_the_host(orc):
    p = W.random_p3d(1000, 1312, box=1.0, vort=1.0)
    vf = cv.vortfunc("gaussian")
    vf.cl_kernel_name_ext = b""
    case("public P3D vel with an empty kernel name: host leg")
    assert_parity("custom vel", capi.P3D_M2M_vel(p, p["coord"], vf, 0.1),
                  orc.P3D_M2M_vel(p, p["coord"], "gaussian", 0.1), TOL)
    leg(cv.LEG_HOST)


def test_disabled_device_runs_on_the_host(orc):
    p = W.random_p3d(1000, 1313, box=1.0, vort=1.0)
    cv.accelerator_disable(0)
    try:
        case("public P3D vel with the device disabled: host leg")
        assert_parity("vel, device disabled", capi.P3D_M2M_vel(p, p["coord"], "winckelmans", 0.1),
                      orc.P3D_M2M_vel(p, p["coord"], "winckelmans", 0.1), TOL)
        leg(cv.LEG_HOST)
    finally:
        cv.accelerator_enable(0)
    capi.P3D_M2M_vel(p, p["coord"], "winckelmans", 0.1)
    leg(cv.LEG_DEVICE)
