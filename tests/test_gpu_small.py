"""Single-source (S2S, S2M) and single-target (M2S) entry points on the GPU
against the oracle's restatement of the reference CPU path (reference
src/P3D.cpp:74-296, src/F3D.cpp:34-128, src/P2D.cpp:71-231).  Tolerance:
relative L2 <= 1e-5 per call (the Gaussian stretching term, whose FP32
reference value carries the reference's own erf error, is checked with the
FP64-limited policy).  The dispatch mode DEVICE sends these small calls to
the device (in the automatic mode they run on the CPU leg, which
tests/test_host_leg.py covers); each test checks that the device ran."""
import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import workloads as W
from gpuutil import assert_parity, assert_parity_limited, case, have_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-5
REGS = ("singular", "winckelmans", "planetary", "gaussian")
VISC = ("winckelmans", "gaussian")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")
    cv.set_accumulation(cv.ACC_COMP)
    prev = cv.set_dispatch(cv.DISPATCH_DEVICE)
    yield
    cv.set_dispatch(prev)


@pytest.fixture(autouse=True)
def on_device():
    yield
    assert cv.last_leg() == cv.LEG_DEVICE, "the last call did not run on the device"


@pytest.fixture(scope="module")
def parts():
    p = W.random_p3d(300, 901, box=1.0, vort=1.0)
    t = W.random_p3d(40, 902, box=1.0, vort=1.0)
    return p, t


@pytest.mark.parametrize("k", REGS)
def test_p3d_s2s(orc, parts, k):
    case(f"P3D S2S {k}")
    p, t = parts
    s = 0.3
    for i in range(8):
        assert_parity(f"S2S vel {k}", cv.P3D_S2S_vel(p[i], t["coord"][i], k, s),
                      orc.P3D_S2S_vel(p[i], t["coord"][i], k, s), TOL)
        d = cv.P3D_S2S_dvort(p[i], t[i], k, s)
        r = orc.P3D_S2S_dvort(p[i], t[i], k, s)
        if k == "gaussian":
            _, d64, _ = orc.f64_P3D(p[i:i + 1], t[i:i + 1], k, s, gauss_mode=1)
            assert_parity_limited(f"S2S dvort {k}", d, r, d64[0], TOL)
        else:
            assert_parity(f"S2S dvort {k}", d, r, TOL)
        if k in VISC:
            assert_parity(f"S2S visc {k}", cv.P3D_S2S_visc_dvort(p[i], t[i], k, s, 0.1),
                          orc.P3D_S2S_visc_dvort(p[i], t[i], k, s, 0.1), TOL)
        got = cv.P3D_S2S_vort(p[i], t["coord"][i], k, s)
        ref = orc.P3D_S2S_vort(p[i], t["coord"][i], k, s)
        if np.any(ref != 0):
            assert_parity(f"S2S vort {k}", got, ref, TOL)
        else:
            assert np.all(got == 0)


def test_p3d_s2s_vort_has_no_cutoff(orc):
    # the single-source vorticity forms keep sources beyond 5 sigma and use
    # the signed radius (reference src/P3D.cpp:146-161)
    case("S2S vort no cut-off")
    from cvortex_amd.types import make_p3d
    srcs = make_p3d([[0.55, 0.0, 0.0]], [[0.3, -0.5, 0.7]], [0.01])
    src = srcs[0]
    for s in (0.1, -0.1):
        got = cv.P3D_S2S_vort(src, np.zeros(3, np.float32), "winckelmans", s)
        ref = orc.P3D_S2S_vort(src, np.zeros(3, np.float32), "winckelmans", s)
        assert np.any(ref != 0)
        assert_parity(f"S2S vort beyond 5 sigma (sigma={s})", got, ref, TOL)
    # while M2S cuts at the box
    assert np.all(cv.P3D_M2S_vort(srcs, np.zeros(3, np.float32), "winckelmans", 0.1) == 0)


@pytest.mark.parametrize("k", REGS)
def test_p3d_s2m_and_m2s(orc, parts, k):
    case(f"P3D S2M/M2S {k}")
    p, t = parts
    s = 0.3
    src = p[:1]
    assert_parity(f"S2M vel {k}", cv.P3D_S2M_vel(p[0], t["coord"], k, s),
                  orc.P3D_M2M_vel(src, t["coord"], k, s), TOL)
    if k != "gaussian":
        assert_parity(f"S2M dvort {k}", cv.P3D_S2M_dvort(p[0], t, k, s),
                      orc.P3D_M2M_dvort(src, t, k, s), TOL)
    if k in VISC:
        assert_parity(f"S2M visc {k}", cv.P3D_S2M_visc_dvort(p[0], t, k, s, 0.1),
                      orc.P3D_M2M_visc_dvort(src, t, k, s, 0.1), TOL)
    ref = np.array([orc.P3D_S2S_vort(p[0], x, k, s) for x in t["coord"]])
    got = cv.P3D_S2M_vort(p[0], t["coord"], k, s)
    if np.any(ref != 0):
        assert_parity(f"S2M vort {k}", got, ref, TOL)
    for i in range(4):
        assert_parity(f"M2S vel {k}", cv.P3D_M2S_vel(p, t["coord"][i], k, s),
                      orc.P3D_M2M_vel(p, t["coord"][i:i + 1], k, s)[0], TOL)
        d = cv.P3D_M2S_dvort(p, t[i], k, s)
        r = orc.P3D_M2M_dvort(p, t[i:i + 1], k, s)[0]
        if k == "gaussian":
            _, d64, _ = orc.f64_P3D(p, t[i:i + 1], k, s, gauss_mode=1)
            assert_parity_limited(f"M2S dvort {k}", d, r, d64[0], TOL)
        else:
            assert_parity(f"M2S dvort {k}", d, r, TOL)
        if k in VISC:
            assert_parity(f"M2S visc {k}", cv.P3D_M2S_visc_dvort(p, t[i], k, s, 0.1),
                          orc.P3D_M2M_visc_dvort(p, t[i:i + 1], k, s, 0.1)[0], TOL)
        w = orc.P3D_M2M_vort(p, t["coord"][i:i + 1], k, s)[0]
        g = cv.P3D_M2S_vort(p, t["coord"][i], k, s)
        if np.any(w != 0):
            assert_parity(f"M2S vort {k}", g, w, TOL)


def test_f3d_small(orc):
    case("F3D S2S/M2S")
    f = W.random_f3d(200, 903, box=1.0, strength=1.0)
    t = W.random_p3d(6, 904, box=1.0, vort=1.0)
    for i in range(6):
        assert_parity("F3D S2S vel", cv.F3D_S2S_vel(f[i], t["coord"][i]),
                      orc.F3D_S2S_vel(f[i], t["coord"][i]), TOL)
        _, d64 = orc.f64_F3D(f[i:i + 1], t[i:i + 1], vel=False)
        assert_parity_limited("F3D S2S dvort", cv.F3D_S2S_dvort(f[i], t[i]),
                              orc.F3D_S2S_dvort(f[i], t[i]), d64[0], TOL)
        assert_parity("F3D M2S vel", cv.F3D_M2S_vel(f, t["coord"][i]),
                      orc.F3D_M2M_vel(f, t["coord"][i:i + 1])[0], TOL)
        _, d64 = orc.f64_F3D(f, t[i:i + 1], vel=False)
        assert_parity_limited("F3D M2S dvort", cv.F3D_M2S_dvort(f, t[i]),
                              orc.F3D_M2M_dvort(f, t[i:i + 1])[0], d64[0], TOL)


@pytest.mark.parametrize("k", REGS)
def test_p2d_small(orc, k):
    case(f"P2D S2S/S2M/M2S {k}")
    q = W.random_p2d(250, 905, box=1.0, vort=1.0)
    r = W.random_p2d(30, 906, box=1.0, vort=1.0)
    s = 0.2
    for i in range(5):
        assert_parity(f"P2D S2S vel {k}", cv.P2D_S2S_vel(q[i], r["coord"][i], k, s),
                      orc.P2D_S2S_vel(q[i], r["coord"][i], k, s), TOL)
        assert_parity(f"P2D M2S vel {k}", cv.P2D_M2S_vel(q, r["coord"][i], k, s),
                      orc.P2D_M2M_vel(q, r["coord"][i:i + 1], k, s)[0], TOL)
        if k in VISC:
            assert_parity(f"P2D S2S visc {k}", [cv.P2D_S2S_visc_dvort(q[i], r[i], k, s, 0.1)],
                          [orc.P2D_S2S_visc_dvort(q[i], r[i], k, s, 0.1)], TOL)
            assert_parity(f"P2D M2S visc {k}", [cv.P2D_M2S_visc_dvort(q, r[i], k, s, 0.1)],
                          orc.P2D_M2M_visc_dvort(q, r[i:i + 1], k, s, 0.1), TOL)
    assert_parity(f"P2D S2M vel {k}", cv.P2D_S2M_vel(q[0], r["coord"], k, s),
                  orc.P2D_M2M_vel(q[:1], r["coord"], k, s), TOL)
    if k in VISC:
        assert_parity(f"P2D S2M visc {k}", cv.P2D_S2M_visc_dvort(q[0], r, k, s, 0.1),
                      orc.P2D_M2M_visc_dvort(q[:1], r, k, s, 0.1), TOL)
