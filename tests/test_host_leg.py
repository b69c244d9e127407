"""The library's CPU leg: the public void cvtx_* functions evaluated on the
host, as the reference does when its accelerator declines a call
(reference src/P3D.cpp:343-498, src/F3D.cpp:162-227, src/P2D.cpp:141-276,
src/accelerators.cpp:107-118).

Every call here runs with the dispatch mode HOST, so these tests run the
same code with or without a GPU; the leg is checked with
cvtx_hip_last_leg().  Results are compared with the oracle's restatement of
the reference CPU path at relative L2 <= 1e-5.  The Gaussian stretching,
whose FP32 reference value carries the reference's own erf error, is held
to the FP64 oracle at 1e-5 and to the restatement at 1.5x its own distance
from FP64 (the policy of the GPU parity tests)."""
import ctypes

import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import capi
from cvortex_amd import workloads as W
from cvortex_amd.types import VortFunc, make_p3d
from gpuutil import assert_parity, assert_parity_limited

TOL = 1e-5
REGS = ("singular", "winckelmans", "planetary", "gaussian")
VISC = ("winckelmans", "gaussian")


@pytest.fixture(autouse=True)
def host_dispatch():
    cv.initialise()
    prev = cv.set_dispatch(cv.DISPATCH_HOST)
    yield
    cv.set_dispatch(prev)


def on_host():
    assert cv.last_leg() == cv.LEG_HOST


@pytest.fixture(scope="module")
def parts():
    p = W.random_p3d(700, 1201, box=1.0, vort=1.0)
    t = W.random_p3d(300, 1202, box=1.0, vort=1.0)
    return p, t


@pytest.mark.parametrize("k", REGS)
def test_p3d_m2m(orc, parts, k):
    p, t = parts
    s = 0.2
    assert_parity(f"vel {k}", capi.P3D_M2M_vel(p, t["coord"], k, s),
                  orc.P3D_M2M_vel(p, t["coord"], k, s), TOL)
    on_host()
    d = capi.P3D_M2M_dvort(p, t, k, s)
    on_host()
    r = orc.P3D_M2M_dvort(p, t, k, s)
    if k == "gaussian":
        _, d64, _ = orc.f64_P3D(p, t, k, s, gauss_mode=1)
        assert_parity_limited(f"dvort {k}", d, r, d64, TOL)
    else:
        assert_parity(f"dvort {k}", d, r, TOL)
    if k in VISC:
        assert_parity(f"visc {k}", capi.P3D_M2M_visc_dvort(p, t, k, s, 0.05),
                      orc.P3D_M2M_visc_dvort(p, t, k, s, 0.05), TOL)
        on_host()
    else:
        assert np.all(capi.P3D_M2M_visc_dvort(p, t, k, s, 0.05) == 0)
    assert_parity(f"vort {k}", capi.P3D_M2M_vort(p, t["coord"], k, s),
                  orc.P3D_M2M_vort(p, t["coord"], k, s), TOL)
    on_host()


def test_p3d_sources_alias_targets(orc, parts):
    # the reference's own agreement test passes one array as both
    # (test/testsamecpugpuresultmany.h:110); self-pairs contribute nothing
    p, _ = parts
    assert_parity("dvort aliased", capi.P3D_M2M_dvort(p, p, "winckelmans", 0.2),
                  orc.P3D_M2M_dvort(p, p, "winckelmans", 0.2), TOL)
    assert_parity("vel at own coordinates", capi.P3D_M2M_vel(p, p["coord"], "singular", 0.2),
                  orc.P3D_M2M_vel(p, p["coord"], "singular", 0.2), TOL)


def test_negative_sigma_and_empty(orc, parts):
    p, t = parts
    d = capi.P3D_M2M_dvort(p, t, "winckelmans", -0.2)
    assert_parity("dvort, negative sigma", d, orc.P3D_M2M_dvort(p, t, "winckelmans", -0.2), TOL)
    assert np.array_equal(d, -capi.P3D_M2M_dvort(p, t, "winckelmans", 0.2))
    # no sources: zero results; no targets: nothing written
    assert np.all(capi.P3D_M2M_vel(p[:0], t["coord"], "winckelmans", 0.2) == 0)
    assert capi.P3D_M2M_vel(p, t["coord"][:0], "winckelmans", 0.2).shape == (0, 3)


def test_relaxation(orc, parts):
    p, _ = parts
    for k in ("winckelmans", "gaussian"):
        got = capi.P3D_pedrizzetti_relaxation(p, 0.3, k, 0.2)
        on_host()
        ref = orc.P3D_pedrizzetti_relaxation(p, 0.3, k, 0.2)
        assert_parity(f"relaxation {k}", got["vorticity"], ref["vorticity"], TOL)
        assert np.array_equal(got["coord"], p["coord"])


def test_f3d(orc):
    f = W.random_f3d(400, 1203, box=1.0, strength=1.0)
    t = W.random_p3d(350, 1204, box=1.0, vort=1.0)
    assert_parity("F3D vel", capi.F3D_M2M_vel(f, t["coord"]), orc.F3D_M2M_vel(f, t["coord"]), TOL)
    on_host()
    _, d64 = orc.f64_F3D(f, t, vel=False)
    assert_parity_limited("F3D dvort", capi.F3D_M2M_dvort(f, t), orc.F3D_M2M_dvort(f, t), d64, TOL)
    on_host()
    dirs = np.random.default_rng(5).normal(size=(350, 3)).astype(np.float32)
    assert_parity("F3D influence matrix", capi.F3D_inf_mtrx(f, t["coord"], dirs),
                  orc.F3D_inf_mtrx(f, t["coord"], dirs), TOL)
    on_host()


def test_f3d_on_the_line_is_zero():
    # a point on a filament's line (its end point, its extension) gets no
    # contribution from it, as in the reference's zeroing of non-finite terms
    from cvortex_amd.types import make_f3d
    f = make_f3d([[0.0, 0.0, 0.0]], [[1.0, 0.0, 0.0]], [1.0])
    pts = np.array([[1.0, 0.0, 0.0], [2.0, 0.0, 0.0], [0.5, 0.0, 0.0]], np.float32)
    assert np.all(capi.F3D_M2M_vel(f, pts) == 0)
    ind = make_p3d(pts, np.ones((3, 3)), np.ones(3))
    assert np.all(np.isfinite(capi.F3D_M2M_dvort(f, ind)))


@pytest.mark.parametrize("k", REGS)
def test_p2d(orc, k):
    q = W.random_p2d(600, 1205, box=1.0, vort=1.0)
    r = W.random_p2d(280, 1206, box=1.0, vort=1.0)
    assert_parity(f"P2D vel {k}", capi.P2D_M2M_vel(q, r["coord"], k, 0.1),
                  orc.P2D_M2M_vel(q, r["coord"], k, 0.1), TOL)
    on_host()
    if k in VISC:
        assert_parity(f"P2D visc {k}", capi.P2D_M2M_visc_dvort(q, r, k, 0.1, 0.02),
                      orc.P2D_M2M_visc_dvort(q, r, k, 0.1, 0.02), TOL)
        on_host()


def test_threaded_by_source_chunks(orc):
    # few targets and many sources: the sum is split over source chunks
    p = W.random_p3d(20000, 1207, box=1.0, vort=1.0)
    t = W.random_p3d(3, 1208, box=1.0, vort=1.0)
    a = capi.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.05)
    assert_parity("3 targets x 20000 sources", a, orc.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.05),
                  TOL)
    # one target per call (the M2S form) gives the same values
    b = np.array([cv.P3D_M2S_vel(p, x, "winckelmans", 0.05) for x in t["coord"]])
    assert np.array_equal(a, b)


def _custom(name, fields):
    """A cvtx_VortFunc with the given kernel name whose function pointers are
    the named fields of the library's own regularisation `fields`."""
    vf = cv.vortfunc(fields)
    vf.cl_kernel_name_ext = name
    return vf


@pytest.mark.parametrize("mode", ("host", "auto"))
def test_custom_regularisation_runs_its_function_pointers(orc, parts, mode):
    # An empty or unknown cl_kernel_name_ext has no device kernel: the call
    # runs through the function pointers (include/cvortex/libcvtx.h:74-94),
    # in every dispatch mode.
    if mode == "auto":
        cv.set_dispatch(cv.DISPATCH_AUTO)
    p, t = parts
    for name in (b"", b"my_kernel"):
        vf = _custom(name, "winckelmans")
        assert_parity("custom vel", capi.P3D_M2M_vel(p, t["coord"], vf, 0.2),
                      orc.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.2), TOL)
        on_host()
        assert_parity("custom dvort", capi.P3D_M2M_dvort(p, t, vf, 0.2),
                      orc.P3D_M2M_dvort(p, t, "winckelmans", 0.2), TOL)
        assert_parity("custom visc", capi.P3D_M2M_visc_dvort(p, t, vf, 0.2, 0.1),
                      orc.P3D_M2M_visc_dvort(p, t, "winckelmans", 0.2, 0.1), TOL)
        on_host()


def test_python_regularisation(orc):
    # A caller-written regularisation: here the oracle's FP32 restatement of
    # the reference's Winckelmans functions behind ctypes callbacks, so the
    # host leg must reproduce the oracle's sums from the caller's pointers.
    f1 = ctypes.CFUNCTYPE(ctypes.c_float, ctypes.c_float)
    f3 = ctypes.CFUNCTYPE(None, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                          ctypes.POINTER(ctypes.c_float))
    g3 = f1(lambda r: orc.vf("winckelmans", "g3", r))
    z3 = f1(lambda r: orc.vf("winckelmans", "zeta3", r))
    e3 = f1(lambda r: orc.vf("winckelmans", "eta3", r))
    g2 = f1(lambda r: orc.vf("winckelmans", "g2", r))
    e2 = f1(lambda r: orc.vf("winckelmans", "eta2", r))

    def comb(r, g, z):
        g[0] = orc.vf("winckelmans", "g3", r)
        z[0] = orc.vf("winckelmans", "zeta3", r)
    c3 = f3(comb)
    vf = VortFunc()
    for fld, fn in (("g_3D", g3), ("g_2D", g2), ("zeta_3D", z3), ("combined_3D", c3),
                    ("eta_3D", e3), ("eta_2D", e2)):
        setattr(vf, fld, ctypes.cast(fn, type(getattr(vf, fld))))
    vf.cl_kernel_name_ext = b""
    p = W.random_p3d(90, 1209, box=1.0, vort=1.0)
    t = W.random_p3d(60, 1210, box=1.0, vort=1.0)
    assert_parity("python g_3D", capi.P3D_M2M_vel(p, t["coord"], vf, 0.2),
                  orc.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.2), 2e-6)
    assert_parity("python combined_3D", capi.P3D_M2M_dvort(p, t, vf, 0.2),
                  orc.P3D_M2M_dvort(p, t, "winckelmans", 0.2), 2e-6)
    assert_parity("python eta_3D", capi.P3D_M2M_visc_dvort(p, t, vf, 0.2, 0.1),
                  orc.P3D_M2M_visc_dvort(p, t, "winckelmans", 0.2, 0.1), 2e-6)
    q = W.random_p2d(80, 1211, box=1.0, vort=1.0)
    assert_parity("python g_2D", capi.P2D_M2M_vel(q, q["coord"], vf, 0.1),
                  orc.P2D_M2M_vel(q, q["coord"], "winckelmans", 0.1), 2e-6)
    assert_parity("python eta_2D", capi.P2D_M2M_visc_dvort(q, q, vf, 0.1, 0.1),
                  orc.P2D_M2M_visc_dvort(q, q, "winckelmans", 0.1, 0.1), 2e-6)


def test_missing_function_pointer_gives_nan():
    # the reference asserts on a missing eta; here: NaN results and a message
    vf = cv.vortfunc("winckelmans")
    vf.cl_kernel_name_ext = b""
    vf.eta_3D = type(vf.eta_3D)()
    p = W.random_p3d(10, 1212, box=1.0, vort=1.0)
    out = capi.P3D_M2M_visc_dvort(p, p, vf, 0.2, 0.1)
    assert np.isnan(out).all()
    assert cv.last_leg() == -1
    assert "eta_3D" in cv.last_error()


def test_dispatch_controls():
    assert cv.get_dispatch() == cv.DISPATCH_HOST
    assert cv.set_dispatch(cv.DISPATCH_DEVICE) == cv.DISPATCH_HOST
    assert cv.set_dispatch(cv.DISPATCH_AUTO) == cv.DISPATCH_DEVICE
    with pytest.raises(cv.CvtxError):
        cv.set_dispatch(7)
    assert cv.host_threads() >= 1


def test_no_enabled_device_runs_on_host(orc, parts):
    # with every accelerator disabled (or none present) the public functions
    # run the CPU leg in any dispatch mode, while the seam reports -1
    cv.set_dispatch(cv.DISPATCH_DEVICE)
    n = cv.num_accelerators()
    states = [cv.accelerator_enabled(i) for i in range(n)]
    for i in range(n):
        cv.accelerator_disable(i)
    try:
        p, t = parts
        assert_parity("vel, no device", capi.P3D_M2M_vel(p, t["coord"], "gaussian", 0.2),
                      orc.P3D_M2M_vel(p, t["coord"], "gaussian", 0.2), TOL)
        on_host()
        with pytest.raises(cv.CvtxError):
            cv.P3D_M2M_vel(p, t["coord"], "gaussian", 0.2)
    finally:
        for i, st in enumerate(states):
            (cv.accelerator_enable if st else cv.accelerator_disable)(i)


@pytest.mark.parametrize("k", REGS)
def test_single_source_and_target_forms(orc, parts, k):
    # S2S, S2M and M2S (reference src/P3D.cpp:74-296, src/P2D.cpp:71-231)
    p, t = parts
    s = 0.3
    for i in range(3):
        assert_parity(f"S2S vel {k}", cv.P3D_S2S_vel(p[i], t["coord"][i], k, s),
                      orc.P3D_S2S_vel(p[i], t["coord"][i], k, s), TOL)
        on_host()
        if k != "gaussian":
            assert_parity(f"S2S dvort {k}", cv.P3D_S2S_dvort(p[i], t[i], k, s),
                          orc.P3D_S2S_dvort(p[i], t[i], k, s), TOL)
        if k in VISC:
            assert_parity(f"S2S visc {k}", cv.P3D_S2S_visc_dvort(p[i], t[i], k, s, 0.1),
                          orc.P3D_S2S_visc_dvort(p[i], t[i], k, s, 0.1), TOL)
        ref = orc.P3D_S2S_vort(p[i], t["coord"][i], k, s)
        got = cv.P3D_S2S_vort(p[i], t["coord"][i], k, s)
        if np.any(ref != 0):
            assert_parity(f"S2S vort {k}", got, ref, TOL)
        assert_parity(f"M2S vel {k}", cv.P3D_M2S_vel(p, t["coord"][i], k, s),
                      orc.P3D_M2M_vel(p, t["coord"][i:i + 1], k, s)[0], TOL)
        on_host()
    assert_parity(f"S2M vel {k}", cv.P3D_S2M_vel(p[0], t["coord"], k, s),
                  orc.P3D_M2M_vel(p[:1], t["coord"], k, s), TOL)
    on_host()
    q = W.random_p2d(200, 1213, box=1.0, vort=1.0)
    assert_parity(f"P2D S2M vel {k}", cv.P2D_S2M_vel(q[0], q["coord"], k, 0.2),
                  orc.P2D_M2M_vel(q[:1], q["coord"], k, 0.2), TOL)
    on_host()
    assert_parity(f"P2D S2S vel {k}", cv.P2D_S2S_vel(q[3], q["coord"][5], k, 0.2),
                  orc.P2D_S2S_vel(q[3], q["coord"][5], k, 0.2), TOL)


def test_small_forms_stay_on_host_below_the_device_minimum(parts):
    # in the automatic mode a single pair, or an S2M/M2S call below
    # SMALL_DEVICE_MIN elements, never goes to the device
    cv.set_dispatch(cv.DISPATCH_AUTO)
    p, t = parts
    cv.P3D_S2S_vel(p[0], t["coord"][0], "winckelmans", 0.2)
    on_host()
    cv.P3D_M2S_vel(p, t["coord"][0], "winckelmans", 0.2)
    on_host()
    cv.F3D_S2S_vel(W.random_f3d(1, 3, box=1.0, strength=1.0)[0], t["coord"][0])
    on_host()
