"""libcvortex.so: loads, exports every symbol the headers declare, has the
reference's struct layouts, and behaves like the reference's controls.
These tests make no device compute call and run without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd.types import P3D_DTYPE, REG_NAMES, VortFunc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for dirpath, _, files in os.walk(os.path.join(ROOT, "include")):
        for fn in files:
            if not fn.endswith(".h"):
                continue
            text = open(os.path.join(dirpath, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            for m in re.finditer(r"\b(cvtx_\w+)\s*\(", text):
                names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def L():
    return cv.lib()


def test_library_is_in_tree_and_loads(L):
    assert os.path.exists(cv.LIB_PATH)
    assert isinstance(L, ctypes.CDLL)


def test_every_declared_symbol_is_exported(L):
    names = declared_symbols()
    assert len(names) >= 40
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_only_cvtx_symbols_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", cv.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    text_syms = [l.split()[-1] for l in out.splitlines() if " T " in l]
    assert text_syms and all(s.startswith("cvtx_") for s in text_syms), text_syms


def test_struct_layouts():
    # reference include/cvortex/libcvtx.h:53-94 on LP64
    assert P3D_DTYPE.itemsize == 28
    assert ctypes.sizeof(VortFunc) == 80


def test_vortfunc_factories(L):
    for name in REG_NAMES:
        vf = cv.vortfunc(name)
        assert vf.cl_kernel_name_ext.decode() == name
        for fld in ("g_3D", "g_2D", "zeta_3D", "combined_3D", "eta_3D", "eta_2D"):
            assert ctypes.cast(getattr(vf, fld), ctypes.c_void_p).value


def test_vortfunc_known_answers(L):
    # the reference's known-answer tests, test/testvortfunc.h:36-67, run
    # against the library's own scalar functions
    s, w, p, g = (cv.vortfunc(n) for n in REG_NAMES)
    assert s.g_3D(0.1) == 1.0 and s.g_3D(3.0) == 1.0
    assert s.zeta_3D(0.1) == 0.0 and s.zeta_3D(10.0) == 0.0
    assert w.g_3D(0.0) == 0.0
    assert abs(w.g_3D(1.0) - 0.61872) < 1e-5
    assert abs(w.g_3D(10.0) - 0.9998168) < 1e-7
    assert w.zeta_3D(0.0) == 7.5
    assert abs(w.zeta_3D(10.0) - 7.2433e-7) < 1e-11
    assert p.g_3D(0.0) == 0.0 and p.g_3D(1.0) == 1.0 and p.g_3D(10.0) == 1.0
    assert p.zeta_3D(0.99) == 3.0 and p.zeta_3D(1.01) == 0.0
    assert g.g_3D(0.0) == 0.0
    for rho, want in ((0.5, 0.030859595), (1.0, 0.198748043), (2.0, 0.738535870),
                      (4.0, 0.998866015), (6.0, 0.999999925), (8.0, 0.999999999)):
        assert abs(g.g_3D(rho) - want) < 1e-6
    assert g.g_3D(10.0) == 1.0
    assert abs(g.zeta_3D(1.0) - 0.483941449) < 1e-6
    assert abs(g.zeta_3D(0.5) - 0.70413065) < 1e-6
    gg, zz = ctypes.c_float(), ctypes.c_float()
    w.combined_3D(0.7, ctypes.byref(gg), ctypes.byref(zz))
    assert gg.value == w.g_3D(0.7) and zz.value == w.zeta_3D(0.7)


def test_vortfunc_matches_oracle_fp64(orc):
    # the library's scalar functions equal the FP64 oracle's exact kernels
    for name in REG_NAMES:
        vf = cv.vortfunc(name)
        for rho in (0.0, 1e-3, 0.05, 0.3, 0.9, 1.1, 2.5, 7.0):
            for fn, fld in (("g3", "g_3D"), ("zeta3", "zeta_3D"), ("g2", "g_2D")):
                want = orc.vf64(name, fn, np.float32(rho), 1)
                got = getattr(vf, fld)(rho)
                assert abs(got - want) <= 2e-7 * abs(want) + 1e-30, (name, fn, rho, got, want)
            if name in ("winckelmans", "gaussian"):
                for fn, fld in (("eta3", "eta_3D"), ("eta2", "eta_2D")):
                    want = orc.vf64(name, fn, np.float32(rho), 1)
                    got = getattr(vf, fld)(rho)
                    assert abs(got - want) <= 2e-7 * abs(want), (name, fn, rho)


def test_accelerator_controls_are_consistent(L):
    # mirrors test/testaccelerators.h:33-66 (with its inverted expectation
    # corrected: enabling a device leaves it enabled)
    cv.initialise()
    n = cv.num_accelerators()
    assert 0 <= n < 64
    for i in range(n):
        assert cv.accelerator_name(i)
        st = cv.accelerator_enabled(i)
        assert st in (0, 1)
        cv.accelerator_disable(i)
        assert cv.accelerator_enabled(i) == 0
        cv.accelerator_enable(i)
        assert cv.accelerator_enabled(i) == 1
        if not st:
            cv.accelerator_disable(i)
    assert cv.accelerator_name(n) is None
    assert cv.accelerator_enabled(-1) == 0
    assert "cvortex" in cv.information()
    cv.initialise()  # repeated initialisation is harmless


def test_unknown_regularisation_is_an_error(L):
    # A regularisation whose name the library does not know has no device
    # kernel: the seam reports -1 (the reference would run its CPU path).
    vf = cv.vortfunc("winckelmans")
    vf.cl_kernel_name_ext = b"custom"
    p = np.zeros(4, P3D_DTYPE)
    with pytest.raises(cv.CvtxError, match="regularisation"):
        cv.P3D_M2M_vel(p, np.zeros((4, 3), np.float32), vf, 0.1)


def test_public_api_invalid_arguments_report_nan(L):
    # The public void API has no error return: invalid arguments (here a
    # null regularisation) give NaN results and a line on stderr, on either
    # leg, while the seam reports -1.  (With no device enabled the public
    # functions compute on the CPU leg: tests/test_host_leg.py.)
    p = np.zeros(3, P3D_DTYPE)
    p["coord"] = np.arange(9, dtype=np.float32).reshape(3, 3)
    from cvortex_amd.types import pointer_array
    ptrs, keep = pointer_array(p)
    out = np.zeros((3, 3), np.float32)
    mes = np.ascontiguousarray(p["coord"])
    for mode in (cv.DISPATCH_HOST, cv.DISPATCH_DEVICE):
        prev = cv.set_dispatch(mode)
        try:
            out[:] = 0
            L.cvtx_P3D_M2M_vel(ctypes.c_void_p(ptrs.ctypes.data), 3, ctypes.c_void_p(mes.ctypes.data),
                               3, ctypes.c_void_p(out.ctypes.data), None, ctypes.c_float(0.1))
            assert np.isnan(out).all()
            assert cv.last_leg() == -1
        finally:
            cv.set_dispatch(prev)
    assert L.cvtx_hip_P3D_M2M_vel(ctypes.c_void_p(ptrs.ctypes.data), 3,
                                  ctypes.c_void_p(mes.ctypes.data), 3,
                                  ctypes.c_void_p(out.ctypes.data), None, ctypes.c_float(0.1)) == -1


def test_redistribution_functions_match_oracle(L, orc):
    # the plugin's host functions (reference src/RedistFunc.cpp:36-96)
    fn_t = ctypes.CFUNCTYPE(ctypes.c_float, ctypes.c_float)
    radii = {"lambda0": 0.5, "lambda1": 1.0, "lambda2": 1.5, "lambda3": 2.0, "m4p": 2.0}
    for k, name in enumerate(cv.REDIST_NAMES):
        r = cv.redistfunc(name)
        assert r.radius == radii[name]
        f = fn_t(r.func)
        for U in np.linspace(0.0, 2.6, 105).astype(np.float32):
            assert f(float(U)) == orc.redist(k, float(U)), (name, U)


def test_foreign_redistributor_is_rejected(L):
    # a cvtx_RedistFunc around a function the library does not own has no
    # device kernel: -1 and a message (the reference would evaluate it)
    from cvortex_amd.types import RedistFunc
    fn_t = ctypes.CFUNCTYPE(ctypes.c_float, ctypes.c_float)
    own = fn_t(lambda U: 1.0 if U < 0.5 else 0.0)
    r = RedistFunc(ctypes.cast(own, ctypes.c_void_p), 0.5)
    p = np.zeros(3, P3D_DTYPE)
    with pytest.raises(cv.CvtxError, match="redistributor"):
        cv.P3D_redistribute_on_grid(p, r, 0.1)
