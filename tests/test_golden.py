"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces them (the FP32 restatement bit for bit, the FP64
oracle to 1e-12).  GPU: the library matches the fixtures' reference outputs
(relative L2 <= 1e-5; where the reference's own FP32 term is the limit --
Gaussian, filament stretching -- against the FP64 fixture as in the other
parity tests)."""
import os

import numpy as np
import pytest

import cvortex_amd as cv
import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REGS = ("singular", "winckelmans", "planetary", "gaussian")


def load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def test_oracle_reproduces_p3d(orc):
    g = load("p3d.npz")
    s, t, sig, nu = g["src"], g["tgt"], float(g["sigma"]), float(g["nu"])
    for k in REGS:
        assert np.array_equal(orc.P3D_M2M_vel(s, t["coord"], k, sig), g[f"ref_vel_{k}"])
        assert np.array_equal(orc.P3D_M2M_dvort(s, t, k, sig), g[f"ref_dvort_{k}"])
        visc = k in ("winckelmans", "gaussian")
        v, d, e = orc.f64_P3D(s, t, k, sig, nu, gauss_mode=1, visc=visc)
        np.testing.assert_allclose(v, g[f"f64_vel_{k}"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(d, g[f"f64_dvort_{k}"], rtol=1e-12, atol=0)
        if visc:
            assert np.array_equal(orc.P3D_M2M_visc_dvort(s, t, k, sig, nu), g[f"ref_visc_{k}"])
            np.testing.assert_allclose(e, g[f"f64_visc_{k}"], rtol=1e-12, atol=0)


def test_oracle_reproduces_f3d_p2d(orc):
    f = load("f3d.npz")
    assert np.array_equal(orc.F3D_M2M_vel(f["fil"], f["tgt"]["coord"]), f["ref_vel"])
    assert np.array_equal(orc.F3D_M2M_dvort(f["fil"], f["tgt"]), f["ref_dvort"])
    g = load("p2d.npz")
    for k in REGS:
        assert np.array_equal(orc.P2D_M2M_vel(g["src"], g["tgt"]["coord"], k, float(g["sigma"])),
                              g[f"ref_vel_{k}"])


@pytest.mark.gpu
def test_gpu_matches_golden():
    from gpuutil import assert_parity, assert_parity_limited, case, have_gpu
    if not have_gpu():
        pytest.fail("no HIP device")
    cv.set_accumulation(cv.ACC_COMP)
    g = load("p3d.npz")
    s, t, sig, nu = g["src"], g["tgt"], float(g["sigma"]), float(g["nu"])
    for k in REGS:
        case(f"golden P3D {k}")
        visc = k in ("winckelmans", "gaussian")
        v, d, e = cv.P3D_M2M_fused(s, t, k, sig, nu, vel=True, dvort=True, visc=visc)
        chk = (lambda n, a, r, x: assert_parity_limited(n, a, r, x, 1e-5)) if k == "gaussian" else \
              (lambda n, a, r, x: assert_parity(n, a, r, 1e-5))
        chk("vel", v, g[f"ref_vel_{k}"], g[f"f64_vel_{k}"])
        chk("dvort", d, g[f"ref_dvort_{k}"], g[f"f64_dvort_{k}"])
        if visc:
            chk("visc", e, g[f"ref_visc_{k}"], g[f"f64_visc_{k}"])
    f = load("f3d.npz")
    case("golden F3D")
    v, d = cv.F3D_M2M_fused(f["fil"], f["tgt"])
    assert_parity("F3D vel", v, f["ref_vel"], 1e-5)
    assert_parity_limited("F3D dvort", d, f["ref_dvort"], f["f64_dvort"], 1e-5)
    g = load("p2d.npz")
    for k in REGS:
        case(f"golden P2D {k}")
        vel = cv.P2D_M2M_vel(g["src"], g["tgt"]["coord"], k, float(g["sigma"]))
        if k == "gaussian":
            assert_parity_limited("P2D vel", vel, g[f"ref_vel_{k}"], g[f"f64_vel_{k}"], 1e-5)
        else:
            assert_parity("P2D vel", vel, g[f"ref_vel_{k}"], 1e-5)
        if k in ("winckelmans", "gaussian"):
            e = cv.P2D_M2M_visc_dvort(g["src"], g["tgt"], k, float(g["sigma"]), float(g["nu"]))
            assert_parity("P2D visc", e, g[f"ref_visc_{k}"], 1e-5)
