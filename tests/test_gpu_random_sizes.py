"""Randomised sizes and scales, all outputs, against the FP64 oracle.

Each case draws a source and a target count log-uniformly in [1, 12000]
(so most launches have partial tiles, partial target blocks and, for the
larger ones, source partitions), a regularisation radius from 0.3 to 3
mean spacings, and evaluates P3D vel/dvort (and visc where the
regularisation has one) through the fused call, P2D vel (and visc), and
F3D vel/dvort on filaments among the same targets.  Tolerance (relative L2
against the FP64 oracle): 1e-5 for P3D and P2D; for F3D 5e-5, or 1.5 x the
error of the oracle's FP32 restatement of the reference where that is
larger: targets here may sit next to a filament, where the stretching's
|r1 x r2| is ill-conditioned in FP32 (reference src/F3D.cpp:68-73; the
worst case measured 3.7e-5)."""
import numpy as np
import pytest

import cvortex_amd as cv
import oracle
from cvortex_amd import workloads as W
from gpuutil import case, have_gpu

pytestmark = pytest.mark.gpu
REGS = ["singular", "winckelmans", "planetary", "gaussian"]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")


def check(name, got, ref, tol):
    got = np.asarray(got, np.float64)
    assert np.all(np.isfinite(got)), f"{name}: non-finite GPU output"
    e = oracle.rel_l2(got, ref)
    print(f"  {name}: relL2 vs FP64 {e:.2e}", flush=True)
    assert e <= tol, f"{name}: relative L2 vs FP64 {e:.3e} > {tol:g}"


@pytest.mark.parametrize("it", range(24))
def test_random_sizes(orc, it):
    rng = np.random.default_rng(7000 + it)
    ns = int(np.exp(rng.uniform(0, np.log(12000))))
    nt = int(np.exp(rng.uniform(0, np.log(12000))))
    reg = REGS[it % 4]
    case(f"random sizes it={it} ns={ns} nt={nt} {reg}")
    sigma = float(rng.uniform(0.3, 3.0) / ns ** (1 / 3))
    nu = 1e-3
    p = W.random_p3d(ns, 1000 + it, box=1.0, vort=1.0, vol=1.0 / ns)
    t = W.random_p3d(nt, 2000 + it, box=1.0, vort=1.0, vol=1.0 / nt)
    visc = reg in ("winckelmans", "gaussian")
    v64, d64, e64 = orc.f64_P3D(p, t, reg, sigma, nu=nu, vel=True, dvort=True, visc=visc)
    vel, dv, vi = cv.P3D_M2M_fused(p, t, reg, sigma, nu, visc=visc)
    check(f"P3D {reg} vel", vel, v64, 1e-5)
    check(f"P3D {reg} dvort", dv, d64, 1e-5)
    if visc:
        check(f"P3D {reg} visc", vi, e64, 1e-5)

    s2 = float(rng.uniform(0.3, 3.0) / ns ** 0.5)
    p2 = W.random_p2d(ns, 3000 + it, box=1.0, vort=1.0, area=1.0 / ns)
    t2 = W.random_p2d(nt, 4000 + it, box=1.0, vort=1.0, area=1.0 / nt)
    r2, re2 = orc.f64_P2D(p2, t2, reg, s2, nu=nu, vel=True, visc=visc)
    check(f"P2D {reg} vel", cv.P2D_M2M_vel(p2, t2["coord"], reg, s2), r2, 1e-5)
    if visc:
        check(f"P2D {reg} visc", cv.P2D_M2M_visc_dvort(p2, t2, reg, s2, nu), re2, 1e-5)

    f = W.random_f3d(max(1, ns // 4), 5000 + it, box=1.0, strength=1.0)
    fv64, fd64 = orc.f64_F3D(f, t)
    fv32, fd32 = orc.F3D_M2M_vel(f, t["coord"]), orc.F3D_M2M_dvort(f, t)
    check("F3D vel", cv.F3D_M2M_vel(f, t["coord"]), fv64,
          max(5e-5, 1.5 * oracle.rel_l2(np.asarray(fv32, np.float64), fv64)))
    check("F3D dvort", cv.F3D_M2M_dvort(f, t), fd64,
          max(5e-5, 1.5 * oracle.rel_l2(np.asarray(fd32, np.float64), fd64)))
