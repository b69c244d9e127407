"""The no-allocation entry point cvtx_hip_p3d_eval_shape_noalloc against the
existing cvtx_hip_p3d_eval_shape with the same launch shape, on two calls in
a row (the second smaller), and against the oracle's restatement of the
reference CPU path.  Its partial sums live in a fixed scratch array; a shape
whose partial sums exceed it is rejected before any launch."""
import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import device as D
from cvortex_amd import workloads as W
from gpuutil import assert_parity, assert_parity_limited, case, have_gpu
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")
    cv.set_accumulation(cv.ACC_COMP)


def run_both(p, t, kernel, sigma, nu, outs, shape):
    """(no-allocation outputs, existing entry point outputs) for one problem."""
    spv, sa = D.p3d_soa(p)
    tpv, ta = D.p3d_soa(t)
    d_spv, d_sa = D.DeviceArray.from_host(spv), D.DeviceArray.from_host(sa)
    d_tpv, d_ta = D.DeviceArray.from_host(tpv), D.DeviceArray.from_host(ta)
    res = []
    for fn in ("noalloc", "shape"):
        o = [D.DeviceArray((len(t), 3), np.float32) if f else None for f in outs]
        for a in o:
            if a is not None:  # stale values would show through a missed write
                a.upload(np.full((len(t), 3), np.nan, np.float32))
        prob = D.p3d_problem(d_spv, d_sa, len(p), d_tpv, d_ta, len(t), *o, kernel, sigma, nu,
                             cv.ACC_COMP)
        if fn == "noalloc":
            D.p3d_eval_noalloc(prob, shape)
        else:
            D.p3d_eval(prob, shape=shape)
        D.synchronize()
        res.append([a.download() if a is not None else None for a in o])
    return res


def test_noalloc_matches_eval_shape_twice(orc):
    # first call: Gaussian vel + dvort, 8192 targets x 65536 sources in 32
    # source partitions (32 x 9 x 8192 x 8 B = 18.9 MB of partial sums)
    p = W.random_p3d(65536, 1401, box=1.0, vort=1.0)
    t = W.random_p3d(8192, 1402, box=1.0, vort=1.0)
    case("noalloc call 1: gaussian vel+dvort 65536 x 8192, T=1, 32 partitions")
    (na, ref) = run_both(p, t, "gaussian", 0.05, 0.0, (True, True, False), (1, 32))
    for name, a, b in zip(("vel", "dvort"), na[:2], ref[:2]):
        assert np.all(np.isfinite(a)), f"{name}: non-finite no-allocation output"
        e = oracle.rel_l2(a, b)
        print(f"  call 1 {name}: relL2 vs cvtx_hip_p3d_eval_shape {e:.2e}, "
              f"bit-identical {np.array_equal(a, b)}", flush=True)
        assert e <= 1e-6
    # second call, smaller: Winckelmans vel + dvort + visc, 3000 targets x
    # 20000 sources in 8 partitions
    p = W.random_p3d(20000, 1403, box=1.0, vort=1.0)
    t = W.random_p3d(3000, 1404, box=1.0, vort=1.0)
    case("noalloc call 2: winckelmans vel+dvort+visc 20000 x 3000, T=2, 8 partitions")
    (na, ref) = run_both(p, t, "winckelmans", 0.05, 0.01, (True, True, True), (2, 8))
    for name, a, b in zip(("vel", "dvort", "visc"), na, ref):
        assert np.all(np.isfinite(a)), f"{name}: non-finite no-allocation output"
        e = oracle.rel_l2(a, b)
        print(f"  call 2 {name}: relL2 vs cvtx_hip_p3d_eval_shape {e:.2e}, "
              f"bit-identical {np.array_equal(a, b)}", flush=True)
        assert e <= 1e-6
    assert_parity("call 2 vel vs reference", na[0], orc.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.05), TOL)
    assert_parity("call 2 dvort vs reference", na[1], orc.P3D_M2M_dvort(p, t, "winckelmans", 0.05), TOL)
    assert_parity("call 2 visc vs reference", na[2],
                  orc.P3D_M2M_visc_dvort(p, t, "winckelmans", 0.05, 0.01), TOL)


def test_noalloc_unpartitioned_and_oracle(orc):
    # one partition needs no scratch at all
    p = W.random_p3d(5000, 1405, box=1.0, vort=1.0)
    case("noalloc: gaussian vel+dvort 5000 x 5000, T=2, 1 partition")
    (na, ref) = run_both(p, p, "gaussian", 0.1, 0.0, (True, True, False), (2, 1))
    assert np.array_equal(na[0], ref[0]) or oracle.rel_l2(na[0], ref[0]) <= 1e-6
    assert_parity("vel vs reference", na[0], orc.P3D_M2M_vel(p, p["coord"], "gaussian", 0.1), TOL)
    _, d64, _ = orc.f64_P3D(p, p, "gaussian", 0.1, gauss_mode=1)
    assert_parity_limited("dvort vs reference", na[1], orc.P3D_M2M_dvort(p, p, "gaussian", 0.1),
                          d64, TOL)


def test_noalloc_rejects_oversized_shape():
    # 2^20 targets x 16 partitions x 9 accumulators x 8 B = 1.2 GB > 64 MiB:
    # rejected on the host before any launch (the outputs keep their values)
    n = 1 << 20
    pv = D.DeviceArray((n, 4), np.float32)
    pv.upload(np.zeros((n, 4), np.float32))
    out = D.DeviceArray((n, 3), np.float32)
    out.upload(np.full((n, 3), 7.0, np.float32))
    prob = D.p3d_problem(pv, pv, n, pv, pv, n, out, out, None, "winckelmans", 0.1, 0.0, cv.ACC_COMP)
    case("noalloc: oversized shape is rejected")
    with pytest.raises(cv.CvtxError, match="scratch"):
        D.p3d_eval_noalloc(prob, (1, 16))
    with pytest.raises(cv.CvtxError, match="shape"):
        D.p3d_eval_noalloc(prob, (3, 1))
    D.synchronize()
    assert np.all(out.download() == 7.0)
