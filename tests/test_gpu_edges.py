"""Edge cases of the GPU path: empty and tiny inputs, sizes that are not
multiples of the 256-source tile, aliasing of sources and targets,
coincident particles, negative regularisation radius, every launch shape,
plain, tile-compensated and compensated accumulation, and the device-resident API against the
host-array API.  Tolerance: relative L2 <= 1e-5 against the oracle's FP32
restatement of the reference (written per assertion)."""
import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import device as D
from cvortex_amd import workloads as W
from gpuutil import assert_parity, assert_parity_gaussian, assert_parity_limited, case, have_gpu
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")
    cv.set_accumulation(cv.ACC_COMP)


def check(orc, k, p, t, sigma, visc_nu=None, acc_note=""):
    """GPU vel/dvort(/visc) of sources p on targets t against the oracle."""
    v, d, s = cv.P3D_M2M_fused(p, t, k, sigma, visc_nu or 0.0, vel=True, dvort=True,
                               visc=visc_nu is not None)
    refs = (orc.P3D_M2M_vel(p, t["coord"], k, sigma), orc.P3D_M2M_dvort(p, t, k, sigma),
            orc.P3D_M2M_visc_dvort(p, t, k, sigma, visc_nu) if visc_nu is not None else None)
    names = (f"vel {k}{acc_note}", f"dvort {k}{acc_note}", f"visc {k}{acc_note}")
    if k == "gaussian":
        ex = orc.f64_P3D(p, t, k, sigma, visc_nu or 0.0, gauss_mode=1, visc=visc_nu is not None)
        for n_, g, r, e in zip(names, (v, d, s), refs, ex):
            if g is not None:
                assert_parity_gaussian(n_, g, r, e, TOL)
    else:
        for n_, g, r in zip(names, (v, d, s), refs):
            if g is not None:
                assert_parity(n_, g, r, TOL)


def test_empty_sources_give_zero():
    case("empty sources")
    p = W.random_p3d(0, 1)
    t = W.random_p3d(300, 2)
    assert np.all(cv.P3D_M2M_vel(p, t["coord"], "winckelmans", 0.3) == 0)
    assert np.all(cv.P3D_M2M_dvort(p, t, "winckelmans", 0.3) == 0)
    f = W.random_f3d(0, 3)
    assert np.all(cv.F3D_M2M_vel(f, t["coord"]) == 0)
    q = W.random_p2d(0, 4)
    assert np.all(cv.P2D_M2M_vel(q, np.ones((5, 2), np.float32), "gaussian", 0.3) == 0)


def test_empty_targets_are_a_no_op():
    case("empty targets")
    p = W.random_p3d(300, 1)
    assert cv.P3D_M2M_vel(p, np.zeros((0, 3), np.float32), "winckelmans", 0.3).shape == (0, 3)
    assert cv.P3D_M2M_dvort(p, W.random_p3d(0, 2), "gaussian", 0.3).shape == (0, 3)


@pytest.mark.parametrize("ns,nt", [(1, 1), (1, 700), (255, 3), (256, 256), (257, 1),
                                   (513, 1031), (4097, 65), (70000, 129)])
def test_ragged_sizes(orc, ns, nt):
    case(f"ragged ns={ns} nt={nt}")
    p = W.random_p3d(ns, 10 + ns, box=1.0, vort=1.0)
    t = W.random_p3d(nt, 20 + nt, box=1.0, vort=1.0)
    for k in ("singular", "winckelmans", "gaussian"):
        check(orc, k, p, t, 0.2)


def test_coincident_particles_are_excluded(orc):
    # duplicated coordinates with different strengths: the reference skips
    # every pair with equal coordinates (src/P3D.cpp:58,94,127)
    case("coincident particles")
    p = W.random_p3d(600, 31, box=1.0, vort=1.0)
    p["coord"][300:] = p["coord"][:300]
    for k in ("singular", "winckelmans", "planetary", "gaussian"):
        check(orc, k, p, p, 0.2, 0.05 if k in ("winckelmans", "gaussian") else None)


def test_negative_radius(orc):
    # CPU semantics: vel uses |sigma|, dvort the signed sigma^3 (P3D.cpp:82,102)
    case("negative sigma")
    p = W.random_p3d(500, 41, box=1.0, vort=1.0)
    for k in ("winckelmans", "gaussian", "singular"):
        check(orc, k, p, p, -0.2)


@pytest.mark.parametrize("T", [1, 2, 4])
@pytest.mark.parametrize("splits", [1, 3, 16])
def test_every_launch_shape(orc, T, splits):
    case(f"shape T={T} splits={splits}")
    p = W.random_p3d(3000, 51, box=1.0, vort=1.0)
    dev = D.P3DDevice(p)
    for k in ("winckelmans", "gaussian"):
        outs = dev.outputs(True, True, True)
        D.p3d_eval(dev.problem(outs, k, 0.1, 0.05), shape=(T, splits))
        got = [o.download() for o in outs]
        refs = (orc.P3D_M2M_vel(p, p["coord"], k, 0.1), orc.P3D_M2M_dvort(p, p, k, 0.1),
                orc.P3D_M2M_visc_dvort(p, p, k, 0.1, 0.05))
        if k == "gaussian":
            ex = orc.f64_P3D(p, p, k, 0.1, 0.05, gauss_mode=1, visc=True)
            for n_, g, r, e in zip(("vel", "dvort", "visc"), got, refs, ex):
                assert_parity_gaussian(f"{n_} {k}", g, r, e, TOL)
        else:
            for n_, g, r in zip(("vel", "dvort", "visc"), got, refs):
                assert_parity(f"{n_} {k}", g, r, TOL)


def test_plain_accumulation(orc):
    case("plain accumulation")
    p = W.random_p3d(5000, 61, box=1.0, vort=1.0)
    prev = cv.set_accumulation(cv.ACC_PLAIN)
    try:
        for k in ("winckelmans", "gaussian"):
            check(orc, k, p, p, 0.1, 0.05, " (plain)")
    finally:
        cv.set_accumulation(prev)


def test_tiles_accumulation(orc):
    # FP32 tile sums folded by TwoSum for every sum, no per-pair FP64
    case("tiles accumulation")
    p = W.random_p3d(5000, 62, box=1.0, vort=1.0)
    prev = cv.set_accumulation(cv.ACC_TILES)
    try:
        for k in ("winckelmans", "gaussian", "planetary", "singular"):
            check(orc, k, p, p, 0.1, 0.05 if k in ("winckelmans", "gaussian") else None,
                  " (tiles)")
    finally:
        cv.set_accumulation(prev)


def test_tiles_accumulation_on_a_large_random_cloud(orc):
    # 2^17 particles with independent strengths: the TwoSum-folded FP32 tile
    # sums stay within 1e-5 of the FP64 sums, where a single FP32 running sum
    # drifts further
    case("tiles accumulation, 2^17 random cloud")
    n = 1 << 17
    p = W.random_p3d(n, 63, box=1.0, vort=1.0)
    idx = np.sort(np.random.default_rng(64).choice(n, size=512, replace=False)).astype(np.int32)
    v64, d64, _ = orc.f64_P3D(p, p, "winckelmans", 0.05, idx=idx)
    prev = cv.set_accumulation(cv.ACC_TILES)
    try:
        v, d, _ = cv.P3D_M2M_fused(p, p, "winckelmans", 0.05, 0.0, vel=True, dvort=True)
    finally:
        cv.set_accumulation(prev)
    ev, ed = oracle.rel_l2(v[idx], v64), oracle.rel_l2(d[idx], d64)
    print(f"tiles vs FP64: vel {ev:.3g} dvort {ed:.3g}", flush=True)
    assert ev <= 1e-5 and ed <= 1e-5, (ev, ed)


@pytest.mark.parametrize("n", [20000, 70000])
def test_target_order_does_not_change_results(n):
    # the spatial target order changes which targets share a wave, not the
    # order in which any target's sources are summed: results are identical
    case(f"target order on/off n={n}")
    p = W.random_p3d(n, 91, box=1.0, vort=1.0)
    L = cv.lib()
    prev = L.cvtx_hip_set_target_order(1)
    try:
        on = cv.P3D_M2M_fused(p, p, "gaussian", 0.05, 0.01, vel=True, dvort=True, visc=True)
        L.cvtx_hip_set_target_order(0)
        off = cv.P3D_M2M_fused(p, p, "gaussian", 0.05, 0.01, vel=True, dvort=True, visc=True)
    finally:
        L.cvtx_hip_set_target_order(prev)
    for a, b in zip(on, off):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("what", ["ring65536", "box20000"])
def test_ordered_partitioned_gaussian_matches_oracle(orc, what):
    # From 4096 targets, compensated Gaussian P3D runs with Morton-ordered
    # targets, one target per lane and many source partitions (64 for the
    # ring of 2^16) whose FP64 partials a second kernel adds.  Sampled
    # targets of the fused and the separate calls against the reference
    # restatement and the FP64 oracle.
    if what == "ring65536":
        p, sigma, nu = W.ring(1 << 16)
    else:
        p, sigma, nu = W.random_p3d(20000, 93, box=1.0, vort=1.0, vol=1e-4), 0.05, 0.01
    n = len(p)
    case(f"ordered, partitioned Gaussian {what}: fused and separate calls")
    fused = cv.P3D_M2M_fused(p, p, "gaussian", sigma, nu, vel=True, dvort=True, visc=True)
    sep = (cv.P3D_M2M_vel(p, p["coord"], "gaussian", sigma),
           cv.P3D_M2M_dvort(p, p, "gaussian", sigma),
           cv.P3D_M2M_visc_dvort(p, p, "gaussian", sigma, nu))
    idx = np.sort(np.random.default_rng(7).choice(n, size=512, replace=False)).astype(np.int32)
    t = p[idx]
    refs = (orc.P3D_M2M_vel(p, t["coord"], "gaussian", sigma),
            orc.P3D_M2M_dvort(p, t, "gaussian", sigma),
            orc.P3D_M2M_visc_dvort(p, t, "gaussian", sigma, nu))
    ex = orc.f64_P3D(p, p, "gaussian", sigma, nu, idx=idx, gauss_mode=1, visc=True)
    for kind, outs in (("fused", fused), ("separate", sep)):
        for name, g, r, e in zip(("vel", "dvort", "visc"), outs, refs, ex):
            assert_parity_gaussian(f"{kind} {name}", g[idx], r, e, TOL)
    for a, b in zip(fused, sep):
        assert oracle.rel_l2(a, b) <= 1e-6


def test_autotuner_records_a_shape_that_agrees(tmp_path):
    # cvtx_hip_autotune_p3d times the candidate launch shapes of a device
    # problem and caches the fastest for its size bucket; later evaluations
    # in the bucket use it.  A shape changes how the source sums are split
    # (tiles per partition, FP64 partials), so results agree with the default
    # shape to rounding.  The cache round-trips through save / load.
    case("autotuner: tune, evaluate, save / load / clear")
    L = cv.lib()
    L.cvtx_hip_tune_clear()
    p = W.random_p3d(20000, 95, box=1.0, vort=1.0)
    dev = D.P3DDevice(p)
    outs = dev.outputs(True, True, False)
    prob = dev.problem(outs, "winckelmans", 0.05)
    D.p3d_eval(prob)
    v0, d0 = outs[0].download(), outs[1].download()
    try:
        T, k, ms = D.autotune_p3d(prob, 1)
        print(f"  winner: {T} targets per lane, {k} partitions, {ms:.3f} ms", flush=True)
        assert T in (1, 2, 4) and k >= 1 and ms > 0
        D.p3d_eval(prob)
        v1, d1 = outs[0].download(), outs[1].download()
        assert oracle.rel_l2(v1, v0) <= 1e-6 and oracle.rel_l2(d1, d0) <= 1e-6
        path = str(tmp_path / "shapes.txt")
        assert L.cvtx_hip_tune_save(path.encode()) == 0
        fields = open(path).read().split()
        assert len(fields) == 10 and int(fields[6]) == T and int(fields[7]) == k
        assert int(fields[9]) == 256
        L.cvtx_hip_tune_clear()
        assert L.cvtx_hip_tune_load(path.encode()) == 1
    finally:
        L.cvtx_hip_tune_clear()


def test_device_api_zero_targets():
    # device problems with sources but no targets are no-ops for every kind
    # (the launch-shape choice sees zero target blocks), and the library
    # keeps working afterwards
    case("device API with zero targets")
    p = W.random_p3d(1000, 97, box=1.0, vort=1.0)
    dev = D.P3DDevice(p)
    v, dv, _ = dev.outputs(True, True, False)
    for k in ("winckelmans", "gaussian"):
        D.p3d_eval(D.p3d_problem(dev.pv, dev.a, 1000, dev.pv, dev.a, 0, v, dv, None, k, 0.1))
    q = D.DeviceArray.from_host(D.p2d_soa(W.random_p2d(1000, 98, box=1.0, vort=1.0)))
    v2 = D.DeviceArray((1000, 2))
    D.p2d_eval(D.P2DProblem(q.addr, 1000, q.addr, 0, v2.addr, None, D.REG_ID["winckelmans"], 0.1,
                            0.0, cv.ACC_COMP))
    fa, fb = D.f3d_soa(W.random_f3d(700, 99, box=1.0, strength=1.0))
    dfa, dfb = D.DeviceArray.from_host(fa), D.DeviceArray.from_host(fb)
    D.f3d_eval(D.F3DProblem(dfa.addr, dfb.addr, 700, dev.pv.addr, dev.a.addr, 0, v.addr, dv.addr,
                            cv.ACC_COMP))
    om = D.DeviceArray((1000, 3))
    D.vort_eval(D.VortProblem(dev.pv.addr, dev.a.addr, 1000, dev.pv.addr, 0, om.addr,
                              D.REG_ID["gaussian"], 0.1, cv.ACC_COMP))
    D.synchronize()
    D.p3d_eval(dev.problem((v, dv, None), "winckelmans", 0.1))
    hv, hd, _ = cv.P3D_M2M_fused(p, p, "winckelmans", 0.1, 0.0, vel=True, dvort=True)
    assert np.array_equal(v.download(), hv) and np.array_equal(dv.download(), hd)


def test_device_api_matches_host_api():
    case("device API == host API")
    p = W.random_p3d(2000, 71, box=1.0, vort=1.0)
    dev = D.P3DDevice(p)
    outs = dev.outputs(True, True, False)
    D.p3d_eval(dev.problem(outs, "winckelmans", 0.1))
    v, d = outs[0].download(), outs[1].download()
    hv, hd, _ = cv.P3D_M2M_fused(p, p, "winckelmans", 0.1, 0.0, vel=True, dvort=True)
    assert np.array_equal(v, hv) and np.array_equal(d, hd)


def test_f3d_p2d_device_shapes(orc):
    case("F3D/P2D device shapes")
    f = W.random_f3d(700, 81, box=1.0, strength=1.0)
    t = W.random_p3d(900, 82, box=1.0, vort=1.0)
    fa, fb = D.f3d_soa(f)
    tpv, ta = D.p3d_soa(t)
    dfa, dfb = D.DeviceArray.from_host(fa), D.DeviceArray.from_host(fb)
    dtp, dta = D.DeviceArray.from_host(tpv), D.DeviceArray.from_host(ta)
    for shape in ((1, 1), (2, 2), (4, 5)):
        v = D.DeviceArray((900, 3))
        dv = D.DeviceArray((900, 3))
        D.f3d_eval(D.F3DProblem(dfa.addr, dfb.addr, 700, dtp.addr, dta.addr, 900, v.addr, dv.addr, 1),
                   shape=shape)
        assert_parity(f"F3D vel {shape}", v.download(), orc.F3D_M2M_vel(f, t["coord"]), TOL)
        _, d64 = orc.f64_F3D(f, t, vel=False)
        assert_parity_limited(f"F3D dvort {shape}", dv.download(), orc.F3D_M2M_dvort(f, t), d64, TOL)
    q = W.random_p2d(800, 83, box=1.0, vort=1.0)
    dq = D.DeviceArray.from_host(D.p2d_soa(q))
    for shape in ((1, 1), (2, 3), (4, 2)):
        v = D.DeviceArray((800, 2))
        e = D.DeviceArray((800,))
        D.p2d_eval(D.P2DProblem(dq.addr, 800, dq.addr, 800, v.addr, e.addr, 1, 0.1, 0.05, 1),
                   shape=shape)
        assert_parity(f"P2D vel {shape}", v.download(), orc.P2D_M2M_vel(q, q["coord"], "winckelmans", 0.1), TOL)
        assert_parity(f"P2D visc {shape}", e.download(),
                      orc.P2D_M2M_visc_dvort(q, q, "winckelmans", 0.1, 0.05), TOL)


def test_tiles_mode_is_the_compensated_mode_for_p2d_f3d_vort():
    # P2D velocity, F3D and the vorticity field have no per-pair FP64 sums:
    # the tile-compensated mode runs their compensated kernels, bit for bit;
    # the P2D viscous exchange is summed in FP32 tiles instead of FP64 pairs
    case("ACC_TILES == ACC_COMP for P2D vel, F3D, vorticity field")
    q = D.DeviceArray.from_host(D.p2d_soa(W.random_p2d(3000, 101, box=1.0, vort=1.0)))
    outs = []
    for acc in (cv.ACC_COMP, cv.ACC_TILES):
        v2 = D.DeviceArray((3000, 2))
        s2 = D.DeviceArray((3000,))
        D.p2d_eval(D.P2DProblem(q.addr, 3000, q.addr, 3000, v2.addr, s2.addr,
                                D.REG_ID["gaussian"], 0.05, 0.01, acc))
        outs.append((v2.download(), s2.download()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert oracle.rel_l2(outs[1][1], outs[0][1]) <= 1e-5
    fa, fb = D.f3d_soa(W.random_f3d(700, 102, box=1.0, strength=1.0))
    t = W.random_p3d(2000, 103, box=1.0, vort=1.0)
    tpv, ta = D.p3d_soa(t)
    dfa, dfb = D.DeviceArray.from_host(fa), D.DeviceArray.from_host(fb)
    dtp, dta = D.DeviceArray.from_host(tpv), D.DeviceArray.from_host(ta)
    outs = []
    for acc in (cv.ACC_COMP, cv.ACC_TILES):
        v = D.DeviceArray((2000, 3))
        dv = D.DeviceArray((2000, 3))
        D.f3d_eval(D.F3DProblem(dfa.addr, dfb.addr, 700, dtp.addr, dta.addr, 2000, v.addr, dv.addr,
                                acc))
        outs.append((v.download(), dv.download()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    outs = []
    for acc in (cv.ACC_COMP, cv.ACC_TILES):
        om = D.DeviceArray((2000, 3))
        D.vort_eval(D.VortProblem(dtp.addr, dta.addr, 2000, dtp.addr, 2000, om.addr,
                                  D.REG_ID["gaussian"], 0.1, acc))
        outs.append(om.download())
    assert np.array_equal(outs[0], outs[1])


def test_workspace_calls_on_several_streams():
    # device-resident calls on different streams share the device's
    # workspace (the FP64 partials of source partitions, the target
    # permutation): the library orders them, so calls queued on two caller
    # streams and the library's own stream, with no synchronisation between
    # them, give the results of one stream bit for bit
    from cvortex_amd import hip as H
    case("workspace ordering across streams")
    n = 20000
    p = W.random_p3d(n, 111, box=1.0, vort=1.0)
    dev = D.P3DDevice(p)
    ref_outs = [dev.outputs(True, True, False) for _ in range(2)]
    probs = [dev.problem(ref_outs[0], "winckelmans", 0.1),
             dev.problem(ref_outs[1], "gaussian", 0.05)]
    for pr in probs:  # the workspace reaches its size on the library stream
        D.p3d_eval(pr, None, shape=(1, 16))
    D.synchronize()
    ref = [[o.download() for o in outs[:2]] for outs in ref_outs]
    a, b = H.Stream(), H.Stream()
    for rep in range(4):
        outs = [dev.outputs(True, True, False) for _ in range(3)]
        p0 = dev.problem(outs[0], "winckelmans", 0.1)
        p1 = dev.problem(outs[1], "gaussian", 0.05)
        p2 = dev.problem(outs[2], "winckelmans", 0.1)
        D.p3d_eval(p0, a.addr, shape=(1, 16))
        D.p3d_eval(p1, b.addr, shape=(1, 16))
        D.p3d_eval(p2, None, shape=(1, 16))
        D.p3d_eval(p1, b.addr, shape=(1, 16))
        D.synchronize()
        for got, want in ((outs[0], ref[0]), (outs[1], ref[1]), (outs[2], ref[0])):
            assert np.array_equal(got[0].download(), want[0])
            assert np.array_equal(got[1].download(), want[1])
