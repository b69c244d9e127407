"""Far-tile pair loops (cvortex_amd/csrc/tile_cull.hpp) on inputs where they
run: spatially coherent, sorted particle sets whose target blocks and source
tiles are mostly far apart, so most tiles take the loop with no range tests
and one reciprocal per four pairs, and the rest the general loop.  Checked
against the oracle's restatement of the reference CPU path (relative L2 <=
1e-5) for every launch shape, and for sets where the far loop must not run
(a non-finite coordinate, every pair in one core)."""
import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import device as D
from cvortex_amd import workloads as W
from gpuutil import assert_parity, case, have_gpu
import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")
    cv.set_accumulation(cv.ACC_COMP)


def sample(n, m, seed):
    return np.sort(np.random.default_rng(seed).choice(n, size=m, replace=False))


def p2d_dev(p, kernel, sigma, nu, visc, shape, acc=cv.ACC_COMP, tgt=None):
    tgt = p if tgt is None else tgt
    n = len(tgt)
    s = D.DeviceArray.from_host(D.p2d_soa(p))
    t = D.DeviceArray.from_host(D.p2d_soa(tgt))
    vel = D.DeviceArray((n, 2))
    vs = D.DeviceArray((n,)) if visc else None
    prob = D.P2DProblem(s.addr, len(p), t.addr, n, vel.addr, vs.addr if visc else None,
                        {"planetary": 2, "gaussian": 3}[kernel], sigma, nu, acc)
    D.p2d_eval(prob, None, shape=shape)
    D.synchronize()
    return vel.download(), (vs.download() if visc else None)


@pytest.mark.parametrize("kernel", ("planetary", "gaussian"))
@pytest.mark.parametrize("T", (1, 2, 4))
def test_p2d_sorted_sheet(orc, kernel, T):
    p, sigma, nu = W.sheet(1 << 16)
    n = len(p)
    idx = sample(n, 384, 7)
    case(f"P2D {kernel} sorted sheet 2^16, T={T}")
    visc = kernel == "gaussian"
    v, e = p2d_dev(p, kernel, sigma, nu, visc, (T, 4))
    assert_parity(f"{kernel} vel T={T}", v[idx], orc.P2D_M2M_vel(p, p["coord"][idx], kernel, sigma), TOL)
    v64, _ = orc.f64_P2D(p, p, kernel, sigma, idx=idx.astype(np.int32))
    assert oracle.rel_l2(v[idx], v64) <= TOL
    if visc:
        ref = orc.P2D_M2M_visc_dvort(p, p[idx], kernel, sigma, nu)
        # uniform strengths and areas: every PSE pair term is exactly zero
        assert np.all(ref == 0) and np.all(e == 0)


def test_p2d_sheet_with_a_nonfinite_target(orc):
    # a NaN coordinate in one target: its block's box says nothing about its
    # pairs, so that block takes the general loop; that target's velocity is
    # NaN, as in the reference, and no other target's is
    p, sigma, _ = W.sheet(1 << 14)
    t = p.copy()
    t["coord"][5000, 1] = np.nan
    case("P2D planetary sheet 2^14, one NaN target coordinate")
    v, _ = p2d_dev(p, "planetary", sigma, 0.0, False, (4, 1), tgt=t)
    ref = orc.P2D_M2M_vel(p, t["coord"], "planetary", sigma)
    nan_rows = np.isnan(v).any(axis=1)
    assert np.array_equal(nan_rows, np.isnan(ref).any(axis=1))
    assert nan_rows.sum() == 1 and nan_rows[5000]
    assert_parity("finite targets", v[~nan_rows], ref[~nan_rows], TOL)


def test_p2d_dense_cluster_and_far_sheet(orc):
    # a dense cluster whose pairs all lie inside the core (general loop)
    # next to a sorted sheet far away (far loop), in one problem
    rng = np.random.default_rng(3)
    a, sigma, _ = W.sheet(1 << 14)
    b = a[:4096].copy()
    b["coord"] = (rng.normal(size=(4096, 2)) * 1e-4 + [5.0, 5.0]).astype(np.float32)
    p = np.concatenate([a, b])
    idx = np.concatenate([sample(len(a), 200, 8), len(a) + sample(len(b), 200, 9)])
    for T in (1, 4):
        case(f"P2D planetary sheet + dense cluster, T={T}")
        v, _ = p2d_dev(p, "planetary", sigma, 0.0, False, (T, 1))
        assert_parity(f"mixed T={T}", v[idx], orc.P2D_M2M_vel(p, p["coord"][idx], "planetary", sigma),
                      TOL)


def f3d_dev(f, t, mask, shape, order=True):
    fa, fb = D.f3d_soa(f)
    pv, a = D.p3d_soa(t)
    dfa, dfb = D.DeviceArray.from_host(fa), D.DeviceArray.from_host(fb)
    dpv, da = D.DeviceArray.from_host(pv), D.DeviceArray.from_host(a)
    vel = D.DeviceArray((len(t), 3)) if mask & cv.OUT_VEL else None
    dv = D.DeviceArray((len(t), 3)) if mask & cv.OUT_DVORT else None
    prob = D.F3DProblem(dfa.addr, dfb.addr, len(f), dpv.addr, da.addr, len(t),
                        vel.addr if vel else None, dv.addr if dv else None, cv.ACC_COMP)
    prev = cv.lib().cvtx_hip_set_target_order(1 if order else 0)
    try:
        D.f3d_eval(prob, None, shape=shape)
        D.synchronize()
    finally:
        cv.lib().cvtx_hip_set_target_order(prev)
    return (vel.download() if vel else None), (dv.download() if dv else None)


@pytest.mark.parametrize("mask", (1, 2, 3))
@pytest.mark.parametrize("T", (1, 2, 4))
def test_f3d_lattice_ordered_targets(orc, mask, T):
    # 64 x 64 lattice on 2^16 targets: ordered target blocks are far from
    # most filament tiles
    f, t = W.lattice(1 << 16, 64, 64)
    idx = sample(len(t), 384, 11)
    case(f"F3D lattice 4096 x 2^16, mask={mask}, T={T}")
    v, d = f3d_dev(f, t, mask, (T, 2))
    if mask & 1:
        assert_parity(f"vel T={T}", v[idx], orc.F3D_M2M_vel(f, t["coord"][idx]), TOL)
    if mask & 2:
        _, d64 = orc.f64_F3D(f, t, idx=idx.astype(np.int32), vel=False)
        from gpuutil import assert_parity_limited
        assert_parity_limited(f"dvort T={T}", d[idx], orc.F3D_M2M_dvort(f, t[idx]), d64, TOL)
    # the unordered launch (general loop almost everywhere) agrees
    v2, d2 = f3d_dev(f, t, mask, (T, 2), order=False)
    for a, b in ((v, v2), (d, d2)):
        if a is not None:
            assert oracle.rel_l2(a, b) <= 1e-6


def test_f3d_targets_on_filament_lines(orc):
    # targets exactly on the extensions of far filaments' lines (and at a
    # filament's end): those pairs contribute nothing, in the far-tile loop
    # as in the general one and in the reference
    from cvortex_amd.types import make_f3d, make_p3d
    n = 64
    st = np.zeros((n, 3), np.float32)
    st[:, 1] = np.arange(n, dtype=np.float32) / 64.0
    en = st.copy()
    en[:, 0] = 1.0 / 128.0
    f = make_f3d(st, en, np.linspace(0.5, 1.5, n))
    # a block of 4096 targets along x on the line of filament 0 (y = 0,
    # z = 0), far beyond its end, plus random targets elsewhere
    rng = np.random.default_rng(12)
    line = np.zeros((4096, 3), np.float32)
    line[:, 0] = np.linspace(3.0, 4.0, 4096, dtype=np.float32)
    other = rng.uniform([-1, -1, 0.5], [2, 2, 1.5], size=(12288, 3)).astype(np.float32)
    x = np.concatenate([line, other])
    t = make_p3d(x, rng.normal(size=(len(x), 3)), np.ones(len(x)))
    idx = np.concatenate([np.arange(0, 4096, 64), 4096 + sample(12288, 64, 13)])
    for T in (1, 2):
        case(f"F3D targets on filament lines, T={T}")
        v, d = f3d_dev(f, t, 3, (T, 1))
        # far from a segment 1/128 long, the reference's FP32 form of the
        # velocity (t2 = r0.r1/|r1| - r0.r2/|r2|) cancels to ~1e-4: the
        # FP64 oracle is the bar there (assert_parity_limited)
        from gpuutil import assert_parity_limited
        v64, d64 = orc.f64_F3D(f, t, idx=idx.astype(np.int32))
        assert np.all(v[:4096, :2] == 0)
        assert_parity_limited(f"vel on lines T={T}", v[idx], orc.F3D_M2M_vel(f, t["coord"][idx]),
                              v64, TOL)
        ref = orc.F3D_M2M_dvort(f, t[idx])
        assert np.all(np.isfinite(d))
        assert_parity_limited(f"dvort on lines T={T}", d[idx], ref, d64, TOL)


@pytest.mark.parametrize("mask", (1, 2, 3))
def test_f3d_polylines_partly_chained(orc, mask):
    # polylines of 1..7 segments: where a filament ends exactly where the
    # next one begins (bit for bit), the far loop shares that point's
    # distance terms between the two.  Links break at arbitrary positions,
    # so groups mix chained and unchained filaments; one link is broken by
    # a single ulp and must not count as chained.
    from cvortex_amd.types import make_f3d, make_p3d
    from gpuutil import assert_parity_limited
    rng = np.random.default_rng(21)
    st, en = [], []
    while len(st) < 3000:
        m = int(rng.integers(1, 8))
        path = (rng.uniform(-1, 1, 3) +
                np.cumsum(rng.normal(scale=0.05, size=(m + 1, 3)), axis=0)).astype(np.float32)
        st.extend(path[:-1])
        en.extend(path[1:])
    st = np.array(st[:3000], np.float32)
    en = np.array(en[:3000], np.float32)
    linked = np.all(en[:-1] == st[1:], axis=1)
    k = int(np.flatnonzero(linked)[5])
    en[k, 0] = np.nextafter(en[k, 0], np.float32(np.inf))
    f = make_f3d(st, en, rng.uniform(0.5, 1.5, len(st)))
    # targets far from every filament (far tiles) and among them (general loop)
    x = np.concatenate([rng.uniform(3.0, 4.0, size=(12288, 3)),
                        rng.uniform(-1.2, 1.2, size=(4096, 3))]).astype(np.float32)
    t = make_p3d(x, rng.normal(size=(len(x), 3)), np.ones(len(x)))
    idx = np.concatenate([sample(12288, 256, 22), 12288 + sample(4096, 128, 23)])
    v64, d64 = orc.f64_F3D(f, t, idx=idx.astype(np.int32))
    for T in (1, 2):
        case(f"F3D partly chained polylines, mask={mask}, T={T}")
        v, d = f3d_dev(f, t, mask, (T, 1))
        if mask & 1:
            assert_parity_limited(f"vel T={T}", v[idx], orc.F3D_M2M_vel(f, t["coord"][idx]), v64,
                                  TOL)
        if mask & 2:
            assert np.all(np.isfinite(d))
            assert_parity_limited(f"dvort T={T}", d[idx], orc.F3D_M2M_dvort(f, t[idx]), d64, TOL)
