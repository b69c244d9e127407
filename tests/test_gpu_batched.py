"""Batched clouds (one persistent launch) and RK4 steps captured as HIP graphs.

Checks: the persistent launch equals one ordinary launch per cloud; both
match the oracle's restatement of the reference on each cloud (sampled
targets); every launch shape gives the same result; a captured RK4 step
replayed several times, with other GPU work and synchronisation between
replays, equals the same RK4 step done eagerly.  Tolerance: relative L2
<= 1e-5 against the oracle (written per assertion); graph vs eager within
1e-6 relative L2 (same kernels, FP32 stage arithmetic)."""
import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import device as D
from cvortex_amd import workloads as W
from gpuutil import assert_parity, case, have_gpu
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")


def small_clouds(n_clouds=24, lo=64, hi=3000, seed=0x5EED0004):
    return W.clouds(n_clouds, lo, hi, seed)


def _dev(p):
    pv, a = D.p3d_soa(p)
    return D.DeviceArray.from_host(pv), D.DeviceArray.from_host(a)


@pytest.mark.parametrize("kernel", ["winckelmans", "gaussian"])
def test_batched_matches_per_cloud_and_oracle(orc, kernel):
    p, off, sig = small_clouds()
    n = len(p)
    case(f"batched {kernel} clouds={len(sig)} n={n}")
    pv, a = _dev(p)
    b = D.Batch(off, sig, kernel, outputs=cv.OUT_VEL | cv.OUT_DVORT | cv.OUT_VISC, nu=1e-3)
    vel, dv, vs = D.DeviceArray((n, 3)), D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    b.eval(pv, a, vel, dv, vs)
    V, Dv, S = vel.download(), dv.download(), vs.download()
    vel2, dv2, vs2 = D.DeviceArray((n, 3)), D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    b.eval_per_cloud(pv, a, vel2, dv2, vs2)
    assert_parity("persistent vs per-cloud vel", V, vel2.download(), 1e-6)
    assert_parity("persistent vs per-cloud dvort", Dv, dv2.download(), 1e-6)
    assert_parity("persistent vs per-cloud visc", S, vs2.download(), 1e-6)
    vel3, dv3, vs3 = D.DeviceArray((n, 3)), D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    b.eval_per_cloud(pv, a, vel3, dv3, vs3, plain=True)
    assert_parity("persistent vs plain per-cloud vel", V, vel3.download(), 1e-6)
    assert_parity("persistent vs plain per-cloud dvort", Dv, dv3.download(), 1e-6)
    assert_parity("persistent vs plain per-cloud visc", S, vs3.download(), 1e-6)
    rng = np.random.default_rng(5)
    for c in range(0, len(sig), 5):
        lo_, hi_ = off[c], off[c + 1]
        cl = np.ascontiguousarray(p[lo_:hi_])
        idx = np.sort(rng.choice(hi_ - lo_, size=min(64, hi_ - lo_), replace=False))
        t = np.ascontiguousarray(cl[idx])
        if kernel == "gaussian":
            v64, d64, s64 = orc.f64_P3D(cl, cl, kernel, float(sig[c]), 1e-3, idx=idx, gauss_mode=1,
                                        visc=True)
        else:
            v64 = orc.P3D_M2M_vel(cl, t["coord"], kernel, float(sig[c]))
            d64 = orc.P3D_M2M_dvort(cl, t, kernel, float(sig[c]))
            s64 = orc.P3D_M2M_visc_dvort(cl, t, kernel, float(sig[c]), 1e-3)
        assert_parity(f"cloud {c} vel", V[lo_:hi_][idx], v64, 1e-5)
        assert_parity(f"cloud {c} dvort", Dv[lo_:hi_][idx], d64, 1e-5)
        assert_parity(f"cloud {c} visc", S[lo_:hi_][idx], s64, 1e-5)
    b.destroy()


@pytest.mark.parametrize("kernel", ["winckelmans", "gaussian"])
def test_batched_tiles_accumulation(orc, kernel):
    # FP32 tile sums folded by TwoSum for every sum, no per-pair FP64
    p, off, sig = small_clouds(16)
    n = len(p)
    case(f"batched tiles accumulation {kernel} n={n}")
    pv, a = _dev(p)
    b = D.Batch(off, sig, kernel, outputs=cv.OUT_VEL | cv.OUT_DVORT | cv.OUT_VISC, nu=1e-3,
                accumulate=cv.ACC_TILES)
    vel, dv, vs = D.DeviceArray((n, 3)), D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    b.eval(pv, a, vel, dv, vs)
    V, Dv, S = vel.download(), dv.download(), vs.download()
    vel2, dv2, vs2 = D.DeviceArray((n, 3)), D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    b.eval_per_cloud(pv, a, vel2, dv2, vs2)
    assert_parity("persistent vs per-cloud vel", V, vel2.download(), 1e-6)
    assert_parity("persistent vs per-cloud dvort", Dv, dv2.download(), 1e-6)
    assert_parity("persistent vs per-cloud visc", S, vs2.download(), 1e-6)
    for c in range(0, len(sig), 3):
        lo_, hi_ = off[c], off[c + 1]
        cl = np.ascontiguousarray(p[lo_:hi_])
        v64, d64, s64 = orc.f64_P3D(cl, cl, kernel, float(sig[c]), 1e-3, gauss_mode=1, visc=True)
        assert_parity(f"cloud {c} vel", V[lo_:hi_], v64, 1e-5)
        assert_parity(f"cloud {c} dvort", Dv[lo_:hi_], d64, 1e-5)
        assert_parity(f"cloud {c} visc", S[lo_:hi_], s64, 1e-5)
    b.destroy()


@pytest.mark.parametrize("T,wpc", [(1, 2), (2, 4), (4, 1), (2, 8)])
def test_batched_shapes(T, wpc):
    p, off, sig = small_clouds(12)
    n = len(p)
    case(f"batched shape T={T} wg/cu={wpc}")
    pv, a = _dev(p)
    ref = D.Batch(off, sig, "winckelmans")
    v0, d0 = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    ref.eval_per_cloud(pv, a, v0, d0)
    b = D.Batch(off, sig, "winckelmans")
    b.set_shape(T, wpc)
    v1, d1 = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    b.eval(pv, a, v1, d1)
    assert_parity("vel", v1.download(), v0.download(), 1e-6)
    assert_parity("dvort", d1.download(), d0.download(), 1e-6)


def test_batched_empty_and_tiny_clouds(orc):
    case("batched empty/tiny clouds")
    sizes = [0, 1, 2, 0, 300, 1, 257, 0]
    off = np.zeros(len(sizes) + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    p = W.random_p3d(int(off[-1]), 77, box=1.0, vort=1.0)
    sig = np.full(len(sizes), 0.1, np.float32)
    pv, a = _dev(p)
    b = D.Batch(off, sig, "winckelmans")
    n = len(p)
    v, d = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    b.eval(pv, a, v, d)
    V = v.download()
    for c in range(len(sizes)):
        cl = np.ascontiguousarray(p[off[c]:off[c + 1]])
        if len(cl) == 0:
            continue
        ref = orc.P3D_M2M_vel(cl, cl["coord"], "winckelmans", 0.1)
        if np.abs(ref).max() == 0:
            assert np.abs(V[off[c]:off[c + 1]]).max() == 0
        else:
            assert_parity(f"cloud {c}", V[off[c]:off[c + 1]], ref, 1e-5)


def _eager_rk4_step(b, pv, a, n, dt):
    """One RK4 step with the batch evaluation and numpy stage updates."""
    y0p, y0a = pv.download(), a.download()
    kv, kd = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    ysp, ysa = D.DeviceArray.from_host(y0p), D.DeviceArray.from_host(y0a)
    S = None
    for c, w in ((0.5 * dt, 1.0), (0.5 * dt, 2.0), (dt, 2.0), (None, 1.0)):
        b.eval(ysp, ysa, kv, kd)
        K = np.concatenate([kv.download(), kd.download()], axis=1).astype(np.float32)
        S = w * K if S is None else S + w * K
        if c is None:
            break
        sp, sa = y0p.copy(), y0a.copy()
        sp[:, :3] += c * K[:, :3]
        sa[:, :3] += c * K[:, 3:]
        ysp.upload(sp)
        ysa.upload(sa)
    y0p[:, :3] += dt / 6.0 * S[:, :3]
    y0a[:, :3] += dt / 6.0 * S[:, 3:]
    return y0p, y0a


def test_rk4_graph_equals_eager():
    p, off, sig = small_clouds(10, 64, 2000)
    n = len(p)
    case(f"RK4 graph vs eager n={n}")
    dt = 1e-3
    pv, a = _dev(p)
    pv_e, a_e = _dev(p)
    bg = D.Batch(off, sig, "winckelmans")
    be = D.Batch(off, sig, "winckelmans")
    bg.rk4_init(pv, a, dt)
    for step in range(3):
        bg.rk4_step(1)
        D.synchronize()
        # unrelated eager GPU work between replays
        junk = D.DeviceArray((n, 3))
        be.eval(pv_e, a_e, junk, D.DeviceArray((n, 3)))
        D.synchronize()
        ep, ea = _eager_rk4_step(be, pv_e, a_e, n, dt)
        pv_e.upload(ep)
        a_e.upload(ea)
        gp, ga = pv.download(), a.download()
        assert_parity(f"step {step} positions", gp[:, :3], ep[:, :3], 1e-6)
        assert_parity(f"step {step} vorticity", ga[:, :3], ea[:, :3], 1e-6)
        assert np.array_equal(gp[:, 3], p["volume"])


def test_rk4_moves_particles():
    # a single vortex ring translates along its axis; RK4 steps through the
    # graph move every particle by about u dt
    n = 512
    th = 2 * np.pi * np.arange(n) / n
    R, G = 1.0, 1.0
    pos = np.stack([R * np.cos(th), R * np.sin(th), np.zeros(n)], 1)
    al = G * 2 * np.pi * R / n * np.stack([-np.sin(th), np.cos(th), np.zeros(n)], 1)
    from cvortex_amd.types import make_p3d
    p = make_p3d(pos, al, np.full(n, 1e-3))
    pv, a = _dev(p)
    case("RK4 ring translation")
    b = D.Batch(np.array([0, n]), np.array([0.1], np.float32), "winckelmans")
    v, d = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    b.eval(pv, a, v, d)
    uz = float(v.download()[:, 2].mean())
    b.rk4_init(pv, a, 0.01)
    b.rk4_step(10)
    D.synchronize()
    z = pv.download()[:, 2]
    assert abs(z.mean() - 10 * 0.01 * uz) < 1e-3 * abs(10 * 0.01 * uz)
    assert z.std() < 1e-4 * abs(z.mean())


def test_batch_autotune_records_its_shape(tmp_path):
    # the library's autotuner picks a shape, leaves the batch in it, and a
    # new batch of the same signature starts in it; the tuned evaluation
    # equals the per-cloud launches
    p, off, sig = small_clouds(16)
    n = len(p)
    pv, a = _dev(p)
    cv.lib().cvtx_hip_tune_clear()
    try:
        b = D.Batch(off, sig, "winckelmans")
        v, d = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
        case("batch autotune: 16 clouds")
        T, wg, wpc, ms = b.autotune(pv, a, v, d, reps=1)
        assert T in (1, 2, 4) and wg in (128, 256) and wpc in (2, 4, 8) and ms > 0
        assert b.shape() == (T, wg, wpc)
        b.eval(pv, a, v, d)
        v0, d0 = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
        b.eval_per_cloud(pv, a, v0, d0)
        assert_parity("tuned vel", v.download(), v0.download(), 1e-6)
        assert_parity("tuned dvort", d.download(), d0.download(), 1e-6)
        b2 = D.Batch(off, sig, "winckelmans")
        assert b2.shape() == (T, wg, wpc)
        path = str(tmp_path / "shapes.txt")
        assert cv.lib().cvtx_hip_tune_save(path.encode()) == 0
        cv.lib().cvtx_hip_tune_clear()
        assert cv.lib().cvtx_hip_tune_load(path.encode()) >= 1
        assert D.Batch(off, sig, "winckelmans").shape() == (T, wg, wpc)
        b.destroy()
        b2.destroy()
    finally:
        cv.lib().cvtx_hip_tune_clear()


@pytest.mark.parametrize("T", (1, 2, 4))
def test_batch_128_wide_workgroups(tmp_path, T):
    # a cached shape with 128-wide workgroups (128-source tiles): same results
    p, off, sig = small_clouds(10)
    n = len(p)
    sizes = np.diff(off).astype(np.int64)
    pairs = int((sizes * sizes).sum())
    line = f"3 1 3 1 {pairs.bit_length() - 1} {len(sig).bit_length() - 1} {T} 4 0.0 128\n"
    path = tmp_path / "shape128.txt"
    path.write_text(line)
    cv.lib().cvtx_hip_tune_clear()
    try:
        assert cv.lib().cvtx_hip_tune_load(str(path).encode()) == 1
        b = D.Batch(off, sig, "winckelmans")
        assert b.shape() == (T, 128, 4)
        pv, a = _dev(p)
        case(f"batch with 128-wide workgroups, T={T}")
        v, d = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
        b.eval(pv, a, v, d)
        v0, d0 = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
        b.eval_per_cloud(pv, a, v0, d0)
        assert_parity("128-wide vel", v.download(), v0.download(), 1e-6)
        assert_parity("128-wide dvort", d.download(), d0.download(), 1e-6)
        b.destroy()
    finally:
        cv.lib().cvtx_hip_tune_clear()


def test_batch_rejected_after_finalise():
    # a batch made before cvtx_finalise refers to freed device state: its
    # calls fail cleanly (no launch) until it is destroyed
    p, off, sig = small_clouds(4)
    n = len(p)
    pv, a = _dev(p)
    b = D.Batch(off, sig, "winckelmans")
    cv.finalise()
    try:
        v, d = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
        with pytest.raises(cv.CvtxError, match="finalise"):
            b.eval(pv, a, v, d)
    finally:
        cv.initialise()
        b.destroy()
    b2 = D.Batch(off, sig, "winckelmans")
    v, d = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    b2.eval(pv, a, v, d)
    b2.destroy()


def test_per_cloud_visc_batch_without_visc_buffer_is_an_error():
    p, off, sig = small_clouds(3)
    n = len(p)
    pv, a = _dev(p)
    b = D.Batch(off, sig, "winckelmans", outputs=cv.OUT_VEL | cv.OUT_DVORT | cv.OUT_VISC, nu=1e-3)
    v, d = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    with pytest.raises(cv.CvtxError, match="null"):
        b.eval_per_cloud(pv, a, v, d, visc=None)
    b.destroy()
