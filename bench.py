#!/usr/bin/env python3
"""Benchmarks of the gfx950 direct-sum Biot-Savart kernels on one GPU.

Default (the headline metric): P3D M2M fused vel + dvort interactions/s at
N = 2^20 -- the thin vortex ring of cvortex_amd.workloads.ring (Gaussian
regularisation sigma = 0.02, nu = 1e-3).  A step is one fused vel + dvort
evaluation of all N targets against all N sources through the
device-resident C ABI (cvtx_hip_p3d_eval), inputs already resident in GPU
memory, compensated accumulation.

Other workloads (--workload):
  cube     N = 8192 uniform cube, Winckelmans sigma = 0.05, vel only
  lattice  16384 straight filaments (128 x 128 sheet) on 2^20 particles,
           F3D vel, dvort and fused
  clouds   512 independent clouds (1024..65536 particles), fused vel + dvort,
           tile-compensated sums, one persistent launch vs one launch per
           cloud, RK4 steps replayed from a captured HIP graph
  sheet    N = 2^22 2D vortex sheet, planetary sigma = 4/sqrt(N), P2D vel
           (+ the viscous term, identically zero for planetary)
  small    per-call latency of the single-source / single-target entry
           points on the CPU leg and on the device

Each prints one JSON line with: interactions/s (pairs per second of the
timed work), the roofline of the pair kernel (algorithmic FP32 flops per
pair x pairs / kernel time, measured with HIP events on the kernel's
stream, against the 157.3 TFLOP/s FP32 vector peak), the relative error
against the FP64 oracle on sampled targets, and the reference's CPU path
(the oracle's restatement, all host threads) timed on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ring]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak
HBM_TBPS = 8.0
# Algorithmic FP32 flops per pair (add/sub/mul/cmp/select 1, FMA 2,
# transcendental 1), fixed per kernel, output set and regularisation.
FLOPS_PER_PAIR = {
    ("p3d", "vel", "winckelmans"): 31, ("p3d", "vel", "gaussian"): 48,
    ("p3d", "dvort", "gaussian"): 64,
    ("p3d", "vel+dvort", "singular"): 56, ("p3d", "vel+dvort", "planetary"): 58,
    ("p3d", "vel+dvort", "winckelmans"): 61, ("p3d", "vel+dvort", "gaussian"): 79,
    ("p3d", "vel+dvort+visc", "winckelmans"): 77, ("p3d", "vel+dvort+visc", "gaussian"): 94,
    ("f3d", "vel", "singular"): 56, ("f3d", "dvort", "singular"): 76,
    ("f3d", "vel+dvort", "singular"): 100,
    ("p2d", "vel", "planetary"): 14, ("p2d", "vel+visc", "gaussian"): 21,
    # vorticity field, counted with the same convention: d 3, box test 5
    # (3 compares, 2 ands), r^2 5, rho^2 1, zeta (exp 2), select 1, 3 FMA 6
    ("vort", "vort", "gaussian"): 23,
}
METRIC = "P3D M2M vel+dvort interactions/s at N=1M; % FP32 VALU peak; rel err vs FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def err_stats(g, r):
    import oracle
    rel = oracle.per_target_rel(g, r, floor=1e-30)
    return {"max": float(rel.max()) if rel.size else 0.0,
            "rms": float(np.sqrt(np.mean(rel * rel))) if rel.size else 0.0,
            "rel_l2": oracle.rel_l2(g, r)}


def roofline(pairs, F, kernel_ms):
    achieved = pairs * F / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel_ms": kernel_ms,
            "flops_per_pair": F}


def traffic_for(key, n):
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tpath):
        return None
    try:
        tj = json.load(open(tpath)).get(key)
        if tj and tj.get("n") == n:
            return tj["bytes_per_launch"]
    except Exception:
        return None
    return None


class Timer:
    """Wall clock and HIP events around `steps` calls of fn on `stream`."""

    def __init__(self, stream):
        self.stream = stream

    def run(self, fn, steps, warmup, per_step=True):
        """(wall ms per step, event ms per step) over `steps` calls after
        `warmup` untimed ones.  per_step: an event is recorded between
        consecutive steps and self.last_steps_ms holds each step's time;
        otherwise only around all of them (an event record costs the GPU a
        few microseconds, which matters for steps of tens of microseconds)."""
        from cvortex_amd import hip as H
        for _ in range(warmup):
            fn()
        H.stream_synchronize(self.stream)
        ev = [H.Event() for _ in range((steps if per_step else 1) + 1)]
        w0 = time.perf_counter()
        ev[0].record(self.stream)
        for k in range(steps):
            fn()
            if per_step:
                ev[k + 1].record(self.stream)
        if not per_step:
            ev[1].record(self.stream)
        ev[-1].synchronize()
        H.stream_synchronize(self.stream)
        wall = time.perf_counter() - w0
        self.last_steps_ms = ([ev[k].elapsed_ms(ev[k + 1]) for k in range(steps)] if per_step
                              else [])
        return wall / steps * 1e3, ev[0].elapsed_ms(ev[-1]) / steps


def n_err(args, default):
    """Error-sample count: the workload's default unless --err-samples is given."""
    return default if args.err_samples < 0 else args.err_samples


def cpu_time(fn, reps=2):
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best, out


def base_line(args, value, unit, ms_step, config, dtype="f32"):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": config}


# --------------------------------------------------------------------------
# C2: ring, fused P3D vel + dvort, N = 2^20 (the headline)
# --------------------------------------------------------------------------
def bench_ring(args, cv, D, W, o, stream):
    n = args.n or (1 << 20)
    p, sigma, nu = W.ring(n)
    dev = D.P3DDevice(p)
    tm = Timer(stream)
    pairs = float(n) * n

    def runner(outs, acc, visc_on):
        prob = dev.problem(outs, "gaussian", sigma, nu if visc_on else 0.0, acc)
        return lambda: D.p3d_eval(prob, stream)

    outs = dev.outputs(True, True, False)
    log(f"ring n={n}: fused vel+dvort, compensated")
    ms_step, ms_k = tm.run(runner(outs, cv.ACC_COMP, False), args.steps, args.warmup)
    steps_ms = list(tm.last_steps_ms)
    log(f"  {ms_step:.1f} ms/step (kernel {ms_k:.1f} ms, median {np.median(steps_ms):.1f} ms)")
    vel, dvort = outs[0].download(), outs[1].download()
    F = FLOPS_PER_PAIR[("p3d", "vel+dvort", "gaussian")]
    variants = {}
    plain = None
    if "plain" in args.variants:
        o2 = dev.outputs(True, True, False)
        s2, k2 = tm.run(runner(o2, cv.ACC_PLAIN, False), max(1, args.steps - 1), 1)
        variants["plain_fp32"] = {"interactions_per_s": pairs / (s2 * 1e-3), "kernel_ms": k2,
                                  "frac_of_peak": roofline(pairs, F, k2)["frac"]}
        plain = (o2[0].download(), o2[1].download())
    tiles = None
    if "acc" in args.variants:
        # FP32 tiles folded by TwoSum for every sum, no per-pair FP64: the
        # ring's near stretching terms are coherent, so this shows what the
        # per-pair FP64 sums of the default buy
        o5 = dev.outputs(True, True, False)
        s5, k5 = tm.run(runner(o5, cv.ACC_TILES, False), max(1, args.steps - 1), 1)
        variants["acc_tiles"] = {"interactions_per_s": pairs / (s5 * 1e-3), "kernel_ms": k5,
                                 "frac_of_peak": roofline(pairs, F, k5)["frac"]}
        tiles = (o5[0].download(), o5[1].download())
        log(f"  ACC_TILES: {k5:.1f} ms")
    visc = None
    if "visc" in args.variants:
        o3 = dev.outputs(True, True, True)
        s3, k3 = tm.run(runner(o3, cv.ACC_COMP, True), max(1, args.steps - 1), 1)
        F3 = FLOPS_PER_PAIR[("p3d", "vel+dvort+visc", "gaussian")]
        variants["vel_dvort_visc"] = {"interactions_per_s": pairs / (s3 * 1e-3), "kernel_ms": k3,
                                      "tflops": roofline(pairs, F3, k3)["achieved"],
                                      "frac_of_peak": roofline(pairs, F3, k3)["frac"]}
        visc = o3[2].download()
    if "single" in args.variants:
        # the velocity and the stretching alone (a caller of the reference
        # API makes two calls), device-resident, each with its own kernel
        for name, outs_, key in (("vel_only", (True, False, False), "vel"),
                                 ("dvort_only", (False, True, False), "dvort")):
            o4 = dev.outputs(*outs_)
            s4, k4 = tm.run(runner(o4, cv.ACC_COMP, False), max(1, args.steps - 1), 1)
            F4 = FLOPS_PER_PAIR[("p3d", "vel" if key == "vel" else "dvort", "gaussian")]
            variants[name] = {"interactions_per_s": pairs / (s4 * 1e-3), "kernel_ms": k4,
                              "flops_per_pair": F4, "frac_of_peak": roofline(pairs, F4, k4)["frac"]}
            log(f"  {name}: {k4:.1f} ms")
        variants["vel_then_dvort_over_fused"] = (variants["vel_only"]["kernel_ms"] +
                                                 variants["dvort_only"]["kernel_ms"]) / ms_k
    if "c_api" in args.variants:
        # the host-array boundary: gather from the caller's pointer arrays,
        # copies over PCIe both ways, kernels, all inside the timed call
        cv.P3D_M2M_fused(p, p, "gaussian", sigma, 0.0, vel=True, dvort=True)
        t0 = time.perf_counter()
        cv.P3D_M2M_fused(p, p, "gaussian", sigma, 0.0, vel=True, dvort=True)
        fused_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        cv.P3D_M2M_vel(p, p["coord"], "gaussian", sigma)
        vel_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        cv.P3D_M2M_dvort(p, p, "gaussian", sigma)
        dvort_s = time.perf_counter() - t0
        variants["c_api_host_arrays"] = {
            "fused_ms": fused_s * 1e3, "vel_ms": vel_s * 1e3, "dvort_ms": dvort_s * 1e3,
            "fused_interactions_per_s": pairs / fused_s,
            "vel_then_dvort_interactions_per_s": pairs / (vel_s + dvort_s)}
        log(f"  C API on host arrays: fused {fused_s * 1e3:.1f} ms, vel {vel_s * 1e3:.1f} ms, "
            f"dvort {dvort_s * 1e3:.1f} ms")
    err = None
    if n_err(args, 4096) > 0:
        rng = np.random.default_rng(0x5EED)
        idx = np.sort(rng.choice(n, size=min(n_err(args, 4096), n), replace=False)).astype(np.int32)
        log(f"FP64 oracle on {len(idx)} sampled targets...")
        v64, d64, e64 = o.f64_P3D(p, p, "gaussian", sigma, nu, idx=idx, gauss_mode=1,
                                  visc=visc is not None)
        err = {"samples": int(len(idx)), "vel": err_stats(vel[idx], v64),
               "dvort": err_stats(dvort[idx], d64)}
        if visc is not None:
            err["visc"] = err_stats(visc[idx], e64)
        if plain is not None:
            err["plain_fp32"] = {"vel": err_stats(plain[0][idx], v64),
                                 "dvort": err_stats(plain[1][idx], d64)}
        if tiles is not None:
            err["acc_tiles"] = {"vel": err_stats(tiles[0][idx], v64),
                                "dvort": err_stats(tiles[1][idx], d64)}
    cpu = None
    if not args.no_cpu:
        m = min(args.cpu_samples, n)
        rng = np.random.default_rng(0x5EED + 1)
        tidx = np.sort(rng.choice(n, size=m, replace=False))
        tg = np.ascontiguousarray(p[tidx])
        log(f"CPU baseline: {m} targets x {n} sources, vel then dvort")
        best, (rv, rd) = cpu_time(lambda: (o.P3D_M2M_vel(p, tg["coord"], "gaussian", sigma),
                                           o.P3D_M2M_dvort(p, tg, "gaussian", sigma)))
        import oracle
        cpu = {"value": m * float(n) / best, "unit": "interactions/s",
               "cores": int(o.L.orc_num_threads()), "kind": "port",
               "sample": f"{m} sampled ring targets x all {n} sources: cvtx_P3D_M2M_vel then "
                         f"cvtx_P3D_M2M_dvort as two passes (the reference has no fused path), "
                         f"best of 2, {best:.2f} s",
               "rel_l2_gpu_vs_cpu": {"vel": oracle.rel_l2(vel[tidx], rv),
                                     "dvort": oracle.rel_l2(dvort[tidx], rd)}}
        log(f"  {cpu['value']:.3e} interactions/s on {cpu['cores']} threads")
    line = base_line(args, pairs / (ms_step * 1e-3), "interactions/s", ms_step,
                     {"workload": "thin vortex ring (R=1, core 0.05), fused P3D vel+dvort, "
                                  "Gaussian sigma=0.02", "n_particles": n,
                      "regularisation": "gaussian", "sigma": sigma,
                      "accumulation": "compensated: FP32 tile sums + TwoSum; FP64 stretching "
                                      "sums for pairs within rho = 14.4; FP64 epilogue"})
    rf = roofline(pairs, F, ms_k)
    rf["kernel_ms_per_step"] = [round(t, 3) for t in steps_ms]
    rf["kernel_ms_median"] = float(np.median(steps_ms))
    rf["traffic"] = traffic_for("p3d_ring_vel_dvort_gaussian", n)
    line.update({"roofline": rf, "cpu_baseline": cpu, "rel_err_vs_fp64": err, "variants": variants})
    return line


# --------------------------------------------------------------------------
# C1: cube, P3D vel, N = 8192
# --------------------------------------------------------------------------
def bench_cube(args, cv, D, W, o, stream):
    n = args.n or 8192
    p, sigma = W.cube(n)
    dev = D.P3DDevice(p)
    tm = Timer(stream)
    pairs = float(n) * n
    vel = D.DeviceArray((n, 3))
    prob = D.p3d_problem(dev.pv, dev.a, n, dev.pv, None, n, vel, None, None, "winckelmans", sigma,
                         0.0, cv.ACC_COMP)
    # an evaluation is tens of microseconds: at least 200 timed steps after
    # 20 warm-ups, so the clock has settled and launch jitter averages out
    steps = max(args.steps, 200)
    ms_step, ms_k = tm.run(lambda: D.p3d_eval(prob, stream), steps, max(args.warmup, 20),
                           per_step=False)
    V = vel.download()
    F = FLOPS_PER_PAIR[("p3d", "vel", "winckelmans")]
    # end to end through the public C API (gather, copies, kernels)
    cv.P3D_M2M_vel(p, p["coord"], "winckelmans", sigma)
    t0 = time.perf_counter()
    for _ in range(10):
        Ve2e = cv.P3D_M2M_vel(p, p["coord"], "winckelmans", sigma)
    e2e_ms = (time.perf_counter() - t0) / 10 * 1e3
    v64 = o.f64_P3D_vel_pts(p, p["coord"], "winckelmans", sigma)
    import oracle
    cpu = None
    if not args.no_cpu:
        best, rv = cpu_time(lambda: o.P3D_M2M_vel(p, p["coord"], "winckelmans", sigma))
        cpu = {"value": pairs / best, "unit": "interactions/s",
               "cores": int(o.L.orc_num_threads()), "kind": "port",
               "sample": f"full {n} x {n} cvtx_P3D_M2M_vel, best of 2, {best:.2f} s",
               "rel_l2_gpu_vs_cpu": oracle.rel_l2(V, rv)}
    algo_bytes = n * 24 + n * 16 + n * 12
    line = base_line(args, pairs / (ms_step * 1e-3), "interactions/s", ms_step,
                     {"workload": "uniform cube [-1,1]^3, P3D vel, Winckelmans sigma=0.05",
                      "n_particles": n, "regularisation": "winckelmans", "sigma": sigma})
    line["steps"] = steps
    rf = roofline(pairs, F, ms_k)
    rf["hbm_gbps_algorithmic"] = algo_bytes / (ms_k * 1e-3) / 1e9
    # HBM bytes of one evaluation (pair kernel + partition reduction),
    # counter-measured; the small problem's rate against the 8 TB/s peak
    rf["traffic"] = traffic_for("p3d_cube_vel_winckelmans", n)
    if rf["traffic"]:
        rf["hbm_gbps_measured"] = rf["traffic"] / (ms_step * 1e-3) / 1e9
        rf["hbm_frac"] = rf["hbm_gbps_measured"] / 8000.0
    line.update({"roofline": rf, "end_to_end_c_api_ms": e2e_ms,
                 "rel_err_vs_fp64": {"samples": n, "vel": err_stats(V, v64),
                                     "c_api_vel": err_stats(Ve2e, v64)},
                 "cpu_baseline": cpu})
    if "plain" in args.variants:
        # one FP32 running sum per component (plain summation), next to the
        # compensated default
        vp = D.DeviceArray((n, 3))
        prob_p = D.p3d_problem(dev.pv, dev.a, n, dev.pv, None, n, vp, None, None, "winckelmans",
                               sigma, 0.0, cv.ACC_PLAIN)
        sp_, kp = tm.run(lambda: D.p3d_eval(prob_p, stream), steps, max(args.warmup, 20),
                         per_step=False)
        line["plain_fp32"] = {"interactions_per_s": pairs / (sp_ * 1e-3), "ms_per_step": sp_,
                              "frac_of_peak": roofline(pairs, F, kp)["frac"],
                              "rel_err_vs_fp64": err_stats(vp.download(), v64)}
        log(f"  plain FP32: {sp_ * 1e3:.1f} us")
    if "shapes" in args.variants:
        # explicit launch shapes (targets per lane, source partitions):
        # microseconds per evaluation over 200 back-to-back evaluations
        by_shape = {}
        for T in (1, 2, 4):
            for k in (8, 16, 32):
                _, kk = tm.run(lambda: D.p3d_eval(prob, stream, shape=(T, k)), 200, 10,
                               per_step=False)
                by_shape[f"T={T},partitions={k}"] = kk * 1e3
        line["us_by_shape"] = by_shape
        log("  " + ", ".join(f"{k}: {v:.1f} us" for k, v in by_shape.items()))
    return line


# --------------------------------------------------------------------------
# C3: lattice, F3D vel / dvort on 2^20 particles
# --------------------------------------------------------------------------
def bench_lattice(args, cv, D, W, o, stream):
    n = args.n or (1 << 20)
    f, t = W.lattice(n)
    nf = len(f)
    fa, fb = D.f3d_soa(f)
    tpv, ta = D.p3d_soa(t)
    dfa, dfb = D.DeviceArray.from_host(fa), D.DeviceArray.from_host(fb)
    dtp, dta = D.DeviceArray.from_host(tpv), D.DeviceArray.from_host(ta)
    vel, dv = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    tm = Timer(stream)
    pairs = float(nf) * n
    res = {}
    for name, v_, d_ in (("vel", vel, None), ("dvort", None, dv), ("vel+dvort", vel, dv)):
        prob = D.F3DProblem(dfa.addr, dfb.addr, nf, dtp.addr, dta.addr, n,
                            v_.addr if v_ else None, d_.addr if d_ else None, cv.ACC_COMP)
        s, k = tm.run(lambda: D.f3d_eval(prob, stream), args.steps, args.warmup)
        F = FLOPS_PER_PAIR[("f3d", name, "singular")]
        res[name] = {"interactions_per_s": pairs / (s * 1e-3), "ms_per_step": s,
                     "roofline": roofline(pairs, F, k)}
        log(f"  F3D {name}: {s:.2f} ms")
    V, Dv = vel.download(), dv.download()
    rng = np.random.default_rng(0x5EED + 3)
    idx = np.sort(rng.choice(n, size=min(n_err(args, 4096) or 4096, n), replace=False)).astype(np.int32)
    v64, d64 = o.f64_F3D(f, t, idx=idx)
    tg = np.ascontiguousarray(t[idx])
    m = min(args.cpu_samples, len(idx))
    import oracle
    cpu = None
    if not args.no_cpu:
        best, (rv, rd) = cpu_time(lambda: (o.F3D_M2M_vel(f, tg["coord"][:m]),
                                           o.F3D_M2M_dvort(f, tg[:m])))
        cpu = {"value": m * float(nf) / best, "unit": "interactions/s",
               "cores": int(o.L.orc_num_threads()), "kind": "port",
               "sample": f"{m} sampled targets x {nf} filaments: "
                         f"cvtx_F3D_M2M_vel then cvtx_F3D_M2M_dvort, best of 2",
               "rel_l2_gpu_vs_cpu": {"vel": oracle.rel_l2(V[idx[:m]], rv),
                                     "dvort": oracle.rel_l2(Dv[idx[:m]], rd)}}
    main = res["vel+dvort"]
    line = base_line(args, main["interactions_per_s"], "interactions/s", main["ms_per_step"],
                     {"workload": "128x128 horseshoe sheet of straight filaments on z=0 over "
                                  "[0,4]x[0,1], 2^20 particles in [0,6]x[-0.5,1.5]x[-0.5,0.5]",
                      "n_filaments": nf, "n_targets": n, "outputs": "F3D vel+dvort (fused)"})
    line.update({"roofline": main["roofline"], "separate": {k: v for k, v in res.items()},
                 "rel_err_vs_fp64": {"samples": int(len(idx)), "vel": err_stats(V[idx], v64),
                                     "dvort": err_stats(Dv[idx], d64)},
                 "cpu_baseline": cpu})
    return line


# --------------------------------------------------------------------------
# C4: 512 clouds, batched persistent launch, RK4 graph
# --------------------------------------------------------------------------
F_CLOUDS = FLOPS_PER_PAIR[("p3d", "vel+dvort", "winckelmans")]


def bench_clouds(args, cv, D, W, o, stream):
    n_clouds = args.n or 512
    p, off, sig = W.clouds(n_clouds)
    n = len(p)
    sizes = np.diff(off)
    pairs = float((sizes.astype(np.float64) ** 2).sum())
    pv, a = D.p3d_soa(p)
    dpv, da = D.DeviceArray.from_host(pv), D.DeviceArray.from_host(a)
    vel, dv = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    tm = Timer(stream)
    # the clouds' strengths are independent random vectors, so their near
    # stretching terms do not add up coherently: FP32 tile sums folded by
    # TwoSum (CVTX_HIP_ACC_TILES) hold the same error bar as the per-pair
    # FP64 sums of CVTX_HIP_ACC_COMP (both are reported)
    b = D.Batch(off, sig, "winckelmans", outputs=cv.OUT_VEL | cv.OUT_DVORT,
                accumulate=cv.ACC_TILES)
    # the library's autotuner picks the persistent launch shape (targets per
    # lane x workgroup size x resident workgroups per CU) and caches it
    shape0 = b.shape()
    t0 = time.perf_counter()
    tuned = b.autotune(dpv, da, vel, dv, reps=1)
    tune_s = time.perf_counter() - t0
    log(f"  autotuned shape T={tuned[0]} wg={tuned[1]} wg/cu={tuned[2]}: {tuned[3]:.1f} ms "
        f"(from {shape0}, {tune_s:.1f} s)")
    ms_step, ms_k = tm.run(lambda: b.eval(dpv, da, vel, dv, stream=stream), args.steps, args.warmup)
    V, Dv = vel.download(), dv.download()
    v2, d2 = D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    pc_step, pc_k = tm.run(lambda: b.eval_per_cloud(dpv, da, v2, d2, stream=stream), args.steps, 1)
    import oracle
    same = {"vel": oracle.rel_l2(V, v2.download()), "dvort": oracle.rel_l2(Dv, d2.download())}
    pp_step, _ = tm.run(lambda: b.eval_per_cloud(dpv, da, v2, d2, stream=stream, plain=True),
                        args.steps, 1)
    same_plain = {"vel": oracle.rel_l2(V, v2.download()), "dvort": oracle.rel_l2(Dv, d2.download())}
    log(f"  persistent {ms_step:.1f} ms, per-cloud tuned {pc_step:.1f} ms, "
        f"per-cloud plain {pp_step:.1f} ms")
    # RK4 graph
    rk = None
    if args.rk4_steps > 0:
        # the RK4 steps move the particles in place: on copies, so the
        # evaluations below see the original state
        rpv, ra = D.DeviceArray.from_host(pv), D.DeviceArray.from_host(a)
        b.rk4_init(rpv, ra, 1e-3)
        b.rk4_step(1, stream)  # warm
        s, k = tm.run(lambda: b.rk4_step(1, stream), args.rk4_steps, 0)
        rk = {"steps": args.rk4_steps, "ms_per_step": s, "gpu_ms_per_step": k,
              "interactions_per_s": 4 * pairs / (s * 1e-3), "dt": 1e-3}
        # the state after warm-up + timed steps: finite, and how far it moved
        xs, al = rpv.download(), ra.download()
        rk["state_finite"] = bool(np.all(np.isfinite(xs)) and np.all(np.isfinite(al)))
        rk["max_displacement"] = float(np.abs(xs[:, :3] - pv[:, :3]).max())
        log(f"  RK4 graph step {s:.1f} ms")
    # accuracy: 256 targets in each of 8 clouds spanning the size range
    order = np.argsort(sizes)
    pick = order[np.linspace(0, n_clouds - 1, 8).astype(int)]
    rng = np.random.default_rng(0x5EED + 4)
    errs_v, errs_d, idxs = [], [], []
    for c in pick:
        cl = np.ascontiguousarray(p[off[c]:off[c + 1]])
        idx = np.sort(rng.choice(len(cl), size=min(256, len(cl)), replace=False)).astype(np.int32)
        idxs.append(idx)
        v64, d64, _ = o.f64_P3D(cl, cl, "winckelmans", float(sig[c]), idx=idx)
        errs_v.append((V[off[c]:off[c + 1]][idx], v64))
        errs_d.append((Dv[off[c]:off[c + 1]][idx], d64))
    # CPU baseline: one evaluation of 8 clouds spanning the sizes up to 8192
    small = order[sizes[order] <= 8192]
    cpick = small[np.linspace(0, len(small) - 1, 8).astype(int)]
    subset = [(np.ascontiguousarray(p[off[c]:off[c + 1]]), float(sig[c])) for c in cpick]
    cpu_pairs = float(sum(len(cl) ** 2 for cl, _ in subset))
    gv = np.concatenate([e[0] for e in errs_v]); rv = np.concatenate([e[1] for e in errs_v])
    gd = np.concatenate([e[0] for e in errs_d]); rd = np.concatenate([e[1] for e in errs_d])
    comp = None
    if "acc" in args.variants:
        # the same batch with per-pair FP64 stretching sums
        bt = D.Batch(off, sig, "winckelmans", outputs=cv.OUT_VEL | cv.OUT_DVORT,
                     accumulate=cv.ACC_COMP)
        tt = bt.autotune(dpv, da, v2, d2, reps=1)
        t_step, t_k = tm.run(lambda: bt.eval(dpv, da, v2, d2, stream=stream),
                             max(1, args.steps - 1), 1)
        Vt, Dt = v2.download(), d2.download()
        tv = np.concatenate([Vt[off[c]:off[c + 1]][i] for c, i in zip(pick, idxs)])
        td = np.concatenate([Dt[off[c]:off[c + 1]][i] for c, i in zip(pick, idxs)])
        comp = {"what": "accumulate=CVTX_HIP_ACC_COMP: FP32 tile sums + TwoSum across tiles, "
                        "FP64 stretching sums per pair",
                 "ms_per_step": t_step, "interactions_per_s": pairs / (t_step * 1e-3),
                 "shape": list(tt[:3]), "roofline": roofline(pairs, F_CLOUDS, t_k),
                 "rel_err_vs_fp64": {"vel": err_stats(tv, rv), "dvort": err_stats(td, rd)}}
        log(f"  ACC_COMP {t_step:.1f} ms (shape {tt[:3]})")
        bt.destroy()
    cpu = None
    if not args.no_cpu:
        best_cpu, _ = cpu_time(lambda: [(o.P3D_M2M_vel(cl, cl["coord"], "winckelmans", s_),
                                         o.P3D_M2M_dvort(cl, cl, "winckelmans", s_))
                                        for cl, s_ in subset], 1)
        cpu = {"value": cpu_pairs / best_cpu, "unit": "interactions/s",
               "cores": int(o.L.orc_num_threads()), "kind": "port",
               "sample": "one evaluation (vel then dvort) of 8 clouds spanning "
                         f"sizes {min(len(c) for c, _ in subset)}.."
                         f"{max(len(c) for c, _ in subset)}, {cpu_pairs:.3g} "
                         f"pairs, {best_cpu:.2f} s"}
    line = base_line(args, pairs / (ms_step * 1e-3), "interactions/s", ms_step,
                     {"workload": "512 clouds, sizes log-uniform in [1024, 65536], unit cubes, "
                                  "Winckelmans sigma_c ~ U[0.01, 0.1], fused vel+dvort",
                      "accumulation": "tile-compensated (CVTX_HIP_ACC_TILES): FP32 tile sums "
                                      "+ TwoSum across tiles, FP64 epilogue",
                      "n_clouds": n_clouds, "n_particles": n, "pairs_per_eval": pairs,
                      "shape": {"targets_per_lane": tuned[0], "workgroup_size": tuned[1],
                                "workgroups_per_cu": tuned[2], "tuned_by": "cvtx_hip_batch_autotune",
                                "tuning_s": tune_s, "shape_before_tuning": list(shape0)}})
    rf = roofline(pairs, F_CLOUDS, ms_k)
    rf["traffic"] = traffic_for("p3d_clouds_vel_dvort_winckelmans", n)
    line.update({"roofline": rf,
                 "per_cloud_plain_launches": {
                     "what": "one launch per cloud, 1 target per lane, no source partitions, "
                             "stream synchronised after each cloud",
                     "ms_per_step": pp_step, "interactions_per_s": pairs / (pp_step * 1e-3),
                     "speedup_of_persistent": pp_step / ms_step,
                     "rel_l2_persistent_vs_per_cloud": same_plain},
                 "per_cloud_tuned_launches": {
                     "what": "one launch per cloud with its default shape (targets per lane, "
                             "source partitions), queued back to back",
                     "ms_per_step": pc_step, "interactions_per_s": pairs / (pc_step * 1e-3),
                     "speedup_of_persistent": pc_step / ms_step,
                     "rel_l2_persistent_vs_per_cloud": same},
                 "rk4_graph": rk,
                 "rel_err_vs_fp64": {"samples": int(len(gv)), "clouds": [int(c) for c in pick],
                                     "vel": err_stats(gv, rv), "dvort": err_stats(gd, rd)},
                 "acc_comp": comp,
                 "cpu_baseline": cpu})
    b.destroy()
    return line


# --------------------------------------------------------------------------
# C5: sheet, P2D vel (+ visc) at N = 2^22
# --------------------------------------------------------------------------
def bench_sheet(args, cv, D, W, o, stream):
    n = args.n or (1 << 22)
    p, sigma, nu = W.sheet(n)
    s = D.p2d_soa(p)
    ds = D.DeviceArray.from_host(s)
    vel, vs = D.DeviceArray((n, 2)), D.DeviceArray((n,))
    tm = Timer(stream)
    pairs = float(n) * n
    prob = D.P2DProblem(ds.addr, n, ds.addr, n, vel.addr, None, 2, sigma, nu, cv.ACC_COMP)
    ms_step, ms_k = tm.run(lambda: D.p2d_eval(prob, stream), args.steps, args.warmup)
    V = vel.download()
    variants = {}
    if "plain" in args.variants:
        v2 = D.DeviceArray((n, 2))
        prob2 = D.P2DProblem(ds.addr, n, ds.addr, n, v2.addr, None, 2, sigma, nu, cv.ACC_PLAIN)
        s2, k2 = tm.run(lambda: D.p2d_eval(prob2, stream), max(1, args.steps - 1), 1)
        variants["plain_fp32"] = {"interactions_per_s": pairs / (s2 * 1e-3), "kernel_ms": k2}
        plainV = v2.download()
    fp64V = None
    if "fp64final" in args.variants:
        # compensated tiles in 8 source partitions, the partitions added in
        # FP64 by the reduction kernel (two targets per lane)
        v4 = D.DeviceArray((n, 2))
        prob4 = D.P2DProblem(ds.addr, n, ds.addr, n, v4.addr, None, 2, sigma, nu, cv.ACC_COMP)
        s4, k4 = tm.run(lambda: D.p2d_eval(prob4, stream, shape=(2, 8)), max(1, args.steps - 1), 1)
        variants["compensated_fp64_partitions"] = {"interactions_per_s": pairs / (s4 * 1e-3),
                                                   "kernel_ms": k4, "shape": "T=2, 8 partitions"}
        fp64V = v4.download()
    if "gaussian" in args.variants:
        # supplementary: the same sheet with the Gaussian kernel, vel + visc fused
        vg, eg = D.DeviceArray((n, 2)), D.DeviceArray((n,))
        prob3 = D.P2DProblem(ds.addr, n, ds.addr, n, vg.addr, eg.addr, 3, sigma, nu, cv.ACC_COMP)
        s3, k3 = tm.run(lambda: D.p2d_eval(prob3, stream), max(1, args.steps - 1), 1)
        F3 = FLOPS_PER_PAIR[("p2d", "vel+visc", "gaussian")]
        variants["gaussian_vel_visc"] = {"interactions_per_s": pairs / (s3 * 1e-3), "kernel_ms": k3,
                                         "frac_of_peak": roofline(pairs, F3, k3)["frac"]}
        log(f"  Gaussian vel+visc: {k3:.1f} ms")
    rng = np.random.default_rng(0x5EED + 5)
    idx = np.sort(rng.choice(n, size=min(n_err(args, 8192) or 8192, n), replace=False)).astype(np.int32)
    v64, _ = o.f64_P2D(p, p, "planetary", sigma, idx=idx)
    err = {"samples": int(len(idx)), "vel": err_stats(V[idx], v64)}
    if "plain" in args.variants:
        err["plain_fp32"] = {"vel": err_stats(plainV[idx], v64)}
    if fp64V is not None:
        err["compensated_fp64_partitions"] = {"vel": err_stats(fp64V[idx], v64)}
    # visc: planetary has no eta -> exactly zero (reference VortFunc.cpp:52-62)
    vz = cv.P2D_M2M_visc_dvort(p[:4096], p[:4096], "planetary", sigma, nu)
    m = min(args.cpu_samples, n)
    tg = np.ascontiguousarray(p[idx[:m]])
    import oracle
    cpu = None
    if not args.no_cpu:
        best, rv = cpu_time(lambda: o.P2D_M2M_vel(p, tg["coord"], "planetary", sigma))
        cpu = {"value": m * float(n) / best, "unit": "interactions/s",
               "cores": int(o.L.orc_num_threads()), "kind": "port",
               "sample": f"{m} sampled targets x {n} sources cvtx_P2D_M2M_vel, "
                         f"best of 2, {best:.2f} s",
               "rel_l2_gpu_vs_cpu": oracle.rel_l2(V[idx[:m]], rv)}
    F = FLOPS_PER_PAIR[("p2d", "vel", "planetary")]
    line = base_line(args, pairs / (ms_step * 1e-3), "interactions/s", ms_step,
                     {"workload": "2D vortex sheet y=0.01 sin(2 pi x), P2D vel, planetary "
                                  "sigma=4/sqrt(N)", "n_particles": n, "sigma": sigma})
    line.update({"roofline": roofline(pairs, F, ms_k), "variants": variants,
                 "rel_err_vs_fp64": err, "visc_exactly_zero": bool(np.all(vz == 0)),
                 "cpu_baseline": cpu})
    return line


# --------------------------------------------------------------------------
# Vorticity field + Pedrizzetti relaxation on the ring, N = 2^20
# --------------------------------------------------------------------------
def bench_relax(args, cv, D, W, o, stream):
    n = args.n or (1 << 20)
    p, sigma, nu = W.ring(n)
    fdt = 0.1
    pv, a = D.p3d_soa(p)
    dpv, da = D.DeviceArray.from_host(pv), D.DeviceArray.from_host(a)
    om = D.DeviceArray((n, 3))
    tm = Timer(stream)
    reg = 3  # gaussian
    prob = D.VortProblem(dpv.addr, da.addr, n, dpv.addr, n, om.addr, reg, sigma, cv.ACC_COMP)
    log(f"relax ring n={n}: vorticity field (sorted, tile skipping)")
    ms_vort, k_vort = tm.run(lambda: D.vort_eval(prob, stream), args.steps, args.warmup)
    W_gpu = om.download()
    # one relaxation step on a copy of the strengths (the timed steps update
    # the copy in place; the step count does not change the work)
    da2 = D.DeviceArray.from_host(a)
    ms_rel, k_rel = tm.run(lambda: D.relax_eval(dpv, da2, n, fdt, "gaussian", sigma, stream=stream),
                           args.steps, 0)
    da3 = D.DeviceArray.from_host(a)
    D.relax_eval(dpv, da3, n, fdt, "gaussian", sigma, stream=stream)
    D.synchronize()
    A_gpu = da3.download()[:, :3]
    # targets per lane and source partitions around the default shape
    # (one target per lane, 64 partitions)
    sweep = {}
    for T, k in ((1, 1), (1, 16), (1, 64), (2, 64), (4, 64)):
        _, kT = tm.run(lambda: D.vort_eval(prob, stream, shape=(T, k)), 2, 1)
        sweep[f"T={T},partitions={k}"] = kT
    log(f"  sorted vort by launch shape: {sweep}")
    # brute force: the same kernel without the spatial sort (every tile evaluated)
    L = cv.lib()
    prev = L.cvtx_hip_set_target_order(0)
    try:
        om2 = D.DeviceArray((n, 3))
        prob2 = D.VortProblem(dpv.addr, da.addr, n, dpv.addr, n, om2.addr, reg, sigma, cv.ACC_COMP)
        ms_b, k_b = tm.run(lambda: D.vort_eval(prob2, stream), max(1, args.steps - 3), 1)
    finally:
        L.cvtx_hip_set_target_order(prev)
    log(f"  vort {k_vort:.2f} ms (sorted) vs {k_b:.1f} ms (brute force); relax step {ms_rel:.2f} ms")
    import oracle
    same = oracle.rel_l2(om2.download(), W_gpu)
    rng = np.random.default_rng(0x5EED + 6)
    idx = np.sort(rng.choice(n, size=min(n_err(args, 2048) or 2048, n), replace=False)).astype(np.int32)
    w64 = o.f64_P3D_vort(p, p["coord"], "gaussian", sigma, idx=idx)
    a64 = p["vorticity"][idx].astype(np.float64)
    aw = np.linalg.norm(w64, axis=1, keepdims=True)
    aa = np.linalg.norm(a64, axis=1, keepdims=True)
    r64 = np.where(aw > 0, (1 - fdt) * a64 + fdt * aa / np.where(aw > 0, aw, 1) * w64, 0.0)
    err = {"samples": int(len(idx)), "vort": err_stats(W_gpu[idx], w64),
           "relaxed_strength": err_stats(A_gpu[idx], r64)}
    cpu = None
    if not args.no_cpu:
        m = min(args.cpu_samples, n)
        tg = np.ascontiguousarray(p["coord"][idx[:m]])
        best, rv = cpu_time(lambda: o.P3D_M2M_vort(p, tg, "gaussian", sigma))
        cpu = {"value": m / best, "unit": "targets/s", "cores": int(o.L.orc_num_threads()),
               "kind": "port",
               "sample": f"cvtx_P3D_M2M_vort at {m} sampled ring particles x all {n} sources "
                         f"(box cut-off test per pair), best of 2, {best:.2f} s",
               "rel_l2_gpu_vs_cpu": oracle.rel_l2(W_gpu[idx[:m]], rv),
               "rel_l2_cpu_vs_fp64": oracle.rel_l2(rv, w64[:m])}
    pairs = float(n) * n
    F = FLOPS_PER_PAIR[("vort", "vort", "gaussian")]
    line = base_line(args, n / (ms_rel * 1e-3), "particles/s", ms_rel,
                     {"workload": "Pedrizzetti relaxation step on the thin vortex ring (R=1, core "
                                  "0.05), Gaussian sigma=0.02, fdt=0.1: vorticity field with the "
                                  "5 sigma box cut-off, then the strength update",
                      "n_particles": n})
    line.update({"roofline": dict(roofline(pairs, F, k_b), note="brute-force kernel (no sort): "
                                  "all 2^40 pairs evaluated"),
                 "vort_sorted_ms": k_vort, "vort_brute_force_ms": k_b,
                 "vort_sorted_ms_by_shape": sweep,
                 "speedup_of_tile_skipping": k_b / k_vort,
                 "rel_l2_sorted_vs_brute_force": same,
                 "rel_err_vs_fp64": err, "cpu_baseline": cpu})
    return line


# --------------------------------------------------------------------------
# F3D influence matrix of the 128 x 128 lattice: 16384 bound segments x 16384
# collocation points (3/4 chord, normal +z)
# --------------------------------------------------------------------------
def bench_influence(args, cv, D, W, o, stream):
    f, _ = W.lattice(n_targets=1)
    nf = len(f)
    nx, ny = 128, 128
    dx, dy = 4.0 / nx, 1.0 / ny
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pts = np.zeros((nx * ny, 4), np.float32)
    pts[:, 0] = ii.ravel() * dx + 0.75 * dx
    pts[:, 1] = (jj.ravel() + 0.5) * dy
    dirs = np.zeros((nx * ny, 4), np.float32)
    dirs[:, 2] = 1.0
    nm = len(pts)
    fa, fb = D.f3d_soa(f)
    dfa, dfb, dm, de = (D.DeviceArray.from_host(x) for x in (fa, fb, pts, dirs))
    out = D.DeviceArray((nm, nf))
    tm = Timer(stream)
    log(f"influence matrix {nm} x {nf} ({4.0 * nm * nf / 2**30:.2f} GiB)")
    ms, k = tm.run(lambda: D.f3d_inf_eval(dfa, dfb, nf, dm, de, nm, out, stream), args.steps,
                   args.warmup)
    log(f"  {ms:.3f} ms/step (kernel {k:.3f} ms)")
    rng = np.random.default_rng(0x5EED + 7)
    rows = np.sort(rng.choice(nm, size=min(n_err(args, 64) or 64, nm), replace=False)).astype(np.int32)
    M = out.download()
    m64 = o.f64_F3D_inf_mtrx(f, pts[:, :3], dirs[:, :3], idx=rows)
    err = {"rows": int(len(rows)), "matrix": err_stats(M[rows], m64)}
    cpu = None
    if not args.no_cpu:
        # a block of rows large enough (about 0.1 s on 16 threads)
        # that thread start-up does not dominate the time
        blk = np.arange(min(4096, nm))
        best, rv = cpu_time(lambda: o.F3D_inf_mtrx(f, pts[blk, :3], dirs[blk, :3]))
        sub = rows[:16]
        r16 = o.F3D_inf_mtrx(f, pts[sub, :3], dirs[sub, :3])
        import oracle
        cpu = {"value": len(blk) * float(nf) / best, "unit": "matrix elements/s",
               "cores": int(o.L.orc_num_threads()), "kind": "port",
               "sample": f"cvtx_F3D_inf_mtrx on the first {len(blk)} rows x {nf} filaments, "
                         "best of 2",
               "rel_l2_gpu_vs_cpu": oracle.rel_l2(M[blk], rv),
               "rel_l2_cpu_vs_fp64": oracle.rel_l2(r16, m64[:len(sub)])}
    elems = float(nm) * nf
    out_bytes = 4.0 * elems
    in_bytes = 32.0 * nf + 32.0 * nm
    achieved = (out_bytes + in_bytes) / (k * 1e-3) / 1e9
    line = base_line(args, elems / (ms * 1e-3), "matrix elements/s", ms,
                     {"workload": "F3D influence matrix of the 128x128 bound-segment lattice "
                                  "([0,4]x[0,1], z=0) at its 3/4-chord collocation points, "
                                  "normals +z", "n_filaments": nf, "n_points": nm})
    line.update({"roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0,
                              "unit": "GB/s", "frac": achieved / 8000.0, "traffic": None,
                              "kernel_ms": k, "bytes_per_launch": out_bytes + in_bytes},
                 "rel_err_vs_fp64": err, "cpu_baseline": cpu})
    return line


# --------------------------------------------------------------------------
# Remeshing: the ring's 2^20 particles redistributed onto a grid (m4')
# --------------------------------------------------------------------------
def bench_remesh(args, cv, D, W, o, stream):
    n = args.n or (1 << 20)
    p, sigma, nu = W.ring(n)
    h = 0.01
    log(f"remesh ring n={n}: m4' onto h={h}")
    # one call as a caller with a preallocated output array makes it:
    # gather, copies, device pipeline, results in host records
    import ctypes
    from cvortex_amd.types import P3D_DTYPE, pointer_array
    ptrs, keep = pointer_array(p)
    buf = np.zeros(2 * n, P3D_DTYPE)
    rf = cv.redistfunc("m4p")
    fn = cv.lib().cvtx_P3D_redistribute_on_grid
    call = lambda: fn(ctypes.c_void_p(ptrs.ctypes.data), n, ctypes.c_void_p(buf.ctypes.data), len(buf),
                      ctypes.byref(rf), h, 0.01)
    for _ in range(args.warmup):
        call()
    t = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        m_out = call()
        t.append((time.perf_counter() - t0) * 1e3)
    if m_out < 0:
        raise RuntimeError(f"remeshing failed: {cv.last_error()}")
    out = buf[:m_out]
    ms = float(np.mean(t))
    log(f"  {ms:.1f} ms per call, {len(out)} particles created")
    a0 = p["vorticity"].astype(np.float64).sum(0)
    a1 = out["vorticity"].astype(np.float64).sum(0)
    x0 = (p["coord"].astype(np.float64).T @ p["vorticity"].astype(np.float64))
    x1 = (out["coord"].astype(np.float64).T @ out["vorticity"].astype(np.float64))
    cons = {"total_strength_rel": float(np.abs(a1 - a0).max() / np.abs(p["vorticity"]).sum()),
            "first_moment_rel": float(np.abs(x1 - x0).max() / np.abs(x0).max())}
    cpu = None
    if not args.no_cpu:
        m = min(65536, n)
        sub = np.ascontiguousarray(p[:: max(1, n // m)][:m])
        best, ref = cpu_time(lambda: o.P3D_redistribute_on_grid(sub, "m4p", h, negligible=0.01), 1)
        got = cv.P3D_redistribute_on_grid(sub, "m4p", h, negligible=0.01)
        # node identity without removal (removal compares FP32-summed strengths
        # with a threshold, so nodes at the threshold may go either way)
        ref0 = o.P3D_redistribute_on_grid(sub, "m4p", h)
        got0 = cv.P3D_redistribute_on_grid(sub, "m4p", h)
        import oracle
        same = bool(len(got0) == len(ref0) and np.array_equal(got0["coord"], ref0["coord"]))
        cpu = {"value": m / best, "unit": "particles/s", "cores": 1, "kind": "port",
               "sample": f"cvtx_P3D_redistribute_on_grid of every {max(1, n // m)}th ring particle "
                         f"({m}), one thread as the restated reference tree, {best:.2f} s",
               "same_nodes_as_gpu_without_removal": same,
               "rel_l2_gpu_vs_cpu_without_removal":
                   oracle.rel_l2(got0["vorticity"], ref0["vorticity"]) if same else None,
               "created_gpu_vs_cpu_with_removal": [int(len(got)), int(len(ref))]}
    line = base_line(args, n / (ms * 1e-3), "particles/s", ms,
                     {"workload": "remeshing of the thin vortex ring (2^20 particles) onto a grid of "
                                  "spacing 0.01 with the m4' kernel, negligible-strength removal "
                                  "at 0.01 x mean; host arrays in and out", "n_particles": n,
                      "n_created": int(len(out)), "entries": int(n * 125)})
    line.update({"roofline": None, "conservation": cons, "cpu_baseline": cpu})
    return line


# --------------------------------------------------------------------------
# single-source / single-target forms (S2S, S2M, M2S): latency per call
# --------------------------------------------------------------------------
def bench_small(args, cv, D, W, o, stream):
    """Per-call latency of the reference's single-source and single-target
    entry points through the public C ABI, on the library's CPU leg and on
    the device (cvtx_hip_set_dispatch), next to the oracle's restatement of
    the reference CPU code.  Each figure is the median over batches of
    calls made from Python through ctypes; `ctypes_call_us` is the cost of a
    trivial library call the same way, included in every figure."""
    import ctypes
    from cvortex_amd.types import P3D, F3D, P2D, V2f, V3f, REG_ID
    L = cv.lib()
    rng = np.random.default_rng(0x5EED + 9)

    def per_call_us(fn, calls):
        best = []
        for _ in range(5):
            t0 = time.perf_counter()
            for _ in range(calls):
                fn()
            best.append((time.perf_counter() - t0) / calls * 1e6)
        return float(np.median(best))

    base_us = per_call_us(lambda: L.cvtx_hip_get_dispatch(), 20000)
    p = W.random_p3d(2, 11, box=1.0, vort=1.0)
    f = W.random_f3d(1, 12, box=1.0, strength=1.0)
    q = W.random_p2d(2, 13, box=1.0, vort=1.0)
    rp = [P3D.from_buffer_copy(np.ascontiguousarray(p[i]).tobytes()) for i in range(2)]
    rf = F3D.from_buffer_copy(np.ascontiguousarray(f[0]).tobytes())
    rq = [P2D.from_buffer_copy(np.ascontiguousarray(q[i]).tobytes()) for i in range(2)]
    mes3 = V3f((ctypes.c_float * 3)(0.3, 0.4, 0.5))
    mes2 = V2f((ctypes.c_float * 2)(0.3, 0.4))
    vf = cv.vortfunc("winckelmans")
    kp = ctypes.byref(vf)
    w = REG_ID["winckelmans"]
    s2s = {
        "P3D_S2S_vel": (lambda: L.cvtx_P3D_S2S_vel(ctypes.byref(rp[0]), mes3, kp, 0.1),
                        lambda: o.L.orc_P3D_S2S_vel(ctypes.byref(rp[0]), mes3, w, 0.1)),
        "P3D_S2S_dvort": (lambda: L.cvtx_P3D_S2S_dvort(ctypes.byref(rp[0]), ctypes.byref(rp[1]), kp, 0.1),
                          lambda: o.L.orc_P3D_S2S_dvort(ctypes.byref(rp[0]), ctypes.byref(rp[1]), w, 0.1)),
        "P3D_S2S_visc_dvort": (lambda: L.cvtx_P3D_S2S_visc_dvort(ctypes.byref(rp[0]), ctypes.byref(rp[1]), kp, 0.1, 0.01),
                               lambda: o.L.orc_P3D_S2S_visc_dvort(ctypes.byref(rp[0]), ctypes.byref(rp[1]), w, 0.1, 0.01)),
        "F3D_S2S_vel": (lambda: L.cvtx_F3D_S2S_vel(ctypes.byref(rf), mes3),
                        lambda: o.L.orc_F3D_S2S_vel(ctypes.byref(rf), mes3)),
        "P2D_S2S_vel": (lambda: L.cvtx_P2D_S2S_vel(ctypes.byref(rq[0]), mes2, kp, 0.1),
                        lambda: o.L.orc_P2D_S2S_vel(ctypes.byref(rq[0]), mes2, w, 0.1)),
    }
    out = {"ctypes_call_us": base_us, "s2s": {}, "s2m_m2s": {}}
    for name, (lib_fn, ref_fn) in s2s.items():
        row = {}
        for mode, label in ((cv.DISPATCH_AUTO, "library_auto_us"), (cv.DISPATCH_DEVICE, "library_device_us")):
            cv.set_dispatch(mode)
            lib_fn()
            row[label] = per_call_us(lib_fn, 2000 if mode == cv.DISPATCH_AUTO else 100)
            if mode == cv.DISPATCH_AUTO:
                row["leg"] = "host" if cv.last_leg() == cv.LEG_HOST else "device"
        cv.set_dispatch(cv.DISPATCH_AUTO)
        row["oracle_cpu_us"] = per_call_us(ref_fn, 2000)
        out["s2s"][name] = row
        log(f"  {name}: auto {row['library_auto_us']:.2f} us ({row['leg']}), device "
            f"{row['library_device_us']:.1f} us, oracle {row['oracle_cpu_us']:.2f} us")
    # S2M (one source, n targets) and M2S (n sources, one target): both legs
    for n in (1024, 16384, 65536, 262144, 1 << 20):
        src = W.random_p3d(n, 20 + n.bit_length(), box=1.0, vort=1.0)
        pts = np.ascontiguousarray(src["coord"])
        rows = {}
        for mode, label in ((cv.DISPATCH_HOST, "host_ms"), (cv.DISPATCH_DEVICE, "device_ms")):
            cv.set_dispatch(mode)
            cv.P3D_S2M_vel(src[0], pts, "winckelmans", 0.05)
            cv.P3D_M2S_vel(src, pts[0], "winckelmans", 0.05)
            reps = 3
            rows.setdefault("S2M_vel", {})[label] = per_call_us(
                lambda: cv.P3D_S2M_vel(src[0], pts, "winckelmans", 0.05), reps) / 1e3
            rows.setdefault("M2S_vel", {})[label] = per_call_us(
                lambda: cv.P3D_M2S_vel(src, pts[0], "winckelmans", 0.05), reps) / 1e3
        cv.set_dispatch(cv.DISPATCH_AUTO)
        cv.P3D_M2S_vel(src, pts[0], "winckelmans", 0.05)
        rows["auto_leg"] = "host" if cv.last_leg() == cv.LEG_HOST else "device"
        out["s2m_m2s"][str(n)] = rows
        log(f"  n={n}: S2M host {rows['S2M_vel']['host_ms']:.3f} ms, device "
            f"{rows['S2M_vel']['device_ms']:.3f} ms; M2S host {rows['M2S_vel']['host_ms']:.3f} ms, "
            f"device {rows['M2S_vel']['device_ms']:.3f} ms; auto: {rows['auto_leg']}")
    s2s_auto = out["s2s"]["P3D_S2S_vel"]["library_auto_us"]
    line = base_line(args, 1e6 / s2s_auto, "calls/s", s2s_auto * 1e-3,
                     {"workload": "single-source / single-target entry points (S2S, S2M, M2S) "
                                  "through the public C ABI, Winckelmans sigma 0.1 / 0.05",
                      "host_threads": cv.host_threads()})
    line.update({"roofline": None, "latency": out, "cpu_baseline": None})
    return line


WORKLOADS = {"ring": bench_ring, "cube": bench_cube, "lattice": bench_lattice,
             "clouds": bench_clouds, "sheet": bench_sheet, "relax": bench_relax,
             "influence": bench_influence, "remesh": bench_remesh, "small": bench_small}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ring", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="size override (clouds: number of clouds)")
    ap.add_argument("--err-samples", type=int, default=-1,
                    help="targets checked against the FP64 oracle (default: per workload)")
    ap.add_argument("--cpu-samples", type=int, default=512)
    ap.add_argument("--rk4-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variants", default="plain,acc,visc,single,c_api,fp64final,gaussian")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; the direct sum would shard its targets "
              "over GPUs with one all-gather of the sources per evaluation. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import cvortex_amd as cv
    from cvortex_amd import device as D
    from cvortex_amd import workloads as W
    import oracle

    cv.initialise()
    if cv.num_accelerators() < 1:
        log("bench.py: no HIP device")
        sys.exit(1)
    log("device:", cv.accelerator_name(0))
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    o = oracle.default()
    stream = cv.lib().cvtx_hip_stream()
    line = WORKLOADS[args.workload](args, cv, D, W, o, stream)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
