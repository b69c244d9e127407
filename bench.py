#!/usr/bin/env python3
"""Benchmark: P3D M2M fused vel + dvort interactions/s at N = 2^20 on one GPU.

Workload (default): the thin vortex ring of cvortex_amd.workloads.ring --
N = 1,048,576 particles, Gaussian regularisation sigma = 0.02, nu = 1e-3.
A step is one fused vel + dvort evaluation of all N targets against all N
sources through the device-resident C ABI (cvtx_hip_p3d_eval), inputs
already resident in GPU memory, compensated accumulation.  The line also
reports the plain-FP32 and the +visc variants on the same inputs, the
relative error against the FP64 oracle on sampled targets, the roofline of
the pair kernel (algorithmic FP32 flops per pair x pairs / kernel time,
against the 157.3 TFLOP/s FP32 vector peak), and the reference CPU path
(oracle restatement, all host cores) timed on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
                       [--err-samples S] [--cpu-samples M] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (packed FP32)
# Algorithmic FP32 flops per pair (add/sub/mul/cmp/select 1, FMA 2,
# transcendental 1), frozen per kernel and regularisation.
FLOPS_PER_PAIR = {
    ("p3d", "vel+dvort", "singular"): 56, ("p3d", "vel+dvort", "planetary"): 58,
    ("p3d", "vel+dvort", "winckelmans"): 61, ("p3d", "vel+dvort", "gaussian"): 79,
    ("p3d", "vel+dvort+visc", "winckelmans"): 77, ("p3d", "vel+dvort+visc", "gaussian"): 94,
    ("p3d", "vel", "winckelmans"): 31, ("p3d", "vel", "gaussian"): 48,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--err-samples", type=int, default=1024)
    ap.add_argument("--cpu-samples", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variants", default="plain,visc")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the direct sum shards over targets "
              "with one source all-gather per evaluation; see DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import cvortex_amd as cv
    from cvortex_amd import device as D
    from cvortex_amd import hip as H
    from cvortex_amd import workloads as W

    cv.initialise()
    if cv.num_accelerators() < 1:
        log("bench.py: no HIP device")
        sys.exit(1)
    log("device:", cv.accelerator_name(0))

    n = args.n
    t0 = time.time()
    p, sigma, nu = W.ring(n)
    log(f"workload ring n={n} built in {time.time() - t0:.1f}s")
    dev = D.P3DDevice(p)
    stream = cv.lib().cvtx_hip_stream()

    def run(outs, acc, visc_on, steps, warmup):
        prob = dev.problem(outs, "gaussian", sigma, nu if visc_on else 0.0, acc)
        for _ in range(warmup):
            D.p3d_eval(prob, stream)
        H.stream_synchronize(stream)
        e0, e1 = H.Event(), H.Event()
        w0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            D.p3d_eval(prob, stream)
        e1.record(stream)
        e1.synchronize()
        H.stream_synchronize(stream)
        wall = time.perf_counter() - w0
        return wall / steps * 1e3, e0.elapsed_ms(e1) / steps

    pairs = float(n) * float(n)
    outs = dev.outputs(True, True, False)
    log("timing fused vel+dvort (compensated)...")
    ms_step, ms_kernel = run(outs, cv.ACC_COMP, False, args.steps, args.warmup)
    log(f"  {ms_step:.1f} ms/step, kernel {ms_kernel:.1f} ms")
    vel = outs[0].download()
    dvort = outs[1].download()
    F = FLOPS_PER_PAIR[("p3d", "vel+dvort", "gaussian")]
    achieved = pairs * F / (ms_kernel * 1e-3) / 1e12

    variants = {}
    vlist = [v for v in args.variants.split(",") if v]
    if "plain" in vlist:
        log("timing plain FP32 accumulation...")
        o2 = dev.outputs(True, True, False)
        s2, k2 = run(o2, cv.ACC_PLAIN, False, max(1, args.steps - 1), 1)
        variants["plain_fp32"] = {"interactions_per_s": pairs / (s2 * 1e-3), "kernel_ms": k2}
        plain_vel, plain_dvort = o2[0].download(), o2[1].download()
    if "visc" in vlist:
        log("timing fused vel+dvort+visc...")
        o3 = dev.outputs(True, True, True)
        s3, k3 = run(o3, cv.ACC_COMP, True, max(1, args.steps - 1), 1)
        F3 = FLOPS_PER_PAIR[("p3d", "vel+dvort+visc", "gaussian")]
        variants["vel_dvort_visc"] = {"interactions_per_s": pairs / (s3 * 1e-3), "kernel_ms": k3,
                                      "tflops": pairs * F3 / (k3 * 1e-3) / 1e12,
                                      "frac_of_peak": pairs * F3 / (k3 * 1e-3) / 1e12 / PEAK_FP32_TFLOPS}

    # ---------------- accuracy vs the FP64 oracle ----------------
    err = None
    cpu = None
    if args.err_samples > 0 or not args.no_cpu:
        import oracle
        if not os.path.exists(oracle.LIB_PATH):
            oracle.build()
        o = oracle.default()
    if args.err_samples > 0:
        rng = np.random.default_rng(0x5EED)
        idx = np.sort(rng.choice(n, size=min(args.err_samples, n), replace=False)).astype(np.int32)
        log(f"FP64 oracle on {len(idx)} sampled targets...")
        t1 = time.time()
        v64, d64, _ = o.f64_P3D(p, p, "gaussian", sigma, idx=idx, gauss_mode=1)
        log(f"  {time.time() - t1:.1f}s")

        def stats(g, r):
            rel = oracle.per_target_rel(g, r, floor=1e-30)
            return {"max": float(rel.max()), "rms": float(np.sqrt(np.mean(rel * rel))),
                    "rel_l2": oracle.rel_l2(g, r)}
        err = {"samples": int(len(idx)), "vel": stats(vel[idx], v64), "dvort": stats(dvort[idx], d64)}
        if "plain" in vlist:
            err["plain_fp32"] = {"vel": stats(plain_vel[idx], v64), "dvort": stats(plain_dvort[idx], d64)}

    # ---------------- CPU baseline: the reference's OpenMP path ----------------
    if not args.no_cpu:
        m = min(args.cpu_samples, n)
        rng = np.random.default_rng(0x5EED + 1)
        tidx = np.sort(rng.choice(n, size=m, replace=False))
        tg = np.ascontiguousarray(p[tidx])
        log(f"CPU baseline: {m} targets x {n} sources, vel then dvort...")
        best = None
        for rep in range(2):
            c0 = time.perf_counter()
            rv = o.P3D_M2M_vel(p, tg["coord"], "gaussian", sigma)
            rd = o.P3D_M2M_dvort(p, tg, "gaussian", sigma)
            dt = time.perf_counter() - c0
            best = dt if best is None else min(best, dt)
        cpu_rate = m * float(n) / best
        cpu = {"value": cpu_rate, "unit": "interactions/s", "cores": int(o.L.orc_num_threads()),
               "kind": "port",
               "sample": f"{m} sampled ring targets x all {n} sources; cvtx_P3D_M2M_vel then "
                         f"cvtx_P3D_M2M_dvort as separate passes (reference has no fused path); "
                         f"best of 2, {best:.1f}s",
               "rel_l2_gpu_vs_cpu": {"vel": oracle.rel_l2(vel[tidx], rv),
                                      "dvort": oracle.rel_l2(dvort[tidx], rd)}}
        log(f"  {cpu_rate:.3e} interactions/s on {cpu['cores']} threads")

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            tj = json.load(open(tpath)).get("p3d_ring_vel_dvort_gaussian")
            if tj and tj.get("n") == n:
                traffic = tj["bytes_per_launch"]
        except Exception:
            traffic = None

    line = {
        "metric": "P3D M2M vel+dvort interactions/s at N=1M; % FP32 VALU peak; rel err vs FP64",
        "value": pairs / (ms_step * 1e-3),
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "thin vortex ring, fused P3D vel+dvort, Gaussian sigma=0.02",
                   "n_particles": n, "regularisation": "gaussian", "sigma": sigma,
                   "accumulation": "compensated (TwoSum tiles, FP64 epilogue)",
                   "flops_per_pair": F},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel_ms": ms_kernel},
        "cpu_baseline": cpu,
        "rel_err_vs_fp64": err,
        "variants": variants,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
