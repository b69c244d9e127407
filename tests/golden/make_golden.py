"""Write the golden fixtures tests/golden/{p3d,f3d,p2d}.npz.

Small seeded problems (64 sources, 32 targets) with the outputs of the
oracle's FP32 restatement of the reference CPU path ("ref_*") and of the
FP64 oracle ("f64_*", exact Gaussian).  The reference ships no golden
vectors (its test/readfloatarray.h reader is unused), so these are
regression anchors for the oracle and fixed expected outputs for the GPU.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from cvortex_amd import workloads as W  # noqa: E402

REGS = ("singular", "winckelmans", "planetary", "gaussian")
SIGMA, NU = 0.25, 0.05


def main():
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    o = oracle.default()
    out = {}
    src = W.random_p3d(64, 901, box=1.0, vort=1.0, vol=0.01)
    tgt = W.random_p3d(32, 902, box=1.0, vort=1.0, vol=0.01)
    out["src"], out["tgt"] = src, tgt
    for k in REGS:
        out[f"ref_vel_{k}"] = o.P3D_M2M_vel(src, tgt["coord"], k, SIGMA)
        out[f"ref_dvort_{k}"] = o.P3D_M2M_dvort(src, tgt, k, SIGMA)
        visc = k in ("winckelmans", "gaussian")
        v, d, s = o.f64_P3D(src, tgt, k, SIGMA, NU, gauss_mode=1, visc=visc)
        out[f"f64_vel_{k}"], out[f"f64_dvort_{k}"] = v, d
        if visc:
            out[f"ref_visc_{k}"] = o.P3D_M2M_visc_dvort(src, tgt, k, SIGMA, NU)
            out[f"f64_visc_{k}"] = s
    np.savez(os.path.join(HERE, "p3d.npz"), sigma=SIGMA, nu=NU, **out)

    fil = W.random_f3d(64, 903, box=1.0, strength=1.0)
    v, d = o.f64_F3D(fil, tgt)
    np.savez(os.path.join(HERE, "f3d.npz"), fil=fil, tgt=tgt,
             ref_vel=o.F3D_M2M_vel(fil, tgt["coord"]), ref_dvort=o.F3D_M2M_dvort(fil, tgt),
             f64_vel=v, f64_dvort=d)

    out = {}
    q = W.random_p2d(64, 904, box=1.0, vort=1.0, area=0.01)
    r = W.random_p2d(32, 905, box=1.0, vort=1.0, area=0.01)
    out["src"], out["tgt"] = q, r
    for k in REGS:
        out[f"ref_vel_{k}"] = o.P2D_M2M_vel(q, r["coord"], k, SIGMA)
        visc = k in ("winckelmans", "gaussian")
        v, e = o.f64_P2D(q, r, k, SIGMA, NU, gauss_mode=1, visc=visc)
        out[f"f64_vel_{k}"] = v
        if visc:
            out[f"ref_visc_{k}"] = o.P2D_M2M_visc_dvort(q, r, k, SIGMA, NU)
            out[f"f64_visc_{k}"] = e
    np.savez(os.path.join(HERE, "p2d.npz"), sigma=SIGMA, nu=NU, **out)


if __name__ == "__main__":
    main()
