"""Device-resident API wrappers (cvtx_hip_p3d_eval and friends).

Particles live in GPU memory as structure-of-arrays float4 buffers (layouts
in include/cvortex_hip.h); these helpers build the problem structs and call
the C ABI.  Calls are asynchronous on the given stream (None = the library's
stream)."""
import ctypes

import numpy as np

from . import CvtxError, last_error, lib
from .hip import DeviceArray
from .types import REG_ID

_VP, _I, _F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float


class P3DProblem(ctypes.Structure):
    _fields_ = [("src_pv", _VP), ("src_a", _VP), ("n_src", _I), ("tgt_pv", _VP), ("tgt_a", _VP),
                ("n_tgt", _I), ("vel", _VP), ("dvort", _VP), ("visc", _VP),
                ("regularisation", _I), ("sigma", _F), ("nu", _F), ("accumulate", _I)]


class P2DProblem(ctypes.Structure):
    _fields_ = [("src", _VP), ("n_src", _I), ("tgt", _VP), ("n_tgt", _I), ("vel", _VP),
                ("visc", _VP), ("regularisation", _I), ("sigma", _F), ("nu", _F),
                ("accumulate", _I)]


class F3DProblem(ctypes.Structure):
    _fields_ = [("fa", _VP), ("fb", _VP), ("n_fil", _I), ("tgt_pv", _VP), ("tgt_a", _VP),
                ("n_tgt", _I), ("vel", _VP), ("dvort", _VP), ("accumulate", _I)]


def p3d_soa(p):
    """cvtx_P3D records -> (pv, a) float32 (n, 4) arrays."""
    n = len(p)
    pv = np.zeros((n, 4), np.float32)
    a = np.zeros((n, 4), np.float32)
    pv[:, :3] = p["coord"]
    pv[:, 3] = p["volume"]
    a[:, :3] = p["vorticity"]
    return pv, a


def p2d_soa(p):
    n = len(p)
    s = np.zeros((n, 4), np.float32)
    s[:, :2] = p["coord"]
    s[:, 2] = p["vorticity"]
    s[:, 3] = p["area"]
    return s


def f3d_soa(f):
    n = len(f)
    fa = np.zeros((n, 4), np.float32)
    fb = np.zeros((n, 4), np.float32)
    fa[:, :3] = f["start"]
    fa[:, 3] = f["strength"]
    fb[:, :3] = f["end"]
    return fa, fb


def _a(d):
    return d.addr if d is not None else None


def p3d_problem(src_pv, src_a, n_src, tgt_pv, tgt_a, n_tgt, vel, dvort, visc, kernel, sigma,
                nu=0.0, accumulate=1):
    return P3DProblem(_a(src_pv), _a(src_a), n_src, _a(tgt_pv), _a(tgt_a), n_tgt, _a(vel),
                      _a(dvort), _a(visc), REG_ID[kernel] if isinstance(kernel, str) else kernel,
                      sigma, nu, accumulate)


def p3d_eval(prob, stream=None, shape=None):
    L = lib()
    if shape is None:
        st = L.cvtx_hip_p3d_eval(ctypes.byref(prob), stream)
    else:
        st = L.cvtx_hip_p3d_eval_shape(ctypes.byref(prob), shape[0], shape[1], stream)
    if st != 0:
        raise CvtxError(f"p3d_eval: {last_error()}")


def p2d_eval(prob, stream=None, shape=None):
    L = lib()
    if shape is None:
        st = L.cvtx_hip_p2d_eval(ctypes.byref(prob), stream)
    else:
        st = L.cvtx_hip_p2d_eval_shape(ctypes.byref(prob), shape[0], shape[1], stream)
    if st != 0:
        raise CvtxError(f"p2d_eval: {last_error()}")


def f3d_eval(prob, stream=None, shape=None):
    L = lib()
    if shape is None:
        st = L.cvtx_hip_f3d_eval(ctypes.byref(prob), stream)
    else:
        st = L.cvtx_hip_f3d_eval_shape(ctypes.byref(prob), shape[0], shape[1], stream)
    if st != 0:
        raise CvtxError(f"f3d_eval: {last_error()}")


class P3DDevice:
    """A P3D particle set resident on the GPU (sources = targets)."""

    def __init__(self, particles):
        pv, a = p3d_soa(particles)
        self.n = len(particles)
        self.pv = DeviceArray.from_host(pv)
        self.a = DeviceArray.from_host(a)

    def outputs(self, vel=True, dvort=True, visc=False):
        mk = lambda f: DeviceArray((self.n, 3), np.float32) if f else None
        return mk(vel), mk(dvort), mk(visc)

    def problem(self, outs, kernel, sigma, nu=0.0, accumulate=1):
        v, d, s = outs
        return p3d_problem(self.pv, self.a, self.n, self.pv, self.a, self.n, v, d, s, kernel,
                           sigma, nu, accumulate)
