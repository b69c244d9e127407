"""Device-resident API wrappers (cvtx_hip_p3d_eval and friends).

Particles live in GPU memory as structure-of-arrays float4 buffers (layouts
in include/cvortex_hip.h); these helpers build the problem structs and call
the C ABI.  Calls are asynchronous on the given stream (None = the library's
stream)."""
import ctypes

import numpy as np

from . import CvtxError, last_error, lib
from .hip import DeviceArray, synchronize  # noqa: F401  (re-exported)
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


class VortProblem(ctypes.Structure):
    _fields_ = [("src_pv", _VP), ("src_a", _VP), ("n_src", _I), ("tgt", _VP), ("n_tgt", _I),
                ("vort", _VP), ("regularisation", _I), ("sigma", _F), ("accumulate", _I)]


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


def p3d_eval_noalloc(prob, shape, stream=None):
    """cvtx_hip_p3d_eval_shape_noalloc: the explicit-shape evaluation with
    its partial sums in the library's fixed scratch array (no allocation)."""
    st = lib().cvtx_hip_p3d_eval_shape_noalloc(ctypes.byref(prob), shape[0], shape[1], stream)
    if st != 0:
        raise CvtxError(f"p3d_eval_noalloc: {last_error()}")


NOALLOC_SCRATCH_BYTES = 64 << 20


def autotune_p3d(prob, reps=2):
    """Time the candidate launch shapes of a device problem and record the
    fastest for its size bucket (cvtx_hip_autotune_p3d).  Returns
    (targets_per_lane, n_split, ms)."""
    L = lib()
    T, k, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_float(0.0)
    st = L.cvtx_hip_autotune_p3d(ctypes.byref(prob), reps, ctypes.byref(T), ctypes.byref(k),
                                 ctypes.byref(ms))
    if st != 0:
        raise CvtxError(f"autotune_p3d: {last_error()}")
    return T.value, k.value, ms.value


def vort_eval(prob, stream=None, shape=None):
    L = lib()
    if shape is None:
        st = L.cvtx_hip_vort_eval(ctypes.byref(prob), stream)
    else:
        st = L.cvtx_hip_vort_eval_shape(ctypes.byref(prob), shape[0], shape[1], stream)
    if st != 0:
        raise CvtxError(f"vort_eval: {last_error()}")


def f3d_inf_eval(fa, fb, n_fil, mes, dirs, n_mes, out, stream=None):
    """Influence matrix of device filaments into out (n_mes, n_fil)."""
    if lib().cvtx_hip_f3d_inf_eval(fa.addr, fb.addr, n_fil, mes.addr, dirs.addr, n_mes, out.addr,
                                   stream) != 0:
        raise CvtxError(f"f3d_inf_eval: {last_error()}")


def relax_eval(pv, a, n, fdt, kernel, sigma, accumulate=1, stream=None):
    """In-place Pedrizzetti relaxation of device strengths `a` (float4 rows)."""
    reg = REG_ID[kernel] if isinstance(kernel, str) else int(kernel)
    if lib().cvtx_hip_relax_eval(pv.addr, a.addr, n, float(fdt), reg, float(sigma), accumulate,
                                 stream) != 0:
        raise CvtxError(f"relax_eval: {last_error()}")


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


class Batch:
    """Many independent P3D clouds evaluated in one persistent launch
    (cvtx_hip_batch_*).  offsets: n+1 particle offsets, sigmas: n radii."""

    def __init__(self, offsets, sigmas, kernel="winckelmans", outputs=3, accumulate=1, nu=0.0):
        self.offsets = np.ascontiguousarray(offsets, np.int32)
        self.sigmas = np.ascontiguousarray(sigmas, np.float32)
        n = len(self.sigmas)
        reg = REG_ID[kernel] if isinstance(kernel, str) else kernel
        self.h = lib().cvtx_hip_batch_create(n, self.offsets.ctypes.data, self.sigmas.ctypes.data,
                                             reg, outputs, accumulate, nu)
        if not self.h:
            raise CvtxError(f"batch_create: {last_error()}")

    def _ck(self, st, what):
        if st != 0:
            raise CvtxError(f"{what}: {last_error()}")

    def set_shape(self, targets_per_lane, workgroups_per_cu):
        self._ck(lib().cvtx_hip_batch_set_shape(self.h, targets_per_lane, workgroups_per_cu), "set_shape")

    def info(self):
        u, g, t = _I(), _I(), _I()
        self._ck(lib().cvtx_hip_batch_info(self.h, ctypes.byref(u), ctypes.byref(g), ctypes.byref(t)), "info")
        return u.value, g.value, t.value

    def shape(self):
        """(targets per lane, workgroup size, resident workgroups per CU)."""
        t, w, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._ck(lib().cvtx_hip_batch_shape(self.h, ctypes.byref(t), ctypes.byref(w),
                                            ctypes.byref(g)), "shape")
        return t.value, w.value, g.value

    def autotune(self, pv, a, vel, dvort, visc=None, reps=1):
        """cvtx_hip_batch_autotune: (targets per lane, workgroup size,
        workgroups per CU, ms) of the fastest shape, now the batch's."""
        t, w, g, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        self._ck(lib().cvtx_hip_batch_autotune(self.h, pv.addr, a.addr, vel.addr, dvort.addr,
                                               visc.addr if visc is not None else None, reps,
                                               ctypes.byref(t), ctypes.byref(w), ctypes.byref(g),
                                               ctypes.byref(ms)), "autotune")
        return t.value, w.value, g.value, ms.value

    def eval(self, pv, a, vel, dvort, visc=None, stream=None):
        self._ck(lib().cvtx_hip_batch_eval(self.h, pv.addr, a.addr, vel.addr, dvort.addr,
                                           visc.addr if visc is not None else None, stream), "batch_eval")

    def eval_per_cloud(self, pv, a, vel, dvort, visc=None, stream=None, plain=False):
        self._ck(lib().cvtx_hip_batch_eval_per_cloud(self.h, pv.addr, a.addr, vel.addr, dvort.addr,
                                                     visc.addr if visc is not None else None,
                                                     1 if plain else 0, stream),
                 "batch_eval_per_cloud")

    def rk4_init(self, pv, a, dt):
        self._ck(lib().cvtx_hip_batch_rk4_init(self.h, pv.addr, a.addr, dt), "rk4_init")

    def rk4_step(self, n_steps=1, stream=None):
        self._ck(lib().cvtx_hip_batch_rk4_step(self.h, n_steps, stream), "rk4_step")

    def destroy(self):
        if self.h:
            lib().cvtx_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
