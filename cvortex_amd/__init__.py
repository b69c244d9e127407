"""Python host interface of libcvortex.so (gfx950 direct-sum Biot-Savart).

The functions mirror the reference cvortex C API (reference
include/cvortex/libcvtx.h): same names without the ``cvtx_`` prefix, same
argument meaning, records as numpy structured arrays (see
:mod:`cvortex_amd.types`).  Every call goes through the C ABI of the in-tree
``libcvortex.so``; there is no Python or CPU implementation of any pair
interaction here, and loading fails loudly when the library has not been
built.

Errors: the public ``cvtx_*`` functions return void and signal failure by
NaN results (plus a line on stderr); the wrappers here call the library's
int-returning seam (``cvtx_hip_*``) instead and raise :class:`CvtxError`
with the library's message when it returns -1.
"""
import ctypes
import os

import numpy as np

from .types import (F3D_DTYPE, P2D_DTYPE, P3D_DTYPE, REG_ID, REG_NAMES, RedistFunc, V2f, V3f,
                    VortFunc, as_records, pointer_array)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcvortex.so")

ACC_PLAIN = 0
ACC_COMP = 1
ACC_TILES = 2
OUT_VEL, OUT_DVORT, OUT_VISC = 1, 2, 4
# which leg the public void functions use (cvtx_hip_set_dispatch)
DISPATCH_AUTO, DISPATCH_DEVICE, DISPATCH_HOST = 0, 1, 2
LEG_DEVICE, LEG_HOST = 0, 1
# single-source / single-target calls from this many elements go to the device
SMALL_DEVICE_MIN = 4096

_lib = None


class CvtxError(RuntimeError):
    pass


_VP = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float

_SIGNATURES = {
    # accelerator controls
    "cvtx_initialise": (None, []),
    "cvtx_finalise": (None, []),
    "cvtx_information": (ctypes.c_char_p, []),
    "cvtx_num_accelerators": (_I, []),
    "cvtx_num_enabled_accelerators": (_I, []),
    "cvtx_accelerator_name": (ctypes.c_char_p, [_I]),
    "cvtx_accelerator_enabled": (_I, [_I]),
    "cvtx_accelerator_enable": (None, [_I]),
    "cvtx_accelerator_disable": (None, [_I]),
    "cvtx_VortFunc_singular": (VortFunc, []),
    "cvtx_VortFunc_winckelmans": (VortFunc, []),
    "cvtx_VortFunc_planetary": (VortFunc, []),
    "cvtx_VortFunc_gaussian": (VortFunc, []),
    # public M2M (void)
    "cvtx_P3D_M2M_vel": (None, [_VP, _I, _VP, _I, _VP, _VP, _F]),
    "cvtx_RedistFunc_lambda0": (RedistFunc, []),
    "cvtx_RedistFunc_lambda1": (RedistFunc, []),
    "cvtx_RedistFunc_lambda2": (RedistFunc, []),
    "cvtx_RedistFunc_lambda3": (RedistFunc, []),
    "cvtx_RedistFunc_m4p": (RedistFunc, []),
    "cvtx_P3D_redistribute_on_grid": (_I, [_VP, _I, _VP, _I, _VP, _F, _F]),
    "cvtx_P2D_redistribute_on_grid": (_I, [_VP, _I, _VP, _I, _VP, _F, _F]),
    # single-source / single-target forms (records by pointer, points by value)
    "cvtx_P3D_S2S_vel": (V3f, [_VP, V3f, _VP, _F]),
    "cvtx_P3D_S2S_dvort": (V3f, [_VP, _VP, _VP, _F]),
    "cvtx_P3D_S2S_visc_dvort": (V3f, [_VP, _VP, _VP, _F, _F]),
    "cvtx_P3D_S2S_vort": (V3f, [_VP, V3f, _VP, _F]),
    "cvtx_P3D_S2M_vel": (None, [_VP, _VP, _I, _VP, _VP, _F]),
    "cvtx_P3D_S2M_dvort": (None, [_VP, _VP, _I, _VP, _VP, _F]),
    "cvtx_P3D_S2M_visc_dvort": (None, [_VP, _VP, _I, _VP, _VP, _F, _F]),
    "cvtx_P3D_S2M_vort": (None, [_VP, _VP, _I, _VP, _VP, _F]),
    "cvtx_P3D_M2S_vel": (V3f, [_VP, _I, V3f, _VP, _F]),
    "cvtx_P3D_M2S_dvort": (V3f, [_VP, _I, _VP, _VP, _F]),
    "cvtx_P3D_M2S_visc_dvort": (V3f, [_VP, _I, _VP, _VP, _F, _F]),
    "cvtx_P3D_M2S_vort": (V3f, [_VP, _I, V3f, _VP, _F]),
    "cvtx_F3D_S2S_vel": (V3f, [_VP, V3f]),
    "cvtx_F3D_S2S_dvort": (V3f, [_VP, _VP]),
    "cvtx_F3D_M2S_vel": (V3f, [_VP, _I, V3f]),
    "cvtx_F3D_M2S_dvort": (V3f, [_VP, _I, _VP]),
    "cvtx_P2D_S2S_vel": (V2f, [_VP, V2f, _VP, _F]),
    "cvtx_P2D_S2M_vel": (None, [_VP, _VP, _I, _VP, _VP, _F]),
    "cvtx_P2D_M2S_vel": (V2f, [_VP, _I, V2f, _VP, _F]),
    "cvtx_P2D_S2S_visc_dvort": (_F, [_VP, _VP, _VP, _F, _F]),
    "cvtx_P2D_S2M_visc_dvort": (None, [_VP, _VP, _I, _VP, _VP, _F, _F]),
    "cvtx_P2D_M2S_visc_dvort": (_F, [_VP, _I, _VP, _VP, _F, _F]),
    "cvtx_P3D_M2M_vort": (None, [_VP, _I, _VP, _I, _VP, _VP, _F]),
    "cvtx_P3D_pedrizzetti_relaxation": (None, [_VP, _I, _F, _VP, _F]),
    "cvtx_P3D_M2M_dvort": (None, [_VP, _I, _VP, _I, _VP, _VP, _F]),
    "cvtx_P3D_M2M_visc_dvort": (None, [_VP, _I, _VP, _I, _VP, _VP, _F, _F]),
    "cvtx_F3D_M2M_vel": (None, [_VP, _I, _VP, _I, _VP]),
    "cvtx_F3D_M2M_dvort": (None, [_VP, _I, _VP, _I, _VP]),
    "cvtx_P2D_M2M_vel": (None, [_VP, _I, _VP, _I, _VP, _VP, _F]),
    "cvtx_P2D_M2M_visc_dvort": (None, [_VP, _I, _VP, _I, _VP, _VP, _F, _F]),
    # seam (int)
    "cvtx_hip_P3D_M2M_vel": (_I, [_VP, _I, _VP, _I, _VP, _VP, _F]),
    "cvtx_hip_P3D_M2M_dvort": (_I, [_VP, _I, _VP, _I, _VP, _VP, _F]),
    "cvtx_hip_P3D_M2M_visc_dvort": (_I, [_VP, _I, _VP, _I, _VP, _VP, _F, _F]),
    "cvtx_hip_P3D_M2M_vort": (_I, [_VP, _I, _VP, _I, _VP, _VP, _F]),
    "cvtx_hip_F3D_inf_mtrx": (_I, [_VP, _I, _VP, _VP, _I, _VP]),
    "cvtx_F3D_inf_mtrx": (None, [_VP, _I, _VP, _VP, _I, _VP]),
    "cvtx_hip_P3D_pedrizzetti_relaxation": (_I, [_VP, _I, _F, _VP, _F]),
    "cvtx_hip_F3D_M2M_vel": (_I, [_VP, _I, _VP, _I, _VP]),
    "cvtx_hip_F3D_M2M_dvort": (_I, [_VP, _I, _VP, _I, _VP]),
    "cvtx_hip_P2D_M2M_vel": (_I, [_VP, _I, _VP, _I, _VP, _VP, _F]),
    "cvtx_hip_P2D_M2M_visc_dvort": (_I, [_VP, _I, _VP, _I, _VP, _VP, _F, _F]),
    "cvtx_hip_P3D_M2M_fused": (_I, [_VP, _I, _VP, _I, _VP, _VP, _VP, _VP, _F, _F]),
    "cvtx_hip_F3D_M2M_fused": (_I, [_VP, _I, _VP, _I, _VP, _VP]),
    "cvtx_hip_P2D_M2M_fused": (_I, [_VP, _I, _VP, _I, _VP, _VP, _VP, _F, _F]),
    # controls
    "cvtx_hip_last_error": (ctypes.c_char_p, []),
    "cvtx_hip_set_accumulation": (_I, [_I]),
    "cvtx_hip_get_accumulation": (_I, []),
    "cvtx_hip_set_target_order": (_I, [_I]),
    "cvtx_hip_stream": (_VP, []),
    "cvtx_hip_set_dispatch": (_I, [_I]),
    "cvtx_hip_get_dispatch": (_I, []),
    "cvtx_hip_last_leg": (_I, []),
    "cvtx_hip_host_threads": (_I, []),
    "cvtx_hip_synchronize": (_I, []),
    # device-resident
    "cvtx_hip_p3d_eval": (_I, [_VP, _VP]),
    "cvtx_hip_vort_eval": (_I, [_VP, _VP]),
    "cvtx_hip_f3d_inf_eval": (_I, [_VP, _VP, _I, _VP, _VP, _I, _VP, _VP]),
    "cvtx_hip_vort_eval_shape": (_I, [_VP, _I, _I, _VP]),
    "cvtx_hip_relax_eval": (_I, [_VP, _VP, _I, _F, _I, _F, _I, _VP]),
    "cvtx_hip_p2d_eval": (_I, [_VP, _VP]),
    "cvtx_hip_f3d_eval": (_I, [_VP, _VP]),
    "cvtx_hip_p3d_eval_shape": (_I, [_VP, _I, _I, _VP]),
    "cvtx_hip_p3d_eval_shape_noalloc": (_I, [_VP, _I, _I, _VP]),
    "cvtx_hip_p2d_eval_shape": (_I, [_VP, _I, _I, _VP]),
    "cvtx_hip_f3d_eval_shape": (_I, [_VP, _I, _I, _VP]),
    "cvtx_hip_autotune_p3d": (_I, [_VP, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                   ctypes.POINTER(_F)]),
    "cvtx_hip_tune_save": (_I, [ctypes.c_char_p]),
    "cvtx_hip_tune_load": (_I, [ctypes.c_char_p]),
    "cvtx_hip_tune_clear": (None, []),
    "cvtx_hip_gaussian_table": (_I, [_VP, _I]),
    "cvtx_hip_batch_create": (_VP, [_I, _VP, _VP, _I, _I, _I, _F]),
    "cvtx_hip_batch_destroy": (None, [_VP]),
    "cvtx_hip_batch_set_shape": (_I, [_VP, _I, _I]),
    "cvtx_hip_batch_info": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "cvtx_hip_batch_shape": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "cvtx_hip_batch_autotune": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _I, ctypes.POINTER(_I),
                                     ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_F)]),
    "cvtx_hip_batch_eval": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "cvtx_hip_batch_eval_per_cloud": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _I, _VP]),
    "cvtx_hip_batch_rk4_init": (_I, [_VP, _VP, _VP, _F]),
    "cvtx_hip_batch_rk4_step": (_I, [_VP, _I, _VP]),
}


def lib():
    """The loaded libcvortex.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CvtxError(f"{LIB_PATH} not found: build it first (python -c "
                            "'import __graft_entry__ as g; g.build()' or make -C cvortex_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    msg = lib().cvtx_hip_last_error()
    return msg.decode() if msg else ""


def _check(status, what):
    if status != 0:
        raise CvtxError(f"{what}: {last_error()}")


# ----------------------------- controls -----------------------------

def initialise():
    lib().cvtx_initialise()


def finalise():
    lib().cvtx_finalise()


def information():
    return lib().cvtx_information().decode()


def num_accelerators():
    return lib().cvtx_num_accelerators()


def num_enabled_accelerators():
    return lib().cvtx_num_enabled_accelerators()


def accelerator_name(i):
    s = lib().cvtx_accelerator_name(i)
    return s.decode() if s is not None else None


def accelerator_enabled(i):
    return lib().cvtx_accelerator_enabled(i)


def accelerator_enable(i):
    lib().cvtx_accelerator_enable(i)


def accelerator_disable(i):
    lib().cvtx_accelerator_disable(i)


def set_accumulation(mode):
    prev = lib().cvtx_hip_set_accumulation(mode)
    if prev < 0:
        raise CvtxError(f"invalid accumulation mode {mode}")
    return prev


def get_accumulation():
    return lib().cvtx_hip_get_accumulation()


def set_dispatch(mode):
    """Leg of the public void functions (DISPATCH_AUTO, _DEVICE, _HOST);
    returns the previous mode."""
    prev = lib().cvtx_hip_set_dispatch(mode)
    if prev < 0:
        raise CvtxError(f"invalid dispatch mode {mode}")
    return prev


def get_dispatch():
    return lib().cvtx_hip_get_dispatch()


def last_leg():
    """LEG_DEVICE or LEG_HOST for the last public void call (-1: none or failed)."""
    return lib().cvtx_hip_last_leg()


def host_threads():
    return lib().cvtx_hip_host_threads()


def VortFunc_singular():
    return lib().cvtx_VortFunc_singular()


def VortFunc_winckelmans():
    return lib().cvtx_VortFunc_winckelmans()


def VortFunc_planetary():
    return lib().cvtx_VortFunc_planetary()


def VortFunc_gaussian():
    return lib().cvtx_VortFunc_gaussian()


def vortfunc(name):
    """cvtx_VortFunc for a regularisation name."""
    return {"singular": VortFunc_singular, "winckelmans": VortFunc_winckelmans,
            "planetary": VortFunc_planetary, "gaussian": VortFunc_gaussian}[name]()


def _kref(kernel):
    if isinstance(kernel, str):
        kernel = vortfunc(kernel)
    return kernel, ctypes.cast(ctypes.pointer(kernel), ctypes.c_void_p)


def _ptrs(records, dtype):
    recs = as_records(records, dtype)
    p, keep = pointer_array(recs)
    return p, keep


def _vp(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


# ------------------------------ P3D ------------------------------

def P3D_M2M_vel(particles, mes, kernel, regularisation_radius):
    """Velocity induced by 3D vortex particles at points mes (n,3).
    Reference include/cvortex/libcvtx.h:213-220."""
    p, keep = _ptrs(particles, P3D_DTYPE)
    mes = np.ascontiguousarray(mes, dtype=np.float32).reshape(-1, 3)
    out = np.empty((mes.shape[0], 3), np.float32)
    k, kp = _kref(kernel)
    _check(lib().cvtx_hip_P3D_M2M_vel(_vp(p), len(p), _vp(mes), mes.shape[0], _vp(out), kp,
                                      float(regularisation_radius)), "P3D_M2M_vel")
    return out


def P3D_M2M_vort(particles, mes, kernel, regularisation_radius):
    """Vorticity field of 3D vortex particles at points mes (n,3), with the
    reference CPU path's 5 sigma box cut-off.  libcvtx.h:241-248."""
    p, keep = _ptrs(particles, P3D_DTYPE)
    mes = np.ascontiguousarray(mes, dtype=np.float32).reshape(-1, 3)
    out = np.empty((mes.shape[0], 3), np.float32)
    k, kp = _kref(kernel)
    _check(lib().cvtx_hip_P3D_M2M_vort(_vp(p), len(p), _vp(mes), mes.shape[0], _vp(out), kp,
                                       float(regularisation_radius)), "P3D_M2M_vort")
    return out


def P3D_pedrizzetti_relaxation(particles, fdt, kernel, regularisation_radius):
    """Pedrizzetti relaxation (libcvtx.h:259-264); returns a relaxed copy of
    the records (the library updates the records it is given in place)."""
    recs = np.array(particles, dtype=P3D_DTYPE, copy=True)
    p, keep = _ptrs(recs, P3D_DTYPE)
    k, kp = _kref(kernel)
    _check(lib().cvtx_hip_P3D_pedrizzetti_relaxation(_vp(p), len(recs), float(fdt), kp,
                                                     float(regularisation_radius)),
           "P3D_pedrizzetti_relaxation")
    return recs


def F3D_inf_mtrx(filaments, mes, dirs):
    """Influence matrix (num_mes, num_filaments): velocity of each filament at
    each point projected on that point's direction.  libcvtx.h:299-305."""
    f, keep = _ptrs(filaments, F3D_DTYPE)
    mes = np.ascontiguousarray(mes, dtype=np.float32).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
    if dirs.shape != mes.shape:
        raise ValueError("one direction per measurement point")
    out = np.empty((mes.shape[0], len(f)), np.float32)
    _check(lib().cvtx_hip_F3D_inf_mtrx(_vp(f), len(f), _vp(mes), _vp(dirs), mes.shape[0],
                                       _vp(out)), "F3D_inf_mtrx")
    return out


def P3D_M2M_dvort(particles, induced, kernel, regularisation_radius):
    """Vortex stretching on `induced` particles.  libcvtx.h:222-229."""
    p, k1 = _ptrs(particles, P3D_DTYPE)
    q, k2 = _ptrs(induced, P3D_DTYPE)
    out = np.empty((len(q), 3), np.float32)
    k, kp = _kref(kernel)
    _check(lib().cvtx_hip_P3D_M2M_dvort(_vp(p), len(p), _vp(q), len(q), _vp(out), kp,
                                        float(regularisation_radius)), "P3D_M2M_dvort")
    return out


def P3D_M2M_visc_dvort(particles, induced, kernel, regularisation_radius, kinematic_visc):
    """Viscous (PSE) vorticity exchange on `induced`.  libcvtx.h:231-239."""
    p, k1 = _ptrs(particles, P3D_DTYPE)
    q, k2 = _ptrs(induced, P3D_DTYPE)
    out = np.empty((len(q), 3), np.float32)
    k, kp = _kref(kernel)
    _check(lib().cvtx_hip_P3D_M2M_visc_dvort(_vp(p), len(p), _vp(q), len(q), _vp(out), kp,
                                             float(regularisation_radius), float(kinematic_visc)),
           "P3D_M2M_visc_dvort")
    return out


def P3D_M2M_fused(particles, induced, kernel, regularisation_radius, kinematic_visc=0.0,
                  vel=True, dvort=True, visc=False):
    """vel, dvort and/or visc_dvort in one pass (None for outputs not asked)."""
    p, k1 = _ptrs(particles, P3D_DTYPE)
    q, k2 = _ptrs(induced, P3D_DTYPE)
    n = len(q)
    outs = [np.empty((n, 3), np.float32) if f else None for f in (vel, dvort, visc)]
    k, kp = _kref(kernel)
    _check(lib().cvtx_hip_P3D_M2M_fused(_vp(p), len(p), _vp(q), n, *(_vp(o) if o is not None
                                                                     else None for o in outs),
                                        kp, float(regularisation_radius), float(kinematic_visc)),
           "P3D_M2M_fused")
    return tuple(outs)


# ------------------------------ F3D ------------------------------

def F3D_M2M_vel(filaments, mes):
    """Velocity induced by straight vortex filaments.  libcvtx.h:285-290."""
    f, keep = _ptrs(filaments, F3D_DTYPE)
    mes = np.ascontiguousarray(mes, dtype=np.float32).reshape(-1, 3)
    out = np.empty((mes.shape[0], 3), np.float32)
    _check(lib().cvtx_hip_F3D_M2M_vel(_vp(f), len(f), _vp(mes), mes.shape[0], _vp(out)),
           "F3D_M2M_vel")
    return out


def F3D_M2M_dvort(filaments, induced):
    """Vortex stretching by filaments on particles.  libcvtx.h:292-297."""
    f, k1 = _ptrs(filaments, F3D_DTYPE)
    q, k2 = _ptrs(induced, P3D_DTYPE)
    out = np.empty((len(q), 3), np.float32)
    _check(lib().cvtx_hip_F3D_M2M_dvort(_vp(f), len(f), _vp(q), len(q), _vp(out)),
           "F3D_M2M_dvort")
    return out


def F3D_M2M_fused(filaments, induced):
    f, k1 = _ptrs(filaments, F3D_DTYPE)
    q, k2 = _ptrs(induced, P3D_DTYPE)
    vel = np.empty((len(q), 3), np.float32)
    dv = np.empty((len(q), 3), np.float32)
    _check(lib().cvtx_hip_F3D_M2M_fused(_vp(f), len(f), _vp(q), len(q), _vp(vel), _vp(dv)),
           "F3D_M2M_fused")
    return vel, dv


# ------------------------------ P2D ------------------------------

def P2D_M2M_vel(particles, mes, kernel, regularisation_radius):
    """Velocity induced by 2D vortex particles.  libcvtx.h:329-336."""
    p, keep = _ptrs(particles, P2D_DTYPE)
    mes = np.ascontiguousarray(mes, dtype=np.float32).reshape(-1, 2)
    out = np.empty((mes.shape[0], 2), np.float32)
    k, kp = _kref(kernel)
    _check(lib().cvtx_hip_P2D_M2M_vel(_vp(p), len(p), _vp(mes), mes.shape[0], _vp(out), kp,
                                      float(regularisation_radius)), "P2D_M2M_vel")
    return out


def P2D_M2M_visc_dvort(particles, induced, kernel, regularisation_radius, kinematic_visc):
    """2D PSE vorticity exchange.  libcvtx.h:362-370."""
    p, k1 = _ptrs(particles, P2D_DTYPE)
    q, k2 = _ptrs(induced, P2D_DTYPE)
    out = np.empty(len(q), np.float32)
    k, kp = _kref(kernel)
    _check(lib().cvtx_hip_P2D_M2M_visc_dvort(_vp(p), len(p), _vp(q), len(q), _vp(out), kp,
                                             float(regularisation_radius), float(kinematic_visc)),
           "P2D_M2M_visc_dvort")
    return out


def P2D_M2M_fused(particles, induced, kernel, regularisation_radius, kinematic_visc):
    p, k1 = _ptrs(particles, P2D_DTYPE)
    q, k2 = _ptrs(induced, P2D_DTYPE)
    vel = np.empty((len(q), 2), np.float32)
    vs = np.empty(len(q), np.float32)
    k, kp = _kref(kernel)
    _check(lib().cvtx_hip_P2D_M2M_fused(_vp(p), len(p), _vp(q), len(q), _vp(vel), _vp(vs), kp,
                                        float(regularisation_radius), float(kinematic_visc)),
           "P2D_M2M_fused")
    return vel, vs


__all__ = [n for n in dir() if not n.startswith("_")] + ["REG_ID", "REG_NAMES"]


# --------------- single-source (S2S, S2M) / single-target (M2S) ---------------
# The reference's small entry points (libcvtx.h:125-211, 266-283, 306-360),
# called through the public C ABI: results are NaN (with a line on stderr)
# when the call fails.

def _rec(r, dtype):
    """Pointer to one record (kept alive by the returned array)."""
    a = as_records(np.atleast_1d(np.asarray(r)), dtype)[:1].copy()
    return ctypes.c_void_p(a.ctypes.data), a


def _v3(x):
    return V3f((ctypes.c_float * 3)(*np.asarray(x, np.float32).ravel()[:3]))


def _v2(x):
    return V2f((ctypes.c_float * 2)(*np.asarray(x, np.float32).ravel()[:2]))


def _arr(v, n):
    return np.array(v.x[:n], np.float32)


def P3D_S2S_vel(p, mes, kernel, sigma):
    rp, k1 = _rec(p, P3D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P3D_S2S_vel(rp, _v3(mes), kp, float(sigma)), 3)


def P3D_S2S_dvort(p, ind, kernel, sigma):
    rp, k1 = _rec(p, P3D_DTYPE)
    ri, k2 = _rec(ind, P3D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P3D_S2S_dvort(rp, ri, kp, float(sigma)), 3)


def P3D_S2S_visc_dvort(p, ind, kernel, sigma, nu):
    rp, k1 = _rec(p, P3D_DTYPE)
    ri, k2 = _rec(ind, P3D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P3D_S2S_visc_dvort(rp, ri, kp, float(sigma), float(nu)), 3)


def P3D_S2S_vort(p, mes, kernel, sigma):
    rp, k1 = _rec(p, P3D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P3D_S2S_vort(rp, _v3(mes), kp, float(sigma)), 3)


def _s2m_points(fn, p, mes, kernel, sigma, dtype, dim):
    rp, k1 = _rec(p, dtype)
    mes = np.ascontiguousarray(mes, np.float32).reshape(-1, dim)
    out = np.empty_like(mes)
    k, kp = _kref(kernel)
    fn(rp, _vp(mes), mes.shape[0], _vp(out), kp, float(sigma))
    return out


def _s2m_records(fn, p, induced, kernel, dtype, width, *extra):
    rp, k1 = _rec(p, dtype)
    q, k2 = _ptrs(induced, dtype)
    out = np.empty((len(q), width) if width > 1 else len(q), np.float32)
    k, kp = _kref(kernel)
    fn(rp, _vp(q), len(q), _vp(out), kp, *[float(x) for x in extra])
    return out


def P3D_S2M_vel(p, mes, kernel, sigma):
    return _s2m_points(lib().cvtx_P3D_S2M_vel, p, mes, kernel, sigma, P3D_DTYPE, 3)


def P3D_S2M_vort(p, mes, kernel, sigma):
    return _s2m_points(lib().cvtx_P3D_S2M_vort, p, mes, kernel, sigma, P3D_DTYPE, 3)


def P3D_S2M_dvort(p, induced, kernel, sigma):
    return _s2m_records(lib().cvtx_P3D_S2M_dvort, p, induced, kernel, P3D_DTYPE, 3, sigma)


def P3D_S2M_visc_dvort(p, induced, kernel, sigma, nu):
    return _s2m_records(lib().cvtx_P3D_S2M_visc_dvort, p, induced, kernel, P3D_DTYPE, 3, sigma, nu)


def P3D_M2S_vel(particles, mes, kernel, sigma):
    p, keep = _ptrs(particles, P3D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P3D_M2S_vel(_vp(p), len(p), _v3(mes), kp, float(sigma)), 3)


def P3D_M2S_vort(particles, mes, kernel, sigma):
    p, keep = _ptrs(particles, P3D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P3D_M2S_vort(_vp(p), len(p), _v3(mes), kp, float(sigma)), 3)


def P3D_M2S_dvort(particles, ind, kernel, sigma):
    p, keep = _ptrs(particles, P3D_DTYPE)
    ri, k2 = _rec(ind, P3D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P3D_M2S_dvort(_vp(p), len(p), ri, kp, float(sigma)), 3)


def P3D_M2S_visc_dvort(particles, ind, kernel, sigma, nu):
    p, keep = _ptrs(particles, P3D_DTYPE)
    ri, k2 = _rec(ind, P3D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P3D_M2S_visc_dvort(_vp(p), len(p), ri, kp, float(sigma), float(nu)), 3)


def F3D_S2S_vel(f, mes):
    rf, k1 = _rec(f, F3D_DTYPE)
    return _arr(lib().cvtx_F3D_S2S_vel(rf, _v3(mes)), 3)


def F3D_S2S_dvort(f, ind):
    rf, k1 = _rec(f, F3D_DTYPE)
    ri, k2 = _rec(ind, P3D_DTYPE)
    return _arr(lib().cvtx_F3D_S2S_dvort(rf, ri), 3)


def F3D_M2S_vel(fils, mes):
    f, keep = _ptrs(fils, F3D_DTYPE)
    return _arr(lib().cvtx_F3D_M2S_vel(_vp(f), len(f), _v3(mes)), 3)


def F3D_M2S_dvort(fils, ind):
    f, keep = _ptrs(fils, F3D_DTYPE)
    ri, k2 = _rec(ind, P3D_DTYPE)
    return _arr(lib().cvtx_F3D_M2S_dvort(_vp(f), len(f), ri), 3)


def P2D_S2S_vel(p, mes, kernel, sigma):
    rp, k1 = _rec(p, P2D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P2D_S2S_vel(rp, _v2(mes), kp, float(sigma)), 2)


def P2D_S2M_vel(p, mes, kernel, sigma):
    return _s2m_points(lib().cvtx_P2D_S2M_vel, p, mes, kernel, sigma, P2D_DTYPE, 2)


def P2D_M2S_vel(particles, mes, kernel, sigma):
    p, keep = _ptrs(particles, P2D_DTYPE)
    k, kp = _kref(kernel)
    return _arr(lib().cvtx_P2D_M2S_vel(_vp(p), len(p), _v2(mes), kp, float(sigma)), 2)


def P2D_S2S_visc_dvort(p, ind, kernel, sigma, nu):
    rp, k1 = _rec(p, P2D_DTYPE)
    ri, k2 = _rec(ind, P2D_DTYPE)
    k, kp = _kref(kernel)
    return float(lib().cvtx_P2D_S2S_visc_dvort(rp, ri, kp, float(sigma), float(nu)))


def P2D_S2M_visc_dvort(p, induced, kernel, sigma, nu):
    return _s2m_records(lib().cvtx_P2D_S2M_visc_dvort, p, induced, kernel, P2D_DTYPE, 1, sigma, nu)


def P2D_M2S_visc_dvort(particles, ind, kernel, sigma, nu):
    p, keep = _ptrs(particles, P2D_DTYPE)
    ri, k2 = _rec(ind, P2D_DTYPE)
    k, kp = _kref(kernel)
    return float(lib().cvtx_P2D_M2S_visc_dvort(_vp(p), len(p), ri, kp, float(sigma), float(nu)))


# ------------------- redistribution on a grid (remeshing) -------------------

REDIST_NAMES = ("lambda0", "lambda1", "lambda2", "lambda3", "m4p")


def redistfunc(name):
    """The library's cvtx_RedistFunc for a kernel name (libcvtx.h:118-123)."""
    if name not in REDIST_NAMES:
        raise ValueError(f"unknown redistribution kernel {name!r}")
    return getattr(lib(), f"cvtx_RedistFunc_{name}")()


def _redistribute(fn, particles, dtype, redistributor, grid_density, negligible, max_out):
    p, keep = _ptrs(particles, dtype)
    r = redistfunc(redistributor) if isinstance(redistributor, str) else redistributor
    rp = ctypes.cast(ctypes.pointer(r), ctypes.c_void_p)
    n = fn(_vp(p), len(p), None, 0, rp, float(grid_density), float(negligible))
    if n < 0:
        raise CvtxError(f"redistribute_on_grid: {last_error()}")
    cap = n if max_out is None else int(max_out)
    out = np.zeros(max(cap, 1), dtype)
    m = fn(_vp(p), len(p), _vp(out), cap, rp, float(grid_density), float(negligible))
    if m < 0:
        raise CvtxError(f"redistribute_on_grid: {last_error()}")
    return out[:m].copy()


def P3D_redistribute_on_grid(particles, redistributor, grid_density, negligible=0.0, max_out=None):
    """Remeshed P3D records (libcvtx.h:250-257); max_out None keeps all."""
    return _redistribute(lib().cvtx_P3D_redistribute_on_grid, particles, P3D_DTYPE, redistributor,
                         grid_density, negligible, max_out)


def P2D_redistribute_on_grid(particles, redistributor, grid_density, negligible=0.0, max_out=None):
    """Remeshed P2D records (libcvtx.h:372-380); max_out None keeps all."""
    return _redistribute(lib().cvtx_P2D_redistribute_on_grid, particles, P2D_DTYPE, redistributor,
                         grid_density, negligible, max_out)
