"""The reference's public void M2M functions, called as a drop-in caller would.

Unlike the wrappers in :mod:`cvortex_amd` (which call the int-returning
seam and raise on failure), these call the public ``cvtx_*`` symbols
(reference include/cvortex/libcvtx.h:213-305, 329-370): the library picks
the device or its CPU leg (``cvtx_hip_set_dispatch``), and a failure gives
NaN results plus a line on stderr.  ``last_leg()`` tells which leg ran.
"""
import ctypes

import numpy as np

from . import _kref, _ptrs, _vp, lib, last_leg  # noqa: F401
from .types import F3D_DTYPE, P2D_DTYPE, P3D_DTYPE

_VP = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float

_SIGS = {
    "cvtx_P3D_M2M_vel": [_VP, _I, _VP, _I, _VP, _VP, _F],
    "cvtx_P3D_M2M_dvort": [_VP, _I, _VP, _I, _VP, _VP, _F],
    "cvtx_P3D_M2M_visc_dvort": [_VP, _I, _VP, _I, _VP, _VP, _F, _F],
    "cvtx_P3D_M2M_vort": [_VP, _I, _VP, _I, _VP, _VP, _F],
    "cvtx_P3D_pedrizzetti_relaxation": [_VP, _I, _F, _VP, _F],
    "cvtx_F3D_M2M_vel": [_VP, _I, _VP, _I, _VP],
    "cvtx_F3D_M2M_dvort": [_VP, _I, _VP, _I, _VP],
    "cvtx_F3D_inf_mtrx": [_VP, _I, _VP, _VP, _I, _VP],
    "cvtx_P2D_M2M_vel": [_VP, _I, _VP, _I, _VP, _VP, _F],
    "cvtx_P2D_M2M_visc_dvort": [_VP, _I, _VP, _I, _VP, _VP, _F, _F],
}


def _fn(name):
    f = getattr(lib(), name)
    f.restype = None
    f.argtypes = _SIGS[name]
    return f


def _pts(x, d):
    return np.ascontiguousarray(x, dtype=np.float32).reshape(-1, d)


def P3D_M2M_vel(particles, mes, kernel, sigma):
    p, keep = _ptrs(particles, P3D_DTYPE)
    mes = _pts(mes, 3)
    out = np.zeros((len(mes), 3), np.float32)
    k, kp = _kref(kernel)
    _fn("cvtx_P3D_M2M_vel")(_vp(p), len(p), _vp(mes), len(mes), _vp(out), kp, float(sigma))
    return out


def P3D_M2M_dvort(particles, induced, kernel, sigma):
    p, k1 = _ptrs(particles, P3D_DTYPE)
    q, k2 = _ptrs(induced, P3D_DTYPE)
    out = np.zeros((len(q), 3), np.float32)
    k, kp = _kref(kernel)
    _fn("cvtx_P3D_M2M_dvort")(_vp(p), len(p), _vp(q), len(q), _vp(out), kp, float(sigma))
    return out


def P3D_M2M_visc_dvort(particles, induced, kernel, sigma, nu):
    p, k1 = _ptrs(particles, P3D_DTYPE)
    q, k2 = _ptrs(induced, P3D_DTYPE)
    out = np.zeros((len(q), 3), np.float32)
    k, kp = _kref(kernel)
    _fn("cvtx_P3D_M2M_visc_dvort")(_vp(p), len(p), _vp(q), len(q), _vp(out), kp, float(sigma),
                                   float(nu))
    return out


def P3D_M2M_vort(particles, mes, kernel, sigma):
    p, keep = _ptrs(particles, P3D_DTYPE)
    mes = _pts(mes, 3)
    out = np.zeros((len(mes), 3), np.float32)
    k, kp = _kref(kernel)
    _fn("cvtx_P3D_M2M_vort")(_vp(p), len(p), _vp(mes), len(mes), _vp(out), kp, float(sigma))
    return out


def P3D_pedrizzetti_relaxation(particles, fdt, kernel, sigma):
    """Returns a relaxed copy of the records."""
    recs = np.array(particles, dtype=P3D_DTYPE, copy=True)
    p, keep = _ptrs(recs, P3D_DTYPE)
    k, kp = _kref(kernel)
    _fn("cvtx_P3D_pedrizzetti_relaxation")(_vp(p), len(recs), float(fdt), kp, float(sigma))
    return recs


def F3D_M2M_vel(filaments, mes):
    f, keep = _ptrs(filaments, F3D_DTYPE)
    mes = _pts(mes, 3)
    out = np.zeros((len(mes), 3), np.float32)
    _fn("cvtx_F3D_M2M_vel")(_vp(f), len(f), _vp(mes), len(mes), _vp(out))
    return out


def F3D_M2M_dvort(filaments, induced):
    f, k1 = _ptrs(filaments, F3D_DTYPE)
    q, k2 = _ptrs(induced, P3D_DTYPE)
    out = np.zeros((len(q), 3), np.float32)
    _fn("cvtx_F3D_M2M_dvort")(_vp(f), len(f), _vp(q), len(q), _vp(out))
    return out


def F3D_inf_mtrx(filaments, mes, dirs):
    f, keep = _ptrs(filaments, F3D_DTYPE)
    mes = _pts(mes, 3)
    dirs = _pts(dirs, 3)
    out = np.zeros((len(mes), len(f)), np.float32)
    _fn("cvtx_F3D_inf_mtrx")(_vp(f), len(f), _vp(mes), _vp(dirs), len(mes), _vp(out))
    return out


def P2D_M2M_vel(particles, mes, kernel, sigma):
    p, keep = _ptrs(particles, P2D_DTYPE)
    mes = _pts(mes, 2)
    out = np.zeros((len(mes), 2), np.float32)
    k, kp = _kref(kernel)
    _fn("cvtx_P2D_M2M_vel")(_vp(p), len(p), _vp(mes), len(mes), _vp(out), kp, float(sigma))
    return out


def P2D_M2M_visc_dvort(particles, induced, kernel, sigma, nu):
    p, k1 = _ptrs(particles, P2D_DTYPE)
    q, k2 = _ptrs(induced, P2D_DTYPE)
    out = np.zeros(len(q), np.float32)
    k, kp = _kref(kernel)
    _fn("cvtx_P2D_M2M_visc_dvort")(_vp(p), len(p), _vp(q), len(q), _vp(out), kp, float(sigma),
                                   float(nu))
    return out
