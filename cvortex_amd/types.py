"""ctypes / numpy mirrors of the cvortex C ABI types.

Layouts match include/cvortex/libcvtx.h (and the reference header
include/cvortex/libcvtx.h:53-94):

    cvtx_P3D  28 bytes  {coord float[3], vorticity float[3], volume float}
    cvtx_F3D  28 bytes  {start float[3], end float[3], strength float}
    cvtx_P2D  16 bytes  {coord float[2], vorticity float, area float}
    bsv_V3f   12 bytes, bsv_V2f 8 bytes
    cvtx_VortFunc 80 bytes (six function pointers + char[32])

Arrays of records are numpy structured arrays; the C API takes arrays of
pointers to records (``const cvtx_P3D**``), built by :func:`pointer_array`.
This module loads no shared library.
"""
import ctypes

import numpy as np

P3D_DTYPE = np.dtype([("coord", "<f4", (3,)), ("vorticity", "<f4", (3,)), ("volume", "<f4")])
F3D_DTYPE = np.dtype([("start", "<f4", (3,)), ("end", "<f4", (3,)), ("strength", "<f4")])
P2D_DTYPE = np.dtype([("coord", "<f4", (2,)), ("vorticity", "<f4"), ("area", "<f4")])

assert P3D_DTYPE.itemsize == 28 and F3D_DTYPE.itemsize == 28 and P2D_DTYPE.itemsize == 16


class V3f(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float * 3)]


class V2f(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float * 2)]


class RedistFunc(ctypes.Structure):
    """cvtx_RedistFunc: interpolation kernel plugin (func pointer + radius)."""
    _fields_ = [("func", ctypes.c_void_p), ("radius", ctypes.c_float)]


class P3D(ctypes.Structure):
    _fields_ = [("coord", V3f), ("vorticity", V3f), ("volume", ctypes.c_float)]


class F3D(ctypes.Structure):
    _fields_ = [("start", V3f), ("end", V3f), ("strength", ctypes.c_float)]


class P2D(ctypes.Structure):
    _fields_ = [("coord", V2f), ("vorticity", ctypes.c_float), ("area", ctypes.c_float)]


FLOAT_FN = ctypes.CFUNCTYPE(ctypes.c_float, ctypes.c_float)
COMBINED_FN = ctypes.CFUNCTYPE(None, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                               ctypes.POINTER(ctypes.c_float))


class VortFunc(ctypes.Structure):
    _fields_ = [
        ("g_3D", FLOAT_FN),
        ("g_2D", FLOAT_FN),
        ("zeta_3D", FLOAT_FN),
        ("combined_3D", COMBINED_FN),
        ("eta_3D", FLOAT_FN),
        ("eta_2D", FLOAT_FN),
        ("cl_kernel_name_ext", ctypes.c_char * 32),
    ]


assert ctypes.sizeof(VortFunc) == 80

REG_NAMES = ("singular", "winckelmans", "planetary", "gaussian")
REG_ID = {name: i for i, name in enumerate(REG_NAMES)}


def as_records(arr, dtype):
    """Return a C-contiguous structured array of `dtype` (no copy if already)."""
    arr = np.asarray(arr)
    if arr.dtype != dtype:
        raise TypeError(f"expected dtype {dtype}, got {arr.dtype}")
    return np.ascontiguousarray(arr)


def pointer_array(records):
    """Array of pointers to each record (the reference's ``const T**`` inputs).

    Returns (ptrs, keepalive): ptrs is a uint64 numpy array whose element i is
    the address of records[i]."""
    n = records.shape[0]
    base = records.ctypes.data
    ptrs = base + np.arange(n, dtype=np.uint64) * np.uint64(records.dtype.itemsize)
    return np.ascontiguousarray(ptrs, dtype=np.uint64), records


def ptr(a):
    """ctypes void pointer to a numpy array's data (None for empty/None)."""
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


def make_p3d(coord, vorticity, volume):
    n = len(coord)
    out = np.zeros(n, P3D_DTYPE)
    out["coord"] = coord
    out["vorticity"] = vorticity
    out["volume"] = volume
    return out


def make_f3d(start, end, strength):
    n = len(start)
    out = np.zeros(n, F3D_DTYPE)
    out["start"] = start
    out["end"] = end
    out["strength"] = strength
    return out


def make_p2d(coord, vorticity, area):
    n = len(coord)
    out = np.zeros(n, P2D_DTYPE)
    out["coord"] = coord
    out["vorticity"] = vorticity
    out["area"] = area
    return out
