"""The Gaussian near-field table used by the kernels is accurate.

The table holds cubic pieces of G(rho) = g(rho)/rho^3 and of H(rho)/3, with
H(rho) = (3G - zeta)/rho^2, on [0, 6).  G and H are integrals,
G = sqrt(2/pi) int_0^1 s^2 exp(-rho^2 s^2 / 2) ds and H the same with s^4,
evaluated here by mpmath quadrature -- independently of the library's
series/erf construction and of the oracle."""
import ctypes
import math

import mpmath
import numpy as np

import cvortex_amd as cv


def exact(rho):
    c = mpmath.sqrt(2 / mpmath.pi)
    G = c * mpmath.quad(lambda s: s * s * mpmath.exp(-rho * rho * s * s / 2), [0, 1])
    H = c * mpmath.quad(lambda s: s ** 4 * mpmath.exp(-rho * rho * s * s / 2), [0, 1])
    return float(G), float(H)


def test_table_accuracy():
    buf = (ctypes.c_float * 768)()
    assert cv.lib().cvtx_hip_gaussian_table(buf, 768) == 768
    tab = np.frombuffer(buf, np.float32).reshape(96, 2, 4).astype(np.float64)
    worst = 0.0
    for rho in np.concatenate([np.linspace(0, 5.999, 241), [1e-6, 1e-3, 0.03, 0.0625, 5.9375]]):
        i = min(int(rho * 16), 95)
        u = rho * 16 - i
        cg, ch = tab[i]
        G = ((cg[3] * u + cg[2]) * u + cg[1]) * u + cg[0]
        H = 3.0 * (((ch[3] * u + ch[2]) * u + ch[1]) * u + ch[0])
        Ge, He = exact(rho)
        worst = max(worst, abs(G - Ge) / Ge, abs(H - He) / He)
    assert worst < 2e-7, worst


def test_table_matches_far_field_at_the_switch():
    # at rho = 6 the far-field form rho^-3 (1 - rho zeta) is exact to 2e-9
    rho = 6.0
    z = math.sqrt(2 / math.pi) * math.exp(-rho * rho / 2)
    Gfar = (1 - rho * z) / rho ** 3
    Ge, He = exact(rho)
    assert abs(Gfar - Ge) / Ge < 1e-8
    assert abs((3 * Gfar - z) / rho ** 2 - He) / He < 1e-8
