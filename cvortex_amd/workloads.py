"""Seeded synthetic workloads for the direct-sum benchmarks and tests.

Random numbers come from a counter-based SplitMix64 stream: uniforms from the
top 53 bits of each 64-bit output, normals from the Box-Muller transform in
FP64, rounded to FP32 only when stored.  Stream seed for workload k is
0x5EED0000 + k.

Workloads (sizes are the defaults; every builder takes the size so that
tests can run the same construction small):

  cube      N particles uniform in [-1,1]^3, alpha ~ N(0,1) * 8/N, volume 8/N,
            targets = the particles' positions; Winckelmans sigma = 0.05.
  ring      N particles on a thin vortex ring: radius 1 in z = 0, angle
            uniform, radial and axial offsets ~ N(0, 0.05^2), vorticity
            tangent with the Gaussian core profile Gamma/(2 pi s^2)
            exp(-rho^2/(2 s^2)) (Gamma = 1, s = 0.05) times the particle
            volume 2 pi R 2 pi s^2 / N; Gaussian sigma = 0.02, nu = 1e-3.
  lattice   a 128 x 128 horseshoe-lattice sheet of straight filaments on
            z = 0 over [0,4] x [0,1]: one spanwise bound segment per cell at
            quarter chord, strength ~ N(1, 0.1^2); targets N particles
            uniform in [0,6] x [-0.5,1.5] x [-0.5,0.5] with alpha ~ N(0,1) 8/N.
  clouds    C independent clouds, sizes round(exp U[ln 1024, ln 65536]),
            each uniform in the unit cube, alpha ~ N(0,1)/N_c, volume 1/N_c,
            Winckelmans sigma_c ~ U[0.01, 0.1].
  sheet     N 2D particles on y = 0.01 sin(2 pi x), x_k = (k + 1/2)/N, with
            N(0, 1e-4^2) jitter on both coordinates, circulation 1/N,
            area 1/N^2; planetary sigma = 4/sqrt(N), nu = 1e-4.
"""
import math

import numpy as np

from .types import F3D_DTYPE, P2D_DTYPE, P3D_DTYPE

SEED_BASE = 0x5EED0000
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based SplitMix64 stream (vectorised draws)."""

    def __init__(self, seed):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def _raw(self, n):
        with np.errstate(over="ignore"):
            idx = np.arange(self.count + 1, self.count + n + 1, dtype=np.uint64)
            z = self.state + idx * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.count += n
        return z

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self._raw(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        t = 2.0 * np.pi * u2
        z = np.concatenate([r * np.cos(t), r * np.sin(t)])
        return z[:n]


def cube(n=8192, seed=SEED_BASE + 1):
    """Uniform cube cloud.  Returns (particles, sigma)."""
    g = SplitMix64(seed)
    p = np.zeros(n, P3D_DTYPE)
    p["coord"] = g.uniform(3 * n, -1.0, 1.0).reshape(n, 3)
    s = 8.0 / n
    p["vorticity"] = (g.normal(3 * n) * s).reshape(n, 3)
    p["volume"] = s
    return p, 0.05


def ring(n=1 << 20, seed=SEED_BASE + 2, R=1.0, core=0.05, gamma=1.0):
    """Thin vortex ring.  Returns (particles, sigma, nu)."""
    g = SplitMix64(seed)
    theta = g.uniform(n, 0.0, 2.0 * np.pi)
    off = g.normal(2 * n).reshape(2, n) * core
    dr, dz = off[0], off[1]
    rad = R + dr
    p = np.zeros(n, P3D_DTYPE)
    p["coord"][:, 0] = rad * np.cos(theta)
    p["coord"][:, 1] = rad * np.sin(theta)
    p["coord"][:, 2] = dz
    vol = 2.0 * np.pi * R * 2.0 * np.pi * core * core / n
    omega = gamma / (2.0 * np.pi * core * core) * np.exp(-(dr * dr + dz * dz) / (2.0 * core * core))
    p["vorticity"][:, 0] = -np.sin(theta) * omega * vol
    p["vorticity"][:, 1] = np.cos(theta) * omega * vol
    p["vorticity"][:, 2] = 0.0
    p["volume"] = vol
    return p, 0.02, 1e-3


def lattice(n_targets=1 << 20, nx=128, ny=128, seed=SEED_BASE + 3):
    """Filament sheet and target particles.  Returns (filaments, targets)."""
    g = SplitMix64(seed)
    dx, dy = 4.0 / nx, 1.0 / ny
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    f = np.zeros(nx * ny, F3D_DTYPE)
    xq = ii * dx + dx / 4.0
    f["start"][:, 0] = xq
    f["start"][:, 1] = jj * dy
    f["end"][:, 0] = xq
    f["end"][:, 1] = (jj + 1) * dy
    f["strength"] = 1.0 + 0.1 * g.normal(nx * ny)
    t = np.zeros(n_targets, P3D_DTYPE)
    u = g.uniform(3 * n_targets).reshape(n_targets, 3)
    t["coord"][:, 0] = 6.0 * u[:, 0]
    t["coord"][:, 1] = -0.5 + 2.0 * u[:, 1]
    t["coord"][:, 2] = -0.5 + u[:, 2]
    s = 8.0 / n_targets
    t["vorticity"] = (g.normal(3 * n_targets) * s).reshape(n_targets, 3)
    t["volume"] = s
    return f, t


def cloud_sizes(n_clouds=512, lo=1024, hi=65536, seed=SEED_BASE + 4):
    g = SplitMix64(seed)
    return np.rint(np.exp(g.uniform(n_clouds, math.log(lo), math.log(hi)))).astype(np.int64)


def clouds(n_clouds=512, lo=1024, hi=65536, seed=SEED_BASE + 4):
    """Independent clouds.  Returns (particles concatenated, offsets[n+1], sigmas)."""
    sizes = cloud_sizes(n_clouds, lo, hi, seed)
    g = SplitMix64(seed + 0x100)
    off = np.zeros(n_clouds + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    tot = int(off[-1])
    p = np.zeros(tot, P3D_DTYPE)
    p["coord"] = g.uniform(3 * tot).reshape(tot, 3)
    nrm = g.normal(3 * tot).reshape(tot, 3)
    per = np.repeat(sizes, sizes).astype(np.float64)
    p["vorticity"] = nrm / per[:, None]
    p["volume"] = 1.0 / per
    sig = g.uniform(n_clouds, 0.01, 0.1).astype(np.float32)
    return p, off, sig


def sheet(n=1 << 22, seed=SEED_BASE + 5):
    """2D vortex sheet.  Returns (particles, sigma, nu)."""
    g = SplitMix64(seed)
    x = (np.arange(n, dtype=np.float64) + 0.5) / n
    y = 0.01 * np.sin(2.0 * np.pi * x)
    j = g.normal(2 * n).reshape(2, n) * 1e-4
    p = np.zeros(n, P2D_DTYPE)
    p["coord"][:, 0] = x + j[0]
    p["coord"][:, 1] = y + j[1]
    p["vorticity"] = 1.0 / n
    p["area"] = 1.0 / (float(n) * n)
    return p, 4.0 / math.sqrt(n), 1e-4


def random_p3d(n, seed, box=10.0, vort=10.0, vol=0.01):
    """Particles like the reference's CPU/GPU agreement tests
    (test/testsamecpugpuresultmany.h:69-79): coordinates and vorticity
    uniform in [0, box] and [0, vort], volume uniform in [0, vol]."""
    g = SplitMix64(seed)
    p = np.zeros(n, P3D_DTYPE)
    p["coord"] = g.uniform(3 * n, 0.0, box).reshape(n, 3)
    p["vorticity"] = g.uniform(3 * n, 0.0, vort).reshape(n, 3)
    p["volume"] = g.uniform(n, 0.0, vol)
    return p


def random_f3d(n, seed, box=10.0, strength=10.0):
    """Filaments like test/testsamecpugpuresultmany.h:80-89."""
    g = SplitMix64(seed)
    f = np.zeros(n, F3D_DTYPE)
    f["start"] = g.uniform(3 * n, 0.0, box).reshape(n, 3)
    f["end"] = g.uniform(3 * n, 0.0, box).reshape(n, 3)
    f["strength"] = g.uniform(n, 0.0, strength)
    return f


def random_p2d(n, seed, box=10.0, vort=10.0, area=0.01):
    g = SplitMix64(seed)
    p = np.zeros(n, P2D_DTYPE)
    p["coord"] = g.uniform(2 * n, 0.0, box).reshape(n, 2)
    p["vorticity"] = g.uniform(n, 0.0, vort)
    p["area"] = g.uniform(n, 0.0, area)
    return p
