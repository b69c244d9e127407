"""F3D influence matrix (cvtx_F3D_inf_mtrx) on the GPU against the oracle's
restatement of the reference CPU path (reference src/F3D.cpp:204-227) and
its FP64 twin.  Tolerance: relative L2 over the whole matrix <= 1e-5; where
the reference's FP32 velocity formula (src/F3D.cpp:42-53, t21 - t22) is
itself less accurate, the GPU must match FP64 to 1e-5 and be no farther
from the reference than 1.5x the reference's own error."""
import numpy as np
import pytest

import cvortex_amd as cv
from cvortex_amd import device as D
from cvortex_amd import workloads as W
from cvortex_amd.types import make_f3d
from gpuutil import assert_parity, assert_parity_limited, case, have_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device: GPU tests need one")


@pytest.mark.parametrize("nf,nm", [(1, 1), (37, 23), (300, 1000), (1025, 257)])
def test_influence_matrix_matches_oracle(orc, nf, nm):
    case(f"inf_mtrx nf={nf} nm={nm}")
    fl = W.random_f3d(nf, 600 + nf, box=1.0, strength=1.0)
    pts = W.random_p3d(nm, 700 + nm, box=1.0, vort=1.0)["coord"]
    dirs = np.random.default_rng(nm).normal(size=(nm, 3)).astype(np.float32)
    got = cv.F3D_inf_mtrx(fl, pts, dirs)
    assert got.shape == (nm, nf)
    assert_parity_limited("inf_mtrx", got, orc.F3D_inf_mtrx(fl, pts, dirs),
                          orc.f64_F3D_inf_mtrx(fl, pts, dirs), TOL)


@pytest.mark.parametrize("broken", [None, 77, 128])
def test_influence_matrix_polyline(orc, broken):
    # consecutive segments that share end points bit for bit (the kernel
    # then computes each shared point's distance terms once per row), with
    # optionally one link broken by one ulp
    case(f"inf_mtrx polyline broken={broken}")
    n = 301
    rng = np.random.default_rng(11)
    pts_line = np.cumsum(rng.normal(scale=0.05, size=(n + 1, 3)), axis=0).astype(np.float32)
    a = pts_line[:-1].copy()
    b = pts_line[1:].copy()
    if broken is not None:
        a[broken, 0] = np.nextafter(a[broken, 0], np.float32(np.inf))
    fl = make_f3d(a, b, rng.normal(size=n).astype(np.float32))
    pts = W.random_p3d(97, 5, box=1.0, vort=1.0)["coord"]
    dirs = rng.normal(size=(97, 3)).astype(np.float32)
    got = cv.F3D_inf_mtrx(fl, pts, dirs)
    assert_parity_limited("inf_mtrx polyline", got, orc.F3D_inf_mtrx(fl, pts, dirs),
                          orc.f64_F3D_inf_mtrx(fl, pts, dirs), TOL)


def test_influence_matrix_vortex_lattice(orc):
    # the use the matrix is made for: a 16 x 16 lattice of square vortex
    # rings on z = 0, collocation points at the panel centres, normals +z
    case("inf_mtrx vortex lattice")
    nx, ny = 16, 16
    dx, dy = 1.0 / nx, 1.0 / ny
    a, b = [], []
    for i in range(nx):
        for j in range(ny):
            x0, y0 = i * dx, j * dy
            corners = [(x0, y0), (x0 + dx, y0), (x0 + dx, y0 + dy), (x0, y0 + dy)]
            for k in range(4):
                a.append((*corners[k], 0.0))
                b.append((*corners[(k + 1) % 4], 0.0))
    fl = make_f3d(np.array(a, np.float32), np.array(b, np.float32), np.ones(len(a), np.float32))
    cx, cy = np.meshgrid((np.arange(nx) + 0.5) * dx, (np.arange(ny) + 0.5) * dy, indexing="ij")
    pts = np.stack([cx.ravel(), cy.ravel(), np.zeros(nx * ny)], 1).astype(np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (nx * ny, 1))
    got = cv.F3D_inf_mtrx(fl, pts, nrm)
    assert_parity_limited("inf_mtrx lattice", got, orc.F3D_inf_mtrx(fl, pts, nrm),
                          orc.f64_F3D_inf_mtrx(fl, pts, nrm), TOL)
    # at its own centre the four sides of a counter-clockwise unit square
    # ring of side L induce 2 sqrt(2) / (pi L) along +z (textbook closed form)
    own = got.reshape(nx * ny, nx * ny, 4).sum(axis=2)
    assert np.allclose(np.diag(own), 2 * np.sqrt(2) / (np.pi * dx), rtol=1e-5)


def test_influence_matrix_empty():
    case("inf_mtrx empty")
    fl = W.random_f3d(0, 1)
    pts = np.zeros((5, 3), np.float32)
    assert cv.F3D_inf_mtrx(fl, pts, pts).shape == (5, 0)
    fl = W.random_f3d(7, 2)
    assert cv.F3D_inf_mtrx(fl, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32)).shape == (0, 7)


def test_influence_matrix_device_api():
    case("inf_mtrx device API")
    fl = W.random_f3d(700, 801, box=1.0, strength=1.0)
    pts = W.random_p3d(900, 802, box=1.0, vort=1.0)["coord"]
    dirs = np.random.default_rng(9).normal(size=(900, 3)).astype(np.float32)
    fa, fb = D.f3d_soa(fl)
    m4 = np.zeros((900, 4), np.float32)
    m4[:, :3] = pts
    e4 = np.zeros((900, 4), np.float32)
    e4[:, :3] = dirs
    dfa, dfb, dm, de = (D.DeviceArray.from_host(x) for x in (fa, fb, m4, e4))
    out = D.DeviceArray((900, 700))
    D.f3d_inf_eval(dfa, dfb, 700, dm, de, 900, out)
    assert_parity("inf_mtrx device vs host", out.download(), cv.F3D_inf_mtrx(fl, pts, dirs), 1e-7)
