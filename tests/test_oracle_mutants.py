"""Every planted mistake in the oracle is caught by the pinning checks.

oracle/refcpu.c has ORC_MUT(k) sites, each changing one formula in a way a
restatement could plausibly get wrong (wrong exponent, wrong constant,
operand order in a cross product, dropped term, flipped sign, swapped
components, off-by-one loop bound).  Each variant is compiled separately
(make -C oracle mutants) and must fail at least one check of tests/pins.py;
the check expected to catch it is listed so that a weakened check shows up.
"""
import os
import subprocess

import pytest

import oracle
import pins

HERE = os.path.dirname(os.path.abspath(oracle.__file__))

EXPECTED = {
    1: "kat_vortfunc",          # Winckelmans g exponent -2.5 -> -1.5
    2: "kat_vortfunc",          # Gaussian zeta constant sqrt(2/pi) -> 1/sqrt(2 pi)
    3: "particle_invariants",   # d x alpha -> alpha x d in the P3D velocity
    4: "dvort_transpose_fd",    # P3D dvort without the gradient term
    5: "pse_conservation",      # PSE sign of the target term
    6: "f3d_vel_quadrature",    # F3D velocity t21 - t22 -> t21 + t22
    7: "f3d_dvort_structure",   # A x alpha -> alpha x A
    8: "p2d_point_vortex",      # 2D velocity components swapped
    9: "moments_3d",            # Winckelmans eta 52.5 -> 50
    10: "ring_centre",          # M2M loop drops the last source
    11: "kernels_2d",           # Gaussian eta_2D exp(-rho^2/2) -> exp(-rho^2)
    12: "kernels_2d",           # Winckelmans g_2D numerator rho^4 + 2 rho^2 -> rho^4 + rho^2
    13: "vort_box_cutoff",      # vorticity cut-off 5 sigma -> 4 sigma
    14: "vort_box_cutoff",      # vorticity cut-off box -> sphere
    15: "relaxation_invariants",  # relaxation keeps fdt alpha instead of (1 - fdt) alpha
    16: "influence_matrix_rows",  # influence matrix stored column-major
    17: "redistribution_moments",  # m4p inner coefficient 2.5 -> 2.0
    18: "redistribution_moments",  # removed strength not redistributed
    19: "f3d_dvort_second_statement",  # F3D dvort isotropic coefficient 3 -> 2
    20: "pse_laplacian",        # 3D PSE prefactor 2 nu / sigma^2 -> nu / sigma^2
    21: "pse_laplacian",        # 3D PSE volumes exchanged: a_s V_s - a_t V_t
    22: "pse_laplacian",        # 2D PSE prefactor 2 nu / sigma^2 -> nu / sigma^2
    23: "pse_laplacian",        # 2D PSE areas exchanged: a_s A_s - a_t A_t
}


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", HERE, "mutants"], check=True)
    return True


@pytest.mark.parametrize("k", sorted(EXPECTED))
def test_mutant_rejected(built, k):
    o = oracle.mutant(k)
    failed = {c.__name__ for c in pins.ALL_CHECKS if c(o)}
    assert failed, f"planted mistake {k} passed every pinning check"
    assert EXPECTED[k] in failed, f"mutant {k} caught only by {sorted(failed)}"
