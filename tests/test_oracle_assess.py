"""Oracle pin: the reference's measured Stokes assess L2 errors (CPU).

The reference's assess tests, built and run (Kokkos-OpenMP) when this path
was surveyed, printed at level 3:
  zero-slip both shells   (test_epsilon_divdiv_stokes_assess.cpp)             vel 2.88619e-05, pre 4.63833e-04
  free-slip both shells   (test_epsilon_divdiv_stokes_assess_freeslip_freeslip.cpp) vel 3.09494e-05, pre 5.53425e-04
  free-slip CMB, zero-slip surface (test_epsilon_divdiv_stokes_assess_freeslip.cpp) vel 3.19586e-05, pre 4.56753e-04
The oracle's exact (direct) solve of the same discrete systems reproduces
them, which fixes the mesh, the one-point epsilon operator with Dirichlet
elimination, the free-slip rotation/projection and normal diagonal, the
rotation penalty (added per node copy), B/B^T (with the centre-wedge routing
and the free-slip projection), the vector mass RHS and the interface sums.
"""
import pytest

import assess as AS
from stokes_direct import solve_assess, solve_assess_dirichlet


def test_assess_dirichlet_level3_matches_reference():
    ev, ep, _ = solve_assess_dirichlet(3)
    rv, rp = AS.REFERENCE_ERRORS_DIRICHLET[3]
    assert abs(ev - rv) / rv < 2e-5, (ev, rv)
    assert abs(ep - rp) / rp < 2e-5, (ep, rp)


@pytest.mark.parametrize("kind", ["freeslip", "freezeroslip"])
def test_assess_freeslip_level3_matches_reference(kind):
    ev, ep = solve_assess(kind, 3)
    rv, rp = (AS.REFERENCE_ERRORS_FREESLIP if kind == "freeslip" else AS.REFERENCE_ERRORS_FREEZEROSLIP)[3]
    assert abs(ev - rv) / rv < 2e-5, (ev, rv)
    assert abs(ep - rp) / rp < 2e-5, (ep, rp)
