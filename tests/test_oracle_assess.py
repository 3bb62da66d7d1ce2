"""Oracle pin: the reference's measured Stokes assess L2 errors (CPU).

The reference's tests/test_epsilon_divdiv_stokes_assess.cpp, built and run
(Kokkos-OpenMP) when this path was surveyed, printed for level 3
vel 2.88619e-05, pre 4.63833e-04.  The oracle's exact (direct) solve of the
same discrete system reproduces them; this fixes mesh, the one-point epsilon
operator with Dirichlet elimination, B/B^T (including the centre-wedge
routing), the vector mass RHS and the interface sums at once.
"""
import assess as AS
from stokes_direct import solve_assess_dirichlet


def test_assess_dirichlet_level3_matches_reference():
    ev, ep, _ = solve_assess_dirichlet(3)
    rv, rp = AS.REFERENCE_ERRORS_DIRICHLET[3]
    assert abs(ev - rv) / rv < 2e-5, (ev, rv)
    assert abs(ep - rp) / rp < 2e-5, (ep, rp)
