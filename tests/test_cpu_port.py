"""The benchmark's CPU baseline (oracle/cpu_port.c, a matrix-free OpenMP
apply in this build's cell formulation) equals the oracle's assembled-matrix
apply, so the baseline times the same operator."""
import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random, smooth_eta


@pytest.mark.parametrize("lat,nr", [(2, 5), (3, 9), (3, 4)])
@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (0, 1), (1, 0)])
def test_cpu_port_matches_oracle(lat, nr, bc):
    om = O.Mesh(lat, nr, 0.55, 1.0)
    eta = smooth_eta(om.coords, om.radii).reshape(-1)
    x = consistent_random(om, 3, 0).reshape(-1)
    ref = O.Eps(om, eta, *bc).apply(x)
    y = O.cpu_port_apply(om, eta, x, *bc)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
