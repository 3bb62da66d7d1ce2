"""Mesh parity (CPU): product mesh vs oracle restatement, and closed forms."""
import numpy as np
import pytest

import oracle_lib as O


@pytest.mark.parametrize("lat", [0, 1, 2, 3, 4, 6])
def test_lateral_coords_bitwise_equal_to_oracle(lat):
    import terraneo_amd as tn

    m = tn.Mesh(lat, 2, 0.55, 1.0)
    c = m.lateral_coords()
    ref = O.lateral_coords(lat)
    assert c.shape == ref.shape
    assert np.array_equal(c.view(np.uint64), ref.view(np.uint64))


def test_icosahedron_closed_form():
    c = O.lateral_coords(0)
    # corners are icosahedron vertices: unit length, 12 distinct, all edges equal
    pts = np.unique(c.reshape(-1, 3).round(14), axis=0)
    assert len(pts) == 12
    assert np.allclose(np.linalg.norm(pts, axis=1), 1.0, atol=1e-15)
    d = np.linalg.norm(pts[:, None] - pts[None], axis=-1)
    edges = np.sort(d[d > 1e-9])
    e = edges[0]
    assert np.allclose(edges[:60], e, rtol=1e-13)  # 30 edges, counted twice
    # edge of the unit icosahedron: 4 / sqrt(10 + 2 sqrt 5)
    assert abs(e - 4 / np.sqrt(10 + 2 * np.sqrt(5))) < 1e-14


def test_midpoint_rule_closed_form():
    c = O.lateral_coords(3)
    assert np.abs(np.linalg.norm(c, axis=-1) - 1).max() <= 4.5e-16
    # every odd node of row 0 is the normalised midpoint of its neighbours
    for d in range(10):
        for j in range(1, 8, 2):
            m = c[d, 0, j - 1] + c[d, 0, j + 1]
            assert np.allclose(c[d, 0, j], m / np.linalg.norm(m), atol=1e-16)


@pytest.mark.parametrize("lat,rad", [(2, 2), (4, 4), (5, 3)])
def test_topology_and_ownership(lat, rad):
    import terraneo_amd as tn

    m = tn.Mesh(lat, rad, 0.55, 1.0)
    om = O.Mesh(lat, (1 << rad) + 1, 0.55, 1.0)
    # unique lateral nodes of an icosahedral grid: 10 * 4^lat + 2
    assert om.ngroups == 10 * 4 ** lat + 2
    owned = m.owned_mask()
    assert np.array_equal(owned.reshape(-1), om.owned)
    assert m.owned_nodes == (10 * 4 ** lat + 2) * m.nr
    # valences: 2 poles with 5 copies, 10 vertices with 3, edge nodes with 2
    cnt = np.bincount(np.bincount(om.groups.ravel()))
    assert cnt[5] == 2 and cnt[3] == 10 and cnt[2] == 20 * ((1 << lat) - 1)  # 30 icosahedral edges minus the 10 diamond diagonals
    g, mem = m.interface_info()
    assert g == 2 + 10 + 20 * ((1 << lat) - 1)
    assert mem == 2 * 5 + 10 * 3 + 2 * 20 * ((1 << lat) - 1)
    assert np.array_equal(m.radii(), om.radii)
