"""Failure reporting of the march kernels' bounded hand-off waits.

The condition is forced from the host: a wait limit of 0 polls makes every
wait whose neighbour has not yet posted give up at once.  The kernels still
run to the end (their output is then not trusted); the host-visible failure
word must turn into an error of the next synchronising call, and be cleared
by it.  With the default limit restored, the same apply is exact again.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from fields import consistent_random, smooth_eta


@pytest.mark.gpu
def test_wait_give_up_is_reported_and_cleared(cuda):
    import torch

    import terraneo_amd as tn

    lib = tn.lib()
    m = tn.Mesh(4, 4, 0.55, 1.0)
    om = O.Mesh(4, m.nr, 0.55, 1.0)
    eta = smooth_eta(om.coords, om.radii).reshape(-1)
    x = consistent_random(om, 3, 0).reshape(-1)
    op = tn.Eps(m, torch.from_numpy(eta).to(cuda), tn.TN_DIRICHLET, tn.TN_DIRICHLET)
    xd = torch.from_numpy(x).to(cuda)
    yd = torch.empty_like(xd)
    assert lib.tn_sync() == 0
    print("case: wait limit 0, 20 applies", flush=True)
    assert lib.tn_set_wait_limit(0) == 0
    try:
        for _ in range(20):
            lib.tn_eps_apply(op.h, tn.api._ptr(xd), tn.api._ptr(yd), None)
        rc = lib.tn_sync()
        msg = lib.tn_last_error()
    finally:
        assert lib.tn_set_wait_limit(1 << 20) == 0
    assert rc != 0
    assert b"hand-off wait" in msg
    # reported once, then cleared
    assert lib.tn_sync() == 0
    print("case: default wait limit, parity", flush=True)
    op.apply(xd, yd)
    assert lib.tn_sync() == 0
    ref = O.Eps(om, eta, tn.TN_DIRICHLET, tn.TN_DIRICHLET).apply(x)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err


def test_wait_limit_rejects_negative():
    import terraneo_amd as tn

    assert tn.lib().tn_set_wait_limit(-1) != 0
    assert b"wait limit" in tn.lib().tn_last_error()
