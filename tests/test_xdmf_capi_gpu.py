"""The C-ABI XDMF writer and checkpoint reader (tn_xdmf_*) against the
Python mirror of the reference's format (terraneo_amd/xdmf.py, whose CPU
tests pin the format): both writers produce byte-identical directories for
the same fields and steps, and the C-ABI reader returns the written fields
(bit-exact for 8-byte output, FP32 rounding for 4-byte output)."""
import os

import numpy as np
import pytest


def _files(d):
    return sorted(os.listdir(d))


@pytest.mark.gpu
@pytest.mark.parametrize("points_bytes,conn_bytes", [(4, 8), (8, 4)])
def test_capi_writer_matches_python_writer(cuda, tmp_path, points_bytes, conn_bytes):
    import torch

    from terraneo_amd import api
    from terraneo_amd.xdmf import XDMFOutput

    m = api.Mesh(3, 3, 0.55, 1.0)
    g = torch.Generator(device=cuda)
    g.manual_seed(0)
    u = torch.randn(3 * m.N, generator=g, device=cuda, dtype=torch.float64)
    T = torch.rand(m.N, generator=g, device=cuda, dtype=torch.float32)
    eta = torch.rand(m.N, generator=g, device=cuda, dtype=torch.float64)
    a, b = tmp_path / "py", tmp_path / "capi"
    py = XDMFOutput(str(a), m, points_bytes, conn_bytes)
    cw = api.XDMFWriter(m, b, points_bytes, conn_bytes)
    for w in (py, cw):
        w.add("velocity", u, vec_dim=3, out_bytes=8)
        w.add("temperature", T, vec_dim=1, out_bytes=4)
        w.add("viscosity", eta, vec_dim=1, out_bytes=4)
    written = []
    for step in range(2):
        py.write()
        cw.write()
        written.append(u.clone())
        u.mul_(1.5)
    assert _files(a) == _files(b)
    for f in _files(a):
        assert (a / f).read_bytes() == (b / f).read_bytes(), f
    # reader: bit-exact for 8-byte data, FP32 rounding of 4-byte data
    u1 = api.read_checkpoint_grid(m, b, "velocity", 1, vec_dim=3)
    assert torch.equal(u1, written[1])
    T0 = api.read_checkpoint_grid(m, b, "temperature", 0, vec_dim=1, dtype=torch.float32)
    assert torch.equal(T0, T)
    e0 = api.read_checkpoint_grid(m, a, "viscosity", 0, vec_dim=1)
    assert torch.equal(e0, eta.float().double())


@pytest.mark.gpu
def test_capi_reader_errors(cuda, tmp_path):
    import torch

    from terraneo_amd import api
    from terraneo_amd._capi import TNError

    m = api.Mesh(2, 2, 0.55, 1.0)
    x = torch.zeros(m.N, dtype=torch.float64, device=cuda)
    w = api.XDMFWriter(m, tmp_path)
    w.add("p", x)
    with pytest.raises(TNError, match="added before"):
        w.add("p", x)
    w.write()
    with pytest.raises(TNError, match="after write"):
        w.add("q", x)
    with pytest.raises(TNError, match="no data 'q'"):
        api.read_checkpoint_grid(m, tmp_path, "q", 0)
    with pytest.raises(TNError, match="vec_dim"):
        api.read_checkpoint_grid(m, tmp_path, "p", 0, vec_dim=3)
    other = api.Mesh(3, 2, 0.55, 1.0)
    with pytest.raises(TNError, match="extents"):
        api.read_checkpoint_grid(other, tmp_path, "p", 0)
