"""tn_dot_noalloc (the owned-node dot product whose partial sums live in a
static device array) against tn_dot through the C ABI: two calls in a row,
the second on a smaller mesh, bit-identical results (same kernels and launch
shapes, fixed-order reductions); argument errors are reported."""
import ctypes as C

import numpy as np
import pytest


def _noalloc(lib):
    f = lib.tn_dot_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]
    return f


@pytest.mark.gpu
def test_dot_noalloc_matches_dot(cuda):
    import torch

    import terraneo_amd as tn
    from terraneo_amd.api import _ptr

    lib = tn.lib()
    f = _noalloc(lib)
    cases = [(5, 5, 3, torch.float64), (3, 3, 1, torch.float64), (4, 4, 3, torch.float32)]
    for lat, rad, ncomp, dt in cases:
        print(f"case lat={lat} rad={rad} ncomp={ncomp} {dt}", flush=True)
        m = tn.Mesh(lat, rad, 0.55, 1.0)
        g = torch.Generator(device=cuda)
        g.manual_seed(lat)
        a = (torch.rand(ncomp * m.N, generator=g, device=cuda, dtype=torch.float64) * 2 - 1).to(dt)
        b = (torch.rand(ncomp * m.N, generator=g, device=cuda, dtype=torch.float64) * 2 - 1).to(dt)
        ref = m.dot(a, b, ncomp)  # tn_dot (also makes the mesh resident)
        out = C.c_double(np.nan)
        code = tn.TN_F64 if dt == torch.float64 else tn.TN_F32
        assert f(m.h, code, ncomp, _ptr(a), _ptr(b), C.byref(out), None) == 0, lib.tn_last_error()
        assert out.value == ref, (out.value, ref)
        # a second call keeps no state from the first
        out2 = C.c_double(np.nan)
        assert f(m.h, code, ncomp, _ptr(b), _ptr(b), C.byref(out2), None) == 0
        assert out2.value == m.dot(b, b, ncomp)


@pytest.mark.gpu
def test_dot_noalloc_rejects_bad_arguments(cuda):
    import torch

    import terraneo_amd as tn
    from terraneo_amd.api import _ptr

    lib = tn.lib()
    f = _noalloc(lib)
    m = tn.Mesh(2, 2, 0.55, 1.0)
    a = torch.ones(3 * m.N, dtype=torch.float64, device=cuda)
    out = C.c_double()
    # not resident yet: no upload (which would allocate) from this entry point
    assert f(m.h, tn.TN_F64, 3, _ptr(a), _ptr(a), C.byref(out), None) != 0
    assert b"not resident" in lib.tn_last_error()
    m.dot(a, a, 3)
    assert f(m.h, tn.TN_F64, 4, _ptr(a), _ptr(a), C.byref(out), None) != 0
    assert b"ncomp" in lib.tn_last_error()
    assert f(m.h, 7, 3, _ptr(a), _ptr(a), C.byref(out), None) != 0
    assert f(m.h, tn.TN_F64, 3, _ptr(a), _ptr(a), C.byref(out), None) == 0
    assert out.value == float(3 * m.owned_nodes)
