"""The C ABI library loads on the CPU and exports every declared entry point."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"\b(tn_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_declared_symbols():
    import terraneo_amd

    lib = ctypes.CDLL(terraneo_amd.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from terraneo_amd import _capi

    names = declared_symbols()
    assert set(names) <= set(_capi.SIGNATURES), sorted(set(names) - set(_capi.SIGNATURES))


def test_error_path_without_device():
    import terraneo_amd as tn

    # invalid arguments are rejected on the host with a message
    h = ctypes.c_void_p()
    rc = tn.lib().tn_mesh_create(3, 0, 0.5, 1.0, ctypes.byref(h))
    assert rc != 0
    assert b"radial level" in tn.lib().tn_last_error()


def test_mesh_subdomain_split_rejected():
    import terraneo_amd as tn

    h = ctypes.c_void_p()
    rc = tn.lib().tn_mesh_create_subdomains(3, 3, 0.55, 1.0, 1, 0, ctypes.byref(h))
    assert rc != 0
    assert b"subdomains" in tn.lib().tn_last_error()
    rc = tn.lib().tn_mesh_create_subdomains(3, 3, 0.55, 1.0, 0, 0, ctypes.byref(h))
    assert rc == 0
    n, nr, N = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
    assert tn.lib().tn_mesh_dims(h, ctypes.byref(n), ctypes.byref(nr), ctypes.byref(N)) == 0
    assert (n.value, nr.value) == (9, 9)
    tn.lib().tn_mesh_destroy(h)
