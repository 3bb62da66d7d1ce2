"""Build libterraneo_amd.so (host C++ + HIP kernels for gfx950) in-tree.

Usage:  python -m terraneo_amd.build      (or terraneo_amd.build.build())
"""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libterraneo_amd.so")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("TN_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 VALU ops (v_pk_*_f32) in any device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(PKG, "..", "include"),
]
# host-side mesh generation must not contract into FMAs (bit-exact coordinates)
PER_FILE = {"mesh.cpp": ["-ffp-contract=off"]}


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + PER_FILE.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lrocsolver", "-lrocblas", "-Wl,--no-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_driver(verbose)
    return LIB


DRIVER_SRC = os.path.join(PKG, "..", "tests", "capi", "capi_driver.cpp")
DRIVER = os.path.join(PKG, "..", "tests", "capi", "capi_driver")


def build_driver(verbose=False):
    """The C++ caller of the C ABI (tests/capi/capi_driver.cpp), header-only use
    of include/terraneo_amd.h, linked against the library just built."""
    if not os.path.exists(DRIVER_SRC):
        return None
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "lib")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(PKG, "..", "include"), DRIVER_SRC,
           "-o", DRIVER + ".tmp", "-L", PKG, "-lterraneo_amd", "-Wl,-rpath,$ORIGIN/../../terraneo_amd",
           "-Wl,-rpath-link," + rocm_lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
