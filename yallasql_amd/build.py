"""Build the gfx950 kernel library (libyallasql_amd.so) in-tree with hipcc.

Usage: python -m yallasql_amd.build   (or yallasql_amd.build.build_library())

Every .hip / .cpp file under csrc/ is compiled to an object in _build/ and
the objects are linked into yallasql_amd/libyallasql_amd.so, next to this
file, so the library travels with the source tree.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libyallasql_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed-FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + HIP_FLAGS + lang + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp"))
    )
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


COMPAT_SRC = os.path.join(ROOT, "tests", "compat", "compat_launchers.cpp")
COMPAT_LIB = os.path.join(ROOT, "tests", "compat", "libcompat_launchers.so")


def build_compat(verbose=True):
    """The test library that drives include/ysql_kernels.hpp's launchers
    (tests/compat/compat_launchers.cpp), linked against the kernel library."""
    if not os.path.exists(COMPAT_SRC):
        return None
    tmp = COMPAT_LIB + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}", "-Wno-unused-result",
           "-I", os.path.join(ROOT, "include"), COMPAT_SRC, "-o", tmp, "-L", PKG, "-lyallasql_amd",
           "-Wl,-rpath,$ORIGIN/../../yallasql_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, COMPAT_LIB)
    if verbose:
        print(f"built {COMPAT_LIB}")
    return COMPAT_LIB


if __name__ == "__main__":
    build_library()
    build_compat()
    sys.exit(0)
