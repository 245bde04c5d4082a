"""Build the HIP library and the pybind11 module in-tree for gfx950.

Outputs (next to this file):
  libpolymlp_amd.so   C-ABI (include/polymlp_amd.h) + HIP kernels
  libmlpcpp*.so       pybind11 module mirroring pypolymlp's libmlpcpp surface

Usage: python -m pypolymlp_amd.build  (or build() from __graft_entry__).
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpolymlp_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-I", CSRC, "-I", os.path.join(ROOT, "include")]

# (source, extra flags)
SOURCES = [
    ("model.cpp", []),
    ("cell.cpp", ["-ffp-contract=off"]),
    ("capi.cpp", []),
    ("normal.cpp", []),
    ("ridge.cpp", []),
    ("profile.cpp", []),
    ("kernels_solve.hip", []),
    ("kernels_resid.hip", []),
    ("kernels_polyz.hip", []),
    ("eval.cpp", []),
    ("eval_noalloc.hip", []),
    ("kernels_eval.hip", []),
    ("kernels_x.hip", []),
    ("kernels_xrows.hip", []),
    ("kernels_nb.hip", []),
    ("kernels_gram.hip", []),
    ("kernels_surrogate.hip", []),
    ("kernels_poly.hip", []),
    ("poly.cpp", []),
    ("kernels_xgroup.hip", []),
    ("xgroup.cpp", []),
    ("surrogate.cpp", []),
    ("kernels_cg.hip", []),
    ("cg.cpp", []),
    ("kernels_api.hip", []),
    ("api.cpp", []),
]


def _compile(src, extra):
    out = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stderr}")
    return out


def build_library(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda a: _compile(*a), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp,
           *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_pybind() -> str:
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "libmlpcpp" + suffix)
    tmp = out + ".tmp"
    inc = sysconfig.get_paths()["include"]
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I", pybind11.get_include(),
           "-I", inc, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           os.path.join(CSRC, "pybind_module.cpp"), "-o", tmp,
           "-L", HERE, "-lpolymlp_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"pybind build failed\n{r.stderr}")
    os.replace(tmp, out)
    return out


def build_all():
    build_library()
    if os.path.exists(os.path.join(CSRC, "pybind_module.cpp")):
        build_pybind()


if __name__ == "__main__":
    build_all()
    print("built", LIB, file=sys.stderr)
