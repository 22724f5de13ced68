"""Build recipe for libcfsa_ns.so (HIP, gfx950) and the C++ driver.

Objects go to <package>/_build/, the shared library to <package>/lib/.
Device code is compiled with FMA contraction off (the FP64 path reproduces
the reference's rounding) and packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
INCLUDE = PKG.parent / "include"
LIB = LIBDIR / "libcfsa_ns.so"
DRIVER = LIBDIR / "cfsa_simulation"

HIP_SOURCES = ["kernels.hip", "solver.hip", "adjoint.hip", "forces.hip", "noalloc.hip",
               "amg_solver.hip", "elem_solver.hip", "capi.cpp"]
HOST_SOURCES = ["mesh.cpp", "pattern.cpp", "vtk.cpp", "yaml_config.cpp", "amg.cpp", "elem_patch.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", f"-I{CSRC}", f"-I{INCLUDE}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
              f"-I{CSRC}", f"-I{INCLUDE}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: str) -> Path:
    path = CSRC / src
    obj = BUILD / (src.replace(".", "_") + ".o")
    if src.endswith(".hip") or src in HIP_SOURCES:
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(path), "-o", str(obj)]
    else:
        cmd = [CXX, *HOST_FLAGS, "-c", str(path), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = True) -> Path:
    """Compile every source (always, no staleness check) and link the library."""
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    sources = HIP_SOURCES + [s for s in HOST_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp),
          *map(str, objs)])
    os.replace(tmp, LIB)
    driver_src = CSRC / "simulation_main.cpp"
    if driver_src.exists():
        _run([CXX, *HOST_FLAGS, str(driver_src), "-o", str(DRIVER), f"-L{LIBDIR}",
              "-lcfsa_ns", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib",
              "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64"])
    if verbose:
        print(f"built {LIB}")
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build()
    sys.exit(0)
