"""Build the CPU oracle (test infrastructure) into oracle/liboracle.so.

The reference itself is not built: its NavierStokes path needs Eigen3,
yaml-cpp and TetGen, none of which exist in this image.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build(verbose: bool = True) -> Path:
    tmp = HERE / "liboracle.so.tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-Wall", str(HERE / "oracle.cpp"), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    tmp.replace(LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
