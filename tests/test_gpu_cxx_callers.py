"""C++ callers written like the reference's own callers, compiled against the
drop-in header (csrc/NavierStokesSolver.hpp over include/cfsa_ns.h) alone and
linked to libcfsa_ns.so, then checked against the CPU oracle run through the
same scenario:

* tests/cxx/ns_example.cpp: the direct-API usage of examples/NavierStokes.cpp
  (lines 43-124);
* tests/cxx/coupled_fluid_step.cpp: the fluid side of CoupledSolver
  (CoupledSolver.cpp:28-48, 86-97, 113-129, 141-175, 186-189), including
  boundary relabelling between steps and the force hand-off;
* tests/cxx/simple_solver.cpp: NavierStokesSimpleSolver
  (NavierStokes.hpp:201-251) against the oracle's simpleStep;
* tests/cxx/step_parts.cpp: the public parts of the step
  (NavierStokes.hpp:101-112) compose pisoStep bit for bit.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from test_gpu_parity import relerr

pytestmark = pytest.mark.gpu
REPO = Path(__file__).resolve().parents[1]
PKG = REPO / "cell-fluid-structure-analysis_amd"


def _build_and_run(name, tmp_path):
    exe = tmp_path / name
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-ffp-contract=off",
           f"-I{REPO / 'include'}", f"-I{PKG / 'csrc'}", str(REPO / "tests" / "cxx" / f"{name}.cpp"),
           "-o", str(exe), f"-L{PKG / 'lib'}", "-lcfsa_ns", f"-Wl,-rpath,{PKG / 'lib'}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    prefix = tmp_path / name
    r = subprocess.run([str(exe), str(prefix)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    return lambda s: np.fromfile(f"{prefix}_{s}.bin", dtype=np.float64)


def test_cxx_example_caller_matches_oracle(O, tmp_path):
    out = _build_and_run("ns_example", tmp_path)
    om = O.OracleMesh.structuredRectangularPrism(2.0, 1.0, 1.0, 12, 6, 6)
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    # the example sets the inlet before initialize: inlet nodes start at it
    o = O.OracleNS(om, 1000.0, 0.003, (1.0, 0.0, 0.0))
    o.setDt(1e-3)
    o.setMeanInletVelocity((1.0, 0.0, 0.0))
    o.setOutlet(0, 0.0)
    o.setRelax(0.5, 0.5)
    for _ in range(12):
        assert o.pisoStep() == 0
    u, p = out("u").reshape(-1, 3), out("p")
    eu, ep = relerr(u, o.getVelocity()), relerr(p, o.getPressure())
    print(f"example caller: rel L2 u {eu:.2e} p {ep:.2e}")
    assert eu <= 1e-8 and ep <= 1e-8
    vtk = sorted(tmp_path.glob("ns_example_t*.vtk"))
    assert [f.name for f in vtk] == ["ns_example_t0.vtk", "ns_example_t10.vtk",
                                     "ns_example_t5.vtk"]


def test_cxx_coupled_fluid_step_matches_oracle(P, O, tmp_path):
    out = _build_and_run("coupled_fluid_step", tmp_path)
    pm = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
    om = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    orig = om.p1BC().copy()
    X = om.vertices()
    o = O.OracleNS(om, 1060.0, 0.004)
    ft = o.forces(2)
    o.setOutlet(0, 0.0)
    o.setMeanInletVelocity((0.5, 0.0, 0.0))
    o.setDt(0.01)
    for step in range(6):
        cx, r = 3.0 + 0.5 * step, 0.35
        lab = orig.copy()
        d2 = (X[:, 0] - cx) * (X[:, 0] - cx) + X[:, 1] * X[:, 1] + X[:, 2] * X[:, 2]
        lab[(orig == 3) & (d2 < r * r)] = 0
        om.setFluidVertexBCs(lab)
        fp, fs = o.forces(0), o.forces(1)
        assert o.pisoStep() == 0
        ft = o.forces(2)
    assert (lab != orig).any()
    u, p = out("u").reshape(-1, 3), out("p")
    eu, ep = relerr(u, o.getVelocity()), relerr(p, o.getPressure())
    print(f"coupled caller: rel L2 u {eu:.2e} p {ep:.2e}")
    assert eu <= 1e-8 and ep <= 1e-8
    for name, want in (("ft", ft), ("fp", fp), ("fs", fs)):
        got = out(name).reshape(-1, 3)
        assert relerr(got, want) <= 1e-8, name
    assert relerr(out("div"), o.computeDivergence(u)) <= 1e-12


def test_cxx_simple_solver_matches_oracle(O, tmp_path):
    out = _build_and_run("simple_solver", tmp_path)
    om = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, 4, 4, 4)
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    o = O.OracleNS(om, 1060.0, 0.004)
    o.setDt(0.01)
    o.setMeanInletVelocity((0.5, 0.0, 0.0))
    o.setOutlet(0, 2.0)
    for _ in range(3):
        assert o.simpleStep() == 0
    u, p = out("u").reshape(-1, 3), out("p")
    assert relerr(u, o.getVelocity()) <= 1e-8 and relerr(p, o.getPressure()) <= 1e-8
    assert abs(out("t")[0] - 0.03) <= 1e-15


def test_cxx_step_parts(tmp_path):
    _build_and_run("step_parts", tmp_path)
