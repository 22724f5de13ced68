"""GPU tests of the paths around the PISO step: mixed precision against FP64,
value-only re-assembly after mesh deformation (coupled-step contract), and
the cfsa_simulation driver built on the C++ host API."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from test_gpu_parity import make_case, relerr

pytestmark = pytest.mark.gpu
REPO = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("kind,n,steps", [("kuhn", 6, 3), ("cyl", 0, 3)])
@pytest.mark.parametrize("precision,msolver", [(1, 0), (2, 0), (2, 1)])
def test_mixed_precision_error_vs_fp64(P, O, kind, n, steps, precision, msolver):
    """FP32 inner solves (inner tol 1e-4) inside FP64 refinement, momentum to
    1e-10 and pressure to 1e-8 (Jacobi-scaled rule, as the FP64 path): the
    velocity stays within 1e-7 and the pressure within 1e-5 of the FP64 path.
    precision 1 refines both solves, 2 only the momentum solve."""
    tols = dict(momentum_tol=1e-10, pressure_tol=1e-8)
    g64, o, pm, om = make_case(P, O, kind, n, **tols)
    g32, _, _, _ = make_case(P, O, kind, n, **tols)
    g64.setOptions(stop_rule=1)
    g32.setOptions(stop_rule=1, precision=precision, momentum_solver=msolver)
    for s in range(steps):
        assert g64.pisoStep(), P.last_error()
        assert g32.pisoStep(), P.last_error()
        st = g32.stepStats()
        assert all(e <= 1e-10 for e in st["momentum_err"]), st
        assert all(e <= 1e-8 for e in st["pressure_err"]), st
        eu = relerr(g32.getVelocity(), g64.getVelocity())
        ep = relerr(g32.getPressure(), g64.getPressure())
        d64 = g64.computeDivergence(g64.getVelocity())
        d32 = g32.computeDivergence(g32.getVelocity())
        ed = np.linalg.norm(d32 - d64) / max(np.linalg.norm(d64), 1e-300)
        print(f"step {s}: velocity {eu:.3e} pressure {ep:.3e} divergence {ed:.3e}")
        assert eu <= 1e-7 and ep <= 1e-5, (s, eu, ep, st)


def test_mixed_krylov_true_residual(P, O):
    g, o, pm, om = make_case(P, O, "kuhn", 5)
    sz = g.sizes()
    b = np.random.default_rng(5).uniform(-1, 1, (sz["nu"], 3))
    x, info = g.krylovSolve("lhs", b, tol=1e-10, precision=1)
    rp, col, val = g.getMatrix("lhs")
    r = b - O.csr_spmv(rp, col, val, x)
    for c in range(3):
        assert np.linalg.norm(r[:, c]) <= 1.01e-10 * np.linalg.norm(b[:, c])
    bp = np.random.default_rng(6).uniform(-1, 1, sz["nv"])
    xp, ip = g.krylovSolve("Lbc", bp, tol=1e-10, precision=1, method=1)
    rp, col, val = g.getMatrix("Lbc")
    assert np.linalg.norm(bp - O.csr_spmv(rp, col, val, xp)) <= 1.01e-10 * np.linalg.norm(bp)


def test_deformation_reassembly_matches_oracle(P, O):
    """Coupled-step contract (CoupledSolver.cpp:113-129): the mesh moves,
    build_sparse_matrices re-assembles into the fixed pattern on the GPU, and
    the next step sees the new geometry; labels stay as set up initially."""
    g, o, pm, om = make_case(P, O, "kuhn", 4)
    assert g.pisoStep() and o.pisoStep() == 0
    X = pm.vertices()
    Xd = X.copy()
    Xd[:, 1] += 0.03 * np.sin(np.pi * X[:, 0]) * X[:, 1] * (1 - X[:, 1])
    Xd[:, 2] += 0.02 * np.cos(np.pi * X[:, 1]) * X[:, 2] * (1 - X[:, 2])
    pm.updateVertices(Xd)
    g.build_sparse_matrices()
    om2 = O.OracleMesh.fromArrays(Xd, pm.tets())
    om2.setFluidVertexBCs(pm.p1BC())
    o2 = O.OracleNS(om2, 1060.0, 0.004)
    o2.setDt(0.01)
    o2.setMeanInletVelocity((0.5, 0.0, 0.0))
    o2.setVelocity(o.getVelocity())
    o2.setPressure(o.getPressure())
    # the GPU re-measures the moved tets with the mesh layer's arithmetic and
    # re-assembles in the reference's summation order: bit-exact
    for name in ["M", "K", "lhs", "L", "Lbc", "Gx", "Gy", "Gz"]:
        vg = g.getMatrix(name)[2]
        vo = o2.getMatrix(name)[2]
        assert np.array_equal(vg, vo), name
    for _ in range(2):
        assert g.pisoStep() and o2.pisoStep() == 0
        assert relerr(g.getVelocity(), o2.getVelocity()) <= 1e-8
        assert relerr(g.getPressure(), o2.getPressure()) <= 1e-8


def test_label_change_takes_effect_next_step(P, O):
    g, o, pm, om = make_case(P, O, "kuhn", 4)
    lab = pm.p1BC().copy()
    interior = np.nonzero(lab == 3)[0][:5]
    lab[interior] = 0  # cells adhering turn interior vertices into walls
    pm.setFluidVertexBCs(lab)
    om.setFluidVertexBCs(lab)
    assert g.pisoStep() and o.pisoStep() == 0
    assert relerr(g.getVelocity(), o.getVelocity()) <= 1e-8
    assert not g.getVelocity()[interior].any()


def test_simulation_driver(P, tmp_path):
    cfg = tmp_path / "box.yaml"
    cfg.write_text(
        "mesh:\n  type: box\n  size: [2.0, 1.0, 1.0]\n  cells: [6, 3, 3]\n"
        "  flow_direction: [1.0, 0.0, 0.0]\n"
        "navier_stokes:\n  enabled: true\n  density: 1060.0\n  viscosity: 0.004\n"
        "  inlet_velocity: [0.5, 0.0, 0.0]\n  dt: 0.01\n"
        f"output:\n  prefix: {tmp_path / 'run'}\n  total_steps: 4\n  output_frequency: 2\n")
    exe = REPO / "cell-fluid-structure-analysis_amd" / "lib" / "cfsa_simulation"
    r = subprocess.run([str(exe), str(cfg)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Simulation completed successfully" in r.stdout
    files = sorted(tmp_path.glob("run_fluid_*.vtk"))
    assert [f.name for f in files] == ["run_fluid_0.vtk", "run_fluid_1.vtk"]
    # same run through the C ABI from Python
    m = P.Mesh.structuredRectangularPrism(2.0, 1.0, 1.0, 6, 3, 3)
    m.setupBoundaryConditions()
    g = P.NavierStokesSolver()
    g.initialize(m, 1060.0, 0.004)
    g.setOutletType(0)
    g.setMeanInletVelocity((0.5, 0.0, 0.0))
    g.setDt(0.01)
    for _ in range(2):
        assert g.pisoStep()
    txt = files[1].read_text().splitlines()
    j = txt.index("VECTORS velocity double")
    nv = m.nVertices()
    u = np.array([[float(v) for v in r.split()] for r in txt[j + 1:j + 1 + nv]])
    np.testing.assert_allclose(u, g.getVelocity()[:nv], rtol=1e-5, atol=1e-10)


def test_node_accessors(P):
    """getVelocityAtNode / getPressureAtNode (NavierStokes.hpp:152-159) return the
    node's entries of the full fields and reject nodes out of range."""
    m = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, 3, 3, 3)
    m.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    g = P.NavierStokesSolver()
    g.initialize(m, 1060.0, 0.004)
    g.setMeanInletVelocity((0.5, 0.0, 0.0))
    g.setDt(0.01)
    for _ in range(2):
        assert g.pisoStep(), P.last_error()
    u, p = g.getVelocity(), g.getPressure()
    sz = g.sizes()
    for i in (0, 1, sz["nv"] - 1, sz["nv"], sz["nu"] - 1):
        assert np.array_equal(g.getVelocityAtNode(i), u[i])
    for i in (0, 5, sz["nv"] - 1):
        assert g.getPressureAtNode(i) == p[i]
    with pytest.raises(Exception):
        g.getVelocityAtNode(sz["nu"])
    with pytest.raises(Exception):
        g.getPressureAtNode(sz["nv"])


def test_face_accessors_after_vertex_move(P, O):
    """After the mesh moves, the face accessors use the face list and vertex
    order fixed when the mesh was built and recompute normal, area and centre
    from it (Mesh::updateGeometry, Mesh.hpp:98-101): the same faces as the
    oracle mesh moved the same way, for internal and boundary faces."""
    g, o, pm, om = make_case(P, O, "kuhn", 3)
    assert g.pisoStep() and o.pisoStep() == 0
    X = pm.vertices()
    Xd = X.copy()
    Xd[:, 0] += 0.04 * np.sin(2 * np.pi * X[:, 1]) * X[:, 0] * (1 - X[:, 0])
    Xd[:, 1] += 0.03 * np.cos(np.pi * X[:, 2]) * X[:, 1] * (1 - X[:, 1])
    pm.updateVertices(Xd)
    om.updateVertices(Xd)
    g.build_sparse_matrices()
    p = np.random.default_rng(8).uniform(-1, 1, g.sizes()["nv"])
    g.setPressure(p)
    v, t, geo = om.faces()
    vh, th, geoh = pm.faces()
    assert np.array_equal(v, vh) and np.array_equal(geo, geoh)
    assert g.nFaces() == len(v)
    for fi in range(0, len(v), 7):
        mp = ((p[v[fi, 0]] + p[v[fi, 1]]) + p[v[fi, 2]]) / 3.0
        assert g.getMeanFacePressure(fi) == mp
        np.testing.assert_array_equal(g.getPressureForceAtFace(fi),
                                      (-geo[fi, :3] * geo[fi, 3]) * mp)
