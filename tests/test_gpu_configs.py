"""The README cylinder workloads (length 30, radius 2, 35 divisions, max
element 0.03: 911,250 tets, 162,946 pressure and 1,258,741 velocity nodes from
the library's own mesher) and mixed precision against the FP64 oracle.

* FP64 steps on the README cylinder at the reference's tolerances and stopping
  rule, two correctors, against the oracle (which re-solves every corrector):
  equal Krylov iteration counts, relative L2 <= 1e-8 (in practice 0).
* Mixed precision on the same mesh (FP32 inner solves, inner tolerance 1e-4,
  FP64 refinement to momentum 1e-10 / pressure 1e-8 on the Jacobi-scaled rule)
  against the GPU FP64 path at the same tolerances: velocity <= 1e-7,
  pressure <= 1e-6; the divergence error is printed.
* Mixed precision against the FP64 oracle on the 6^3 Kuhn channel and the
  small cylinder (length 10, radius 1, 16 divisions, max element 0.1) at the
  same tolerances: velocity <= 1e-7, pressure <= 1e-6.
"""
import numpy as np
import pytest

from test_gpu_parity import relerr

pytestmark = pytest.mark.gpu

RHO, MU = 1060.0, 0.004


def _parabolic(X, R, U=0.5):
    g = np.zeros((len(X), 3))
    g[:, 0] = U * np.clip(1.0 - (X[:, 1] ** 2 + X[:, 2] ** 2) / R ** 2, 0.0, None)
    return g


@pytest.fixture(scope="module")
def readme_cylinder(P):
    m = P.Mesh.cylinder(30.0, 2.0, 35, 0.03)
    m.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    X = np.vstack([m.vertices(), m.edgeNodes()])
    return m, _parabolic(X, 2.0)


def _gpu(P, m, prof, dt, ncorr, **opts):
    g = P.NavierStokesSolver()
    g.initialize(m, RHO, MU)
    g.setDt(dt)
    g.setNumberOfCorrections(ncorr)
    g.setInletProfile(prof)
    if opts:
        g.setOptions(**opts)
    return g


def _div_err(g, gref):
    d = g.computeDivergence(g.getVelocity())
    dr = gref.computeDivergence(gref.getVelocity())
    return np.linalg.norm(d - dr) / max(np.linalg.norm(dr), 1e-300)


def test_readme_cylinder_fp64_matches_oracle(P, O, readme_cylinder):
    """Two FP64 PISO steps with the reference's defaults (BiCGSTAB momentum
    1e-6, BiCGSTAB pressure 1e-8, |r| <= tol |b|), dt 0.01, 2 correctors."""
    m, prof = readme_cylinder
    g = _gpu(P, m, prof, 0.01, 2)
    om = O.OracleMesh.fromArrays(m.vertices(), m.tets())
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    np.testing.assert_array_equal(om.p1BC(), m.p1BC())
    o = O.OracleNS(om, RHO, MU)
    o.setDt(0.01)
    o.setNumberOfCorrections(2)
    o.setInletProfile(prof)
    for step in range(2):
        assert g.pisoStep(), P.last_error()
        assert o.pisoStep() == 0
        st, so = g.stepStats(), o.stats()
        assert st["momentum_iters"] == so["momentum_iters"], (step, st, so)
        assert st["pressure_iters"] == so["pressure_iters"][:2], (step, st, so)
        eu = relerr(g.getVelocity(), o.getVelocity())
        ep = relerr(g.getPressure(), o.getPressure())
        print(f"step {step}: iters {st['momentum_iters']} / {st['pressure_iters']}, "
              f"rel L2 u {eu:.2e} p {ep:.2e}", flush=True)
        assert eu <= 1e-8 and ep <= 1e-8, (step, eu, ep)


@pytest.mark.parametrize("precision,msolver", [(2, 1), (1, 1), (2, 0)])
def test_readme_cylinder_mixed_vs_fp64(P, readme_cylinder, precision, msolver):
    """Mixed precision on the README cylinder (precision 2: FP32 inner
    momentum solves; 1: momentum and pressure; momentum inner solver CG (1)
    or BiCGSTAB (0)) against the FP64 path, 3 steps."""
    m, prof = readme_cylinder
    tols = dict(stop_rule=1, momentum_tol=1e-10, pressure_tol=1e-8, pressure_solver=1)
    g64 = _gpu(P, m, prof, 0.01, 2, **tols)
    g32 = _gpu(P, m, prof, 0.01, 2, precision=precision, momentum_solver=msolver,
               inner_tol=1e-4, **tols)
    for step in range(3):
        assert g64.pisoStep(), P.last_error()
        assert g32.pisoStep(), P.last_error()
        st = g32.stepStats()
        assert all(e <= 1e-10 for e in st["momentum_err"]), st
        assert all(e <= 1e-8 for e in st["pressure_err"]), st
        eu = relerr(g32.getVelocity(), g64.getVelocity())
        ep = relerr(g32.getPressure(), g64.getPressure())
        ed = _div_err(g32, g64)
        print(f"precision {precision} solver {msolver} step {step}: velocity {eu:.2e} "
              f"pressure {ep:.2e} divergence {ed:.2e}", flush=True)
        assert eu <= 1e-7 and ep <= 1e-6, (step, eu, ep)


@pytest.mark.parametrize("kind", ["kuhn6", "cyl_config1"])
@pytest.mark.parametrize("precision,msolver", [(2, 1), (1, 0)])
def test_mixed_vs_oracle(P, O, kind, precision, msolver):
    """Mixed precision against the FP64 oracle (Jacobi-scaled rule, momentum
    1e-10, pressure 1e-8 on both sides), 3 steps."""
    if kind == "kuhn6":
        pm = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, 6, 6, 6)
        R, dt, ncorr = None, 0.01, 1
    else:
        pm = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
        R, dt, ncorr = 1.0, 0.01, 2
    pm.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    X = np.vstack([pm.vertices(), pm.edgeNodes()])
    if R is None:
        prof = np.zeros((len(X), 3))
        prof[:, 0] = 0.5 * 16 * X[:, 1] * (1 - X[:, 1]) * X[:, 2] * (1 - X[:, 2])
    else:
        prof = _parabolic(X, R)
    g = _gpu(P, pm, prof, dt, ncorr, stop_rule=1, momentum_tol=1e-10, pressure_tol=1e-8,
             pressure_solver=1, precision=precision, momentum_solver=msolver)
    om = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    o = O.OracleNS(om, RHO, MU)
    o.setDt(dt)
    o.setNumberOfCorrections(ncorr)
    o.setInletProfile(prof)
    o.setSolver(1e-10, 1e-8, 1000, 1)
    o.setStopScaled(True)
    for step in range(3):
        assert g.pisoStep(), P.last_error()
        assert o.pisoStep() == 0
        u, uo = g.getVelocity(), o.getVelocity()
        eu = relerr(u, uo)
        ep = relerr(g.getPressure(), o.getPressure())
        d, do = g.computeDivergence(u), o.computeDivergence(uo)
        ed = np.linalg.norm(d - do) / max(np.linalg.norm(do), 1e-300)
        print(f"{kind} precision {precision} solver {msolver} step {step}: velocity {eu:.2e} "
              f"pressure {ep:.2e} divergence {ed:.2e}", flush=True)
        assert eu <= 1e-7 and ep <= 1e-6, (step, eu, ep)
