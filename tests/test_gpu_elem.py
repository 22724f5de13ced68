"""The FP32 momentum operator applied tet by tet from element patches
(elem_spmv = 1) on the GPU:

* its product equals the assembled FP64 operator's (NavierStokes.cpp:460-495)
  to FP32 rounding, row by row against |A| |x|, as the FP32 CSR copy does; it
  is deterministic; it follows a moved mesh (re-assembly) and relabelled
  boundary nodes;
* mixed-precision PISO steps with it match the FP64 oracle (velocity 1e-7,
  pressure 1e-6 at momentum 1e-10 / pressure 1e-8, Jacobi-scaled rule) with CG
  and BiCGSTAB inner solves;
* graph replay and eager launches give the same bits.
"""
import numpy as np
import pytest

from test_gpu_parity import relerr

pytestmark = pytest.mark.gpu

RHO, MU = 1060.0, 0.004


def _mesh(P, kind):
    if kind == "cyl":
        pm = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
    else:
        n = int(kind[4:])
        pm = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
    pm.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    return pm


def _solver(P, pm, **opts):
    g = P.NavierStokesSolver()
    g.initialize(pm, RHO, MU)
    g.setDt(0.01)
    g.setMeanInletVelocity((0.5, 0.0, 0.0))
    if opts:
        g.setOptions(**opts)
    return g


def _check_product(g, x, label):
    rp, col, val = g.getMatrix("lhs")
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    absx = np.abs(x)
    bound = np.zeros((n, 3))
    np.add.at(bound, rows, np.abs(val)[:, None] * absx[col])
    y64 = g.applyMatrix("lhs", x, 3)
    g.setOptions(elem_spmv=0)
    yc = g.applyMatrix("lhs32", x, 3)
    g.setOptions(elem_spmv=1)
    ye = g.applyMatrix("lhs32", x, 3)
    ye2 = g.applyMatrix("lhs32", x, 3)
    assert np.array_equal(ye, ye2)
    ee = (np.abs(ye - y64) / bound).max()
    ec = (np.abs(yc - y64) / bound).max()
    print(f"{label}: |y - y64| / (|A||x|) element {ee:.2e}, CSR FP32 {ec:.2e}", flush=True)
    assert ee <= 2e-6
    return ye


@pytest.mark.parametrize("kind", ["kuhn8", "cyl"])
def test_elem_product_matches_assembled(P, kind):
    pm = _mesh(P, kind)
    g = _solver(P, pm, precision=2)
    nu = g.sizes()["nu"]
    x = np.random.default_rng(7).standard_normal((nu, 3))
    _check_product(g, x, kind)
    # a moved mesh: re-assembly refreshes the per-tet geometry
    X = pm.vertices()
    Xd = X.copy()
    Xd[:, 1] += 0.02 * np.sin(np.pi * X[:, 0]) * X[:, 1] * (1 - X[:, 1])
    pm.updateVertices(Xd)
    g.build_sparse_matrices()
    _check_product(g, x, kind + " moved")
    # relabelled boundary nodes: the penalty rows follow
    bc = pm.p1BC().copy()
    flip = np.nonzero(bc == 3)[0][:5]
    bc[flip] = 0
    pm.setFluidVertexBCs(bc)
    g.build_sparse_matrices()
    ye = _check_product(g, x, kind + " relabelled")
    assert np.all(np.abs(ye[flip]) >= 1e9 * np.abs(x[flip]) - 1e3)


@pytest.mark.parametrize("kind,precision,msolver", [("kuhn6", 2, 1), ("kuhn6", 1, 0),
                                                    ("cyl", 2, 1), ("cyl", 1, 1)])
def test_elem_mixed_steps_vs_oracle(P, O, kind, precision, msolver):
    pm = _mesh(P, kind)
    ncorr = 2 if kind == "cyl" else 1
    g = _solver(P, pm, stop_rule=1, momentum_tol=1e-10, pressure_tol=1e-8, pressure_solver=1,
                precision=precision, momentum_solver=msolver, elem_spmv=1)
    g.setNumberOfCorrections(ncorr)
    om = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    o = O.OracleNS(om, RHO, MU)
    o.setDt(0.01)
    o.setMeanInletVelocity((0.5, 0.0, 0.0))
    o.setNumberOfCorrections(ncorr)
    o.setSolver(1e-10, 1e-8, 1000, 1)
    o.setStopScaled(True)
    for step in range(3):
        assert g.pisoStep(), P.last_error()
        assert o.pisoStep() == 0
        st = g.stepStats()
        eu = relerr(g.getVelocity(), o.getVelocity())
        ep = relerr(g.getPressure(), o.getPressure())
        print(f"{kind} precision {precision} solver {msolver} step {step}: inner iterations "
              f"{st['momentum_iters']} in {st['momentum_refinements']} sweeps; velocity "
              f"{eu:.2e} pressure {ep:.2e}", flush=True)
        assert eu <= 1e-7 and ep <= 1e-6, (step, eu, ep)


def test_elem_graphs_bit_identical(P):
    pm = _mesh(P, "cyl")
    outs = []
    for graphs in (1, 0, 1):
        g = _solver(P, pm, stop_rule=1, momentum_tol=1e-10, pressure_tol=1e-8, precision=1,
                    momentum_solver=1, elem_spmv=1, graphs=graphs)
        for _ in range(2):
            assert g.pisoStep(), P.last_error()
        outs.append((g.getVelocity(), g.getPressure(), g.stepStats()["momentum_iters"]))
    for u, p, it in outs[1:]:
        assert np.array_equal(u, outs[0][0]) and np.array_equal(p, outs[0][1])
        assert it == outs[0][2]
