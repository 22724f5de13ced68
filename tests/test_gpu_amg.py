"""AMG-preconditioned CG for the pressure system (pressure_solver 2) on the
GPU:

* one device V-cycle (cfsa_ns_amg_apply) against the NumPy restatement in
  tests/amg_ref.py on the hierarchy of the same operator, built by the
  library's host set-up (which tests/test_amg_setup.py checks on the CPU);
* AMG-PCG solves of L_bc: the Jacobi-scaled residual of the returned x,
  agreement with the Jacobi-PCG solution, a fraction of its iterations, and
  the iteration count of the restated CG (within one);
* PISO steps with the AMG pressure solve against the FP64 oracle (Jacobi-CG
  pressure) at tight tolerances, FP64 and mixed momentum;
* graph replay and eager launches give the same bits; rebuild policy;
* the adjoint with AMG pressure solves equals the Jacobi-CG adjoint.
"""
import numpy as np
import pytest

from amg_ref import VCycle, csr, pcg
from test_gpu_parity import relerr

pytestmark = pytest.mark.gpu

RHO, MU = 1060.0, 0.004


def _mesh(P, kind):
    if kind.startswith("cyl"):
        pm = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
    else:
        n = int(kind[4:6])
        pm = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
    pm.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    return pm


def _solver(P, pm, kind, **opts):
    g = P.NavierStokesSolver()
    g.initialize(pm, RHO, MU)
    g.setDt(0.01)
    g.setMeanInletVelocity((0.5, 0.0, 0.0))
    if kind.endswith("n"):
        g.setOutletType(P.OUTLET_NEUMANN)
        g.setReferencePressureNode(5)
    if opts:
        g.setOptions(**opts)
    return g


CASES = ["kuhn08", "kuhn12n", "cyl"]


@pytest.mark.parametrize("kind,layout", [(k, 0) for k in CASES] + [("cyl", 2)])
def test_vcycle_matches_restatement(P, kind, layout):
    """layout 2 puts every coarse operator, P and R in sliced-ELL (the layout
    rows longer than a CSR-stream window use)."""
    pm = _mesh(P, kind)
    g = _solver(P, pm, kind, pressure_solver=2, spmv_layout=layout)
    rp, col, val = g.getMatrix("Lbc")
    n = len(rp) - 1
    H = P.HostAmg(rp, col, val)
    V = VCycle(H)
    rng = np.random.default_rng(11)
    for trial in range(2):
        r = rng.standard_normal(n)
        z = g.amgApply(r)
        zw = V.apply(r)
        e = relerr(z, zw)
        print(f"{kind}: levels {H.levels()[0]}, V-cycle rel err {e:.2e}", flush=True)
        assert e <= 1e-12, e
        # the FP32 V-cycle of the mixed mode (precision 1): FP32 rounding
        e32 = relerr(g.amgApply(r, precision=1), zw)
        print(f"{kind}: FP32 V-cycle rel err {e32:.2e}", flush=True)
        assert e32 <= 1e-5, e32
    info = g.amgInfo()
    L = H.levels()[0]
    assert info["levels"] == L and info["builds"] == 1
    assert info["rows"] == [H.matrix(l, "A")[3][0] for l in range(L)]


@pytest.mark.parametrize("kind,precision", [(k, 0) for k in CASES] + [("cyl", 1)])
def test_amg_pcg_solves(P, kind, precision):
    """precision 1: FP64 CG around the FP32 V-cycle."""
    pm = _mesh(P, kind)
    g = _solver(P, pm, kind, stop_rule=1)
    rp, col, val = g.getMatrix("Lbc")
    n = len(rp) - 1
    A = csr(rp, col, val, (n, n))
    d = A.diagonal()
    b = A @ np.random.default_rng(5).standard_normal(n)
    tol = 1e-10
    xa, ra = g.krylovSolve("Lbc", b, method=2, tol=tol, maxit=1000, precision=precision)
    xj, rj = g.krylovSolve("Lbc", b, method=1, tol=tol, maxit=5000)
    assert ra["ok"] and rj["ok"]
    res = np.linalg.norm((b - A @ xa) / d) / np.linalg.norm(b / d)
    _, itw = pcg(A, b, VCycle(P.HostAmg(rp, col, val)).apply, tol)
    print(f"{kind}: AMG-PCG {ra['iters'][0]} iterations (restated {itw}), Jacobi-PCG "
          f"{rj['iters'][0]}; scaled residual {res:.2e}, vs Jacobi {relerr(xa, xj):.2e}",
          flush=True)
    assert res <= 1.01 * tol
    assert abs(ra["iters"][0] - itw) <= (1 if precision == 0 else 3)
    assert 4 * ra["iters"][0] <= rj["iters"][0]
    assert relerr(xa, xj) <= 1e-6


@pytest.mark.parametrize("kind,precision,simple",
                         [("kuhn06", 0, False), ("cyl", 0, False), ("cyl", 2, False),
                          ("cyl", 1, False), ("kuhn10n", 0, False), ("kuhn08", 0, True)])
def test_piso_steps_vs_oracle(P, O, kind, precision, simple):
    """pressure_solver 2 against the FP64 oracle (Jacobi-CG pressure), both at
    momentum 1e-10 / pressure 1e-10 on the Jacobi-scaled rule, 3 steps (kuhn10n:
    Neumann outlet with the reference node pinned; simple: simpleStep)."""
    pm = _mesh(P, kind)
    ncorr = 2 if kind == "cyl" else 1
    g = _solver(P, pm, kind, stop_rule=1, momentum_tol=1e-10, pressure_tol=1e-10,
                pressure_solver=2, precision=precision, momentum_solver=1)
    g.setNumberOfCorrections(ncorr)
    om = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    o = O.OracleNS(om, RHO, MU)
    o.setDt(0.01)
    o.setMeanInletVelocity((0.5, 0.0, 0.0))
    o.setNumberOfCorrections(ncorr)
    if kind.endswith("n"):
        o.setOutlet(1, 0.0)
        o.setReferencePressureNode(5)
    o.setSolver(1e-10, 1e-10, 5000, 1)
    o.setMomentumSolver(1)
    o.setStopScaled(True)
    for step in range(3):
        if simple:
            assert g.simpleStep(), P.last_error()
            assert o.simpleStep() == 0
        else:
            assert g.pisoStep(), P.last_error()
            assert o.pisoStep() == 0
        st, so = g.stepStats(), o.stats()
        eu = relerr(g.getVelocity(), o.getVelocity())
        ep = relerr(g.getPressure(), o.getPressure())
        print(f"{kind} precision {precision} step {step}: pressure iters AMG "
              f"{st['pressure_iters'][:1]} vs Jacobi {so['pressure_iters'][:1]}; "
              f"velocity {eu:.2e} pressure {ep:.2e}", flush=True)
        assert eu <= 1e-7 and ep <= 1e-6, (step, eu, ep)


@pytest.mark.parametrize("precision", [0, 1])
def test_graphs_and_repeats_bit_identical(P, precision):
    pm = _mesh(P, "cyl")
    outs = []
    for graphs in (1, 0, 1):
        g = _solver(P, pm, "cyl", pressure_solver=2, stop_rule=1, graphs=graphs,
                    precision=precision)
        for _ in range(2):
            assert g.pisoStep(), P.last_error()
        outs.append((g.getVelocity(), g.getPressure(), g.stepStats()["pressure_iters"]))
    for u, p, it in outs[1:]:
        assert np.array_equal(u, outs[0][0]) and np.array_equal(p, outs[0][1])
        assert it == outs[0][2]


def test_rebuild_policy(P):
    pm = _mesh(P, "kuhn06")
    X = pm.vertices()
    g = _solver(P, pm, "kuhn06", pressure_solver=2, stop_rule=1)
    assert g.pisoStep(), P.last_error()
    assert g.amgInfo()["builds"] == 1
    g.build_sparse_matrices()  # same geometry: same values, hierarchy kept
    assert g.pisoStep()
    assert g.amgInfo()["builds"] == 1
    for k, reuse in enumerate((0, 1)):
        g.setOptions(amg_reuse=reuse)
        Xd = X.copy()
        Xd[:, 1] += (0.01 + 0.01 * k) * np.sin(np.pi * X[:, 0]) * X[:, 1] * (1 - X[:, 1])
        pm.updateVertices(Xd)
        g.build_sparse_matrices()
        assert g.pisoStep(), P.last_error()
        assert g.stepStats()["pressure_err"][0] <= 1e-8
        # reuse 0: rebuilt on the change; reuse 1: the first change keeps it
        assert g.amgInfo()["builds"] == 2, (reuse, g.amgInfo())


def test_adjoint_with_amg_pressure(P):
    """The adjoint's forward and transposed pressure solves through the AMG
    preconditioner (L_bc is symmetric, so the same hierarchy serves both) give
    the Jacobi-CG adjoint's J and gradient, at the tight tolerance gradients need."""
    from test_gpu_adjoint import inlet_basis
    pm = _mesh(P, "cyl")
    X = np.vstack([pm.vertices(), pm.edgeNodes()])
    B = inlet_basis(X, ax_lo=(0.0, 0.0), scale=1.0)
    c = 0.5 + np.random.default_rng(0).normal(0, 0.05, 4)
    out = []
    for ps in (1, 2):
        g = P.NavierStokesSolver()
        g.initialize(pm, RHO, MU)
        g.setDt(0.01)
        g.setNumberOfCorrections(2)
        g.setOptions(momentum_tol=1e-13, pressure_tol=1e-13, stop_rule=1, max_iter=20000,
                     pressure_solver=ps)
        out.append(g.adjointGradient(3, B, c))
    (J1, dc1, dmu1), (J2, dc2, dmu2) = out
    print(f"adjoint: J {J1:.12e} / {J2:.12e}, dc rel {relerr(dc2, dc1):.2e}, "
          f"dmu rel {abs(dmu2 - dmu1) / abs(dmu1):.2e}", flush=True)
    assert abs(J2 - J1) <= 1e-10 * abs(J1)
    assert relerr(dc2, dc1) <= 1e-7
    assert abs(dmu2 - dmu1) <= 1e-6 * abs(dmu1)
