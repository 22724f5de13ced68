"""The public parts of the step (NavierStokes.hpp:101-112) through the C ABI:

* solveMomentumPredictor, then solvePressurePoisson / correctVelocity once per
  corrector, gives the fields of pisoStep bit for bit (correctVelocity already
  sets the Dirichlet values that pisoStep re-imposes, NavierStokes.cpp:727-733,
  815-835) and leaves the time where it was;
* normalizePressire subtracts the reference node's pressure (and is a no-op
  without a reference node, NavierStokes.cpp:804-813);
* solvePressureBiCGSTAB solves the unpenalised Poisson matrix with Jacobi
  BiCGSTAB at 1e-8 / 1000 iterations (NavierStokes.cpp:773-802): the oracle's
  solve of the same matrix takes the same iterations and gives the same x.
"""
import numpy as np
import pytest

from test_gpu_parity import relerr

pytestmark = pytest.mark.gpu

RHO, MU = 1060.0, 0.004


def _pair(P, kind, **opts):
    if kind == "cyl":
        pm = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
    else:
        pm = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, 5, 5, 5)
    pm.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    out = []
    for _ in range(2):
        g = P.NavierStokesSolver()
        g.initialize(pm, RHO, MU)
        g.setDt(0.01)
        g.setMeanInletVelocity((0.5, 0.0, 0.0))
        g.setNumberOfCorrections(2 if kind == "cyl" else 1)
        if opts:
            g.setOptions(**opts)
        out.append(g)
    return pm, out


@pytest.mark.parametrize("kind,opts", [("kuhn", {}), ("cyl", {}),
                                       ("cyl", {"pressure_solver": 2, "stop_rule": 1}),
                                       ("kuhn", {"precision": 1})])
def test_parts_compose_piso_step(P, kind, opts):
    pm, (a, b) = _pair(P, kind, **opts)
    ncorr = 2 if kind == "cyl" else 1
    for step in range(2):
        t0 = b.getTime()
        assert a.pisoStep(), P.last_error()
        assert b.solveMomentumPredictor(), P.last_error()
        for _ in range(ncorr):
            assert b.solvePressurePoisson(), P.last_error()
            assert b.correctVelocity()
        assert b.getTime() == t0
        print(f"{kind} {opts} step {step}: velocity "
              f"{relerr(b.getVelocity(), a.getVelocity()):.1e}", flush=True)
        assert np.array_equal(b.getVelocity(), a.getVelocity()), step
        assert np.array_equal(b.getPressure(), a.getPressure()), step
        assert np.array_equal(b.getVelocityStar(), a.getVelocityStar()), step
        assert np.array_equal(b.getPressureCorrection(), a.getPressureCorrection()), step


def test_normalize_pressure(P):
    pm, (a, b) = _pair(P, "kuhn")
    assert a.pisoStep(), P.last_error()
    p = a.getPressure()
    assert a.normalizePressire()  # no reference node: unchanged
    assert np.array_equal(a.getPressure(), p)
    a.setReferencePressureNode(7)
    assert a.normalizePressire()
    q = a.getPressure()
    assert q[7] == 0.0
    assert np.array_equal(q, p - p[7])


@pytest.mark.parametrize("kind", ["kuhn", "cyl"])
def test_bicgstab_on_poisson_matrix(P, O, kind):
    pm, (g, _) = _pair(P, kind)
    rp, col, val = g.getMatrix("L")
    n = len(rp) - 1
    rng = np.random.default_rng(3)
    # L is the Neumann Laplacian (singular): a right-hand side in its range
    b = O.csr_spmv(rp, col, val, rng.uniform(-1, 1, n))
    ok, x, it = g.solvePressureBiCGSTAB(b)
    xo, io = O.csr_krylov(rp, col, val, b, method=0, tol=1e-8, maxit=1000)
    print(f"{kind}: BiCGSTAB on L: {it} iterations (oracle {io['iters']}), "
          f"x rel {relerr(x, xo):.1e}", flush=True)
    assert ok and it == io["iters"]
    assert relerr(x, xo) <= 1e-12
    r = b - O.csr_spmv(rp, col, val, x)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b)
    # the fields are untouched by the solve
    u, p = g.getVelocity(), g.getPressure()
    g.solvePressureBiCGSTAB(b)
    assert np.array_equal(g.getVelocity(), u) and np.array_equal(g.getPressure(), p)
    with pytest.raises(ValueError):
        g.solvePressureBiCGSTAB(b[:-1])
