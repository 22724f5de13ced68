"""GPU adjoint of the PISO step: against the oracle's FP64 CPU adjoint, against
central finite differences (h = 1e-6) of the GPU forward, and through the
torch.autograd.Function wrapper.  Solves use the Jacobi-scaled stopping rule
at 1e-13 so that finite differences resolve the gradient (the reference's
rule is dominated by the 1e10 inlet penalty rows)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-13


def inlet_basis(X, ax_lo=(0.5, 0.5), scale=0.5):
    y, z = X[:, 1] - ax_lo[0], X[:, 2] - ax_lo[1]
    base = np.clip(1 - (y * y + z * z) / scale, 0, None)
    th = np.arctan2(z, y)
    B = np.zeros((4, len(X), 3))
    for k, f in enumerate([np.ones_like(th), np.cos(th), np.sin(th), np.cos(2 * th)]):
        B[k, :, 0] = base * f
    return B


def gpu_solver(P, pm, mu, ncorr=1, tol=TOL):
    g = P.NavierStokesSolver()
    g.initialize(pm, 1060.0, mu)
    g.setDt(0.01)
    g.setNumberOfCorrections(ncorr)
    g.setOptions(momentum_tol=tol, pressure_tol=tol, stop_rule=1, max_iter=5000)
    return g


@pytest.fixture(scope="module")
def setup(P, O):
    pm = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, 4, 4, 4)
    pm.setupBoundaryConditions()
    om = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    om.setupBoundaryConditions()
    X = np.vstack([pm.vertices(), pm.edgeNodes()])
    B = inlet_basis(X)
    c = 0.5 + np.random.default_rng(0).normal(0, 0.05, 4)
    return pm, om, B, c


@pytest.mark.parametrize("ncorr", [1, 2])
def test_adjoint_matches_oracle_adjoint(P, O, setup, ncorr):
    pm, om, B, c = setup
    steps = 4
    g = gpu_solver(P, pm, 0.004, ncorr)
    J, dc, dmu = g.adjointGradient(steps, B, c)
    o = O.OracleNS(om, 1060.0, 0.004)
    o.setDt(0.01)
    o.setSolver(TOL, TOL, 5000, 0)
    o.setStopScaled(True)
    o.setNumberOfCorrections(ncorr)
    Jo, dco, dmuo = o.adjointGradient(steps, B, c)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    np.testing.assert_allclose(dc, dco, rtol=1e-7, atol=1e-9 * np.abs(dco).max())
    assert abs(dmu - dmuo) <= 1e-7 * abs(dmuo)


def test_adjoint_matches_central_differences(P, setup):
    """Normwise: |fd_k - dJ/dc_k| <= 1e-5 max_j |dJ/dc_j|.  The FD quotient
    carries an absolute noise of about tol * |J| / h from the forward solves,
    so a component far below the largest one is not resolved relatively."""
    pm, om, B, c = setup
    steps = 4
    g = gpu_solver(P, pm, 0.004, tol=1e-14)
    J, dc, dmu = g.adjointGradient(steps, B, c)
    h = 1e-6
    for k in range(4):
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        fd = (g.adjointGradient(steps, B, cp)[0] - g.adjointGradient(steps, B, cm)[0]) / (2 * h)
        assert abs(fd - dc[k]) <= 1e-5 * np.abs(dc).max(), (k, fd, dc[k])
    mu, hm = 0.004, 1e-6  # h = 1e-6 for every parameter (2.5e-4 relative for mu)
    Jp = gpu_solver(P, pm, mu + hm, tol=1e-14).adjointGradient(steps, B, c)[0]
    Jm = gpu_solver(P, pm, mu - hm, tol=1e-14).adjointGradient(steps, B, c)[0]
    fdmu = (Jp - Jm) / (2 * hm)
    assert abs(fdmu - dmu) <= 1e-5 * abs(dmu), (fdmu, dmu)


def test_torch_autograd_wrapper(P, setup):
    import torch
    from adjoint_autograd import PisoRollout, outlet_energy
    pm, om, B, c = setup
    steps = 3
    g = gpu_solver(P, pm, 0.004)
    coef = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(0.004, dtype=torch.float64, requires_grad=True)
    u = PisoRollout.apply(coef, mu, g, B, steps)
    lab = np.r_[pm.p1BC(), pm.p2BC()]
    J = outlet_energy(u, torch.from_numpy(lab == 2))
    J.backward()
    J2, dc, dmu = gpu_solver(P, pm, 0.004).adjointGradient(steps, B, c)
    assert abs(J.item() - J2) <= 1e-12 * abs(J2)
    np.testing.assert_allclose(coef.grad.numpy(), dc, rtol=1e-10, atol=1e-14)
    assert abs(mu.grad.item() - dmu) <= 1e-10 * abs(dmu)
