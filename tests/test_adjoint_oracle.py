"""Pins the oracle's discrete adjoint of the PISO step (the CPU reference
gradient for the GPU adjoint) against central finite differences.

J = 1/2 sum over Outlet velocity nodes of |u|^2 after a few steps; the
parameters are four inlet-profile coefficients and the viscosity.  Forward
solves stop on the Jacobi-scaled residual at 1e-14, because the reference's
rule (|r| <= tol |b|) is dominated by the 1e10 inlet penalty rows and leaves the
interior unconverged far above finite-difference resolution."""
import numpy as np
import pytest


def inlet_basis(m):
    X = np.vstack([m.vertices(), m.edgeNodes()])
    y, z = X[:, 1] - 0.5, X[:, 2] - 0.5
    base = np.clip(1 - (y * y + z * z) / 0.5, 0, None)
    th = np.arctan2(z, y)
    B = np.zeros((4, len(X), 3))
    for k, f in enumerate([np.ones_like(th), np.cos(th), np.sin(th), np.cos(2 * th)]):
        B[k, :, 0] = base * f
    return B


def make(O, m, mu, ncorr=1, relax=(1.0, 1.0), p_out=0.0):
    ns = O.OracleNS(m, 1060.0, mu)
    ns.setDt(0.01)
    ns.setSolver(1e-14, 1e-14, 5000, 0)
    ns.setStopScaled(True)
    ns.setNumberOfCorrections(ncorr)
    ns.setRelax(*relax)
    ns.setOutlet(0, p_out)
    return ns


@pytest.mark.parametrize("ncorr,relax,p_out", [(1, (1.0, 1.0), 0.0), (2, (0.7, 0.5), 2.0)])
def test_adjoint_matches_central_differences(O, ncorr, relax, p_out):
    m = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, 3, 3, 3)
    m.setupBoundaryConditions()
    B = inlet_basis(m)
    c = 0.5 + np.random.default_rng(0).normal(0, 0.05, 4)
    mu = 0.004
    steps = 4
    J, dc, dmu = make(O, m, mu, ncorr, relax, p_out).adjointGradient(steps, B, c)
    h = 1e-6
    for k in range(4):
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        fd = (make(O, m, mu, ncorr, relax, p_out).adjointGradient(steps, B, cp)[0]
              - make(O, m, mu, ncorr, relax, p_out).adjointGradient(steps, B, cm)[0]) / (2 * h)
        assert abs(fd - dc[k]) <= 1e-5 * max(abs(dc[k]), 1e-3 * np.abs(dc).max()), (k, fd, dc[k])
    hm = h  # the same absolute step as the coefficients
    fdmu = (make(O, m, mu + hm, ncorr, relax, p_out).adjointGradient(steps, B, c)[0]
            - make(O, m, mu - hm, ncorr, relax, p_out).adjointGradient(steps, B, c)[0]) / (2 * hm)
    assert abs(fdmu - dmu) <= 1e-5 * abs(dmu), (fdmu, dmu)


def test_functional_is_quadratic_in_a_uniform_scaling_for_stokes_limit(O):
    """With zero pressure and a single step from rest, J(s c) for tiny s is
    dominated by the linear response: J(2s)/J(s) -> 4."""
    m = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, 3, 3, 3)
    m.setupBoundaryConditions()
    B = inlet_basis(m)
    c = np.array([1.0, 0.0, 0.0, 0.0])
    j1 = make(O, m, 0.004).adjointGradient(1, B, 1e-6 * c)[0]
    j2 = make(O, m, 0.004).adjointGradient(1, B, 2e-6 * c)[0]
    assert abs(j2 / j1 - 4.0) < 1e-4
