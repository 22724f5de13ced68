"""Adjoint at the README cylinder's size (length 30, radius 2, 35 divisions,
max element 0.03; 911,250 tets): gradient of the outlet kinetic energy after
10 PISO steps (2 correctors, dt 0.01) with respect to 4 inlet Fourier
coefficients c ~ 0.5 + N(0, 0.05) (numpy default_rng(0)) and the viscosity,
against central differences with h = 1e-6 of the GPU forward.

The difference quotient's signal for the viscosity is ~1e-7 of J, so the
forward solves run at 1e-14 (Jacobi-scaled rule); at 1e-12 the quotient is
noise-limited (~3e-4).  A five-point stencil with 10 h (truncation ~4e-11)
checks the viscosity component more tightly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 2.0
STEPS = 10
TOL = 1e-14


def fourier_basis(X):
    y, z = X[:, 1], X[:, 2]
    base = np.clip(1.0 - (y * y + z * z) / R ** 2, 0.0, None)
    th = np.arctan2(z, y)
    B = np.zeros((4, len(X), 3))
    for k, f in enumerate([np.ones_like(th), np.cos(th), np.sin(th), np.cos(2 * th)]):
        B[k, :, 0] = base * f
    return B


@pytest.fixture(scope="module")
def cyl(P):
    m = P.Mesh.cylinder(30.0, R, 35, 0.03)
    m.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    X = np.vstack([m.vertices(), m.edgeNodes()])
    g = P.NavierStokesSolver()
    g.initialize(m, 1060.0, 0.004)
    g.setDt(0.01)
    g.setNumberOfCorrections(2)
    g.setOptions(momentum_tol=TOL, pressure_tol=TOL, stop_rule=1, max_iter=20000,
                 pressure_solver=1)
    lab = np.r_[m.p1BC(), m.p2BC()]
    return m, g, fourier_basis(X), lab == 2


def test_cylinder_adjoint_vs_central_differences(cyl):
    m, g, B, outlet = cyl
    assert g.sizes()["nt"] == 911250
    c = 0.5 + np.random.default_rng(0).normal(0.0, 0.05, 4)
    J, dc, dmu = g.adjointGradient(STEPS, B, c)

    def J_of(coef):
        u = g.adjointForward(STEPS, np.tensordot(coef, B, axes=1))
        return 0.5 * float((u[outlet] ** 2).sum())

    assert abs(J_of(c) - J) <= 1e-12 * abs(J)
    h = 1e-6
    scale = np.abs(dc).max()
    for k in range(4):
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        fd = (J_of(cp) - J_of(cm)) / (2 * h)
        print(f"c{k}: adjoint {dc[k]:.10e} fd {fd:.10e}")
        assert abs(fd - dc[k]) <= 1e-5 * scale, (k, fd, dc[k])
    mu = 0.004
    Jm = {}
    for s in (-20, -10, -1, 1, 10, 20):
        g.setViscosity(mu + s * h)
        Jm[s] = J_of(c)
    g.setViscosity(mu)
    fd = (Jm[1] - Jm[-1]) / (2 * h)
    H = 10 * h
    fd5 = (Jm[-20] - 8 * Jm[-10] + 8 * Jm[10] - Jm[20]) / (12 * H)
    print(f"mu: adjoint {dmu:.10e} fd {fd:.10e} five-point {fd5:.10e}")
    assert abs(fd - dmu) <= 1e-5 * abs(dmu), (fd, dmu)
    assert abs(fd5 - dmu) <= 2e-6 * abs(dmu), (fd5, dmu)
