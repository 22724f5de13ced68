"""The allocation-free matrix entry point (cfsa_ns_apply_matrix_noalloc)
against cfsa_ns_apply_matrix through the C ABI: the same bits on two calls in
a row, the second with a smaller input, on one solver and then on a second
solver with a smaller mesh (every call shares the GPU's one scratch array),
and a clear error for inputs beyond the scratch array's size."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(P, n):
    m = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
    m.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    g = P.NavierStokesSolver()
    g.initialize(m, 1060.0, 0.004)
    g.setDt(0.01)
    g.setMeanInletVelocity((0.5, 0.0, 0.0))
    return m, g


def _noalloc(P, g, which, x, nrhs):
    sz = g.sizes()
    if which >= 5:
        y = np.empty((sz["nv"], 3), np.float64)
    else:
        n = sz["nu"] if which <= 2 else sz["nv"]
        y = np.empty((n, 3) if nrhs == 3 else (n,), np.float64)
    x = np.ascontiguousarray(x, np.float64)
    rc = P.library().cfsa_ns_apply_matrix_noalloc(g.handle, which, x, y, nrhs)
    return rc, y


def test_noalloc_matches_apply_matrix(P):
    rng = np.random.default_rng(7)
    m1, g1 = _solver(P, 6)
    assert g1.pisoStep(), P.last_error()
    nu, nv = g1.sizes()["nu"], g1.sizes()["nv"]
    print(f"noalloc: 6^3 solver nu={nu} nv={nv}", flush=True)
    # first call: the momentum operator on three right-hand sides
    x = rng.uniform(-1, 1, (nu, 3))
    rc, y = _noalloc(P, g1, 2, x, 3)
    assert rc == 0, P.last_error()
    assert np.array_equal(y, g1.applyMatrix(2, x, 3))
    # second call, smaller input: the pressure operator with its penalties
    xp = rng.uniform(-1, 1, nv)
    rc, yp = _noalloc(P, g1, 4, xp, 1)
    assert rc == 0, P.last_error()
    assert np.array_equal(yp, g1.applyMatrix(4, xp, 1))
    # the gradient operators (x has nv entries, y nv x 3)
    rc, yg = _noalloc(P, g1, 5, xp, 1)
    assert rc == 0, P.last_error()
    assert np.array_equal(yg, g1.applyMatrix(5, xp, 1))
    # a second solver on a smaller mesh shares the same array
    m2, g2 = _solver(P, 3)
    nu2 = g2.sizes()["nu"]
    x2 = rng.uniform(-1, 1, nu2)
    rc, y2 = _noalloc(P, g2, 0, x2, 1)
    assert rc == 0, P.last_error()
    assert np.array_equal(y2, g2.applyMatrix(0, x2, 1))
    # and the first solver again, after a time-step change (values re-derived in place)
    g1.setDt(0.02)
    rc, y3 = _noalloc(P, g1, 2, x, 3)
    assert rc == 0, P.last_error()
    assert np.array_equal(y3, g1.applyMatrix(2, x, 3))
    assert not np.array_equal(y3, y)


def test_noalloc_rejects_oversize(P):
    # 58^3 cells: about 1.6 M velocity nodes, so x and y of three right-hand
    # sides need about 9.6 Mi doubles, beyond the 8 Mi of the scratch array
    m, g = _solver(P, 58)
    nu = g.sizes()["nu"]
    assert 2 * 3 * nu > 8 << 20
    x = np.zeros((nu, 3))
    rc, _ = _noalloc(P, g, 2, x, 3)
    assert rc == 5
    assert "64 MiB" in P.last_error()
    # one right-hand side fits
    x1 = np.random.default_rng(1).uniform(-1, 1, nu)
    rc, y1 = _noalloc(P, g, 2, x1, 1)
    assert rc == 0, P.last_error()
    assert np.array_equal(y1, g.applyMatrix(2, x1, 1))
