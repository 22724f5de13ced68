"""GPU checks at the benchmark's full size (96^3 Kuhn channel: 5,308,416 tets,
7,189,057 velocity nodes, 204,763,393 momentum nonzeros), where the CPU oracle
is too slow to compare step by step.  They use properties that do not depend on
size: operator identities (K 1 = 0, L 1 = 0, G_c 1 = 0, lumped-mass row sums),
bit equality of the two Krylov SpMV layouts and of graph replay against eager
launches, determinism, and mixed precision against FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 96


@pytest.fixture(scope="module")
def channel(P):
    m = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, N, N, N)
    m.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    X = np.vstack([m.vertices(), m.edgeNodes()])
    U = 100 * 0.004 / 1060.0
    prof = np.zeros((len(X), 3))
    prof[:, 0] = U * 16.0 * X[:, 1] * (1 - X[:, 1]) * X[:, 2] * (1 - X[:, 2])
    return m, prof


def solver(P, m, prof, **opts):
    g = P.NavierStokesSolver()
    g.initialize(m, 1060.0, 0.004)
    g.setDt(1e-3)
    g.setInletProfile(prof)
    g.setOptions(**opts)
    return g


def test_sizes(P, channel):
    m, prof = channel
    g = solver(P, m, prof)
    sz = g.sizes()
    assert (sz["nt"], sz["nv"], sz["nu"], sz["nnzA"], sz["nnzP"]) == (
        5308416, 912673, 7189057, 204763393, 13465441)


def test_operator_identities(P, channel):
    m, prof = channel
    g = solver(P, m, prof)
    sz = g.sizes()
    ones_u = np.ones((sz["nu"], 3))
    ones_p = np.ones(sz["nv"])
    rpK, _, valK = g.getMatrix("K")
    Ku = g.applyMatrix("K", ones_u, 3)
    # per row |K 1| <= 1e-12 sum_j |K_ij|
    absrow = np.add.reduceat(np.abs(valK), rpK[:-1])
    assert np.all(np.abs(Ku[:, 0]) <= 1e-12 * absrow + 1e-300)
    rpL, _, valL = g.getMatrix("L")
    L1 = g.applyMatrix("L", ones_p)
    assert np.all(np.abs(L1) <= 1e-12 * np.add.reduceat(np.abs(valL), rpL[:-1]) + 1e-300)
    G1 = g.applyMatrix("Gx", ones_p)
    assert np.abs(G1).max() <= 1e-15
    # lumped mass: 0.45 rho V per vertex and (2/15 + 0.15) rho V per edge node,
    # summed over the node's tets; every Kuhn tet has volume h^3 / 6
    V = (1.0 / N) ** 3 / 6.0
    ilm = g.getInvLumpedMass()
    nv = sz["nv"]
    T = m.tets()
    E = m.tetEdgeNodes()
    inc = 0.45 * 1060.0 * V * np.bincount(T.ravel(), minlength=sz["nu"]).astype(float)
    inc += (2.0 / 15.0 + 0.15) * 1060.0 * V * np.bincount(nv + E.ravel(), minlength=sz["nu"])
    np.testing.assert_allclose(1.0 / ilm, inc, rtol=1e-12)


def test_spmv_layouts_bit_identical(P, channel):
    m, prof = channel
    g = solver(P, m, prof)
    sz = g.sizes()
    rng = np.random.default_rng(0)
    x3 = rng.uniform(-1, 1, (sz["nu"], 3))
    xp = rng.uniform(-1, 1, sz["nv"])
    out = []
    for layout in (1, 2):
        g.setOptions(spmv_layout=layout)
        out.append((g.applyMatrix("lhs", x3, 3), g.applyMatrix("Lbc", xp)))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_step_graphs_and_repeat_bit_identical(P, channel):
    """Two FP64 steps: graph replay, a second graph run and eager launches
    give the same bits and iteration counts."""
    m, prof = channel
    res = []
    for graphs in (1, 1, 0):
        g = solver(P, m, prof, graphs=graphs, stop_rule=1, momentum_tol=1e-8,
                    pressure_solver=1)
        for _ in range(2):
            assert g.pisoStep(), P.last_error()
        res.append((g.getVelocity(), g.getPressure(), g.stepStats()))
        del g
    for u, p, st in res[1:]:
        assert np.array_equal(res[0][0], u) and np.array_equal(res[0][1], p)
        assert st["momentum_iters"] == res[0][2]["momentum_iters"]
        assert st["pressure_iters"] == res[0][2]["pressure_iters"]


def test_mixed_matches_fp64(P, channel):
    """Mixed precision (FP32 CG inner solves) against FP64 after 2 steps at
    momentum 1e-10 / pressure 1e-8: velocity within 1e-7, pressure 1e-6."""
    m, prof = channel
    tols = dict(stop_rule=1, momentum_tol=1e-10, pressure_tol=1e-8, pressure_solver=1)
    g64 = solver(P, m, prof, **tols)
    for _ in range(2):
        assert g64.pisoStep(), P.last_error()
    u64, p64 = g64.getVelocity(), g64.getPressure()
    del g64
    g32 = solver(P, m, prof, precision=2, momentum_solver=1, **tols)
    for _ in range(2):
        assert g32.pisoStep(), P.last_error()
    eu = np.linalg.norm(g32.getVelocity() - u64) / np.linalg.norm(u64)
    ep = np.linalg.norm(g32.getPressure() - p64) / np.linalg.norm(p64)
    print(f"velocity {eu:.3e} pressure {ep:.3e}")
    assert eu <= 1e-7 and ep <= 1e-6, (eu, ep)
