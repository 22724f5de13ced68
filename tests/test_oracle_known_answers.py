"""Pins the CPU oracle against closed forms, invariants and brute force.

The reference ships no tests or golden vectors, so every function of the
oracle is pinned here by mathematics that a plausible slip (dropped term,
wrong sign, transposed operand, wrong weight) would break.
"""
import numpy as np
import pytest


def _prism(O, n, L=1.0, W=1.0, H=1.0):
    m = O.OracleMesh.structuredRectangularPrism(L, W, H, n, n, n)
    m.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    return m


def _dense(rp, col, val, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return A


# ---------------------------------------------------------------- B1 geometry
@pytest.mark.parametrize("n", [1, 2, 3])
def test_gradients_and_volumes(O, n):
    m = _prism(O, n, 1.0, 2.0, 3.0)
    X, T, V, G = m.vertices(), m.tets(), m.volumes(), m.gradients()
    h = np.array([1.0 / n, 2.0 / n, 3.0 / n])
    np.testing.assert_allclose(V, np.prod(h) / 6.0, rtol=1e-14)
    assert abs(V.sum() - 6.0) < 1e-12
    np.testing.assert_allclose(G.sum(axis=1), 0.0, atol=1e-12)
    for t in range(len(T)):
        x0 = X[T[t, 0]]
        for i in range(4):
            for j in range(1, 4):
                want = (1.0 if i == j else 0.0) - (1.0 if i == 0 else 0.0)
                assert abs(G[t, i] @ (X[T[t, j]] - x0) - want) < 1e-12


# --------------------------------------------------- B9 structured counts
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_kuhn_counts(O, n):
    m = _prism(O, n)
    nv, nt, ne, nf = m.sizes()
    assert nv == (n + 1) ** 3
    assert nt == 6 * n ** 3
    assert ne == 7 * n ** 3 + 9 * n ** 2 + 3 * n
    # Euler: V - E + F - T = 1 for a ball
    assert nv - ne + nf - nt == 1


def test_unit_cube_labels_by_hand(O):
    """1x1x1 Kuhn cube: every vertex is on the boundary.  x=0 face centres lie
    on the inlet plane, x=1 on the outlet plane, side faces have centres at
    x = 1/3 or 2/3 (>= tol 0.01) so stay Wall; the last face (sorted-triple
    order) touching a vertex decides."""
    m = _prism(O, 1)
    X = m.vertices()
    v, t, g = m.faces()
    bnd = t[:, 1] == -1
    lab = np.full(len(X), 3, np.uint8)
    for f in np.nonzero(bnd)[0]:
        c = g[f, 4:7]
        for vi in v[f]:
            lab[vi] = 0
        if abs(c[0] - 0.0) < 0.01:
            lab[v[f]] = 1
        elif abs(c[0] - 1.0) < 0.01:
            lab[v[f]] = 2
    np.testing.assert_array_equal(m.p1BC(), lab)
    # faces come in ascending sorted-triple order
    keys = [tuple(sorted(r)) for r in v]
    assert keys == sorted(keys)


# ------------------------------------------------------------- B2-B6 operators
@pytest.fixture(scope="module")
def ns3(O):
    m = _prism(O, 3, 1.0, 1.5, 2.0)
    ns = O.OracleNS(m, 1060.0, 0.004)
    ns.setDt(0.01)
    return m, ns


def test_mass_invariants(O, ns3):
    m, ns = ns3
    sz = ns.sizes()
    M = _dense(*ns.getMatrix("M"), sz["nu"])
    Vt = m.volumes().sum()
    assert abs(M.sum() - 3.5 * 1060.0 * Vt) < 1e-9 * M.sum()
    np.testing.assert_allclose(M, M.T, rtol=0, atol=1e-15 * abs(M).max())
    # lumped row sums: per tet 0.45 rho V for a vertex, (2/15 + 0.15) rho V for an edge
    T, E, V = m.tets(), m.tetEdgeNodes(), m.volumes()
    want = np.zeros(sz["nu"])
    for t in range(len(T)):
        want[T[t]] += 0.45 * 1060.0 * V[t]
        want[sz["nv"] + E[t]] += (2.0 / 15.0 + 0.15) * 1060.0 * V[t]
    np.testing.assert_allclose(M.sum(axis=1), want, rtol=1e-13)
    np.testing.assert_allclose(ns.getInvLumpedMass(), 1.0 / want, rtol=1e-13)


def test_single_tet_local_blocks(O):
    """One Kuhn tet of the unit cube: assemble by hand from the reference's
    formulas and compare entry by entry."""
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]], float)
    m = O.OracleMesh.fromArrays(X, np.array([[0, 1, 2, 3]], np.int32))
    m.setupBoundaryConditions()
    rho, mu = 2.0, 0.5
    ns = O.OracleNS(m, rho, mu)
    V = 1.0 / 6.0
    g = np.linalg.inv(np.c_[np.ones(4), X])[1:].T  # rows = grad phi_i
    E = m.tetEdgeNodes()[0]
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    ge = np.array([g[a] + g[b] for a, b in pairs])
    nodes = list(range(4)) + [4 + e for e in E]
    Gall = np.vstack([g, ge])
    Kw = (1.0 / mu) * V * Gall @ Gall.T
    Mw = np.zeros((10, 10))
    mass = rho * V
    for a in range(10):
        for b in range(10):
            if a < 4 and b < 4:
                Mw[a, b] = mass / 10 if a == b else mass / 20
            elif a >= 4 and b >= 4:
                Mw[a, b] = mass / 15 if a == b else mass / 60
            else:
                Mw[a, b] = mass / 30
    Kd = _dense(*ns.getMatrix("K"), 10)
    Md = _dense(*ns.getMatrix("M"), 10)
    P = np.zeros((10, 10))
    for a, na in enumerate(nodes):
        P[na, a] = 1
    np.testing.assert_allclose(Kd, P @ Kw @ P.T, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(Md, P @ Mw @ P.T, rtol=1e-14)
    L = _dense(*ns.getMatrix("L"), 4)
    np.testing.assert_allclose(L, V * g @ g.T, atol=1e-14)
    for c, name in enumerate(["Gx", "Gy", "Gz"]):
        Gc = _dense(*ns.getMatrix(name), 4)
        np.testing.assert_allclose(Gc, np.tile(V / 4 * g[:, c], (4, 1)), atol=1e-15)


def test_stiffness_poisson_gradient_invariants(O, ns3):
    m, ns = ns3
    sz = ns.sizes()
    K = _dense(*ns.getMatrix("K"), sz["nu"])
    L = _dense(*ns.getMatrix("L"), sz["nv"])
    np.testing.assert_allclose(K @ np.ones(sz["nu"]), 0.0, atol=1e-10 * abs(K).max())
    np.testing.assert_allclose(K, K.T, atol=1e-12 * abs(K).max())
    np.testing.assert_allclose(L @ np.ones(sz["nv"]), 0.0, atol=1e-13 * abs(L).max())
    assert np.linalg.eigvalsh(L).min() > -1e-10
    X = m.vertices()
    lab = m.p1BC()
    a = np.array([0.3, -1.2, 0.7])
    p = X @ a
    # patch test: (L p)_i = 0 at interior vertices for linear p
    Lp = L @ p
    assert np.abs(Lp[lab == 3]).max() < 1e-12
    # G_c p = a_c * sum_{T ni i} V_T / 4 for linear p
    T, V = m.tets(), m.volumes()
    wsum = np.zeros(sz["nv"])
    for t in range(len(T)):
        wsum[T[t]] += V[t] / 4
    for c, name in enumerate(["Gx", "Gy", "Gz"]):
        G = _dense(*ns.getMatrix(name), sz["nv"])
        np.testing.assert_allclose(G @ np.ones(sz["nv"]), 0.0, atol=1e-14)
        np.testing.assert_allclose(G @ p, a[c] * wsum, rtol=1e-12, atol=1e-14)


def test_divergence_and_convection_linear_fields(O, ns3):
    m, ns = ns3
    sz = ns.sizes()
    nv, nu = sz["nv"], sz["nu"]
    X = np.vstack([m.vertices(), m.edgeNodes()])
    A = np.array([[0.2, -0.4, 0.1], [0.5, 0.3, -0.2], [-0.1, 0.6, 0.4]])
    b = np.array([0.05, -0.1, 0.2])
    u = X @ A.T + b
    T, V, E = m.tets(), m.volumes(), m.tetEdgeNodes()
    wsum = np.zeros(nv)
    for t in range(len(T)):
        wsum[T[t]] += V[t] / 4
    np.testing.assert_allclose(ns.computeDivergence(u), np.trace(A) * wsum, rtol=1e-12)
    # divergence equals sum_c G_c u_c on the vertices (B6)
    G = [_dense(*ns.getMatrix(nm), nv) for nm in ("Gx", "Gy", "Gz")]
    np.testing.assert_allclose(ns.computeDivergence(u), sum(G[c] @ u[:nv, c] for c in range(3)),
                               rtol=1e-12, atol=1e-14)
    ns.setVelocity(u)
    conv = ns.computeAdvectionRHS()
    want = np.zeros((nu, 3))
    Xv = m.vertices()
    for t in range(len(T)):
        xc = Xv[T[t]].mean(axis=0)
        c = A @ (A @ xc + b)
        want[T[t]] -= V[t] / 4 * c
        want[nv + E[t]] -= V[t] / 24 * c
    np.testing.assert_allclose(conv, want, rtol=1e-11, atol=1e-15)


# ---------------------------------------------------- B10 Krylov restatements
def test_eigen_dot_reduction_order(O):
    rng = np.random.default_rng(0)
    for n in [1, 2, 3, 4, 5, 6, 7, 8, 9, 17, 1000, 1001]:
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        p = a * b
        if n < 2:
            want = p[0]
        else:
            k2, k1 = (n // 4) * 4, (n // 2) * 2
            if k1 == 2:
                lanes = [p[0], p[1]]
            else:
                l0, l1 = p[0], p[1]
                l2, l3 = p[2], p[3]
                for i in range(4, k2, 4):
                    l0, l1, l2, l3 = l0 + p[i], l1 + p[i + 1], l2 + p[i + 2], l3 + p[i + 3]
                lanes = [l0 + l2, l1 + l3]
                if k1 > k2:
                    lanes = [lanes[0] + p[k2], lanes[1] + p[k2 + 1]]
            want = lanes[0] + lanes[1]
            for i in range(k1, n):
                want = want + p[i]
        assert O.eigen_dot(a, b) == want, n
        assert abs(O.eigen_dot(a, b) - np.dot(a, b)) < 1e-12 * max(1, abs(np.dot(a, b)))


def test_dot2_correctly_rounded_and_order_free(O):
    """Dot2 against the exactly rounded value of the exact inner product
    (rational arithmetic); a permutation of the terms changes nothing; and a
    cancellation that defeats the plain loop is resolved."""
    from fractions import Fraction
    rng = np.random.default_rng(11)
    for trial in range(30):
        n = int(rng.integers(1, 600))
        a = rng.standard_normal(n) * 10.0 ** rng.integers(-4, 5, n)
        b = rng.standard_normal(n) * 10.0 ** rng.integers(-4, 5, n)
        exact = sum((Fraction(x) * Fraction(y) for x, y in zip(a, b)), Fraction(0))
        want = float(exact)
        assert O.dot2(a, b) == want, trial
        perm = rng.permutation(n)
        assert O.dot2(a[perm], b[perm]) == want, trial
    a = np.array([1e16, 1.0, -1e16])
    b = np.ones(3)
    assert O.dot2(a, b) == 1.0 and O.eigen_dot(a, b) == 0.0
    assert O.dot2(np.zeros(0), np.zeros(0)) == 0.0


def test_step_with_eigen_dot_order_agrees(O):
    """The Krylov restatements with Eigen's packet-order inner products and with
    Dot2 converge to the same step at tight tolerances (the inner-product
    rounding only moves the iterates at the level of the tolerance)."""
    out = []
    try:
        for mode in (1, 0):
            O.set_dot_mode(mode)
            m = _prism(O, 3, 1.0, 1.0, 1.0)
            m.setupBoundaryConditions()
            ns = O.OracleNS(m, 1060.0, 0.004)
            ns.setDt(0.01)
            ns.setMeanInletVelocity((0.5, 0.0, 0.0))
            ns.setSolver(1e-13, 1e-13, 5000, 0)
            ns.setStopScaled(True)
            for _ in range(3):
                assert ns.pisoStep() == 0
            out.append((ns.getVelocity(), ns.getPressure()))
    finally:
        O.set_dot_mode(0)
    for a, b in zip(out[0], out[1]):
        assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b)


def test_krylov_matches_dense_solve(O, ns3):
    m, ns = ns3
    sz = ns.sizes()
    rng = np.random.default_rng(1)
    for which, n in (("lhs", sz["nu"]), ("Lbc", sz["nv"])):
        A = _dense(*ns.getMatrix(which), n)
        b = rng.standard_normal(n)
        xd = np.linalg.solve(A, b)
        for method in (0, 1):
            x, info = ns.krylovSolve(which, b, method=method, tol=1e-13, maxit=5000)
            assert info["err"] <= 1e-13
            np.testing.assert_allclose(x, xd, rtol=0, atol=1e-9 * abs(xd).max())
        x, info = ns.krylovSolve(which, np.zeros(n), method=0)
        assert info["iters"] == 0 and not x.any()


def test_csr_spmv_order(O):
    rp = np.array([0, 3, 4], np.int32)
    col = np.array([0, 1, 2, 2], np.int32)
    val = np.array([1e16, 1.0, -1e16, 2.0])
    x = np.array([1.0, 1.0, 1.0])
    y = O.csr_spmv(rp, col, val, x)
    # ((0 + 1e16) + 1) + (-1e16) = 0 in ascending-column order
    assert y[0] == 0.0 and y[1] == 2.0


def test_colouring_is_a_greedy_proper_colouring(O):
    m = _prism(O, 3)
    T = m.tets()
    col, nc = O.colour_tets(T, m.sizes()[0])
    assert nc == col.max() + 1
    share = {}
    for t, tv in enumerate(T):
        for v in tv:
            share.setdefault(v, []).append(t)
    for t, tv in enumerate(T):
        nb = {o for v in tv for o in share[v] if o != t}
        assert all(col[o] != col[t] for o in nb)
        earlier = {col[o] for o in nb if o < t}
        assert all(c in earlier for c in range(col[t]))


# ------------------------------------------------------------ whole step
def test_quiescent_step_stays_at_rest(O):
    """Zero inlet velocity and zero outlet pressure: u = 0, p = 0 is a fixed
    point of the step (every right-hand side vanishes)."""
    m = _prism(O, 3)
    ns = O.OracleNS(m, 1060.0, 0.004)
    ns.setDt(0.01)
    assert ns.pisoStep() == 0
    assert not ns.getVelocity().any() and not ns.getPressure().any()


def test_step_enforces_boundary_values(O):
    m = _prism(O, 3)
    ns = O.OracleNS(m, 1060.0, 0.004)
    ns.setDt(0.01)
    ns.setMeanInletVelocity((0.5, 0.0, 0.0))
    assert ns.pisoStep() == 0
    u, p = ns.getVelocity(), ns.getPressure()
    lab = np.r_[m.p1BC(), m.p2BC()]
    np.testing.assert_array_equal(u[lab == 1], np.tile([0.5, 0.0, 0.0], ((lab == 1).sum(), 1)))
    assert not u[lab == 0].any()
    assert not p[m.p1BC() == 2].any()
    assert np.isfinite(u).all() and np.abs(u).max() < 10


# ------------------------------------------------------- boundary forces (f1)
def test_forces_closed_forms(O):
    """Linear velocity u = A x + b: every tet has grad u = A, so each boundary
    face carries traction 2 mu sym(A) n area, one third to each vertex.  A
    uniform pressure p0 gives -p0 n area / 3 per face vertex."""
    m = _prism(O, 2, 1.0, 1.5, 2.0)
    mu = 0.004
    ns = O.OracleNS(m, 1060.0, mu)
    sz = ns.sizes()
    X = np.vstack([m.vertices(), m.edgeNodes()])
    A = np.array([[0.2, -0.4, 0.1], [0.5, 0.3, -0.2], [-0.1, 0.6, 0.4]])
    ns.setVelocity(X @ A.T + 0.3)
    ns.setPressure(np.full(sz["nv"], 7.0))
    v, t, g = m.faces()
    S = 2 * mu * 0.5 * (A + A.T)
    want_s = np.zeros((sz["nv"], 3))
    want_p = np.zeros((sz["nv"], 3))
    for f in np.nonzero(t[:, 1] == -1)[0]:
        n, area = g[f, :3], g[f, 3]
        for vi in v[f]:
            want_s[vi] += S @ n * area / 3
            want_p[vi] += -7.0 * n * area / 3
    np.testing.assert_allclose(ns.forces(1), want_s, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(ns.forces(0), want_p, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(ns.forces(2), want_p + want_s, rtol=1e-12, atol=1e-15)
    # boundary normals are unit vectors, areas sum to the box surface
    assert np.allclose(np.linalg.norm(g[t[:, 1] == -1, :3], axis=1), 1.0)
    assert abs(g[t[:, 1] == -1, 3].sum() - 2 * (1.5 + 2.0 + 3.0)) < 1e-12


# ------------------------------------------------ BiCGSTAB restart branch
def _eigen_bicgstab(rp, col, val, b, tol, maxit, scale):
    """Eigen's BiCGSTAB with the Jacobi preconditioner, restated from its
    published algorithm (x0 = 0, restart when |rho| < scale eps^2 |r0|^2,
    iteration count reset on the first restart), with exactly rounded inner
    products (what the oracle's Dot2 sums give) and row sums in column order."""
    from fractions import Fraction
    n = len(b)

    def dot(a, c):
        return float(sum((Fraction(x) * Fraction(y) for x, y in zip(a, c)), Fraction(0)))

    def spmv(x):
        y = []
        for i in range(n):
            s = 0.0
            for k in range(rp[i], rp[i + 1]):
                s += val[k] * x[col[k]]
            y.append(s)
        return y

    D = [1.0] * n
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            if col[k] == i and val[k] != 0.0:
                D[i] = 1.0 / val[k]
    x = [0.0] * n
    Ax = spmv(x)
    r = [b[i] - Ax[i] for i in range(n)]
    r0 = list(r)
    r0sq = dot(r0, r0)
    bsq = dot(b, b)
    rho = alpha = w = 1.0
    v = [0.0] * n
    p = [0.0] * n
    tol2 = tol * tol * bsq
    eps2 = np.finfo(float).eps ** 2
    i = restarts = 0
    while dot(r, r) > tol2 and i < maxit:
        rho_old = rho
        rho = dot(r0, r)
        if abs(rho) < scale * eps2 * r0sq:
            Ax = spmv(x)
            r = [b[k] - Ax[k] for k in range(n)]
            r0 = list(r)
            rho = r0sq = dot(r, r)
            if restarts == 0:
                i = 0
            restarts += 1
        beta = (rho / rho_old) * (alpha / w)
        p = [r[k] + beta * (p[k] - w * v[k]) for k in range(n)]
        y = [D[k] * p[k] for k in range(n)]
        v = spmv(y)
        alpha = rho / dot(r0, v)
        s = [r[k] - alpha * v[k] for k in range(n)]
        z = [D[k] * s[k] for k in range(n)]
        t = spmv(z)
        tt = dot(t, t)
        w = dot(t, s) / tt if tt > 0.0 else 0.0
        x = [x[k] + (alpha * y[k] + w * z[k]) for k in range(n)]
        r = [s[k] - w * t[k] for k in range(n)]
        i += 1
    return np.array(x), i, restarts


def test_bicgstab_restart_branch(O):
    """A nonsymmetric convection-diffusion-like system; with the restart
    threshold raised the solve passes through Eigen's restart branch several
    times.  The oracle must take the same iterations and restarts as the
    restated algorithm and return the same bits, and still solve the system
    (dense LU) to its tolerance."""
    n = 40
    rng = np.random.default_rng(5)
    A = np.diag(4.0 + rng.uniform(0, 1, n))
    for i in range(n - 1):
        A[i, i + 1] = -1.0 + 0.6
        A[i + 1, i] = -1.0 - 0.6
    A[0, n - 1] = 0.3
    rp = np.zeros(n + 1, np.int32)
    col, val = [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        col += nz.tolist()
        val += A[i, nz].tolist()
        rp[i + 1] = len(col)
    col = np.array(col, np.int32)
    val = np.array(val)
    b = rng.uniform(-1, 1, n)
    tol = 1e-12
    try:
        for scale in (1.0, 1e27):
            O.set_restart_scale(scale)
            x, info = O.csr_krylov(rp, col, val, b, method=0, tol=tol, maxit=500)
            xr, it, rs = _eigen_bicgstab(rp.tolist(), col.tolist(), val.tolist(), b.tolist(),
                                         tol, 500, scale)
            assert O.last_restarts() == rs
            assert info["iters"] == it
            np.testing.assert_array_equal(x, xr)
            assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
            if scale > 1.0:
                assert rs >= 2
            else:
                assert rs == 0
    finally:
        O.set_restart_scale(1.0)
