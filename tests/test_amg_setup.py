"""Host set-up of the smoothed-aggregation hierarchy (cfsa_amg_build, the code
the solver runs for pressure_solver 2) on the oracle's pressure operators,
checked against what the mathematics fixes:

* strength and aggregation follow the three greedy passes (an independent
  restatement here); penalty rows (no strong connection) stay unaggregated;
* P = (I - 1.2/lmax D^-1 A) T with T the normalised aggregate indicator and
  R = P^T and A_c = R A P (scipy products);
* the spectral radius estimate is within 10% below (and at most 0.1% above)
  the largest eigenvalue of D^-1 A, and the Chebyshev interval's top lmax
  covers it (so the prolongator weight 1.2 / lmax keeps omega rho <= 1.2);
* the smoothing coefficients give the error polynomial
  T_2((theta - l)/delta) / T_2(theta/delta) of the degree-2 Chebyshev
  smoother (theta, delta: centre and half width of [lmax/30, lmax]);
* the coarsest inverse inverts the coarsest operator;
* the V-cycle (tests/amg_ref.py) is symmetric positive definite and, as a CG
  preconditioner, needs a fraction of the Jacobi iterations.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from amg_ref import VCycle, csr, pcg

THETA = 0.06


def _lbc(O, kind):
    if kind == "cyl":
        import cfsa_ns as P
        pm = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
        m = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    else:
        n = {"kuhn8": 8, "kuhn10n": 10}[kind]
        m = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
    m.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    ns = O.OracleNS(m, 1060.0, 0.004)
    ns.setDt(1e-3)
    if kind == "kuhn10n":
        ns.setOutlet(1, 0.0)
        ns.setReferencePressureNode(5)
    else:
        ns.setOutlet(0, 0.0)
    rp, col, val = ns.getMatrix("Lbc")
    return rp, col, val


@pytest.fixture(scope="module", params=["kuhn8", "kuhn10n", "cyl"])
def case(request, O, P):
    rp, col, val = _lbc(O, request.param)
    n = len(rp) - 1
    A = csr(rp, col, val, (n, n))
    H = P.HostAmg(rp, col, val, theta=THETA, max_coarse=60)
    return request.param, A, H


def _aggregate(A, theta):
    """The three greedy passes over strong connections, written out."""
    n = A.shape[0]
    d = np.abs(A.diagonal())
    C = A.tocoo()
    keep = (C.row != C.col) & (np.abs(C.data) >= theta * np.sqrt(d[C.row] * d[C.col]))
    S = sp.csr_matrix((np.ones(keep.sum()), (C.row[keep], C.col[keep])), shape=A.shape)
    S.sort_indices()
    nb = [S.indices[S.indptr[i]:S.indptr[i + 1]] for i in range(n)]
    agg = -np.ones(n, np.int64)
    na = 0
    for i in range(n):
        if agg[i] < 0 and len(nb[i]) and np.all(agg[nb[i]] < 0):
            agg[i] = na
            agg[nb[i]] = na
            na += 1
    for i in range(n):
        if agg[i] < 0:
            a = agg[nb[i]]
            a = a[a >= 0]
            if len(a):
                agg[i] = a[0]
    for i in range(n):
        if agg[i] < 0 and len(nb[i]):
            agg[i] = na
            for j in nb[i]:
                if agg[j] < 0:
                    agg[j] = na
            na += 1
    return agg, na


def test_levels_and_galerkin_products(case):
    kind, A0, H = case
    L, nc, cjac = H.levels()
    assert L >= 2 and not cjac and nc <= 60
    A = A0
    for l in range(L - 1):
        rp, ci, v, shape = H.matrix(l, "A")
        Al = csr(rp, ci, v, shape)
        assert abs(Al - A).max() <= 1e-12 * abs(A).max()
        prm = H.params(l)
        agg, na = _aggregate(A, THETA)
        np.testing.assert_array_equal(prm["agg"], agg)
        # penalty rows (diagonal >= 1e11) have no strong connection
        big = A.diagonal() >= 1e11
        assert np.all(agg[big] == -1)
        n = A.shape[0]
        cnt = np.bincount(agg[agg >= 0], minlength=na)
        rows = np.nonzero(agg >= 0)[0]
        T = sp.csr_matrix((1.0 / np.sqrt(cnt[agg[rows]]), (rows, agg[rows])), shape=(n, na))
        Dinv = sp.diags(1.0 / A.diagonal())
        Pw = T - (1.2 / prm["lmax"]) * (Dinv @ (A @ T))
        Pl = csr(*H.matrix(l, "P")[:3], H.matrix(l, "P")[3])
        Rl = csr(*H.matrix(l, "R")[:3], H.matrix(l, "R")[3])
        assert Pl.shape == (n, na)
        assert abs(Pl - Pw).max() <= 1e-13 * abs(Pw).max()
        assert (Rl != Pl.T).nnz == 0
        Ac = (Pl.T @ A @ Pl).tocsr()
        Ah = csr(*H.matrix(l + 1, "A")[:3], H.matrix(l + 1, "A")[3])
        assert abs(Ah - Ac).max() <= 1e-12 * abs(Ac).max(), (kind, l)
        A = Ah


def test_no_stored_zeros(case):
    """Every operator of the hierarchy keeps only nonzero entries; level 0 holds
    exactly the nonzero entries of the input (the Kuhn Laplacian stores about
    half of its entries as exact zeros), in the input's order."""
    kind, A0, H = case
    L, _, _ = H.levels()
    for l in range(L):
        for w in ("A", "P", "R") if l + 1 < L else ("A",):
            rp, ci, v, shape = H.matrix(l, w)
            assert np.all(v != 0.0), (kind, l, w)
            assert np.all(np.diff(rp) >= 0)
    rp, ci, v, _ = H.matrix(0, "A")
    keep = A0.data != 0.0
    np.testing.assert_array_equal(v, A0.data[keep])
    np.testing.assert_array_equal(ci, A0.indices[keep])
    if kind.startswith("kuhn"):
        assert keep.mean() < 0.6


def test_spectral_bounds_and_chebyshev_polynomial(case):
    kind, A0, H = case
    L, _, _ = H.levels()
    for l in range(L):
        Al = csr(*H.matrix(l, "A")[:3], H.matrix(l, "A")[3])
        prm = H.params(l)
        d = Al.diagonal()
        Ds = sp.diags(1.0 / np.sqrt(d))
        S = (Ds @ Al @ Ds).tocsr()
        if S.shape[0] > 400:
            lam = sla.eigsh(S, k=1, which="LA", return_eigenvectors=False, tol=1e-10)[0]
        else:
            lam = np.linalg.eigvalsh(S.toarray())[-1]
        # the power iteration measures |D^-1 A x| / |x| of a non-symmetric
        # matrix: close to the spectral radius from below, possibly a hair above
        assert prm["rho"] <= lam * (1 + 1e-3) and prm["rho"] >= 0.9 * lam, (l, prm["rho"], lam)
        assert prm["lmax"] >= lam * (1 - 1e-12), (l, prm["lmax"], lam)
        c0, c1, c2 = prm["coef"]
        lmax = prm["lmax"]
        lmin = lmax / 30.0
        th, de = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
        t2 = lambda t: 2 * t * t - 1
        for lv in np.linspace(0.0, lmax, 17):
            # x0 = 0: e1 = (1 - c0 l) e0, e2 = (1 + c1 - c2 l) e1 - c1 e0
            q = (1 + c1 - c2 * lv) * (1 - c0 * lv) - c1
            want = t2((th - lv) / de) / t2(th / de)
            assert abs(q - want) <= 1e-12, (l, lv, q, want)


def test_coarse_inverse(case):
    kind, A0, H = case
    L, nc, _ = H.levels()
    Ac = csr(*H.matrix(L - 1, "A")[:3], H.matrix(L - 1, "A")[3]).toarray()
    C = H.coarseInverse()
    np.testing.assert_allclose(C @ Ac, np.eye(nc), atol=1e-9)


@pytest.mark.parametrize("kind,outlet", [("kuhn6", 0), ("kuhn6n", 1)])
def test_single_level_inverse_with_penalty_rows(P, O, kind, outlet):
    """Operators below max_coarse are one level inverted densely, 1e12 penalty
    rows beside rows of order h included: C A = I."""
    m = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, 6, 6, 6)
    m.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    ns = O.OracleNS(m, 1060.0, 0.004)
    ns.setOutlet(outlet, 0.0)
    if outlet:
        ns.setReferencePressureNode(5)
    rp, col, val = ns.getMatrix("Lbc")
    n = len(rp) - 1
    H = P.HostAmg(rp, col, val)
    assert H.levels() == (1, n, False)
    A = csr(rp, col, val, (n, n)).toarray()
    np.testing.assert_allclose(H.coarseInverse() @ A, np.eye(n), atol=1e-8)


def test_semidefinite_coarse_pseudo_inverse(P):
    """A pure Neumann Laplacian (constants in the null space): the dropped
    pivot leaves a pseudo-inverse that still solves every consistent system."""
    n = 50
    main = np.full(n, 2.0)
    main[0] = main[-1] = 1.0
    A = sp.diags([main, -np.ones(n - 1), -np.ones(n - 1)], [0, -1, 1]).tocsr()
    A.sort_indices()
    H = P.HostAmg(A.indptr, A.indices, A.data)
    C = H.coarseInverse()
    b = np.random.default_rng(1).standard_normal(n)
    b -= b.mean()
    x = C @ b
    np.testing.assert_allclose(A @ x, b, atol=1e-9)


def test_vcycle_spd_and_effective(case):
    kind, A0, H = case
    V = VCycle(H)
    n = A0.shape[0]
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    Mu, Mv = V.apply(u), V.apply(v)
    assert abs(Mu @ v - u @ Mv) <= 1e-10 * np.sqrt((Mu @ u) * (Mv @ v))
    assert Mu @ u > 0 and Mv @ v > 0
    b = A0 @ rng.standard_normal(n)
    d = A0.diagonal()
    xa, ita = pcg(A0, b, V.apply, 1e-10)
    xj, itj = pcg(A0, b, lambda r: r / d, 1e-10)
    print(f"{kind}: n {n}, levels {H.levels()[0]}, AMG-CG {ita} vs Jacobi-CG {itj} iterations")
    assert ita * 4 <= itj, (ita, itj)
    assert np.linalg.norm((A0 @ xa - b) / d) <= 1e-10 * np.linalg.norm(b / d)


def test_deterministic_and_bad_input(P, O):
    rp, col, val = _lbc(O, "kuhn8")
    H1 = P.HostAmg(rp, col, val)
    H2 = P.HostAmg(rp, col, val)
    assert H1.levels() == H2.levels()
    for l in range(H1.levels()[0] - 1):
        for w in "APR":
            a, b = H1.matrix(l, w), H2.matrix(l, w)
            for x, y in zip(a[:3], b[:3]):
                assert np.array_equal(x, y)
    bad = col.copy()
    bad[1], bad[2] = bad[2], bad[1]  # columns out of order in row 0
    with pytest.raises(P.CfsaError):
        P.HostAmg(rp, bad, val)
