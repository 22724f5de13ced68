"""Pins of the oracle's mesh numbering, labels, face orientation and PISO
step composition against values derived by hand from the reference's rules
and against an independent dense restatement of the step.

* P2 edge nodes: numbered by first appearance scanning tets in order, local
  edges (01, 02, 03, 12, 13, 23) (src/Mesh/Mesh.cpp:641-677).
* P2 labels: the shared endpoint label when both endpoints agree and are not
  Internal, else Internal (include/Mesh/Mesh.hpp:198-215).
* Face normals: ac x ab of the owner tet's local face, normalised
  (Mesh.cpp:461-467); internal faces flipped (with vertids[1], vertids[2]
  swapped) when the normal does not point from the owner to the neighbour;
  the nested else-if never touches boundary faces (Mesh.cpp:615-639).
* Step: momentum rhs (M u)/dt - G p (P1 rows) + conv + penalties, lhs
  M/dt + K + penalties (NavierStokes.cpp:416-524); pressure rhs -div/dt with
  outlet penalties, p += relax_p p', outlet reset, normalisation
  (:586-668, :804-813, :1011-1024); correction of P1 rows only, P2 rows keep
  u* (:702-771); final BC reset (:815-839).
"""
import numpy as np
import pytest

INLET, OUTLET, WALL, INTERNAL = 1, 2, 0, 3


def _cube(O, n=1, bc=True):
    m = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
    if bc:
        m.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    return m


# ------------------------------------------------------------ P2 numbering
def test_p2_numbering_unit_cube_by_hand(O):
    """Unit cube, tets {0,1,3,7} {0,3,2,7} {0,2,6,7} {0,6,4,7} {0,4,5,7}
    {0,5,1,7} (Mesh.cpp:195-206); edges numbered as first met."""
    m = _cube(O, 1)
    want = np.array([[0, 1, 2, 3, 4, 5],
                     [1, 6, 2, 7, 5, 8],
                     [6, 9, 2, 10, 8, 11],
                     [9, 12, 2, 13, 11, 14],
                     [12, 15, 2, 16, 14, 17],
                     [15, 0, 2, 18, 17, 4]], np.int32)
    np.testing.assert_array_equal(m.tetEdgeNodes(), want)
    ends = [(0, 1), (0, 3), (0, 7), (1, 3), (1, 7), (3, 7), (0, 2), (2, 3), (2, 7), (0, 6),
            (2, 6), (6, 7), (0, 4), (4, 6), (4, 7), (0, 5), (4, 5), (5, 7), (1, 5)]
    X = m.vertices()
    mid = np.array([0.5 * (X[a] + X[b]) for a, b in ends])
    np.testing.assert_array_equal(m.edgeNodes(), mid)


@pytest.mark.parametrize("n", [2, 3])
def test_p2_numbering_is_first_appearance(O, n):
    """Scanning tets and their local edges in order, every edge not seen
    before carries the next id, and every later sighting the same id."""
    m = _cube(O, n)
    T, E = m.tets(), m.tetEdgeNodes()
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    seen = {}
    nxt = 0
    for t in range(len(T)):
        for k, (a, b) in enumerate(pairs):
            key = tuple(sorted((int(T[t, a]), int(T[t, b]))))
            if key not in seen:
                assert E[t, k] == nxt, (t, k)
                seen[key] = nxt
                nxt += 1
            else:
                assert E[t, k] == seen[key], (t, k)
    assert nxt == m.sizes()[2]
    # a sorted-pair (std::map order) numbering would differ from this one
    order = sorted(seen, key=lambda k: k)
    assert [seen[k] for k in order] != list(range(len(order)))


# ------------------------------------------------------------ labels
def test_unit_cube_p1_labels_by_hand(O):
    """Boundary faces in sorted-triple order: 013 015 023 026 045 046 137 157
    237 267 457 467; x=0 faces (026, 046) Inlet, x=1 faces (137, 157)
    Outlet, the rest Wall (centres at x = 1/3 or 2/3); the last face touching
    a vertex decides: 0 <- 046 I, 1 <- 157 O, 2 <- 267 W, 3 <- 237 W,
    4, 6, 7 <- 467 W, 5 <- 457 W."""
    m = _cube(O, 1)
    np.testing.assert_array_equal(m.p1BC(), [1, 2, 0, 0, 0, 0, 0, 0])
    # P2: edges with an Inlet/Outlet/Wall mix or an Internal end are Internal
    want = [3, 3, 3, 3, 3, 0, 3, 0, 0, 3, 0, 0, 3, 0, 0, 3, 0, 0, 3]
    np.testing.assert_array_equal(m.p2BC(), want)


def test_p2_label_rule_every_combination(O):
    """P1 labels chosen so the 19 edges of the unit cube meet I/I, O/O, W/W,
    Internal/Internal and every mixed pair."""
    m = _cube(O, 1, bc=False)
    m.setFluidVertexBCs(np.array([1, 1, 2, 2, 0, 0, 3, 3], np.uint8))
    # edges: 01 03 07 13 17 37 02 23 27 06 26 67 04 46 47 05 45 57 15
    want = [1, 3, 3, 3, 3, 3, 3, 2, 3, 3, 3, 3, 3, 3, 3, 3, 0, 3, 3]
    np.testing.assert_array_equal(m.p2BC(), want)


# ------------------------------------------------------------ face normals
_LOCAL_FACES = [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]


def test_face_normals_unit_cube_by_hand(O):
    m = _cube(O, 1)
    v, t, g = m.faces()
    key = {tuple(sorted(r)): i for i, r in enumerate(v)}
    # boundary face 137 of tet 0 = {0,1,3,7}, local face {1,2,3}: ac x ab of
    # a = x1, b = x3, c = x7 points into the cube and stays unflipped
    f = key[(1, 3, 7)]
    assert list(v[f]) == [1, 3, 7] and list(t[f]) == [0, -1, 0, -1]
    np.testing.assert_allclose(g[f, :3], [-1.0, 0.0, 0.0], atol=1e-15)
    assert abs(g[f, 3] - 0.5) < 1e-15
    np.testing.assert_allclose(g[f, 4:], [1.0, 2.0 / 3.0, 1.0 / 3.0], atol=1e-15)
    # internal face 037 of tet 0 (local face 1 -> [0, 7, 3]) and tet 1
    # (local face 2): ac x ab = (1, -1, 0)/sqrt2 points from tet 1's centroid
    # (1/2, 3/4, 1/4) towards tet 0's (3/4, 1/2, 1/4), so it is flipped and
    # vertids become [0, 3, 7]
    f = key[(0, 3, 7)]
    assert list(t[f]) == [0, 1, 1, 2]
    assert list(v[f]) == [0, 3, 7]
    s = 1.0 / np.sqrt(2.0)
    np.testing.assert_allclose(g[f, :3], [-s, s, 0.0], atol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_face_orientation_rules(O, n):
    m = O.OracleMesh.structuredRectangularPrism(1.0, 1.5, 2.0, n, n, n)
    X, T = m.vertices(), m.tets()
    v, t, g = m.faces()
    cen = X[T].mean(axis=1)
    nflip = 0
    for f in range(len(v)):
        ta, tb, fa, _ = t[f]
        lf = _LOCAL_FACES[fa]
        made = [T[ta, lf[0]], T[ta, lf[1]], T[ta, lf[2]]]
        a, b, c = X[v[f, 0]], X[v[f, 1]], X[v[f, 2]]
        nrm = np.cross(c - a, b - a)
        np.testing.assert_allclose(g[f, :3], nrm / np.linalg.norm(nrm), atol=1e-14)
        if tb < 0:
            # boundary faces keep the owner's local order, whatever way it faces
            assert list(v[f]) == made
        else:
            assert g[f, :3] @ (cen[tb] - cen[ta]) > 0
            swapped = [made[0], made[2], made[1]]
            assert list(v[f]) in (made, swapped)
            nflip += list(v[f]) == swapped
    assert nflip > 0


# ------------------------------------------------------------ step composition
def _dense(rp, col, val, n, m=None):
    A = np.zeros((n, m or n))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return A


def _solve(A, b):
    """Dense solve of a penalised system, Jacobi-scaled so that the 1e10 /
    1e12 penalty rows do not spoil the conditioning of the LU."""
    d = np.abs(np.diag(A))
    return np.linalg.solve(A / d[:, None], b / d[:, None] if b.ndim == 2 else b / d)


def _convection(T, E, G, V, u, nv):
    """computeAdvectionRHS (NavierStokes.cpp:526-584), written out per tet."""
    out = np.zeros_like(u)
    for t in range(len(T)):
        ue = 0.25 * u[T[t]].sum(axis=0)
        gu = np.einsum("ja,jb->ab", u[T[t]], G[t])  # sum_j u_j (x) grad phi_j
        c = gu @ ue
        out[T[t]] -= (V[t] / 4.0) * c
        out[nv + E[t]] -= (V[t] / 24.0) * c
    return out


def dense_piso_step(ops, mesh, u0, p0, uin, dt, outlet, p_out, relax_u, relax_p, ref,
                    mutate=None):
    """One PISO step (one corrector) with dense direct solves."""
    M, K, Gx, Gy, Gz, L, ilm = ops
    T, E, G, V, b1, b2 = mesh
    nv, nu = len(b1), len(b1) + len(b2)
    lab = np.r_[b1, b2]
    dirich = (lab == INLET) | (lab == WALL)
    pen = 1e10
    ubc = np.where((lab == INLET)[:, None], np.asarray(uin)[None, :], 0.0)
    conv = _convection(T, E, G, V, u0, nv)
    if mutate == "conv_sign":
        conv = -conv
    gp = np.stack([Gx @ p0, Gy @ p0, Gz @ p0], axis=1)
    rhs = (M @ u0) / dt
    if mutate == "grad_sign":
        rhs[:nv] += gp
    else:
        rhs[:nv] -= gp
    rhs += conv
    rhs[dirich] += pen * ubc[dirich]
    lhs = M / dt + K
    lhs[np.arange(nu)[dirich], np.arange(nu)[dirich]] += pen
    us = _solve(lhs, rhs)
    div = Gx @ us[:nv, 0] + Gy @ us[:nv, 1] + Gz @ us[:nv, 2]
    rp = -div / dt
    Lbc = L.copy()
    if outlet == 0:
        out = b1 == OUTLET
        Lbc[np.arange(nv)[out], np.arange(nv)[out]] += 1e12
        rp[out] += 1e12 * p_out
    else:
        Lbc[ref, ref] += 1e12
    pc = _solve(Lbc, rp)
    p = p0 + relax_p * pc
    if mutate == "normalise_first" and ref >= 0:
        p -= p[ref]
    if outlet == 0:
        p[b1 == OUTLET] = p_out
    if ref >= 0 and mutate != "normalise_first":
        p -= p[ref]
    g = np.stack([Gx @ pc, Gy @ pc, Gz @ pc], axis=1)
    u = us.copy()
    free1 = ~dirich[:nv] & (ilm[:nv] > 1e-14)
    u[:nv][free1] = us[:nv][free1] + relax_u * (-dt * ilm[:nv][free1, None] * g[free1])
    if mutate == "correct_p2":
        # a plausible slip: correcting edge nodes with their endpoints' mean gradient
        pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
        for t in range(len(T)):
            for k, (a, b) in enumerate(pairs):
                e = nv + E[t, k]
                if not dirich[e]:
                    gm = 0.5 * (g[T[t, a]] + g[T[t, b]])
                    u[e] = us[e] + relax_u * (-dt * ilm[e] * gm)
    u[dirich] = ubc[dirich]
    return u, p


def _case(O, n, outlet, with_ref, rng_seed=3):
    om = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    # density 1 and viscosity 1 keep convection (which carries no density),
    # the viscous term and the mass term of comparable size
    rho, mu, dt = 1.0, 1.0, 0.1
    uin = (0.5, 0.1, -0.05)
    o = O.OracleNS(om, rho, mu)
    ref = -1
    if with_ref:
        ref = int(np.nonzero(om.p1BC() == INTERNAL)[0][0])
        o.setReferencePressureNode(ref)
    o.setDt(dt)
    o.setMeanInletVelocity(uin)
    p_out = 0.3 if outlet == 0 else 0.0
    o.setOutlet(outlet, p_out)
    o.setRelax(0.8, 0.7)
    o.setSolver(1e-14, 1e-14, 20000, 0)
    o.setStopScaled(True)
    sz = o.sizes()
    nv, nu = sz["nv"], sz["nu"]
    X = np.vstack([om.vertices(), om.edgeNodes()])
    rng = np.random.default_rng(rng_seed)
    # a swirling field with O(1) gradients, so convection is far from zero
    u0 = np.stack([0.5 + 0.4 * np.sin(3 * X[:, 1]) * X[:, 2], 0.6 * X[:, 0] * X[:, 2],
                   -0.5 * X[:, 1] * X[:, 0]], axis=1) + 0.05 * rng.standard_normal((nu, 3))
    p0 = 20.0 * rng.standard_normal(nv)
    o.setVelocity(u0)
    o.setPressure(p0)
    ops = [_dense(*o.getMatrix(k), nu) for k in ("M", "K")]
    ops += [_dense(*o.getMatrix(k), nv) for k in ("Gx", "Gy", "Gz", "L")]
    ops.append(o.getInvLumpedMass())
    mesh = (om.tets(), om.tetEdgeNodes(), om.gradients(), om.volumes(), om.p1BC(), om.p2BC())
    args = (ops, mesh, u0, p0, uin, dt, outlet, p_out, 0.8, 0.7, ref)
    return o, args


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("n,outlet,with_ref", [(2, 0, False), (3, 0, True), (3, 1, True)])
def test_piso_step_matches_dense_restatement(O, n, outlet, with_ref):
    """Dirichlet outlet, Dirichlet outlet with a reference node (the outlet
    reset comes before the normalisation), Neumann outlet with the reference
    node pinned by the 1e12 penalty."""
    o, args = _case(O, n, outlet, with_ref)
    u_want, p_want = dense_piso_step(*args)
    assert o.pisoStep() == 0
    u, p = o.getVelocity(), o.getPressure()
    eu, ep = _rel(u, u_want), _rel(p, p_want)
    assert eu < 1e-9 and ep < 1e-9, (eu, ep)
    # the check has teeth: each plausible slip moves the step far beyond it
    muts = ("conv_sign", "grad_sign", "correct_p2")
    if outlet == 0 and with_ref:
        muts += ("normalise_first",)
    for mut in muts:
        um, pm = dense_piso_step(*args, mutate=mut)
        assert max(_rel(u, um), _rel(p, pm)) > 1e-5, mut


def test_uniform_flow_steps(O):
    """All walls Inlet at U, outlet vertices at x = 1 with p_out = 0.7, u0 = U,
    p0 = 0.  Convection vanishes (grad u = 0) and K U = 0, so u* = U and
    div u* = 0.  The outlet penalty fixes the correction itself, so p' = 0.7
    everywhere (harmonic with boundary value 0.7), G p' = 0 and u = U.  Step 1:
    p = 0 + 0.7 = 0.7 everywhere.  Step 2: p' = 0.7 again, p = 1.4, then the
    outlet reset puts 0.7 back on the outlet vertices only."""
    om = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, 3, 3, 3)
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    lab = om.p1BC().copy()
    lab[lab == WALL] = INLET
    om.setFluidVertexBCs(lab)
    out = lab == OUTLET
    U = np.array([0.4, -0.2, 0.1])
    o = O.OracleNS(om, 1060.0, 0.004, U)
    o.setDt(0.05)
    o.setMeanInletVelocity(U)
    o.setOutlet(0, 0.7)
    o.setSolver(1e-13, 1e-13, 20000, 0)
    o.setStopScaled(True)
    nu, nv = o.sizes()["nu"], o.sizes()["nv"]
    o.setVelocity(np.tile(U, (nu, 1)))
    o.setPressure(np.zeros(nv))
    want_p = [np.full(nv, 0.7), np.where(out, 0.7, 1.4)]
    for step in range(2):
        assert o.pisoStep() == 0
        np.testing.assert_allclose(o.getVelocity(), np.tile(U, (nu, 1)), rtol=0, atol=1e-10)
        np.testing.assert_allclose(o.getPressureCorrection(), 0.7, rtol=0, atol=1e-9)
        np.testing.assert_allclose(o.getPressure(), want_p[step], rtol=0, atol=1e-9)
