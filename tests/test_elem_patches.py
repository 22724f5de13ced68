"""Element patches of the momentum operator (elem_spmv), host set-up on the CPU.

* structure: every velocity node is owned by exactly one patch; a patch's node
  table starts with its owned rows; every tet touching an owned row is a visit
  of that patch, and only those; the local ids of a visit name the tet's 10
  nodes; the LDS slots of a patch's owned (row, tet) pairs are a permutation
  of 0 .. slots-1, lie in the row's range of the slot starts and run in
  ascending tet order per row; the LDS limits of the kernel hold;
* the two phases of the device kernel restated in FP64 NumPy (per-visit
  element products written to slots, then per-row sums of the slots in order,
  plus the Inlet / Wall penalty) reproduce the oracle's assembled momentum
  operator lhs = M/dt + K + penalties (NavierStokes.cpp:135-310, 460-495) to
  rounding, on the Kuhn cube and on the cylinder mesh.
"""
import numpy as np
import pytest

RHO, MU, DT = 1060.0, 0.004, 1e-3
EV = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
SLOTS, NODES, ROWS = 4096, 2048, 1024


def _mesh(P, kind):
    if kind == "cyl":
        return P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
    n = int(kind[4:])
    return P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)


@pytest.fixture(scope="module", params=["kuhn6", "kuhn13", "cyl"])
def case(request, P):
    pm = _mesh(P, request.param)
    pm.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    X, tets, ten = pm.vertices(), pm.tets(), pm.tetEdgeNodes()
    nv = len(X)
    E = P.HostElemPatches(nv, tets, ten, X)
    return request.param, pm, E


def test_patch_structure(case):
    kind, pm, E = case
    tets, ten = pm.tets(), pm.tetEdgeNodes()
    nv = len(pm.vertices())
    nodes10 = np.hstack([tets, nv + ten]).astype(np.int64)
    nu = nv + len(pm.edgeNodes())
    assert np.array_equal(np.sort(E.rows), np.arange(nu))
    assert E.max_rows <= ROWS and E.max_nodes <= NODES and E.max_slots <= SLOTS
    inc = np.bincount(nodes10.ravel(), minlength=nu)
    lid, sl = E.local_ids(), E.slots()
    pos = np.empty(nu, np.int64)
    pos[E.rows] = np.arange(nu)
    assert np.array_equal(np.sort(E.tet_order), np.arange(len(tets)))
    for p in range(E.npatch):
        table = E.nodes[E.node_off[p]:E.node_off[p + 1]]
        owned = E.rows[E.row_off[p]:E.row_off[p + 1]]
        nr = len(owned)
        assert np.array_equal(table[:nr], owned) and len(np.unique(table)) == len(table)
        vs = slice(E.visit_off[p], E.visit_off[p + 1])
        vt = E.vtet[vs]
        assert np.array_equal(E.tets()[vs], vt)
        # the visits are exactly the tets touching an owned row, in the patch
        # numbering of the tets; the non-owned nodes follow the patch order
        want = np.nonzero(np.isin(nodes10, owned).any(1))[0]
        assert np.array_equal(np.sort(vt), want)
        assert np.all(np.diff(E.tet_numbers()[vs]) > 0)
        assert np.all(np.diff(pos[table[nr:]]) > 0)
        assert np.array_equal(table[lid[vs]], nodes10[vt])
        st = E.starts[E.row_off[p] + p:E.row_off[p + 1] + p + 1].astype(np.int64)
        assert st[0] == 0 and np.array_equal(np.diff(st), inc[owned])
        own = lid[vs] < nr
        assert np.all(sl[vs][~own] == 0)
        s = sl[vs][own]
        assert np.array_equal(np.sort(s), np.arange(st[-1]))
        r = lid[vs][own]
        assert np.all((s >= st[r]) & (s < st[r + 1]))
        # per owned row, the slots follow ascending tet order
        row_of = lid[vs][own]
        tet_of = np.broadcast_to(vt[:, None], own.shape)[own]
        order = np.lexsort((tet_of, row_of))
        assert np.array_equal(s[order], np.arange(st[-1]))
        if p > 3 and kind.startswith("kuhn13"):
            break


def test_two_phase_product_matches_assembled_operator(case, O):
    kind, pm, E = case
    om = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    o = O.OracleNS(om, RHO, MU)
    o.setDt(DT)
    rp, col, val = o.getMatrix("lhs")
    nv = len(pm.vertices())
    nu = len(rp) - 1
    tets, ten = om.tets(), om.tetEdgeNodes()
    nodes10 = np.hstack([tets, nv + ten]).astype(np.int64)
    V, G = om.volumes(), om.gradients().reshape(-1, 4, 3)
    bc = np.concatenate([om.p1BC(), om.p2BC()])
    x = np.random.default_rng(3).standard_normal((nu, 3))
    lid, sl = E.local_ids(), E.slots()
    y = np.zeros((nu, 3))
    for p in range(E.npatch):
        table = E.nodes[E.node_off[p]:E.node_off[p + 1]]
        owned = E.rows[E.row_off[p]:E.row_off[p + 1]]
        nr = len(owned)
        vs = slice(E.visit_off[p], E.visit_off[p + 1])
        t = E.vtet[vs]
        xe = x[table[lid[vs]]]                      # visits x 10 x 3
        a = np.sqrt(V[t] / MU)[:, None, None] * G[t]  # visits x 4 x 3
        m = (RHO * V[t] / DT)[:, None]
        sv, se = xe[:, :4].sum(1), xe[:, 4:].sum(1)
        z = xe[:, :4].copy()
        for k, (i, j) in enumerate(EV):
            z[:, i] += xe[:, 4 + k]
            z[:, j] += xe[:, 4 + k]
        w = np.einsum("vak,vac->vkc", a, z)
        sa = np.einsum("vak,vkc->vac", a, w)
        ye = np.empty_like(xe)
        ye[:, :4] = m[:, None] * xe[:, :4] / 20 + (m * (sv / 20 + se / 30))[:, None] + sa
        for k, (i, j) in enumerate(EV):
            ye[:, 4 + k] = m * xe[:, 4 + k] / 20 + m * (se / 60 + sv / 30) + sa[:, i] + sa[:, j]
        st = E.starts[E.row_off[p] + p:E.row_off[p + 1] + p + 1].astype(np.int64)
        cnt = np.zeros((st[-1], 3))
        own = lid[vs] < nr
        cnt[sl[vs][own]] = ye[own]
        for i in range(nr):
            y[owned[i]] = cnt[st[i]:st[i + 1]].sum(0)
    pen = (bc == 0) | (bc == 1)
    y[pen] += 1e10 * x[pen]
    yc = O.csr_spmv(rp, col, val, x)
    err = np.abs(y - yc).max() / np.abs(yc).max()
    print(f"{kind}: {E.npatch} patches, max rows/nodes/slots {E.max_rows}/{E.max_nodes}/"
          f"{E.max_slots}; visits per tet {len(E.vtet) / len(tets):.2f}; rel err {err:.1e}")
    assert err <= 1e-14
    inner = ~pen
    assert np.abs(y - yc)[inner].max() <= 1e-12 * np.abs(yc[inner]).max()


def test_bad_input(P):
    with pytest.raises(Exception):
        P.HostElemPatches(4, np.array([[0, 1, 2, 7]]), np.zeros((1, 6)), np.zeros((4, 3)))
