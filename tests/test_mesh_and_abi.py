"""Host-side checks that need no GPU: the product mesh layer against the
oracle (bit-exact numbering, labels, geometry), the structured-mesh counts,
and that libcfsa_ns.so exports every entry point include/cfsa_ns.h declares."""
import re
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]


def _pair(P, O, kind, n=2):
    if kind == "kuhn":
        pm = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
        om = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
    elif kind == "box":
        pm = P.Mesh.structuredRectangularPrism(2.0, 1.0, 0.5, 2 * n, n, n)
        om = O.OracleMesh.structuredRectangularPrism(2.0, 1.0, 0.5, 2 * n, n, n)
    else:
        pm = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
        om = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    return pm, om


# kuhn 16 (24,576 tets) is large enough for the multi-threaded set-up loops and
# sorts; the smaller cases run them on one thread
@pytest.mark.parametrize("kind,n", [("kuhn", 1), ("kuhn", 2), ("kuhn", 3), ("kuhn", 5),
                                    ("kuhn", 16), ("box", 2), ("cyl", 0)])
def test_mesh_layer_bit_exact(P, O, kind, n):
    pm, om = _pair(P, O, kind, n)
    for d, cov in [((1.0, 0.0, 0.0), 1.0), ((-1.0, 0.0, 0.0), 2.5)]:
        pm.setupBoundaryConditions(d, cov)
        om.setupBoundaryConditions(d, cov)
        for name in ["vertices", "tets", "volumes", "gradients", "tetEdgeNodes", "edgeNodes",
                     "p1BC", "p2BC"]:
            a, b = getattr(pm, name)(), getattr(om, name)()
            assert a.shape == b.shape and np.array_equal(a, b), name
        for a, b in zip(pm.faces(), om.faces()):
            assert np.array_equal(a, b)


def test_vertex_update_matches_oracle(P, O):
    """Moving the vertices (cfsa_mesh_update_vertices, multi-threaded geometry)
    gives the oracle's volumes, gradients and face geometry after the same move
    (Mesh.hpp:98-101), and a moved mesh's volumes and gradients equal those of a
    mesh built from the moved coordinates (face centres differ in rounding:
    the update sums the stored, possibly flipped, vertex order)."""
    pm = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, 16, 16, 16)
    om = O.OracleMesh.structuredRectangularPrism(1.0, 1.0, 1.0, 16, 16, 16)
    x = pm.vertices()
    rng = np.random.default_rng(3)
    moved = x + 1e-3 * np.sin(3.0 * x[:, [1, 2, 0]]) + 1e-5 * rng.standard_normal(x.shape)
    pm.updateVertices(moved)
    om.updateVertices(moved)
    for name in ["vertices", "volumes", "gradients"]:
        assert np.array_equal(getattr(pm, name)(), getattr(om, name)()), name
    for a, b in zip(pm.faces(), om.faces()):
        assert np.array_equal(a, b)
    fresh = O.OracleMesh.fromArrays(moved, pm.tets())
    assert np.array_equal(pm.volumes(), fresh.volumes())
    assert np.array_equal(pm.gradients(), fresh.gradients())


@pytest.mark.parametrize("n", [3, 4, 6, 8])
def test_structured_counts_cubic_fit(P, n):
    """Exact counts of the Kuhn prism; at n = 96 these give 912,673 vertices,
    5,308,416 tets, 6,276,384 edge nodes, 204,763,393 / 13,465,441 nnz."""
    m = P.Mesh.structuredRectangularPrism(1, 1, 1, n, n, n)
    nv, nt, ne, nf, nu = m.sizes()
    assert (nv, nt, ne) == ((n + 1) ** 3, 6 * n ** 3, 7 * n ** 3 + 9 * n ** 2 + 3 * n)
    # nnz of the 10-node clique and 4-vertex clique patterns from the edges
    ten, T = m.tetEdgeNodes(), m.tets()
    pairs2 = set()
    pairs1 = set()
    for t in range(nt):
        nodes = list(T[t]) + [nv + e for e in ten[t]]
        for a in nodes:
            for b in nodes:
                pairs2.add((a, b))
        for a in T[t]:
            for b in T[t]:
                pairs1.add((a, b))
    assert len(pairs2) == 230 * n ** 3 + 138 * n ** 2 + 24 * n + 1
    assert len(pairs1) == 15 * n ** 3 + 21 * n ** 2 + 9 * n + 1


def test_cylinder_mesher_is_valid(P):
    m = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
    nv, nt, ne, nf, nu = m.sizes()
    assert nt == 22500 and nv == 4641
    V = m.volumes()
    assert V.min() > 0
    assert abs(V.sum() - np.pi * 10.0) / (np.pi * 10.0) < 0.05  # inscribed polygon
    v, t, g = m.faces()
    bnd = t[:, 1] == -1
    # closed surface: every boundary edge shared by exactly two boundary faces
    edges = {}
    for f in v[bnd]:
        for a, b in ((0, 1), (1, 2), (2, 0)):
            k = tuple(sorted((f[a], f[b])))
            edges[k] = edges.get(k, 0) + 1
    assert set(edges.values()) == {2}
    assert nv - ne + nf - nt == 1


def test_msh_reader_roundtrip(P, tmp_path):
    m = P.Mesh.structuredRectangularPrism(1, 1, 1, 2, 2, 2)
    X, T = m.vertices(), m.tets()
    path = tmp_path / "cube.msh"
    with open(path, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n")
        f.write(f"1 {len(X)} 1 {len(X)}\n3 1 0 {len(X)}\n")
        for i in range(len(X)):
            f.write(f"{i + 1}\n")
        for x in X:
            f.write(" ".join(repr(float(c)) for c in x) + "\n")
        f.write("$EndNodes\n$Elements\n")
        f.write(f"2 {len(T) + 1} 1 {len(T) + 1}\n2 1 2 1\n1 1 2 3\n3 1 4 {len(T)}\n")
        for i, t in enumerate(T):
            f.write(f"{i + 2} {t[0] + 1} {t[1] + 1} {t[2] + 1} {t[3] + 1}\n")
        f.write("$EndElements\n")
    m2 = P.Mesh.fromMshFile(path)
    assert np.array_equal(m2.tets(), T) and np.array_equal(m2.vertices(), X)
    with pytest.raises(P.CfsaError):
        P.Mesh.fromMshFile(tmp_path / "missing.msh")


def test_mesh_vtk_format(P, tmp_path):
    m = P.Mesh.structuredRectangularPrism(1, 1, 1, 1, 1, 1)
    m.setupBoundaryConditions()
    path = tmp_path / "m.vtk"
    m.exportVtk(path, scale=2.0)
    lines = path.read_text().splitlines()
    assert lines[:5] == ["# vtk DataFile Version 3.0", "Shape export", "ASCII",
                         "DATASET UNSTRUCTURED_GRID", "POINTS 8 double"]
    assert lines[5] == "0 0 0" and lines[6] == "2 0 0"
    i = lines.index("CELLS 6 30")
    assert lines[i + 1] == "4 0 1 3 7"
    assert lines[i + 7] == "CELL_TYPES 6" and lines[i + 8] == "10"
    j = lines.index("POINT_DATA 8")
    assert lines[j - 1] == "" and lines[j + 1] == "SCALARS FluidBoundaryType int 1"
    assert [int(x) for x in lines[j + 3:j + 11]] == m.p1BC().tolist()


def test_library_exports_every_declared_symbol(P):
    header = (REPO / "include" / "cfsa_ns.h").read_text()
    names = set(re.findall(r"\b(cfsa_[a-z0-9_]+)\s*\(", header))
    lib = P.library()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.cfsa_abi_version() == 4


def test_abi_rejects_bad_arguments(P):
    import ctypes as C
    lib = P.library()
    assert lib.cfsa_mesh_structured_prism(1.0, 1.0, 1.0, 0, 1, 1, C.byref(C.c_void_p())) == 5
    assert "nx, ny, nz" in P.last_error()
    assert lib.cfsa_mesh_sizes(None, None) == 5
    with pytest.raises(P.CfsaError):
        P.Mesh.fromArrays(np.zeros((3, 3)), np.array([[0, 1, 2, 3]], np.int32))
