"""GPU parity: every PISO building block and whole steps of the HIP library
against the CPU oracle on the same mesh and inputs.

Bars: integer work (patterns, colouring, incidence) bit-exact; kernels whose
floating-point order reproduces the reference (row-gather assembly, SpMV,
convection, divergence) bit-exact; the optional coloured scatter assembly
within 1e-13 relative (it sums element contributions in colour order instead
of tet order); Krylov solves and whole steps: equal iteration counts and
velocity / pressure relative L2 error <= 1e-8 per step (the FP64 tolerance of
the step parity requirement).  Inner products are order-robust (Dot2) on both
sides, so in practice the FP64 iterates agree bit for bit; the tests report
the largest difference seen.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BLOOD = dict(rho=1060.0, mu=0.004)


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def make_case(P, O, kind, n=4, rho=1060.0, mu=0.004, dt=0.01, uin=(0.5, 0.0, 0.0),
              ncorr=1, outlet=0, p_out=0.0, relax=(1.0, 1.0), pressure_solver=0,
              momentum_tol=1e-6, pressure_tol=1e-8, inlet_before_init=False, momentum_solver=0):
    if kind == "kuhn":
        pm = P.Mesh.structuredRectangularPrism(1.0, 1.0, 1.0, n, n, n)
    elif kind == "box":
        pm = P.Mesh.structuredRectangularPrism(2.0, 1.0, 1.0, 2 * n, n, n)
    else:
        pm = P.Mesh.cylinder(10.0, 1.0, 16, 0.1)
    om = O.OracleMesh.fromArrays(pm.vertices(), pm.tets())
    pm.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    om.setupBoundaryConditions((1.0, 0.0, 0.0), 1.0)
    g = P.NavierStokesSolver()
    inlet0 = uin if inlet_before_init else (0.0, 0.0, 0.0)
    if inlet_before_init:
        g.setMeanInletVelocity(uin)
    g.setOutletType(outlet)
    g.initialize(pm, rho, mu)
    g.setDt(dt)
    g.setMeanInletVelocity(uin)
    g.setOutletPressure(p_out)
    g.setNumberOfCorrections(ncorr)
    g.setRelaxU(relax[0])
    g.setRelaxP(relax[1])
    g.setOptions(momentum_tol=momentum_tol, pressure_tol=pressure_tol,
                 pressure_solver=pressure_solver, momentum_solver=momentum_solver)
    o = O.OracleNS(om, rho, mu, inlet0) if outlet == 0 else None
    if o is None:
        # the oracle picks the Neumann reference node at initialize, like the reference
        o = _oracle_neumann(O, om, rho, mu, inlet0)
    o.setDt(dt)
    o.setMeanInletVelocity(uin)
    o.setOutlet(outlet, p_out)
    o.setNumberOfCorrections(ncorr)
    o.setRelax(*relax)
    o.setSolver(momentum_tol, pressure_tol, 1000, pressure_solver)
    o.setMomentumSolver(momentum_solver)
    return g, o, pm, om


def _oracle_neumann(O, om, rho, mu, inlet0):
    o = O.OracleNS(om, rho, mu, inlet0)
    lab = om.p1BC()
    ref = int(np.nonzero(lab == 3)[0][0])
    o.setReferencePressureNode(ref)
    return o


@pytest.fixture(scope="module")
def kuhn4(P, O):
    return make_case(P, O, "kuhn", 4)


@pytest.fixture(scope="module")
def cyl(P, O):
    return make_case(P, O, "cyl")


# ------------------------------------------------------------------ operators
@pytest.mark.parametrize("case", ["kuhn4", "cyl"])
def test_patterns_values_colouring(P, O, case, request):
    g, o, pm, om = request.getfixturevalue(case)
    for name in ["M", "K", "lhs", "L", "Lbc", "Gx", "Gy", "Gz"]:
        rg, cg, vg = g.getMatrix(name)
        ro, co, vo = o.getMatrix(name)
        assert np.array_equal(rg, ro) and np.array_equal(cg, co), name
        assert np.array_equal(vg, vo), name
    assert np.array_equal(g.getInvLumpedMass(), o.getInvLumpedMass())
    col, nc = O.colour_tets(pm.tets(), pm.nVertices())
    assert np.array_equal(g.colouring(), col) and g.numColours() == nc
    ptr, tet = g.incidence()
    T = np.c_[pm.tets(), pm.nVertices() + pm.tetEdgeNodes()]
    for node in np.random.default_rng(0).integers(0, len(ptr) - 1, 50):
        want = np.nonzero((T == node).any(axis=1))[0]
        assert np.array_equal(tet[ptr[node]:ptr[node + 1]], want)


@pytest.mark.parametrize("case", ["kuhn4", "cyl"])
def test_coloured_assembly_close(P, O, case, request):
    g, o, pm, om = request.getfixturevalue(case)
    g.setOptions(assembly=1)
    try:
        g.build_sparse_matrices()
        for name in ["M", "K", "L", "Gx", "Gy", "Gz"]:
            vg = g.getMatrix(name)[2]
            vo = o.getMatrix(name)[2]
            np.testing.assert_allclose(vg, vo, rtol=1e-13, atol=1e-15 * np.abs(vo).max(),
                                       err_msg=name)
    finally:
        g.setOptions(assembly=0)
        g.build_sparse_matrices()
    assert np.array_equal(g.getMatrix("K")[2], o.getMatrix("K")[2])


@pytest.mark.parametrize("case", ["kuhn4", "cyl"])
def test_row_gather_forms_bit_identical(P, O, case, request):
    """The default row gather (six rows per wavefront, LDS accumulators) and its
    fallback (one thread per row, assembly = 2) write the same bits, also for
    the lumped-mass inverse, and match the oracle."""
    g, o, pm, om = request.getfixturevalue(case)
    names = ["M", "K", "L", "Gx", "Gy", "Gz"]
    try:
        g.setOptions(assembly=2)
        g.build_sparse_matrices()
        thread = {n: g.getMatrix(n)[2] for n in names}
        ilm2 = g.getInvLumpedMass()
    finally:
        g.setOptions(assembly=0)
        g.build_sparse_matrices()
    for n in names:
        assert np.array_equal(thread[n], g.getMatrix(n)[2]), n
        assert np.array_equal(thread[n], o.getMatrix(n)[2]), n
    assert np.array_equal(ilm2, g.getInvLumpedMass())
    assert np.array_equal(ilm2, o.getInvLumpedMass())
    with pytest.raises(Exception):
        g.setOptions(assembly=3)


@pytest.mark.parametrize("case", ["kuhn4", "cyl"])
@pytest.mark.parametrize("layout", [1, 2])
def test_spmv_bit_exact(P, O, case, layout, request):
    """Both Krylov SpMV layouts (CSR-stream, sliced-ELL) and the CSR-stream
    products of M, K, L, G reproduce the sequential CSR sums bit for bit."""
    g, o, pm, om = request.getfixturevalue(case)
    g.setOptions(spmv_layout=layout)
    rng = np.random.default_rng(0)
    sz = g.sizes()
    for name, nrhs in [("lhs", 3), ("lhs", 1), ("M", 3), ("K", 1), ("L", 1), ("Lbc", 1)]:
        n = sz["nu"] if name in ("M", "K", "lhs") else sz["nv"]
        x = rng.uniform(-1, 1, (n, 3) if nrhs == 3 else n)
        rp, col, val = g.getMatrix(name)
        assert np.array_equal(g.applyMatrix(name, x, nrhs), O.csr_spmv(rp, col, val, x)), name
    g.setOptions(spmv_layout=0)
    p = rng.uniform(-1, 1, sz["nv"])
    y = g.applyMatrix("Gx", p)
    for c, name in enumerate(["Gx", "Gy", "Gz"]):
        rp, col, val = g.getMatrix(name)
        assert np.array_equal(y[:, c], O.csr_spmv(rp, col, val, p)), name


@pytest.mark.parametrize("case", ["kuhn4", "cyl"])
def test_advection_divergence_bit_exact(P, O, case, request):
    g, o, pm, om = request.getfixturevalue(case)
    u = np.random.default_rng(1).uniform(-1, 1, (g.sizes()["nu"], 3))
    g.setVelocity(u)
    o.setVelocity(u)
    assert np.array_equal(g.computeAdvectionRHS(), o.computeAdvectionRHS())
    assert np.array_equal(g.computeDivergence(u), o.computeDivergence(u))


@pytest.mark.parametrize("case", ["kuhn4", "cyl"])
@pytest.mark.parametrize("layout", [1, 2])
def test_krylov_parity(P, O, case, layout, request):
    g, o, pm, om = request.getfixturevalue(case)
    g.setOptions(spmv_layout=layout)
    sz = g.sizes()
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, (sz["nu"], 3))
    x, info = g.krylovSolve("lhs", b, method=0, tol=1e-10)
    for c in range(3):
        xo, io = o.krylovSolve("lhs", b[:, c], method=0, tol=1e-10)
        assert info["iters"][c] == io["iters"], (c, info, io)
        assert relerr(x[:, c], xo) <= 1e-12, (c, relerr(x[:, c], xo))
    bp = rng.uniform(-1, 1, sz["nv"])
    for method in (0, 1):
        xp, ip = g.krylovSolve("Lbc", bp, method=method, tol=1e-9)
        xo, io = o.krylovSolve("Lbc", bp, method=method, tol=1e-9)
        assert ip["iters"][0] == io["iters"], (method, ip, io)
        assert relerr(xp, xo) <= 1e-12, (method, relerr(xp, xo))
    x0, i0 = g.krylovSolve("Lbc", np.zeros(sz["nv"]))
    assert i0["iters"][0] == 0 and not x0.any()
    # batched Jacobi CG on the (SPD) momentum matrix, 3 right-hand sides
    x, info = g.krylovSolve("lhs", b, method=1, tol=1e-10)
    for c in range(3):
        xo, io = o.krylovSolve("lhs", b[:, c], method=1, tol=1e-10)
        assert info["iters"][c] == io["iters"], (c, info, io)
        assert relerr(x[:, c], xo) <= 1e-12, (c, relerr(x[:, c], xo))
    g.setOptions(spmv_layout=0)


def test_layout_switch_after_mixed_buffers(P, O):
    """FP32 operator copies made while the momentum operator is in CSR order
    are resized when a later call switches to the (padded) sliced-ELL layout,
    for FP64 and FP32 SpMV and mixed solves alike (the order a benchmark
    that times both layouts takes)."""
    g, o, pm, om = make_case(P, O, "kuhn", 4)
    sz = g.sizes()
    b = np.random.default_rng(4).uniform(-1, 1, (sz["nu"], 3))
    g.setOptions(spmv_layout=1)
    g.timeMomentumSpmv(2, 1)
    x1, i1 = g.krylovSolve("lhs", b, tol=1e-10, precision=1)
    g.setOptions(spmv_layout=2)
    g.timeMomentumSpmv(2, 0)
    g.timeMomentumSpmv(2, 1)
    x2, i2 = g.krylovSolve("lhs", b, tol=1e-10, precision=1)
    assert np.array_equal(x1, x2) and i1["iters"] == i2["iters"]
    g.setOptions(spmv_layout=0)
    assert np.array_equal(g.applyMatrix("lhs", b, 3), O.csr_spmv(*g.getMatrix("lhs"), b))


@pytest.mark.parametrize("scale", [1e27, 1e29])
def test_bicgstab_restart_branch_parity(P, O, kuhn4, scale):
    """Eigen's BiCGSTAB restart branch (|rho| < eps^2 |r0|^2: r = b - A x,
    r0 = r, the iteration count reset on the first restart) with the
    threshold raised on both sides so that a small solve passes through it:
    the host-side restart of the batched 3-rhs and the 1-rhs device solves
    takes the oracle's iterations and restarts and returns its solution."""
    g, o, pm, om = kuhn4
    sz = g.sizes()
    rng = np.random.default_rng(11)
    b3 = rng.uniform(-1, 1, (sz["nu"], 3))
    bp = rng.uniform(-1, 1, sz["nv"])
    g.setOptions(restart_scale=scale)
    O.set_restart_scale(scale)
    try:
        total = 0
        x, info = g.krylovSolve("lhs", b3, method=0, tol=1e-10)
        for c in range(3):
            xo, io = o.krylovSolve("lhs", b3[:, c], method=0, tol=1e-10)
            rs = O.last_restarts()
            assert (info["iters"][c], info["restarts"][c]) == (io["iters"], rs), (c, info, io, rs)
            assert relerr(x[:, c], xo) <= 1e-12, (c, relerr(x[:, c], xo))
            total += rs
        xp, ip = g.krylovSolve("Lbc", bp, method=0, tol=1e-9)
        xo, io = o.krylovSolve("Lbc", bp, method=0, tol=1e-9)
        rs = O.last_restarts()
        assert (ip["iters"][0], ip["restarts"][0]) == (io["iters"], rs), (ip, io, rs)
        assert relerr(xp, xo) <= 1e-12
        total += rs
        print(f"restart scale {scale:g}: restarts lhs {info['restarts']} Lbc {rs}")
        assert total >= 2
    finally:
        g.setOptions(restart_scale=1.0)
        O.set_restart_scale(1.0)


# ------------------------------------------------------------------ steps
STEP_CASES = {
    "kuhn4": dict(kind="kuhn", n=4, steps=5),
    "kuhn6_re100": dict(kind="kuhn", n=6, rho=1.0, mu=0.01, dt=1e-3, uin=(1.0, 0, 0), steps=3),
    "box3_ncorr2_relax": dict(kind="box", n=3, ncorr=2, relax=(0.5, 0.5), steps=4),
    "kuhn4_neumann": dict(kind="kuhn", n=4, outlet=1, steps=3),
    "kuhn4_cg_pout": dict(kind="kuhn", n=4, pressure_solver=1, p_out=3.0, steps=3),
    "kuhn4_inlet_first": dict(kind="kuhn", n=4, inlet_before_init=True, steps=3),
    "cyl_config1": dict(kind="cyl", steps=20, pressure_solver=1),
    "cyl_bicgstab": dict(kind="cyl", steps=5),
    "kuhn5_momentum_cg": dict(kind="kuhn", n=5, momentum_solver=1, pressure_solver=1, steps=3),
}


@pytest.mark.parametrize("name", list(STEP_CASES))
def test_piso_step_parity(P, O, name):
    spec = dict(STEP_CASES[name])
    steps = spec.pop("steps")
    g, o, pm, om = make_case(P, O, **spec)
    for s in range(steps):
        assert g.pisoStepStatus() == 0, (name, s, P.last_error())
        assert o.pisoStep() == 0, (name, s)
        sg, so = g.stepStats(), o.stats()
        assert sg["momentum_iters"] == so["momentum_iters"], (name, s, sg, so)
        nc = spec.get("ncorr", 1)
        assert sg["pressure_iters"] == so["pressure_iters"][:nc], (name, s, sg, so)
        eu = relerr(g.getVelocity(), o.getVelocity())
        ep = relerr(g.getPressure(), o.getPressure())
        assert eu <= 1e-8 and ep <= 1e-8, (name, s, eu, ep)
    assert abs(g.getTime() - steps * spec.get("dt", 0.01)) < 1e-12


@pytest.mark.parametrize("ncorr,outlet", [(2, 0), (3, 0), (2, 1)])
def test_corrector_reuse_bit_identical(P, O, ncorr, outlet):
    """With several correctors, correctors after the first reuse the first one's
    pressure correction (reuse_pressure_correction = 1, the default): every input
    of the reference's re-solve is unchanged, so re-solving (option 0) and the
    oracle, which re-solves, give the same bits and iteration counts, with one
    pressure solve per step instead of ncorr."""
    out = []
    for reuse in (1, 0):
        g, o, pm, om = make_case(P, O, "box", 3, ncorr=ncorr, relax=(0.5, 0.7),
                                 pressure_solver=1, p_out=1.5, outlet=outlet)
        g.setOptions(reuse_pressure_correction=reuse)
        for s in range(3):
            assert g.pisoStepStatus() == 0, P.last_error()
            st = g.stepStats()
            assert st["pressure_solves"] == (1 if reuse else ncorr)
            assert len(st["pressure_iters"]) == ncorr
            assert o.pisoStep() == 0
            assert st["pressure_iters"] == o.stats()["pressure_iters"][:ncorr], (s, st)
        out.append((g.getVelocity(), g.getPressure()))
        assert relerr(out[-1][0], o.getVelocity()) <= 1e-8
        assert relerr(out[-1][1], o.getPressure()) <= 1e-8
        with pytest.raises(Exception):
            g.setOptions(reuse_pressure_correction=2)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_corrector_reuse_mixed_pressure(P, O):
    """Reuse with FP32 inner pressure solves (precision 1): the same bits as
    re-solving, one pressure solve per step."""
    out = []
    for reuse in (1, 0):
        g, o, pm, om = make_case(P, O, "kuhn", 4, ncorr=2, pressure_solver=1)
        g.setOptions(precision=1, reuse_pressure_correction=reuse, stop_rule=1,
                     momentum_tol=1e-10, pressure_tol=1e-10)
        for _ in range(3):
            assert g.pisoStepStatus() == 0, P.last_error()
            assert g.stepStats()["pressure_solves"] == (1 if reuse else 2)
        out.append((g.getVelocity(), g.getPressure()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_simple_step_parity(P, O):
    g, o, pm, om = make_case(P, O, "kuhn", 4, p_out=2.0)
    for _ in range(3):
        assert g.simpleStep() and o.simpleStep() == 0
        assert relerr(g.getVelocity(), o.getVelocity()) <= 1e-8
        assert relerr(g.getPressure(), o.getPressure()) <= 1e-8


def test_inlet_profile_parity(P, O):
    g, o, pm, om = make_case(P, O, "kuhn", 4)
    X = np.vstack([pm.vertices(), pm.edgeNodes()])
    prof = np.zeros((len(X), 3))
    prof[:, 0] = 16 * X[:, 1] * (1 - X[:, 1]) * X[:, 2] * (1 - X[:, 2]) * 0.5
    g.setInletProfile(prof)
    o.setInletProfile(prof)
    for _ in range(3):
        assert g.pisoStep() and o.pisoStep() == 0
        assert relerr(g.getVelocity(), o.getVelocity()) <= 1e-8
    lab = np.r_[pm.p1BC(), pm.p2BC()]
    np.testing.assert_array_equal(g.getVelocity()[lab == 1], prof[lab == 1])


@pytest.mark.parametrize("msolver,precision", [(0, 0), (1, 2), (0, 1)])
def test_repeat_runs_bitwise_identical(P, O, msolver, precision):
    """Repeated runs, and graph replay against eager launches, give the same
    bits (each step replays its captured chunks many times, with eager work
    and host synchronisation in between)."""
    out = []
    for graphs in (1, 1, 0):
        g, o, pm, om = make_case(P, O, "kuhn", 4, momentum_solver=msolver)
        g.setOptions(graphs=graphs, precision=precision)
        for _ in range(3):
            assert g.pisoStep()
        out.append((g.getVelocity(), g.getPressure(), g.stepStats()))
    for u, p, st in out[1:]:
        assert np.array_equal(out[0][0], u) and np.array_equal(out[0][1], p)
        assert st["momentum_iters"] == out[0][2]["momentum_iters"]


# ------------------------------------------------------------------ failures
def test_step_before_initialize(P):
    g = P.NavierStokesSolver()
    assert g.pisoStepStatus() == 1
    assert "not initialized" in P.last_error()


def test_momentum_noconvergence_reported(P, O):
    # step 1 from rest: the y/z right-hand sides are zero and x is penalty
    # dominated, so momentum converges; step 2 needs more than one iteration
    g, o, pm, om = make_case(P, O, "kuhn", 4)
    assert g.pisoStep()
    g.setOptions(max_iter=1)
    assert g.pisoStepStatus() == 2
    assert "Momentum" in P.last_error()


def test_nan_reported(P, O):
    g, o, pm, om = make_case(P, O, "kuhn", 4)
    u = g.getVelocity()
    lab = np.r_[pm.p1BC(), pm.p2BC()]
    u[lab == 3] = np.nan
    g.setVelocity(u)
    st = g.pisoStepStatus()
    assert st in (2, 4)
    assert g.hasNans() or st == 2


def test_vtk_export_matches_fields(P, O, tmp_path):
    """io::exportToVtk (VtkExport.cpp:11-16,49-84,238-268) after two PISO steps,
    line by line against the text rendered here from the oracle's own state:
    ostream's default formatting is printf's %g (6 significant digits), points
    and velocity scaled, pressure not, labels as integers."""
    g, o, pm, om = make_case(P, O, "kuhn", 2)
    for _ in range(2):
        assert g.pisoStep() and o.pisoStep() == 0
    scale = 2.0
    path = tmp_path / "f.vtk"
    g.exportVtk(path, scale=scale)
    got = path.read_text().split("\n")
    X, T, bc = om.vertices(), om.tets(), om.p1BC()
    u, p = o.getVelocity(), o.getPressure()
    nv, nt = len(X), len(T)
    f = lambda v: "%g" % v
    want = ["# vtk DataFile Version 3.0", "Shape export", "ASCII", "DATASET UNSTRUCTURED_GRID",
            f"POINTS {nv} double"]
    want += [" ".join(f(c * scale) for c in x) for x in X]
    want += [f"CELLS {nt} {5 * nt}"] + ["4 " + " ".join(str(int(i)) for i in t) for t in T]
    want += [f"CELL_TYPES {nt}"] + ["10"] * nt
    want += ["", f"POINT_DATA {nv}", "SCALARS FluidBoundaryType int 1", "LOOKUP_TABLE default"]
    want += [str(int(b)) for b in bc]
    want += ["", "SCALARS pressure double 1", "LOOKUP_TABLE default"] + [f(v) for v in p]
    want += ["", "VECTORS velocity double"]
    want += [" ".join(f(c * scale) for c in x) for x in u[:nv]] + [""]
    assert got == want


def test_boundary_forces_bit_exact(P, O):
    g, o, pm, om = make_case(P, O, "kuhn", 4)
    for _ in range(2):
        assert g.pisoStep() and o.pisoStep() == 0
    u = np.random.default_rng(3).uniform(-1, 1, (g.sizes()["nu"], 3))
    p = np.random.default_rng(4).uniform(-1, 1, g.sizes()["nv"])
    for obj in (g, o):
        obj.setVelocity(u)
        obj.setPressure(p)
    assert np.array_equal(g.computePressureForces(), o.forces(0))
    assert np.array_equal(g.computeShearStress(), o.forces(1))
    assert np.array_equal(g.computeTotalFluidForces(), o.forces(2))
    v, t, geo = pm.faces()
    assert g.nFaces() == len(v)
    for fi in [0, len(v) // 2, len(v) - 1]:
        mp = p[v[fi]].sum() / 3.0
        assert g.getMeanFacePressure(fi) == mp
        np.testing.assert_array_equal(g.getPressureForceAtFace(fi), (-geo[fi, :3] * geo[fi, 3]) * mp)
