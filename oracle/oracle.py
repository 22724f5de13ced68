"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product library never does.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
_lib = C.CDLL(str(LIB_PATH))
_p = C.c_void_p
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("orc_last_error", C.c_char_p)
_sig("orc_eigen_dot", C.c_double, _f64p, _f64p, C.c_int64)
_sig("orc_dot2", C.c_double, _f64p, _f64p, C.c_int64)
_sig("orc_set_dot_mode", None, C.c_int)
_sig("orc_get_dot_mode", C.c_int)
_sig("orc_csr_spmv", None, C.c_int64, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int)
_sig("orc_mesh_structured_prism", C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
     C.c_int, C.POINTER(_p))
_sig("orc_mesh_from_arrays", C.c_int, C.c_int64, _f64p, C.c_int64, _i32p, C.POINTER(_p))
_sig("orc_mesh_setup_bc", C.c_int, _p, _f64p, C.c_double)
_sig("orc_mesh_update_vertices", C.c_int, _p, _f64p)
_sig("orc_mesh_set_p1_bc", C.c_int, _p, _u8p)
_sig("orc_mesh_sizes", None, _p, C.POINTER(C.c_int64))
for _n, _t in [("vertices", _f64p), ("tets", _i32p), ("volumes", _f64p), ("gradients", _f64p),
               ("tet_edge_nodes", _i32p), ("edge_nodes", _f64p), ("p1_bc", _u8p),
               ("p2_bc", _u8p)]:
    _sig(f"orc_mesh_get_{_n}", None, _p, _t)
_sig("orc_mesh_get_faces", None, _p, _i32p, _i32p, _f64p)
_sig("orc_mesh_destroy", None, _p)
_sig("orc_colour_tets", C.c_int, _i32p, C.c_int64, C.c_int64, _i32p)
_sig("orc_ns_create", C.c_int, _p, C.c_double, C.c_double, _f64p, C.POINTER(_p))
_sig("orc_ns_destroy", None, _p)
_sig("orc_ns_set_dt", None, _p, C.c_double)
_sig("orc_ns_set_mean_inlet", None, _p, _f64p)
_sig("orc_ns_set_inlet_profile", None, _p, _f64p)
_sig("orc_ns_set_outlet", None, _p, C.c_int, C.c_double)
_sig("orc_ns_set_corrections", None, _p, C.c_int)
_sig("orc_ns_set_relax", None, _p, C.c_double, C.c_double)
_sig("orc_ns_set_reference_node", None, _p, C.c_int)
_sig("orc_ns_set_solver", None, _p, C.c_double, C.c_double, C.c_int, C.c_int)
_sig("orc_ns_piso_step", C.c_int, _p)
_sig("orc_ns_simple_step", C.c_int, _p)
_sig("orc_ns_sizes", None, _p, C.POINTER(C.c_int64))
for _n in ["velocity", "velocity_star", "pressure", "pressure_correction", "inv_lumped_mass"]:
    _sig(f"orc_ns_get_{_n}", None, _p, _f64p)
_sig("orc_ns_set_velocity", None, _p, _f64p)
_sig("orc_ns_set_pressure", None, _p, _f64p)
_sig("orc_ns_get_matrix", None, _p, C.c_int, _p, _p, _p)
_sig("orc_ns_compute_advection", None, _p, _f64p)
_sig("orc_ns_compute_divergence", None, _p, _f64p, _f64p)
_sig("orc_ns_get_stats", None, _p, _i32p, _f64p)
_sig("orc_ns_krylov_solve", C.c_int, _p, C.c_int, C.c_int, _f64p, _f64p, C.c_double, C.c_int,
     C.POINTER(C.c_int32), C.POINTER(C.c_double))

_sig("orc_ns_set_stop_scaled", None, _p, C.c_int)
_sig("orc_ns_set_momentum_solver", None, _p, C.c_int)
_sig("orc_ns_forces", None, _p, C.c_int, _f64p)
_sig("orc_adjoint_gradient", C.c_int, _p, C.c_int, C.c_int, _f64p, _f64p,
     C.POINTER(C.c_double), _f64p, C.POINTER(C.c_double))

MATRIX = {"M": 0, "K": 1, "lhs": 2, "L": 3, "Lbc": 4, "Gx": 5, "Gy": 6, "Gz": 7}


def eigen_dot(a, b) -> float:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return _lib.orc_eigen_dot(a, b, a.size)


def dot2(a, b) -> float:
    """Order-robust inner product (Ogita-Rump-Oishi Dot2), sequential."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return _lib.orc_dot2(a, b, a.size)


def set_dot_mode(mode: int) -> None:
    """Inner products of the Krylov restatements: 0 Dot2, 1 Eigen packet order."""
    _lib.orc_set_dot_mode(int(mode))


def get_dot_mode() -> int:
    return int(_lib.orc_get_dot_mode())


def csr_spmv(rp, col, val, x):
    """Sequential CSR product in Eigen's accumulation order."""
    x = np.ascontiguousarray(x, np.float64)
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    n = len(rp) - 1
    y = np.empty((n, nrhs) if nrhs > 1 else (n,), np.float64)
    _lib.orc_csr_spmv(n, np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(col, np.int32),
                      np.ascontiguousarray(val, np.float64), x, y, nrhs)
    return y


_sig("orc_csr_krylov", C.c_int, C.c_int64, _i32p, _i32p, _f64p, C.c_int, _f64p, _f64p,
     C.c_double, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double))


_sig("orc_set_restart_scale", None, C.c_double)
_sig("orc_last_restarts", C.c_int)


def set_restart_scale(s: float) -> None:
    """Multiplier of eps^2 in the BiCGSTAB restart test (1 = Eigen's rule)."""
    _lib.orc_set_restart_scale(float(s))


def last_restarts() -> int:
    return _lib.orc_last_restarts()


def csr_krylov(rp, col, val, b, method=0, tol=1e-8, maxit=1000):
    """Eigen-form BiCGSTAB (0) or Jacobi CG (1) on a CSR matrix."""
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    it, er = C.c_int32(), C.c_double()
    _lib.orc_csr_krylov(len(rp) - 1, np.ascontiguousarray(rp, np.int32),
                        np.ascontiguousarray(col, np.int32), np.ascontiguousarray(val, np.float64),
                        int(method), b, x, float(tol), int(maxit), C.byref(it), C.byref(er))
    return x, {"iters": it.value, "err": er.value}


def colour_tets(tets, nv):
    tets = np.ascontiguousarray(tets, np.int32)
    out = np.empty(len(tets), np.int32)
    n = _lib.orc_colour_tets(tets, len(tets), int(nv), out)
    return out, n


class OracleMesh:
    def __init__(self, h):
        self._h = h

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            _lib.orc_mesh_destroy(h)

    @staticmethod
    def structuredRectangularPrism(L, W, H, nx, ny, nz):
        h = _p()
        if _lib.orc_mesh_structured_prism(L, W, H, nx, ny, nz, C.byref(h)):
            raise RuntimeError(_lib.orc_last_error().decode())
        return OracleMesh(h)

    @staticmethod
    def fromArrays(coords, tets):
        coords = np.ascontiguousarray(coords, np.float64)
        tets = np.ascontiguousarray(tets, np.int32)
        h = _p()
        if _lib.orc_mesh_from_arrays(len(coords), coords, len(tets), tets, C.byref(h)):
            raise RuntimeError(_lib.orc_last_error().decode())
        return OracleMesh(h)

    def setupBoundaryConditions(self, direction=(1.0, 0.0, 0.0), coverage=1.0):
        _lib.orc_mesh_setup_bc(self._h, np.asarray(direction, np.float64), float(coverage))

    def updateVertices(self, coords):
        coords = np.ascontiguousarray(coords, np.float64)
        assert coords.shape == (self.sizes()[0], 3)
        _lib.orc_mesh_update_vertices(self._h, coords)

    def setFluidVertexBCs(self, bc):
        _lib.orc_mesh_set_p1_bc(self._h, np.ascontiguousarray(bc, np.uint8))

    def sizes(self):
        out = (C.c_int64 * 4)()
        _lib.orc_mesh_sizes(self._h, out)
        return tuple(out)

    def _get(self, name, shape, dtype):
        out = np.empty(shape, dtype)
        getattr(_lib, f"orc_mesh_get_{name}")(self._h, out)
        return out

    def vertices(self):
        return self._get("vertices", (self.sizes()[0], 3), np.float64)

    def tets(self):
        return self._get("tets", (self.sizes()[1], 4), np.int32)

    def volumes(self):
        return self._get("volumes", (self.sizes()[1],), np.float64)

    def gradients(self):
        return self._get("gradients", (self.sizes()[1], 4, 3), np.float64)

    def tetEdgeNodes(self):
        return self._get("tet_edge_nodes", (self.sizes()[1], 6), np.int32)

    def edgeNodes(self):
        return self._get("edge_nodes", (self.sizes()[2], 3), np.float64)

    def p1BC(self):
        return self._get("p1_bc", (self.sizes()[0],), np.uint8)

    def p2BC(self):
        return self._get("p2_bc", (self.sizes()[2],), np.uint8)

    def faces(self):
        nf = self.sizes()[3]
        v = np.empty((nf, 3), np.int32)
        t = np.empty((nf, 4), np.int32)
        g = np.empty((nf, 7), np.float64)
        _lib.orc_mesh_get_faces(self._h, v, t, g)
        return v, t, g


class OracleNS:
    """CPU restatement of NavierStokesSolver (FP64, Eigen semantics)."""

    def __init__(self, mesh: OracleMesh, density, viscosity, inlet0=(0.0, 0.0, 0.0)):
        h = _p()
        if _lib.orc_ns_create(mesh._h, float(density), float(viscosity),
                              np.asarray(inlet0, np.float64), C.byref(h)):
            raise RuntimeError(_lib.orc_last_error().decode())
        self._h = h
        self._mesh = mesh

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            _lib.orc_ns_destroy(h)

    def setDt(self, dt):
        _lib.orc_ns_set_dt(self._h, float(dt))

    def setMeanInletVelocity(self, v):
        _lib.orc_ns_set_mean_inlet(self._h, np.asarray(v, np.float64))

    def setInletProfile(self, g):
        _lib.orc_ns_set_inlet_profile(self._h, np.ascontiguousarray(g, np.float64))

    def setOutlet(self, kind=0, pressure=0.0):
        _lib.orc_ns_set_outlet(self._h, int(kind), float(pressure))

    def setNumberOfCorrections(self, n):
        _lib.orc_ns_set_corrections(self._h, int(n))

    def setRelax(self, ru, rp):
        _lib.orc_ns_set_relax(self._h, float(ru), float(rp))

    def setReferencePressureNode(self, n):
        _lib.orc_ns_set_reference_node(self._h, int(n))

    def setSolver(self, momentum_tol=1e-6, pressure_tol=1e-8, max_iter=1000, pressure_solver=0):
        _lib.orc_ns_set_solver(self._h, momentum_tol, pressure_tol, max_iter, pressure_solver)

    def forces(self, which):
        """0 computePressureForces, 1 computeShearStress, 2 computeTotalFluidForces."""
        out = np.empty((self.sizes()["nv"], 3), np.float64)
        _lib.orc_ns_forces(self._h, int(which), out)
        return out

    def setStopScaled(self, on=True):
        _lib.orc_ns_set_stop_scaled(self._h, int(bool(on)))

    def setMomentumSolver(self, method):
        """Extension: 0 BiCGSTAB (reference), 1 Jacobi CG."""
        _lib.orc_ns_set_momentum_solver(self._h, int(method))

    def adjointGradient(self, nsteps, basis, coef):
        """J, dJ/dcoef, dJ/dviscosity of J = 1/2 sum_outlet |u|^2 after nsteps."""
        basis = np.ascontiguousarray(basis, np.float64)
        coef = np.ascontiguousarray(coef, np.float64)
        J = C.c_double()
        dmu = C.c_double()
        dc = np.empty(len(coef), np.float64)
        st = _lib.orc_adjoint_gradient(self._h, int(nsteps), len(coef), basis, coef, C.byref(J),
                                       dc, C.byref(dmu))
        if st:
            raise RuntimeError(f"adjoint failed with status {st}")
        return J.value, dc, dmu.value

    def pisoStep(self) -> int:
        return _lib.orc_ns_piso_step(self._h)

    def simpleStep(self) -> int:
        return _lib.orc_ns_simple_step(self._h)

    def sizes(self):
        out = (C.c_int64 * 6)()
        _lib.orc_ns_sizes(self._h, out)
        return dict(zip(["nv", "nt", "ne", "nu", "nnzA", "nnzP"], out))

    def _arr(self, name, shape):
        out = np.empty(shape, np.float64)
        getattr(_lib, f"orc_ns_get_{name}")(self._h, out)
        return out

    def getVelocity(self):
        return self._arr("velocity", (self.sizes()["nu"], 3))

    def setVelocity(self, u):
        _lib.orc_ns_set_velocity(self._h, np.ascontiguousarray(u, np.float64))

    def getVelocityStar(self):
        return self._arr("velocity_star", (self.sizes()["nu"], 3))

    def getPressure(self):
        return self._arr("pressure", (self.sizes()["nv"],))

    def setPressure(self, p):
        _lib.orc_ns_set_pressure(self._h, np.ascontiguousarray(p, np.float64))

    def getPressureCorrection(self):
        return self._arr("pressure_correction", (self.sizes()["nv"],))

    def getInvLumpedMass(self):
        return self._arr("inv_lumped_mass", (self.sizes()["nu"],))

    def getMatrix(self, which):
        w = MATRIX[which] if isinstance(which, str) else int(which)
        sz = self.sizes()
        n, nnz = (sz["nu"], sz["nnzA"]) if w <= 2 else (sz["nv"], sz["nnzP"])
        rp = np.empty(n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        _lib.orc_ns_get_matrix(self._h, w, rp.ctypes.data_as(_p), col.ctypes.data_as(_p),
                               val.ctypes.data_as(_p))
        return rp, col, val

    def computeAdvectionRHS(self):
        return self._arr("velocity", (0,)) if False else self._adv()

    def _adv(self):
        out = np.empty((self.sizes()["nu"], 3), np.float64)
        _lib.orc_ns_compute_advection(self._h, out)
        return out

    def computeDivergence(self, u):
        out = np.empty(self.sizes()["nv"], np.float64)
        _lib.orc_ns_compute_divergence(self._h, np.ascontiguousarray(u, np.float64), out)
        return out

    def stats(self):
        it = np.empty(11, np.int32)
        er = np.empty(11, np.float64)
        _lib.orc_ns_get_stats(self._h, it, er)
        return {"momentum_iters": it[:3].tolist(), "momentum_err": er[:3].tolist(),
                "pressure_iters": it[3:].tolist(), "pressure_err": er[3:].tolist()}

    def krylovSolve(self, which, b, method=0, tol=1e-8, maxit=1000):
        w = MATRIX[which] if isinstance(which, str) else int(which)
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        it = C.c_int32()
        er = C.c_double()
        _lib.orc_ns_krylov_solve(self._h, w, int(method), b, x, float(tol), int(maxit),
                                 C.byref(it), C.byref(er))
        return x, {"iters": it.value, "err": er.value}
