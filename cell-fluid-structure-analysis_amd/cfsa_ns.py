"""ctypes binding of libcfsa_ns.so (the C ABI in include/cfsa_ns.h).

Method names follow the reference C++ classes (Mesh, NavierStokesSolver in
sudo-shaka/Cell-Fluid-Structure-Analysis) so harness code reads like code
written against the reference API.  There is no fallback: if the HIP library
is not built, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("CFSA_NS_LIB", _PKG / "lib" / "libcfsa_ns.so"))

if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
_lib = C.CDLL(str(LIB_PATH))

_p = C.c_void_p
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")

WALL, INLET, OUTLET, INTERNAL, UNDEFINED = 0, 1, 2, 3, 4
OUTLET_DIRICHLET, OUTLET_NEUMANN, OUTLET_UNDEFINED = 0, 1, 2
STATUS = {0: "OK", 1: "NOT_INIT", 2: "MOMENTUM_NOCONV", 3: "PRESSURE_NOCONV", 4: "NAN",
          5: "INVALID", 6: "DEVICE", 7: "IO"}
MATRIX = {"M": 0, "K": 1, "lhs": 2, "L": 3, "Lbc": 4, "Gx": 5, "Gy": 6, "Gz": 7, "lhs32": 8}


class SolverOptions(C.Structure):
    _fields_ = [("momentum_tol", C.c_double), ("pressure_tol", C.c_double),
                ("max_iter", C.c_int), ("pressure_solver", C.c_int),
                ("precision", C.c_int), ("inner_tol", C.c_double),
                ("max_refine", C.c_int), ("check_every", C.c_int), ("timing", C.c_int),
                ("stop_rule", C.c_int), ("assembly", C.c_int),
                ("spmv_layout", C.c_int), ("momentum_solver", C.c_int),
                ("check_every_1", C.c_int), ("graphs", C.c_int), ("initial_guess", C.c_int),
                ("reuse_pressure_correction", C.c_int), ("restart_scale", C.c_double),
                ("amg_reuse", C.c_int), ("elem_spmv", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("momentum_iters", C.c_int * 3), ("momentum_err", C.c_double * 3),
                ("momentum_refinements", C.c_int), ("n_pressure", C.c_int),
                ("pressure_iters", C.c_int * 8), ("pressure_err", C.c_double * 8),
                ("pressure_refinements", C.c_int * 8),
                ("ms_conv", C.c_double), ("ms_rhs", C.c_double), ("ms_momentum", C.c_double),
                ("ms_pressure", C.c_double), ("ms_correct", C.c_double),
                ("ms_total", C.c_double), ("spmv_launches", C.c_int), ("spmv_ms", C.c_double),
                ("pressure_solves", C.c_int)]

    def as_dict(self) -> dict:
        n = self.n_pressure
        return {"momentum_iters": list(self.momentum_iters),
                "momentum_err": list(self.momentum_err),
                "momentum_refinements": self.momentum_refinements,
                "pressure_iters": list(self.pressure_iters)[:n],
                "pressure_err": list(self.pressure_err)[:n],
                "pressure_refinements": list(self.pressure_refinements)[:n],
                "ms_momentum": self.ms_momentum, "ms_pressure": self.ms_pressure,
                "ms_total": self.ms_total, "spmv_launches": self.spmv_launches,
                "spmv_ms": self.spmv_ms, "pressure_solves": self.pressure_solves}


class SolveResult(C.Structure):
    _fields_ = [("iters", C.c_int * 3), ("err", C.c_double * 3), ("restarts", C.c_int * 3),
                ("refinements", C.c_int), ("inner_iters", C.c_int), ("ok", C.c_int)]


class SimConfig(C.Structure):
    _fields_ = [("mesh_type", C.c_int), ("mesh_file", C.c_char * 1024), ("obj_scale", C.c_double),
                ("cylinder_length", C.c_double), ("cylinder_radius", C.c_double),
                ("cylinder_divisions", C.c_int), ("icosphere_radius", C.c_double),
                ("icosphere_recursion", C.c_int), ("max_element_size", C.c_double),
                ("flow_direction", C.c_double * 3), ("inlet_outlet_coverage", C.c_double),
                ("box_size", C.c_double * 3), ("box_cells", C.c_int * 3),
                ("ns_enabled", C.c_int), ("density", C.c_double), ("viscosity", C.c_double),
                ("is_newtonian", C.c_int), ("is_laminar", C.c_int),
                ("inlet_velocity", C.c_double * 3), ("outlet_pressure", C.c_double),
                ("use_dirichlet_pressure", C.c_int), ("dt", C.c_double), ("relax_p", C.c_double),
                ("relax_u", C.c_double), ("n_corrections", C.c_int),
                ("pressure_solver", C.c_int), ("precision", C.c_int),
                ("momentum_tol", C.c_double), ("pressure_tol", C.c_double),
                ("inlet_profile", C.c_int), ("output_prefix", C.c_char * 512),
                ("total_steps", C.c_int), ("output_frequency", C.c_int),
                ("vtk_scale", C.c_double), ("output_fluid", C.c_int), ("output_solid", C.c_int),
                ("output_dpm", C.c_int), ("solid_enabled", C.c_int), ("dpm_enabled", C.c_int),
                ("fsi_enabled", C.c_int)]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if isinstance(v, bytes):
                v = v.decode()
            elif hasattr(v, "__len__"):
                v = list(v)
            out[name] = v
        return out


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("cfsa_last_error", C.c_char_p)
_sig("cfsa_abi_version", C.c_int)
_sig("cfsa_mesh_structured_prism", C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
     C.c_int, C.c_int, C.POINTER(_p))
_sig("cfsa_mesh_cylinder", C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.POINTER(_p))
_sig("cfsa_mesh_from_msh", C.c_int, C.c_char_p, C.POINTER(_p))
_sig("cfsa_mesh_from_arrays", C.c_int, C.c_int64, _f64p, C.c_int64, _i32p, C.POINTER(_p))
_sig("cfsa_mesh_setup_boundary_conditions", C.c_int, _p, _f64p, C.c_double)
_sig("cfsa_mesh_set_p1_bc", C.c_int, _p, _u8p)
_sig("cfsa_mesh_update_vertices", C.c_int, _p, _f64p)
_sig("cfsa_mesh_sizes", C.c_int, _p, C.POINTER(C.c_int64))
for _n, _t in [("vertices", _f64p), ("tets", _i32p), ("volumes", _f64p), ("gradients", _f64p),
               ("tet_edge_nodes", _i32p), ("edge_nodes", _f64p), ("edge_endpoints", _i32p),
               ("p1_bc", _u8p), ("p2_bc", _u8p)]:
    _sig(f"cfsa_mesh_get_{_n}", C.c_int, _p, _t)
_sig("cfsa_mesh_get_faces", C.c_int, _p, _i32p, _i32p, _f64p)
_sig("cfsa_mesh_get_tet_faces", C.c_int, _p, _i32p, _i32p)
_sig("cfsa_mesh_get_min_edge_length", C.c_int, _p, C.POINTER(C.c_double))
_sig("cfsa_mesh_export_vtk", C.c_int, _p, C.c_char_p, C.c_double)
_sig("cfsa_mesh_destroy", None, _p)

_sig("cfsa_ns_create", C.c_int, C.POINTER(_p))
_sig("cfsa_ns_destroy", None, _p)
_sig("cfsa_ns_initialize", C.c_int, _p, _p, C.c_double, C.c_double)
_sig("cfsa_ns_initialize_arrays", C.c_int, _p, C.c_int64, _f64p, C.c_int64, _i32p, C.c_int64,
     _i32p, _u8p, _u8p, C.c_double, C.c_double)
_sig("cfsa_ns_update_mesh", C.c_int, _p, _p, _p, _p)
_sig("cfsa_ns_build_sparse_matrices", C.c_int, _p)
_sig("cfsa_ns_piso_step", C.c_int, _p)
_sig("cfsa_ns_simple_step", C.c_int, _p)
_sig("cfsa_ns_set_mean_inlet_velocity", C.c_int, _p, _f64p)
_sig("cfsa_ns_set_inlet_profile", C.c_int, _p, _p)
for _n, _t in [("outlet_type", C.c_int), ("outlet_pressure", C.c_double),
               ("inlet_type", C.c_int), ("number_of_corrections", C.c_int),
               ("relax_u", C.c_double), ("relax_p", C.c_double),
               ("reference_pressure_node", C.c_int), ("dt", C.c_double),
               ("tolerance", C.c_double), ("time", C.c_double)]:
    _sig(f"cfsa_ns_set_{_n}", C.c_int, _p, _t)
_sig("cfsa_ns_set_options", C.c_int, _p, C.POINTER(SolverOptions))
_sig("cfsa_ns_get_options", C.c_int, _p, C.POINTER(SolverOptions))
_sig("cfsa_ns_get_sizes", C.c_int, _p, C.POINTER(C.c_int64))
for _n in ["velocity", "velocity_star", "pressure", "pressure_correction", "inv_lumped_mass"]:
    _sig(f"cfsa_ns_get_{_n}", C.c_int, _p, _f64p)
_sig("cfsa_ns_set_velocity", C.c_int, _p, _f64p)
_sig("cfsa_ns_set_pressure", C.c_int, _p, _f64p)
_sig("cfsa_ns_get_velocity_at_node", C.c_int, _p, C.c_int64, _f64p)
_sig("cfsa_ns_get_fields_async", C.c_int, _p, _p, _p)
_sig("cfsa_ns_wait_fields", C.c_int, _p)
_sig("cfsa_host_alloc", C.c_int, C.c_size_t, C.POINTER(_p))
_sig("cfsa_host_free", None, _p)
_sig("cfsa_ns_get_pressure_at_node", C.c_int, _p, C.c_int64, C.POINTER(C.c_double))
_sig("cfsa_ns_get_scalar", C.c_int, _p, C.c_int, C.POINTER(C.c_double))
_sig("cfsa_ns_get_reference_pressure_node", C.c_int, _p, C.POINTER(C.c_int))
_sig("cfsa_ns_has_nans", C.c_int, _p, C.POINTER(C.c_int))
_sig("cfsa_ns_get_step_stats", C.c_int, _p, C.POINTER(StepStats))
_sig("cfsa_ns_export_vtk", C.c_int, _p, C.c_char_p, C.c_double)
_sig("cfsa_ns_get_matrix", C.c_int, _p, C.c_int, _p, _p, _p)
_sig("cfsa_ns_get_num_colours", C.c_int, _p, C.POINTER(C.c_int))
_sig("cfsa_ns_get_colouring", C.c_int, _p, _i32p)
_sig("cfsa_ns_get_incidence", C.c_int, _p, _i32p, _i32p)
_sig("cfsa_ns_compute_advection", C.c_int, _p, _f64p)
_sig("cfsa_ns_compute_divergence", C.c_int, _p, _f64p, _f64p)
_sig("cfsa_ns_apply_matrix", C.c_int, _p, C.c_int, _f64p, _f64p, C.c_int)
_sig("cfsa_ns_apply_matrix_noalloc", C.c_int, _p, C.c_int, _f64p, _f64p, C.c_int)
_sig("cfsa_ns_krylov_solve", C.c_int, _p, C.c_int, C.c_int, C.c_int, _f64p, _f64p, C.c_double,
     C.c_int, C.c_int, C.POINTER(SolveResult))
_sig("cfsa_ns_time_momentum_spmv", C.c_int, _p, C.c_int, C.c_int, C.POINTER(C.c_double))
_sig("cfsa_ns_amg_apply", C.c_int, _p, _f64p, _f64p, C.c_int)
for _n in ["solve_momentum_predictor", "solve_pressure_poisson", "correct_velocity",
           "normalize_pressure"]:
    _sig(f"cfsa_ns_{_n}", C.c_int, _p)
_sig("cfsa_ns_solve_pressure_bicgstab", C.c_int, _p, _f64p, _f64p, C.POINTER(SolveResult))
_sig("cfsa_ns_amg_info", C.c_int, _p, C.c_int, _p, _p, C.POINTER(C.c_double),
     C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("cfsa_amg_build", C.c_int, C.c_int64, _i32p, _i32p, _f64p, C.c_double, C.c_int,
     C.POINTER(_p))
_sig("cfsa_amg_free", C.c_int, _p)
_sig("cfsa_amg_num_levels", C.c_int, _p, C.POINTER(C.c_int), C.POINTER(C.c_int64),
     C.POINTER(C.c_int))
_sig("cfsa_amg_level_size", C.c_int, _p, C.c_int, C.c_int, C.POINTER(C.c_int64),
     C.POINTER(C.c_int64), C.POINTER(C.c_int64))
_sig("cfsa_amg_level_get", C.c_int, _p, C.c_int, C.c_int, _i32p, _i32p, _f64p)
_sig("cfsa_amg_level_params", C.c_int, _p, C.c_int, C.POINTER(C.c_double),
     C.POINTER(C.c_double), _f64p, _p)
_sig("cfsa_amg_coarse_inverse", C.c_int, _p, _f64p)
_sig("cfsa_elem_patches_build", C.c_int, C.c_int64, C.c_int64, C.c_int64, _i32p, _i32p, _f64p,
     C.POINTER(_p))
_sig("cfsa_elem_patches_free", C.c_int, _p)
_sig("cfsa_ns_momentum_poly", C.c_int, _p,
     np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS"))
_sig("cfsa_ns_momentum_poly2", C.c_int, _p,
     np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS"))
_sig("cfsa_ns_elem_info", C.c_int, _p, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"),
     C.POINTER(C.c_double))
_sig("cfsa_elem_patches_sizes", C.c_int, _p, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"))
_sig("cfsa_elem_patches_get", C.c_int, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p)
_sig("cfsa_ns_get_stream", C.c_int, _p, C.POINTER(_p))
for _n in ["pressure_forces", "shear_stress", "total_fluid_forces"]:
    _sig(f"cfsa_ns_compute_{_n}", C.c_int, _p, _f64p)
_sig("cfsa_ns_num_faces", C.c_int, _p, C.POINTER(C.c_int64))
_sig("cfsa_ns_get_mean_face_pressure", C.c_int, _p, C.c_int64, C.POINTER(C.c_double))
_sig("cfsa_ns_get_pressure_force_at_face", C.c_int, _p, C.c_int64, _f64p)
_sig("cfsa_ns_set_viscosity", C.c_int, _p, C.c_double)
_sig("cfsa_ns_adjoint_forward", C.c_int, _p, C.c_int, _f64p, _p)
_sig("cfsa_ns_adjoint_backward", C.c_int, _p, _p, _f64p, C.POINTER(C.c_double))
_sig("cfsa_ns_adjoint_gradient", C.c_int, _p, C.c_int, C.c_int, _f64p, _f64p,
     C.POINTER(C.c_double), _f64p, C.POINTER(C.c_double))
_sig("cfsa_config_load", C.c_int, C.c_char_p, C.POINTER(SimConfig))
_sig("cfsa_config_parse", C.c_int, C.c_char_p, C.POINTER(SimConfig))


def loadConfig(path) -> dict:
    """config::loadConfig (src/IO/YamlConfig.cpp:200-232)."""
    c = SimConfig()
    _check(_lib.cfsa_config_load(str(path).encode(), C.byref(c)))
    return c.as_dict()


def parseConfig(text: str) -> dict:
    c = SimConfig()
    _check(_lib.cfsa_config_parse(text.encode(), C.byref(c)))
    return c.as_dict()


class CfsaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def _check(code: int) -> int:
    if code != 0:
        raise CfsaError(code, _lib.cfsa_last_error().decode())
    return code


def last_error() -> str:
    return _lib.cfsa_last_error().decode()


def library():
    return _lib


class PinnedArray:
    """A float64 numpy array in page-locked host memory (cfsa_host_alloc)."""

    def __init__(self, shape):
        n = int(np.prod(shape))
        self._p = _p()
        _check(_lib.cfsa_host_alloc(n * 8, C.byref(self._p)))
        buf = (C.c_double * n).from_address(self._p.value) if n else (C.c_double * 0)()
        self.array = np.frombuffer(buf, dtype=np.float64).reshape(shape)

    def __del__(self):
        p, self._p = getattr(self, "_p", None), None
        if p and p.value and _lib is not None:
            self.array = None
            _lib.cfsa_host_free(p)


# ---------------------------------------------------------------------------
class Mesh:
    """Reference Mesh (include/Mesh/Mesh.hpp:35-220) on the host."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.cfsa_mesh_destroy(h)

    @property
    def handle(self):
        return self._h

    @staticmethod
    def structuredRectangularPrism(length, width, height, nx=1, ny=1, nz=1) -> "Mesh":
        h = _p()
        _check(_lib.cfsa_mesh_structured_prism(length, width, height, nx, ny, nz, C.byref(h)))
        return Mesh(h)

    @staticmethod
    def cylinder(length, radius, divisions, max_element_size) -> "Mesh":
        h = _p()
        _check(_lib.cfsa_mesh_cylinder(length, radius, divisions, max_element_size, C.byref(h)))
        return Mesh(h)

    @staticmethod
    def fromMshFile(path) -> "Mesh":
        h = _p()
        _check(_lib.cfsa_mesh_from_msh(str(path).encode(), C.byref(h)))
        return Mesh(h)

    @staticmethod
    def fromArrays(coords, tets) -> "Mesh":
        coords = np.ascontiguousarray(coords, np.float64)
        tets = np.ascontiguousarray(tets, np.int32)
        h = _p()
        _check(_lib.cfsa_mesh_from_arrays(len(coords), coords, len(tets), tets, C.byref(h)))
        return Mesh(h)

    def setupBoundaryConditions(self, direction=(1.0, 0.0, 0.0), coverage=1.0):
        _check(_lib.cfsa_mesh_setup_boundary_conditions(
            self._h, np.asarray(direction, np.float64), float(coverage)))

    def setFluidVertexBCs(self, p1_bc):
        _check(_lib.cfsa_mesh_set_p1_bc(self._h, np.ascontiguousarray(p1_bc, np.uint8)))

    def updateVertices(self, coords):
        _check(_lib.cfsa_mesh_update_vertices(self._h, np.ascontiguousarray(coords, np.float64)))

    def sizes(self):
        out = (C.c_int64 * 5)()
        _check(_lib.cfsa_mesh_sizes(self._h, out))
        return tuple(out)

    def nVertices(self):
        return self.sizes()[0]

    def nTets(self):
        return self.sizes()[1]

    def getNumberOfEdgeNodes(self):
        return self.sizes()[2]

    def nFaces(self):
        return self.sizes()[3]

    def _get(self, name, shape, dtype):
        out = np.empty(shape, dtype)
        _check(getattr(_lib, f"cfsa_mesh_get_{name}")(self._h, out))
        return out

    def vertices(self):
        return self._get("vertices", (self.nVertices(), 3), np.float64)

    def tets(self):
        return self._get("tets", (self.nTets(), 4), np.int32)

    def volumes(self):
        return self._get("volumes", (self.nTets(),), np.float64)

    def gradients(self):
        return self._get("gradients", (self.nTets(), 4, 3), np.float64)

    def tetEdgeNodes(self):
        return self._get("tet_edge_nodes", (self.nTets(), 6), np.int32)

    def edgeNodes(self):
        return self._get("edge_nodes", (self.getNumberOfEdgeNodes(), 3), np.float64)

    def edgeEndpoints(self):
        return self._get("edge_endpoints", (self.getNumberOfEdgeNodes(), 2), np.int32)

    def p1BC(self):
        return self._get("p1_bc", (self.nVertices(),), np.uint8)

    def p2BC(self):
        return self._get("p2_bc", (self.getNumberOfEdgeNodes(),), np.uint8)

    def faces(self):
        nf = self.nFaces()
        v = np.empty((nf, 3), np.int32)
        t = np.empty((nf, 4), np.int32)
        g = np.empty((nf, 7), np.float64)
        _check(_lib.cfsa_mesh_get_faces(self._h, v, t, g))
        return v, t, g

    def tetFaces(self):
        f = np.empty((self.nTets(), 4), np.int32)
        n = np.empty((self.nTets(), 4), np.int32)
        _check(_lib.cfsa_mesh_get_tet_faces(self._h, f, n))
        return f, n

    def getMinEdgeLength(self):
        out = C.c_double()
        _check(_lib.cfsa_mesh_get_min_edge_length(self._h, C.byref(out)))
        return out.value

    def exportVtk(self, path, scale=1.0):
        _check(_lib.cfsa_mesh_export_vtk(self._h, str(path).encode(), float(scale)))


# ---------------------------------------------------------------------------
class NavierStokesSolver:
    """Reference NavierStokesSolver (include/FEM/NavierStokes.hpp:53-193)."""

    def __init__(self):
        h = _p()
        _check(_lib.cfsa_ns_create(C.byref(h)))
        self._h = h
        self._mesh = None

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.cfsa_ns_destroy(h)

    @property
    def handle(self):
        return self._h

    # --- setup ---
    def initialize(self, mesh: Mesh, density=1000.0, viscosity=0.001):
        _check(_lib.cfsa_ns_initialize(self._h, mesh.handle, float(density), float(viscosity)))
        self._mesh = mesh  # the handle must outlive the solver

    def initializeArrays(self, coords, tets, tet_edge_nodes, p1_bc, p2_bc, density, viscosity):
        coords = np.ascontiguousarray(coords, np.float64)
        tets = np.ascontiguousarray(tets, np.int32)
        ten = np.ascontiguousarray(tet_edge_nodes, np.int32)
        _check(_lib.cfsa_ns_initialize_arrays(
            self._h, len(coords), coords, len(tets), tets, int(ten.max()) + 1 if ten.size else 0,
            ten, np.ascontiguousarray(p1_bc, np.uint8), np.ascontiguousarray(p2_bc, np.uint8),
            float(density), float(viscosity)))

    def updateMesh(self, coords=None, p1_bc=None, p2_bc=None):
        keep = []

        def ptr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a.ctypes.data_as(_p)
        _check(_lib.cfsa_ns_update_mesh(self._h, ptr(coords, np.float64), ptr(p1_bc, np.uint8),
                                        ptr(p2_bc, np.uint8)))

    def build_sparse_matrices(self):
        _check(_lib.cfsa_ns_build_sparse_matrices(self._h))

    # --- stepping (bool like the reference; status code via *_status) ---
    def pisoStep(self) -> bool:
        return _lib.cfsa_ns_piso_step(self._h) == 0

    def pisoStepStatus(self) -> int:
        return _lib.cfsa_ns_piso_step(self._h)

    def simpleStep(self) -> bool:
        return _lib.cfsa_ns_simple_step(self._h) == 0

    # --- the parts of the step (NavierStokes.cpp:416-524, 586-668, 702-771, 804-813) ---
    def solveMomentumPredictor(self) -> bool:
        return _lib.cfsa_ns_solve_momentum_predictor(self._h) == 0

    def solvePressurePoisson(self) -> bool:
        return _lib.cfsa_ns_solve_pressure_poisson(self._h) == 0

    def correctVelocity(self) -> bool:
        return _lib.cfsa_ns_correct_velocity(self._h) == 0

    def normalizePressire(self) -> bool:
        return _lib.cfsa_ns_normalize_pressure(self._h) == 0

    def solvePressureBiCGSTAB(self, rhs):
        """NavierStokes.cpp:773-802; returns (ok, x, iterations)."""
        rhs = np.ascontiguousarray(rhs, np.float64)
        if rhs.shape != (self.sizes()["nv"],):
            raise ValueError("solvePressureBiCGSTAB: rhs has one entry per vertex")
        x = np.empty_like(rhs)
        r = SolveResult()
        st = _lib.cfsa_ns_solve_pressure_bicgstab(self._h, rhs, x, C.byref(r))
        if st not in (0, 3):
            _check(st)
        return st == 0, x, r.iters[0]

    # --- setters ---
    def setMeanInletVelocity(self, v):
        _check(_lib.cfsa_ns_set_mean_inlet_velocity(self._h, np.asarray(v, np.float64)))

    def setInletProfile(self, g):
        if g is None:
            _check(_lib.cfsa_ns_set_inlet_profile(self._h, None))
        else:
            g = np.ascontiguousarray(g, np.float64)
            _check(_lib.cfsa_ns_set_inlet_profile(self._h, g.ctypes.data_as(_p)))

    def setOutletType(self, t):
        _check(_lib.cfsa_ns_set_outlet_type(self._h, int(t)))

    def setOutletPressure(self, p):
        _check(_lib.cfsa_ns_set_outlet_pressure(self._h, float(p)))

    def setInletType(self, t):
        _check(_lib.cfsa_ns_set_inlet_type(self._h, int(t)))

    def setNumberOfCorrections(self, n):
        _check(_lib.cfsa_ns_set_number_of_corrections(self._h, int(n)))

    def setRelaxU(self, r):
        _check(_lib.cfsa_ns_set_relax_u(self._h, float(r)))

    def setRelaxP(self, r):
        _check(_lib.cfsa_ns_set_relax_p(self._h, float(r)))

    def setReferencePressureNode(self, n):
        _check(_lib.cfsa_ns_set_reference_pressure_node(self._h, int(n)))

    def setDt(self, dt):
        _check(_lib.cfsa_ns_set_dt(self._h, float(dt)))

    def setTolerance(self, t):
        _check(_lib.cfsa_ns_set_tolerance(self._h, float(t)))

    def setTime(self, t):
        _check(_lib.cfsa_ns_set_time(self._h, float(t)))

    def options(self) -> SolverOptions:
        o = SolverOptions()
        _check(_lib.cfsa_ns_get_options(self._h, C.byref(o)))
        return o

    def setOptions(self, **kw):
        o = self.options()
        for k, v in kw.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        _check(_lib.cfsa_ns_set_options(self._h, C.byref(o)))

    # --- getters ---
    def sizes(self):
        out = (C.c_int64 * 6)()
        _check(_lib.cfsa_ns_get_sizes(self._h, out))
        return dict(zip(["nv", "nt", "ne", "nu", "nnzA", "nnzP"], out))

    def _arr(self, name, shape):
        out = np.empty(shape, np.float64)
        _check(getattr(_lib, f"cfsa_ns_get_{name}")(self._h, out))
        return out

    def getVelocity(self):
        return self._arr("velocity", (self.sizes()["nu"], 3))

    def setVelocity(self, u):
        _check(_lib.cfsa_ns_set_velocity(self._h, np.ascontiguousarray(u, np.float64)))

    def getVelocityStar(self):
        return self._arr("velocity_star", (self.sizes()["nu"], 3))

    def getPressure(self):
        return self._arr("pressure", (self.sizes()["nv"],))

    def setPressure(self, p):
        _check(_lib.cfsa_ns_set_pressure(self._h, np.ascontiguousarray(p, np.float64)))

    def getPressureCorrection(self):
        return self._arr("pressure_correction", (self.sizes()["nv"],))

    def getInvLumpedMass(self):
        return self._arr("inv_lumped_mass", (self.sizes()["nu"],))

    def getFieldsAsync(self, u_out, p_out):
        """Start copying u and p into caller arrays (pinned ones from
        PinnedArray overlap later steps); waitFields() completes the copy."""
        _check(_lib.cfsa_ns_get_fields_async(self._h, u_out.ctypes.data if u_out is not None else None,
                                             p_out.ctypes.data if p_out is not None else None))

    def waitFields(self):
        _check(_lib.cfsa_ns_wait_fields(self._h))

    def getVelocityAtNode(self, i):
        out = np.empty(3, np.float64)
        _check(_lib.cfsa_ns_get_velocity_at_node(self._h, int(i), out))
        return out

    def getPressureAtNode(self, i):
        out = C.c_double()
        _check(_lib.cfsa_ns_get_pressure_at_node(self._h, int(i), C.byref(out)))
        return out.value

    def _scalar(self, k):
        out = C.c_double()
        _check(_lib.cfsa_ns_get_scalar(self._h, k, C.byref(out)))
        return out.value

    def getTime(self):
        return self._scalar(0)

    def getDt(self):
        return self._scalar(1)

    def getViscosity(self):
        return self._scalar(2)

    def getDensity(self):
        return self._scalar(3)

    def getTolerance(self):
        return self._scalar(4)

    def getReferencePressureNode(self):
        out = C.c_int()
        _check(_lib.cfsa_ns_get_reference_pressure_node(self._h, C.byref(out)))
        return out.value

    def hasNans(self) -> bool:
        out = C.c_int()
        _check(_lib.cfsa_ns_has_nans(self._h, C.byref(out)))
        return bool(out.value)

    def stepStats(self) -> dict:
        s = StepStats()
        _check(_lib.cfsa_ns_get_step_stats(self._h, C.byref(s)))
        return s.as_dict()

    def exportVtk(self, path, scale=1.0):
        _check(_lib.cfsa_ns_export_vtk(self._h, str(path).encode(), float(scale)))

    # --- parity / diagnostics ---
    def getMatrix(self, which):
        """CSR (rowptr, col, val) of operator `which` (see MATRIX)."""
        w = MATRIX[which] if isinstance(which, str) else int(which)
        sz = self.sizes()
        n, nnz = (sz["nu"], sz["nnzA"]) if w <= 2 else (sz["nv"], sz["nnzP"])
        rp = np.empty(n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        _check(_lib.cfsa_ns_get_matrix(self._h, w, rp.ctypes.data_as(_p),
                                       col.ctypes.data_as(_p), val.ctypes.data_as(_p)))
        return rp, col, val

    def numColours(self):
        out = C.c_int()
        _check(_lib.cfsa_ns_get_num_colours(self._h, C.byref(out)))
        return out.value

    def colouring(self):
        out = np.empty(self.sizes()["nt"], np.int32)
        _check(_lib.cfsa_ns_get_colouring(self._h, out))
        return out

    def incidence(self):
        sz = self.sizes()
        ptr = np.empty(sz["nu"] + 1, np.int32)
        tet = np.empty(sz["nt"] * 10, np.int32)
        _check(_lib.cfsa_ns_get_incidence(self._h, ptr, tet))
        return ptr, tet

    def computeAdvectionRHS(self):
        return self._arr("velocity", (0,)) if False else self._advection()

    def _advection(self):
        out = np.empty((self.sizes()["nu"], 3), np.float64)
        _check(_lib.cfsa_ns_compute_advection(self._h, out))
        return out

    def computeDivergence(self, u):
        out = np.empty(self.sizes()["nv"], np.float64)
        _check(_lib.cfsa_ns_compute_divergence(self._h, np.ascontiguousarray(u, np.float64), out))
        return out

    def applyMatrix(self, which, x, nrhs=1):
        w = MATRIX[which] if isinstance(which, str) else int(which)
        x = np.ascontiguousarray(x, np.float64)
        sz = self.sizes()
        if 5 <= w <= 7:
            y = np.empty((sz["nv"], 3), np.float64)
        else:
            n = sz["nu"] if w <= 2 or w == 8 else sz["nv"]
            y = np.empty((n, nrhs) if nrhs == 3 else (n,), np.float64)
        _check(_lib.cfsa_ns_apply_matrix(self._h, w, x, y, nrhs))
        return y

    def krylovSolve(self, which, b, method=0, tol=1e-8, maxit=1000, precision=0):
        w = MATRIX[which] if isinstance(which, str) else int(which)
        b = np.ascontiguousarray(b, np.float64)
        nrhs = 3 if b.ndim == 2 else 1
        x = np.empty_like(b)
        r = SolveResult()
        _check(_lib.cfsa_ns_krylov_solve(self._h, w, int(method), nrhs, b, x, float(tol),
                                         int(maxit), int(precision), C.byref(r)))
        return x, {"iters": list(r.iters)[:nrhs], "err": list(r.err)[:nrhs],
                   "restarts": list(r.restarts)[:nrhs], "refinements": r.refinements,
                   "ok": bool(r.ok)}

    def amgApply(self, r, precision=0):
        """One V-cycle z = M^-1 r of the pressure operator's AMG hierarchy
        (precision 1: the FP32 V-cycle of the mixed mode)."""
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty_like(r)
        _check(_lib.cfsa_ns_amg_apply(self._h, r, z, int(precision)))
        return z

    def amgInfo(self):
        cap = 32
        rows, nnz = (C.c_int64 * cap)(), (C.c_int64 * cap)()
        ms, builds, levels = C.c_double(), C.c_int(), C.c_int()
        _check(_lib.cfsa_ns_amg_info(self._h, cap, C.cast(rows, _p), C.cast(nnz, _p),
                                     C.byref(ms), C.byref(builds), C.byref(levels)))
        L = levels.value
        return {"levels": L, "rows": list(rows)[:L], "nnz": list(nnz)[:L],
                "setup_ms": ms.value, "builds": builds.value}

    def momentumPoly(self):
        """The polynomial preconditioners of momentum_solver 2 and 3 for the
        current operator: M^-1 = (g0 I - g1 D^-1 A) D^-1 and, as `a`,
        M^-1 = (a0 I + a1 D^-1 A + a2 (D^-1 A)^2) D^-1."""
        out = np.zeros(5, np.float64)
        _check(_lib.cfsa_ns_momentum_poly(self._h, out))
        a = np.zeros(3, np.float64)
        _check(_lib.cfsa_ns_momentum_poly2(self._h, a))
        return {"rho": out[0], "lmax": out[1], "g0": out[2], "g1": out[3],
                "estimates": int(out[4]), "a": a}

    def elemInfo(self):
        out = np.zeros(4, np.int64)
        ms = C.c_double()
        _check(_lib.cfsa_ns_elem_info(self._h, out, C.byref(ms)))
        return {"patches": int(out[0]), "visits": int(out[1]), "node_entries": int(out[2]),
                "built": bool(out[3]), "setup_ms": ms.value}

    def timeMomentumSpmv(self, reps=20, precision=0):
        out = C.c_double()
        _check(_lib.cfsa_ns_time_momentum_spmv(self._h, int(reps), int(precision), C.byref(out)))
        return out.value

    def stream(self):
        out = _p()
        _check(_lib.cfsa_ns_get_stream(self._h, C.byref(out)))
        return out.value

    # --- boundary forces (NavierStokes.cpp:357-372, 841-965) ---
    def _forces(self, name):
        out = np.empty((self.sizes()["nv"], 3), np.float64)
        _check(getattr(_lib, f"cfsa_ns_compute_{name}")(self._h, out))
        return out

    def computePressureForces(self):
        return self._forces("pressure_forces")

    def computeShearStress(self):
        return self._forces("shear_stress")

    def computeTotalFluidForces(self):
        return self._forces("total_fluid_forces")

    def nFaces(self):
        out = C.c_int64()
        _check(_lib.cfsa_ns_num_faces(self._h, C.byref(out)))
        return out.value

    def getMeanFacePressure(self, fi):
        out = C.c_double()
        _check(_lib.cfsa_ns_get_mean_face_pressure(self._h, int(fi), C.byref(out)))
        return out.value

    def getPressureForceAtFace(self, fi):
        out = np.empty(3, np.float64)
        _check(_lib.cfsa_ns_get_pressure_force_at_face(self._h, int(fi), out))
        return out

    # --- adjoint (extension) ---
    def setViscosity(self, mu):
        _check(_lib.cfsa_ns_set_viscosity(self._h, float(mu)))

    def adjointForward(self, nsteps, g):
        g = np.ascontiguousarray(g, np.float64)
        u = np.empty_like(g)
        _check(_lib.cfsa_ns_adjoint_forward(self._h, int(nsteps), g, u.ctypes.data_as(_p)))
        return u

    def adjointBackward(self, lam=None):
        nu = self.sizes()["nu"]
        dg = np.empty((nu, 3), np.float64)
        dmu = C.c_double()
        keep = None
        if lam is not None:
            keep = np.ascontiguousarray(lam, np.float64)
        _check(_lib.cfsa_ns_adjoint_backward(
            self._h, None if keep is None else keep.ctypes.data_as(_p), dg, C.byref(dmu)))
        return dg, dmu.value

    def adjointGradient(self, nsteps, basis, coef):
        basis = np.ascontiguousarray(basis, np.float64)
        coef = np.ascontiguousarray(coef, np.float64)
        J, dmu = C.c_double(), C.c_double()
        dc = np.empty(len(coef), np.float64)
        _check(_lib.cfsa_ns_adjoint_gradient(self._h, int(nsteps), len(coef), basis, coef,
                                             C.byref(J), dc, C.byref(dmu)))
        return J.value, dc, dmu.value


class HostAmg:
    """The library's smoothed-aggregation set-up run on a host CSR matrix
    (cfsa_amg_build): levels, operators, prolongations and coefficients, for
    inspection and tests."""

    def __init__(self, rowptr, col, val, theta=0.06, max_coarse=1000):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        h = _p()
        _check(_lib.cfsa_amg_build(len(rowptr) - 1, rowptr, col, val, float(theta),
                                   int(max_coarse), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cfsa_amg_free(self._h)
            self._h = None

    def levels(self):
        L, nc, cj = C.c_int(), C.c_int64(), C.c_int()
        _check(_lib.cfsa_amg_num_levels(self._h, C.byref(L), C.byref(nc), C.byref(cj)))
        return L.value, nc.value, bool(cj.value)

    def matrix(self, level, which):
        """which: 'A', 'P' or 'R'; returns (rowptr, col, val, shape)."""
        w = {"A": 0, "P": 1, "R": 2}[which]
        r, c, z = C.c_int64(), C.c_int64(), C.c_int64()
        _check(_lib.cfsa_amg_level_size(self._h, int(level), w, C.byref(r), C.byref(c),
                                        C.byref(z)))
        rp = np.empty(r.value + 1, np.int32)
        ci = np.empty(z.value, np.int32)
        v = np.empty(z.value, np.float64)
        _check(_lib.cfsa_amg_level_get(self._h, int(level), w, rp, ci, v))
        return rp, ci, v, (r.value, c.value)

    def params(self, level):
        rho, lmax = C.c_double(), C.c_double()
        coef = np.empty(3, np.float64)
        n = self.matrix(level, "A")[3][0]
        agg = np.full(n, -2, np.int32)
        _check(_lib.cfsa_amg_level_params(self._h, int(level), C.byref(rho), C.byref(lmax), coef,
                                          agg.ctypes.data_as(_p)))
        return {"rho": rho.value, "lmax": lmax.value, "coef": coef, "agg": agg}

    def coarseInverse(self):
        _, nc, _ = self.levels()
        inv = np.empty(nc * nc, np.float64)
        _check(_lib.cfsa_amg_coarse_inverse(self._h, inv))
        return inv.reshape(nc, nc)


class HostElemPatches:
    """The element patches the FP32 momentum operator uses with elem_spmv = 1
    (cfsa_elem_patches_build), built on the host from a mesh, for inspection
    and tests."""

    def __init__(self, nv, tets, ten, xyz):
        tets = np.ascontiguousarray(tets, np.int32)
        ten = np.ascontiguousarray(ten, np.int32)
        xyz = np.ascontiguousarray(xyz, np.float64)
        ne = int(ten.max()) + 1 if ten.size else 0
        h = _p()
        _check(_lib.cfsa_elem_patches_build(int(nv), len(tets), ne, tets, ten, xyz, C.byref(h)))
        self._h = h
        sz = np.zeros(8, np.int64)
        _check(_lib.cfsa_elem_patches_sizes(self._h, sz))
        npatch, nvisit, nnodes, nrows, nstarts = (int(v) for v in sz[:5])
        self.max_rows, self.max_nodes, self.max_slots = (int(v) for v in sz[5:])
        self.node_off = np.empty(npatch + 1, np.int32)
        self.nodes = np.empty(nnodes, np.int32)
        self.row_off = np.empty(npatch + 1, np.int32)
        self.rows = np.empty(nrows, np.int32)
        self.starts = np.empty(nstarts, np.uint16)
        self.visit_off = np.empty(npatch + 1, np.int32)
        self.vtet = np.empty(nvisit, np.int32)
        self.vstat = np.empty((nvisit, 8), np.uint32)
        self.tet_order = np.empty(len(tets), np.int32)
        arrs = [self.node_off, self.nodes, self.row_off, self.rows, self.starts, self.visit_off,
                self.vtet, self.vstat, self.tet_order]
        _check(_lib.cfsa_elem_patches_get(self._h, *[a.ctypes.data_as(_p) for a in arrs]))
        self.npatch = npatch

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cfsa_elem_patches_free(self._h)
            self._h = None

    def _words(self, k):
        w = self.vstat[:, 4 * k:4 * k + 4].astype(object)
        return w[:, 0] | (w[:, 1] << 32) | (w[:, 2] << 64) | (w[:, 3] << 96)

    def local_ids(self):
        """visits x 10 local node ids (11 bits each at bit 11 l of words 0-3)"""
        big = self._words(0)
        return np.array([[(int(v) >> (11 * l)) & 0x7FF for l in range(10)] for v in big],
                        np.int64).reshape(-1, 10)

    def slots(self):
        """visits x 10 LDS slots of the owned nodes (12 bits each at bit 12 l of
        words 4-7; 0 for a node the patch does not own)"""
        big = self._words(1)
        return np.array([[(int(v) >> (12 * l)) & 0xFFF for l in range(10)] for v in big],
                        np.int64).reshape(-1, 10)

    def tet_numbers(self):
        """the tet of each visit in patch numbering (tet_order): bits 0-17 at bit
        110 of words 0-3, bits 18-25 at bit 120 of words 4-7"""
        a = self.vstat[:, 3].astype(np.int64) >> 14
        b = self.vstat[:, 7].astype(np.int64) >> 24
        return a | (b << 18)

    def tets(self):
        """the mesh tet of each visit, from its record"""
        return self.tet_order[self.tet_numbers()].astype(np.int64)
