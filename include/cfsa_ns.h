/*
 * C ABI of the MI355X Navier-Stokes PISO library (libcfsa_ns.so).
 *
 * The library replaces the reference's C++ class NavierStokesSolver
 * (include/FEM/NavierStokes.hpp:53-193 of sudo-shaka/Cell-Fluid-Structure-
 * Analysis) and the parts of Mesh (include/Mesh/Mesh.hpp:35-220) that the
 * solver reads.  Every entry point is extern "C", takes plain pointers and
 * sizes, never throws, and returns CFSA_OK or an error code; the message of
 * the last failure on the calling thread is returned by cfsa_last_error().
 *
 * Array layouts match the reference's std::vector storage:
 *   velocity   double[nu][3]  (P1 vertices 0..nv-1, then P2 edge nodes;
 *                              std::vector<Eigen::Vector3d>, NavierStokes.hpp:66)
 *   pressure   double[nv]     (std::vector<double>, NavierStokes.hpp:68)
 *   labels     uint8 FluidBCType values Wall=0 Inlet=1 Outlet=2 Internal=3
 *              Undefined=4 (NavierStokes.hpp:13-19)
 * Handles are not thread safe (the reference solver is single threaded).
 */
#ifndef CFSA_NS_H
#define CFSA_NS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFSA_ABI_VERSION 4

enum cfsa_status {
  CFSA_OK = 0,
  CFSA_ERR_NOT_INIT = 1,         /* pisoStep before initialize (NavierStokes.cpp:375-378) */
  CFSA_ERR_MOMENTUM_NOCONV = 2,  /* BiCGSTAB NoConvergence (NavierStokes.cpp:511-515) */
  CFSA_ERR_PRESSURE_NOCONV = 3,  /* NavierStokes.cpp:646-649 */
  CFSA_ERR_NAN = 4,              /* hasNans() after the step (NavierStokes.hpp:116-138) */
  CFSA_ERR_INVALID = 5,          /* bad argument */
  CFSA_ERR_DEVICE = 6,           /* HIP runtime failure */
  CFSA_ERR_IO = 7                /* file could not be read or written */
};

/* enum values of the reference (NavierStokes.hpp:13-31) */
enum cfsa_fluid_bc { CFSA_WALL = 0, CFSA_INLET = 1, CFSA_OUTLET = 2, CFSA_INTERNAL = 3, CFSA_UNDEFINED = 4 };
enum cfsa_outlet_type { CFSA_OUTLET_DIRICHLET = 0, CFSA_OUTLET_NEUMANN = 1, CFSA_OUTLET_UNDEFINED = 2 };
enum cfsa_inlet_type { CFSA_INLET_UNIFORM = 0, CFSA_INLET_PULSATILE = 1, CFSA_INLET_UNDEFINED = 2 };

typedef struct cfsa_mesh cfsa_mesh;
typedef struct cfsa_ns cfsa_ns;

/* Solver options.  The defaults reproduce the reference exactly; everything
 * else is an extension (tolerances of NavierStokes.cpp:499-500,634-635). */
typedef struct cfsa_solver_options {
  double momentum_tol;   /* 1e-6 */
  double pressure_tol;   /* 1e-8 */
  int max_iter;          /* 1000 */
  int pressure_solver;   /* 0 BiCGSTAB (reference), 1 Jacobi PCG, 2 CG preconditioned by one
                            V-cycle of a smoothed-aggregation AMG hierarchy of L_bc (built on
                            the host from the current values).  The CG is FP64 in every
                            precision mode; with precision 1 the V-cycle runs in FP32 */
  int precision;         /* 0 FP64, 1 FP32 inner solves in FP64 iterative refinement for the
                            momentum and pressure solves, 2 for the momentum solve only */
  double inner_tol;      /* relative tolerance of each FP32 inner solve, 1e-4 */
  int max_refine;        /* refinement sweeps, 40 */
  int check_every;       /* Krylov iterations between host convergence reads, 8 */
  int timing;            /* 0 off, 1 per-phase events, 2 also per-SpMV events */
  int stop_rule;         /* 0 |r| <= tol |b| (reference), 1 |D^-1 r| <= tol |D^-1 b| with
                            D the Jacobi diagonal (for adjoint / finite-difference runs,
                            where the 1e10 penalty rows dominate |b|) */
  int assembly;          /* 0 row gather in ascending tet order (sums every entry in the
                            reference's order: bit-exact matrices), six rows per
                            wavefront; 1 graph-coloured element scatter (same values up
                            to rounding order); 2 the row gather with one thread per row
                            (same bits as 0; 0 falls back to it for rows too long for
                            LDS) */
  int spmv_layout;       /* Krylov SpMV: 0 auto (momentum CSR-stream, pressure sliced-ELL),
                            1 CSR-stream, 2 sliced-ELL; every layout gives the same bits.
                            The AMG hierarchy's coarser operators (pressure_solver 2) are
                            CSR-stream unless a row exceeds 2048 nonzeros, or sliced-ELL
                            with 2, as set when the hierarchy is built */
  int momentum_solver;   /* 0 BiCGSTAB (reference), 1 Jacobi PCG (the momentum matrix
                            M/dt + K + penalty diagonal is symmetric positive definite),
                            2 CG preconditioned by the degree-1 Chebyshev polynomial in
                            D^-1 A, M^-1 = (g0 I - g1 D^-1 A) D^-1 (two products per
                            iteration, half the iterations; g0, g1 from a power-iteration
                            estimate of the spectral radius of D^-1 A, refreshed with the
                            lhs: cfsa_ns_momentum_poly), 3 CG preconditioned by the
                            degree-2 polynomial of the same Chebyshev iteration, M^-1 =
                            (a0 I + a1 D^-1 A + a2 (D^-1 A)^2) D^-1 (three products per
                            iteration, about two thirds of degree 1's iterations;
                            cfsa_ns_momentum_poly2) */
  int check_every_1;     /* the same for single right-hand-side (pressure) solves, 32: their
                            iterations are short, so host reads cost relatively more */
  int graphs;            /* 1 (default): each chunk of check_every Krylov iterations is
                            captured once as a HIP graph and replayed; 0: launched eagerly.
                            Same results either way.  Chunks with per-SpMV timing
                            (timing = 2) always run eagerly. */
  int initial_guess;     /* 0 (reference): the momentum solve starts from zero; 1: the
                            mixed-precision momentum refinement starts from u^n */
  int reuse_pressure_correction; /* 1 (default): with several PISO correctors, correctors
                            after the first reuse the first one's pressure correction p'.
                            The reference re-solves (NavierStokes.cpp:392-405), but every
                            input of that solve (u*, L_bc, right-hand side) is unchanged
                            between correctors, so the result is the same bits; p still
                            accumulates relax_p p' per corrector.  0: re-solve. */
  double restart_scale;  /* 1 (Eigen's rule): BiCGSTAB restarts when |rho| < restart_scale
                            eps^2 |r0|^2 (Eigen's restart branch, restated from its published
                            algorithm).  Larger values make that branch reachable on small
                            systems, for tests; > 0. */
  int amg_reuse;         /* pressure_solver 2: changes of the L_bc values (moved mesh, new
                            labels) for which the AMG hierarchy is kept, the finest level
                            always smoothing with the current operator; 0 (default) rebuilds
                            on every change.  Re-assembly to identical values never rebuilds. */
  int elem_spmv;         /* precision 1 / 2: 1 applies the FP32 momentum operator of the inner
                            solves tet by tet (patches of tets, from the per-tet volume and
                            gradients, no assembled FP32 copy read); 0 (default) uses the
                            assembled operator's FP32 copy.  The FP64 path is unaffected. */
} cfsa_solver_options;

typedef struct cfsa_step_stats {
  int momentum_iters[3];
  double momentum_err[3];
  int momentum_refinements;
  int n_pressure;
  int pressure_iters[8];
  double pressure_err[8];
  int pressure_refinements[8];
  double ms_conv, ms_rhs, ms_momentum, ms_pressure, ms_correct, ms_total;
  int spmv_launches;
  double spmv_ms;
  int pressure_solves;  /* pressure solves run this step */
} cfsa_step_stats;

typedef struct cfsa_solve_result {
  int iters[3];
  double err[3];
  int restarts[3];
  int refinements;
  int inner_iters;
  int ok;
} cfsa_solve_result;

const char *cfsa_last_error(void);
/* version of this C ABI (no counterpart in the reference: lets a binding check
 * it was built against the same struct layouts) */
int cfsa_abi_version(void);

/* ------------------------------------------------------------------------
 * Mesh (reference Mesh, include/Mesh/Mesh.hpp:35-220)
 * ---------------------------------------------------------------------- */
/* Mesh::structuredRectangularPrism (Mesh.hpp:87-93, Mesh.cpp:146-225) */
int cfsa_mesh_structured_prism(double length, double width, double height, int nx, int ny,
                               int nz, cfsa_mesh **out);
/* Cylinder along +x; replaces Mesh::fromPolyhedron(Polyhedron::cylendar(...),
 * max_element_size) (src/main.cpp:25-33), which needs TetGen. */
int cfsa_mesh_cylinder(double length, double radius, int divisions, double max_volume,
                       cfsa_mesh **out);
/* Mesh::fromMshFile (Mesh.hpp:82-86, Mesh.cpp:16-144), Gmsh MSH 4.1 ASCII */
int cfsa_mesh_from_msh(const char *path, cfsa_mesh **out);
/* vertices + tets from any source; runs the same connectivity/geometry/P2 setup */
int cfsa_mesh_from_arrays(int64_t nv, const double *coords, int64_t nt, const int32_t *tets,
                          cfsa_mesh **out);
/* Mesh::setupBoundaryConditions (Mesh.hpp:94-96, Mesh.cpp:325-417) */
int cfsa_mesh_setup_boundary_conditions(cfsa_mesh *m, const double dir[3], double coverage);
/* Mesh::setFluidVertexBC for all vertices, then setP2BoundariesFromP1Boundaries
 * (Mesh.hpp:180-215) */
int cfsa_mesh_set_p1_bc(cfsa_mesh *m, const uint8_t *p1_bc);
/* move vertices and Mesh::updateGeometry (Mesh.hpp:98-101) */
int cfsa_mesh_update_vertices(cfsa_mesh *m, const double *coords);
/* out[0..4] = nVertices, nTets, getNumberOfEdgeNodes, nFaces, nv + ne */
int cfsa_mesh_sizes(const cfsa_mesh *m, int64_t out[5]);
int cfsa_mesh_get_vertices(const cfsa_mesh *m, double *out);           /* nv x 3 */
int cfsa_mesh_get_tets(const cfsa_mesh *m, int32_t *out);              /* nt x 4 */
int cfsa_mesh_get_volumes(const cfsa_mesh *m, double *out);            /* nt */
int cfsa_mesh_get_gradients(const cfsa_mesh *m, double *out);          /* nt x 4 x 3 */
int cfsa_mesh_get_tet_edge_nodes(const cfsa_mesh *m, int32_t *out);    /* nt x 6 */
int cfsa_mesh_get_edge_nodes(const cfsa_mesh *m, double *out);         /* ne x 3 */
int cfsa_mesh_get_edge_endpoints(const cfsa_mesh *m, int32_t *out);    /* ne x 2 */
int cfsa_mesh_get_p1_bc(const cfsa_mesh *m, uint8_t *out);             /* nv */
int cfsa_mesh_get_p2_bc(const cfsa_mesh *m, uint8_t *out);             /* ne */
/* Mesh::getFaces (include/Mesh/Mesh.hpp:123, Face :23-33; built by
 * Mesh::buildConnectivity, src/Mesh/Mesh.cpp:494-588): faces in sorted-triple
 * order: verts nf x 3; tets nf x 4 (tet_a, tet_b, face_id_a, face_id_b); geom
 * nf x 7 (normal, area, center); any may be NULL */
int cfsa_mesh_get_faces(const cfsa_mesh *m, int32_t *verts, int32_t *tets, double *geom);
int cfsa_mesh_get_tet_faces(const cfsa_mesh *m, int32_t *faceids, int32_t *neighbors);
int cfsa_mesh_get_min_edge_length(const cfsa_mesh *m, double *out);
/* io::exportToVtk(filename, mesh, scale) (VtkExport.cpp:49-84) */
int cfsa_mesh_export_vtk(const cfsa_mesh *m, const char *path, double scale);
void cfsa_mesh_destroy(cfsa_mesh *m);

/* ------------------------------------------------------------------------
 * Navier-Stokes PISO solver (reference NavierStokesSolver)
 * ---------------------------------------------------------------------- */
int cfsa_ns_create(cfsa_ns **out);
void cfsa_ns_destroy(cfsa_ns *h);
/* NavierStokesSolver::initialize(mesh_ptr, fluid) (NavierStokes.hpp:91-92,
 * NavierStokes.cpp:10-68).  The mesh handle must outlive the solver: like the
 * reference's shared_ptr<Mesh>, it is re-read by build_sparse_matrices. */
int cfsa_ns_initialize(cfsa_ns *h, const cfsa_mesh *mesh, double density, double viscosity);
/* Same from raw arrays (for FFI callers that own their mesh). */
int cfsa_ns_initialize_arrays(cfsa_ns *h, int64_t nv, const double *coords, int64_t nt,
                              const int32_t *tets, int64_t ne, const int32_t *tet_edge_nodes,
                              const uint8_t *p1_bc, const uint8_t *p2_bc, double density,
                              double viscosity);
/* Push moved vertices and/or new labels (NULL = unchanged).  Labels and
 * element geometry take effect at the next step, the operators M, K, G, L at
 * the next build_sparse_matrices, as in the reference where they are read
 * live from the shared mesh. */
int cfsa_ns_update_mesh(cfsa_ns *h, const double *coords, const uint8_t *p1_bc,
                        const uint8_t *p2_bc);
/* NavierStokesSolver::build_sparse_matrices (NavierStokes.hpp:97-102) */
int cfsa_ns_build_sparse_matrices(cfsa_ns *h);
/* NavierStokesSolver::pisoStep (NavierStokes.cpp:374-414); CFSA_OK == true */
int cfsa_ns_piso_step(cfsa_ns *h);
/* NavierStokesSolver::simpleStep (NavierStokes.cpp:968-1009) */
int cfsa_ns_simple_step(cfsa_ns *h);
/* the parts of the step, public in the reference: solveMomentumPredictor
 * (NavierStokes.cpp:416-524: u* from u and p), solvePressurePoisson (:586-668:
 * p' from u*, p += relax_p p', outlet reset, normalisation), correctVelocity
 * (:702-771: u from u* and p'), normalizePressire (:804-813).  Status codes as
 * cfsa_ns_piso_step's; the time does not advance. */
int cfsa_ns_solve_momentum_predictor(cfsa_ns *h);
int cfsa_ns_solve_pressure_poisson(cfsa_ns *h);
int cfsa_ns_correct_velocity(cfsa_ns *h);
int cfsa_ns_normalize_pressure(cfsa_ns *h);
/* NavierStokesSolver::solvePressureBiCGSTAB (NavierStokes.cpp:773-802): x = L^-1
 * rhs by Jacobi BiCGSTAB on the Poisson matrix L (no penalty rows), tolerance
 * 1e-8, 1000 iterations, from zero; CFSA_ERR_PRESSURE_NOCONV when it does not
 * converge (the reference returns false); rhs and x hold nv values; res may be
 * NULL */
int cfsa_ns_solve_pressure_bicgstab(cfsa_ns *h, const double *rhs, double *x,
                                    cfsa_solve_result *res);

/* setters, NavierStokes.hpp:167-186 */
int cfsa_ns_set_mean_inlet_velocity(cfsa_ns *h, const double v[3]);
/* per-velocity-node inlet values nu x 3 (extension; NULL restores uniform) */
int cfsa_ns_set_inlet_profile(cfsa_ns *h, const double *g);
int cfsa_ns_set_outlet_type(cfsa_ns *h, int type);
int cfsa_ns_set_outlet_pressure(cfsa_ns *h, double p);
int cfsa_ns_set_inlet_type(cfsa_ns *h, int type);
int cfsa_ns_set_number_of_corrections(cfsa_ns *h, int n);
int cfsa_ns_set_relax_u(cfsa_ns *h, double r);
int cfsa_ns_set_relax_p(cfsa_ns *h, double r);
int cfsa_ns_set_reference_pressure_node(cfsa_ns *h, int node);
int cfsa_ns_set_dt(cfsa_ns *h, double dt);
int cfsa_ns_set_tolerance(cfsa_ns *h, double tol);
int cfsa_ns_set_time(cfsa_ns *h, double t);
int cfsa_ns_set_options(cfsa_ns *h, const cfsa_solver_options *opt);
int cfsa_ns_get_options(const cfsa_ns *h, cfsa_solver_options *opt);

/* getters, NavierStokes.hpp:146-189 */
/* out[0..5] = nv, nt, ne, nu, nnz of the P2 pattern, nnz of the P1 pattern */
int cfsa_ns_get_sizes(const cfsa_ns *h, int64_t out[6]);
int cfsa_ns_get_velocity(cfsa_ns *h, double *out);         /* nu x 3 */
int cfsa_ns_set_velocity(cfsa_ns *h, const double *in);    /* nu x 3 */
int cfsa_ns_get_velocity_star(cfsa_ns *h, double *out);    /* nu x 3 */
int cfsa_ns_get_pressure(cfsa_ns *h, double *out);         /* nv */
int cfsa_ns_set_pressure(cfsa_ns *h, const double *in);    /* nv */
int cfsa_ns_get_pressure_correction(cfsa_ns *h, double *out);
int cfsa_ns_get_velocity_at_node(cfsa_ns *h, int64_t node, double out[3]);
/* Read-back of the reference's host-side fields (getVelocity / getPressure
 * return host vectors, NavierStokes.hpp:188-189) without stalling the solver:
 * u (nu x 3) and p (nv) are snapshotted on the device and copied to u_out /
 * p_out (either may be NULL) on a second stream, overlapping later steps when
 * the buffers are pinned (cfsa_host_alloc); cfsa_ns_wait_fields blocks until
 * the copies are done.  The buffers must stay valid until then. */
int cfsa_ns_get_fields_async(cfsa_ns *h, double *u_out, double *p_out);
int cfsa_ns_wait_fields(cfsa_ns *h);
/* pinned (page-locked) host memory for cfsa_ns_get_fields_async */
int cfsa_host_alloc(size_t bytes, void **out);
void cfsa_host_free(void *p);
int cfsa_ns_get_pressure_at_node(cfsa_ns *h, int64_t node, double *out);
/* which: 0 time, 1 dt, 2 viscosity, 3 density, 4 tolerance, 5 outlet pressure */
int cfsa_ns_get_scalar(const cfsa_ns *h, int which, double *out);
int cfsa_ns_get_reference_pressure_node(const cfsa_ns *h, int *out);
int cfsa_ns_has_nans(cfsa_ns *h, int *out);
int cfsa_ns_get_step_stats(const cfsa_ns *h, cfsa_step_stats *out);
/* io::exportToVtk(filename, ns_solver, scale) (VtkExport.cpp:238-268) */
int cfsa_ns_export_vtk(cfsa_ns *h, const char *path, double scale);

/* boundary forces per P1 vertex (nv x 3): computePressureForces,
 * computeShearStress, computeTotalFluidForces (NavierStokes.cpp:841-965) */
int cfsa_ns_compute_pressure_forces(cfsa_ns *h, double *out);
int cfsa_ns_compute_shear_stress(cfsa_ns *h, double *out);
int cfsa_ns_compute_total_fluid_forces(cfsa_ns *h, double *out);
/* faces in sorted-triple order: getMeanFacePressure / getPressureForceAtFace
 * (NavierStokes.cpp:357-372); cfsa_ns_num_faces gives their count */
int cfsa_ns_num_faces(cfsa_ns *h, int64_t *out);
int cfsa_ns_get_mean_face_pressure(cfsa_ns *h, int64_t face, double *out);
int cfsa_ns_get_pressure_force_at_face(cfsa_ns *h, int64_t face, double out[3]);

/* ---- parity / diagnostics (operators the reference builds internally) ---- */
/* inv_lumped_mass_ (NavierStokes.cpp:290-306), nu entries */
int cfsa_ns_get_inv_lumped_mass(cfsa_ns *h, double *out);
/* matrix which: 0 M, 1 K, 2 momentum lhs M/dt+K with penalties, 3 L, 4 L with
 * pressure penalties, 5 Gx, 6 Gy, 7 Gz.  rowptr n+1, col/val nnz; any may be NULL */
int cfsa_ns_get_matrix(cfsa_ns *h, int which, int32_t *rowptr, int32_t *col, double *val);
int cfsa_ns_get_num_colours(const cfsa_ns *h, int *out);
int cfsa_ns_get_colouring(const cfsa_ns *h, int32_t *colour);  /* nt */
int cfsa_ns_get_incidence(cfsa_ns *h, int32_t *ptr, int32_t *tet); /* nu+1, 10 nt */
/* computeAdvectionRHS on the current velocity (NavierStokes.cpp:526-584) */
int cfsa_ns_compute_advection(cfsa_ns *h, double *out);
/* computeDivergence(u, div) (NavierStokes.cpp:670-700); u is nu x 3 */
int cfsa_ns_compute_divergence(cfsa_ns *h, const double *u, double *out);
/* y = A x; nrhs 1 or 3 (interleaved); for which 5..7 y = [Gx x, Gy x, Gz x];
 * which 8: the FP32 momentum operator of the mixed-precision inner solves
 * (nrhs 3; the assembled operator's FP32 copy, or its element patches with
 * elem_spmv = 1), x rounded to FP32 and y widened back */
int cfsa_ns_apply_matrix(cfsa_ns *h, int which, const double *x, double *y, int nrhs);
/* cfsa_ns_apply_matrix without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  x and y
 * are staged in a static 64 MiB device array per GPU, so x and y together must
 * fit in 8 Mi doubles (x padded to a multiple of 32; CFSA_ERR_INVALID beyond).
 * Every call on one GPU shares that GPU's array, which keeps the last call's
 * contents: calls that overlap on one GPU race.  The operators must not need
 * a fresh device allocation: true after initialize unless the precision or
 * spmv_layout option changed since (then one allocating call first; otherwise
 * CFSA_ERR_INVALID).  Same results as cfsa_ns_apply_matrix, bit for bit. */
int cfsa_ns_apply_matrix_noalloc(cfsa_ns *h, int which, const double *x, double *y,
                                 int nrhs);
/* solve matrix 2 (nrhs 1 or 3) or 4 (nrhs 1) with method 0 BiCGSTAB / 1 PCG /
 * 2 AMG-preconditioned CG (matrix 4 only) / 3 and 4 CG with the polynomial
 * preconditioners of momentum_solver 2 and 3 (matrix 2, nrhs 3) */
int cfsa_ns_krylov_solve(cfsa_ns *h, int which, int method, int nrhs, const double *b,
                         double *x, double tol, int maxit, int precision,
                         cfsa_solve_result *res);
/* AMG preconditioner of the pressure operator (extension, no counterpart in the
 * reference): one V-cycle z = M^-1 r (nv values each), building the hierarchy
 * from the current L_bc if needed; precision 0 FP64, 1 the FP32 V-cycle of the
 * mixed precision = 1 mode */
int cfsa_ns_amg_apply(cfsa_ns *h, const double *r, double *z, int precision);
/* the solver's current hierarchy: *levels (0 if none built), rows and nonzeros
 * of up to cap levels, set-up time of the last build (ms), builds so far; any
 * output pointer may be NULL */
int cfsa_ns_amg_info(cfsa_ns *h, int cap, int64_t *rows, int64_t *nnz, double *setup_ms,
                     int *builds, int *levels);
/* the polynomial preconditioner of momentum_solver 2 (no counterpart in the
 * reference) for the current momentum operator, estimating the spectral
 * radius of D^-1 A first if the operator changed: out = rho estimate, lmax =
 * 1.1 rho, g0, g1 (M^-1 = (g0 I - g1 D^-1 A) D^-1), estimates so far */
int cfsa_ns_momentum_poly(cfsa_ns *h, double out[5]);
/* the degree-2 polynomial preconditioner of momentum_solver 3 (no counterpart in
 * the reference) for the current momentum operator: out = a0, a1, a2 with M^-1 =
 * (a0 I + a1 D^-1 A + a2 (D^-1 A)^2) D^-1, from the same estimate as
 * cfsa_ns_momentum_poly */
int cfsa_ns_momentum_poly2(cfsa_ns *h, double out[3]);
/* the element patches of the FP32 momentum operator (elem_spmv = 1; no
 * counterpart in the reference): out = patches, visits, node-table entries,
 * 1 if built (all 0 before the first mixed-precision solve that uses them);
 * setup_ms (may be NULL) the host set-up time */
int cfsa_ns_elem_info(cfsa_ns *h, int64_t out[4], double *setup_ms);
/* Host-side hierarchy of any square CSR matrix (sorted columns per row), for
 * inspection and tests: the set-up the solver runs (strength threshold theta,
 * at most max_coarse rows on the coarsest level) */
typedef struct cfsa_amg cfsa_amg;
int cfsa_amg_build(int64_t n, const int32_t *rowptr, const int32_t *col, const double *val,
                   double theta, int max_coarse, cfsa_amg **out);
int cfsa_amg_free(cfsa_amg *h);
/* levels (>= 1), rows of the coarsest level, 1 if the coarsest level is a
 * Jacobi step instead of the dense inverse */
int cfsa_amg_num_levels(const cfsa_amg *h, int *levels, int64_t *coarse_n, int *coarse_jacobi);
/* which: 0 the level operator A, 1 the prolongation P (rows x cols = n x nc),
 * 2 the restriction R = P^T; P and R exist on every level but the last */
int cfsa_amg_level_size(const cfsa_amg *h, int level, int which, int64_t *rows, int64_t *cols,
                        int64_t *nnz);
int cfsa_amg_level_get(const cfsa_amg *h, int level, int which, int32_t *rowptr, int32_t *col,
                       double *val);
/* spectral radius estimate of D^-1 A, Chebyshev upper bound lmax, coefficients
 * (c0, c1, c2) and, if agg is not NULL, the aggregate of each row (-1: none;
 * empty on the last level) */
int cfsa_amg_level_params(const cfsa_amg *h, int level, double *rho, double *lmax,
                          double coef[3], int32_t *agg);

/* ---- element patches of the momentum operator (elem_spmv), for inspection
 * and tests: no counterpart in the reference, whose operator is only the
 * assembled matrix (NavierStokes.cpp:460-495).  Rows (velocity nodes, the nv
 * vertices then the ne edge nodes) are cut into patches in the Morton order of
 * the node positions; a patch owns its rows and visits every tet touching one
 * of them. ---- */
typedef struct cfsa_elem_patches cfsa_elem_patches;
/* tets nt x 4, tet edge nodes nt x 6 (0..ne-1), vertex coordinates nv x 3 */
int cfsa_elem_patches_build(int64_t nv, int64_t nt, int64_t ne, const int32_t *tets,
                            const int32_t *ten, const double *xyz, cfsa_elem_patches **out);
int cfsa_elem_patches_free(cfsa_elem_patches *h);
/* out: patches, visits, node-table entries, owned rows (= nv + ne), slot-start
 * entries (owned rows + patches), largest owned rows / nodes / slots of a patch */
int cfsa_elem_patches_sizes(const cfsa_elem_patches *h, int64_t out[8]);
/* node_off / row_off / visit_off: patches + 1 offsets; nodes: each patch's node
 * table (owned rows first); rows: owned rows; starts: per patch owned rows + 1
 * slot starts; vtet: the tet of each visit; vstat: 8 words per visit, words
 * 0-3 the 10 local node ids (12 bits each, id l at bit 12 l of the 128-bit
 * little-endian value), words 4-6 the 10 slot ranks (8 bits each, 255 for a
 * node the patch does not own), word 7 the tet in the patches' numbering;
 * tet_order (nt): tet j of that numbering is mesh tet tet_order[j].  Any pointer
 * may be NULL. */
int cfsa_elem_patches_get(const cfsa_elem_patches *h, int32_t *node_off, int32_t *nodes,
                          int32_t *row_off, int32_t *rows, uint16_t *starts, int32_t *visit_off,
                          int32_t *vtet, uint32_t *vstat, int32_t *tet_order);
/* dense (pseudo-)inverse of the coarsest operator, row-major coarse_n^2 */
int cfsa_amg_coarse_inverse(const cfsa_amg *h, double *inv);

/* mean duration (ms) of the momentum-matrix SpMV kernel over `reps` launches */
int cfsa_ns_time_momentum_spmv(cfsa_ns *h, int reps, int precision, double *ms);
/* the hipStream_t all solver work is issued on */
int cfsa_ns_get_stream(const cfsa_ns *h, void **stream);

/* ------------------------------------------------------------------------
 * Discrete adjoint of the PISO step (extension; no counterpart in the
 * reference).  Dirichlet outlet and n_corrections >= 1.  Linear solves use the
 * solver options (use stop_rule 1 and tight tolerances for gradients that
 * must match finite differences).
 * ---------------------------------------------------------------------- */
/* set the viscosity and re-assemble K (the reference takes it from Fluid at
 * initialize only) */
int cfsa_ns_set_viscosity(cfsa_ns *h, double viscosity);
/* Run nsteps PISO steps from rest (u = 0 with Inlet nodes at g, p = 0) with
 * per-node inlet values g (nu x 3), keeping a device checkpoint of (u, p) per
 * step; u_final (nu x 3) may be NULL. */
int cfsa_ns_adjoint_forward(cfsa_ns *h, int nsteps, const double *g, double *u_final);
/* Reverse sweep of the last adjoint_forward: given lambda = dL/du_final
 * (nu x 3; NULL means the outlet kinetic energy J = 1/2 sum_Outlet |u|^2),
 * returns dL/dg (nu x 3) and dL/dviscosity.  Each step's u* and p' are
 * recomputed from its checkpoint. */
int cfsa_ns_adjoint_backward(cfsa_ns *h, const double *lambda_u, double *dLdg,
                             double *dLdviscosity);
/* J = 1/2 sum over Outlet velocity nodes of |u|^2 after nsteps from rest with
 * g = sum_k coef[k] basis[k] (basis nk x nu x 3); returns J, dJ/dcoef[nk],
 * dJ/dviscosity. */
int cfsa_ns_adjoint_gradient(cfsa_ns *h, int nsteps, int nk, const double *basis,
                             const double *coef, double *J, double *dJdcoef,
                             double *dJdviscosity);

/* ------------------------------------------------------------------------
 * Configuration: config::loadConfig (include/IO/YamlConfig.hpp:8-157,
 * src/IO/YamlConfig.cpp:30-96,180-232) for the mesh, navier_stokes and output
 * sections; solid / dpm / fsi sections report only `enabled`.
 * ---------------------------------------------------------------------- */
typedef struct cfsa_sim_config {
  /* mesh (MeshConfig) */
  int mesh_type; /* 0 obj, 1 cylinder, 2 icosphere, 3 box (extension), 4 msh (extension) */
  char mesh_file[1024];
  double obj_scale;
  double cylinder_length, cylinder_radius;
  int cylinder_divisions;
  double icosphere_radius;
  int icosphere_recursion;
  double max_element_size;
  double flow_direction[3];
  double inlet_outlet_coverage;
  double box_size[3];
  int box_cells[3];
  /* navier_stokes (NavierStokesConfig) */
  int ns_enabled;
  double density, viscosity;
  int is_newtonian, is_laminar;
  double inlet_velocity[3];
  double outlet_pressure;
  int use_dirichlet_pressure;
  double dt, relax_p, relax_u;
  int n_corrections;   /* extension, default 1 */
  int pressure_solver; /* extension: 0 bicgstab, 1 cg, 2 amg (AMG-preconditioned CG) */
  int precision;       /* extension: 0 fp64, 1 mixed */
  double momentum_tol, pressure_tol; /* extension, 1e-6 / 1e-8 */
  int inlet_profile;   /* extension: 0 uniform, 1 parabolic */
  /* output (OutputConfig) */
  char output_prefix[512];
  int total_steps, output_frequency;
  double vtk_scale;
  int output_fluid, output_solid, output_dpm;
  /* other sections */
  int solid_enabled, dpm_enabled, fsi_enabled;
} cfsa_sim_config;

int cfsa_config_load(const char *path, cfsa_sim_config *out);
int cfsa_config_parse(const char *yaml_text, cfsa_sim_config *out);

#ifdef __cplusplus
}
#endif

#endif /* CFSA_NS_H */
