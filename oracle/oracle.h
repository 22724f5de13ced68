/*
 * CPU FP64 oracle of the reference Navier-Stokes PISO path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libcfsa_ns.so) never links or calls it.
 *
 * It restates, in plain sequential C++, the reference files
 *   src/Mesh/Mesh.cpp:146-225, 325-417, 448-677   (mesh, BCs, P2 numbering)
 *   include/Mesh/Mesh.hpp:198-215                  (P2 labels)
 *   src/FEM/NavierStokes.cpp:10-356, 374-839, 968-1024 (assembly, PISO)
 * plus the third-party semantics the reference relies on (Eigen >= 3.3,
 * absent here; restated from its published algorithms):
 *   SparseMatrix::setFromTriplets  duplicates summed in insertion order,
 *                                  explicit zeros kept, sorted indices;
 *   sparse * dense product         res(i) += A(i,j) * x(j), ascending j;
 *   dot / squaredNorm              2-wide packet reduction of an SSE2 build;
 *   BiCGSTAB + DiagonalPreconditioner (bicgstab.h), ConjugateGradient stop rule.
 */
#ifndef CFSA_ORACLE_H
#define CFSA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh orc_mesh;
typedef struct orc_ns orc_ns;

const char *orc_last_error(void);

/* ---- mesh ---- */
int orc_mesh_structured_prism(double L, double W, double H, int nx, int ny, int nz,
                              orc_mesh **out);
int orc_mesh_from_arrays(int64_t nv, const double *coords, int64_t nt, const int32_t *tets,
                         orc_mesh **out);
int orc_mesh_update_vertices(orc_mesh *m, const double *coords);
int orc_mesh_setup_bc(orc_mesh *m, const double dir[3], double coverage);
int orc_mesh_set_p1_bc(orc_mesh *m, const uint8_t *bc);
void orc_mesh_sizes(const orc_mesh *m, int64_t out[4]); /* nv nt ne nfaces */
void orc_mesh_get_vertices(const orc_mesh *m, double *out);
void orc_mesh_get_tets(const orc_mesh *m, int32_t *out);
void orc_mesh_get_volumes(const orc_mesh *m, double *out);
void orc_mesh_get_gradients(const orc_mesh *m, double *out);
void orc_mesh_get_tet_edge_nodes(const orc_mesh *m, int32_t *out);
void orc_mesh_get_edge_nodes(const orc_mesh *m, double *out);
void orc_mesh_get_p1_bc(const orc_mesh *m, uint8_t *out);
void orc_mesh_get_p2_bc(const orc_mesh *m, uint8_t *out);
void orc_mesh_get_faces(const orc_mesh *m, int32_t *verts, int32_t *tets, double *geom);
void orc_mesh_destroy(orc_mesh *m);

/* greedy first-fit tet colouring in tet order, conflict = shared vertex */
int orc_colour_tets(const int32_t *tets, int64_t nt, int64_t nv, int32_t *colour);

/* ---- solver ---- */
int orc_ns_create(const orc_mesh *m, double density, double viscosity,
                  const double inlet0[3], orc_ns **out);
void orc_ns_destroy(orc_ns *h);
void orc_ns_set_dt(orc_ns *h, double dt);
void orc_ns_set_mean_inlet(orc_ns *h, const double v[3]);
void orc_ns_set_inlet_profile(orc_ns *h, const double *g); /* nu x 3 or NULL */
void orc_ns_set_outlet(orc_ns *h, int type, double pressure);
void orc_ns_set_corrections(orc_ns *h, int n);
void orc_ns_set_relax(orc_ns *h, double relax_u, double relax_p);
void orc_ns_set_reference_node(orc_ns *h, int node);
void orc_ns_set_solver(orc_ns *h, double momentum_tol, double pressure_tol, int max_iter,
                       int pressure_solver);
/* boundary forces per vertex (nv x 3): 0 pressure, 1 shear, 2 total
 * (NavierStokes.cpp:841-965) */
void orc_ns_forces(orc_ns *h, int which, double *out);
/* extension: stop Krylov solves on |D^-1 r| <= tol |D^-1 b| (Jacobi-scaled) */
void orc_ns_set_stop_scaled(orc_ns *h, int on);
/* extension: momentum solver 0 BiCGSTAB (reference), 1 Jacobi CG */
void orc_ns_set_momentum_solver(orc_ns *h, int method);
int orc_ns_piso_step(orc_ns *h);   /* 0 ok, 2 momentum, 3 pressure, 4 NaN */
int orc_ns_simple_step(orc_ns *h);
void orc_ns_sizes(const orc_ns *h, int64_t out[6]); /* nv nt ne nu nnzA nnzP */
void orc_ns_get_velocity(const orc_ns *h, double *out);
void orc_ns_set_velocity(orc_ns *h, const double *in);
void orc_ns_get_velocity_star(const orc_ns *h, double *out);
void orc_ns_get_pressure(const orc_ns *h, double *out);
void orc_ns_set_pressure(orc_ns *h, const double *in);
void orc_ns_get_pressure_correction(const orc_ns *h, double *out);
void orc_ns_get_inv_lumped_mass(const orc_ns *h, double *out);
/* which: 0 M, 1 K, 2 lhs, 3 L, 4 L_bc, 5 Gx, 6 Gy, 7 Gz */
void orc_ns_get_matrix(orc_ns *h, int which, int32_t *rowptr, int32_t *col, double *val);
void orc_ns_compute_advection(orc_ns *h, double *out);
void orc_ns_compute_divergence(orc_ns *h, const double *u, double *out);
/* iterations: momentum[3], pressure per corrector (up to 8), and errors */
void orc_ns_get_stats(const orc_ns *h, int32_t *iters, double *errs);
/* Eigen-form BiCGSTAB (method 0) or Jacobi CG (1) on matrix 2 or 4 */
int orc_ns_krylov_solve(orc_ns *h, int which, int method, const double *b, double *x,
                        double tol, int maxit, int32_t *iters, double *err);

/* the same solvers on a caller-provided CSR matrix */
/* multiplier of eps^2 in the BiCGSTAB restart test (1 = Eigen's rule; a test
 * knob) and the restart count of the last BiCGSTAB solve of orc_csr_krylov or
 * orc_ns_krylov_solve */
void orc_set_restart_scale(double s);
int orc_last_restarts(void);
int orc_csr_krylov(int64_t n, const int32_t *rp, const int32_t *col, const double *val,
                   int method, const double *b, double *x, double tol, int maxit,
                   int32_t *iters, double *err);

/* Eigen SSE2 reduction order, exposed for the known-answer tests */
double orc_eigen_dot(const double *a, const double *b, int64_t n);
/* Order-robust inner product (Dot2) used by the Krylov restatements. */
double orc_dot2(const double *a, const double *b, int64_t n);
/* Inner products of the Krylov restatements: 0 Dot2 (default), 1 Eigen's
 * 2-wide packet order (a build-dependent order of the reference's Eigen). */
void orc_set_dot_mode(int mode);
int orc_get_dot_mode(void);
/* y = A x in Eigen's sparse*dense accumulation order (nrhs interleaved) */
void orc_csr_spmv(int64_t n, const int32_t *rp, const int32_t *col, const double *val,
                  const double *x, double *y, int nrhs);

/* ---- discrete adjoint of the PISO step (FP64 CPU reference gradient) ---- */
/* J = 1/2 sum over Outlet-labelled velocity nodes of |u|^2 after nsteps steps
 * started from u0 = 0, p0 = 0 (inlet nodes at the inlet values), with the
 * inlet values g = sum_k c_k basis_k (basis nk x nu x 3).  Returns J and
 * dJ/dc (nk) and dJ/dviscosity. */
int orc_adjoint_gradient(orc_ns *h, int nsteps, int nk, const double *basis,
                         const double *coef, double *J, double *dJdc, double *dJdmu);

#ifdef __cplusplus
}
#endif
#endif
