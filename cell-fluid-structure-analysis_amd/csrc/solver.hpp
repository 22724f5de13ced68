// GPU Navier-Stokes PISO solver: device state and orchestration behind the
// C ABI.  State, defaults and step order mirror NavierStokesSolver
// (include/FEM/NavierStokes.hpp:53-193, src/FEM/NavierStokes.cpp:10-1024).
#pragma once

#include <string>
#include <vector>

#include "dev.hpp"
#include "mesh.hpp"
#include "ops.hpp"

namespace cfsa {

enum Status {
  ST_OK = 0,
  ST_NOT_INIT = 1,
  ST_MOMENTUM_NOCONV = 2,
  ST_PRESSURE_NOCONV = 3,
  ST_NAN = 4,
  ST_INVALID = 5,
  ST_DEVICE = 6,
  ST_IO = 7,
};

// Non-zero while an allocation-free entry point runs on this thread: any
// device allocation or free a DBuf would make then throws instead.
inline thread_local int g_no_alloc = 0;
struct NoAllocScope {
  NoAllocScope() { ++g_no_alloc; }
  ~NoAllocScope() { --g_no_alloc; }
  NoAllocScope(const NoAllocScope &) = delete;
  NoAllocScope &operator=(const NoAllocScope &) = delete;
};
inline void check_may_allocate() {
  if (g_no_alloc)
    throw std::runtime_error(
        "device allocation needed inside an allocation-free entry point (run a step, "
        "build_sparse_matrices or an allocating entry point once first)");
}

// Device address of the allocation-free entry point's static scratch array
// on the current device (noalloc.hip).
double *noalloc_scratch_current_device();

template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) {
      check_may_allocate();
      (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (p && count == n) return;
    check_may_allocate();
    release();
    n = count;
    if (count) CFSA_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void upload(const T *h, size_t count) {
    alloc(count);
    if (count) CFSA_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void download(T *h, size_t count) const {
    if (count) CFSA_HIP(hipMemcpy(h, p, count * sizeof(T), hipMemcpyDeviceToHost));
  }
  void zero(hipStream_t s) {
    if (n) CFSA_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
};

struct SolverOptions {
  double momentum_tol = 1e-6;  // NavierStokes.cpp:500
  double pressure_tol = 1e-8;  // NavierStokes.cpp:635
  int max_iter = 1000;         // NavierStokes.cpp:499,634
  int pressure_solver = 0;     // 0 BiCGSTAB (reference), 1 Jacobi PCG, 2 AMG-preconditioned
                               // CG (FP64 in every precision mode)
  int momentum_solver = 0;     // 0 BiCGSTAB (reference), 1 Jacobi PCG (lhs is SPD), 2 CG
                               // with the degree-1 Chebyshev polynomial preconditioner,
                               // 3 with the degree-2 one
  int precision = 0;           // 0 FP64, 1 FP32 inner + FP64 refinement, 2 momentum only
  double inner_tol = 1e-4;     // relative tolerance of each FP32 inner solve
  int max_refine = 40;         // outer refinement sweeps (mixed)
  int check_every = 8;         // Krylov iterations between host status reads (3 rhs)
  int check_every_1 = 32;      // the same for single-rhs solves (pressure: short iterations)
  int timing = 0;              // 1: per-phase events; 2: also per-SpMV events
  int stop_scaled = 0;         // 1: stop on |D^-1 r| <= tol |D^-1 b| (Jacobi-scaled)
  int assembly = 0;            // 0 row gather (reference summation order), 1 coloured,
                               // 2 row gather with one thread per row (same bits as 0)
  int spmv_layout = 0;         // Krylov SpMV: 0 auto, 1 CSR-stream, 2 sliced-ELL
  int graphs = 1;              // replay Krylov iteration chunks as captured HIP graphs
  int initial_guess = 0;       // momentum solve from 0 (reference) or 1: from u^n (mixed)
  // Correctors after the first re-solve the first one's system (u*, L_bc and the
  // right-hand side do not change between correctors), so 1 reuses its p'
  // instead: the same bits, one pressure solve per step.  0 re-solves.
  int reuse_correction = 1;
  // multiplier of eps^2 in BiCGSTAB's restart test (1 = Eigen's rule)
  double restart_scale = 1.0;
  // mixed precision: 1 applies the FP32 momentum operator tet by tet from the
  // per-tet geometry (element patches) instead of its assembled CSR copy
  int elem_spmv = 0;
  // pressure_solver 2: changes of the L_bc values for which the AMG hierarchy is
  // kept (the finest level always smooths with the current operator); 0
  // rebuilds it whenever the values change
  int amg_reuse = 0;
};

struct SolveResult {
  int iters[3] = {0, 0, 0};
  double err[3] = {0, 0, 0};
  int restarts[3] = {0, 0, 0};
  int refinements = 0;
  int inner_iters = 0;
  bool ok = true;
};

struct StepStats {
  int momentum_iters[3];
  double momentum_err[3];
  int momentum_refinements;
  int n_pressure;
  int pressure_iters[8];
  double pressure_err[8];
  int pressure_refinements[8];
  double ms_conv, ms_rhs, ms_momentum, ms_pressure, ms_correct, ms_total;
  int spmv_launches;
  double spmv_ms;  // summed event time of momentum-matrix SpMV launches
  int pressure_solves;  // pressure solves run (correctors that reused p' excluded)
};

struct MeshInput {
  int64_t nv = 0, nt = 0, ne = 0;
  const double *coords = nullptr;     // nv x 3
  const int32_t *tets = nullptr;      // nt x 4
  const int32_t *ten = nullptr;       // nt x 6 edge-node ids (0..ne-1)
  const uint8_t *p1_bc = nullptr;     // nv
  const uint8_t *p2_bc = nullptr;     // ne
  // faces in sorted-triple order with the orientation the mesh chose when it
  // was built (nullptr: build them from coords)
  const std::vector<Face> *faces = nullptr;
};

template <class T>
struct KrylovWork {
  DBuf<T> r, r0, p, v, s, t, y, z;
  void alloc(size_t n) {
    r.alloc(n); r0.alloc(n); p.alloc(n); v.alloc(n);
    s.alloc(n); t.alloc(n); y.alloc(n); z.alloc(n);
  }
};

class NSSolver {
 public:
  NSSolver();
  ~NSSolver();

  // ---- reference state and defaults (NavierStokes.hpp:54-83) ----
  bool initialized = false;
  int reference_node = -1;
  double tolerance = 1e-12;
  int n_corrections = 1;
  double relax_u = 1.0, relax_p = 1.0, dt = 1e-4, time = 0.0;
  double density = 1000.0, viscosity = 0.001;
  double mean_inlet[3] = {0.0, 0.0, 0.0};
  int inlet_type = 0;     // InletType::Uniform
  double outlet_pressure = 0.0;
  int outlet_type = 0;    // OutletType::DirichletPressure
  bool inlet_profile = false;
  SolverOptions opt;
  StepStats last{};
  std::string last_error;

  // ---- sizes ----
  int nv = 0, nt = 0, ne = 0, nu = 0;
  int64_t nnzA = 0, nnzP = 0;
  int ncolours = 0;

  void initialize(const MeshInput &m, double rho, double mu);
  void update_mesh(const double *coords, const uint8_t *p1_bc, const uint8_t *p2_bc);
  void build_sparse_matrices();
  int piso_step();
  int simple_step();
  // the step's parts as the reference exposes them (NavierStokes.cpp:416-524,
  // 586-668, 702-771, 804-813): status codes as piso_step's
  int momentum_predictor_step();
  int pressure_poisson_step();
  void correct_velocity_step();
  void normalize_pressure_step();
  // solvePressureBiCGSTAB (NavierStokes.cpp:773-802): BiCGSTAB with Jacobi on
  // the Poisson matrix L itself (no penalty rows), tolerance 1e-8, 1000
  // iterations, the reference's stopping rule; x (nv) gets the solution
  SolveResult solve_poisson_bicgstab(const double *rhs, double *x);
  bool has_nans();

  // setters with side effects on cached operators
  void set_dt(double v);
  void set_mean_inlet(const double v[3]);
  void set_inlet_profile(const double *g);  // nu x 3 host array, nullptr = uniform
  void set_outlet_type(int t);
  void set_reference_node(int r);

  // host <-> device field access
  void get_velocity(double *out);
  void get_velocity_at(int64_t node, double out[3]);  // one node's 3 values only
  // Snapshot u and p on the device, then copy them to host buffers on a
  // second stream (overlapping later work when the buffers are pinned);
  // fields_wait blocks until the copies of the last fields_async are done.
  void fields_async(double *u_out, double *p_out);
  void fields_wait();
  void get_pressure_at(int64_t node, double *out);
  void set_velocity(const double *in);
  void get_velocity_star(double *out);
  void get_pressure(double *out);
  void set_pressure(const double *in);
  void get_pressure_correction(double *out);
  void get_inv_lumped_mass(double *out);
  void get_colouring(int32_t *colour) const;
  void get_matrix(int which, int32_t *rowptr, int32_t *col, double *val);
  void get_incidence(int32_t *ptr, int32_t *tet);
  void get_mesh_arrays(double *coords, int32_t *tets, uint8_t *p1_bc) const;
  void compute_advection(double *out);             // computeAdvectionRHS
  void compute_divergence(const double *u, double *out);  // computeDivergence
  void apply_matrix(int which, const double *x, double *y, int nrhs);
  // apply_matrix with x and y staged in the library's static device scratch
  // array instead of per-call allocations (noalloc.hip); no device memory is
  // allocated or freed
  void apply_matrix_noalloc(int which, const double *x, double *y, int nrhs);
  SolveResult krylov_solve(int which, int method, int nrhs, const double *b, double *x,
                           double tol, int maxit, int precision);
  // one V-cycle of the pressure operator's AMG hierarchy applied to r (host
  // arrays of nv values), building the hierarchy if needed; precision 1: the
  // FP32 V-cycle of the mixed mode
  void amg_apply(const double *r, double *z, int precision);
  // levels of the AMG hierarchy (0 if none), rows and nonzeros per level (up to
  // cap entries), set-up time of the last build in ms, builds so far
  int amg_info(int cap, int64_t *rows, int64_t *nnz, double *setup_ms, int *builds);
  // element patches: patches, visits, node-table entries, built; set-up time
  void elem_info(int64_t out[4], double *setup_ms) const {
    out[0] = elem_built_ ? elem_dev_.npatch : 0;
    out[1] = elem_built_ ? elem_dev_.nvisit : 0;
    out[2] = elem_built_ ? (int64_t)d_enodes.n : 0;
    out[3] = elem_built_ ? 1 : 0;
    if (setup_ms) *setup_ms = elem_setup_ms_;
  }
  // momentum_solver 2: the polynomial preconditioner M^-1 = (g0 I - g1 D^-1 A) D^-1
  // of the current lhs (estimating the spectral radius of D^-1 A if needed):
  // out = {rho estimate, lmax, g0, g1, estimates so far}
  void momentum_poly(double out[5]) {
    prepare_operators();
    ensure_mom_poly();
    out[0] = mom_rho_;
    out[1] = mom_lmax_;
    out[2] = mom_g0_;
    out[3] = mom_g1_;
    out[4] = mom_poly_builds_;
  }
  // momentum_solver 3: M^-1 = (a0 I + a1 D^-1 A + a2 (D^-1 A)^2) D^-1, out = {a0, a1, a2}
  void momentum_poly2(double out[3]) {
    prepare_operators();
    ensure_mom_poly();
    for (int k = 0; k < 3; ++k) out[k] = mom_a_[k];
  }
  hipStream_t stream() const { return stream_; }
  // Momentum-matrix SpMV launches repeated `reps` times between two events.
  double time_momentum_spmv(int reps, int precision);

  // ---- boundary forces (forces.hip); which: 0 pressure, 1 shear, 2 total ----
  void compute_forces(int which, double *out);
  void face_pressure(int64_t fi, double *mean, double force[3]);
  int64_t n_faces() {
    ensure_faces();
    return (int64_t)all_faces_.size();
  }

  // ---- discrete adjoint (adjoint.hip) ----
  void set_viscosity(double mu);  // re-assembles K
  int adjoint_forward(int nsteps, const double *g, double *u_final);
  int adjoint_backward(const double *lambda_u, double *dLdg, double *dLdmu);
  int adjoint_gradient(int nsteps, int nk, const double *basis, const double *coef, double *J,
                       double *dJdc, double *dJdmu);

 private:
  hipStream_t stream_ = nullptr;
  // field read-back: snapshot buffers, copy stream, snapshot / copied events
  hipStream_t copy_stream_ = nullptr;
  hipEvent_t ev_snap_ = nullptr, ev_copied_ = nullptr;
  bool copy_pending_ = false;
  DBuf<double> d_usnap, d_psnap;
  // this device's static scratch array (noalloc.hip), looked up at construction
  double *scratch_ = nullptr;
  // mesh on device
  DBuf<int> d_tets, d_ten, d_order, d_incp, d_inct;
  DBuf<uint8_t> d_loc;       // row-gather slot map (assembly 0): local index per incidence
  DBuf<int> d_amap, d_pmap;  // P2 slots [10 nt][10], P1 slots of vertex rows [4 nt][4]
  int maxrowA_ = 0, maxrowP_ = 0;  // longest P2 / P1 pattern row
  DBuf<double> d_vol, d_grad, d_xyz;
  DBuf<uint8_t> d_bc;
  std::vector<int32_t> h_colour_offsets, h_order;
  std::vector<uint8_t> h_bc;
  std::vector<double> h_coords;
  std::vector<int32_t> h_tets, h_ten;
  // patterns
  DBuf<int> d_Arp, d_Acol, d_Adiag, d_Prp, d_Pcol, d_Pdiag;
  DBuf<int> d_Aw1, d_Aw3, d_Aw3f, d_Pw1, d_Pw3, d_Pw3f;  // CSR-stream windows
  DBuf<double> d_wpart;                  // CSR-stream inner-product scratch
  DevCsr A_{}, P_{};
  // sliced-ELL copies of the two Krylov operators' patterns (lhs, L_bc)
  DBuf<int> d_Asptr, d_Arlen, d_Aperm, d_Ascol, d_Asrc;
  DBuf<int> d_Psptr, d_Prlen, d_Pperm, d_Pscol, d_Psrc;
  DevCsr As_{}, Ps_{};
  int64_t storedA = 0, storedP = 0;
  bool sellA_built_ = false;  // the momentum operator's sliced-ELL copy exists
  // operators (CSR slot order; the *S and *32 arrays are in sliced-ELL order)
  DBuf<double> d_M, d_K, d_lhs, d_L, d_Lbc, d_G3, d_ilm, d_invdA, d_invdL;
  DBuf<double> d_lhsS, d_LbcS;
  DBuf<float> d_lhs32, d_Lbc32, d_invdA32, d_invdL32;
  bool lhs_dirty = true, lbc_dirty = true;
  int sellA_ = -1, sellP_ = -1;  // layout the Krylov operator arrays hold (1 sliced-ELL)
  // the Krylov operators in their current layout (valid after prepare_operators)
  const DevCsr &momA() const { return sellA_ == 1 ? As_ : A_; }
  // the FP32 momentum operator of the inner solves: momA(), or momA() with
  // the element patches once they hold the current operator (elem_spmv)
  const DevCsr &momA32() {
    if (!(opt.elem_spmv && elem_built_ && elem_fresh_)) return momA();
    elemA_ = momA();
    elemA_.elem = &elem_dev_;
    return elemA_;
  }
  const double *momV() const { return sellA_ == 1 ? d_lhsS.p : d_lhs.p; }
  // builds the length-sorted sliced-ELL copy of the momentum pattern when first used
  void ensure_sell_A();
  const DevCsr &preA() const { return sellP_ == 1 ? Ps_ : P_; }
  const double *preV() const { return sellP_ == 1 ? d_LbcS.p : d_Lbc.p; }
  double lhs_dt = 0.0;
  // fields
  DBuf<double> d_u, d_us, d_p, d_pc, d_ginlet, d_conv, d_cT, d_wT, d_div, d_rhs,
      d_rhsp, d_gp, d_scalar, d_xref;
  DBuf<int> d_flag, d_mask;
  // Krylov
  KrylovWork<double> w3_, w1_;
  // element patches of the momentum operator (elem_spmv; elem_solver.hip)
  DBuf<int> d_enode_off, d_enodes, d_erow_off, d_erows, d_evisit_off, d_etorder;
  DBuf<uint16_t> d_estarts;
  DBuf<uint32_t> d_evstat;
  DBuf<float> d_evgeo, d_epen, d_einvd32;  // d_einvd32: FP32 Jacobi scale in patch order
  ElemPatchDev elem_dev_{};
  DevCsr elemA_{};
  bool elem_built_ = false, elem_fresh_ = false;
  // momentum_solver 2: spectral radius estimate of D^-1 A (power iteration on the
  // FP64 operator, refreshed with the lhs; later estimates start from the last
  // vector) and the polynomial preconditioner's coefficients
  bool mom_poly_fresh_ = false;
  int mom_poly_builds_ = 0;
  double mom_rho_ = 0.0, mom_lmax_ = 0.0, mom_g0_ = 0.0, mom_g1_ = 0.0;
  double mom_a_[3] = {0.0, 0.0, 0.0};  // degree 2, power basis in D^-1 A
  DBuf<double> d_pow_x, d_pow_v;
  int64_t elem_nvisit_ = 0, elem_nrows_ = 0;
  double elem_setup_ms_ = 0.0;
  void ensure_elem();
  void ensure_mom_poly();
  DBuf<double> d_bsb, d_bsx, d_bsinvd;  // solve_poisson_bicgstab operands (kept: graph keys)
  KrylovWork<float> f3_, f1_;
  DBuf<float> d_b32, d_x32;
  DBuf<double> d_r64, d_part_a, d_part_b, d_part_c, d_part_r;
  DBuf<KrylovState> d_state;
  KrylovState *h_state = nullptr;
  std::vector<hipEvent_t> ev_pool_;
  int ev_used_ = 0;
  // captured Krylov iteration chunks, keyed by every argument they bake in
  struct ChunkGraph {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<ChunkGraph> graphs_;
  void clear_graphs();
  template <class F>
  void run_chunk(const std::vector<uintptr_t> &key, F &&launch);
  // AMG hierarchy of the pressure operator (pressure_solver 2; amg_solver.hip)
  struct AmgDevMat {  // one operator: CSR-stream, or sliced-ELL for rows too long for it
    DBuf<int> rp, ci, win, sptr, rlen, perm;
    DBuf<double> v;
    DBuf<float> v32;  // FP32 copy for the mixed-mode V-cycle
  };
  struct AmgDevLevel {
    AmgDevMat A, P, R;
    DBuf<double> dinv, b, x, y, r;
    DBuf<float> dinv32, b32, x32, y32, r32;
  };
  std::vector<AmgDevLevel> amg_lv_;
  std::vector<AmgLevelView<double>> amg_view_;
  std::vector<AmgLevelView<float>> amg_view32_;
  DBuf<double> amg_cinv_;
  DBuf<float> amg_cinv32_;
  bool amg32_built_ = false;  // FP32 copies of the current hierarchy exist
  int amg_nc_ = 0;
  bool amg_cjac_ = false, amg_built_ = false, amg_lbc_changed_ = true;
  int amg_stale_ = 0, amg_builds_ = 0;
  double amg_setup_ms_ = 0.0;
  uint64_t amg_print_ = 0;          // fingerprint of the L_bc values last seen
  DBuf<uint64_t> d_print_;
  std::vector<int64_t> amg_rows_, amg_nnz_;
  // level 0's own copy of L_bc without its stored zeros (amg_lv_[0].A), in the
  // layout of the Krylov pressure operator; used by the V-cycle and the CG's
  // product while it holds the current values (not after a change the
  // hierarchy is kept through, amg_reuse)
  DevCsr amg_top_A_{};
  bool amg_top_fresh_ = false;
  const DevCsr &amg_fine_A() { return amg_top_fresh_ ? amg_top_A_ : preA(); }
  const double *amg_fine_v() { return amg_top_fresh_ ? amg_lv_[0].A.v.p : preV(); }
  void ensure_amg();
  void ensure_amg32();  // FP32 copies of the hierarchy (device conversions)
  template <class T>
  AmgLevelView<T> *amg_views();
  template <class T>
  const T *amg_cinv();
  template <class T>
  SolveResult pcg_amg(const double *b, double *x, double tol, int maxit);
  // adjoint state
  DBuf<int> d_Ptslot;  // slot of (j, i) for each slot (i, j) of the P1 pattern
  DBuf<double> d_G3T, d_lu, d_lus, d_lb, d_lg, d_lgp, d_lpc, d_lrp, d_lp, d_lp_new, d_contrib,
      d_tmp3;
  std::vector<DBuf<double>> ckpt_u, ckpt_p;
  bool g3t_dirty = true;
  int adj_steps_ = 0;
  void ensure_adjoint_buffers(int nsteps);
  // boundary faces for the force kernels
  DBuf<int> d_bfv, d_bft, d_vfp, d_vff;
  DBuf<double> d_bfg, d_fp, d_fs, d_forces;
  std::vector<Face> all_faces_;
  Mesh face_mesh_;              // face list (orientation fixed once) + current geometry
  bool face_topology_ = false;  // face_mesh_ holds the faces of this mesh
  int nbf_ = 0;
  bool faces_ready_ = false;
  void ensure_faces();

  void assemble();
  void prepare_operators();
  void upload_inlet();
  int momentum_predictor();
  void repeat_pressure_correction();
  int pressure_poisson();
  void correct_velocity();
  int finish_step();
  void ensure_mixed_buffers();
  void set_state_params(double tol, int maxit, int nrhs);
  void set_mask(int m0, int m1, int m2);
  void read_state();
  void spmv_event_begin();
  void spmv_event_end();

  template <class T, int K>
  SolveResult bicgstab(const DevCsr &A, const T *val, const T *invd, const T *b, T *x,
                       double tol, int maxit, KrylovWork<T> &w);
  template <class T, int K>
  SolveResult pcg_poly(const DevCsr &A, const T *val, const T *invd, const T *b, T *x,
                       double tol, int maxit, KrylovWork<T> &w, int degree);
  template <class T, int K>
  SolveResult pcg(const DevCsr &A, const T *val, const T *invd, const T *b, T *x,
                  double tol, int maxit, KrylovWork<T> &w);
  template <int K>
  SolveResult refine(const DevCsr &A, const double *val64, const double *invd64,
                     const float *val32, const float *invd32, const double *b, double *x,
                     double tol, int maxit, int method, bool keep_x = false);
  SolveResult solve_momentum(const double *b, double *x);
  SolveResult solve_pressure(const double *b, double *x);
};

// the V-cycle views and coarse inverses per value type (amg_solver.hip)
template <>
AmgLevelView<double> *NSSolver::amg_views<double>();
template <>
AmgLevelView<float> *NSSolver::amg_views<float>();
template <>
const double *NSSolver::amg_cinv<double>();
template <>
const float *NSSolver::amg_cinv<float>();

}  // namespace cfsa
