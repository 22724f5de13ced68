// Discrete adjoint of the PISO step on the GPU (build extension).
//
// Forward map of one step (Dirichlet outlet, n_corr >= 1), reference
// NavierStokes.cpp:374-771:
//   b_c  = (M u_c)/dt - E G_c p + conv_c(u) + 1e10 g_c            (Inlet rows)
//   u*_c = A^-1 b_c,      A = M/dt + K1/viscosity + penalties
//   p'   = Lbc^-1 ( -(sum_c G_c u*_c)/dt + penalties )
//   u'   = g (Inlet) | 0 (Wall) | u* - relax_u dt ilm G p' (free P1) | u* (free P2)
//   p_new = p + n_corr relax_p p' (non-outlet rows) | p_out (outlet rows)
// The reverse sweep maps (lambda_u', lambda_p') to (lambda_u, lambda_p) and
// accumulates dJ/dg and dJ/dviscosity.  A, M, K and Lbc are symmetric, so the
// adjoint solves reuse the forward operators and solvers; G^T products use
// the transposed values of the (symmetric) P1 pattern.  Each step's u* and p'
// are recomputed from per-step checkpoints of (u, p) kept on the device.
#include <cmath>
#include <cstring>

#include "solver.hpp"

#include "dd.hpp"

namespace cfsa {

namespace {

constexpr double kPen = 1e10;

// velocity outputs -> lambda_u*, lambda_(G p'), lambda_g
__global__ void k_adj_velocity_out(int nu, int nv, const uint8_t *__restrict__ bc,
                                   const double *__restrict__ ilm,
                                   const double *__restrict__ lu, double relax_u, double dt,
                                   double *__restrict__ lus, double *__restrict__ lgp,
                                   double *__restrict__ lg) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nu) return;
  const uint8_t b = bc[i];
  const size_t o = (size_t)i * 3;
  double s = 0.0;
  const bool free_row = b != kInlet && b != kWall;
  if (i < nv && free_row && ilm[i] > 1e-14) s = -relax_u * dt * ilm[i];
  for (int c = 0; c < 3; ++c) {
    const double l = lu[o + c];
    if (b == kInlet) lg[o + c] += l;
    lus[o + c] = free_row ? l : 0.0;
    if (i < nv) lgp[o + c] = s * l;
  }
}

// pressure outputs: lambda_p' += n_corr relax_p lambda_p_new; lambda_p carried
// through non-outlet rows
__global__ void k_adj_pressure_out(int nv, const uint8_t *__restrict__ bc, int dirichlet,
                                   double coef, const double *__restrict__ lp,
                                   double *__restrict__ lpc, double *__restrict__ lp_new) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nv) return;
  if (dirichlet && bc[i] == kOutlet) {
    lp_new[i] = 0.0;
    return;
  }
  lp_new[i] = lp[i];
  lpc[i] += coef * lp[i];
}

__global__ void k_scale(int n, double a, const double *__restrict__ x, double *__restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

// per tet: lambda_c = -sum_i w_i lambda_b_i; contributions to the 4 vertices
//   (grad phi_j . u_bar) lambda_c + (Gu^T lambda_c) / 4
__global__ __launch_bounds__(kBlock) void k_conv_adj_tet(
    int nt, int nv, const int *__restrict__ tets, const int *__restrict__ ten,
    const double *__restrict__ grad, const double *__restrict__ vol,
    const double *__restrict__ u, const double *__restrict__ lb, double *__restrict__ contrib) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  const double v = vol[t];
  double lc[3] = {0, 0, 0};
  for (int a = 0; a < 4; ++a) {
    const size_t n = (size_t)tets[(size_t)t * 4 + a] * 3;
    for (int c = 0; c < 3; ++c) lc[c] -= (v / 4.0) * lb[n + c];
  }
  for (int e = 0; e < 6; ++e) {
    const size_t n = ((size_t)nv + ten[(size_t)t * 6 + e]) * 3;
    for (int c = 0; c < 3; ++c) lc[c] -= (v / 24.0) * lb[n + c];
  }
  double uu[4][3], g[4][3], ue[3] = {0, 0, 0}, Gu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int j = 0; j < 4; ++j) {
    const size_t n = (size_t)tets[(size_t)t * 4 + j] * 3;
    for (int c = 0; c < 3; ++c) {
      uu[j][c] = u[n + c];
      g[j][c] = grad[(size_t)t * 12 + j * 3 + c];
      ue[c] += 0.25 * uu[j][c];
    }
  }
  for (int j = 0; j < 4; ++j)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Gu[a][b] += uu[j][a] * g[j][b];
  double gt[3];
  for (int b = 0; b < 3; ++b) gt[b] = 0.25 * (Gu[0][b] * lc[0] + Gu[1][b] * lc[1] + Gu[2][b] * lc[2]);
  for (int j = 0; j < 4; ++j) {
    const double s = g[j][0] * ue[0] + g[j][1] * ue[1] + g[j][2] * ue[2];
    for (int c = 0; c < 3; ++c) contrib[((size_t)t * 4 + j) * 3 + c] = s * lc[c] + gt[c];
  }
}

__global__ __launch_bounds__(kBlock) void k_conv_adj_gather(
    int nv, const int *__restrict__ ip, const int *__restrict__ it, const int *__restrict__ tets,
    const double *__restrict__ contrib, double *__restrict__ lu) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nv) return;
  double a[3] = {0, 0, 0};
  for (int q = ip[i]; q < ip[i + 1]; ++q) {
    const int t = it[q];
    int j = 0;
    while (j < 3 && tets[(size_t)t * 4 + j] != i) ++j;
    for (int c = 0; c < 3; ++c) a[c] += contrib[((size_t)t * 4 + j) * 3 + c];
  }
  for (int c = 0; c < 3; ++c) lu[(size_t)i * 3 + c] += a[c];
}

// lambda_u^N = u on Outlet velocity nodes, 0 elsewhere
__global__ void k_adj_seed(int nu, const uint8_t *__restrict__ bc, const double *__restrict__ u,
                           double *__restrict__ lu) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nu) return;
  for (int c = 0; c < 3; ++c) lu[(size_t)i * 3 + c] = bc[i] == kOutlet ? u[(size_t)i * 3 + c] : 0.0;
}

__global__ void k_adj_inlet_add(int nu, const uint8_t *__restrict__ bc,
                                const double *__restrict__ lu, double *__restrict__ lg) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nu || bc[i] != kInlet) return;
  for (int c = 0; c < 3; ++c) lg[(size_t)i * 3 + c] += lu[(size_t)i * 3 + c];
}

__global__ void k_gather_slots3(int64_t nnz, const int *__restrict__ tslot,
                                const double *__restrict__ g, double *__restrict__ gt) {
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * kBlock)
    for (int c = 0; c < 3; ++c) gt[k * 3 + c] = g[(size_t)tslot[k] * 3 + c];
}

}  // namespace

void NSSolver::ensure_adjoint_buffers(int nsteps) {
  const size_t U = (size_t)nu * 3;
  if (d_lu.n != U) {
    d_lu.alloc(U);
    d_lus.alloc(U);
    d_lb.alloc(U);
    d_lg.alloc(U);
    d_lgp.alloc((size_t)nv * 3);
    d_lpc.alloc((size_t)nv);
    d_lrp.alloc((size_t)nv);
    d_lp.alloc((size_t)nv);
    d_lp_new.alloc((size_t)nv);
    d_contrib.alloc((size_t)nt * 12);
    d_tmp3.alloc(U);
  }
  if (d_G3T.n != (size_t)nnzP * 3) {
    d_G3T.alloc((size_t)nnzP * 3);
    g3t_dirty = true;
  }
  if ((int)ckpt_u.size() < nsteps) {
    ckpt_u.resize((size_t)nsteps);
    ckpt_p.resize((size_t)nsteps);
  }
  for (int n = 0; n < nsteps; ++n) {
    ckpt_u[(size_t)n].alloc(U);
    ckpt_p[(size_t)n].alloc((size_t)nv);
  }
  if (g3t_dirty) {
    hipLaunchKernelGGL(k_gather_slots3, dim3(1024), dim3(kBlock), 0, stream_, nnzP, d_Ptslot.p,
                       d_G3.p, d_G3T.p);
    CFSA_LAUNCH_CHECK();
    g3t_dirty = false;
  }
}

void NSSolver::set_viscosity(double mu) {
  if (!initialized) throw std::runtime_error("set_viscosity before initialize");
  viscosity = mu;
  assemble();
  g3t_dirty = true;
  CFSA_HIP(hipStreamSynchronize(stream_));
}

// Forward N steps from rest with inlet values g, keeping checkpoints.
int NSSolver::adjoint_forward(int nsteps, const double *g, double *u_final) {
  if (!initialized) throw std::runtime_error("adjoint_forward before initialize");
  if (nsteps < 1) throw std::invalid_argument("adjoint_forward: nsteps >= 1");
  if (outlet_type != 0 || n_corrections < 1)
    throw std::invalid_argument("adjoint: Dirichlet outlet and at least one corrector required");
  ensure_adjoint_buffers(nsteps);
  set_inlet_profile(g);
  d_u.zero(stream_);
  d_p.zero(stream_);
  set_inlet_nodes(stream_, nu, d_bc.p, d_ginlet.p, d_u.p);
  adj_steps_ = 0;
  for (int n = 0; n < nsteps; ++n) {
    CFSA_HIP(hipMemcpyAsync(ckpt_u[(size_t)n].p, d_u.p, (size_t)nu * 3 * sizeof(double),
                            hipMemcpyDeviceToDevice, stream_));
    CFSA_HIP(hipMemcpyAsync(ckpt_p[(size_t)n].p, d_p.p, (size_t)nv * sizeof(double),
                            hipMemcpyDeviceToDevice, stream_));
    const int st = piso_step();
    if (st != ST_OK) return st;
    adj_steps_ = n + 1;
  }
  if (u_final) get_velocity(u_final);
  return ST_OK;
}

// Reverse sweep: lambda_u (nu x 3) = dL/du at the final state of the last
// adjoint_forward; returns dL/dg (nu x 3) and dL/dviscosity.
int NSSolver::adjoint_backward(const double *lambda_u, double *dLdg, double *dLdmu) {
  if (adj_steps_ < 1) throw std::runtime_error("adjoint_backward without adjoint_forward");
  hipStream_t s = stream_;
  const size_t U = (size_t)nu * 3;
  if (lambda_u)
    CFSA_HIP(hipMemcpy(d_lu.p, lambda_u, U * sizeof(double), hipMemcpyHostToDevice));
  else
    hipLaunchKernelGGL(k_adj_seed, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, d_bc.p,
                       d_u.p, d_lu.p);
  d_lp.zero(s);
  d_lg.zero(s);
  double dnu = 0.0;
  const double ncorr_relax = n_corrections * relax_p;
  int Gd = 0;
  for (int n = adj_steps_ - 1; n >= 0; --n) {
    // recompute u* and p' of step n from its checkpoint
    CFSA_HIP(hipMemcpyAsync(d_u.p, ckpt_u[(size_t)n].p, U * sizeof(double),
                            hipMemcpyDeviceToDevice, s));
    CFSA_HIP(hipMemcpyAsync(d_p.p, ckpt_p[(size_t)n].p, (size_t)nv * sizeof(double),
                            hipMemcpyDeviceToDevice, s));
    prepare_operators();
    std::memset(&last, 0, sizeof(last));
    if (int st = momentum_predictor()) return st;
    if (int st = pressure_poisson()) return st;
    // velocity and pressure outputs
    hipLaunchKernelGGL(k_adj_velocity_out, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, nv,
                       d_bc.p, d_ilm.p, d_lu.p, relax_u, dt, d_lus.p, d_lgp.p, d_lg.p);
    CFSA_LAUNCH_CHECK();
    spmv_gradT_in3(s, P_, d_G3T.p, d_lgp.p, 3, d_lpc.p, 1.0, 0.0);
    hipLaunchKernelGGL(k_adj_pressure_out, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, s, nv,
                       d_bc.p, 1, ncorr_relax, d_lp.p, d_lpc.p, d_lp_new.p);
    CFSA_LAUNCH_CHECK();
    // pressure solve; lambda_d = -lambda_rp / dt; lambda_u*_c(P1) += G_c^T lambda_d
    SolveResult rp = solve_pressure(d_lpc.p, d_lrp.p);
    if (!rp.ok) return ST_PRESSURE_NOCONV;
    hipLaunchKernelGGL(k_scale, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, s, nv, -1.0 / dt,
                       d_lrp.p, d_lpc.p);
    CFSA_LAUNCH_CHECK();
    spmv_grad3_acc(s, P_, d_G3T.p, d_lpc.p, d_lus.p);
    // momentum solves (A symmetric): lambda_b = A^-1 lambda_u*
    SolveResult rm = solve_momentum(d_lus.p, d_lb.p);
    if (!rm.ok) return ST_MOMENTUM_NOCONV;
    // dJ/dviscosity += sum_c lambda_b_c . (K u*_c) / viscosity
    spmv3_dot<double>(s, A_, d_K.p, d_us.p, d_tmp3.p, d_lb.p, nullptr, d_part_r.p, &Gd);
    hipLaunchKernelGGL((k_reduce_pairs<3>), dim3(1), dim3(kBlock), 0, s, d_part_r.p, Gd, d_scalar.p);
    CFSA_LAUNCH_CHECK();
    double dots[3];
    CFSA_HIP(hipMemcpyAsync(dots, d_scalar.p, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    // right-hand side: lambda_u = M lambda_b / dt + C(u)^T lambda_b;
    // lambda_p = lambda_p_new - sum_c G_c^T lambda_b_c(P1); lambda_g += 1e10 lambda_b (Inlet)
    momentum_rhs_adjoint(s, A_, d_M.p, d_lb.p, d_bc.p, dt, kPen, d_lu.p, d_lg.p);
    spmv_gradT_in3(s, P_, d_G3T.p, d_lb.p, 3, d_lp_new.p, -1.0, 1.0);
    hipLaunchKernelGGL(k_conv_adj_tet, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, s, nt, nv,
                       d_tets.p, d_ten.p, d_grad.p, d_vol.p, d_u.p, d_lb.p, d_contrib.p);
    CFSA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_conv_adj_gather, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, s, nv,
                       d_incp.p, d_inct.p, d_tets.p, d_contrib.p, d_lu.p);
    CFSA_LAUNCH_CHECK();
    std::swap(d_lp.p, d_lp_new.p);
    CFSA_HIP(hipStreamSynchronize(s));
    dnu += (dots[0] + dots[1] + dots[2]) / viscosity;
  }
  // u^0 has the inlet values on Inlet nodes
  hipLaunchKernelGGL(k_adj_inlet_add, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, d_bc.p,
                     d_lu.p, d_lg.p);
  CFSA_LAUNCH_CHECK();
  CFSA_HIP(hipStreamSynchronize(s));
  if (dLdg) d_lg.download(dLdg, U);
  if (dLdmu) *dLdmu = dnu;
  return ST_OK;
}

int NSSolver::adjoint_gradient(int nsteps, int nk, const double *basis, const double *coef,
                               double *J, double *dJdc, double *dJdmu) {
  const size_t U = (size_t)nu * 3;
  std::vector<double> g(U, 0.0), uN(U), lg(U), pN((size_t)nv);
  for (int k = 0; k < nk; ++k)
    for (size_t i = 0; i < U; ++i) g[i] += coef[k] * basis[(size_t)k * U + i];
  if (int st = adjoint_forward(nsteps, g.data(), uN.data())) return st;
  get_pressure(pN.data());
  double j = 0.0;
  for (int i = 0; i < nu; ++i)
    if (h_bc[(size_t)i] == kOutlet)
      for (int c = 0; c < 3; ++c) j += 0.5 * uN[(size_t)i * 3 + c] * uN[(size_t)i * 3 + c];
  if (int st = adjoint_backward(nullptr, lg.data(), dJdmu)) return st;
  if (J) *J = j;
  for (int k = 0; k < nk; ++k) {
    double s = 0.0;
    for (size_t i = 0; i < U; ++i) s += lg[i] * basis[(size_t)k * U + i];
    dJdc[k] = s;
  }
  // leave the solver at the final forward state
  CFSA_HIP(hipMemcpy(d_u.p, uN.data(), U * sizeof(double), hipMemcpyHostToDevice));
  CFSA_HIP(hipMemcpy(d_p.p, pN.data(), (size_t)nv * sizeof(double), hipMemcpyHostToDevice));
  return ST_OK;
}

}  // namespace cfsa
