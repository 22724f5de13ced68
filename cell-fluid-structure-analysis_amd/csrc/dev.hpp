// Shared definitions for the HIP side of the PISO library (gfx950 only).
//
// Conventions used by every kernel:
//  * velocity-like vectors are node-interleaved [n][3] (the reference stores
//    std::vector<Eigen::Vector3d>, NavierStokes.hpp:66-67, so host copies are
//    plain memcpy);
//  * CSR patterns use int32 row pointers and column indices, ascending columns;
//  * reductions are deterministic: a fixed grid writes per-block partials and a
//    single-block kernel sums them in a fixed order, so repeated runs are
//    bitwise identical;
//  * the FP64 path is compiled without FMA contraction so element kernels and
//    row-sequential SpMV reproduce the reference's rounding exactly.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace cfsa {

#define CFSA_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw std::runtime_error(std::string("HIP error '") +                   \
                               hipGetErrorString(e_) + "' at " #expr " (" +   \
                               __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define CFSA_LAUNCH_CHECK() CFSA_HIP(hipGetLastError())

constexpr int kBlock = 256;         // threads per block, 4 wave64
constexpr int kVecGrid = 1024;      // fixed grid of vector/reduction kernels
constexpr int kSpmvGrid = 2048;     // fixed grid of CSR-stream SpMV kernels
constexpr int kMaxPartials = 4096;  // >= max(kVecGrid, kSpmvGrid); partials hold pairs

// Fluid boundary labels (FluidBCType, NavierStokes.hpp:13-19).
constexpr uint8_t kWall = 0, kInlet = 1, kOutlet = 2, kInternal = 3;

// Krylov solver status per right-hand side.
enum KStatus : int {
  KS_RUNNING = 0,
  KS_DONE = 1,     // loop ended (converged or iteration limit)
  KS_RESTART = 2,  // BiCGSTAB rho breakdown: host performs the restart
  KS_RESUME = 3,   // restart done, continue at the beta update
};

// Device-resident state of a batched (up to 3 right-hand sides) Krylov solve.
// Scalars are FP64 on both the FP64 and the mixed path.
// The Jacobi-PCG solve reuses the fields: rho = |r|^2_{M^-1} (absNew),
// rho_old = absOld, tol2 = stopping threshold, rr = |r|^2, r0v = p.Ap.
struct KrylovState {
  double rho[3], rho_old[3], alpha[3], omega[3], beta[3];
  double r0_sq[3], rhs_sq[3], tol2[3], rr[3], err[3];
  double r0v[3], tt[3], ts[3];  // last reduced inner products (diagnostics)
  // PCG: rho and iteration count in two slots by iteration parity, so that a
  // kernel reads one slot while its block 0 writes the other
  double rho_s[2][3];
  int it_s[2][3];
  // PCG: the active flags as k_cg_xr saw them (what k_cg_p updates x for)
  int act_s[2][3];
  double tol;
  double restart_scale;  // multiplier of eps^2 in the BiCGSTAB restart test
  int iters[3], restarts[3], status[3], active[3];
  int maxit, nrhs, any_active, pad;
};

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace cfsa
