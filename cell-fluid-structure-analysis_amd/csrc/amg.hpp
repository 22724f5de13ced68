// Smoothed-aggregation algebraic multigrid for the pressure Poisson operator
// L_bc (P1 Laplacian plus penalty rows): host set-up of the hierarchy.
//
// The reference solves the pressure system with Jacobi-preconditioned
// BiCGSTAB (src/FEM/NavierStokes.cpp:633-649); this is an extension used as the
// preconditioner of a CG solve (pressure_solver = 2).  Set-up, per level:
//
//  * strength of connection: j != i with |a_ij| >= theta sqrt(|a_ii a_jj|);
//  * aggregation in three greedy passes over the rows in index order (Vanek,
//    Mandel & Brezina, "Algebraic multigrid by smoothed aggregation for second
//    and fourth order elliptic problems", Computing 56, 1996): a row whose
//    strong neighbours are all free starts an aggregate with them; a free row
//    joins the aggregate of its first aggregated strong neighbour; a remaining
//    row starts an aggregate with its free strong neighbours.  Rows without
//    strong connections (the 1e12 penalty rows) stay unaggregated: the coarse
//    levels do not see them and the smoother handles them;
//  * tentative prolongator T (constant per aggregate, columns normalised),
//    smoothed as P = (I - omega D^-1 A) T with omega = 1.2 / lmax, lmax the
//    smoothers' upper bound of the spectral radius of D^-1 A (below; the
//    textbook 4 / (3 rho) coarsens the next level too aggressively here);
//  * Galerkin coarse operator A_c = P^T A P, restriction R = P^T;
//  * until at most max_coarse rows remain (or aggregation stalls); the
//    coarsest operator is inverted densely (Cholesky; pivots that vanish are
//    dropped, which gives a pseudo-inverse for a semi-definite operator).
//
// Each level is smoothed by a degree-2 Chebyshev polynomial in D^-1 A over
// [lmax / 30, lmax], lmax = 1.1 rho (capped by the Gershgorin bound), before
// and after the coarse correction; the polynomial is symmetric in the A inner
// product, so the V-cycle is a symmetric positive definite preconditioner.
//
// Entries that are exactly zero (half of the Kuhn meshes' P1 Laplacian, and
// what products make of them) are not stored on any level: a sum that skips a
// zero term is the same number, so values are unchanged and products cost less.
//
// Every loop is deterministic: sparse products run row-parallel with a fixed
// per-row accumulation order and sorted columns, so the hierarchy does not
// depend on the thread count.
#pragma once

#include <cstdint>
#include <vector>

namespace cfsa {

struct HostCsr {
  int64_t n = 0;  // rows
  int64_t m = 0;  // columns
  std::vector<int32_t> rp, ci;
  std::vector<double> v;
  int64_t nnz() const { return rp.empty() ? 0 : rp.back(); }
};

struct AmgParams {
  double theta = 0.06;   // strength threshold
  int max_coarse = 1000;  // rows of the coarsest (dense) level at most
  int max_levels = 12;
};

struct AmgLevel {
  HostCsr A;                   // level operator (level 0: the fine matrix)
  HostCsr P, R;                // n x nc and nc x n (all levels but the last)
  std::vector<int32_t> agg;    // aggregate of each row, -1 if none
  std::vector<double> dinv;    // 1 / a_ii (0 for a zero diagonal)
  double rho = 0.0;            // spectral radius estimate of D^-1 A
  double lmax = 0.0;           // Chebyshev interval [lmax / 30, lmax]
  double c0 = 0.0, c1 = 0.0, c2 = 0.0;  // degree-2 Chebyshev coefficients
};

struct AmgHierarchy {
  std::vector<AmgLevel> levels;     // levels.back() is the coarsest
  std::vector<double> coarse_inv;   // dense (pseudo-)inverse, row-major
  int64_t coarse_n = 0;
  bool coarse_jacobi = false;       // coarsening stalled above the dense limit
};

// Degree-2 Chebyshev coefficients of a level over [lmax / 30, lmax]:
//   x1 = x0 + c0 D^-1 (b - A x0)
//   x2 = x1 + c1 (x1 - x0) + c2 D^-1 (b - A x1)
void amg_chebyshev(double lmax, double *c0, double *c1, double *c2);

// Builds the hierarchy of the square matrix A (sorted columns per row).
AmgHierarchy amg_build(const HostCsr &A, const AmgParams &prm = AmgParams());

// y = A x (row-sequential sums, host)
void host_spmv(const HostCsr &A, const double *x, double *y);

}  // namespace cfsa
