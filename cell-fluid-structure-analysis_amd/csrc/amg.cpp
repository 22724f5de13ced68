// Host set-up of the smoothed-aggregation hierarchy (see amg.hpp).
#include "amg.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "parallel.hpp"

namespace cfsa {

namespace {

// Largest coarsest level inverted densely; a hierarchy whose coarsening stalls
// above it ends in a Jacobi step instead.
constexpr int64_t kMaxDense = 1500;

double diag_of(const HostCsr &A, int64_t i) {
  for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k)
    if (A.ci[(size_t)k] == i) return A.v[(size_t)k];
  return 0.0;
}

// A without its stored zeros.  The Kuhn meshes' P1 Laplacian stores half of its
// entries as exact zeros (the cotangent weight of an edge opposite a right
// dihedral angle), and products carry them on: every operator of the hierarchy
// is kept free of them.  A row sum that skips zero terms is the same number
// (s + 0 = s), so this changes no value, only the work per product.
HostCsr drop_zeros(const HostCsr &A) {
  HostCsr B;
  B.n = A.n;
  B.m = A.m;
  B.rp.assign((size_t)A.n + 1, 0);
  B.ci.reserve(A.ci.size());
  B.v.reserve(A.v.size());
  for (int64_t i = 0; i < A.n; ++i) {
    for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k)
      if (A.v[(size_t)k] != 0.0) {
        B.ci.push_back(A.ci[(size_t)k]);
        B.v.push_back(A.v[(size_t)k]);
      }
    B.rp[(size_t)i + 1] = (int32_t)B.ci.size();
  }
  return B;
}

// C = A B, rows in parallel.  Row i accumulates a_ik b_kj over k in A's column
// order and, inside, j in B's column order; columns are emitted sorted, entries
// that come out exactly zero are not stored.  The result does not depend on
// the thread count.
HostCsr spgemm(const HostCsr &A, const HostCsr &B) {
  if (A.m != B.n) throw std::logic_error("spgemm: inner sizes differ");
  HostCsr C;
  C.n = A.n;
  C.m = B.m;
  C.rp.assign((size_t)A.n + 1, 0);
  const int nt = host_threads();
  const int64_t chunk = (A.n + nt - 1) / nt;
  std::vector<std::vector<int32_t>> tc((size_t)nt);
  std::vector<std::vector<double>> tv((size_t)nt);
  run_threads(nt, [&](int t) {
    const int64_t a = std::min(A.n, t * chunk), b = std::min(A.n, a + chunk);
    if (a >= b) return;
    std::vector<double> acc((size_t)B.m, 0.0);
    std::vector<int64_t> mark((size_t)B.m, -1);
    std::vector<int32_t> touched;
    auto &oc = tc[(size_t)t];
    auto &ov = tv[(size_t)t];
    for (int64_t i = a; i < b; ++i) {
      touched.clear();
      for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k) {
        const int32_t r = A.ci[(size_t)k];
        const double av = A.v[(size_t)k];
        for (int32_t l = B.rp[(size_t)r]; l < B.rp[(size_t)r + 1]; ++l) {
          const int32_t j = B.ci[(size_t)l];
          const double pr = av * B.v[(size_t)l];
          if (mark[(size_t)j] != i) {
            mark[(size_t)j] = i;
            acc[(size_t)j] = pr;
            touched.push_back(j);
          } else {
            acc[(size_t)j] += pr;
          }
        }
      }
      std::sort(touched.begin(), touched.end());
      int32_t kept = 0;
      for (int32_t j : touched)
        if (acc[(size_t)j] != 0.0) {
          oc.push_back(j);
          ov.push_back(acc[(size_t)j]);
          ++kept;
        }
      C.rp[(size_t)i + 1] = kept;
    }
  });
  for (int64_t i = 0; i < A.n; ++i) {
    const int64_t s = (int64_t)C.rp[(size_t)i] + C.rp[(size_t)i + 1];
    if (s > INT32_MAX) throw std::runtime_error("AMG: operator too large for int32 indices");
    C.rp[(size_t)i + 1] = (int32_t)s;
  }
  C.ci.reserve((size_t)C.nnz());
  C.v.reserve((size_t)C.nnz());
  for (int t = 0; t < nt; ++t) {
    C.ci.insert(C.ci.end(), tc[(size_t)t].begin(), tc[(size_t)t].end());
    C.v.insert(C.v.end(), tv[(size_t)t].begin(), tv[(size_t)t].end());
  }
  return C;
}

// B = A^T (rows of B in ascending column order of A: sorted columns)
HostCsr transpose(const HostCsr &A) {
  HostCsr B;
  B.n = A.m;
  B.m = A.n;
  B.rp.assign((size_t)B.n + 1, 0);
  for (int32_t j : A.ci) B.rp[(size_t)j + 1]++;
  for (int64_t j = 0; j < B.n; ++j) B.rp[(size_t)j + 1] += B.rp[(size_t)j];
  B.ci.resize(A.ci.size());
  B.v.resize(A.v.size());
  std::vector<int32_t> pos(B.rp.begin(), B.rp.end() - 1);
  for (int64_t i = 0; i < A.n; ++i)
    for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k) {
      const int32_t j = A.ci[(size_t)k];
      const int32_t o = pos[(size_t)j]++;
      B.ci[(size_t)o] = (int32_t)i;
      B.v[(size_t)o] = A.v[(size_t)k];
    }
  return B;
}

// Strong neighbours of every row: j != i, |a_ij| >= theta sqrt(|a_ii a_jj|).
void strength(const HostCsr &A, const std::vector<double> &d, double theta,
              std::vector<int32_t> &sp, std::vector<int32_t> &sc) {
  const int64_t n = A.n;
  sp.assign((size_t)n + 1, 0);
  auto strong = [&](int64_t i, int32_t k) {
    const int32_t j = A.ci[(size_t)k];
    if (j == i) return false;
    return std::fabs(A.v[(size_t)k]) >= theta * std::sqrt(std::fabs(d[(size_t)i] * d[(size_t)j]));
  };
  parallel_chunks(n, [&](int, int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      int32_t c = 0;
      for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k) c += strong(i, k);
      sp[(size_t)i + 1] = c;
    }
  });
  for (int64_t i = 0; i < n; ++i) sp[(size_t)i + 1] += sp[(size_t)i];
  sc.resize((size_t)sp[(size_t)n]);
  parallel_chunks(n, [&](int, int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      int32_t o = sp[(size_t)i];
      for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k)
        if (strong(i, k)) sc[(size_t)o++] = A.ci[(size_t)k];
    }
  });
}

// Three greedy passes in row order; returns the number of aggregates.
int64_t aggregate(int64_t n, const std::vector<int32_t> &sp, const std::vector<int32_t> &sc,
                  std::vector<int32_t> &agg) {
  agg.assign((size_t)n, -1);
  int32_t na = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t a = sp[(size_t)i], b = sp[(size_t)i + 1];
    if (agg[(size_t)i] >= 0 || a == b) continue;
    bool free = true;
    for (int32_t k = a; k < b && free; ++k) free = agg[(size_t)sc[(size_t)k]] < 0;
    if (!free) continue;
    agg[(size_t)i] = na;
    for (int32_t k = a; k < b; ++k) agg[(size_t)sc[(size_t)k]] = na;
    ++na;
  }
  for (int64_t i = 0; i < n; ++i) {
    if (agg[(size_t)i] >= 0) continue;
    for (int32_t k = sp[(size_t)i]; k < sp[(size_t)i + 1]; ++k) {
      const int32_t g = agg[(size_t)sc[(size_t)k]];
      if (g >= 0) {
        agg[(size_t)i] = g;
        break;
      }
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    const int32_t a = sp[(size_t)i], b = sp[(size_t)i + 1];
    if (agg[(size_t)i] >= 0 || a == b) continue;
    agg[(size_t)i] = na;
    for (int32_t k = a; k < b; ++k)
      if (agg[(size_t)sc[(size_t)k]] < 0) agg[(size_t)sc[(size_t)k]] = na;
    ++na;
  }
  return na;
}

// Power-iteration estimate of the spectral radius of D^-1 A (20 steps from a
// fixed pseudo-random start vector).
double spectral_radius(const HostCsr &A, const std::vector<double> &dinv) {
  const int64_t n = A.n;
  std::vector<double> x((size_t)n), y((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t h = (uint32_t)i * 2654435761u;
    x[(size_t)i] = (double)(h >> 8) / 16777216.0 - 0.5;
  }
  auto norm = [](const std::vector<double> &v) {
    double s = 0.0;
    for (double e : v) s += e * e;
    return std::sqrt(s);
  };
  double nx = norm(x), rho = 0.0;
  if (nx == 0.0) return 0.0;
  for (double &e : x) e /= nx;
  for (int it = 0; it < 20; ++it) {
    host_spmv(A, x.data(), y.data());
    for (int64_t i = 0; i < n; ++i) y[(size_t)i] *= dinv[(size_t)i];
    rho = norm(y);
    if (rho == 0.0) return 0.0;
    for (int64_t i = 0; i < n; ++i) x[(size_t)i] = y[(size_t)i] / rho;
  }
  return rho;
}

// Dense (pseudo-)inverse of a symmetric positive semi-definite matrix by
// Cholesky; a pivot below 1e-12 of its row's own diagonal entry (a direction
// of the null space) is dropped, its unknown left at zero by every solve.  The
// test is per row: penalty rows (1e12) beside ordinary rows (~h) must not make
// the ordinary pivots look negligible.
std::vector<double> dense_inverse(const HostCsr &A) {
  const int64_t n = A.n;
  std::vector<double> a((size_t)(n * n), 0.0), diag((size_t)n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k) {
      a[(size_t)(i * n + A.ci[(size_t)k])] = A.v[(size_t)k];
      if (A.ci[(size_t)k] == i) diag[(size_t)i] = std::fabs(A.v[(size_t)k]);
    }
  // lower Cholesky factor in place
  std::vector<char> dropped((size_t)n, 0);
  for (int64_t j = 0; j < n; ++j) {
    double d = a[(size_t)(j * n + j)];
    for (int64_t k = 0; k < j; ++k) d -= a[(size_t)(j * n + k)] * a[(size_t)(j * n + k)];
    if (!(d > 1e-12 * diag[(size_t)j])) {
      dropped[(size_t)j] = 1;
      for (int64_t i = j; i < n; ++i) a[(size_t)(i * n + j)] = 0.0;
      continue;
    }
    const double l = std::sqrt(d);
    a[(size_t)(j * n + j)] = l;
    for (int64_t i = j + 1; i < n; ++i) {
      double s = a[(size_t)(i * n + j)];
      for (int64_t k = 0; k < j; ++k) s -= a[(size_t)(i * n + k)] * a[(size_t)(j * n + k)];
      a[(size_t)(i * n + j)] = s / l;
    }
  }
  std::vector<double> inv((size_t)(n * n), 0.0);
  // one column of the inverse per solve: L y = e_c, then L^T x = y (dropped
  // pivots: zero); columns in parallel, each thread with its own work vector
  parallel_chunks(n, [&](int, int64_t c0, int64_t c1) {
    std::vector<double> y((size_t)n);
    for (int64_t c = c0; c < c1; ++c) {
      for (int64_t i = 0; i < n; ++i) {
        if (dropped[(size_t)i]) {
          y[(size_t)i] = 0.0;
          continue;
        }
        double s = i == c ? 1.0 : 0.0;
        for (int64_t k = 0; k < i; ++k) s -= a[(size_t)(i * n + k)] * y[(size_t)k];
        y[(size_t)i] = s / a[(size_t)(i * n + i)];
      }
      for (int64_t i = n - 1; i >= 0; --i) {
        if (dropped[(size_t)i]) {
          y[(size_t)i] = 0.0;
          continue;
        }
        double s = y[(size_t)i];
        for (int64_t k = i + 1; k < n; ++k) s -= a[(size_t)(k * n + i)] * y[(size_t)k];
        y[(size_t)i] = s / a[(size_t)(i * n + i)];
      }
      for (int64_t i = 0; i < n; ++i) inv[(size_t)(i * n + c)] = y[(size_t)i];
    }
  }, 64);
  return inv;
}

void finish_level(AmgLevel &L) {
  const HostCsr &A = L.A;
  L.dinv.resize((size_t)A.n);
  double gersh = 0.0;
  for (int64_t i = 0; i < A.n; ++i) {
    const double d = diag_of(A, i);
    L.dinv[(size_t)i] = d != 0.0 ? 1.0 / d : 0.0;
    double s = 0.0;
    for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k) s += std::fabs(A.v[(size_t)k]);
    gersh = std::max(gersh, s * std::fabs(L.dinv[(size_t)i]));
  }
  L.rho = spectral_radius(A, L.dinv);
  L.lmax = std::min(1.1 * L.rho, gersh);
  amg_chebyshev(L.lmax, &L.c0, &L.c1, &L.c2);
}

}  // namespace

void amg_chebyshev(double lmax, double *c0, double *c1, double *c2) {
  if (!(lmax > 0.0)) {
    *c0 = *c1 = *c2 = 0.0;
    return;
  }
  const double lmin = lmax / 30.0;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  const double r0 = 1.0 / sigma, r1 = 1.0 / (2.0 * sigma - r0);
  *c0 = 1.0 / theta;
  *c1 = r1 * r0;
  *c2 = 2.0 * r1 / delta;
}

void host_spmv(const HostCsr &A, const double *x, double *y) {
  parallel_chunks(A.n, [&](int, int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      double s = 0.0;
      for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k)
        s += A.v[(size_t)k] * x[A.ci[(size_t)k]];
      y[i] = s;
    }
  });
}

AmgHierarchy amg_build(const HostCsr &A0, const AmgParams &prm) {
  if (A0.n != A0.m) throw std::invalid_argument("AMG: the operator must be square");
  AmgHierarchy H;
  H.levels.emplace_back();
  H.levels.back().A = drop_zeros(A0);
  for (;;) {
    AmgLevel &L = H.levels.back();
    finish_level(L);
    const int64_t n = L.A.n;
    if (n <= prm.max_coarse || (int)H.levels.size() >= prm.max_levels) break;
    std::vector<double> d((size_t)n);
    for (int64_t i = 0; i < n; ++i) d[(size_t)i] = diag_of(L.A, i);
    std::vector<int32_t> sp, sc;
    strength(L.A, d, prm.theta, sp, sc);
    const int64_t na = aggregate(n, sp, sc, L.agg);
    if (na == 0 || na * 10 > n * 9) {
      L.agg.clear();
      break;
    }
    // tentative prolongator: one entry per aggregated row, columns normalised
    std::vector<int32_t> cnt((size_t)na, 0);
    for (int32_t g : L.agg)
      if (g >= 0) cnt[(size_t)g]++;
    HostCsr T;
    T.n = n;
    T.m = na;
    T.rp.assign((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; ++i) T.rp[(size_t)i + 1] = T.rp[(size_t)i] + (L.agg[(size_t)i] >= 0);
    for (int64_t i = 0; i < n; ++i)
      if (L.agg[(size_t)i] >= 0) {
        T.ci.push_back(L.agg[(size_t)i]);
        T.v.push_back(1.0 / std::sqrt((double)cnt[(size_t)L.agg[(size_t)i]]));
      }
    // P = T - omega D^-1 (A T), omega = 1.2 / lmax (lmax bounds the spectral
    // radius of D^-1 A from above, so omega rho <= 1.2).  The textbook weight
    // 4 / (3 rho) sits on a cliff for these Laplacians: with omega rho above
    // ~1.27 the smoothed P makes the next level's strong couplings aggregate
    // too coarsely (Kuhn 64^3: level 2 of ~3,300 instead of ~9,000 rows and 36
    // instead of 19 CG iterations; 96^3: 55 instead of 24); 1.0-1.25 all give
    // the short counts.
    const HostCsr AT = spgemm(L.A, T);
    const double omega = L.lmax > 0.0 ? 1.2 / L.lmax : 0.0;
    HostCsr &P = L.P;
    P.n = n;
    P.m = na;
    P.rp.assign((size_t)n + 1, 0);
    P.ci.reserve((size_t)AT.nnz() + (size_t)n);
    P.v.reserve((size_t)AT.nnz() + (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      const int32_t g = L.agg[(size_t)i];
      const double tv = g >= 0 ? T.v[(size_t)T.rp[(size_t)i]] : 0.0;
      const double s = omega * L.dinv[(size_t)i];
      bool placed = g < 0;
      for (int32_t k = AT.rp[(size_t)i]; k < AT.rp[(size_t)i + 1]; ++k) {
        const int32_t j = AT.ci[(size_t)k];
        if (!placed && g < j) {
          P.ci.push_back(g);
          P.v.push_back(tv);
          placed = true;
        }
        double v = -s * AT.v[(size_t)k];
        if (j == g) {
          v = tv + v;
          placed = true;
        }
        if (v == 0.0) continue;
        P.ci.push_back(j);
        P.v.push_back(v);
      }
      if (!placed) {
        P.ci.push_back(g);
        P.v.push_back(tv);
      }
      P.rp[(size_t)i + 1] = (int32_t)P.ci.size();
    }
    L.R = transpose(P);
    HostCsr Ac = spgemm(L.R, spgemm(L.A, P));
    H.levels.emplace_back();
    H.levels.back().A = std::move(Ac);
  }
  const HostCsr &Ac = H.levels.back().A;
  H.coarse_n = Ac.n;
  if (Ac.n <= kMaxDense)
    H.coarse_inv = dense_inverse(Ac);
  else
    H.coarse_jacobi = true;
  return H;
}

}  // namespace cfsa
