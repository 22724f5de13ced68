// CPU FP64 oracle (test infrastructure; see oracle.h).  Written as a direct,
// sequential restatement of the reference: containers and loop orders follow
// the reference functions cited per block, not the GPU library.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct V3 {
  double x = 0, y = 0, z = 0;
  double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 mul(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
V3 divs(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
// Eigen cross3_impl (generic): (l1 r2 - l2 r1, l2 r0 - l0 r2, l0 r1 - l1 r0)
V3 cross(V3 l, V3 r) {
  return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x};
}
// Vector3d dot / squaredNorm: (p0 + p1) + p2
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 normalized(V3 a) {
  double z = dot(a, a);
  if (z > 0.0) return divs(a, std::sqrt(z));
  return a;
}

constexpr uint8_t WALL = 0, INLET = 1, OUTLET = 2, INTERNAL = 3, UNDEFINED = 4;

}  // namespace

// ===========================================================================
// Mesh restatement (reference src/Mesh/Mesh.cpp)
// ===========================================================================
struct OTet {
  double volume = 0.0;
  std::array<int, 4> vertids{};
  std::array<int, 4> faceids{{-1, -1, -1, -1}};
  std::array<int, 4> neighbors{{-1, -1, -1, -1}};
  V3 centroid;
};
struct OFace {
  bool is_ecm = false;
  int tet_a = -1, tet_b = -1, face_id_a = -1, face_id_b = -1;
  double area = 0.0;
  std::array<int, 3> vertids{};
  V3 normal, center;
};

struct orc_mesh {
  std::vector<V3> vertices;
  std::vector<V3> edge_nodes;
  std::vector<OFace> faces;
  std::vector<OTet> tets;
  std::vector<uint8_t> p1_bc, p2_bc;
  std::vector<std::array<V3, 4>> grads;
  std::map<std::pair<int, int>, int> edge_to_node_id;
  std::vector<std::array<int, 6>> tet_edge_nodes;

  // Mesh.cpp:494-588
  void buildConnectivity() {
    faces.clear();
    std::map<std::array<int, 3>, OFace> face_map;
    static const int lf[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    for (int t = 0; t < (int)tets.size(); ++t) {
      for (int f = 0; f < 4; ++f) {
        std::array<int, 3> fv = {tets[t].vertids[lf[f][0]], tets[t].vertids[lf[f][1]],
                                 tets[t].vertids[lf[f][2]]};
        std::array<int, 3> key = fv;
        std::sort(key.begin(), key.end());
        auto it = face_map.find(key);
        if (it == face_map.end()) {
          OFace face;
          face.vertids = fv;
          face.tet_a = t;
          face.face_id_a = f;
          face_map[key] = face;
        } else {
          it->second.tet_b = t;
          it->second.face_id_b = f;
        }
      }
    }
    std::map<std::array<int, 3>, int> index;
    for (auto &kv : face_map) {
      index[kv.first] = (int)faces.size();
      faces.push_back(kv.second);
    }
    for (auto &tet : tets)
      for (int f = 0; f < 4; ++f) {
        std::array<int, 3> key = {tet.vertids[lf[f][0]], tet.vertids[lf[f][1]],
                                  tet.vertids[lf[f][2]]};
        std::sort(key.begin(), key.end());
        tet.faceids[f] = index[key];
      }
    for (const auto &face : faces)
      if (face.tet_a >= 0 && face.tet_b >= 0) {
        tets[face.tet_a].neighbors[face.face_id_a] = face.tet_b;
        tets[face.tet_b].neighbors[face.face_id_b] = face.tet_a;
      }
  }

  // Mesh.cpp:448-471
  void computeGeometry() {
    for (auto &tet : tets) {
      const V3 &a = vertices[tet.vertids[0]], &b = vertices[tet.vertids[1]],
               &c = vertices[tet.vertids[2]], &d = vertices[tet.vertids[3]];
      tet.centroid = divs(add(add(add(a, b), c), d), 4.0);
      tet.volume = std::fabs(dot(cross(sub(b, a), sub(c, a)), sub(d, a)) / 6.0);
    }
    for (auto &face : faces) {
      const V3 &a = vertices[face.vertids[0]], &b = vertices[face.vertids[1]],
               &c = vertices[face.vertids[2]];
      V3 ab = sub(b, a), ac = sub(c, a);
      face.normal = normalized(cross(ac, ab));
      V3 cr = cross(ab, ac);
      face.area = 0.5 * std::sqrt(dot(cr, cr));
      face.center = divs(add(add(a, b), c), 3.0);
    }
    computeShapeFunctionGradients();
  }

  // Mesh.cpp:590-614
  void computeShapeFunctionGradients() {
    grads.resize(tets.size());
    for (size_t t = 0; t < tets.size(); ++t) {
      const V3 &p0 = vertices[tets[t].vertids[0]], &p1 = vertices[tets[t].vertids[1]],
               &p2 = vertices[tets[t].vertids[2]], &p3 = vertices[tets[t].vertids[3]];
      V3 e1 = sub(p1, p0), e2 = sub(p2, p0), e3 = sub(p3, p0);
      double sixv = dot(e1, cross(e2, e3));
      double inv = 1.0 / sixv;
      grads[t][0] = mul(inv, cross(sub(p3, p1), sub(p2, p1)));
      grads[t][1] = mul(inv, cross(sub(p2, p0), sub(p3, p0)));
      grads[t][2] = mul(inv, cross(sub(p3, p0), sub(p1, p0)));
      grads[t][3] = mul(inv, cross(sub(p1, p0), sub(p2, p0)));
    }
  }

  // Mesh.cpp:615-639
  void ensureConsistentFaceNormals() {
    const double eps = 1e-12;
    for (auto &face : faces) {
      if (face.tet_a >= 0 && face.tet_b >= 0) {
        V3 desired = normalized(sub(tets[face.tet_b].centroid, tets[face.tet_a].centroid));
        if (dot(face.normal, desired) <= eps) {
          face.normal = neg(face.normal);
          std::swap(face.vertids[1], face.vertids[2]);
        } else if (face.tet_a >= 0) {
          V3 out = normalized(sub(face.center, tets[face.tet_a].centroid));
          if (dot(face.normal, out) <= eps) {
            face.normal = neg(face.normal);
            std::swap(face.vertids[1], face.vertids[2]);
          }
        }
      }
    }
  }

  // Mesh.cpp:641-677
  void buildP2EdgeNodes() {
    edge_nodes.clear();
    edge_to_node_id.clear();
    tet_edge_nodes.assign(tets.size(), std::array<int, 6>{});
    int id = 0;
    static const int ep[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (size_t t = 0; t < tets.size(); ++t)
      for (int e = 0; e < 6; ++e) {
        int v1 = tets[t].vertids[ep[e][0]], v2 = tets[t].vertids[ep[e][1]];
        if (v1 > v2) std::swap(v1, v2);
        auto key = std::make_pair(v1, v2);
        auto it = edge_to_node_id.find(key);
        if (it == edge_to_node_id.end()) {
          edge_nodes.push_back(mul(0.5, add(vertices[v1], vertices[v2])));
          edge_to_node_id[key] = id;
          tet_edge_nodes[t][e] = id;
          ++id;
        } else {
          tet_edge_nodes[t][e] = it->second;
        }
      }
    p2_bc.assign(edge_nodes.size(), UNDEFINED);
  }

  void finish() {
    buildConnectivity();
    computeGeometry();
    ensureConsistentFaceNormals();
    computeShapeFunctionGradients();
    buildP2EdgeNodes();
    p1_bc.assign(vertices.size(), UNDEFINED);
  }

  // Mesh.hpp:198-215
  void setP2BoundariesFromP1Boundaries() {
    for (const auto &edge : edge_to_node_id) {
      const uint8_t a = p1_bc[edge.first.first], b = p1_bc[edge.first.second];
      p2_bc[edge.second] = (a == b && a != INTERNAL) ? a : INTERNAL;
    }
  }

  // Mesh.cpp:325-417
  void setupBoundaryConditions(const double dir[3], double coverage) {
    if (faces.empty() || vertices.size() < 3) return;
    V3 mn = vertices[0], mx = vertices[0];
    for (const auto &v : vertices)
      for (int c = 0; c < 3; ++c) {
        mn[c] = std::min(mn[c], v[c]);
        mx[c] = std::max(mx[c], v[c]);
      }
    V3 size = sub(mx, mn);
    int axis = 0;
    if (size.y > size.x && size.y > size.z)
      axis = 1;
    else if (size.z > size.x && size.z > size.y)
      axis = 2;
    const double tol = size[axis] * 0.01 * coverage;
    const double inlet = dir[axis] > 0 ? mn[axis] : mx[axis];
    const double outlet = dir[axis] > 0 ? mx[axis] : mn[axis];
    p1_bc.assign(vertices.size(), INTERNAL);
    for (auto &face : faces) {
      if (face.tet_b != -1) {
        face.is_ecm = false;
        continue;
      }
      face.is_ecm = true;
      for (int v : face.vertids) p1_bc[v] = WALL;
      double di = std::abs(face.center[axis] - inlet);
      double dout = std::abs(face.center[axis] - outlet);
      if (di < tol) {
        face.is_ecm = false;
        for (int v : face.vertids) p1_bc[v] = INLET;
      } else if (dout < tol) {
        face.is_ecm = false;
        for (int v : face.vertids) p1_bc[v] = OUTLET;
      }
    }
    setP2BoundariesFromP1Boundaries();
  }
};

// ===========================================================================
// Sparse matrices with Eigen semantics
// ===========================================================================
namespace {

struct Trip {
  int r, c;
  double v;
};

struct Csr {
  int n = 0;
  std::vector<int> rp, col;
  std::vector<double> val;
  int find(int r, int c) const {
    for (int k = rp[r]; k < rp[r + 1]; ++k)
      if (col[k] == c) return k;
    return -1;
  }
};

// SparseMatrix::setFromTriplets: entries of a row in triplet order, duplicates
// folded into the first occurrence left to right, then columns sorted.
Csr from_triplets(int n, const std::vector<Trip> &trips) {
  std::vector<std::vector<std::pair<int, double>>> rows(n);
  std::vector<int> wi(n, -1);
  for (const auto &t : trips) rows[t.r].push_back({t.c, t.v});
  Csr A;
  A.n = n;
  A.rp.assign(n + 1, 0);
  for (int r = 0; r < n; ++r) {
    std::vector<std::pair<int, double>> coll;
    for (const auto &e : rows[r]) {
      if (wi[e.first] >= 0) {
        coll[wi[e.first]].second = coll[wi[e.first]].second + e.second;
      } else {
        wi[e.first] = (int)coll.size();
        coll.push_back(e);
      }
    }
    for (const auto &e : coll) wi[e.first] = -1;
    std::sort(coll.begin(), coll.end(),
              [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
                return a.first < b.first;
              });
    for (const auto &e : coll) {
      A.col.push_back(e.first);
      A.val.push_back(e.second);
    }
    A.rp[r + 1] = (int)A.col.size();
  }
  return A;
}

// sparse * dense: res = 0; res(i) += A(i,j) * x(j), ascending j per row
std::vector<double> spmv(const Csr &A, const std::vector<double> &x) {
  std::vector<double> y(A.n, 0.0);
  for (int i = 0; i < A.n; ++i) {
    double s = 0.0;
    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.val[k] * x[A.col[k]];
    y[i] = s;
  }
  return y;
}

// Eigen redux over a dynamic-size double vector with 2-wide packets (SSE2):
// two packet accumulators over blocks of 4, lane-wise combine, one optional
// trailing packet, horizontal add, then the remaining scalar.
double eigen_dot(const double *a, const double *b, size_t n) {
  if (n == 0) return 0.0;
  const size_t ps = 2;
  const size_t as2 = (n / (2 * ps)) * (2 * ps);
  const size_t as = (n / ps) * ps;
  if (as == 0) return a[0] * b[0];
  double p0[2] = {a[0] * b[0], a[1] * b[1]};
  if (as > ps) {
    double p1[2] = {a[2] * b[2], a[3] * b[3]};
    for (size_t i = 2 * ps; i < as2; i += 2 * ps) {
      p0[0] = p0[0] + a[i] * b[i];
      p0[1] = p0[1] + a[i + 1] * b[i + 1];
      p1[0] = p1[0] + a[i + 2] * b[i + 2];
      p1[1] = p1[1] + a[i + 3] * b[i + 3];
    }
    p0[0] = p0[0] + p1[0];
    p0[1] = p0[1] + p1[1];
    if (as > as2) {
      p0[0] = p0[0] + a[as2] * b[as2];
      p0[1] = p0[1] + a[as2 + 1] * b[as2 + 1];
    }
  }
  double res = p0[0] + p0[1];
  for (size_t i = as; i < n; ++i) res = res + a[i] * b[i];
  return res;
}
// Dot2 of Ogita, Rump & Oishi ("Accurate Sum and Dot Product", SIAM J. Sci.
// Comput. 26(6), 2005): exact products (TwoProduct via fma) and error-free
// additions (TwoSum) accumulated in (p, s); the result p + s is as accurate as
// a sum in twice the working precision, so its rounding does not depend on
// the order of the terms (up to values within ~n*u^2*sum|a_i b_i| of a
// rounding boundary).  Used, sequentially, for every Krylov inner product.
double dot2(const double *a, const double *b, size_t n) {
  double p = 0.0, s = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double h = a[i] * b[i];
    const double r = std::fma(a[i], b[i], -h);
    const double t = p + h;
    const double z = t - p;
    const double q = (p - (t - z)) + (h - z);
    p = t;
    s = s + (q + r);
  }
  return p + s;
}

// 0: Dot2 (default; what the GPU library evaluates), 1: Eigen's packet order
int g_dot_mode = 0;

double edot(const std::vector<double> &a, const std::vector<double> &b) {
  return g_dot_mode == 1 ? eigen_dot(a.data(), b.data(), a.size())
                         : dot2(a.data(), b.data(), a.size());
}

struct KResult {
  int iters = 0;
  double err = 0.0;
  bool ok = true;
  int restarts = 0;
};

// Multiplier of Eigen's eps^2 in the BiCGSTAB restart test (1 = Eigen); a
// test knob that makes the restart branch reachable on small systems.
double g_restart_scale = 1.0;
int g_last_restarts = 0;

std::vector<double> inv_diag(const Csr &A) {
  std::vector<double> d(A.n, 1.0);
  for (int j = 0; j < A.n; ++j) {
    int k = A.find(j, j);
    if (k >= 0 && A.val[k] != 0.0) d[j] = 1.0 / A.val[k];
  }
  return d;
}

// Eigen::BiCGSTAB<SparseMatrix, DiagonalPreconditioner>::solve with x0 = 0
// (Eigen/src/IterativeLinearSolvers/BiCGSTAB.h, internal::bicgstab)
// Squared norm used by the stopping test: |v|^2 (reference) or, with the
// scaled rule, |D^-1 v|^2 where D^-1 is the Jacobi scaling.
double stop_norm2(const std::vector<double> &v, const std::vector<double> &D, bool scaled) {
  if (!scaled) return edot(v, v);
  std::vector<double> w(v.size());
  for (size_t k = 0; k < v.size(); ++k) w[k] = D[k] * v[k];
  return edot(w, w);
}

KResult bicgstab(const Csr &A, const std::vector<double> &rhs, std::vector<double> &x,
                 double tol, int maxit, bool scaled = false) {
  const int n = A.n;
  KResult res;
  x.assign(n, 0.0);
  const std::vector<double> D = inv_diag(A);
  std::vector<double> Ax = spmv(A, x);
  std::vector<double> r(n), r0;
  for (int i = 0; i < n; ++i) r[i] = rhs[i] - Ax[i];
  r0 = r;
  double r0_sqnorm = edot(r0, r0);
  const double rhs_sqnorm = edot(rhs, rhs);
  if (rhs_sqnorm == 0) {
    x.assign(n, 0.0);
    res.iters = 0;
    res.err = 0.0;
    return res;
  }
  double rho = 1, alpha = 1, w = 1;
  std::vector<double> v(n, 0.0), p(n, 0.0), y(n), z(n), s(n), t(n);
  const double b_stop = scaled ? stop_norm2(rhs, D, true) : rhs_sqnorm;
  const double tol2 = tol * tol * b_stop;
  const double eps = std::numeric_limits<double>::epsilon();
  const double eps2 = eps * eps;
  int i = 0, restarts = 0;
  while (stop_norm2(r, D, scaled) > tol2 && i < maxit) {
    double rho_old = rho;
    rho = edot(r0, r);
    if (std::abs(rho) < g_restart_scale * eps2 * r0_sqnorm) {
      std::vector<double> Axr = spmv(A, x);
      for (int k = 0; k < n; ++k) r[k] = rhs[k] - Axr[k];
      r0 = r;
      rho = r0_sqnorm = edot(r, r);
      if (restarts++ == 0) i = 0;
    }
    double beta = (rho / rho_old) * (alpha / w);
    for (int k = 0; k < n; ++k) p[k] = r[k] + beta * (p[k] - w * v[k]);
    for (int k = 0; k < n; ++k) y[k] = D[k] * p[k];
    v = spmv(A, y);
    alpha = rho / edot(r0, v);
    for (int k = 0; k < n; ++k) s[k] = r[k] - alpha * v[k];
    for (int k = 0; k < n; ++k) z[k] = D[k] * s[k];
    t = spmv(A, z);
    double tmp = edot(t, t);
    if (tmp > 0.0)
      w = edot(t, s) / tmp;
    else
      w = 0.0;
    for (int k = 0; k < n; ++k) x[k] = x[k] + (alpha * y[k] + w * z[k]);
    for (int k = 0; k < n; ++k) r[k] = s[k] - w * t[k];
    ++i;
  }
  res.err = std::sqrt(stop_norm2(r, D, scaled) / b_stop);
  res.iters = i;
  res.ok = res.err <= tol;
  res.restarts = restarts;
  return res;
}

// Jacobi PCG with the loop and stopping rule of Eigen's conjugate_gradient
// (full matrix product).
KResult pcg(const Csr &A, const std::vector<double> &rhs, std::vector<double> &x, double tol,
            int maxit, bool scaled = false) {
  const int n = A.n;
  KResult res;
  x.assign(n, 0.0);
  const std::vector<double> D = inv_diag(A);
  std::vector<double> r = rhs;
  const double rhsNorm2 = stop_norm2(rhs, D, scaled);
  if (rhsNorm2 == 0) {
    res.iters = 0;
    res.err = 0;
    return res;
  }
  const double threshold = std::max(tol * tol * rhsNorm2, std::numeric_limits<double>::min());
  double rn2 = stop_norm2(r, D, scaled);
  if (rn2 < threshold) {
    res.iters = 0;
    res.err = std::sqrt(rn2 / rhsNorm2);
    return res;
  }
  std::vector<double> p(n), z(n), q;
  for (int k = 0; k < n; ++k) p[k] = D[k] * r[k];
  double absNew = edot(r, p);
  int i = 0;
  while (i < maxit) {
    q = spmv(A, p);
    double alpha = absNew / edot(p, q);
    for (int k = 0; k < n; ++k) x[k] = x[k] + alpha * p[k];
    for (int k = 0; k < n; ++k) r[k] = r[k] - alpha * q[k];
    rn2 = stop_norm2(r, D, scaled);
    if (rn2 < threshold) break;
    for (int k = 0; k < n; ++k) z[k] = D[k] * r[k];
    double absOld = absNew;
    absNew = edot(r, z);
    double beta = absNew / absOld;
    for (int k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
    ++i;
  }
  res.err = std::sqrt(rn2 / rhsNorm2);
  res.iters = i;
  res.ok = res.err <= tol;
  return res;
}

}  // namespace

// ===========================================================================
// Navier-Stokes restatement (reference src/FEM/NavierStokes.cpp)
// ===========================================================================
struct orc_ns {
  const orc_mesh *mesh = nullptr;
  int nv = 0, nt = 0, ne = 0, nu = 0;
  double rho = 1000, mu = 0.001;
  int reference_node = -1, n_corrections = 1;
  double relax_u = 1.0, relax_p = 1.0, dt = 1e-4, time = 0.0;
  std::vector<V3> velocity, velocity_star, g;  // g: inlet value per node
  std::vector<double> pressure, pressure_correction, inv_lumped_mass;
  double outlet_pressure = 0.0;
  int outlet_type = 0;
  double momentum_tol = 1e-6, pressure_tol = 1e-8;
  int max_iter = 1000, pressure_solver = 0;
  int momentum_solver = 0;  // extension: 1 Jacobi CG on the (SPD) momentum matrix
  bool stop_scaled = false;  // extension: Jacobi-scaled residual stopping rule
  Csr M, K, Gx, Gy, Gz, L;
  int it_mom[3] = {0, 0, 0};
  double err_mom[3] = {0, 0, 0};
  int it_p[8] = {0};
  double err_p[8] = {0};
  int n_p = 0;

  uint8_t bc1(int i) const { return mesh->p1_bc[i]; }
  uint8_t bc2(int k) const { return mesh->p2_bc[k]; }

  // :70-133
  void buildGradientMatrices() {
    std::vector<Trip> tx, ty, tz;
    for (int t = 0; t < nt; ++t) {
      const auto &tet = mesh->tets[t];
      const auto &gr = mesh->grads[t];
      const double vol = tet.volume;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          const double contrib = vol / 4.0;
          tx.push_back({tet.vertids[i], tet.vertids[j], contrib * gr[j].x});
          ty.push_back({tet.vertids[i], tet.vertids[j], contrib * gr[j].y});
          tz.push_back({tet.vertids[i], tet.vertids[j], contrib * gr[j].z});
        }
    }
    Gx = from_triplets(nv, tx);
    Gy = from_triplets(nv, ty);
    Gz = from_triplets(nv, tz);
  }
  // :135-225
  void buildStiffnessMatrix() {
    std::vector<Trip> tr;
    const double m = 1.0 / mu;
    static const int ev[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int t = 0; t < nt; ++t) {
      const auto &tet = mesh->tets[t];
      const auto &gr = mesh->grads[t];
      const auto &en = mesh->tet_edge_nodes[t];
      const double vol = tet.volume;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          tr.push_back({tet.vertids[i], tet.vertids[j], m * vol * dot(gr[i], gr[j])});
      for (int e = 0; e < 6; ++e) {
        const int ei = nv + en[e];
        V3 ge = add(gr[ev[e][0]], gr[ev[e][1]]);
        for (int j = 0; j < 4; ++j) {
          const double k = m * vol * dot(ge, gr[j]);
          tr.push_back({ei, tet.vertids[j], k});
          tr.push_back({tet.vertids[j], ei, k});
        }
        for (int f = 0; f < 6; ++f) {
          V3 gf = add(gr[ev[f][0]], gr[ev[f][1]]);
          tr.push_back({ei, nv + en[f], m * vol * dot(ge, gf)});
        }
      }
    }
    K = from_triplets(nu, tr);
  }
  // :227-310
  void buildMassMatrix() {
    std::vector<Trip> tr;
    for (int t = 0; t < nt; ++t) {
      const auto &tet = mesh->tets[t];
      const auto &en = mesh->tet_edge_nodes[t];
      const double mass = rho * tet.volume;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          tr.push_back({tet.vertids[i], tet.vertids[j], (i == j) ? mass / 10.0 : mass / 20.0});
      for (int e = 0; e < 6; ++e) {
        const int ei = nv + en[e];
        for (int i = 0; i < 4; ++i) {
          tr.push_back({ei, tet.vertids[i], mass / 30.0});
          tr.push_back({tet.vertids[i], ei, mass / 30.0});
        }
        for (int f = 0; f < 6; ++f)
          tr.push_back({ei, nv + en[f], (e == f) ? mass / 15.0 : mass / 60.0});
      }
    }
    M = from_triplets(nu, tr);
    std::vector<double> ones(nu, 1.0);
    std::vector<double> rs = spmv(M, ones);
    inv_lumped_mass.assign(nu, 0.0);
    for (int i = 0; i < nu; ++i) inv_lumped_mass[i] = std::abs(rs[i]) > 1e-14 ? 1.0 / rs[i] : 1.0;
  }
  // :316-356
  void buildPoissonMatrix() {
    std::vector<Trip> tr;
    for (int t = 0; t < nt; ++t) {
      const auto &tet = mesh->tets[t];
      const auto &gr = mesh->grads[t];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          tr.push_back({tet.vertids[i], tet.vertids[j], tet.volume * dot(gr[i], gr[j])});
    }
    L = from_triplets(nv, tr);
  }

  // :10-68
  void initialize(const double inlet0[3]) {
    nv = (int)mesh->vertices.size();
    nt = (int)mesh->tets.size();
    ne = (int)mesh->edge_nodes.size();
    nu = nv + ne;
    velocity.assign(nu, V3{});
    velocity_star.assign(nu, V3{});
    pressure.assign(nv, 0.0);
    pressure_correction.assign(nv, 0.0);
    g.assign(nu, V3{inlet0[0], inlet0[1], inlet0[2]});
    buildMassMatrix();
    buildStiffnessMatrix();
    buildGradientMatrices();
    buildPoissonMatrix();
    for (int i = 0; i < nv; ++i)
      if (bc1(i) == INLET) velocity[i] = g[i];
    for (int k = 0; k < ne; ++k)
      if (bc2(k) == INLET) velocity[nv + k] = g[nv + k];
    if (reference_node < 0 && outlet_type == 1) {
      for (int i = 0; i < nv; ++i)
        if (bc1(i) == INTERNAL) {
          reference_node = i;
          break;
        }
      if (reference_node < 0 && nv > 0) reference_node = 0;
    }
  }

  uint8_t bc(int i) const { return i < nv ? bc1(i) : bc2(i - nv); }

  // :526-584
  std::vector<V3> computeAdvectionRHS() {
    std::vector<V3> conv(nu, V3{});
    for (int t = 0; t < nt; ++t) {
      const auto &tet = mesh->tets[t];
      const auto &gr = mesh->grads[t];
      const auto &en = mesh->tet_edge_nodes[t];
      const double vol = tet.volume;
      V3 ue{};
      for (int i = 0; i < 4; ++i) ue = add(ue, mul(0.25, velocity[tet.vertids[i]]));
      double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int j = 0; j < 4; ++j)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) G[a][b] = G[a][b] + velocity[tet.vertids[j]][a] * gr[j][b];
      V3 c;
      for (int a = 0; a < 3; ++a) c[a] = (G[a][0] * ue.x + G[a][1] * ue.y) + G[a][2] * ue.z;
      for (int i = 0; i < 4; ++i)
        conv[tet.vertids[i]] = sub(conv[tet.vertids[i]], mul(vol / 4.0, c));
      for (int i = 0; i < 6; ++i) conv[nv + en[i]] = sub(conv[nv + en[i]], mul(vol / 24.0, c));
    }
    return conv;
  }

  // :670-700
  std::vector<double> computeDivergence(const std::vector<V3> &u) {
    std::vector<double> div(nv, 0.0);
    for (int t = 0; t < nt; ++t) {
      const auto &tet = mesh->tets[t];
      const auto &gr = mesh->grads[t];
      double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int j = 0; j < 4; ++j)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) G[a][b] = G[a][b] + u[tet.vertids[j]][a] * gr[j][b];
      const double d = (G[0][0] + G[1][1]) + G[2][2];
      for (int i = 0; i < 4; ++i) div[tet.vertids[i]] = div[tet.vertids[i]] + (tet.volume / 4.0) * d;
    }
    return div;
  }

  Csr lhs_matrix() const {
    Csr A = M;
    for (size_t k = 0; k < A.val.size(); ++k) A.val[k] = M.val[k] / dt + K.val[k];
    for (int i = 0; i < nu; ++i)
      if (bc(i) == INLET || bc(i) == WALL) A.val[A.find(i, i)] += 1e10;
    return A;
  }
  Csr lbc_matrix() const {
    Csr A = L;
    if (outlet_type == 1) {
      if (reference_node >= 0 && reference_node < nv)
        A.val[A.find(reference_node, reference_node)] += 1e12;
    } else if (outlet_type == 0) {
      for (int i = 0; i < nv; ++i)
        if (bc1(i) == OUTLET) A.val[A.find(i, i)] += 1e12;
    }
    return A;
  }

  // :416-524
  int solveMomentumPredictor() {
    std::vector<V3> conv = computeAdvectionRHS();
    const Csr lhs = lhs_matrix();
    for (int comp = 0; comp < 3; ++comp) {
      std::vector<double> rhs(nu, 0.0), uc(nu);
      for (int i = 0; i < nu; ++i) uc[i] = velocity[i][comp];
      std::vector<double> Mu = spmv(M, uc);
      for (int i = 0; i < nu; ++i) rhs[i] = rhs[i] + Mu[i] / dt;
      const Csr &G = comp == 0 ? Gx : (comp == 1 ? Gy : Gz);
      std::vector<double> gp = spmv(G, pressure);
      for (int i = 0; i < nv; ++i) rhs[i] = rhs[i] - gp[i];
      for (int i = 0; i < nu; ++i) rhs[i] = rhs[i] + conv[i][comp];
      for (int i = 0; i < nu; ++i) {
        const uint8_t b = bc(i);
        if (b == INLET || b == WALL) {
          const double up = (b == INLET) ? g[i][comp] : 0.0;
          rhs[i] = rhs[i] + 1e10 * up;
        }
      }
      std::vector<double> sol;
      KResult r = momentum_solver == 1 ? pcg(lhs, rhs, sol, momentum_tol, max_iter, stop_scaled)
                                       : bicgstab(lhs, rhs, sol, momentum_tol, max_iter, stop_scaled);
      it_mom[comp] = r.iters;
      err_mom[comp] = r.err;
      if (!r.ok) return 2;
      for (int i = 0; i < nu; ++i) velocity_star[i][comp] = sol[i];
    }
    return 0;
  }

  // :586-668
  int solvePressurePoisson() {
    std::vector<double> div = computeDivergence(velocity_star);
    std::vector<double> rhs(nv);
    for (int i = 0; i < nv; ++i) rhs[i] = -div[i] / dt;
    const Csr A = lbc_matrix();
    if (outlet_type == 1) {
      if (reference_node >= 0 && reference_node < nv)
        rhs[reference_node] = rhs[reference_node] + 1e12 * 0.0;
    } else if (outlet_type == 0) {
      for (int i = 0; i < nv; ++i)
        if (bc1(i) == OUTLET) rhs[i] = rhs[i] + 1e12 * outlet_pressure;
    }
    std::vector<double> x;
    KResult r = pressure_solver == 1 ? pcg(A, rhs, x, pressure_tol, max_iter, stop_scaled)
                                     : bicgstab(A, rhs, x, pressure_tol, max_iter, stop_scaled);
    const int k = std::min(n_p, 7);
    it_p[k] = r.iters;
    err_p[k] = r.err;
    n_p++;
    if (!r.ok) return 3;
    pressure_correction = x;
    for (int i = 0; i < nv; ++i) pressure[i] = pressure[i] + relax_p * pressure_correction[i];
    reinforcePressureBCs();
    if (reference_node >= 0 && reference_node < nv) {
      const double ref = pressure[reference_node];
      for (auto &p : pressure) p = p - ref;
    }
    return 0;
  }

  void reinforcePressureBCs() {
    if (outlet_type == 0)
      for (int i = 0; i < nv; ++i)
        if (bc1(i) == OUTLET) pressure[i] = outlet_pressure;
  }

  // :702-771
  void correctVelocity() {
    std::vector<double> gx = spmv(Gx, pressure_correction), gy = spmv(Gy, pressure_correction),
                        gz = spmv(Gz, pressure_correction);
    for (int i = 0; i < nv; ++i) {
      const uint8_t b = bc1(i);
      if (b == INLET) {
        velocity[i] = g[i];
        continue;
      } else if (b == WALL) {
        velocity[i] = V3{};
        continue;
      }
      if (inv_lumped_mass[i] > 1e-14) {
        const double s = -dt * inv_lumped_mass[i];
        V3 corr = mul(s, V3{gx[i], gy[i], gz[i]});
        velocity[i] = add(velocity_star[i], mul(relax_u, corr));
      } else {
        velocity[i] = velocity_star[i];
      }
    }
    for (int k = 0; k < ne; ++k) {
      const int idx = nv + k;
      const uint8_t b = bc2(k);
      if (b == INLET) {
        velocity[idx] = g[idx];
        continue;
      } else if (b == WALL) {
        velocity[idx] = V3{};
        continue;
      }
      velocity[idx] = velocity_star[idx];
    }
  }

  void reenforceVelocityBCs() {
    for (int i = 0; i < nu; ++i) {
      const uint8_t b = bc(i);
      if (b == INLET)
        velocity[i] = g[i];
      else if (b == WALL)
        velocity[i] = V3{};
    }
  }

  bool hasNans() const {
    for (const auto &v : velocity)
      if (std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z)) return true;
    for (const auto &v : velocity_star)
      if (std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z)) return true;
    for (double p : pressure)
      if (std::isnan(p)) return true;
    for (double p : pressure_correction)
      if (std::isnan(p)) return true;
    return false;
  }

  // :374-414
  int pisoStep() {
    n_p = 0;
    if (int s = solveMomentumPredictor()) return s;
    for (int c = 0; c < n_corrections; ++c) {
      if (int s = solvePressurePoisson()) return s;
      correctVelocity();
    }
    reenforceVelocityBCs();
    time += dt;
    return hasNans() ? 4 : 0;
  }

  // :841-883
  std::vector<V3> computePressureForces() const {
    std::vector<V3> out(nv, V3{});
    for (const auto &face : mesh->faces) {
      if (face.tet_b != -1) continue;
      double fp = 0.0;
      for (int v : face.vertids) fp += pressure[v];
      fp /= 3.0;
      V3 ff = mul(face.area, mul(-fp, face.normal));
      const double w = 1.0 / 3.0;
      for (int v : face.vertids) out[v] = add(out[v], mul(w, ff));
    }
    return out;
  }
  // :885-944
  std::vector<V3> computeShearStress() const {
    std::vector<V3> out(nv, V3{});
    for (const auto &face : mesh->faces) {
      if (face.tet_b != -1) continue;
      const auto &tet = mesh->tets[face.tet_a];
      const auto &gr = mesh->grads[face.tet_a];
      double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int i = 0; i < 4; ++i)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) G[a][b] = G[a][b] + velocity[tet.vertids[i]][a] * gr[i][b];
      V3 tr;
      for (int a = 0; a < 3; ++a) {
        double tau[3];
        for (int b = 0; b < 3; ++b) tau[b] = (2.0 * mu) * (0.5 * (G[a][b] + G[b][a]));
        tr[a] = (tau[0] * face.normal.x + tau[1] * face.normal.y) + tau[2] * face.normal.z;
      }
      V3 ff = mul(face.area, tr);
      const double w = 1.0 / 3.0;
      for (int v : face.vertids) out[v] = add(out[v], mul(w, ff));
    }
    return out;
  }

  // ---------------------------------------------------------------------
  // Discrete adjoint of pisoStep (build extension; Dirichlet outlet, at
  // least one corrector).  The step is
  //   b_c  = (M u_c)/dt - E G_c p + conv_c(u) + 1e10 g_c (Inlet rows)
  //   u*_c = A^-1 b_c,   A = M/dt + K + penalties,  K = K1 / viscosity
  //   p'   = Lbc^-1 (-(sum_c G_c u*_c)/dt + penalties)
  //   u'   = g (Inlet), 0 (Wall), u* - relax_u dt ilm G p' (free P1), u* (free P2)
  //   p_new = p + ncorr relax_p p' (non-outlet rows), p_out (outlet rows)
  // and this routine maps (lambda_u', lambda_p') to (lambda_u, lambda_p),
  // accumulating dJ/dg and dJ/dviscosity.  A, M, K, Lbc are symmetric, so
  // the adjoint solves reuse the forward operators.
  static std::vector<double> spmv_t(const Csr &A, const std::vector<double> &x) {
    std::vector<double> y(A.n, 0.0);
    for (int i = 0; i < A.n; ++i)
      for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) y[A.col[k]] += A.val[k] * x[i];
    return y;
  }
  KResult solve(const Csr &A, const std::vector<double> &b, std::vector<double> &x, double tol,
                int method) {
    return method == 1 ? pcg(A, b, x, tol, max_iter, stop_scaled)
                       : bicgstab(A, b, x, tol, max_iter, stop_scaled);
  }

  int adjointStep(const std::vector<V3> &u_in, const std::vector<double> &p_in,
                  std::vector<V3> &lu, std::vector<double> &lp, std::vector<V3> &lg,
                  double &dnu) {
    velocity = u_in;
    pressure = p_in;
    if (int s = solveMomentumPredictor()) return s;
    n_p = 0;
    if (int s = solvePressurePoisson()) return s;
    const std::vector<V3> us = velocity_star;
    const std::vector<double> pc = pressure_correction;
    const Csr *G[3] = {&Gx, &Gy, &Gz};
    // velocity outputs
    std::vector<V3> lus(nu, V3{});
    std::vector<double> lgp[3] = {std::vector<double>(nv, 0.0), std::vector<double>(nv, 0.0),
                                  std::vector<double>(nv, 0.0)};
    for (int i = 0; i < nu; ++i) {
      const uint8_t b = bc(i);
      if (b == INLET) {
        lg[i] = add(lg[i], lu[i]);
      } else if (b == WALL) {
      } else {
        lus[i] = add(lus[i], lu[i]);
        if (i < nv && inv_lumped_mass[i] > 1e-14) {
          const double s = -relax_u * dt * inv_lumped_mass[i];
          for (int c = 0; c < 3; ++c) lgp[c][i] = s * lu[i][c];
        }
      }
    }
    std::vector<double> lpc(nv, 0.0);
    for (int c = 0; c < 3; ++c) {
      std::vector<double> t = spmv_t(*G[c], lgp[c]);
      for (int i = 0; i < nv; ++i) lpc[i] += t[i];
    }
    // pressure outputs
    std::vector<double> lp_new(nv, 0.0);
    for (int i = 0; i < nv; ++i) {
      if (outlet_type == 0 && bc1(i) == OUTLET) continue;
      lp_new[i] = lp[i];
      lpc[i] += n_corrections * relax_p * lp[i];
    }
    // pressure solve and divergence
    std::vector<double> lrp;
    solve(lbc_matrix(), lpc, lrp, pressure_tol, pressure_solver);
    std::vector<double> ld(nv);
    for (int i = 0; i < nv; ++i) ld[i] = -lrp[i] / dt;
    for (int c = 0; c < 3; ++c) {
      std::vector<double> t = spmv_t(*G[c], ld);
      for (int i = 0; i < nv; ++i) lus[i][c] += t[i];
    }
    // momentum solves, viscosity dependence of K
    const Csr lhs = lhs_matrix();
    std::vector<V3> lu_new(nu, V3{});
    std::vector<V3> lb(nu);
    for (int c = 0; c < 3; ++c) {
      std::vector<double> rhs(nu), x, usc(nu);
      for (int i = 0; i < nu; ++i) {
        rhs[i] = lus[i][c];
        usc[i] = us[i][c];
      }
      solve(lhs, rhs, x, momentum_tol, 0);
      std::vector<double> Kus = spmv(K, usc);
      double s = 0.0;
      for (int i = 0; i < nu; ++i) s += x[i] * Kus[i];
      dnu += s / mu;
      std::vector<double> Mx = spmv(M, x);
      for (int i = 0; i < nu; ++i) {
        lb[i][c] = x[i];
        lu_new[i][c] += Mx[i] / dt;
        if (bc(i) == INLET) lg[i][c] += 1e10 * x[i];
      }
      std::vector<double> xp(x.begin(), x.begin() + nv);
      std::vector<double> t = spmv_t(*G[c], xp);
      for (int i = 0; i < nv; ++i) lp_new[i] -= t[i];
    }
    // convection: conv_i = -sum_T w_iT c_T(u), c_T = (grad u_T) u_bar_T
    for (int t = 0; t < nt; ++t) {
      const auto &tet = mesh->tets[t];
      const auto &gr = mesh->grads[t];
      const auto &en = mesh->tet_edge_nodes[t];
      const double vol = tet.volume;
      V3 lc{};
      for (int i = 0; i < 4; ++i) lc = sub(lc, mul(vol / 4.0, lb[tet.vertids[i]]));
      for (int e = 0; e < 6; ++e) lc = sub(lc, mul(vol / 24.0, lb[nv + en[e]]));
      V3 ue{};
      for (int i = 0; i < 4; ++i) ue = add(ue, mul(0.25, u_in[tet.vertids[i]]));
      double Gu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int j = 0; j < 4; ++j)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) Gu[a][b] += u_in[tet.vertids[j]][a] * gr[j][b];
      V3 gt{};  // Gu^T lc / 4
      for (int b = 0; b < 3; ++b)
        gt[b] = 0.25 * (Gu[0][b] * lc.x + Gu[1][b] * lc.y + Gu[2][b] * lc.z);
      for (int j = 0; j < 4; ++j) {
        const double s = dot(gr[j], ue);
        lu_new[tet.vertids[j]] = add(lu_new[tet.vertids[j]], add(mul(s, lc), gt));
      }
    }
    lu = lu_new;
    lp = lp_new;
    return 0;
  }

  // J = 1/2 sum_{Outlet velocity nodes} |u|^2 after nsteps from u0 = 0 with
  // inlet nodes at g, p0 = 0; g = sum_k coef_k basis_k.
  int adjointGradient(int nsteps, int nk, const double *basis, const double *coef, double *J,
                      double *dJdc, double *dJdmu) {
    g.assign(nu, V3{});
    for (int k = 0; k < nk; ++k)
      for (int i = 0; i < nu; ++i)
        for (int c = 0; c < 3; ++c) g[i][c] += coef[k] * basis[((size_t)k * nu + i) * 3 + c];
    velocity.assign(nu, V3{});
    for (int i = 0; i < nu; ++i)
      if (bc(i) == INLET) velocity[i] = g[i];
    pressure.assign(nv, 0.0);
    std::vector<std::vector<V3>> cu;
    std::vector<std::vector<double>> cp;
    for (int n = 0; n < nsteps; ++n) {
      cu.push_back(velocity);
      cp.push_back(pressure);
      if (int s = pisoStep()) return s;
    }
    double j = 0.0;
    std::vector<V3> lu(nu, V3{});
    for (int i = 0; i < nu; ++i)
      if (bc(i) == OUTLET) {
        j += 0.5 * dot(velocity[i], velocity[i]);
        lu[i] = velocity[i];
      }
    std::vector<double> lp(nv, 0.0);
    std::vector<V3> lg(nu, V3{});
    double dnu = 0.0;
    for (int n = nsteps - 1; n >= 0; --n)
      if (int s = adjointStep(cu[n], cp[n], lu, lp, lg, dnu)) return s;
    for (int i = 0; i < nu; ++i)
      if (bc(i) == INLET) lg[i] = add(lg[i], lu[i]);
    *J = j;
    for (int k = 0; k < nk; ++k) {
      double s = 0.0;
      for (int i = 0; i < nu; ++i)
        for (int c = 0; c < 3; ++c) s += lg[i][c] * basis[((size_t)k * nu + i) * 3 + c];
      dJdc[k] = s;
    }
    *dJdmu = dnu;
    return 0;
  }

  // :968-1009
  int simpleStep() {
    n_p = 0;
    if (int s = solveMomentumPredictor()) return s;
    if (int s = solvePressurePoisson()) return s;
    reinforcePressureBCs();
    correctVelocity();
    reenforceVelocityBCs();
    time += dt;
    return hasNans() ? 4 : 0;
  }
};

// ===========================================================================
// C API
// ===========================================================================
namespace {
template <class F>
int guard(F &&f) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_err = e.what();
    return 5;
  }
}
void csr_out(const Csr &A, int32_t *rp, int32_t *col, double *val) {
  if (rp) std::copy(A.rp.begin(), A.rp.end(), rp);
  if (col) std::copy(A.col.begin(), A.col.end(), col);
  if (val) std::copy(A.val.begin(), A.val.end(), val);
}
}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

double orc_dot2(const double *a, const double *b, int64_t n) { return dot2(a, b, (size_t)n); }
void orc_set_dot_mode(int mode) { g_dot_mode = mode == 1 ? 1 : 0; }
int orc_get_dot_mode(void) { return g_dot_mode; }
double orc_eigen_dot(const double *a, const double *b, int64_t n) {
  return eigen_dot(a, b, (size_t)n);
}

// y = A x with Eigen's product order (res(i) += A(i,j) x(j), ascending j),
// nrhs interleaved right-hand sides
void orc_csr_spmv(int64_t n, const int32_t *rp, const int32_t *col, const double *val,
                  const double *x, double *y, int nrhs) {
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < nrhs; ++c) {
      double s = 0.0;
      for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * x[(int64_t)col[k] * nrhs + c];
      y[i * nrhs + c] = s;
    }
}

// Mesh.cpp:146-225
int orc_mesh_structured_prism(double L, double W, double H, int nx, int ny, int nz,
                              orc_mesh **out) {
  return guard([&]() -> int {
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("nx, ny, nz must be >= 1");
    auto *m = new orc_mesh();
    const int nxp = nx + 1, nyp = ny + 1, nzp = nz + 1;
    const double dx = L / nx, dy = W / ny, dz = H / nz;
    auto idx = [&](int i, int j, int k) { return i + nxp * (j + nyp * k); };
    for (int k = 0; k < nzp; ++k)
      for (int j = 0; j < nyp; ++j)
        for (int i = 0; i < nxp; ++i)
          m->vertices.push_back(V3{(double)i * dx, (double)j * dy, (double)k * dz});
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          int v0 = idx(i, j, k), v1 = idx(i + 1, j, k), v2 = idx(i, j + 1, k),
              v3 = idx(i + 1, j + 1, k), v4 = idx(i, j, k + 1), v5 = idx(i + 1, j, k + 1),
              v6 = idx(i, j + 1, k + 1), v7 = idx(i + 1, j + 1, k + 1);
          int t[6][4] = {{v0, v1, v3, v7}, {v0, v3, v2, v7}, {v0, v2, v6, v7},
                         {v0, v6, v4, v7}, {v0, v4, v5, v7}, {v0, v5, v1, v7}};
          for (auto &q : t) {
            OTet tet;
            tet.vertids = {q[0], q[1], q[2], q[3]};
            m->tets.push_back(tet);
          }
        }
    m->finish();
    *out = m;
    return 0;
  });
}

int orc_mesh_from_arrays(int64_t nv, const double *coords, int64_t nt, const int32_t *tets,
                         orc_mesh **out) {
  return guard([&]() -> int {
    auto *m = new orc_mesh();
    for (int64_t i = 0; i < nv; ++i)
      m->vertices.push_back(V3{coords[i * 3], coords[i * 3 + 1], coords[i * 3 + 2]});
    for (int64_t t = 0; t < nt; ++t) {
      OTet tet;
      tet.vertids = {tets[t * 4], tets[t * 4 + 1], tets[t * 4 + 2], tets[t * 4 + 3]};
      m->tets.push_back(tet);
    }
    m->finish();
    *out = m;
    return 0;
  });
}

// Mesh::updateGeometry after the vertices move (Mesh.hpp:98-101): element and
// face geometry from the stored (possibly flipped) face vertex order.
int orc_mesh_update_vertices(orc_mesh *m, const double *coords) {
  for (size_t i = 0; i < m->vertices.size(); ++i)
    m->vertices[i] = V3{coords[i * 3], coords[i * 3 + 1], coords[i * 3 + 2]};
  m->computeGeometry();
  m->computeShapeFunctionGradients();
  return 0;
}

int orc_mesh_setup_bc(orc_mesh *m, const double dir[3], double coverage) {
  m->setupBoundaryConditions(dir, coverage);
  return 0;
}
int orc_mesh_set_p1_bc(orc_mesh *m, const uint8_t *bc) {
  m->p1_bc.assign(bc, bc + m->vertices.size());
  m->setP2BoundariesFromP1Boundaries();
  return 0;
}
void orc_mesh_sizes(const orc_mesh *m, int64_t out[4]) {
  out[0] = (int64_t)m->vertices.size();
  out[1] = (int64_t)m->tets.size();
  out[2] = (int64_t)m->edge_nodes.size();
  out[3] = (int64_t)m->faces.size();
}
void orc_mesh_get_vertices(const orc_mesh *m, double *out) {
  for (size_t i = 0; i < m->vertices.size(); ++i)
    for (int c = 0; c < 3; ++c) out[i * 3 + c] = m->vertices[i][c];
}
void orc_mesh_get_tets(const orc_mesh *m, int32_t *out) {
  for (size_t i = 0; i < m->tets.size(); ++i)
    for (int c = 0; c < 4; ++c) out[i * 4 + c] = m->tets[i].vertids[c];
}
void orc_mesh_get_volumes(const orc_mesh *m, double *out) {
  for (size_t i = 0; i < m->tets.size(); ++i) out[i] = m->tets[i].volume;
}
void orc_mesh_get_gradients(const orc_mesh *m, double *out) {
  for (size_t t = 0; t < m->grads.size(); ++t)
    for (int i = 0; i < 4; ++i)
      for (int c = 0; c < 3; ++c) out[t * 12 + i * 3 + c] = m->grads[t][i][c];
}
void orc_mesh_get_tet_edge_nodes(const orc_mesh *m, int32_t *out) {
  for (size_t t = 0; t < m->tet_edge_nodes.size(); ++t)
    for (int e = 0; e < 6; ++e) out[t * 6 + e] = m->tet_edge_nodes[t][e];
}
void orc_mesh_get_edge_nodes(const orc_mesh *m, double *out) {
  for (size_t i = 0; i < m->edge_nodes.size(); ++i)
    for (int c = 0; c < 3; ++c) out[i * 3 + c] = m->edge_nodes[i][c];
}
void orc_mesh_get_p1_bc(const orc_mesh *m, uint8_t *out) {
  std::copy(m->p1_bc.begin(), m->p1_bc.end(), out);
}
void orc_mesh_get_p2_bc(const orc_mesh *m, uint8_t *out) {
  std::copy(m->p2_bc.begin(), m->p2_bc.end(), out);
}
void orc_mesh_get_faces(const orc_mesh *m, int32_t *verts, int32_t *tets, double *geom) {
  for (size_t f = 0; f < m->faces.size(); ++f) {
    const auto &F = m->faces[f];
    if (verts)
      for (int a = 0; a < 3; ++a) verts[f * 3 + a] = F.vertids[a];
    if (tets) {
      tets[f * 4] = F.tet_a;
      tets[f * 4 + 1] = F.tet_b;
      tets[f * 4 + 2] = F.face_id_a;
      tets[f * 4 + 3] = F.face_id_b;
    }
    if (geom) {
      for (int a = 0; a < 3; ++a) geom[f * 7 + a] = F.normal[a];
      geom[f * 7 + 3] = F.area;
      for (int a = 0; a < 3; ++a) geom[f * 7 + 4 + a] = F.center[a];
    }
  }
}
void orc_mesh_destroy(orc_mesh *m) { delete m; }

// Greedy colouring by definition: each tet takes the smallest colour not used
// by an earlier tet that shares one of its vertices.
int orc_colour_tets(const int32_t *tets, int64_t nt, int64_t nv, int32_t *colour) {
  std::vector<std::vector<int>> vt(nv);
  for (int64_t t = 0; t < nt; ++t)
    for (int a = 0; a < 4; ++a) vt[tets[t * 4 + a]].push_back((int)t);
  int maxc = -1;
  for (int64_t t = 0; t < nt; ++t) {
    std::vector<char> used(512, 0);
    for (int a = 0; a < 4; ++a)
      for (int o : vt[tets[t * 4 + a]])
        if (o < t) used[colour[o]] = 1;
    int c = 0;
    while (used[c]) ++c;
    colour[t] = c;
    maxc = std::max(maxc, c);
  }
  return maxc + 1;
}

int orc_ns_create(const orc_mesh *m, double density, double viscosity, const double inlet0[3],
                  orc_ns **out) {
  return guard([&]() -> int {
    auto *h = new orc_ns();
    h->mesh = m;
    h->rho = density;
    h->mu = viscosity;
    const double z[3] = {0, 0, 0};
    h->initialize(inlet0 ? inlet0 : z);
    *out = h;
    return 0;
  });
}
void orc_ns_destroy(orc_ns *h) { delete h; }
void orc_ns_set_dt(orc_ns *h, double dt) { h->dt = dt; }
void orc_ns_set_mean_inlet(orc_ns *h, const double v[3]) {
  h->g.assign(h->nu, V3{v[0], v[1], v[2]});
}
void orc_ns_set_inlet_profile(orc_ns *h, const double *g) {
  for (int i = 0; i < h->nu; ++i) h->g[i] = V3{g[i * 3], g[i * 3 + 1], g[i * 3 + 2]};
}
void orc_ns_set_outlet(orc_ns *h, int type, double pressure) {
  h->outlet_type = type;
  h->outlet_pressure = pressure;
}
void orc_ns_set_corrections(orc_ns *h, int n) { h->n_corrections = n; }
void orc_ns_set_relax(orc_ns *h, double ru, double rp) {
  h->relax_u = ru;
  h->relax_p = rp;
}
void orc_ns_set_reference_node(orc_ns *h, int node) { h->reference_node = node; }
void orc_ns_set_solver(orc_ns *h, double mt, double pt, int maxit, int ps) {
  h->momentum_tol = mt;
  h->pressure_tol = pt;
  h->max_iter = maxit;
  h->pressure_solver = ps;
}
void orc_ns_set_stop_scaled(orc_ns *h, int on) { h->stop_scaled = on != 0; }
void orc_ns_set_momentum_solver(orc_ns *h, int method) { h->momentum_solver = method; }
// which: 0 pressure, 1 shear, 2 total (nv x 3)
void orc_ns_forces(orc_ns *h, int which, double *out) {
  std::vector<V3> p = h->computePressureForces(), s = h->computeShearStress();
  for (int i = 0; i < h->nv; ++i) {
    V3 v = which == 0 ? p[i] : (which == 1 ? s[i] : add(p[i], s[i]));
    for (int c = 0; c < 3; ++c) out[i * 3 + c] = v[c];
  }
}
int orc_adjoint_gradient(orc_ns *h, int nsteps, int nk, const double *basis, const double *coef,
                         double *J, double *dJdc, double *dJdmu) {
  return guard([&]() -> int {
    return h->adjointGradient(nsteps, nk, basis, coef, J, dJdc, dJdmu);
  });
}
int orc_ns_piso_step(orc_ns *h) { return h->pisoStep(); }
int orc_ns_simple_step(orc_ns *h) { return h->simpleStep(); }
void orc_ns_sizes(const orc_ns *h, int64_t out[6]) {
  out[0] = h->nv;
  out[1] = h->nt;
  out[2] = h->ne;
  out[3] = h->nu;
  out[4] = (int64_t)h->M.val.size();
  out[5] = (int64_t)h->L.val.size();
}
void orc_ns_get_velocity(const orc_ns *h, double *out) {
  for (int i = 0; i < h->nu; ++i)
    for (int c = 0; c < 3; ++c) out[i * 3 + c] = h->velocity[i][c];
}
void orc_ns_set_velocity(orc_ns *h, const double *in) {
  for (int i = 0; i < h->nu; ++i) h->velocity[i] = V3{in[i * 3], in[i * 3 + 1], in[i * 3 + 2]};
}
void orc_ns_get_velocity_star(const orc_ns *h, double *out) {
  for (int i = 0; i < h->nu; ++i)
    for (int c = 0; c < 3; ++c) out[i * 3 + c] = h->velocity_star[i][c];
}
void orc_ns_get_pressure(const orc_ns *h, double *out) {
  std::copy(h->pressure.begin(), h->pressure.end(), out);
}
void orc_ns_set_pressure(orc_ns *h, const double *in) {
  h->pressure.assign(in, in + h->nv);
}
void orc_ns_get_pressure_correction(const orc_ns *h, double *out) {
  std::copy(h->pressure_correction.begin(), h->pressure_correction.end(), out);
}
void orc_ns_get_inv_lumped_mass(const orc_ns *h, double *out) {
  std::copy(h->inv_lumped_mass.begin(), h->inv_lumped_mass.end(), out);
}
void orc_ns_get_matrix(orc_ns *h, int which, int32_t *rp, int32_t *col, double *val) {
  switch (which) {
    case 0: csr_out(h->M, rp, col, val); break;
    case 1: csr_out(h->K, rp, col, val); break;
    case 2: csr_out(h->lhs_matrix(), rp, col, val); break;
    case 3: csr_out(h->L, rp, col, val); break;
    case 4: csr_out(h->lbc_matrix(), rp, col, val); break;
    case 5: csr_out(h->Gx, rp, col, val); break;
    case 6: csr_out(h->Gy, rp, col, val); break;
    case 7: csr_out(h->Gz, rp, col, val); break;
    default: break;
  }
}
void orc_ns_compute_advection(orc_ns *h, double *out) {
  std::vector<V3> c = h->computeAdvectionRHS();
  for (int i = 0; i < h->nu; ++i)
    for (int k = 0; k < 3; ++k) out[i * 3 + k] = c[i][k];
}
void orc_ns_compute_divergence(orc_ns *h, const double *u, double *out) {
  std::vector<V3> uu(h->nu);
  for (int i = 0; i < h->nu; ++i) uu[i] = V3{u[i * 3], u[i * 3 + 1], u[i * 3 + 2]};
  std::vector<double> d = h->computeDivergence(uu);
  std::copy(d.begin(), d.end(), out);
}
void orc_ns_get_stats(const orc_ns *h, int32_t *iters, double *errs) {
  for (int c = 0; c < 3; ++c) {
    iters[c] = h->it_mom[c];
    errs[c] = h->err_mom[c];
  }
  for (int k = 0; k < 8; ++k) {
    iters[3 + k] = h->it_p[k];
    errs[3 + k] = h->err_p[k];
  }
}
// Eigen-form BiCGSTAB (method 0) / Jacobi CG (1) on a caller-provided CSR.
int orc_csr_krylov(int64_t n, const int32_t *rp, const int32_t *col, const double *val,
                   int method, const double *b, double *x, double tol, int maxit,
                   int32_t *iters, double *err) {
  Csr A;
  A.n = (int)n;
  A.rp.assign(rp, rp + n + 1);
  A.col.assign(col, col + rp[n]);
  A.val.assign(val, val + rp[n]);
  std::vector<double> rhs(b, b + n), sol;
  KResult r = method == 1 ? pcg(A, rhs, sol, tol, maxit) : bicgstab(A, rhs, sol, tol, maxit);
  std::copy(sol.begin(), sol.end(), x);
  if (iters) *iters = r.iters;
  if (err) *err = r.err;
  g_last_restarts = r.restarts;
  return r.ok ? 0 : 1;
}

void orc_set_restart_scale(double s) { g_restart_scale = s; }
int orc_last_restarts(void) { return g_last_restarts; }

int orc_ns_krylov_solve(orc_ns *h, int which, int method, const double *b, double *x,
                        double tol, int maxit, int32_t *iters, double *err) {
  const Csr A = which == 2 ? h->lhs_matrix() : h->lbc_matrix();
  std::vector<double> rhs(b, b + A.n), sol;
  KResult r = method == 1 ? pcg(A, rhs, sol, tol, maxit) : bicgstab(A, rhs, sol, tol, maxit);
  std::copy(sol.begin(), sol.end(), x);
  if (iters) *iters = r.iters;
  if (err) *err = r.err;
  g_last_restarts = r.restarts;
  return r.ok ? 0 : 1;
}

}  // extern "C"
