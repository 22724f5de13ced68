// Host mesh layer.  See mesh.hpp for the design; reference behaviour cited per
// function.  Vector arithmetic is spelled out component by component in the
// operation order Eigen uses for Vector3d (cross: Geometry/OrthoMethods; dot
// and squaredNorm: left-to-right sum of three products).
#include "mesh.hpp"

#include "parallel.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace cfsa {
namespace {

struct V3 {
  double x, y, z;
};
inline V3 load3(const double *p) { return {p[0], p[1], p[2]}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 l, V3 r) {
  return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x};
}
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 scale(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 normalized(V3 a) {
  double z = dot(a, a);
  if (z > 0.0) {
    double s = std::sqrt(z);
    return {a.x / s, a.y / s, a.z / s};
  }
  return a;
}

// Local faces of a tet, reference Mesh.cpp:498-503.
constexpr int kLocalFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
// Local edges of a tet, reference Mesh.cpp:653-654.
constexpr int kLocalEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

}  // namespace

// ---------------------------------------------------------------------------
// Construction
// ---------------------------------------------------------------------------

// reference: Mesh::generateStructuredRectangularPrism (Mesh.cpp:146-225)
Mesh Mesh::structuredRectangularPrism(double length, double width,
                                      double height, int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1)
    throw std::invalid_argument("nx, ny, nz must be >= 1");
  Mesh m;
  const int64_t nxp = nx + 1, nyp = ny + 1, nzp = nz + 1;
  const double dx = length / static_cast<double>(nx);
  const double dy = width / static_cast<double>(ny);
  const double dz = height / static_cast<double>(nz);
  m.vertices.resize((size_t)(nxp * nyp * nzp) * 3);
  size_t w = 0;
  for (int64_t k = 0; k < nzp; ++k)
    for (int64_t j = 0; j < nyp; ++j)
      for (int64_t i = 0; i < nxp; ++i) {
        m.vertices[w++] = static_cast<double>(i) * dx;
        m.vertices[w++] = static_cast<double>(j) * dy;
        m.vertices[w++] = static_cast<double>(k) * dz;
      }
  auto idx = [nxp, nyp](int64_t i, int64_t j, int64_t k) {
    return (int32_t)(i + nxp * (j + nyp * k));
  };
  m.tets.resize((size_t)nx * ny * nz * 24);
  w = 0;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int32_t v0 = idx(i, j, k), v1 = idx(i + 1, j, k),
                      v2 = idx(i, j + 1, k), v3 = idx(i + 1, j + 1, k),
                      v4 = idx(i, j, k + 1), v5 = idx(i + 1, j, k + 1),
                      v6 = idx(i, j + 1, k + 1), v7 = idx(i + 1, j + 1, k + 1);
        // Kuhn split around the v0-v7 body diagonal, same tet order.
        const int32_t t[6][4] = {{v0, v1, v3, v7}, {v0, v3, v2, v7},
                                 {v0, v2, v6, v7}, {v0, v6, v4, v7},
                                 {v0, v4, v5, v7}, {v0, v5, v1, v7}};
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < 4; ++b) m.tets[w++] = t[a][b];
      }
  m.finalize();
  return m;
}

Mesh Mesh::fromArrays(const double *coords, int64_t nv, const int32_t *tets,
                      int64_t nt) {
  if (nv <= 0 || nt <= 0) throw std::invalid_argument("empty mesh");
  Mesh m;
  m.vertices.assign(coords, coords + nv * 3);
  m.tets.assign(tets, tets + nt * 4);
  for (int64_t i = 0; i < nt * 4; ++i)
    if (tets[i] < 0 || tets[i] >= nv)
      throw std::invalid_argument("tet references a vertex out of range");
  m.finalize();
  if (m.hasDegenerateTet())
    throw std::runtime_error("[Mesh] degenerate tetrahedral elements found.");
  return m;
}

// reference: Mesh::generateFromMshFile (Mesh.cpp:16-144), Gmsh MSH 4.1 ASCII:
// node blocks list all tags then all coordinates; element type 4 = tet4.
Mesh Mesh::fromMshFile(const std::string &path) {
  std::ifstream ifs(path);
  if (!ifs.is_open())
    throw std::runtime_error("[Mesh] Could not open msh file: " + path);
  std::vector<double> verts;
  std::vector<int32_t> tets;
  std::map<long, int32_t> tag_to_index;
  std::string line;
  while (std::getline(ifs, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty()) continue;
    if (line == "$Nodes") {
      if (!std::getline(ifs, line)) break;
      std::istringstream hs(line);
      long n_blocks = 0, n_nodes = 0, min_tag = 0, max_tag = 0;
      hs >> n_blocks >> n_nodes >> min_tag >> max_tag;
      verts.reserve((size_t)std::max<long>(n_nodes, 0) * 3);
      for (long b = 0; b < n_blocks; ++b) {
        if (!std::getline(ifs, line)) break;
        std::istringstream bs(line);
        long dim = 0, tag = 0, param = 0, n_in = 0;
        bs >> dim >> tag >> param >> n_in;
        std::vector<long> tags((size_t)std::max<long>(n_in, 0));
        for (long i = 0; i < n_in; ++i) {
          if (!std::getline(ifs, line))
            throw std::runtime_error("Unexpected EOF in Node Tags");
          tags[(size_t)i] = std::stol(line);
        }
        for (long i = 0; i < n_in; ++i) {
          if (!std::getline(ifs, line))
            throw std::runtime_error("Unexpected EOF in Node Coords");
          std::istringstream cs(line);
          double x = 0, y = 0, z = 0;
          cs >> x >> y >> z;
          tag_to_index[tags[(size_t)i]] = (int32_t)(verts.size() / 3);
          verts.push_back(x);
          verts.push_back(y);
          verts.push_back(z);
        }
      }
    } else if (line == "$Elements") {
      if (!std::getline(ifs, line)) break;
      std::istringstream hs(line);
      long n_blocks = 0, n_el = 0, min_tag = 0, max_tag = 0;
      hs >> n_blocks >> n_el >> min_tag >> max_tag;
      for (long b = 0; b < n_blocks; ++b) {
        if (!std::getline(ifs, line)) break;
        std::istringstream bs(line);
        long dim = 0, tag = 0, etype = 0, n_in = 0;
        bs >> dim >> tag >> etype >> n_in;
        for (long i = 0; i < n_in; ++i) {
          if (!std::getline(ifs, line)) break;
          if (etype != 4) continue;
          std::istringstream es(line);
          long et = 0, n[4] = {0, 0, 0, 0};
          es >> et >> n[0] >> n[1] >> n[2] >> n[3];
          for (int c = 0; c < 4; ++c) {
            auto it = tag_to_index.find(n[c]);
            if (it == tag_to_index.end())
              throw std::runtime_error(
                  "[Mesh] Element references unknown node id: " +
                  std::to_string(n[0]));
            tets.push_back(it->second);
          }
        }
      }
    }
  }
  if (verts.empty() || tets.empty())
    throw std::runtime_error("[Mesh] Failed to read vertices or tets from msh.");
  return fromArrays(verts.data(), (int64_t)verts.size() / 3, tets.data(),
                    (int64_t)tets.size() / 4);
}

Mesh Mesh::cylinder(double length, double radius, int divisions,
                    double max_volume) {
  if (!(length > 0) || !(radius > 0) || !(max_volume > 0))
    throw std::invalid_argument("cylinder: length, radius, max_volume > 0");
  // Ring count from the TetGen max-volume switch (edge ~ 0.42 * a^(1/3)) and
  // the reference's circumferential division count (outer ring has 6m points).
  const double h = 0.42 * std::cbrt(max_volume);
  int m = (int)std::lround(radius / h);
  m = std::max(m, (divisions + 5) / 6);
  m = std::max(m, 1);
  int layers = (int)std::lround(length * m / radius);
  layers = std::max(layers, 1);

  const int64_t P = 1 + 3LL * m * (m + 1);  // points per cross-section
  std::vector<double> disc((size_t)P * 2);
  auto ring_start = [](int k) -> int64_t {
    return k == 0 ? 0 : 1 + 3LL * (k - 1) * k;
  };
  disc[0] = 0.0;
  disc[1] = 0.0;
  for (int k = 1; k <= m; ++k) {
    const double r = radius * (double)k / (double)m;
    const int nk = 6 * k;
    for (int j = 0; j < nk; ++j) {
      const double th = 2.0 * M_PI * (double)j / (double)nk;
      const int64_t p = ring_start(k) + j;
      disc[(size_t)p * 2 + 0] = r * std::cos(th);
      disc[(size_t)p * 2 + 1] = r * std::sin(th);
    }
  }
  std::vector<std::array<int64_t, 3>> tris;
  tris.reserve((size_t)6 * m * m);
  for (int k = 1; k <= m; ++k) {
    const int nk = 6 * k, nk1 = 6 * (k - 1);
    for (int s = 0; s < 6; ++s) {
      auto outer = [&](int i) { return ring_start(k) + (s * k + i) % nk; };
      auto inner = [&](int i) -> int64_t {
        return k == 1 ? 0 : ring_start(k - 1) + (s * (k - 1) + i) % nk1;
      };
      for (int i = 0; i < k; ++i) tris.push_back({outer(i), outer(i + 1), inner(i)});
      for (int i = 0; i + 1 < k; ++i)
        tris.push_back({inner(i), outer(i + 1), inner(i + 1)});
    }
  }
  Mesh me;
  me.vertices.resize((size_t)P * (layers + 1) * 3);
  for (int l = 0; l <= layers; ++l) {
    const double x = length * (double)l / (double)layers;
    for (int64_t p = 0; p < P; ++p) {
      const size_t o = ((size_t)l * P + p) * 3;
      me.vertices[o + 0] = x;
      me.vertices[o + 1] = disc[(size_t)p * 2 + 0];
      me.vertices[o + 2] = disc[(size_t)p * 2 + 1];
    }
  }
  // Prism -> 3 tets: rotate/reflect so the smallest global index is vertex 0,
  // then cut the opposite quad along the diagonal through its smallest vertex.
  static const int perm[6][6] = {{0, 1, 2, 3, 4, 5}, {1, 2, 0, 4, 5, 3},
                                 {2, 0, 1, 5, 3, 4}, {3, 5, 4, 0, 2, 1},
                                 {4, 3, 5, 1, 0, 2}, {5, 4, 3, 2, 1, 0}};
  me.tets.reserve((size_t)tris.size() * layers * 12);
  for (int l = 0; l < layers; ++l) {
    for (const auto &t : tris) {
      int64_t V[6] = {l * P + t[0],       l * P + t[1],       l * P + t[2],
                      (l + 1) * P + t[0], (l + 1) * P + t[1], (l + 1) * P + t[2]};
      int imin = 0;
      for (int q = 1; q < 6; ++q)
        if (V[q] < V[imin]) imin = q;
      int64_t W[6];
      for (int q = 0; q < 6; ++q) W[q] = V[perm[imin][q]];
      int64_t tt[3][4];
      if (std::min(W[1], W[5]) < std::min(W[2], W[4])) {
        int64_t a[3][4] = {{W[0], W[1], W[2], W[5]},
                           {W[0], W[1], W[5], W[4]},
                           {W[0], W[4], W[5], W[3]}};
        std::memcpy(tt, a, sizeof(tt));
      } else {
        int64_t a[3][4] = {{W[0], W[1], W[2], W[4]},
                           {W[0], W[4], W[2], W[5]},
                           {W[0], W[4], W[5], W[3]}};
        std::memcpy(tt, a, sizeof(tt));
      }
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 4; ++b) me.tets.push_back((int32_t)tt[a][b]);
    }
  }
  me.finalize();
  if (me.hasDegenerateTet())
    throw std::runtime_error("[Mesh] cylinder mesher produced a degenerate tet");
  return me;
}

void Mesh::finalize() {
  // Same call order as the reference constructors (Mesh.cpp:216-224).
  buildConnectivity();
  computeGeometry();
  ensureConsistentFaceNormals();
  computeShapeFunctionGradients();
  buildP2EdgeNodes();
  computeMinEdgeLength();
  p1_bc.assign((size_t)nVertices(), BC_UNDEFINED);
}

// ---------------------------------------------------------------------------
// Connectivity: reference Mesh::buildConnectivity (Mesh.cpp:494-588).  Faces
// are ordered by their sorted vertex triple (the std::map order of the
// reference); the owner is the first (tet, local face) in tet order, the
// neighbour the last other one.  Implemented as one sort of 4*nt records.
// ---------------------------------------------------------------------------
void Mesh::buildConnectivity() {
  const int64_t nt = nTets();
  struct Rec {
    int32_t k[3];
    int32_t tf;  // tet * 4 + local face
  };
  std::vector<Rec> recs((size_t)nt * 4);
  parallel_chunks(nt, [&](int, int64_t t0, int64_t t1) {
    for (int64_t t = t0; t < t1; ++t) {
      const int32_t *v = &tets[(size_t)t * 4];
      for (int f = 0; f < 4; ++f) {
        int32_t a = v[kLocalFace[f][0]], b = v[kLocalFace[f][1]],
                c = v[kLocalFace[f][2]];
        if (a > b) std::swap(a, b);
        if (b > c) std::swap(b, c);
        if (a > b) std::swap(a, b);
        recs[(size_t)t * 4 + f] = {{a, b, c}, (int32_t)(t * 4 + f)};
      }
    }
  });
  parallel_sort(recs, [](const Rec &x, const Rec &y) {
    if (x.k[0] != y.k[0]) return x.k[0] < y.k[0];
    if (x.k[1] != y.k[1]) return x.k[1] < y.k[1];
    if (x.k[2] != y.k[2]) return x.k[2] < y.k[2];
    return x.tf < y.tf;
  });
  faces.clear();
  {
    size_t nf = recs.empty() ? 0 : 1;
    for (size_t q = 1; q < recs.size(); ++q)
      nf += recs[q].k[0] != recs[q - 1].k[0] || recs[q].k[1] != recs[q - 1].k[1] ||
            recs[q].k[2] != recs[q - 1].k[2];
    faces.reserve(nf);
  }
  tet_faceids.assign((size_t)nt * 4, -1);
  tet_neighbors.assign((size_t)nt * 4, -1);
  size_t i = 0;
  while (i < recs.size()) {
    size_t j = i + 1;
    while (j < recs.size() && recs[j].k[0] == recs[i].k[0] &&
           recs[j].k[1] == recs[i].k[1] && recs[j].k[2] == recs[i].k[2])
      ++j;
    Face f;
    const int32_t ta = recs[i].tf / 4, fa = recs[i].tf % 4;
    const int32_t *v = &tets[(size_t)ta * 4];
    f.vertids = {v[kLocalFace[fa][0]], v[kLocalFace[fa][1]], v[kLocalFace[fa][2]]};
    f.tet_a = ta;
    f.face_id_a = fa;
    for (size_t q = i + 1; q < j; ++q) {
      f.tet_b = recs[q].tf / 4;
      f.face_id_b = recs[q].tf % 4;
    }
    const int32_t fid = (int32_t)faces.size();
    for (size_t q = i; q < j; ++q) tet_faceids[(size_t)recs[q].tf] = fid;
    faces.push_back(f);
    i = j;
  }
  for (const Face &f : faces) {
    if (f.tet_a >= 0 && f.tet_b >= 0 && f.face_id_a >= 0 && f.face_id_a < 4 &&
        f.face_id_b >= 0 && f.face_id_b < 4) {
      tet_neighbors[(size_t)f.tet_a * 4 + f.face_id_a] = f.tet_b;
      tet_neighbors[(size_t)f.tet_b * 4 + f.face_id_b] = f.tet_a;
    }
  }
}

// reference: Mesh::computeGeometry (Mesh.cpp:448-471)
void Mesh::computeGeometry() {
  const int64_t nt = nTets();
  volumes.resize((size_t)nt);
  centroids.resize((size_t)nt * 3);
  parallel_chunks(nt, [&](int, int64_t t0, int64_t t1) {
    for (int64_t t = t0; t < t1; ++t) {
      const int32_t *v = &tets[(size_t)t * 4];
      V3 a = load3(&vertices[(size_t)v[0] * 3]), b = load3(&vertices[(size_t)v[1] * 3]),
         c = load3(&vertices[(size_t)v[2] * 3]), d = load3(&vertices[(size_t)v[3] * 3]);
      centroids[(size_t)t * 3 + 0] = (((a.x + b.x) + c.x) + d.x) / 4.0;
      centroids[(size_t)t * 3 + 1] = (((a.y + b.y) + c.y) + d.y) / 4.0;
      centroids[(size_t)t * 3 + 2] = (((a.z + b.z) + c.z) + d.z) / 4.0;
      volumes[(size_t)t] = std::fabs(dot(cross(sub(b, a), sub(c, a)), sub(d, a)) / 6.0);
    }
  });
  parallel_chunks((int64_t)faces.size(), [&](int, int64_t f0, int64_t f1) {
    for (int64_t k = f0; k < f1; ++k) {
      Face &f = faces[(size_t)k];
      V3 a = load3(&vertices[(size_t)f.vertids[0] * 3]),
         b = load3(&vertices[(size_t)f.vertids[1] * 3]),
         c = load3(&vertices[(size_t)f.vertids[2] * 3]);
      V3 ab = sub(b, a), ac = sub(c, a);
      V3 n = normalized(cross(ac, ab));
      f.normal = {n.x, n.y, n.z};
      V3 cr = cross(ab, ac);
      f.area = 0.5 * std::sqrt(dot(cr, cr));
      f.center = {((a.x + b.x) + c.x) / 3.0, ((a.y + b.y) + c.y) / 3.0,
                  ((a.z + b.z) + c.z) / 3.0};
    }
  });
  computeShapeFunctionGradients();
}

// reference: Mesh::ensureConsistentFaceNormals (Mesh.cpp:615-639).  Only
// internal faces are visited (the outward check is nested under the
// internal-face branch in the reference).
void Mesh::ensureConsistentFaceNormals() {
  const double eps = 1e-12;
  for (Face &f : faces) {
    if (f.tet_a >= 0 && f.tet_b >= 0) {
      V3 ca = load3(&centroids[(size_t)f.tet_a * 3]);
      V3 cb = load3(&centroids[(size_t)f.tet_b * 3]);
      V3 n = {f.normal[0], f.normal[1], f.normal[2]};
      V3 desired = normalized(sub(cb, ca));
      if (dot(n, desired) <= eps) {
        f.normal = {-n.x, -n.y, -n.z};
        std::swap(f.vertids[1], f.vertids[2]);
      } else {
        V3 c = {f.center[0], f.center[1], f.center[2]};
        V3 out = normalized(sub(c, ca));
        if (dot(n, out) <= eps) {
          f.normal = {-n.x, -n.y, -n.z};
          std::swap(f.vertids[1], f.vertids[2]);
        }
      }
    }
  }
}

// reference: Mesh::computeShapeFunctionGradients (Mesh.cpp:590-614); signed
// 6V so the gradients are correct for either tet orientation.
void Mesh::computeShapeFunctionGradients() {
  gradients.resize((size_t)nTets() * 12);
  std::vector<double> vtmp((size_t)nTets());
  tetGeometry(vertices.data(), tets.data(), nTets(), vtmp.data(), gradients.data());
}

void Mesh::tetGeometry(const double *coords, const int32_t *tets, int64_t nt,
                       double *vol, double *grad) {
  parallel_chunks(nt, [&](int, int64_t t0, int64_t t1) {
    for (int64_t t = t0; t < t1; ++t) {
      const int32_t *v = &tets[(size_t)t * 4];
      V3 p0 = load3(&coords[(size_t)v[0] * 3]), p1 = load3(&coords[(size_t)v[1] * 3]),
         p2 = load3(&coords[(size_t)v[2] * 3]), p3 = load3(&coords[(size_t)v[3] * 3]);
      // volume: reference Mesh.cpp:455
      vol[t] = std::fabs(dot(cross(sub(p1, p0), sub(p2, p0)), sub(p3, p0)) / 6.0);
      V3 e1 = sub(p1, p0), e2 = sub(p2, p0), e3 = sub(p3, p0);
      const double sixv = dot(e1, cross(e2, e3));
      const double inv = 1.0 / sixv;
      V3 g[4] = {scale(inv, cross(sub(p3, p1), sub(p2, p1))),
                 scale(inv, cross(sub(p2, p0), sub(p3, p0))),
                 scale(inv, cross(sub(p3, p0), sub(p1, p0))),
                 scale(inv, cross(sub(p1, p0), sub(p2, p0)))};
      double *o = &grad[(size_t)t * 12];
      for (int i = 0; i < 4; ++i) {
        o[i * 3 + 0] = g[i].x;
        o[i * 3 + 1] = g[i].y;
        o[i * 3 + 2] = g[i].z;
      }
    }
  });
}

// reference: Mesh::buildP2EdgeNodes (Mesh.cpp:641-677).  Edge nodes are
// numbered by first appearance scanning tets in order and local edges in
// (01,02,03,12,13,23) order.  One sort of 6*nt records finds each edge's first
// appearance; a second orders the unique edges by it.
void Mesh::buildP2EdgeNodes() {
  const int64_t nt = nTets();
  struct Rec {
    int32_t lo, hi;
    int64_t pos;  // tet * 6 + local edge
  };
  std::vector<Rec> recs((size_t)nt * 6);
  parallel_chunks(nt, [&](int, int64_t t0, int64_t t1) {
    for (int64_t t = t0; t < t1; ++t) {
      const int32_t *v = &tets[(size_t)t * 4];
      for (int e = 0; e < 6; ++e) {
        int32_t a = v[kLocalEdge[e][0]], b = v[kLocalEdge[e][1]];
        if (a > b) std::swap(a, b);
        recs[(size_t)t * 6 + e] = {a, b, t * 6 + e};
      }
    }
  });
  parallel_sort(recs, [](const Rec &x, const Rec &y) {
    if (x.lo != y.lo) return x.lo < y.lo;
    if (x.hi != y.hi) return x.hi < y.hi;
    return x.pos < y.pos;
  });
  // unique edges with their first-appearance position
  std::vector<std::pair<int64_t, size_t>> first;  // (first pos, start in recs)
  for (size_t i = 0; i < recs.size();) {
    size_t j = i + 1;
    while (j < recs.size() && recs[j].lo == recs[i].lo && recs[j].hi == recs[i].hi) ++j;
    first.push_back({recs[i].pos, i});
    i = j;
  }
  std::vector<size_t> order(first.size());
  std::iota(order.begin(), order.end(), 0);
  parallel_sort(order, [&](size_t a, size_t b) { return first[a].first < first[b].first; });
  const size_t ne = first.size();
  edge_nodes.resize(ne * 3);
  edge_endpoints.resize(ne * 2);
  tet_edge_nodes.assign((size_t)nt * 6, -1);
  // every edge writes its own entries (and the tet slots of its own records)
  parallel_chunks((int64_t)ne, [&](int, int64_t i0, int64_t i1) {
    for (size_t id = (size_t)i0; id < (size_t)i1; ++id) {
      const size_t start = first[order[id]].second;
      const int32_t lo = recs[start].lo, hi = recs[start].hi;
      edge_endpoints[id * 2 + 0] = lo;
      edge_endpoints[id * 2 + 1] = hi;
      for (int c = 0; c < 3; ++c)
        edge_nodes[id * 3 + c] =
            0.5 * (vertices[(size_t)lo * 3 + c] + vertices[(size_t)hi * 3 + c]);
      for (size_t q = start; q < recs.size() && recs[q].lo == lo && recs[q].hi == hi; ++q)
        tet_edge_nodes[(size_t)recs[q].pos] = (int32_t)id;
    }
  });
  p2_bc.assign(ne, BC_UNDEFINED);
}

// reference: Mesh::computeMinEdgeLength (Mesh.cpp:419-446)
void Mesh::computeMinEdgeLength() {
  const int64_t nt = nTets();
  // per-thread minima; the minimum does not depend on the split
  std::vector<double> part((size_t)host_threads(), std::numeric_limits<double>::max());
  parallel_chunks(nt, [&](int th, int64_t t0, int64_t t1) {
    double m = std::numeric_limits<double>::max();
    for (int64_t t = t0; t < t1; ++t) {
      const int32_t *v = &tets[(size_t)t * 4];
      for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) {
          V3 d = sub(load3(&vertices[(size_t)v[i] * 3]), load3(&vertices[(size_t)v[j] * 3]));
          m = std::min(m, dot(d, d));
        }
    }
    part[(size_t)th] = m;
  });
  const double mn = *std::min_element(part.begin(), part.end());
  min_edge_length = (mn == std::numeric_limits<double>::max()) ? 0.0 : std::sqrt(mn);
}

bool Mesh::hasDegenerateTet() const {
  for (double v : volumes)
    if (v <= 0.0) return true;
  return false;
}

// reference: Mesh::setupBoundaryConditions (Mesh.cpp:325-417).  Every vertex
// starts Internal; boundary faces are visited in face order, each setting its
// vertices to Wall and then to Inlet/Outlet when the face centre lies within
// the tolerance of the inlet/outlet plane.  The last face to touch a vertex
// decides its label.
void Mesh::setupBoundaryConditions(const double dir[3], double percent_coverage,
                                   Mesh &mesh) {
  if (mesh.nFaces() == 0 || mesh.nVertices() < 3) return;
  const int64_t nv = mesh.nVertices();
  double mn[3] = {mesh.vertices[0], mesh.vertices[1], mesh.vertices[2]};
  double mx[3] = {mn[0], mn[1], mn[2]};
  for (int64_t v = 0; v < nv; ++v)
    for (int c = 0; c < 3; ++c) {
      mn[c] = std::min(mn[c], mesh.vertices[(size_t)v * 3 + c]);
      mx[c] = std::max(mx[c], mesh.vertices[(size_t)v * 3 + c]);
    }
  const double size[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
  int axis = 0;
  if (size[1] > size[0] && size[1] > size[2])
    axis = 1;
  else if (size[2] > size[0] && size[2] > size[1])
    axis = 2;
  const double tol = size[axis] * 0.01 * percent_coverage;
  const double inlet_pos = dir[axis] > 0 ? mn[axis] : mx[axis];
  const double outlet_pos = dir[axis] > 0 ? mx[axis] : mn[axis];
  mesh.p1_bc.assign((size_t)nv, BC_INTERNAL);
  for (Face &f : mesh.faces) {
    if (f.tet_b != -1) {
      f.is_ecm = false;
      continue;
    }
    f.is_ecm = true;
    for (int32_t v : f.vertids) mesh.p1_bc[(size_t)v] = BC_WALL;
    const double di = std::fabs(f.center[axis] - inlet_pos);
    const double dout = std::fabs(f.center[axis] - outlet_pos);
    if (di < tol) {
      f.is_ecm = false;
      for (int32_t v : f.vertids) mesh.p1_bc[(size_t)v] = BC_INLET;
    } else if (dout < tol) {
      f.is_ecm = false;
      for (int32_t v : f.vertids) mesh.p1_bc[(size_t)v] = BC_OUTLET;
    }
  }
  mesh.setP2BoundariesFromP1Boundaries();
}

void Mesh::setP2BoundariesFromP1Boundaries() {
  const int64_t ne = nEdgeNodes();
  p2_bc.resize((size_t)ne);
  for (int64_t e = 0; e < ne; ++e) {
    const uint8_t a = p1_bc[(size_t)edge_endpoints[(size_t)e * 2]];
    const uint8_t b = p1_bc[(size_t)edge_endpoints[(size_t)e * 2 + 1]];
    p2_bc[(size_t)e] = (a == b && a != BC_INTERNAL) ? a : (uint8_t)BC_INTERNAL;
  }
}

}  // namespace cfsa
