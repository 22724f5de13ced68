// Fluid boundary forces on the GPU, reference NavierStokesSolver::
// computePressureForces / computeShearStress / computeTotalFluidForces
// (src/FEM/NavierStokes.cpp:841-965), getMeanFacePressure /
// getPressureForceAtFace (:357-372).
//
// A per-face kernel evaluates each boundary face's pressure and viscous force
// with the reference's expression order; a per-vertex gather adds one third of
// every adjacent face's force in ascending face order, the order of the
// reference's scatter loop, so the results match it bit for bit.
#include <algorithm>
#include <memory>

#include "mesh.hpp"
#include "solver.hpp"

namespace cfsa {

namespace {

__global__ __launch_bounds__(kBlock) void k_face_forces(
    int nbf, const int *__restrict__ fv, const int *__restrict__ ftet,
    const double *__restrict__ fgeom, const int *__restrict__ tets,
    const double *__restrict__ grad, const double *__restrict__ u,
    const double *__restrict__ p, double mu, double *__restrict__ fp,
    double *__restrict__ fs) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= nbf) return;
  const double n[3] = {fgeom[(size_t)f * 4], fgeom[(size_t)f * 4 + 1], fgeom[(size_t)f * 4 + 2]};
  const double area = fgeom[(size_t)f * 4 + 3];
  // pressure: -(mean face pressure) * normal * area
  double pm = 0.0;
  for (int a = 0; a < 3; ++a) pm += p[fv[(size_t)f * 3 + a]];
  pm /= 3.0;
  for (int c = 0; c < 3; ++c) fp[(size_t)f * 3 + c] = (-pm * n[c]) * area;
  // viscous: tau = 2 mu sym(grad u) of the owner tet, traction tau . n
  const int t = ftet[f];
  double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < 4; ++i) {
    const size_t v = (size_t)tets[(size_t)t * 4 + i] * 3;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) G[a][b] = G[a][b] + u[v + a] * grad[(size_t)t * 12 + i * 3 + b];
  }
  const double tm = 2.0 * mu;
  for (int a = 0; a < 3; ++a) {
    double tau[3];
    for (int b = 0; b < 3; ++b) tau[b] = tm * (0.5 * (G[a][b] + G[b][a]));
    fs[(size_t)f * 3 + a] = ((tau[0] * n[0] + tau[1] * n[1]) + tau[2] * n[2]) * area;
  }
}

// out[v] = sum over v's boundary faces (ascending) of F_f * (1/3); which:
// 0 pressure, 1 shear, 2 total = pressure + shear
__global__ __launch_bounds__(kBlock) void k_vertex_forces(
    int nv, const int *__restrict__ vp, const int *__restrict__ vf,
    const double *__restrict__ fp, const double *__restrict__ fs, int which,
    double *__restrict__ out) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= nv) return;
  const double w = 1.0 / 3.0;
  double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
  for (int q = vp[v]; q < vp[v + 1]; ++q) {
    const size_t f = (size_t)vf[q] * 3;
    for (int c = 0; c < 3; ++c) {
      a[c] = a[c] + fp[f + c] * w;
      b[c] = b[c] + fs[f + c] * w;
    }
  }
  for (int c = 0; c < 3; ++c)
    out[(size_t)v * 3 + c] = which == 0 ? a[c] : (which == 1 ? b[c] : a[c] + b[c]);
}

}  // namespace

// Faces in sorted-triple order with the reference's normals.  Their list and
// vertex order (which fixes each normal's orientation) are set once per mesh:
// taken from the mesh handle, or built with the reference's connectivity and
// flip rule at the first use (Mesh.cpp:494-639).  After vertices move, only
// normal, area and centre are recomputed from the stored vertex order, as the
// reference's Mesh::updateGeometry does (Mesh.hpp:98-101), without flipping.
void NSSolver::ensure_faces() {
  if (faces_ready_) return;
  Mesh &m = face_mesh_;
  m.vertices = h_coords;
  if (!face_topology_) {
    m.tets = h_tets;
    m.buildConnectivity();
    m.computeGeometry();
    m.ensureConsistentFaceNormals();
    face_topology_ = true;
  } else {
    m.computeGeometry();
  }
  all_faces_.clear();
  std::vector<int32_t> bv, bt;
  std::vector<double> bg;
  for (const Face &f : m.faces) {
    all_faces_.push_back(f);
    if (f.tet_b != -1) continue;
    for (int a = 0; a < 3; ++a) bv.push_back(f.vertids[(size_t)a]);
    bt.push_back(f.tet_a);
    for (int a = 0; a < 3; ++a) bg.push_back(f.normal[(size_t)a]);
    bg.push_back(f.area);
  }
  nbf_ = (int)bt.size();
  // vertex -> boundary faces, ascending face order
  std::vector<int32_t> vp((size_t)nv + 1, 0), vf((size_t)nbf_ * 3);
  for (int f = 0; f < nbf_; ++f)
    for (int a = 0; a < 3; ++a) vp[(size_t)bv[(size_t)f * 3 + a] + 1]++;
  for (int v = 0; v < nv; ++v) vp[(size_t)v + 1] += vp[(size_t)v];
  std::vector<int32_t> fill(vp.begin(), vp.end() - 1);
  for (int f = 0; f < nbf_; ++f)
    for (int a = 0; a < 3; ++a) vf[(size_t)fill[(size_t)bv[(size_t)f * 3 + a]]++] = f;
  d_bfv.upload(bv.data(), bv.size());
  d_bft.upload(bt.data(), bt.size());
  d_bfg.upload(bg.data(), bg.size());
  d_vfp.upload(vp.data(), vp.size());
  d_vff.upload(vf.data(), vf.size());
  d_fp.alloc((size_t)std::max(nbf_, 1) * 3);
  d_fs.alloc((size_t)std::max(nbf_, 1) * 3);
  d_forces.alloc((size_t)nv * 3);
  faces_ready_ = true;
}

void NSSolver::compute_forces(int which, double *out) {
  if (!initialized) throw std::runtime_error("[NS] Error: mesh not initialized");
  ensure_faces();
  if (nbf_ > 0)
    hipLaunchKernelGGL(k_face_forces, dim3(div_up(nbf_, kBlock)), dim3(kBlock), 0, stream_,
                       nbf_, d_bfv.p, d_bft.p, d_bfg.p, d_tets.p, d_grad.p, d_u.p, d_p.p,
                       viscosity, d_fp.p, d_fs.p);
  CFSA_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_vertex_forces, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, stream_, nv,
                     d_vfp.p, d_vff.p, d_fp.p, d_fs.p, which, d_forces.p);
  CFSA_LAUNCH_CHECK();
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_forces.download(out, (size_t)nv * 3);
}

// getMeanFacePressure (NavierStokes.cpp:357-365) and getPressureForceAtFace
// (:367-372) over the full sorted face list.
void NSSolver::face_pressure(int64_t fi, double *mean, double force[3]) {
  ensure_faces();
  if (fi < 0 || fi >= (int64_t)all_faces_.size()) throw std::invalid_argument("face out of range");
  const Face &f = all_faces_[(size_t)fi];
  double s = 0.0;
  for (int a = 0; a < 3; ++a) {
    double pv = 0.0;  // only the face's three vertex pressures are read
    get_pressure_at(f.vertids[(size_t)a], &pv);
    s += pv;
  }
  const double avg = s / 3.0;
  if (mean) *mean = avg;
  if (force)
    for (int c = 0; c < 3; ++c) force[c] = (-f.normal[(size_t)c] * f.area) * avg;
}

}  // namespace cfsa
