// C++ host API with the reference's class and method names, over the C ABI of
// libcfsa_ns.so.  A caller written against the reference
//   auto mesh = std::make_shared<Mesh>(Mesh::structuredRectangularPrism(...));
//   Mesh::setupBoundaryConditions(dir, coverage, *mesh);
//   auto ns = std::make_shared<NavierStokesSolver>();
//   ns->initialize(mesh, fluid); ns->setDt(dt); ns->pisoStep();
//   auto forces = ns->computeTotalFluidForces();
//   io::exportToVtk(name, *ns);
// compiles against this header with `using namespace cfsa::api;`.  Differences
// a caller sees: vectors are std::array<double, 3> (Vector3d here) instead of
// Eigen::Vector3d, and getVelocity / getPressure return vectors held in pinned
// host memory (Field3 / Field1) that stay valid until the next step plus one
// (the step after next reuses the buffer).
//
// Reference: include/FEM/NavierStokes.hpp:13-193, include/Mesh/Mesh.hpp:35-220,
// include/Numerics/FEMIntegrator.hpp:10-92, include/IO/VtkExport.hpp:20-21.
#pragma once

#include <array>
#include <cstdint>
#include <iostream>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "cfsa_ns.h"

namespace cfsa {
namespace api {

using Vector3d = std::array<double, 3>;

enum class FluidBCType { Wall, Inlet, Outlet, Internal, Undefined };
enum class OutletType { DirichletPressure, Neumann, Undefined };
enum class InletType { Uniform, Pulsitile, Undefined };
enum class ViscosityModel { Newtonian, Carreau };
enum class TurbulenceModel { Laminar, RANS, LES };

// NavierStokes.hpp:44-51
struct Fluid {
  double density = 1000;
  double viscosity = 0.001;
  TurbulenceModel turbuelence_model = TurbulenceModel::Laminar;
  ViscosityModel viscosity_model = ViscosityModel::Newtonian;
  std::vector<double> effective_viscosity;
};

inline void check(int status) {
  if (status != CFSA_OK) throw std::runtime_error(cfsa_last_error());
}

// Allocator for page-locked host memory (device-to-host copies land by DMA).
template <class T>
struct PinnedAllocator {
  using value_type = T;
  PinnedAllocator() = default;
  template <class U>
  PinnedAllocator(const PinnedAllocator<U> &) {}
  T *allocate(size_t n) {
    void *p = nullptr;
    if (cfsa_host_alloc(n * sizeof(T), &p) != CFSA_OK || (!p && n)) throw std::bad_alloc();
    return static_cast<T *>(p);
  }
  void deallocate(T *p, size_t) { cfsa_host_free(p); }
  template <class U>
  bool operator==(const PinnedAllocator<U> &) const { return true; }
  template <class U>
  bool operator!=(const PinnedAllocator<U> &) const { return false; }
};
using Field3 = std::vector<Vector3d, PinnedAllocator<Vector3d>>;
using Field1 = std::vector<double, PinnedAllocator<double>>;

// Mesh::Face (Mesh.hpp:23-33)
struct Face {
  int tet_a = -1, tet_b = -1, face_id_a = -1, face_id_b = -1;
  double area = 0.0;
  std::array<int, 3> vertids{};
  Vector3d normal{}, center{};
};

class Mesh {
 public:
  static Mesh structuredRectangularPrism(double length, double width, double height,
                                         int nx = 1, int ny = 1, int nz = 1) {
    cfsa_mesh *h = nullptr;
    check(cfsa_mesh_structured_prism(length, width, height, nx, ny, nz, &h));
    return Mesh(h);
  }
  static Mesh cylinder(double length, double radius, int divisions, double max_element_size) {
    cfsa_mesh *h = nullptr;
    check(cfsa_mesh_cylinder(length, radius, divisions, max_element_size, &h));
    return Mesh(h);
  }
  static Mesh fromMshFile(const std::string &filename) {
    cfsa_mesh *h = nullptr;
    check(cfsa_mesh_from_msh(filename.c_str(), &h));
    return Mesh(h);
  }
  static void setupBoundaryConditions(const Vector3d &inlet_to_outlet_direction,
                                      double percent_inlet_to_outlet_coverage, Mesh &mesh) {
    check(cfsa_mesh_setup_boundary_conditions(mesh.st_->h, inlet_to_outlet_direction.data(),
                                              percent_inlet_to_outlet_coverage));
    mesh.st_->labels.clear();
  }
  size_t nVertices() const { return (size_t)sizes()[0]; }
  size_t nTets() const { return (size_t)sizes()[1]; }
  size_t getNumberOfEdgeNodes() const { return (size_t)sizes()[2]; }
  size_t nFaces() const { return (size_t)sizes()[3]; }
  size_t getP1plusP2DegreesOfFreedom() const { return (size_t)sizes()[4]; }
  Vector3d getVertexPositon(size_t vi) const {
    const std::vector<double> &v = vertices();
    return {v.at(3 * vi), v.at(3 * vi + 1), v.at(3 * vi + 2)};
  }
  // fluid labels: per-vertex writes are kept on the host and pushed (with the
  // P2 rule applied) by setP2BoundariesFromP1Boundaries or the solver's next
  // call that reads the mesh
  FluidBCType getFluidVertexBC(size_t vi) const { return (FluidBCType)labels().at(vi); }
  void setFluidVertexBC(size_t vi, FluidBCType bc) {
    labels().at(vi) = (uint8_t)bc;
    st_->dirty = true;
  }
  void setP2BoundariesFromP1Boundaries() { sync(); }
  FluidBCType getP2FluidVertexBC(size_t k) const {
    sync();
    std::vector<uint8_t> b(getNumberOfEdgeNodes());
    check(cfsa_mesh_get_p2_bc(st_->h, b.data()));
    return (FluidBCType)b.at(k);
  }
  std::vector<uint8_t> p1_bc() const { return labels(); }
  const std::vector<double> &vertices() const {
    if (st_->coords.empty()) {
      st_->coords.resize(nVertices() * 3);
      check(cfsa_mesh_get_vertices(st_->h, st_->coords.data()));
    }
    return st_->coords;
  }
  std::vector<double> edgeNodes() const {
    std::vector<double> v(getNumberOfEdgeNodes() * 3);
    check(cfsa_mesh_get_edge_nodes(st_->h, v.data()));
    return v;
  }
  std::vector<Face> getFaces() const {
    const size_t nf = nFaces();
    std::vector<int32_t> v(nf * 3), t(nf * 4);
    std::vector<double> g(nf * 7);
    check(cfsa_mesh_get_faces(st_->h, v.data(), t.data(), g.data()));
    std::vector<Face> out(nf);
    for (size_t f = 0; f < nf; ++f) {
      Face &F = out[f];
      F.tet_a = t[4 * f];
      F.tet_b = t[4 * f + 1];
      F.face_id_a = t[4 * f + 2];
      F.face_id_b = t[4 * f + 3];
      for (int c = 0; c < 3; ++c) {
        F.vertids[c] = v[3 * f + c];
        F.normal[c] = g[7 * f + c];
        F.center[c] = g[7 * f + 4 + c];
      }
      F.area = g[7 * f + 3];
    }
    return out;
  }
  // Mesh::updateGeometry after moving the vertices (Mesh.hpp:98-101)
  void updateVertices(const std::vector<double> &coords) {
    check(cfsa_mesh_update_vertices(st_->h, coords.data()));
    st_->coords.clear();
  }
  void setFluidVertexBCs(const std::vector<uint8_t> &bc) {
    check(cfsa_mesh_set_p1_bc(st_->h, bc.data()));
    st_->labels = bc;
    st_->dirty = false;
  }
  // pushes labels written with setFluidVertexBC to the library
  void sync() const {
    if (!st_->dirty) return;
    check(cfsa_mesh_set_p1_bc(st_->h, st_->labels.data()));
    st_->dirty = false;
  }
  const cfsa_mesh *handle() const {
    sync();
    return st_->h;
  }

 private:
  struct State {
    cfsa_mesh *h = nullptr;
    std::vector<uint8_t> labels;  // host copy of the P1 labels (empty: not read yet)
    std::vector<double> coords;   // host copy of the vertices (empty: not read yet)
    bool dirty = false;
    ~State() { cfsa_mesh_destroy(h); }
  };
  explicit Mesh(cfsa_mesh *h) : st_(std::make_shared<State>()) { st_->h = h; }
  std::array<int64_t, 5> sizes() const {
    std::array<int64_t, 5> s{};
    check(cfsa_mesh_sizes(st_->h, s.data()));
    return s;
  }
  std::vector<uint8_t> &labels() const {
    if (st_->labels.empty()) {
      st_->labels.resize(nVertices());
      check(cfsa_mesh_get_p1_bc(st_->h, st_->labels.data()));
    }
    return st_->labels;
  }
  std::shared_ptr<State> st_;
};

class NavierStokesSolver {
 public:
  NavierStokesSolver() { check(cfsa_ns_create(&h_)); }
  ~NavierStokesSolver() {
    (void)cfsa_ns_wait_fields(h_);
    cfsa_ns_destroy(h_);
  }
  NavierStokesSolver(const NavierStokesSolver &) = delete;
  NavierStokesSolver &operator=(const NavierStokesSolver &) = delete;

  // NavierStokes.cpp:10-68
  void initialize(const std::shared_ptr<Mesh> mesh_ptr, const Fluid &fluid) {
    mesh_ = mesh_ptr;
    check(cfsa_ns_wait_fields(h_));
    check(cfsa_ns_initialize(h_, mesh_ptr->handle(), fluid.density, fluid.viscosity));
    int64_t sz[6];
    check(cfsa_ns_get_sizes(h_, sz));
    nv_ = (size_t)sz[0];
    nu_ = (size_t)sz[3];
    // sized to every velocity node and filled with the viscosity, as the
    // reference does (NavierStokes.cpp:20); the solver does not read it
    effective_viscosity_ = fluid.effective_viscosity;
    effective_viscosity_.resize(nu_, fluid.viscosity);
    for (auto &b : buf_) {
      b.u.assign(nu_, Vector3d{0.0, 0.0, 0.0});
      b.p.assign(nv_, 0.0);
    }
    step_ = 0;
    state_[0] = state_[1] = Buf::kStale;
  }
  void build_sparse_matrices() {
    mesh_->sync();
    check(cfsa_ns_build_sparse_matrices(h_));
  }
  bool pisoStep() {
    if (mesh_) mesh_->sync();
    const int st = cfsa_ns_piso_step(h_);
    if (st != CFSA_OK && st != CFSA_ERR_MOMENTUM_NOCONV && st != CFSA_ERR_PRESSURE_NOCONV &&
        st != CFSA_ERR_NAN && st != CFSA_ERR_NOT_INIT)
      throw std::runtime_error(cfsa_last_error());
    after_step();
    return st == CFSA_OK;
  }
  bool simpleStep() {
    if (mesh_) mesh_->sync();
    const int st = cfsa_ns_simple_step(h_);
    after_step();
    return st == CFSA_OK;
  }
  // the parts of the step (NavierStokes.cpp:416-524, 586-668, 702-771, 804-813)
  bool solveMomentumPredictor() { return part(cfsa_ns_solve_momentum_predictor); }
  bool solvePressurePoisson() { return part(cfsa_ns_solve_pressure_poisson); }
  bool correctVelocity() { return part(cfsa_ns_correct_velocity); }
  bool normalizePressire() { return part(cfsa_ns_normalize_pressure); }
  // NavierStokes.cpp:773-802: x = L^-1 rhs (BiCGSTAB, 1e-8, 1000 iterations)
  bool solvePressureBiCGSTAB(const std::vector<double> &rhs, std::vector<double> &x) {
    if (rhs.size() != nv_) throw std::invalid_argument("solvePressureBiCGSTAB: rhs has nv entries");
    x.assign(nv_, 0.0);
    const int st = cfsa_ns_solve_pressure_bicgstab(h_, rhs.data(), x.data(), nullptr);
    if (st == CFSA_ERR_PRESSURE_NOCONV) {
      std::cerr << "[NS] BiCGSTAB failed" << std::endl;
      return false;
    }
    check(st);
    return true;
  }
  // NavierStokes.hpp:92-95: the GPU assembles the four operators in one pass,
  // so each of these re-assembles all of them (the same values)
  void buildGradientMatrices() { build_sparse_matrices(); }
  void buildMassMatrix() { build_sparse_matrices(); }
  void buildStiffnessMatrix() { build_sparse_matrices(); }
  void buildPoissonMatrix() { build_sparse_matrices(); }
  bool hasNans() {
    int out = 0;
    check(cfsa_ns_has_nans(h_, &out));
    return out != 0;
  }
  // NavierStokes.cpp:526-584 and :670-700
  std::vector<Vector3d> computeAdvectionRHS() {
    std::vector<Vector3d> out(nu_);
    check(cfsa_ns_compute_advection(h_, out.empty() ? nullptr : out[0].data()));
    return out;
  }
  template <class Vec>
  void computeDivergence(const Vec &u, std::vector<double> &div_out) {
    if (u.size() != nu_) throw std::invalid_argument("computeDivergence: u has nu entries");
    div_out.assign(nv_, 0.0);
    check(cfsa_ns_compute_divergence(h_, u[0].data(), div_out.data()));
  }

  // boundary forces per P1 vertex (NavierStokes.cpp:841-965)
  std::vector<Vector3d> computePressureForces() const { return forces(0); }
  std::vector<Vector3d> computeShearStress() const { return forces(1); }
  std::vector<Vector3d> computeTotalFluidForces() const { return forces(2); }
  // faces in the mesh's sorted-triple order (NavierStokes.cpp:357-372)
  double getMeanFacePressure(size_t fi) const {
    double out = 0.0;
    check(cfsa_ns_get_mean_face_pressure(h_, (int64_t)fi, &out));
    return out;
  }
  Vector3d getPressureForceAtFace(size_t fi) const {
    Vector3d out{};
    check(cfsa_ns_get_pressure_force_at_face(h_, (int64_t)fi, out.data()));
    return out;
  }

  // setters (NavierStokes.hpp:167-186)
  void setMeanInletVelocity(const Vector3d &v) { check(cfsa_ns_set_mean_inlet_velocity(h_, v.data())); }
  void setInletProfile(const std::vector<double> &g) { check(cfsa_ns_set_inlet_profile(h_, g.data())); }
  void setOutletType(OutletType t) { check(cfsa_ns_set_outlet_type(h_, (int)t)); }
  void setOutletPressure(double p) { check(cfsa_ns_set_outlet_pressure(h_, p)); }
  void setInletType(InletType t) { check(cfsa_ns_set_inlet_type(h_, (int)t)); }
  void setNumberOfCorrections(int n) { check(cfsa_ns_set_number_of_corrections(h_, n)); }
  void setRelaxU(double r) { check(cfsa_ns_set_relax_u(h_, r)); }
  void setRelaxP(double r) { check(cfsa_ns_set_relax_p(h_, r)); }
  void setReferencePressureNode(int node) { check(cfsa_ns_set_reference_pressure_node(h_, node)); }
  void setDt(double dt) { check(cfsa_ns_set_dt(h_, dt)); }
  void setTolerance(double tol) { check(cfsa_ns_set_tolerance(h_, tol)); }
  void setOptions(const cfsa_solver_options &o) { check(cfsa_ns_set_options(h_, &o)); }
  cfsa_solver_options options() const {
    cfsa_solver_options o{};
    check(cfsa_ns_get_options(h_, &o));
    return o;
  }

  // getters (NavierStokes.hpp:146-189)
  double getViscosity() const { return scalar(2); }
  double getDensity() const { return scalar(3); }
  double getTolerance() const { return scalar(4); }
  double getDt() const { return scalar(1); }
  double getTime() const { return scalar(0); }
  const std::vector<double> &getEffectiveViscosity() const { return effective_viscosity_; }
  // Whole fields: the first call reads them from the device into pinned host
  // memory; after that every step starts copying its fields on a second
  // stream as it returns, so a read after the step waits only for what of
  // that copy is left.
  const Field3 &getVelocity() const { return fields().u; }
  const Field1 &getPressure() const { return fields().p; }
  // One node: from the current host copy when it is complete, otherwise the
  // node's values alone (no whole-field copy).
  Vector3d getVelocityAtNode(size_t ni) const {
    if (ni >= nu_) throw std::out_of_range("getVelocityAtNode");
    if (state_[cur()] == Buf::kReady) return buf_[cur()].u[ni];
    Vector3d v{};
    check(cfsa_ns_get_velocity_at_node(h_, (int64_t)ni, v.data()));
    return v;
  }
  double getPressureAtNode(size_t ni) const {
    if (ni >= nv_) throw std::out_of_range("getPressureAtNode");
    if (state_[cur()] == Buf::kReady) return buf_[cur()].p[ni];
    double v = 0.0;
    check(cfsa_ns_get_pressure_at_node(h_, (int64_t)ni, &v));
    return v;
  }
  const Mesh &getMesh() const { return *mesh_; }
  cfsa_step_stats stepStats() const {
    cfsa_step_stats s{};
    check(cfsa_ns_get_step_stats(h_, &s));
    return s;
  }
  cfsa_ns *handle() const { return h_; }

 private:
  struct Buf {
    enum State { kStale, kPending, kReady };
    Field3 u;
    Field1 p;
  };
  size_t cur() const { return step_ & 1; }
  bool part(int (*f)(cfsa_ns *)) {
    if (mesh_) mesh_->sync();
    const int st = f(h_);
    if (st != CFSA_OK && st != CFSA_ERR_MOMENTUM_NOCONV && st != CFSA_ERR_PRESSURE_NOCONV &&
        st != CFSA_ERR_NOT_INIT)
      throw std::runtime_error(cfsa_last_error());
    after_step();  // the fields changed: host copies are stale
    return st == CFSA_OK;
  }
  void after_step() {
    ++step_;
    state_[cur()] = Buf::kStale;
    if (reading_) {
      // start this step's copy into the buffer the caller is not holding
      check(cfsa_ns_get_fields_async(h_, buf_[cur()].u.empty() ? nullptr : buf_[cur()].u[0].data(),
                                     buf_[cur()].p.data()));
      state_[cur()] = Buf::kPending;
      state_[cur() ^ 1] = Buf::kStale;
    }
  }
  const Buf &fields() const {
    reading_ = true;
    Buf &b = buf_[cur()];
    if (state_[cur()] == Buf::kStale) {
      check(cfsa_ns_get_fields_async(h_, b.u.empty() ? nullptr : b.u[0].data(), b.p.data()));
      state_[cur()] = Buf::kPending;
    }
    if (state_[cur()] == Buf::kPending) {
      check(cfsa_ns_wait_fields(h_));
      state_[cur()] = Buf::kReady;
    }
    return b;
  }
  std::vector<Vector3d> forces(int which) const {
    if (mesh_) mesh_->sync();
    std::vector<Vector3d> out(nv_);
    const int st = which == 0   ? cfsa_ns_compute_pressure_forces(h_, out[0].data())
                   : which == 1 ? cfsa_ns_compute_shear_stress(h_, out[0].data())
                                : cfsa_ns_compute_total_fluid_forces(h_, out[0].data());
    check(st);
    return out;
  }
  double scalar(int which) const {
    double v = 0;
    check(cfsa_ns_get_scalar(h_, which, &v));
    return v;
  }
  cfsa_ns *h_ = nullptr;
  std::shared_ptr<Mesh> mesh_;
  size_t nv_ = 0, nu_ = 0;
  std::vector<double> effective_viscosity_;
  // double-buffered host copies of (u, p): step k's copy lands in buffer k % 2
  mutable Buf buf_[2];
  mutable Buf::State state_[2] = {Buf::kStale, Buf::kStale};
  mutable bool reading_ = false;
  uint64_t step_ = 0;
};

// NavierStokesIntegrator (FEMIntegrator.hpp:10-92)
class NavierStokesIntegrator {
 public:
  explicit NavierStokesIntegrator(std::shared_ptr<NavierStokesSolver> solver)
      : solver_(std::move(solver)) {}
  bool step() {
    if (!solver_) return false;
    const bool ok = solver_->pisoStep();
    if (ok) {
      time_ = solver_->getTime();
      step_count_++;
    }
    return ok;
  }
  bool advanceSteps(int n) {
    for (int i = 0; i < n; ++i)
      if (!step()) return false;
    return true;
  }
  bool runUntil(double target_time) {
    while (time_ < target_time)
      if (!step()) return false;
    return true;
  }
  double getTime() const { return time_; }
  int getStepCount() const { return step_count_; }
  std::shared_ptr<NavierStokesSolver> getSolver() { return solver_; }

 private:
  std::shared_ptr<NavierStokesSolver> solver_;
  double time_ = 0.0;
  int step_count_ = 0;
};

// NavierStokesSimpleSolver (NavierStokes.hpp:201-251): simpleStep() time
// stepping over a shared solver; its time is the solver's own
class NavierStokesSimpleSolver {
 public:
  explicit NavierStokesSimpleSolver(std::shared_ptr<NavierStokesSolver> ns_solver)
      : solver_(std::move(ns_solver)) {}
  bool step() { return solver_ ? solver_->simpleStep() : false; }
  bool advanceSteps(int n) {
    for (int i = 0; i < n; ++i)
      if (!step()) return false;
    return true;
  }
  bool runUntil(double target_time) {
    while (solver_->getTime() < target_time)
      if (!step()) return false;
    return true;
  }
  double getTime() const { return solver_->getTime(); }
  std::shared_ptr<NavierStokesSolver> getSolver() { return solver_; }

 private:
  std::shared_ptr<NavierStokesSolver> solver_;
};

}  // namespace api

namespace io {
// io::exportToVtk(filename, ns_solver, scale) (VtkExport.cpp:238-268)
inline void exportToVtk(const std::string &filename, const api::NavierStokesSolver &ns,
                        double scale = 1.0) {
  if (cfsa_ns_export_vtk(ns.handle(), filename.c_str(), scale) != CFSA_OK)
    std::cerr << cfsa_last_error() << std::endl;
}
inline void exportToVtk(const std::string &filename, const api::Mesh &mesh,
                        double scale = 1.0) {
  if (cfsa_mesh_export_vtk(mesh.handle(), filename.c_str(), scale) != CFSA_OK)
    std::cerr << cfsa_last_error() << std::endl;
}
}  // namespace io
}  // namespace cfsa
