// C++ host API with the reference's class and method names, over the C ABI of
// libcfsa_ns.so.  A caller written against the reference
//   auto mesh = std::make_shared<Mesh>(Mesh::structuredRectangularPrism(...));
//   Mesh::setupBoundaryConditions(dir, coverage, *mesh);
//   auto ns = std::make_shared<NavierStokesSolver>();
//   ns->initialize(mesh, fluid); ns->setDt(dt); ns->pisoStep();
//   io::exportToVtk(name, *ns);
// compiles against this header with `using namespace cfsa::api;` (velocity
// vectors are std::array<double, 3> instead of Eigen::Vector3d).
//
// Reference: include/FEM/NavierStokes.hpp:13-193, include/Mesh/Mesh.hpp:35-220,
// include/Numerics/FEMIntegrator.hpp:10-92, include/IO/VtkExport.hpp:20-21.
#pragma once

#include <array>
#include <cstdint>
#include <iostream>
#include <memory>
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

struct Fluid {
  double density = 1000;
  double viscosity = 0.001;
  TurbulenceModel turbuelence_model = TurbulenceModel::Laminar;
  ViscosityModel viscosity_model = ViscosityModel::Newtonian;
};

inline void check(int status) {
  if (status != CFSA_OK) throw std::runtime_error(cfsa_last_error());
}

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
    check(cfsa_mesh_setup_boundary_conditions(mesh.h_.get(), inlet_to_outlet_direction.data(),
                                              percent_inlet_to_outlet_coverage));
  }
  size_t nVertices() const { return (size_t)sizes()[0]; }
  size_t nTets() const { return (size_t)sizes()[1]; }
  size_t getNumberOfEdgeNodes() const { return (size_t)sizes()[2]; }
  size_t nFaces() const { return (size_t)sizes()[3]; }
  size_t getP1plusP2DegreesOfFreedom() const { return (size_t)sizes()[4]; }
  FluidBCType getFluidVertexBC(size_t vi) const { return (FluidBCType)p1_bc().at(vi); }
  std::vector<uint8_t> p1_bc() const {
    std::vector<uint8_t> b(nVertices());
    check(cfsa_mesh_get_p1_bc(h_.get(), b.data()));
    return b;
  }
  std::vector<double> vertices() const {
    std::vector<double> v(nVertices() * 3);
    check(cfsa_mesh_get_vertices(h_.get(), v.data()));
    return v;
  }
  std::vector<double> edgeNodes() const {
    std::vector<double> v(getNumberOfEdgeNodes() * 3);
    check(cfsa_mesh_get_edge_nodes(h_.get(), v.data()));
    return v;
  }
  // Mesh::updateGeometry after moving the vertices (Mesh.hpp:98-101)
  void updateVertices(const std::vector<double> &coords) {
    check(cfsa_mesh_update_vertices(h_.get(), coords.data()));
  }
  void setFluidVertexBCs(const std::vector<uint8_t> &bc) {
    check(cfsa_mesh_set_p1_bc(h_.get(), bc.data()));
  }
  const cfsa_mesh *handle() const { return h_.get(); }

 private:
  explicit Mesh(cfsa_mesh *h) : h_(h, cfsa_mesh_destroy) {}
  std::array<int64_t, 5> sizes() const {
    std::array<int64_t, 5> s{};
    check(cfsa_mesh_sizes(h_.get(), s.data()));
    return s;
  }
  std::shared_ptr<cfsa_mesh> h_;
};

class NavierStokesSolver {
 public:
  NavierStokesSolver() { check(cfsa_ns_create(&h_)); }
  ~NavierStokesSolver() { cfsa_ns_destroy(h_); }
  NavierStokesSolver(const NavierStokesSolver &) = delete;
  NavierStokesSolver &operator=(const NavierStokesSolver &) = delete;

  void initialize(const std::shared_ptr<Mesh> mesh_ptr, const Fluid &fluid) {
    mesh_ = mesh_ptr;
    check(cfsa_ns_initialize(h_, mesh_ptr->handle(), fluid.density, fluid.viscosity));
    dirty_ = true;
  }
  void build_sparse_matrices() { check(cfsa_ns_build_sparse_matrices(h_)); }
  bool pisoStep() {
    dirty_ = true;
    const int st = cfsa_ns_piso_step(h_);
    if (st != CFSA_OK && st != CFSA_ERR_MOMENTUM_NOCONV && st != CFSA_ERR_PRESSURE_NOCONV &&
        st != CFSA_ERR_NAN && st != CFSA_ERR_NOT_INIT)
      throw std::runtime_error(cfsa_last_error());
    return st == CFSA_OK;
  }
  bool simpleStep() {
    dirty_ = true;
    return cfsa_ns_simple_step(h_) == CFSA_OK;
  }
  bool hasNans() {
    int out = 0;
    check(cfsa_ns_has_nans(h_, &out));
    return out != 0;
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

  // getters (NavierStokes.hpp:146-189); field copies are cached per step
  double getViscosity() const { return scalar(2); }
  double getDensity() const { return scalar(3); }
  double getTolerance() const { return scalar(4); }
  double getDt() const { return scalar(1); }
  double getTime() const { return scalar(0); }
  const std::vector<Vector3d> &getVelocity() const {
    refresh();
    return velocity_;
  }
  const std::vector<double> &getPressure() const {
    refresh();
    return pressure_;
  }
  const Vector3d &getVelocityAtNode(size_t ni) const { return getVelocity().at(ni); }
  double getPressureAtNode(size_t ni) const { return getPressure().at(ni); }
  const Mesh &getMesh() const { return *mesh_; }
  cfsa_step_stats stepStats() const {
    cfsa_step_stats s{};
    check(cfsa_ns_get_step_stats(h_, &s));
    return s;
  }
  cfsa_ns *handle() const { return h_; }

 private:
  double scalar(int which) const {
    double v = 0;
    check(cfsa_ns_get_scalar(h_, which, &v));
    return v;
  }
  void refresh() const {
    if (!dirty_) return;
    int64_t sz[6];
    check(cfsa_ns_get_sizes(h_, sz));
    velocity_.resize((size_t)sz[3]);
    pressure_.resize((size_t)sz[0]);
    check(cfsa_ns_get_velocity(h_, velocity_.empty() ? nullptr : velocity_[0].data()));
    check(cfsa_ns_get_pressure(h_, pressure_.data()));
    dirty_ = false;
  }
  cfsa_ns *h_ = nullptr;
  std::shared_ptr<Mesh> mesh_;
  mutable bool dirty_ = true;
  mutable std::vector<Vector3d> velocity_;
  mutable std::vector<double> pressure_;
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
