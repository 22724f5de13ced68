// cfsa_simulation <config.yaml>: the fluid path of the reference's
// `simulation` program (src/main.cpp:14-55 createMesh, :60-209 runSimulation,
// NS wiring :105-123, loop :155-185) on the GPU PISO library.  Solid and DPM
// sections are reported and skipped: only the Navier-Stokes solver is part of
// this library.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "NavierStokesSolver.hpp"

using namespace cfsa::api;

static std::shared_ptr<Mesh> createMesh(const cfsa_sim_config &c) {
  std::shared_ptr<Mesh> mesh;
  switch (c.mesh_type) {
    case 1:
      std::cout << "Creating cylindrical mesh (L=" << c.cylinder_length
                << ", R=" << c.cylinder_radius << ")" << std::endl;
      mesh = std::make_shared<Mesh>(Mesh::cylinder(c.cylinder_length, c.cylinder_radius,
                                                   c.cylinder_divisions, c.max_element_size));
      break;
    case 3:
      mesh = std::make_shared<Mesh>(Mesh::structuredRectangularPrism(
          c.box_size[0], c.box_size[1], c.box_size[2], c.box_cells[0], c.box_cells[1],
          c.box_cells[2]));
      break;
    case 4:
      mesh = std::make_shared<Mesh>(Mesh::fromMshFile(c.mesh_file));
      break;
    default:
      throw std::runtime_error("mesh type needs TetGen (obj / icosphere); use cylinder, box or msh");
  }
  std::cout << "  Mesh created: " << mesh->nVertices() << " vertices, " << mesh->nTets()
            << " tets, " << mesh->nFaces() << " faces" << std::endl;
  Vector3d dir = {c.flow_direction[0], c.flow_direction[1], c.flow_direction[2]};
  Mesh::setupBoundaryConditions(dir, c.inlet_outlet_coverage, *mesh);
  return mesh;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <config.yaml>" << std::endl;
    return 1;
  }
  cfsa_sim_config cfg;
  if (cfsa_config_load(argv[1], &cfg) != CFSA_OK) {
    std::cerr << "Error: " << cfsa_last_error() << std::endl;
    return 1;
  }
  try {
    if (!cfg.ns_enabled) {
      std::cerr << "ERROR: No solver enabled in configuration!" << std::endl;
      return 1;
    }
    if (cfg.solid_enabled || cfg.dpm_enabled)
      std::cout << "[cfsa] solid / dpm sections are not part of this library; running the "
                   "fluid solver only"
                << std::endl;
    auto mesh = createMesh(cfg);
    auto ns = std::make_shared<NavierStokesSolver>();
    Fluid fluid;
    fluid.density = cfg.density;
    fluid.viscosity = cfg.viscosity;
    // initialize before the inlet velocity is set, then force the Dirichlet
    // outlet (CoupledSolver.cpp:28-48, main.cpp:118-122); relax_u / relax_p /
    // use_dirichlet_pressure are not applied on this path in the reference.
    ns->initialize(mesh, fluid);
    ns->setOutletType(OutletType::DirichletPressure);
    ns->setMeanInletVelocity({cfg.inlet_velocity[0], cfg.inlet_velocity[1], cfg.inlet_velocity[2]});
    ns->setOutletPressure(cfg.outlet_pressure);
    ns->setDt(cfg.dt);
    ns->setNumberOfCorrections(cfg.n_corrections);
    cfsa_solver_options opt = ns->options();
    opt.momentum_tol = cfg.momentum_tol;
    opt.pressure_tol = cfg.pressure_tol;
    opt.pressure_solver = cfg.pressure_solver;
    opt.precision = cfg.precision;
    ns->setOptions(opt);
    if (cfg.inlet_profile == 1) {
      // parabolic (Poiseuille) profile across the inlet: peak = |inlet_velocity|
      // on the axis of the flow direction, zero on the bounding radius
      const auto X = mesh->vertices();
      const auto E = mesh->edgeNodes();
      std::vector<double> pts(X);
      pts.insert(pts.end(), E.begin(), E.end());
      const size_t nu = pts.size() / 3;
      int ax = 0;
      for (int a = 1; a < 3; ++a)
        if (std::fabs(cfg.flow_direction[a]) > std::fabs(cfg.flow_direction[ax])) ax = a;
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (size_t i = 0; i < nu; ++i)
        for (int a = 0; a < 3; ++a) {
          lo[a] = std::min(lo[a], pts[i * 3 + a]);
          hi[a] = std::max(hi[a], pts[i * 3 + a]);
        }
      double ctr[3], rad = 0;
      for (int a = 0; a < 3; ++a) ctr[a] = 0.5 * (lo[a] + hi[a]);
      for (int a = 0; a < 3; ++a)
        if (a != ax) rad = std::max(rad, 0.5 * (hi[a] - lo[a]));
      std::vector<double> g(nu * 3, 0.0);
      for (size_t i = 0; i < nu; ++i) {
        double r2 = 0;
        for (int a = 0; a < 3; ++a)
          if (a != ax) r2 += (pts[i * 3 + a] - ctr[a]) * (pts[i * 3 + a] - ctr[a]);
        const double s = std::max(0.0, 1.0 - r2 / (rad * rad));
        for (int a = 0; a < 3; ++a) g[i * 3 + a] = s * cfg.inlet_velocity[a];
      }
      ns->setInletProfile(g);
    }
    const int per = std::max(1, cfg.output_frequency);
    const int num_outputs = cfg.total_steps / per;
    for (int i = 0; i < num_outputs; ++i) {
      std::cout << "\n========== Output Step " << i << " / " << num_outputs << " ==========" << std::endl;
      if (cfg.output_fluid)
        cfsa::io::exportToVtk(std::string(cfg.output_prefix) + "_fluid_" + std::to_string(i) + ".vtk",
                              *ns, cfg.vtk_scale);
      for (int j = 0; j < per; ++j)
        if (!ns->pisoStep())
          throw std::runtime_error("[Coupled Solver] Navier Stokes PISO step failed");
      double max_vel = 0.0;
      for (const auto &v : ns->getVelocity())
        max_vel = std::max(max_vel, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
      std::cout << "  Max fluid velocity: " << max_vel << " m/s" << std::endl;
    }
    std::cout << "\n  Simulation completed successfully!" << std::endl;
  } catch (const std::exception &e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
