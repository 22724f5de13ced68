// A direct-API caller shaped like the reference's examples/NavierStokes.cpp
// (lines 43-124: set the inlet before initialize, Fluid, setters, the
// integrator loop with periodic field statistics and VTK output), built
// against NavierStokesSolver.hpp alone.  The mesh is a 6^3 Kuhn cube (the
// example's OBJ mesh needs TetGen).  Writes the final velocity and pressure to
// <prefix>_u.bin / <prefix>_p.bin for comparison with the CPU oracle.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>

#include "NavierStokesSolver.hpp"

using namespace cfsa::api;

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string prefix = argv[1];
  const double inlet_velocity = 1.0;
  const double outlet_pressure = 0.0;
  const double density = 1000.0;
  const double viscosity = 0.003;
  const double dt = 1e-3;
  const int nsteps = 12;
  const int output_interval = 5;

  auto mesh = std::make_shared<Mesh>(Mesh::structuredRectangularPrism(2.0, 1.0, 1.0, 12, 6, 6));
  Vector3d flow_direction{1, 0, 0};
  Mesh::setupBoundaryConditions(flow_direction, 1, *mesh);

  auto ns_solver = std::make_shared<NavierStokesSolver>();
  ns_solver->setMeanInletVelocity(Vector3d{inlet_velocity, 0.0, 0.0});
  Fluid water;
  water.density = density;
  water.viscosity = viscosity;
  ns_solver->initialize(mesh, water);
  ns_solver->setDt(dt);
  ns_solver->setOutletType(OutletType::DirichletPressure);
  ns_solver->setOutletPressure(outlet_pressure);
  ns_solver->setInletType(InletType::Uniform);
  ns_solver->setRelaxU(0.5);
  ns_solver->setRelaxP(0.5);
  if (ns_solver->getEffectiveViscosity().size() != mesh->getP1plusP2DegreesOfFreedom()) return 3;

  NavierStokesIntegrator integrator(ns_solver);
  int step = 0;
  while (integrator.getStepCount() < nsteps) {
    if (!integrator.step()) {
      std::cerr << "[Error] Simulation failed at step " << step << std::endl;
      return 1;
    }
    if (step % output_interval == 0) {
      double max_vel = 0.0, min_p = 1e9, max_p = -1e9;
      const auto &velocities = ns_solver->getVelocity();
      const auto &pressures = ns_solver->getPressure();
      for (const auto &v : velocities)
        max_vel = std::max(max_vel, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
      for (const auto &p : pressures) {
        min_p = std::min(min_p, p);
        max_p = std::max(max_p, p);
      }
      std::cout << "  Step " << step << ": t = " << integrator.getTime()
                << " s, max velocity = " << max_vel << ", pressure range [" << min_p << ", "
                << max_p << "]" << std::endl;
      cfsa::io::exportToVtk(prefix + "_t" + std::to_string(step) + ".vtk", *ns_solver);
    }
    step++;
  }
  const auto &u = ns_solver->getVelocity();
  const auto &p = ns_solver->getPressure();
  // the per-node getters agree with the whole fields
  for (size_t i : {size_t(0), u.size() / 2, u.size() - 1})
    if (ns_solver->getVelocityAtNode(i) != u[i]) return 4;
  if (ns_solver->getPressureAtNode(p.size() / 2) != p[p.size() / 2]) return 5;
  FILE *f = std::fopen((prefix + "_u.bin").c_str(), "wb");
  std::fwrite(u.data(), sizeof(double), u.size() * 3, f);
  std::fclose(f);
  f = std::fopen((prefix + "_p.bin").c_str(), "wb");
  std::fwrite(p.data(), sizeof(double), p.size(), f);
  std::fclose(f);
  std::cout << "[Complete] steps " << integrator.getStepCount() << " time "
            << integrator.getTime() << std::endl;
  return 0;
}
