// A caller of NavierStokesSimpleSolver (NavierStokes.hpp:201-251) built against
// NavierStokesSolver.hpp alone: a 4^3 Kuhn cube, outlet pressure 2, advanced
// with advanceSteps and runUntil; writes velocity, pressure and the solver's
// time to <prefix>_{u,p,t}.bin for comparison with the CPU oracle's simpleStep.
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>

#include "NavierStokesSolver.hpp"

using namespace cfsa::api;

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string prefix = argv[1];
  auto mesh = std::make_shared<Mesh>(Mesh::structuredRectangularPrism(1.0, 1.0, 1.0, 4, 4, 4));
  Mesh::setupBoundaryConditions(Vector3d{1, 0, 0}, 1, *mesh);
  auto ns = std::make_shared<NavierStokesSolver>();
  Fluid fluid;
  fluid.density = 1060.0;
  fluid.viscosity = 0.004;
  ns->initialize(mesh, fluid);
  ns->setDt(0.01);
  ns->setMeanInletVelocity(Vector3d{0.5, 0.0, 0.0});
  ns->setOutletPressure(2.0);
  NavierStokesSimpleSolver simple(ns);
  if (!simple.advanceSteps(2)) return 1;
  // one more step: runUntil stops at the first time at or past the target
  if (!simple.runUntil(simple.getTime() + 0.005)) return 1;
  if (simple.getSolver() != ns) return 3;
  const auto &u = ns->getVelocity();
  const auto &p = ns->getPressure();
  const double t = simple.getTime();
  FILE *f = std::fopen((prefix + "_u.bin").c_str(), "wb");
  std::fwrite(u.data(), sizeof(double), u.size() * 3, f);
  std::fclose(f);
  f = std::fopen((prefix + "_p.bin").c_str(), "wb");
  std::fwrite(p.data(), sizeof(double), p.size(), f);
  std::fclose(f);
  f = std::fopen((prefix + "_t.bin").c_str(), "wb");
  std::fwrite(&t, sizeof(double), 1, f);
  std::fclose(f);
  std::cout << "simple solver: t = " << t << std::endl;
  return 0;
}
