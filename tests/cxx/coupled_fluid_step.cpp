// The fluid side of the reference's CoupledSolver, built against
// NavierStokesSolver.hpp alone: initializeNavierStokesSolver
// (CoupledSolver.cpp:28-48: initialize, computeTotalFluidForces, then the
// Dirichlet outlet), the runner's setters (src/main.cpp:120-122), and
// integrateStep (:113-129) with boundary updates as
// updateBoundariesFromParticlePositions makes them (:141-175: Internal
// vertices inside a cell become Wall, then setP2BoundariesFromP1Boundaries),
// build_sparse_matrices, fluidStep (:86-97, a failed step throws) and the
// force hand-off (computeTotalFluidForces; computePressureForces and
// computeShearStress as interpolateFluidForcesToParticles calls them).
// Writes u, p and the last forces to <prefix>_{u,p,ft,fp,fs}.bin.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "NavierStokesSolver.hpp"

using namespace cfsa::api;

template <class V>
static void dump(const std::string &name, const V &v, size_t doubles) {
  FILE *f = std::fopen(name.c_str(), "wb");
  std::fwrite(v.data(), sizeof(double), doubles, f);
  std::fclose(f);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string prefix = argv[1];
  auto mesh = std::make_shared<Mesh>(Mesh::cylinder(10.0, 1.0, 16, 0.1));
  Mesh::setupBoundaryConditions(Vector3d{1, 0, 0}, 1.0, *mesh);
  const size_t nv = mesh->nVertices();
  std::vector<FluidBCType> original(nv);
  for (size_t vi = 0; vi < nv; ++vi) original[vi] = mesh->getFluidVertexBC(vi);

  Fluid blood;
  blood.density = 1060.0;
  blood.viscosity = 0.004;
  auto fluid_solver = std::make_shared<NavierStokesSolver>();
  fluid_solver->initialize(mesh, blood);
  std::vector<Vector3d> fluid_forces = fluid_solver->computeTotalFluidForces();
  fluid_solver->setOutletType(OutletType::DirichletPressure);
  fluid_solver->setMeanInletVelocity(Vector3d{0.5, 0.0, 0.0});
  fluid_solver->setOutletPressure(0.0);
  fluid_solver->setDt(0.01);
  if (fluid_forces.size() != nv) return 3;

  std::vector<Vector3d> pressure_forces, shear_forces;
  for (int step = 0; step < 6; ++step) {
    // a "cell" sphere that moves along the axis: the Internal vertices inside
    // it become walls, everything else keeps its original label
    const double cx = 3.0 + 0.5 * step, r = 0.35;
    for (size_t vi = 0; vi < nv; ++vi) mesh->setFluidVertexBC(vi, original[vi]);
    for (size_t vi = 0; vi < nv; ++vi) {
      if (mesh->getFluidVertexBC(vi) != FluidBCType::Internal) continue;
      const Vector3d x = mesh->getVertexPositon(vi);
      const double d2 = (x[0] - cx) * (x[0] - cx) + x[1] * x[1] + x[2] * x[2];
      if (d2 < r * r) mesh->setFluidVertexBC(vi, FluidBCType::Wall);
    }
    mesh->setP2BoundariesFromP1Boundaries();
    fluid_solver->build_sparse_matrices();
    pressure_forces = fluid_solver->computePressureForces();
    shear_forces = fluid_solver->computeShearStress();
    if (!fluid_solver->pisoStep())
      throw std::runtime_error("[Coupled Solver] Navier Stokes PISO step failed\n");
    fluid_forces = fluid_solver->computeTotalFluidForces();
    std::cout << "[Coupled Solver] Fluid Step complete." << std::endl;
  }
  // one face-level accessor pair on a boundary face
  const auto faces = mesh->getFaces();
  size_t fb = 0;
  while (fb < faces.size() && faces[fb].tet_b >= 0) ++fb;
  const double pm = fluid_solver->getMeanFacePressure(fb);
  const Vector3d pf = fluid_solver->getPressureForceAtFace(fb);
  const double a = faces[fb].area;
  for (int c = 0; c < 3; ++c)
    if (std::fabs(pf[c] + faces[fb].normal[c] * a * pm) > 1e-12 * (1 + std::fabs(pf[c])))
      return 4;
  std::vector<double> div;
  fluid_solver->computeDivergence(fluid_solver->getVelocity(), div);
  if (div.size() != nv) return 5;
  dump(prefix + "_u.bin", fluid_solver->getVelocity(), 3 * mesh->getP1plusP2DegreesOfFreedom());
  dump(prefix + "_p.bin", fluid_solver->getPressure(), nv);
  dump(prefix + "_ft.bin", fluid_forces, 3 * nv);
  dump(prefix + "_fp.bin", pressure_forces, 3 * nv);
  dump(prefix + "_fs.bin", shear_forces, 3 * nv);
  dump(prefix + "_div.bin", div, nv);
  return 0;
}
