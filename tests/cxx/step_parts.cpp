// A caller of the public parts of the step (NavierStokes.hpp:101-112) built
// against NavierStokesSolver.hpp alone: on a 4^3 Kuhn cube with two
// correctors, solveMomentumPredictor followed by solvePressurePoisson /
// correctVelocity twice gives pisoStep's velocity and pressure bit for bit,
// reading the fields between the parts (the host copies follow each part);
// normalizePressire subtracts the reference node's pressure; and
// solvePressureBiCGSTAB returns x = 0 for rhs = 0 and solves L x = L x0.
// Exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <vector>

#include "NavierStokesSolver.hpp"

using namespace cfsa::api;

namespace {
int fails = 0;
void expect(bool ok, const char *what) {
  if (!ok) {
    std::cout << "FAIL: " << what << std::endl;
    ++fails;
  }
}
template <class V, class W>
bool same(const V &a, const W &b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (!(a[i] == b[i])) return false;
  return true;
}
}  // namespace

int main() {
  auto mesh = std::make_shared<Mesh>(Mesh::structuredRectangularPrism(1.0, 1.0, 1.0, 4, 4, 4));
  Mesh::setupBoundaryConditions(Vector3d{1, 0, 0}, 1, *mesh);
  Fluid fluid;
  fluid.density = 1060.0;
  fluid.viscosity = 0.004;
  NavierStokesSolver a, b;
  for (NavierStokesSolver *s : {&a, &b}) {
    s->initialize(mesh, fluid);
    s->setDt(0.01);
    s->setMeanInletVelocity(Vector3d{0.5, 0.0, 0.0});
    s->setNumberOfCorrections(2);
  }
  for (int step = 0; step < 2; ++step) {
    expect(a.pisoStep(), "pisoStep");
    const Field3 u0 = b.getVelocity();
    expect(b.solveMomentumPredictor(), "solveMomentumPredictor");
    for (int k = 0; k < 2; ++k) {
      expect(b.solvePressurePoisson(), "solvePressurePoisson");
      expect(b.correctVelocity(), "correctVelocity");
      if (step == 0 && k == 0) expect(!same(b.getVelocity(), u0), "velocity read after a part");
    }
    expect(same(b.getVelocity(), a.getVelocity()), "velocity equals pisoStep's");
    expect(same(b.getPressure(), a.getPressure()), "pressure equals pisoStep's");
  }
  expect(b.getTime() == 0.0 && a.getTime() > 0.0, "the parts leave the time alone");

  const Field1 p = a.getPressure();
  a.setReferencePressureNode(3);
  expect(a.normalizePressire(), "normalizePressire");
  const Field1 &q = a.getPressure();
  bool sub = q.size() == p.size() && q[3] == 0.0;
  for (size_t i = 0; sub && i < p.size(); ++i) sub = q[i] == p[i] - p[3];
  expect(sub, "normalizePressire subtracts the reference node's pressure");

  const size_t nv = p.size();
  std::vector<double> rhs(nv, 0.0), x(nv, 1.0);
  expect(a.solvePressureBiCGSTAB(rhs, x), "BiCGSTAB with rhs 0");
  bool zero = x.size() == nv;
  for (double v : x) zero = zero && v == 0.0;
  expect(zero, "x = 0 for rhs = 0");
  // rhs = L x0 with L the Poisson matrix read through the C ABI: a consistent
  // right-hand side for the singular Neumann Laplacian
  std::vector<double> x0(nv);
  for (size_t i = 0; i < nv; ++i) x0[i] = std::sin(0.37 * (double)i);
  int64_t sz[6] = {};
  if (cfsa_ns_get_sizes(a.handle(), sz) != CFSA_OK || sz[0] != (int64_t)nv) return 4;
  std::vector<int32_t> rp(nv + 1);
  if (cfsa_ns_get_matrix(a.handle(), 3, rp.data(), nullptr, nullptr) != CFSA_OK) return 4;
  std::vector<int32_t> ci((size_t)rp[nv]);
  std::vector<double> val((size_t)rp[nv]);
  if (cfsa_ns_get_matrix(a.handle(), 3, rp.data(), ci.data(), val.data()) != CFSA_OK) return 4;
  for (size_t i = 0; i < nv; ++i) {
    double s = 0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += val[(size_t)k] * x0[(size_t)ci[(size_t)k]];
    rhs[i] = s;
  }
  expect(a.solvePressureBiCGSTAB(rhs, x), "BiCGSTAB on L x0");
  double rn = 0, bn = 0;
  for (size_t i = 0; i < nv; ++i) {
    double s = 0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += val[(size_t)k] * x[(size_t)ci[(size_t)k]];
    rn += (rhs[i] - s) * (rhs[i] - s);
    bn += rhs[i] * rhs[i];
  }
  std::cout << "BiCGSTAB relative residual " << std::sqrt(rn / bn) << std::endl;
  expect(std::sqrt(rn / bn) <= 1e-8, "BiCGSTAB residual");
  std::cout << (fails ? "step parts: failures" : "step parts: ok") << std::endl;
  return fails ? 1 : 0;
}
