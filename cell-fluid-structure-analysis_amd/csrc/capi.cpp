// extern "C" boundary of libcfsa_ns.so; declarations and reference citations
// are in include/cfsa_ns.h.  No exception crosses this layer.
#include "../../include/cfsa_ns.h"

#include <cstring>
#include <iostream>
#include <memory>
#include <string>

#include "amg.hpp"
#include "elem_patch.hpp"
#include "mesh.hpp"
#include "solver.hpp"
#include "vtk.hpp"
#include "yaml_config.hpp"

struct cfsa_mesh {
  cfsa::Mesh m;
  uint64_t version = 0;  // bumped whenever labels or vertices change
};
struct cfsa_amg {
  cfsa::AmgHierarchy H;
};
struct cfsa_elem_patches {
  cfsa::ElemPatchHost E;
};
struct cfsa_ns {
  cfsa::NSSolver s;
  const cfsa_mesh *mesh = nullptr;  // set when initialized from a handle
  uint64_t synced = 0;              // mesh version last pushed to the device
};

namespace {

thread_local std::string g_last_error;

template <class F>
int guard(F &&f) {
  try {
    return f();
  } catch (const std::invalid_argument &e) {
    g_last_error = e.what();
    return CFSA_ERR_INVALID;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    const std::string w = e.what();
    return w.rfind("HIP error", 0) == 0 ? CFSA_ERR_DEVICE : CFSA_ERR_INVALID;
  } catch (...) {
    g_last_error = "unknown exception";
    return CFSA_ERR_INVALID;
  }
}

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define CHECK_PTR(p)                                             \
  do {                                                           \
    if (!(p)) return fail(CFSA_ERR_INVALID, "null argument " #p); \
  } while (0)

cfsa::SolverOptions to_opts(const cfsa_solver_options &o) {
  cfsa::SolverOptions s;
  s.momentum_tol = o.momentum_tol;
  s.pressure_tol = o.pressure_tol;
  s.max_iter = o.max_iter;
  s.pressure_solver = o.pressure_solver;
  s.precision = o.precision;
  s.inner_tol = o.inner_tol;
  s.max_refine = o.max_refine;
  s.check_every = o.check_every;
  s.check_every_1 = o.check_every_1;
  s.timing = o.timing;
  s.stop_scaled = o.stop_rule;
  s.assembly = o.assembly;
  s.spmv_layout = o.spmv_layout;
  s.momentum_solver = o.momentum_solver;
  s.graphs = o.graphs;
  s.initial_guess = o.initial_guess;
  s.reuse_correction = o.reuse_pressure_correction;
  s.restart_scale = o.restart_scale;
  s.amg_reuse = o.amg_reuse;
  s.elem_spmv = o.elem_spmv;
  return s;
}

void to_result(const cfsa::SolveResult &r, cfsa_solve_result *o) {
  for (int c = 0; c < 3; ++c) {
    o->iters[c] = r.iters[c];
    o->err[c] = r.err[c];
    o->restarts[c] = r.restarts[c];
  }
  o->refinements = r.refinements;
  o->inner_iters = r.inner_iters;
  o->ok = r.ok ? 1 : 0;
}

// The reference reads labels and element geometry live from the shared Mesh
// on every step; push them when the mesh handle changed since the last sync.
void sync_mesh(cfsa_ns *h) {
  if (!h->mesh || !h->s.initialized || h->mesh->version == h->synced) return;
  const cfsa::Mesh &m = h->mesh->m;
  h->s.update_mesh(m.vertices.data(), m.p1_bc.data(), m.p2_bc.data());
  h->synced = h->mesh->version;
}

int check_init(const cfsa_ns *h) {
  if (!h->s.initialized) return fail(CFSA_ERR_NOT_INIT, "[NS] Error: solver not initialized");
  return CFSA_OK;
}

}  // namespace

extern "C" {

const char *cfsa_last_error(void) { return g_last_error.c_str(); }
int cfsa_abi_version(void) { return CFSA_ABI_VERSION; }

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------
int cfsa_mesh_structured_prism(double length, double width, double height, int nx, int ny,
                               int nz, cfsa_mesh **out) {
  CHECK_PTR(out);
  return guard([&]() -> int {
    auto m = std::make_unique<cfsa_mesh>();
    m->m = cfsa::Mesh::structuredRectangularPrism(length, width, height, nx, ny, nz);
    *out = m.release();
    return CFSA_OK;
  });
}

int cfsa_mesh_cylinder(double length, double radius, int divisions, double max_volume,
                       cfsa_mesh **out) {
  CHECK_PTR(out);
  return guard([&]() -> int {
    auto m = std::make_unique<cfsa_mesh>();
    m->m = cfsa::Mesh::cylinder(length, radius, divisions, max_volume);
    *out = m.release();
    return CFSA_OK;
  });
}

int cfsa_mesh_from_msh(const char *path, cfsa_mesh **out) {
  CHECK_PTR(path);
  CHECK_PTR(out);
  return guard([&]() -> int {
    auto m = std::make_unique<cfsa_mesh>();
    try {
      m->m = cfsa::Mesh::fromMshFile(path);
    } catch (const std::runtime_error &e) {
      return fail(CFSA_ERR_IO, e.what());
    }
    *out = m.release();
    return CFSA_OK;
  });
}

int cfsa_mesh_from_arrays(int64_t nv, const double *coords, int64_t nt, const int32_t *tets,
                          cfsa_mesh **out) {
  CHECK_PTR(coords);
  CHECK_PTR(tets);
  CHECK_PTR(out);
  return guard([&]() -> int {
    auto m = std::make_unique<cfsa_mesh>();
    m->m = cfsa::Mesh::fromArrays(coords, nv, tets, nt);
    *out = m.release();
    return CFSA_OK;
  });
}

int cfsa_mesh_setup_boundary_conditions(cfsa_mesh *m, const double dir[3], double coverage) {
  CHECK_PTR(m);
  CHECK_PTR(dir);
  return guard([&]() -> int {
    cfsa::Mesh::setupBoundaryConditions(dir, coverage, m->m);
    m->version++;
    return CFSA_OK;
  });
}

int cfsa_mesh_set_p1_bc(cfsa_mesh *m, const uint8_t *bc) {
  CHECK_PTR(m);
  CHECK_PTR(bc);
  return guard([&]() -> int {
    for (int64_t i = 0; i < m->m.nVertices(); ++i)
      if (bc[i] > 4) return fail(CFSA_ERR_INVALID, "label out of range");
    m->m.p1_bc.assign(bc, bc + m->m.nVertices());
    m->m.setP2BoundariesFromP1Boundaries();
    m->version++;
    return CFSA_OK;
  });
}

int cfsa_mesh_update_vertices(cfsa_mesh *m, const double *coords) {
  CHECK_PTR(m);
  CHECK_PTR(coords);
  return guard([&]() -> int {
    m->m.vertices.assign(coords, coords + m->m.nVertices() * 3);
    m->m.updateGeometry();
    m->version++;
    return CFSA_OK;
  });
}

int cfsa_mesh_sizes(const cfsa_mesh *m, int64_t out[5]) {
  CHECK_PTR(m);
  CHECK_PTR(out);
  out[0] = m->m.nVertices();
  out[1] = m->m.nTets();
  out[2] = m->m.nEdgeNodes();
  out[3] = m->m.nFaces();
  out[4] = out[0] + out[2];
  return CFSA_OK;
}

#define MESH_COPY(fn, T, field)                                      \
  int fn(const cfsa_mesh *m, T *out) {                               \
    CHECK_PTR(m);                                                    \
    CHECK_PTR(out);                                                  \
    std::memcpy(out, m->m.field.data(), m->m.field.size() * sizeof(T)); \
    return CFSA_OK;                                                  \
  }
MESH_COPY(cfsa_mesh_get_vertices, double, vertices)
MESH_COPY(cfsa_mesh_get_tets, int32_t, tets)
MESH_COPY(cfsa_mesh_get_volumes, double, volumes)
MESH_COPY(cfsa_mesh_get_gradients, double, gradients)
MESH_COPY(cfsa_mesh_get_tet_edge_nodes, int32_t, tet_edge_nodes)
MESH_COPY(cfsa_mesh_get_edge_nodes, double, edge_nodes)
MESH_COPY(cfsa_mesh_get_edge_endpoints, int32_t, edge_endpoints)
MESH_COPY(cfsa_mesh_get_p1_bc, uint8_t, p1_bc)
MESH_COPY(cfsa_mesh_get_p2_bc, uint8_t, p2_bc)

int cfsa_mesh_get_faces(const cfsa_mesh *m, int32_t *verts, int32_t *tets, double *geom) {
  CHECK_PTR(m);
  const auto &F = m->m.faces;
  for (size_t f = 0; f < F.size(); ++f) {
    if (verts)
      for (int a = 0; a < 3; ++a) verts[f * 3 + a] = F[f].vertids[(size_t)a];
    if (tets) {
      tets[f * 4 + 0] = F[f].tet_a;
      tets[f * 4 + 1] = F[f].tet_b;
      tets[f * 4 + 2] = F[f].face_id_a;
      tets[f * 4 + 3] = F[f].face_id_b;
    }
    if (geom) {
      for (int a = 0; a < 3; ++a) geom[f * 7 + a] = F[f].normal[(size_t)a];
      geom[f * 7 + 3] = F[f].area;
      for (int a = 0; a < 3; ++a) geom[f * 7 + 4 + a] = F[f].center[(size_t)a];
    }
  }
  return CFSA_OK;
}

int cfsa_mesh_get_tet_faces(const cfsa_mesh *m, int32_t *faceids, int32_t *neighbors) {
  CHECK_PTR(m);
  if (faceids)
    std::memcpy(faceids, m->m.tet_faceids.data(), m->m.tet_faceids.size() * sizeof(int32_t));
  if (neighbors)
    std::memcpy(neighbors, m->m.tet_neighbors.data(),
                m->m.tet_neighbors.size() * sizeof(int32_t));
  return CFSA_OK;
}

int cfsa_mesh_get_min_edge_length(const cfsa_mesh *m, double *out) {
  CHECK_PTR(m);
  CHECK_PTR(out);
  *out = m->m.min_edge_length;
  return CFSA_OK;
}

int cfsa_mesh_export_vtk(const cfsa_mesh *m, const char *path, double scale) {
  CHECK_PTR(m);
  CHECK_PTR(path);
  return guard([&]() -> int {
    if (!cfsa::write_mesh_vtk(path, m->m.vertices.data(), m->m.nVertices(), m->m.tets.data(),
                              m->m.nTets(), m->m.p1_bc.data(), scale))
      return fail(CFSA_ERR_IO, std::string("[IO] Failed to open file: ") + path);
    return CFSA_OK;
  });
}

void cfsa_mesh_destroy(cfsa_mesh *m) { delete m; }

// ---------------------------------------------------------------------------
// Solver
// ---------------------------------------------------------------------------
int cfsa_ns_create(cfsa_ns **out) {
  CHECK_PTR(out);
  return guard([&]() -> int {
    *out = new cfsa_ns();
    return CFSA_OK;
  });
}

void cfsa_ns_destroy(cfsa_ns *h) {
  try {
    delete h;
  } catch (...) {
  }
}

int cfsa_ns_initialize(cfsa_ns *h, const cfsa_mesh *mesh, double density, double viscosity) {
  CHECK_PTR(h);
  CHECK_PTR(mesh);
  return guard([&]() -> int {
    const cfsa::Mesh &m = mesh->m;
    cfsa::MeshInput in;
    in.nv = m.nVertices();
    in.nt = m.nTets();
    in.ne = m.nEdgeNodes();
    in.coords = m.vertices.data();
    in.tets = m.tets.data();
    in.ten = m.tet_edge_nodes.data();
    in.p1_bc = m.p1_bc.data();
    in.p2_bc = m.p2_bc.data();
    in.faces = &m.faces;
    h->s.initialize(in, density, viscosity);
    h->mesh = mesh;
    h->synced = mesh->version;
    return CFSA_OK;
  });
}

int cfsa_ns_initialize_arrays(cfsa_ns *h, int64_t nv, const double *coords, int64_t nt,
                              const int32_t *tets, int64_t ne, const int32_t *ten,
                              const uint8_t *p1_bc, const uint8_t *p2_bc, double density,
                              double viscosity) {
  CHECK_PTR(h);
  return guard([&]() -> int {
    cfsa::MeshInput in;
    in.nv = nv;
    in.nt = nt;
    in.ne = ne;
    in.coords = coords;
    in.tets = tets;
    in.ten = ten;
    in.p1_bc = p1_bc;
    in.p2_bc = p2_bc;
    h->s.initialize(in, density, viscosity);
    h->mesh = nullptr;
    return CFSA_OK;
  });
}

int cfsa_ns_update_mesh(cfsa_ns *h, const double *coords, const uint8_t *p1_bc,
                        const uint8_t *p2_bc) {
  CHECK_PTR(h);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.update_mesh(coords, p1_bc, p2_bc);
    return CFSA_OK;
  });
}

int cfsa_ns_build_sparse_matrices(cfsa_ns *h) {
  CHECK_PTR(h);
  if (!h->s.initialized) {
    std::cerr << "[NS] Error: mesh not initialized" << std::endl;
    return fail(CFSA_ERR_NOT_INIT, "[NS] Error: mesh not initialized");
  }
  return guard([&]() -> int {
    // the reference re-reads the mesh here; geometry and labels are pushed
    // when the mesh changed since the last push
    sync_mesh(h);
    h->s.build_sparse_matrices();
    return CFSA_OK;
  });
}

int cfsa_ns_piso_step(cfsa_ns *h) {
  CHECK_PTR(h);
  return guard([&]() -> int {
    sync_mesh(h);
    const int st = h->s.piso_step();
    if (st != CFSA_OK && st != CFSA_ERR_NAN) g_last_error = h->s.last_error;
    if (st == CFSA_ERR_NAN) g_last_error = "[NS] NaN in velocity or pressure after the step";
    return st;
  });
}

int cfsa_ns_simple_step(cfsa_ns *h) {
  CHECK_PTR(h);
  return guard([&]() -> int {
    sync_mesh(h);
    const int st = h->s.simple_step();
    if (st != CFSA_OK) g_last_error = h->s.last_error;
    return st;
  });
}

namespace {
int substep(cfsa_ns *h, int (cfsa::NSSolver::*f)()) {
  CHECK_PTR(h);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    sync_mesh(h);
    const int st = (h->s.*f)();
    if (st != CFSA_OK) g_last_error = h->s.last_error;
    return st;
  });
}
}  // namespace
int cfsa_ns_solve_momentum_predictor(cfsa_ns *h) {
  return substep(h, &cfsa::NSSolver::momentum_predictor_step);
}
int cfsa_ns_solve_pressure_poisson(cfsa_ns *h) {
  return substep(h, &cfsa::NSSolver::pressure_poisson_step);
}
int cfsa_ns_correct_velocity(cfsa_ns *h) {
  CHECK_PTR(h);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    sync_mesh(h);
    h->s.correct_velocity_step();
    return CFSA_OK;
  });
}
int cfsa_ns_normalize_pressure(cfsa_ns *h) {
  CHECK_PTR(h);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.normalize_pressure_step();
    return CFSA_OK;
  });
}
int cfsa_ns_solve_pressure_bicgstab(cfsa_ns *h, const double *rhs, double *x,
                                    cfsa_solve_result *res) {
  CHECK_PTR(h);
  CHECK_PTR(rhs);
  CHECK_PTR(x);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    sync_mesh(h);
    const cfsa::SolveResult r = h->s.solve_poisson_bicgstab(rhs, x);
    if (res) to_result(r, res);
    if (!r.ok) return fail(CFSA_ERR_PRESSURE_NOCONV, "[NS] BiCGSTAB failed");
    return CFSA_OK;
  });
}

int cfsa_ns_set_mean_inlet_velocity(cfsa_ns *h, const double v[3]) {
  CHECK_PTR(h);
  CHECK_PTR(v);
  return guard([&]() -> int {
    h->s.set_mean_inlet(v);
    return CFSA_OK;
  });
}
int cfsa_ns_set_inlet_profile(cfsa_ns *h, const double *g) {
  CHECK_PTR(h);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.set_inlet_profile(g);
    return CFSA_OK;
  });
}
int cfsa_ns_set_outlet_type(cfsa_ns *h, int type) {
  CHECK_PTR(h);
  if (type < 0 || type > 2) return fail(CFSA_ERR_INVALID, "outlet type out of range");
  h->s.set_outlet_type(type);
  return CFSA_OK;
}
int cfsa_ns_set_outlet_pressure(cfsa_ns *h, double p) {
  CHECK_PTR(h);
  if (h->s.outlet_type != 0)
    std::cout << "[NS Solver] Warning: setting outlet pressure without dirichlet outlet being "
                 "the outlet type"
              << std::endl;
  h->s.outlet_pressure = p;
  return CFSA_OK;
}
int cfsa_ns_set_inlet_type(cfsa_ns *h, int type) {
  CHECK_PTR(h);
  h->s.inlet_type = type;
  return CFSA_OK;
}
int cfsa_ns_set_number_of_corrections(cfsa_ns *h, int n) {
  CHECK_PTR(h);
  h->s.n_corrections = n;
  return CFSA_OK;
}
int cfsa_ns_set_relax_u(cfsa_ns *h, double r) {
  CHECK_PTR(h);
  h->s.relax_u = r;
  return CFSA_OK;
}
int cfsa_ns_set_relax_p(cfsa_ns *h, double r) {
  CHECK_PTR(h);
  h->s.relax_p = r;
  return CFSA_OK;
}
int cfsa_ns_set_reference_pressure_node(cfsa_ns *h, int node) {
  CHECK_PTR(h);
  h->s.set_reference_node(node);
  return CFSA_OK;
}
int cfsa_ns_set_dt(cfsa_ns *h, double dt) {
  CHECK_PTR(h);
  h->s.set_dt(dt);
  return CFSA_OK;
}
int cfsa_ns_set_tolerance(cfsa_ns *h, double tol) {
  CHECK_PTR(h);
  h->s.tolerance = tol;
  return CFSA_OK;
}
int cfsa_ns_set_time(cfsa_ns *h, double t) {
  CHECK_PTR(h);
  h->s.time = t;
  return CFSA_OK;
}
int cfsa_ns_set_options(cfsa_ns *h, const cfsa_solver_options *opt) {
  CHECK_PTR(h);
  CHECK_PTR(opt);
  if (opt->max_iter < 0 || opt->check_every < 1 || opt->check_every_1 < 1 || opt->precision < 0 || opt->precision > 2 ||
      opt->pressure_solver < 0 || opt->pressure_solver > 2 || !(opt->momentum_tol >= 0) ||
      !(opt->pressure_tol >= 0) || !(opt->inner_tol > 0) || opt->max_refine < 0 ||
      opt->stop_rule < 0 || opt->stop_rule > 1 || opt->assembly < 0 || opt->assembly > 2 ||
      opt->spmv_layout < 0 || opt->spmv_layout > 2 || opt->momentum_solver < 0 ||
      opt->momentum_solver > 3 || opt->graphs < 0 || opt->graphs > 1 ||
      opt->initial_guess < 0 || opt->initial_guess > 1 || opt->reuse_pressure_correction < 0 ||
      opt->reuse_pressure_correction > 1 || !(opt->restart_scale > 0) || opt->amg_reuse < 0 ||
      opt->elem_spmv < 0 || opt->elem_spmv > 1)
    return fail(CFSA_ERR_INVALID, "invalid solver options");
  h->s.opt = to_opts(*opt);
  return CFSA_OK;
}
int cfsa_ns_get_options(const cfsa_ns *h, cfsa_solver_options *o) {
  CHECK_PTR(h);
  CHECK_PTR(o);
  const auto &s = h->s.opt;
  o->momentum_tol = s.momentum_tol;
  o->pressure_tol = s.pressure_tol;
  o->max_iter = s.max_iter;
  o->pressure_solver = s.pressure_solver;
  o->precision = s.precision;
  o->inner_tol = s.inner_tol;
  o->max_refine = s.max_refine;
  o->check_every = s.check_every;
  o->check_every_1 = s.check_every_1;
  o->timing = s.timing;
  o->stop_rule = s.stop_scaled;
  o->assembly = s.assembly;
  o->spmv_layout = s.spmv_layout;
  o->momentum_solver = s.momentum_solver;
  o->graphs = s.graphs;
  o->initial_guess = s.initial_guess;
  o->reuse_pressure_correction = s.reuse_correction;
  o->restart_scale = s.restart_scale;
  o->amg_reuse = s.amg_reuse;
  o->elem_spmv = s.elem_spmv;
  return CFSA_OK;
}

int cfsa_ns_get_sizes(const cfsa_ns *h, int64_t out[6]) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  out[0] = h->s.nv;
  out[1] = h->s.nt;
  out[2] = h->s.ne;
  out[3] = h->s.nu;
  out[4] = h->s.nnzA;
  out[5] = h->s.nnzP;
  return CFSA_OK;
}

#define NS_FIELD(fn, method, T)                  \
  int fn(cfsa_ns *h, T *p) {                     \
    CHECK_PTR(h);                                \
    CHECK_PTR(p);                                \
    if (int e = check_init(h)) return e;         \
    return guard([&]() -> int {                           \
      h->s.method(p);                            \
      return CFSA_OK;                            \
    });                                          \
  }
NS_FIELD(cfsa_ns_get_velocity, get_velocity, double)
NS_FIELD(cfsa_ns_set_velocity, set_velocity, const double)
NS_FIELD(cfsa_ns_get_velocity_star, get_velocity_star, double)
NS_FIELD(cfsa_ns_get_pressure, get_pressure, double)
NS_FIELD(cfsa_ns_set_pressure, set_pressure, const double)
NS_FIELD(cfsa_ns_get_pressure_correction, get_pressure_correction, double)
NS_FIELD(cfsa_ns_get_inv_lumped_mass, get_inv_lumped_mass, double)
NS_FIELD(cfsa_ns_compute_advection, compute_advection, double)

int cfsa_ns_get_velocity_at_node(cfsa_ns *h, int64_t node, double out[3]) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  if (node < 0 || node >= h->s.nu) return fail(CFSA_ERR_INVALID, "node out of range");
  return guard([&]() -> int {
    h->s.get_velocity_at(node, out);
    return CFSA_OK;
  });
}
int cfsa_ns_get_fields_async(cfsa_ns *h, double *u_out, double *p_out) {
  CHECK_PTR(h);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.fields_async(u_out, p_out);
    return CFSA_OK;
  });
}
int cfsa_ns_wait_fields(cfsa_ns *h) {
  CHECK_PTR(h);
  return guard([&]() -> int {
    h->s.fields_wait();
    return CFSA_OK;
  });
}
int cfsa_host_alloc(size_t bytes, void **out) {
  CHECK_PTR(out);
  *out = nullptr;
  if (bytes == 0) return CFSA_OK;
  const hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *out = nullptr;
    return fail(CFSA_ERR_DEVICE, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  }
  return CFSA_OK;
}
void cfsa_host_free(void *p) {
  if (p) (void)hipHostFree(p);
}
int cfsa_ns_get_pressure_at_node(cfsa_ns *h, int64_t node, double *out) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  if (node < 0 || node >= h->s.nv) return fail(CFSA_ERR_INVALID, "node out of range");
  return guard([&]() -> int {
    h->s.get_pressure_at(node, out);
    return CFSA_OK;
  });
}
int cfsa_ns_get_scalar(const cfsa_ns *h, int which, double *out) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  const auto &s = h->s;
  switch (which) {
    case 0: *out = s.time; break;
    case 1: *out = s.dt; break;
    case 2: *out = s.viscosity; break;
    case 3: *out = s.density; break;
    case 4: *out = s.tolerance; break;
    case 5: *out = s.outlet_pressure; break;
    default: return fail(CFSA_ERR_INVALID, "unknown scalar id");
  }
  return CFSA_OK;
}
int cfsa_ns_get_reference_pressure_node(const cfsa_ns *h, int *out) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  *out = h->s.reference_node;
  return CFSA_OK;
}
int cfsa_ns_has_nans(cfsa_ns *h, int *out) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    *out = h->s.has_nans() ? 1 : 0;
    return CFSA_OK;
  });
}
int cfsa_ns_get_step_stats(const cfsa_ns *h, cfsa_step_stats *o) {
  CHECK_PTR(h);
  CHECK_PTR(o);
  static_assert(sizeof(cfsa_step_stats) == sizeof(cfsa::StepStats), "stats layout");
  std::memcpy(o, &h->s.last, sizeof(*o));
  return CFSA_OK;
}

int cfsa_ns_export_vtk(cfsa_ns *h, const char *path, double scale) {
  CHECK_PTR(h);
  CHECK_PTR(path);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    const int64_t nv = h->s.nv, nt = h->s.nt;
    std::vector<double> coords((size_t)nv * 3), u((size_t)h->s.nu * 3), p((size_t)nv);
    std::vector<int32_t> tets((size_t)nt * 4);
    std::vector<uint8_t> bc((size_t)nv);
    h->s.get_mesh_arrays(coords.data(), tets.data(), bc.data());
    h->s.get_velocity(u.data());
    h->s.get_pressure(p.data());
    if (!cfsa::write_mesh_vtk(path, coords.data(), nv, tets.data(), nt, bc.data(), scale))
      return fail(CFSA_ERR_IO, std::string("[IO] Failed to open file: ") + path);
    if (!cfsa::append_ns_fields_vtk(path, p.data(), u.data(), nv, scale))
      return fail(CFSA_ERR_IO, std::string("[IO] Failed to open VTK file for appending: ") + path);
    return CFSA_OK;
  });
}

int cfsa_ns_get_matrix(cfsa_ns *h, int which, int32_t *rowptr, int32_t *col, double *val) {
  CHECK_PTR(h);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.get_matrix(which, rowptr, col, val);
    return CFSA_OK;
  });
}
int cfsa_ns_get_num_colours(const cfsa_ns *h, int *out) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  *out = h->s.ncolours;
  return CFSA_OK;
}
int cfsa_ns_get_colouring(const cfsa_ns *h, int32_t *colour) {
  CHECK_PTR(h);
  CHECK_PTR(colour);
  if (!h->s.initialized) return fail(CFSA_ERR_NOT_INIT, "solver not initialized");
  h->s.get_colouring(colour);
  return CFSA_OK;
}
int cfsa_ns_get_incidence(cfsa_ns *h, int32_t *ptr, int32_t *tet) {
  CHECK_PTR(h);
  CHECK_PTR(ptr);
  CHECK_PTR(tet);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.get_incidence(ptr, tet);
    return CFSA_OK;
  });
}
int cfsa_ns_compute_divergence(cfsa_ns *h, const double *u, double *out) {
  CHECK_PTR(h);
  CHECK_PTR(u);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.compute_divergence(u, out);
    return CFSA_OK;
  });
}
int cfsa_ns_apply_matrix(cfsa_ns *h, int which, const double *x, double *y, int nrhs) {
  CHECK_PTR(h);
  CHECK_PTR(x);
  CHECK_PTR(y);
  if (int e = check_init(h)) return e;
  if (which < 0 || which > 8 || (nrhs != 1 && nrhs != 3) ||
      (which >= 3 && which <= 4 && nrhs != 1) || (which == 8 && nrhs != 3))
    return fail(CFSA_ERR_INVALID, "apply_matrix: bad matrix id or nrhs");
  return guard([&]() -> int {
    h->s.apply_matrix(which, x, y, nrhs);
    return CFSA_OK;
  });
}
int cfsa_ns_apply_matrix_noalloc(cfsa_ns *h, int which, const double *x, double *y,
                                 int nrhs) {
  CHECK_PTR(h);
  CHECK_PTR(x);
  CHECK_PTR(y);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.apply_matrix_noalloc(which, x, y, nrhs);
    return CFSA_OK;
  });
}
int cfsa_ns_krylov_solve(cfsa_ns *h, int which, int method, int nrhs, const double *b,
                         double *x, double tol, int maxit, int precision,
                         cfsa_solve_result *res) {
  CHECK_PTR(h);
  CHECK_PTR(b);
  CHECK_PTR(x);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    cfsa::SolveResult r = h->s.krylov_solve(which, method, nrhs, b, x, tol, maxit, precision);
    if (res) to_result(r, res);
    return CFSA_OK;
  });
}
int cfsa_ns_amg_apply(cfsa_ns *h, const double *r, double *z, int precision) {
  CHECK_PTR(h);
  CHECK_PTR(r);
  CHECK_PTR(z);
  if (int e = check_init(h)) return e;
  if (precision != 0 && precision != 1) return fail(CFSA_ERR_INVALID, "precision 0 or 1");
  return guard([&]() -> int {
    h->s.amg_apply(r, z, precision);
    return CFSA_OK;
  });
}
int cfsa_ns_amg_info(cfsa_ns *h, int cap, int64_t *rows, int64_t *nnz, double *setup_ms,
                     int *builds, int *levels) {
  CHECK_PTR(h);
  if (cap < 0) return fail(CFSA_ERR_INVALID, "cap < 0");
  return guard([&]() -> int {
    const int L = h->s.amg_info(cap, rows, nnz, setup_ms, builds);
    if (levels) *levels = L;
    return CFSA_OK;
  });
}
int cfsa_ns_momentum_poly(cfsa_ns *h, double out[5]) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.momentum_poly(out);
    return CFSA_OK;
  });
}
int cfsa_ns_momentum_poly2(cfsa_ns *h, double out[3]) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.momentum_poly2(out);
    return CFSA_OK;
  });
}
int cfsa_ns_elem_info(cfsa_ns *h, int64_t out[4], double *setup_ms) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  return guard([&]() -> int {
    h->s.elem_info(out, setup_ms);
    return CFSA_OK;
  });
}
int cfsa_amg_build(int64_t n, const int32_t *rowptr, const int32_t *col, const double *val,
                   double theta, int max_coarse, cfsa_amg **out) {
  CHECK_PTR(rowptr);
  CHECK_PTR(out);
  *out = nullptr;
  if (n < 1 || n >= INT32_MAX) return fail(CFSA_ERR_INVALID, "matrix size out of range");
  if (!(theta >= 0) || max_coarse < 1) return fail(CFSA_ERR_INVALID, "invalid AMG parameters");
  return guard([&]() -> int {
    cfsa::HostCsr A;
    A.n = A.m = n;
    A.rp.assign(rowptr, rowptr + n + 1);
    if (A.rp[0] != 0) throw std::invalid_argument("rowptr[0] must be 0");
    for (int64_t i = 0; i < n; ++i)
      if (A.rp[(size_t)i + 1] < A.rp[(size_t)i]) throw std::invalid_argument("rowptr decreases");
    const int64_t nnz = A.rp[(size_t)n];
    if (nnz > 0 && (!col || !val)) throw std::invalid_argument("null column or value array");
    A.ci.assign(col, col + nnz);
    A.v.assign(val, val + nnz);
    for (int64_t i = 0; i < n; ++i)
      for (int32_t k = A.rp[(size_t)i]; k < A.rp[(size_t)i + 1]; ++k) {
        const int32_t j = A.ci[(size_t)k];
        if (j < 0 || j >= n || (k > A.rp[(size_t)i] && j <= A.ci[(size_t)k - 1]))
          throw std::invalid_argument("columns must be in range and ascending per row");
      }
    cfsa::AmgParams prm;
    prm.theta = theta;
    prm.max_coarse = max_coarse;
    auto h = std::make_unique<cfsa_amg>();
    h->H = cfsa::amg_build(A, prm);
    *out = h.release();
    return CFSA_OK;
  });
}
int cfsa_amg_free(cfsa_amg *h) {
  delete h;
  return CFSA_OK;
}

int cfsa_elem_patches_build(int64_t nv, int64_t nt, int64_t ne, const int32_t *tets,
                            const int32_t *ten, const double *xyz, cfsa_elem_patches **out) {
  CHECK_PTR(tets);
  CHECK_PTR(ten);
  CHECK_PTR(xyz);
  CHECK_PTR(out);
  *out = nullptr;
  if (nv < 1 || nt < 1 || ne < 0 || nv + ne >= INT32_MAX || nt * 10 >= INT32_MAX)
    return fail(CFSA_ERR_INVALID, "mesh sizes out of range");
  return guard([&]() -> int {
    for (int64_t i = 0; i < nt * 4; ++i)
      if (tets[i] < 0 || tets[i] >= nv) throw std::invalid_argument("tet vertex out of range");
    for (int64_t i = 0; i < nt * 6; ++i)
      if (ten[i] < 0 || ten[i] >= ne) throw std::invalid_argument("edge node out of range");
    auto h = std::make_unique<cfsa_elem_patches>();
    h->E = cfsa::build_elem_patches(nv + ne, nv, nt, tets, ten, xyz);
    *out = h.release();
    return CFSA_OK;
  });
}
int cfsa_elem_patches_free(cfsa_elem_patches *h) {
  delete h;
  return CFSA_OK;
}
int cfsa_elem_patches_sizes(const cfsa_elem_patches *h, int64_t out[8]) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  const cfsa::ElemPatchHost &E = h->E;
  out[0] = E.npatch;
  out[1] = E.nvisit;
  out[2] = (int64_t)E.nodes.size();
  out[3] = (int64_t)E.rows.size();
  out[4] = (int64_t)E.starts.size();
  out[5] = E.max_rows;
  out[6] = E.max_nodes;
  out[7] = E.max_slots;
  return CFSA_OK;
}
int cfsa_elem_patches_get(const cfsa_elem_patches *h, int32_t *node_off, int32_t *nodes,
                          int32_t *row_off, int32_t *rows, uint16_t *starts, int32_t *visit_off,
                          int32_t *vtet, uint32_t *vstat, int32_t *tet_order) {
  CHECK_PTR(h);
  const cfsa::ElemPatchHost &E = h->E;
  auto put = [](const auto &v, auto *o) {
    if (o) std::copy(v.begin(), v.end(), o);
  };
  put(E.node_off, node_off);
  put(E.nodes, nodes);
  put(E.row_off, row_off);
  put(E.rows, rows);
  put(E.starts, starts);
  put(E.visit_off, visit_off);
  put(E.vtet, vtet);
  put(E.vstat, vstat);
  put(E.tet_order, tet_order);
  return CFSA_OK;
}
namespace {
const cfsa::HostCsr *amg_matrix(const cfsa_amg *h, int level, int which) {
  if (level < 0 || level >= (int)h->H.levels.size()) return nullptr;
  const cfsa::AmgLevel &L = h->H.levels[(size_t)level];
  if (which == 0) return &L.A;
  if (level == (int)h->H.levels.size() - 1) return nullptr;
  return which == 1 ? &L.P : which == 2 ? &L.R : nullptr;
}
}  // namespace
int cfsa_amg_num_levels(const cfsa_amg *h, int *levels, int64_t *coarse_n, int *coarse_jacobi) {
  CHECK_PTR(h);
  if (levels) *levels = (int)h->H.levels.size();
  if (coarse_n) *coarse_n = h->H.coarse_n;
  if (coarse_jacobi) *coarse_jacobi = h->H.coarse_jacobi ? 1 : 0;
  return CFSA_OK;
}
int cfsa_amg_level_size(const cfsa_amg *h, int level, int which, int64_t *rows, int64_t *cols,
                        int64_t *nnz) {
  CHECK_PTR(h);
  const cfsa::HostCsr *M = amg_matrix(h, level, which);
  if (!M) return fail(CFSA_ERR_INVALID, "no such AMG level matrix");
  if (rows) *rows = M->n;
  if (cols) *cols = M->m;
  if (nnz) *nnz = M->nnz();
  return CFSA_OK;
}
int cfsa_amg_level_get(const cfsa_amg *h, int level, int which, int32_t *rowptr, int32_t *col,
                       double *val) {
  CHECK_PTR(h);
  const cfsa::HostCsr *M = amg_matrix(h, level, which);
  if (!M) return fail(CFSA_ERR_INVALID, "no such AMG level matrix");
  if (rowptr) std::memcpy(rowptr, M->rp.data(), M->rp.size() * sizeof(int32_t));
  if (col) std::memcpy(col, M->ci.data(), M->ci.size() * sizeof(int32_t));
  if (val) std::memcpy(val, M->v.data(), M->v.size() * sizeof(double));
  return CFSA_OK;
}
int cfsa_amg_level_params(const cfsa_amg *h, int level, double *rho, double *lmax,
                          double coef[3], int32_t *agg) {
  CHECK_PTR(h);
  if (level < 0 || level >= (int)h->H.levels.size())
    return fail(CFSA_ERR_INVALID, "no such AMG level");
  const cfsa::AmgLevel &L = h->H.levels[(size_t)level];
  if (rho) *rho = L.rho;
  if (lmax) *lmax = L.lmax;
  if (coef) {
    coef[0] = L.c0;
    coef[1] = L.c1;
    coef[2] = L.c2;
  }
  if (agg && !L.agg.empty()) std::memcpy(agg, L.agg.data(), L.agg.size() * sizeof(int32_t));
  return CFSA_OK;
}
int cfsa_amg_coarse_inverse(const cfsa_amg *h, double *inv) {
  CHECK_PTR(h);
  CHECK_PTR(inv);
  if (h->H.coarse_jacobi) return fail(CFSA_ERR_INVALID, "the coarsest level has no dense inverse");
  std::memcpy(inv, h->H.coarse_inv.data(), h->H.coarse_inv.size() * sizeof(double));
  return CFSA_OK;
}
int cfsa_ns_time_momentum_spmv(cfsa_ns *h, int reps, int precision, double *ms) {
  CHECK_PTR(h);
  CHECK_PTR(ms);
  if (int e = check_init(h)) return e;
  if (reps < 1) return fail(CFSA_ERR_INVALID, "reps < 1");
  return guard([&]() -> int {
    *ms = h->s.time_momentum_spmv(reps, precision);
    return CFSA_OK;
  });
}
int cfsa_ns_get_stream(const cfsa_ns *h, void **stream) {
  CHECK_PTR(h);
  CHECK_PTR(stream);
  *stream = (void *)h->s.stream();
  return CFSA_OK;
}

// ---------------------------------------------------------------------------
// Boundary forces
// ---------------------------------------------------------------------------
static int forces(cfsa_ns *h, int which, double *out) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    sync_mesh(h);
    h->s.compute_forces(which, out);
    return CFSA_OK;
  });
}
int cfsa_ns_compute_pressure_forces(cfsa_ns *h, double *out) { return forces(h, 0, out); }
int cfsa_ns_compute_shear_stress(cfsa_ns *h, double *out) { return forces(h, 1, out); }
int cfsa_ns_compute_total_fluid_forces(cfsa_ns *h, double *out) { return forces(h, 2, out); }
int cfsa_ns_num_faces(cfsa_ns *h, int64_t *out) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    *out = h->s.n_faces();
    return CFSA_OK;
  });
}
int cfsa_ns_get_mean_face_pressure(cfsa_ns *h, int64_t face, double *out) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.face_pressure(face, out, nullptr);
    return CFSA_OK;
  });
}
int cfsa_ns_get_pressure_force_at_face(cfsa_ns *h, int64_t face, double out[3]) {
  CHECK_PTR(h);
  CHECK_PTR(out);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    h->s.face_pressure(face, nullptr, out);
    return CFSA_OK;
  });
}

// ---------------------------------------------------------------------------
// Adjoint
// ---------------------------------------------------------------------------
int cfsa_ns_set_viscosity(cfsa_ns *h, double viscosity) {
  CHECK_PTR(h);
  if (int e = check_init(h)) return e;
  if (!(viscosity > 0)) return fail(CFSA_ERR_INVALID, "viscosity must be > 0");
  return guard([&]() -> int {
    h->s.set_viscosity(viscosity);
    return CFSA_OK;
  });
}
int cfsa_ns_adjoint_forward(cfsa_ns *h, int nsteps, const double *g, double *u_final) {
  CHECK_PTR(h);
  CHECK_PTR(g);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    sync_mesh(h);
    const int st = h->s.adjoint_forward(nsteps, g, u_final);
    if (st != CFSA_OK) g_last_error = h->s.last_error;
    return st;
  });
}
int cfsa_ns_adjoint_backward(cfsa_ns *h, const double *lambda_u, double *dLdg,
                             double *dLdviscosity) {
  CHECK_PTR(h);
  if (int e = check_init(h)) return e;
  return guard([&]() -> int {
    const int st = h->s.adjoint_backward(lambda_u, dLdg, dLdviscosity);
    if (st != CFSA_OK) g_last_error = h->s.last_error;
    return st;
  });
}
int cfsa_ns_adjoint_gradient(cfsa_ns *h, int nsteps, int nk, const double *basis,
                             const double *coef, double *J, double *dJdcoef,
                             double *dJdviscosity) {
  CHECK_PTR(h);
  CHECK_PTR(basis);
  CHECK_PTR(coef);
  CHECK_PTR(dJdcoef);
  if (int e = check_init(h)) return e;
  if (nk < 1) return fail(CFSA_ERR_INVALID, "nk >= 1");
  return guard([&]() -> int {
    sync_mesh(h);
    const int st = h->s.adjoint_gradient(nsteps, nk, basis, coef, J, dJdcoef, dJdviscosity);
    if (st != CFSA_OK) g_last_error = h->s.last_error;
    return st;
  });
}

// ---------------------------------------------------------------------------
// Configuration
// ---------------------------------------------------------------------------
static void to_c_config(const cfsa::SimulationConfig &c, cfsa_sim_config *o) {
  std::memset(o, 0, sizeof(*o));
  const auto &m = c.mesh;
  o->mesh_type = m.type;
  const std::string &file = m.type == 4 ? m.msh_file_path : m.obj_file_path;
  std::strncpy(o->mesh_file, file.c_str(), sizeof(o->mesh_file) - 1);
  o->obj_scale = m.obj_scale;
  o->cylinder_length = m.cylinder_length;
  o->cylinder_radius = m.cylinder_radius;
  o->cylinder_divisions = m.cylinder_divisions;
  o->icosphere_radius = m.icosphere_radius;
  o->icosphere_recursion = m.icosphere_recursion;
  o->max_element_size = m.max_element_size;
  for (int i = 0; i < 3; ++i) {
    o->flow_direction[i] = m.flow_direction[i];
    o->box_size[i] = m.box_size[i];
    o->box_cells[i] = m.box_cells[i];
  }
  o->inlet_outlet_coverage = m.inlet_outlet_coverage;
  const auto &n = c.navier_stokes;
  o->ns_enabled = n.enabled;
  o->density = n.density;
  o->viscosity = n.viscosity;
  o->is_newtonian = n.is_newtonian;
  o->is_laminar = n.is_laminar;
  for (int i = 0; i < 3; ++i) o->inlet_velocity[i] = n.inlet_velocity[i];
  o->outlet_pressure = n.outlet_pressure;
  o->use_dirichlet_pressure = n.use_dirichlet_pressure;
  o->dt = n.dt;
  o->relax_p = n.relax_p;
  o->relax_u = n.relax_u;
  o->n_corrections = n.n_corrections;
  o->pressure_solver = n.pressure_solver;
  o->precision = n.precision;
  o->momentum_tol = n.momentum_tol;
  o->pressure_tol = n.pressure_tol;
  o->inlet_profile = n.inlet_profile;
  const auto &out = c.output;
  std::strncpy(o->output_prefix, out.output_prefix.c_str(), sizeof(o->output_prefix) - 1);
  o->total_steps = out.total_steps;
  o->output_frequency = out.output_frequency;
  o->vtk_scale = out.vtk_scale;
  o->output_fluid = out.output_fluid;
  o->output_solid = out.output_solid;
  o->output_dpm = out.output_dpm;
  o->solid_enabled = c.solid_enabled;
  o->dpm_enabled = c.dpm_enabled;
  o->fsi_enabled = c.fsi_enabled;
}

int cfsa_config_load(const char *path, cfsa_sim_config *out) {
  CHECK_PTR(path);
  CHECK_PTR(out);
  try {
    to_c_config(cfsa::load_config(path), out);
    return CFSA_OK;
  } catch (const std::exception &e) {
    return fail(std::string(e.what()).find("bad file") != std::string::npos ? CFSA_ERR_IO
                                                                             : CFSA_ERR_INVALID,
                e.what());
  }
}

int cfsa_config_parse(const char *text, cfsa_sim_config *out) {
  CHECK_PTR(text);
  CHECK_PTR(out);
  try {
    to_c_config(cfsa::parse_config(text), out);
    return CFSA_OK;
  } catch (const std::exception &e) {
    return fail(CFSA_ERR_INVALID, e.what());
  }
}

}  // extern "C"
