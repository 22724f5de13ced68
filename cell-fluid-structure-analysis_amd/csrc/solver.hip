// PISO orchestration on the GPU.  Every device computation of the step runs
// here through hand-written HIP kernels; the host only sequences launches and
// reads back Krylov convergence flags every few iterations.
#include "solver.hpp"

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>

#include "amg.hpp"
#include "dd.hpp"
#include "mesh.hpp"
#include "pattern.hpp"

namespace cfsa {

namespace {

__global__ void k_set_params(KrylovState *st, double tol, int maxit, int nrhs,
                             double restart_scale) {
  st->tol = tol;
  st->restart_scale = restart_scale;
  st->maxit = maxit;
  st->nrhs = nrhs;
}
__global__ void k_set_mask(int *m, int a, int b, int c) {
  m[0] = a;
  m[1] = b;
  m[2] = c;
}
// r32 = masked ? float(r64) : 0, interleaved [n][K]
__global__ void k_demote_masked(int n, int K, const double *__restrict__ in,
                                const int *__restrict__ mask, float *__restrict__ out) {
  const int64_t N = (int64_t)n * K;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * kBlock)
    out[i] = mask[i % K] ? (float)in[i] : 0.0f;
}
__global__ void k_accumulate(int64_t N, const float *__restrict__ d, double *__restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * kBlock)
    x[i] = x[i] + (double)d[i];
}
// The element patches number the velocity nodes in patch order (node j of
// that order is node perm[j]); the FP32 inner solves on them run in that order.
// out[j][c] = masked ? float(in[perm[j]][c]) : 0
__global__ void k_demote_perm(int n, const double *__restrict__ in, const int *__restrict__ mask,
                              const int *__restrict__ perm, float *__restrict__ out) {
  const int64_t N = (int64_t)n * 3;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * kBlock) {
    const int c = (int)(i % 3);
    out[i] = mask[c] ? (float)in[(int64_t)perm[i / 3] * 3 + c] : 0.0f;
  }
}
// x[perm[j]][c] += d[j][c]
__global__ void k_accumulate_perm(int n, const float *__restrict__ d, const int *__restrict__ perm,
                                  double *__restrict__ x) {
  const int64_t N = (int64_t)n * 3;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * kBlock) {
    const int64_t o = (int64_t)perm[i / 3] * 3 + i % 3;
    x[o] = x[o] + (double)d[i];
  }
}
// out[j] = in[perm[j]] (K = 1), or its inverse with inverse = true
template <class T>
__global__ void k_permute(int n, const T *__restrict__ in, const int *__restrict__ perm,
                          bool inverse, T *__restrict__ out) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if (inverse)
      out[perm[i]] = in[i];
    else
      out[i] = in[perm[i]];
  }
}
constexpr double kMomentumPenalty = 1e10;  // NavierStokes.cpp:463
constexpr double kPressurePenalty = 1e12;  // NavierStokes.cpp:611,619

}  // namespace

NSSolver::NSSolver() {
  CFSA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  CFSA_HIP(hipHostMalloc((void **)&h_state, sizeof(KrylovState), hipHostMallocDefault));
  std::memset(h_state, 0, sizeof(KrylovState));
  d_state.alloc(1);
  d_part_a.alloc((size_t)12 * kMaxPartials);
  d_part_b.alloc((size_t)12 * kMaxPartials);
  d_part_c.alloc((size_t)12 * kMaxPartials);
  d_part_r.alloc((size_t)12 * kMaxPartials);
  d_mask.alloc(3);
  d_flag.alloc(1);
  d_scalar.alloc(8);
  scratch_ = noalloc_scratch_current_device();
}

NSSolver::~NSSolver() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (copy_stream_) {
    (void)hipStreamSynchronize(copy_stream_);
    (void)hipStreamDestroy(copy_stream_);
  }
  if (ev_snap_) (void)hipEventDestroy(ev_snap_);
  if (ev_copied_) (void)hipEventDestroy(ev_copied_);
  clear_graphs();
  for (auto e : ev_pool_) (void)hipEventDestroy(e);
  if (h_state) (void)hipHostFree(h_state);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// ---------------------------------------------------------------------------
// Setup: reference NavierStokesSolver::initialize (NavierStokes.cpp:10-68)
// ---------------------------------------------------------------------------
void NSSolver::initialize(const MeshInput &m, double rho, double mu) {
  fields_wait();  // a pending host copy reads the snapshot buffers
  clear_graphs();
  amg_built_ = false;  // a hierarchy belongs to the previous mesh's operator
  amg_top_fresh_ = false;
  elem_built_ = elem_fresh_ = false;  // so do the element patches
  amg_lbc_changed_ = true;
  amg_lv_.clear();
  amg_view_.clear();
  if (m.nv <= 0 || m.nt <= 0 || m.ne < 0 || !m.coords || !m.tets || !m.ten || !m.p1_bc ||
      (m.ne > 0 && !m.p2_bc))
    throw std::invalid_argument("initialize: incomplete mesh arrays");
  if (m.nv + m.ne >= INT32_MAX || m.nt * 10 >= INT32_MAX)
    throw std::invalid_argument("initialize: mesh exceeds int32 indexing");
  nv = (int)m.nv;
  nt = (int)m.nt;
  ne = (int)m.ne;
  nu = nv + ne;
  density = rho;
  viscosity = mu;
  h_coords.assign(m.coords, m.coords + (size_t)nv * 3);
  h_tets.assign(m.tets, m.tets + (size_t)nt * 4);
  h_ten.assign(m.ten, m.ten + (size_t)nt * 6);
  for (size_t i = 0; i < h_tets.size(); ++i)
    if (h_tets[i] < 0 || h_tets[i] >= nv) throw std::invalid_argument("tet vertex out of range");
  for (size_t i = 0; i < h_ten.size(); ++i)
    if (h_ten[i] < 0 || h_ten[i] >= ne) throw std::invalid_argument("edge node out of range");
  h_bc.resize((size_t)nu);
  std::memcpy(h_bc.data(), m.p1_bc, (size_t)nv);
  if (ne) std::memcpy(h_bc.data() + nv, m.p2_bc, (size_t)ne);

  std::vector<double> vol((size_t)nt), grad((size_t)nt * 12);
  Mesh::tetGeometry(h_coords.data(), h_tets.data(), nt, vol.data(), grad.data());

  std::vector<int32_t> elem10((size_t)nt * 10);
  for (int t = 0; t < nt; ++t) {
    for (int a = 0; a < 4; ++a) elem10[(size_t)t * 10 + a] = h_tets[(size_t)t * 4 + a];
    for (int e = 0; e < 6; ++e) elem10[(size_t)t * 10 + 4 + e] = nv + h_ten[(size_t)t * 6 + e];
  }
  Incidence inc = build_incidence(elem10, 10, nu);
  CsrPattern PA = build_pattern(elem10, 10, inc.ptr.data(), inc.tet.data(), nu);
  std::vector<int32_t>().swap(elem10);
  CsrPattern PP = build_pattern(h_tets, 4, inc.ptr.data(), inc.tet.data(), nv);
  for (int i = 0; i < nu; ++i)
    if (PA.diag[(size_t)i] < 0) throw std::runtime_error("velocity node without element");
  Colouring C = colour_tets(h_tets.data(), nt, nv);
  ncolours = C.ncolours();
  h_colour_offsets = C.offsets;
  h_order = C.order;

  d_tets.upload(h_tets.data(), h_tets.size());
  d_ten.upload(h_ten.data(), h_ten.size());
  d_vol.upload(vol.data(), vol.size());
  d_grad.upload(grad.data(), grad.size());
  d_bc.upload(h_bc.data(), h_bc.size());
  d_order.upload(C.order.data(), C.order.size());
  // boundary faces: the mesh's own list when given, else built at first use
  faces_ready_ = false;
  face_mesh_ = Mesh();
  face_topology_ = m.faces != nullptr;
  if (face_topology_) {
    face_mesh_.faces = *m.faces;
    face_mesh_.tets = h_tets;
  }
  d_amap.release();  // the row-gather slot map belongs to the previous patterns
  d_pmap.release();
  d_loc.release();
  d_incp.upload(inc.ptr.data(), inc.ptr.size());
  d_inct.upload(inc.tet.data(), inc.tet.size());
  d_Arp.upload(PA.rowptr.data(), PA.rowptr.size());
  d_Acol.upload(PA.col.data(), PA.col.size());
  d_Adiag.upload(PA.diag.data(), PA.diag.size());
  d_Prp.upload(PP.rowptr.data(), PP.rowptr.size());
  d_Pcol.upload(PP.col.data(), PP.col.size());
  d_Pdiag.upload(PP.diag.data(), PP.diag.size());
  {
    // the P1 pattern is symmetric: map every slot (i, j) to the slot (j, i)
    std::vector<int32_t> tslot((size_t)PP.nnz());
    for (int i = 0; i < nv; ++i)
      for (int32_t k = PP.rowptr[(size_t)i]; k < PP.rowptr[(size_t)i + 1]; ++k) {
        const int32_t j = PP.col[(size_t)k];
        const int32_t *lo = PP.col.data() + PP.rowptr[(size_t)j];
        const int32_t *hi = PP.col.data() + PP.rowptr[(size_t)j + 1];
        const int32_t *it = std::lower_bound(lo, hi, i);
        if (it == hi || *it != i) throw std::runtime_error("P1 pattern is not symmetric");
        tslot[(size_t)k] = (int32_t)(it - PP.col.data());
      }
    d_Ptslot.upload(tslot.data(), tslot.size());
  }
  nnzA = PA.nnz();
  nnzP = PP.nnz();
  maxrowA_ = maxrowP_ = 0;
  for (int64_t i = 0; i < nu; ++i)
    maxrowA_ = std::max(maxrowA_, PA.rowptr[(size_t)i + 1] - PA.rowptr[(size_t)i]);
  for (int64_t i = 0; i < nv; ++i)
    maxrowP_ = std::max(maxrowP_, PP.rowptr[(size_t)i + 1] - PP.rowptr[(size_t)i]);
  A_ = DevCsr{nu, nnzA, d_Arp.p, d_Acol.p};
  P_ = DevCsr{nv, nnzP, d_Prp.p, d_Pcol.p};
  {
    const std::vector<int32_t> a1 = stream_windows(PA, kStreamB<double, 1>, kStreamRows);
    const std::vector<int32_t> a3 = stream_windows(PA, kStreamB<double, 3>, kStreamRows);
    const std::vector<int32_t> a3f = stream_windows(PA, kStreamB<float, 3>, kStreamRows);
    const std::vector<int32_t> p1 = stream_windows(PP, kStreamB<double, 1>, kStreamRows);
    const std::vector<int32_t> p3 = stream_windows(PP, kStreamB<double, 3>, kStreamRows);
    const std::vector<int32_t> p3f = stream_windows(PP, kStreamB<float, 3>, kStreamRows);
    d_Aw1.upload(a1.data(), a1.size());
    d_Aw3.upload(a3.data(), a3.size());
    d_Aw3f.upload(a3f.data(), a3f.size());
    d_Pw1.upload(p1.data(), p1.size());
    d_Pw3.upload(p3.data(), p3.size());
    d_Pw3f.upload(p3f.data(), p3f.size());
    A_.win1 = d_Aw1.p;
    A_.win3 = d_Aw3.p;
    A_.win3f = d_Aw3f.p;
    A_.nwin1 = (int)a1.size() - 1;
    A_.nwin3 = (int)a3.size() - 1;
    A_.nwin3f = (int)a3f.size() - 1;
    P_.win1 = d_Pw1.p;
    P_.win3 = d_Pw3.p;
    P_.win3f = d_Pw3f.p;
    P_.nwin1 = (int)p1.size() - 1;
    P_.nwin3 = (int)p3.size() - 1;
    P_.nwin3f = (int)p3f.size() - 1;
    const int64_t cap = std::max({(int64_t)2 * A_.nwin1, (int64_t)6 * A_.nwin3,
                                  (int64_t)2 * P_.nwin1, (int64_t)6 * P_.nwin3});
    d_wpart.alloc((size_t)(2 * cap));
    for (DevCsr *X : {&A_, &P_}) {
      X->wpart = d_wpart.p;
      X->wpart_cap = cap;
    }
  }
  // the momentum operator's sliced-ELL copy (spmv_layout 2) is built when
  // first used (prepare_operators)
  for (DBuf<int> *b : {&d_Asptr, &d_Arlen, &d_Aperm, &d_Ascol, &d_Asrc}) b->release();
  d_lhsS.release();
  storedA = 0;
  sellA_built_ = false;
  sellA_ = -1;
  {
    SellPattern SP = build_sell(PP);
    d_Psptr.upload(SP.sptr.data(), SP.sptr.size());
    d_Prlen.upload(SP.rlen.data(), SP.rlen.size());
    d_Pperm.upload(SP.perm.data(), SP.perm.size());
    d_Pscol.upload(SP.col.data(), SP.col.size());
    d_Psrc.upload(SP.src.data(), SP.src.size());
    storedP = SP.stored();
    Ps_ = DevCsr{nv, nnzP, nullptr, d_Pscol.p, d_Psptr.p, d_Prlen.p, d_Pperm.p, SP.nslices};
  }

  d_M.alloc((size_t)nnzA);
  d_K.alloc((size_t)nnzA);
  d_lhs.alloc((size_t)nnzA);
  d_L.alloc((size_t)nnzP);
  d_Lbc.alloc((size_t)nnzP);
  d_G3.alloc((size_t)nnzP * 3);
  d_ilm.alloc((size_t)nu);
  d_invdA.alloc((size_t)nu);
  d_invdL.alloc((size_t)nv);
  d_u.alloc((size_t)nu * 3);
  d_us.alloc((size_t)nu * 3);
  d_p.alloc((size_t)nv);
  d_pc.alloc((size_t)nv);
  d_ginlet.alloc((size_t)nu * 3);
  d_conv.alloc((size_t)nu * 3);
  d_cT.alloc((size_t)nt * 3);
  d_wT.alloc((size_t)nt);
  d_div.alloc((size_t)nv);
  d_rhs.alloc((size_t)nu * 3);
  d_rhsp.alloc((size_t)nv);
  d_gp.alloc((size_t)nv * 3);
  d_xref.alloc((size_t)nu * 3);
  w3_.alloc((size_t)nu * 3);
  w1_.alloc((size_t)nv);
  d_r64.alloc((size_t)nu * 3);

  assemble();
  d_u.zero(stream_);
  d_us.zero(stream_);
  d_p.zero(stream_);
  d_pc.zero(stream_);
  upload_inlet();
  set_inlet_nodes(stream_, nu, d_bc.p, d_ginlet.p, d_u.p);
  // Neumann outlets pin the first Internal vertex (NavierStokes.cpp:51-67)
  if (reference_node < 0 && outlet_type == 1) {
    for (int i = 0; i < nv; ++i)
      if (h_bc[(size_t)i] == kInternal) {
        reference_node = i;
        break;
      }
    if (reference_node < 0 && nv > 0) {
      reference_node = 0;
      std::cout << "[NS] Warning: No internal node found, using vertex 0 as pressure reference"
                << std::endl;
    }
  }
  lhs_dirty = lbc_dirty = true;
  initialized = true;
  // allocate and fill the Krylov operators' layout copies now, so later
  // entry points (the allocation-free one included) find them in place
  prepare_operators();
  CFSA_HIP(hipStreamSynchronize(stream_));
}

void NSSolver::update_mesh(const double *coords, const uint8_t *p1_bc, const uint8_t *p2_bc) {
  if (!initialized) throw std::runtime_error("update_mesh before initialize");
  if (coords) {
    // moved vertices: re-measure the tets on the GPU (identical arithmetic to
    // the host mesh layer)
    h_coords.assign(coords, coords + (size_t)nv * 3);
    d_xyz.upload(h_coords.data(), h_coords.size());
    tet_geometry(stream_, nt, d_xyz.p, d_tets.p, d_vol.p, d_grad.p);
  }
  if (p1_bc) std::memcpy(h_bc.data(), p1_bc, (size_t)nv);
  if (p2_bc && ne) std::memcpy(h_bc.data() + nv, p2_bc, (size_t)ne);
  if (p1_bc || p2_bc) d_bc.upload(h_bc.data(), h_bc.size());
  if (coords) faces_ready_ = false;
  lhs_dirty = lbc_dirty = true;
}

// Assembly of M, K (P2 clique pattern) and L, Gx/y/z (P1 pattern) into the fixed
// CSR patterns; reference build order G, M, L, K (NavierStokes.hpp:97-102).
// assembly 0 (default): row gather through the slot map, six rows per wavefront
// with LDS accumulators, bit-exact; 2: row gather with one thread per row,
// bit-exact; 1: graph-coloured, atomic-free element scatter (colour-order sums).
void NSSolver::assemble() {
  const double inv_mu = 1.0 / viscosity;  // NavierStokes.cpp:153
  if (opt.assembly != 1 && d_amap.n == 0) {
    // slot map of the row gather, built once (the patterns never change)
    const size_t ninc = (size_t)nt * 10, nvinc = (size_t)nt * 4;
    d_loc.alloc(ninc);
    d_amap.alloc(ninc * 10);
    d_pmap.alloc(nvinc * 4);
    build_slot_map(stream_, nu, d_incp.p, d_inct.p, d_tets.p, d_ten.p, nv, A_, P_, d_loc.p,
                   d_amap.p, d_pmap.p);
  }
  // assembly 0: six rows per wavefront with LDS accumulators, unless the longest
  // row does not fit; 2 (and that fallback): one thread per row
  if (opt.assembly == 0 &&
      assemble_waves(stream_, nu, nv, maxrowA_, maxrowP_, d_incp.p, d_inct.p, d_loc.p, d_amap.p,
                     d_pmap.p, A_, P_, d_vol.p, d_grad.p, density, inv_mu, d_M.p, d_K.p, d_L.p,
                     d_G3.p, d_ilm.p)) {
    // every slot and every ilm entry written
    lhs_dirty = lbc_dirty = true;
    g3t_dirty = true;
    return;
  }
  d_M.zero(stream_);
  d_K.zero(stream_);
  d_L.zero(stream_);
  d_G3.zero(stream_);
  if (opt.assembly != 1) {
    assemble_mapped(stream_, nu, nv, d_incp.p, d_inct.p, d_loc.p, d_amap.p, d_pmap.p, d_vol.p,
                    d_grad.p, density, inv_mu, d_M.p, d_K.p, d_L.p, d_G3.p);
  } else {
    for (int c = 0; c < ncolours; ++c) {
      const int a = h_colour_offsets[(size_t)c], b = h_colour_offsets[(size_t)c + 1];
      assemble_colour(stream_, b - a, d_order.p + a, d_tets.p, d_ten.p, nv, d_vol.p,
                      d_grad.p, density, inv_mu, A_, d_M.p, d_K.p, P_, d_L.p, d_G3.p);
    }
  }
  lumped_mass_inverse(stream_, A_, d_M.p, d_ilm.p);
  lhs_dirty = lbc_dirty = true;
  g3t_dirty = true;
}

void NSSolver::build_sparse_matrices() {
  if (!initialized) throw std::runtime_error("build_sparse_matrices before initialize");
  assemble();
  CFSA_HIP(hipStreamSynchronize(stream_));
}

// length-sorted sliced-ELL copy of the momentum pattern (spmv_layout 2), built
// from the device pattern when first used
void NSSolver::ensure_sell_A() {
  if (sellA_built_) return;
  CsrPattern PA;
  PA.n = nu;
  PA.rowptr.resize((size_t)nu + 1);
  PA.col.resize((size_t)nnzA);
  d_Arp.download(PA.rowptr.data(), PA.rowptr.size());
  d_Acol.download(PA.col.data(), PA.col.size());
  SellPattern SA = build_sell(PA);
  d_Asptr.upload(SA.sptr.data(), SA.sptr.size());
  d_Arlen.upload(SA.rlen.data(), SA.rlen.size());
  d_Aperm.upload(SA.perm.data(), SA.perm.size());
  d_Ascol.upload(SA.col.data(), SA.col.size());
  d_Asrc.upload(SA.src.data(), SA.src.size());
  storedA = SA.stored();
  As_ = DevCsr{nu, nnzA, nullptr, d_Ascol.p, d_Asptr.p, d_Arlen.p, d_Aperm.p, SA.nslices};
  sellA_built_ = true;
}

void NSSolver::ensure_mixed_buffers() {
  const size_t na = (size_t)std::max<int64_t>(storedA, nnzA);
  const size_t np = (size_t)std::max<int64_t>(storedP, nnzP);
  if (d_lhs32.n == na && d_Lbc32.n == np) return;
  d_lhs32.alloc(na);
  d_Lbc32.alloc(np);
  d_invdA32.alloc((size_t)nu);
  d_invdL32.alloc((size_t)nv);
  f3_.alloc((size_t)nu * 3);
  f1_.alloc((size_t)nv);
  d_b32.alloc((size_t)nu * 3);
  d_x32.alloc((size_t)nu * 3);
  lhs_dirty = lbc_dirty = true;
}

// lhs = M/dt + K with 1e10 on Inlet/Wall diagonals (NavierStokes.cpp:460-495);
// L_bc = L with 1e12 on Outlet (or reference-node) diagonals (:603-626).
void NSSolver::prepare_operators() {
  // Krylov SpMV layout: the momentum operator streams best as CSR-stream (its
  // gathers of neighbouring columns share cache lines), the P1 pressure
  // operator as sliced-ELL; option spmv_layout forces one of them.
  const int la = opt.spmv_layout == 2 ? 1 : 0;
  const int lp = opt.spmv_layout == 1 ? 0 : 1;
  if (la) ensure_sell_A();
  // FP32 copies, once they exist, are kept in step with every layout the
  // operators are filled into (the sliced-ELL copy stores padded entries, more
  // than the CSR slots the FP32 arrays were first sized for)
  if (opt.precision != 0 || d_lhs32.n > 0) ensure_mixed_buffers();
  if (la != sellA_) lhs_dirty = true;
  if (lp != sellP_) lbc_dirty = true;
  const bool mixed = d_lhs32.n > 0;
  if (lhs_dirty || lhs_dt != dt) {
    build_lhs(stream_, nnzA, d_M.p, d_K.p, dt, d_lhs.p);
    add_diag_penalty(stream_, nu, d_Adiag.p, d_bc.p, 0, -1, kMomentumPenalty, d_lhs.p);
    inv_diag(stream_, nu, d_Adiag.p, d_lhs.p, d_invdA.p);
    if (la) {
      if (d_lhsS.n != (size_t)storedA) d_lhsS.alloc((size_t)storedA);
      sell_fill<double>(stream_, storedA, d_Asrc.p, d_lhs.p, d_lhsS.p);
    }
    elem_fresh_ = false;  // the element patches' geometry and penalties follow the lhs
    mom_poly_fresh_ = false;  // and so does the polynomial preconditioner
    if (mixed) {
      if (d_lhs32.n < (size_t)(la ? storedA : nnzA))
        throw std::logic_error("FP32 momentum operator smaller than its layout");
      if (la)
        sell_fill<float>(stream_, storedA, d_Asrc.p, d_lhs.p, d_lhs32.p);
      else
        convert<float, double>(stream_, nnzA, d_lhs.p, d_lhs32.p);
      convert<float, double>(stream_, nu, d_invdA.p, d_invdA32.p);
    }
    sellA_ = la;
    lhs_dirty = false;
    lhs_dt = dt;
  }
  if (mixed && opt.elem_spmv && !elem_fresh_) {
    ensure_elem();
    elem_refresh(stream_, nt, elem_nrows_, d_etorder.p, d_erows.p, d_vol.p, d_grad.p,
                 density / dt, 1.0 / viscosity, d_bc.p, kMomentumPenalty, d_evgeo.p, d_epen.p);
    hipLaunchKernelGGL((k_permute<float>), dim3(1024), dim3(kBlock), 0, stream_, nu,
                       d_invdA32.p, d_erows.p, false, d_einvd32.p);
    CFSA_LAUNCH_CHECK();
    elem_fresh_ = true;
  }
  if (lbc_dirty) {
    CFSA_HIP(hipMemcpyAsync(d_Lbc.p, d_L.p, (size_t)nnzP * sizeof(double),
                            hipMemcpyDeviceToDevice, stream_));
    if (outlet_type == 0)
      add_diag_penalty(stream_, nv, d_Pdiag.p, d_bc.p, 1, -1, kPressurePenalty, d_Lbc.p);
    else if (outlet_type == 1 && reference_node >= 0 && reference_node < nv)
      add_diag_penalty(stream_, nv, d_Pdiag.p, d_bc.p, 2, reference_node, kPressurePenalty,
                       d_Lbc.p);
    inv_diag(stream_, nv, d_Pdiag.p, d_Lbc.p, d_invdL.p);
    if (lp) {
      if (d_LbcS.n != (size_t)storedP) d_LbcS.alloc((size_t)storedP);
      sell_fill<double>(stream_, storedP, d_Psrc.p, d_Lbc.p, d_LbcS.p);
    }
    if (mixed) {
      if (d_Lbc32.n < (size_t)(lp ? storedP : nnzP))
        throw std::logic_error("FP32 pressure operator smaller than its layout");
      if (lp)
        sell_fill<float>(stream_, storedP, d_Psrc.p, d_Lbc.p, d_Lbc32.p);
      else
        convert<float, double>(stream_, nnzP, d_Lbc.p, d_Lbc32.p);
      convert<float, double>(stream_, nv, d_invdL.p, d_invdL32.p);
    }
    sellP_ = lp;
    lbc_dirty = false;
    amg_lbc_changed_ = true;
  }
}

void NSSolver::upload_inlet() {
  if (!inlet_profile) fill3(stream_, nu, mean_inlet, d_ginlet.p);
}

void NSSolver::set_dt(double v) { dt = v; }
void NSSolver::set_mean_inlet(const double v[3]) {
  mean_inlet[0] = v[0];
  mean_inlet[1] = v[1];
  mean_inlet[2] = v[2];
  inlet_profile = false;
  if (initialized) {
    upload_inlet();
    CFSA_HIP(hipStreamSynchronize(stream_));
  }
}
void NSSolver::set_inlet_profile(const double *g) {
  if (!initialized) throw std::runtime_error("set_inlet_profile needs an initialized solver");
  if (!g) {
    inlet_profile = false;
    upload_inlet();
    CFSA_HIP(hipStreamSynchronize(stream_));
    return;
  }
  inlet_profile = true;
  CFSA_HIP(hipMemcpy(d_ginlet.p, g, (size_t)nu * 3 * sizeof(double), hipMemcpyHostToDevice));
}
void NSSolver::set_outlet_type(int t) {
  outlet_type = t;
  lbc_dirty = true;
}
void NSSolver::set_reference_node(int r) {
  reference_node = r;
  lbc_dirty = true;
}

// ---------------------------------------------------------------------------
// Krylov drivers
// ---------------------------------------------------------------------------
void NSSolver::set_state_params(double tol, int maxit, int nrhs) {
  hipLaunchKernelGGL(k_set_params, dim3(1), dim3(1), 0, stream_, d_state.p, tol, maxit, nrhs,
                     opt.restart_scale);
  CFSA_LAUNCH_CHECK();
}
void NSSolver::set_mask(int a, int b, int c) {
  hipLaunchKernelGGL(k_set_mask, dim3(1), dim3(1), 0, stream_, d_mask.p, a, b, c);
  CFSA_LAUNCH_CHECK();
}
void NSSolver::read_state() {
  CFSA_HIP(hipMemcpyAsync(h_state, d_state.p, sizeof(KrylovState), hipMemcpyDeviceToHost,
                          stream_));
  CFSA_HIP(hipStreamSynchronize(stream_));
}

void NSSolver::spmv_event_begin() {
  if (opt.timing < 2) return;
  while ((int)ev_pool_.size() < ev_used_ + 2) {
    hipEvent_t e;
    CFSA_HIP(hipEventCreate(&e));
    ev_pool_.push_back(e);
  }
  CFSA_HIP(hipEventRecord(ev_pool_[(size_t)ev_used_], stream_));
}
void NSSolver::spmv_event_end() {
  if (opt.timing < 2) return;
  CFSA_HIP(hipEventRecord(ev_pool_[(size_t)ev_used_ + 1], stream_));
  ev_used_ += 2;
}

void NSSolver::clear_graphs() {
  for (auto &g : graphs_)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  graphs_.clear();
}

// Runs one chunk of Krylov iterations: replays the graph captured for this key
// (all pointers, sizes and grids the chunk's launches use), capturing it on
// first use.  The chunk holds only kernel launches (no memset or copy nodes).
template <class F>
void NSSolver::run_chunk(const std::vector<uintptr_t> &key, F &&launch) {
  if (!opt.graphs) {
    launch();
    return;
  }
  for (auto &g : graphs_)
    if (g.key == key) {
      CFSA_HIP(hipGraphLaunch(g.exec, stream_));
      return;
    }
  if (graphs_.size() >= 16) clear_graphs();
  hipGraph_t graph = nullptr;
  CFSA_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  try {
    launch();
  } catch (...) {
    (void)hipStreamEndCapture(stream_, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  CFSA_HIP(hipStreamEndCapture(stream_, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  CFSA_HIP(e);
  graphs_.push_back({key, exec});
  CFSA_HIP(hipGraphLaunch(exec, stream_));
}

template <class T, int K>
SolveResult NSSolver::bicgstab(const DevCsr &A, const T *val, const T *invd, const T *b,
                               T *x, double tol, int maxit, KrylovWork<T> &w) {
  hipStream_t s = stream_;
  const int n = A.n;
  KrylovState *st = d_state.p;
  const int *act = st->active;
  const bool time_spmv = (K == 3) && opt.timing >= 2;
  set_state_params(tol, maxit, K);
  int Gv = 0, Gb = 0, Gc = 0, Gr = 0;
  const T *sc = opt.stop_scaled ? invd : nullptr;
  bicg_init<T, K>(s, n, b, x, w.r.p, w.r0.p, w.p.p, w.v.p, sc, d_part_a.p, &Gv);
  bicg_s_init<K>(s, st, d_part_a.p, Gv);
  const int chunk = std::max(1, K == 1 ? opt.check_every_1 : opt.check_every);
  const long max_chunks = (2L * maxit + 16) / chunk + 16;
  // SpMV launches are timed (per-SpMV events) on the second iteration of each
  // chunk only (the first follows the host's state read); chunks run as captured graphs unless they are timed
  auto body = [&](bool timed) {
    for (int j = 0; j < chunk; ++j) {
      const bool ev = timed && j == std::min(1, chunk - 1);
      bicg_s_begin<K>(s, st, d_part_a.p, Gv);
      bicg_v1<T, K>(s, n, st, w.r.p, w.p.p, w.v.p, invd, w.y.p);
      if (ev) spmv_event_begin();
      if (K == 3)
        spmv3_dot<T>(s, A, val, w.y.p, w.v.p, w.r0.p, act, d_part_b.p, &Gb);
      else
        spmv1_dot<T>(s, A, val, w.y.p, w.v.p, w.r0.p, act, d_part_b.p, &Gb);
      if (ev) spmv_event_end();
      bicg_s_alpha<K>(s, st, d_part_b.p, Gb);
      bicg_v2<T, K>(s, n, st, w.r.p, w.v.p, invd, w.s.p, w.z.p);
      if (ev) spmv_event_begin();
      if (K == 3)
        spmv3_dot2<T>(s, A, val, w.z.p, w.t.p, w.s.p, act, d_part_c.p, &Gc);
      else
        spmv1_dot2<T>(s, A, val, w.z.p, w.t.p, w.s.p, act, d_part_c.p, &Gc);
      if (ev) spmv_event_end();
      bicg_s_omega<K>(s, st, d_part_c.p, Gc);
      bicg_v3<T, K>(s, n, st, w.y.p, w.z.p, w.s.p, w.t.p, w.r0.p, x, w.r.p, sc, d_part_a.p,
                    &Gv);
    }
  };
  const std::vector<uintptr_t> key = {
      1, sizeof(T), K, (uintptr_t)n, (uintptr_t)chunk, (uintptr_t)A.rowptr,
      (uintptr_t)A.col, (uintptr_t)A.sptr, (uintptr_t)A.elem, (uintptr_t)val, (uintptr_t)invd,
      (uintptr_t)x,
      (uintptr_t)w.r.p, (uintptr_t)w.r0.p, (uintptr_t)w.p.p, (uintptr_t)w.v.p,
      (uintptr_t)w.y.p, (uintptr_t)w.s.p, (uintptr_t)w.z.p, (uintptr_t)w.t.p, (uintptr_t)sc,
      (uintptr_t)st, (uintptr_t)d_part_a.p, (uintptr_t)d_part_b.p, (uintptr_t)d_part_c.p,
      (uintptr_t)Gv};
  for (long ch = 0;; ++ch) {
    if (ch > max_chunks) throw std::runtime_error("BiCGSTAB driver did not terminate");
    if (time_spmv)
      body(true);
    else
      run_chunk(key, [&] { body(false); });
    read_state();
    bool all_done = true;
    int rs[3] = {0, 0, 0};
    for (int c = 0; c < K; ++c) {
      if (h_state->status[c] == KS_RESTART) rs[c] = 1;
      if (h_state->status[c] != KS_DONE) all_done = false;
    }
    if (all_done) break;
    if (rs[0] | rs[1] | rs[2]) {
      // r = b - A x; r0 = r; rho = r0_sqnorm = |r|^2 (Eigen restart branch)
      set_mask(rs[0], rs[1], rs[2]);
      if (K == 3)
        spmv3_residual<T>(s, A, val, x, b, w.r.p, w.r0.p, nullptr, d_mask.p, d_part_r.p, &Gr);
      else
        spmv1_residual<T>(s, A, val, x, b, w.r.p, w.r0.p, nullptr, d_mask.p, d_part_r.p, &Gr);
      bicg_s_restart<K>(s, st, d_part_r.p, Gr, d_mask.p);
    }
  }
  SolveResult res;
  for (int c = 0; c < K; ++c) {
    res.iters[c] = h_state->iters[c];
    res.err[c] = h_state->err[c];
    res.restarts[c] = h_state->restarts[c];
    if (!(h_state->err[c] <= tol)) res.ok = false;
  }
  return res;
}

template <class T, int K>
SolveResult NSSolver::pcg(const DevCsr &A, const T *val, const T *invd, const T *b, T *x,
                          double tol, int maxit, KrylovWork<T> &w) {
  hipStream_t s = stream_;
  const int n = A.n;
  KrylovState *st = d_state.p;
  const int *act = st->active;
  const bool time_spmv = (K == 3) && opt.timing >= 2;
  set_state_params(tol, maxit, K);
  int Gv = 0, Gb = 0;
  cg_init<T, K>(s, n, b, invd, x, w.r.p, w.p.p, opt.stop_scaled, d_part_a.p, &Gv);
  cg_s_init<K>(s, st, d_part_a.p, Gv);
  // even, so that every chunk starts on parity slot 0
  int chunk = std::max(1, K == 1 ? opt.check_every_1 : opt.check_every);
  chunk += chunk & 1;
  const long max_chunks = (long)maxit / chunk + 16;
  auto body = [&](bool timed) {
    for (int j = 0; j < chunk; ++j) {
      const bool ev = timed && j == std::min(1, chunk - 1);
      if (ev) spmv_event_begin();
      if (K == 3)
        spmv3_dot<T>(s, A, val, w.p.p, w.v.p, w.p.p, act, d_part_b.p, &Gb);
      else
        spmv1_dot<T>(s, A, val, w.p.p, w.v.p, w.p.p, act, d_part_b.p, &Gb);
      if (ev) spmv_event_end();
      const int par = j & 1;
      cg_xr<T, K>(s, n, st, d_part_b.p, Gb, par, w.v.p, invd, w.r.p, opt.stop_scaled, d_part_a.p,
                  &Gv);
      cg_p<T, K>(s, n, st, d_part_a.p, Gv, par, w.r.p, invd, w.p.p, x);
    }
  };
  const std::vector<uintptr_t> key = {
      2, sizeof(T), K, (uintptr_t)n, (uintptr_t)chunk, (uintptr_t)A.rowptr,
      (uintptr_t)A.col, (uintptr_t)A.sptr, (uintptr_t)A.elem, (uintptr_t)val, (uintptr_t)invd,
      (uintptr_t)x,
      (uintptr_t)w.r.p, (uintptr_t)w.p.p, (uintptr_t)w.v.p,
      (uintptr_t)st, (uintptr_t)d_part_a.p, (uintptr_t)d_part_b.p,
      (uintptr_t)opt.stop_scaled, (uintptr_t)Gv};
  for (long ch = 0;; ++ch) {
    if (ch > max_chunks) throw std::runtime_error("PCG driver did not terminate");
    read_state();
    bool done = true;
    for (int c = 0; c < K; ++c) done = done && h_state->status[c] == KS_DONE;
    if (done) break;
    if (time_spmv)
      body(true);
    else
      run_chunk(key, [&] { body(false); });
  }
  SolveResult res;
  res.ok = true;
  for (int c = 0; c < K; ++c) {
    res.iters[c] = h_state->iters[c];
    res.err[c] = h_state->err[c];
    if (!(h_state->err[c] <= tol)) res.ok = false;
  }
  return res;
}

// Spectral radius of D^-1 A for the polynomial preconditioner (momentum_solver
// 2): power iteration on three columns of the FP64 momentum operator, 30 steps
// from a fixed pseudo-random start, 8 from the last vector after the lhs
// changed; the estimate is |D^-1 A x| for the unit x of the last step, the
// largest of the three columns.  lmax = 1.1 rho; the degree-2 Chebyshev
// iteration over [lmax / 30, lmax] from zero (the coefficients of the AMG
// smoother, amg_chebyshev), x1 = c0 D^-1 r, x2 = x1 + c1 x1 + c2 D^-1 (r - A x1),
// is x2 = (g0 I - g1 D^-1 A) D^-1 r with g0 = (1 + c1) c0 + c2, g1 = c2 c0.  Its
// error polynomial stays below 1 on (0, lmax (1 + 1/30)), so M^-1 is positive
// definite unless the estimate undershoots the spectral radius by 8%.
void NSSolver::ensure_mom_poly() {
  if (mom_poly_fresh_) return;
  const int n = nu;
  const bool warm = d_pow_x.n == (size_t)n * 3;
  if (!warm) {
    d_pow_x.alloc((size_t)n * 3);
    d_pow_v.alloc((size_t)n * 3);
    power_init3(stream_, n, d_pow_x.p);
  }
  const int steps = warm ? 8 : 30;
  std::vector<double> part;
  double rho = 0.0;
  for (int it = 0; it <= steps; ++it) {
    spmv3<double>(stream_, momA(), momV(), d_pow_x.p, d_pow_v.p, nullptr);
    const int G = power_v3(stream_, n, d_pow_v.p, d_invdA.p, d_pow_v.p, d_part_r.p);
    part.resize((size_t)6 * G);
    CFSA_HIP(hipMemcpyAsync(part.data(), d_part_r.p, part.size() * sizeof(double),
                            hipMemcpyDeviceToHost, stream_));
    CFSA_HIP(hipStreamSynchronize(stream_));
    double sc[3];
    rho = 0.0;
    for (int c = 0; c < 3; ++c) {
      double s2 = 0.0;
      for (int g = 0; g < G; ++g)
        s2 += part[2 * ((size_t)c * G + g)] + part[2 * ((size_t)c * G + g) + 1];
      const double nrm = std::sqrt(s2);
      if (!(nrm > 0.0) || !std::isfinite(nrm))
        throw std::runtime_error("momentum power iteration: zero or non-finite vector");
      sc[c] = 1.0 / nrm;
      rho = std::max(rho, nrm);
    }
    power_scale3(stream_, n, d_pow_v.p, sc, d_pow_x.p);
  }
  double c0 = 0.0, c1 = 0.0, c2 = 0.0;
  mom_rho_ = rho;
  mom_lmax_ = 1.1 * rho;
  amg_chebyshev(mom_lmax_, &c0, &c1, &c2);
  mom_g0_ = (1.0 + c1) * c0 + c2;
  mom_g1_ = c2 * c0;
  // degree 2: one more step of the same Chebyshev iteration over [lmax / 30,
  // lmax] on D^-1 A z = w from z = 0, z_{k+1} = z_k + e (z_k - z_{k-1}) +
  // f (w - D^-1 A z_k), kept as coefficients of powers of D^-1 A:
  // z_1 = c0 w, z_2 = (g0 - g1 D^-1 A) w, z_3 = (a0 + a1 D^-1 A + a2 (D^-1 A)^2) w
  {
    const double lmin = mom_lmax_ / 30.0;
    const double theta = 0.5 * (mom_lmax_ + lmin), delta = 0.5 * (mom_lmax_ - lmin);
    const double sigma = theta / delta;
    const double r0 = 1.0 / sigma, r1 = 1.0 / (2.0 * sigma - r0), r2 = 1.0 / (2.0 * sigma - r1);
    const double e = r2 * r1, f = 2.0 * r2 / delta;
    const double p1 = c0, q0 = mom_g0_, q1 = -mom_g1_;
    mom_a_[0] = q0 + e * (q0 - p1) + f;
    mom_a_[1] = q1 + e * q1 - f * q0;
    mom_a_[2] = -f * q1;
  }
  mom_poly_fresh_ = true;
  ++mom_poly_builds_;
  clear_graphs();  // captured chunks hold the coefficients
}

// CG with the polynomial preconditioner M^-1 = (g0 I - g1 D^-1 A) D^-1
// (momentum_solver 2, degree 1; three right-hand sides).  Eigen's
// conjugate_gradient step and stopping rule with z = M^-1 r: one iteration is
// q = A p with p.q; alpha, r -= alpha q and w = D^-1 r; z = g0 w - g1 D^-1 (A w)
// with the stopping norm and r.z; the convergence test, x += alpha p and
// p = z + beta p.  Two products per iteration, half the iterations of Jacobi
// CG.  Degree 2 (momentum_solver 3): u = a1 w + a2 D^-1 (A w), then
// z = a0 w + D^-1 (A u); three products per iteration, about two thirds of the
// iterations of degree 1.
template <class T, int K>
SolveResult NSSolver::pcg_poly(const DevCsr &A, const T *val, const T *invd, const T *b, T *x,
                               double tol, int maxit, KrylovWork<T> &w, int degree) {
  if (degree != 1 && degree != 2) throw std::invalid_argument("pcg_poly: degree 1 or 2");
  static_assert(K == 3, "the polynomial preconditioner runs on the momentum operator");
  ensure_mom_poly();
  hipStream_t s = stream_;
  const int n = A.n;
  KrylovState *st = d_state.p;
  const int *act = st->active;
  const bool time_spmv = opt.timing >= 2;
  set_state_params(tol, maxit, K);
  const double g0 = mom_g0_, g1 = mom_g1_;
  const double a0 = mom_a_[0], a1 = mom_a_[1], a2 = mom_a_[2];
  const bool elem = elem_product(A, sizeof(T));
  int Gv = 0, Gb = 0;
  // z = M^-1 r from r (and w = D^-1 r when stored) into zout, with the
  // stopping and r.z pairs; u: degree 2's intermediate vector
  auto precondition = [&](const T *wst, T *zout, const int *mask) {
    if (degree == 1) {
      spmv3_polyz<T>(s, A, val, wst, zout, w.r.p, invd, g0, g1, opt.stop_scaled, mask,
                     d_part_a.p, &Gv);
    } else {
      spmv3_polyu<T>(s, A, val, wst, w.t.p, w.r.p, invd, a1, a2, mask);
      spmv3_polyz<T>(s, A, val, w.t.p, zout, w.r.p, invd, a0, -1.0, opt.stop_scaled, mask,
                     d_part_a.p, &Gv, 1, elem ? nullptr : wst);
    }
  };
  // p = z0 = M^-1 b with the initial stopping and rho pairs
  cgp_init<T, K>(s, n, b, invd, x, w.r.p, w.y.p);
  precondition(w.y.p, w.p.p, nullptr);
  cg_s_init<K>(s, st, d_part_a.p, Gv);
  int chunk = std::max(1, opt.check_every);
  chunk += chunk & 1;  // even, so that every chunk starts on parity slot 0
  const long max_chunks = (long)maxit / chunk + 16;
  // the element-patch product forms w = D^-1 r from r itself: w is not stored
  T *wst = elem ? nullptr : w.y.p;
  auto body = [&](bool timed) {
    for (int j = 0; j < chunk; ++j) {
      const bool ev = timed && j == std::min(1, chunk - 1);
      if (ev) spmv_event_begin();
      spmv3_dot<T>(s, A, val, w.p.p, w.v.p, w.p.p, act, d_part_b.p, &Gb);
      if (ev) spmv_event_end();
      const int par = j & 1;
      cgp_xr<T, K>(s, n, st, d_part_b.p, Gb, par, w.v.p, invd, w.r.p, wst);
      precondition(wst, w.z.p, act);
      cg_p<T, K>(s, n, st, d_part_a.p, Gv, par, w.r.p, invd, w.p.p, x, w.z.p);
    }
  };
  uint64_t gb0 = 0, gb1 = 0, ab[3] = {0, 0, 0};
  std::memcpy(&gb0, &g0, sizeof g0);
  std::memcpy(&gb1, &g1, sizeof g1);
  std::memcpy(ab, mom_a_, sizeof ab);
  const std::vector<uintptr_t> key = {
      5, sizeof(T), K, (uintptr_t)n, (uintptr_t)chunk, (uintptr_t)A.rowptr,
      (uintptr_t)A.col, (uintptr_t)A.sptr, (uintptr_t)A.elem, (uintptr_t)val, (uintptr_t)invd,
      (uintptr_t)x, (uintptr_t)w.r.p, (uintptr_t)w.p.p, (uintptr_t)w.v.p, (uintptr_t)w.y.p,
      (uintptr_t)w.z.p, (uintptr_t)w.t.p, (uintptr_t)st, (uintptr_t)d_part_a.p,
      (uintptr_t)d_part_b.p, (uintptr_t)opt.stop_scaled, (uintptr_t)Gv, (uintptr_t)gb0,
      (uintptr_t)gb1, (uintptr_t)degree, ab[0], ab[1], ab[2]};
  for (long ch = 0;; ++ch) {
    if (ch > max_chunks) throw std::runtime_error("polynomial PCG driver did not terminate");
    read_state();
    bool done = true;
    for (int c = 0; c < K; ++c) done = done && h_state->status[c] == KS_DONE;
    if (done) break;
    if (time_spmv)
      body(true);
    else
      run_chunk(key, [&] { body(false); });
  }
  SolveResult res;
  res.ok = true;
  for (int c = 0; c < K; ++c) {
    res.iters[c] = h_state->iters[c];
    res.err[c] = h_state->err[c];
    if (!(h_state->err[c] <= tol)) res.ok = false;
  }
  return res;
}
template SolveResult NSSolver::pcg_poly<double, 3>(const DevCsr &, const double *,
                                                   const double *, const double *, double *,
                                                   double, int, KrylovWork<double> &, int);
template SolveResult NSSolver::pcg_poly<float, 3>(const DevCsr &, const float *, const float *,
                                                  const float *, float *, double, int,
                                                  KrylovWork<float> &, int);

// CG preconditioned by one V-cycle of the pressure operator's AMG hierarchy
// (amg.hpp), Eigen's conjugate_gradient step and stopping rule otherwise.  One
// iteration: q = A p with p.q; alpha, r -= alpha q and the first smoothing step
// of the finest level; the V-cycle z = M^-1 r with r.z; the convergence test,
// x += alpha p and p = z + beta p.  The CG is FP64; the V-cycle runs in T
// (float in the mixed precision = 1 mode).  Chunks of four iterations replay as
// graphs.
template <class T>
SolveResult NSSolver::pcg_amg(const double *b, double *x, double tol, int maxit) {
  ensure_amg();
  hipStream_t s = stream_;
  const int n = nv;
  KrylovState *st = d_state.p;
  const int *act = st->active;
  KrylovWork<double> &w = w1_;
  AmgLevelView<T> *lv = amg_views<T>();
  const T *cinv = amg_cinv<T>();
  AmgLevelView<T> &top = lv[0];
  const DevCsr &A = amg_fine_A();
  const double *Av = amg_fine_v();
  // T = double: the V-cycle's right-hand side is the CG residual itself
  T *b0 = std::is_same<T, double>::value ? nullptr : top.b;
  const int L = (int)amg_view_.size();
  set_state_params(tol, maxit, 1);
  int Ga = 0, Gb = 0, Gc = 0;
  amg_cg_init<T>(s, n, b, d_invdL.p, top.dinv, top.c0, x, w.r.p, b0, top.x, opt.stop_scaled,
                 d_part_a.p, &Ga);
  amg_vcycle<T>(s, lv, L, cinv, amg_nc_, amg_cjac_, w.r.p, nullptr, d_part_c.p, &Gc);
  amg_cg_s_init(s, st, d_part_a.p, Ga, d_part_c.p, Gc);
  amg_cg_p0<T>(s, n, top.y, w.p.p);
  constexpr int chunk = 4;  // even: every chunk starts on parity slot 0
  const long max_chunks = (long)maxit / chunk + 16;
  auto body = [&]() {
    for (int j = 0; j < chunk; ++j) {
      const int par = j & 1;
      spmv1_dot<double>(s, A, Av, w.p.p, w.v.p, w.p.p, act, d_part_b.p, &Gb);
      amg_cg_xr<T>(s, n, st, d_part_b.p, Gb, par, w.v.p, d_invdL.p, top.dinv, top.c0, w.r.p, b0,
                   top.x, opt.stop_scaled, d_part_a.p, &Ga);
      amg_vcycle<T>(s, lv, L, cinv, amg_nc_, amg_cjac_, w.r.p, act, d_part_c.p, &Gc);
      amg_cg_p<T>(s, n, st, d_part_a.p, Ga, d_part_c.p, Gc, par, top.y, w.p.p, x);
    }
  };
  const std::vector<uintptr_t> key = {
      3, sizeof(T), (uintptr_t)n, (uintptr_t)chunk, (uintptr_t)A.rowptr, (uintptr_t)A.sptr,
      (uintptr_t)Av, (uintptr_t)top.Av, (uintptr_t)d_invdL.p, (uintptr_t)top.dinv,
      (uintptr_t)x, (uintptr_t)w.r.p, (uintptr_t)w.p.p, (uintptr_t)w.v.p, (uintptr_t)st,
      (uintptr_t)d_part_a.p, (uintptr_t)d_part_b.p, (uintptr_t)d_part_c.p,
      (uintptr_t)opt.stop_scaled, (uintptr_t)amg_builds_, (uintptr_t)top.b, (uintptr_t)top.x,
      (uintptr_t)top.y};
  for (long ch = 0;; ++ch) {
    if (ch > max_chunks) throw std::runtime_error("AMG-PCG driver did not terminate");
    read_state();
    if (h_state->status[0] == KS_DONE) break;
    run_chunk(key, body);
  }
  SolveResult res;
  res.iters[0] = h_state->iters[0];
  res.err[0] = h_state->err[0];
  res.ok = h_state->err[0] <= tol;
  return res;
}
template SolveResult NSSolver::pcg_amg<double>(const double *, double *, double, int);
template SolveResult NSSolver::pcg_amg<float>(const double *, double *, double, int);

// FP64 iterative refinement around FP32 inner Krylov solves: the residual
// b - A x, its norm and the stopping test are FP64; matrix values, vectors and
// SpMV accumulation of the inner solve are FP32 (inner scalars stay FP64).
template <int K>
SolveResult NSSolver::refine(const DevCsr &A, const double *val64, const double *invd64,
                             const float *val32, const float *invd32, const double *b,
                             double *x, double tol, int maxit, int method, bool keep_x) {
  hipStream_t s = stream_;
  const int n = A.n;
  KrylovWork<double> &w64 = (K == 3) ? w3_ : w1_;
  KrylovWork<float> &w32 = (K == 3) ? f3_ : f1_;
  // the FP32 momentum operator: its CSR copy or the element patches (whose
  // inner solves run in patch order, with the Jacobi scale in that order)
  const DevCsr &A32 = (K == 3) ? momA32() : A;
  const int *perm = A32.elem ? d_erows.p : nullptr;
  if (perm) invd32 = d_einvd32.p;
  if (!keep_x) CFSA_HIP(hipMemsetAsync(x, 0, (size_t)n * K * sizeof(double), s));
  int Gr = 0;
  set_mask(1, K == 3, K == 3);
  // the outer test uses the same stopping rule as the FP64 solvers
  const double *sc = opt.stop_scaled ? invd64 : nullptr;
  auto residual = [&]() {
    if (K == 3)
      spmv3_residual<double>(s, A, val64, x, b, d_r64.p, w64.r0.p, sc, d_mask.p, d_part_r.p,
                             &Gr);
    else
      spmv1_residual<double>(s, A, val64, x, b, d_r64.p, w64.r0.p, sc, d_mask.p, d_part_r.p,
                             &Gr);
    hipLaunchKernelGGL((k_reduce_pairs<K>), dim3(1), dim3(kBlock), 0, s, d_part_r.p, Gr,
                       d_scalar.p);
    CFSA_LAUNCH_CHECK();
    double out[3] = {0, 0, 0};
    CFSA_HIP(hipMemcpyAsync(out, d_scalar.p, K * sizeof(double), hipMemcpyDeviceToHost, s));
    CFSA_HIP(hipStreamSynchronize(s));
    return std::vector<double>(out, out + 3);
  };
  // stopping norms of b itself: the residual of x = 0, r = b, without the SpMV
  auto residual_zero = [&]() {
    residual_x0<K>(s, n, b, d_r64.p, w64.r0.p, sc, d_part_r.p, &Gr);
    hipLaunchKernelGGL((k_reduce_pairs<K>), dim3(1), dim3(kBlock), 0, s, d_part_r.p, Gr,
                       d_scalar.p);
    CFSA_LAUNCH_CHECK();
    double out[3] = {0, 0, 0};
    CFSA_HIP(hipMemcpyAsync(out, d_scalar.p, K * sizeof(double), hipMemcpyDeviceToHost, s));
    CFSA_HIP(hipStreamSynchronize(s));
    return std::vector<double>(out, out + 3);
  };
  std::vector<double> bb = residual_zero(), rr;
  if (keep_x)
    rr = residual();  // the residual of the initial guess
  else
    rr = bb;
  SolveResult res;
  int conv[3] = {1, 1, 1};
  for (int k = 0;; ++k) {
    bool all = true;
    for (int c = 0; c < K; ++c) {
      conv[c] = (bb[c] == 0.0) || (rr[c] <= tol * tol * bb[c]);
      all = all && conv[c];
    }
    if (all || k >= opt.max_refine) {
      res.refinements = k;
      break;
    }
    set_mask(!conv[0], K == 3 && !conv[1], K == 3 && !conv[2]);
    if (perm)
      hipLaunchKernelGGL(k_demote_perm, dim3(2048), dim3(kBlock), 0, s, n, d_r64.p, d_mask.p,
                         perm, d_b32.p);
    else
      hipLaunchKernelGGL(k_demote_masked, dim3(2048), dim3(kBlock), 0, s, n, K, d_r64.p,
                         d_mask.p, d_b32.p);
    CFSA_LAUNCH_CHECK();
    // Inner tolerance: inner_tol, or looser when less than that is still needed
    // (0.7 x the reduction the least-converged active rhs needs: the FP32
    // recursive residual tracks the FP64 one to a few per cent at that point,
    // so the margin finishes the solve in this sweep), at most 0.1.
    double need = 1.0;
    for (int c = 0; c < K; ++c)
      if (!conv[c]) need = std::min(need, tol * std::sqrt(bb[c] / rr[c]));
    const double tin = std::min(0.1, std::max(opt.inner_tol, 0.7 * need));
    SolveResult in;
    if (method == 1)
      in = pcg<float, K>(A32, val32, invd32, d_b32.p, d_x32.p, tin, maxit, w32);
    else if (method == 2 || method == 3) {
      if constexpr (K == 3)
        in = pcg_poly<float, K>(A32, val32, invd32, d_b32.p, d_x32.p, tin, maxit, w32,
                                method - 1);
      else
        throw std::logic_error("the polynomial preconditioner needs the momentum operator");
    } else
      in = bicgstab<float, K>(A32, val32, invd32, d_b32.p, d_x32.p, tin, maxit, w32);
    for (int c = 0; c < K; ++c) res.iters[c] += in.iters[c];
    if (perm)
      hipLaunchKernelGGL(k_accumulate_perm, dim3(2048), dim3(kBlock), 0, s, n, d_x32.p, perm, x);
    else
      hipLaunchKernelGGL(k_accumulate, dim3(2048), dim3(kBlock), 0, s, (int64_t)n * K,
                         d_x32.p, x);
    CFSA_LAUNCH_CHECK();
    set_mask(1, K == 3, K == 3);
    rr = residual();
  }
  for (int c = 0; c < K; ++c) {
    res.err[c] = bb[c] == 0.0 ? 0.0 : std::sqrt(rr[c] / bb[c]);
    if (!(res.err[c] <= tol)) res.ok = false;
  }
  return res;
}

// precision 1: FP32 inner solves for both systems; 2: for the momentum solve only
SolveResult NSSolver::solve_momentum(const double *b, double *x) {
  if (opt.precision == 1 || opt.precision == 2) {
    // initial_guess 1: start the refinement from the current velocity u^n
    const bool guess = opt.initial_guess == 1 && x != d_u.p;
    if (guess)
      CFSA_HIP(hipMemcpyAsync(x, d_u.p, (size_t)nu * 3 * sizeof(double),
                              hipMemcpyDeviceToDevice, stream_));
    return refine<3>(momA(), momV(), d_invdA.p, d_lhs32.p, d_invdA32.p, b, x,
                     opt.momentum_tol, opt.max_iter, opt.momentum_solver, guess);
  }
  if (opt.momentum_solver == 1)
    return pcg<double, 3>(momA(), momV(), d_invdA.p, b, x, opt.momentum_tol, opt.max_iter,
                          w3_);
  if (opt.momentum_solver == 2 || opt.momentum_solver == 3)
    return pcg_poly<double, 3>(momA(), momV(), d_invdA.p, b, x, opt.momentum_tol, opt.max_iter,
                               w3_, opt.momentum_solver - 1);
  return bicgstab<double, 3>(momA(), momV(), d_invdA.p, b, x, opt.momentum_tol, opt.max_iter,
                             w3_);
}

SolveResult NSSolver::solve_pressure(const double *b, double *x) {
  // AMG: the CG is FP64; the mixed mode for both solves runs the V-cycle in FP32
  if (opt.pressure_solver == 2)
    return opt.precision == 1 ? pcg_amg<float>(b, x, opt.pressure_tol, opt.max_iter)
                              : pcg_amg<double>(b, x, opt.pressure_tol, opt.max_iter);
  if (opt.precision == 1)
    return refine<1>(preA(), preV(), d_invdL.p, d_Lbc32.p, d_invdL32.p, b, x,
                     opt.pressure_tol, opt.max_iter, opt.pressure_solver);
  if (opt.pressure_solver == 1)
    return pcg<double, 1>(preA(), preV(), d_invdL.p, b, x, opt.pressure_tol, opt.max_iter, w1_);
  return bicgstab<double, 1>(preA(), preV(), d_invdL.p, b, x, opt.pressure_tol, opt.max_iter,
                             w1_);
}

// ---------------------------------------------------------------------------
// PISO step (NavierStokes.cpp:374-414)
// ---------------------------------------------------------------------------
int NSSolver::momentum_predictor() {
  // computeAdvectionRHS (:526-584) from the P1 vertex velocities
  conv_tet(stream_, nt, d_tets.p, d_grad.p, d_u.p, d_cT.p);
  conv_gather(stream_, nu, nv, d_incp.p, d_inct.p, d_vol.p, d_cT.p, d_conv.p);
  // G_c p for the P1 rows (:441-452)
  spmv_grad3(stream_, P_, d_G3.p, d_p.p, d_gp.p);
  momentum_rhs(stream_, A_, d_M.p, d_u.p, d_gp.p, d_conv.p, d_bc.p, d_ginlet.p, dt, nv,
               kMomentumPenalty, d_rhs.p);
  SolveResult r = solve_momentum(d_rhs.p, d_us.p);
  for (int c = 0; c < 3; ++c) {
    last.momentum_iters[c] = r.iters[c];
    last.momentum_err[c] = r.err[c];
  }
  last.momentum_refinements = r.refinements;
  if (!r.ok) {
    int bad = 0;
    while (bad < 3 && r.err[bad] <= opt.momentum_tol) ++bad;
    last_error = "[NS] Momentum solve failed for component " + std::to_string(bad);
    std::cerr << last_error << std::endl;
    std::cerr << "[NS] Momentum predictor failed" << std::endl;
    return ST_MOMENTUM_NOCONV;
  }
  return ST_OK;
}

int NSSolver::pressure_poisson() {
  // computeDivergence (:670-700) and rhs = -div/dt with penalties (:598-626)
  div_tet(stream_, nt, d_tets.p, d_grad.p, d_vol.p, d_us.p, d_wT.p);
  const int mode = outlet_type == 0 ? 1 : (outlet_type == 1 ? 2 : 0);
  div_gather_prhs(stream_, nv, d_incp.p, d_inct.p, d_wT.p, d_bc.p, mode, reference_node, dt,
                  outlet_pressure, kPressurePenalty, d_div.p, d_rhsp.p);
  SolveResult r = solve_pressure(d_rhsp.p, d_pc.p);
  last.pressure_solves++;
  const int k = std::min(last.n_pressure, 7);
  last.pressure_iters[k] = r.iters[0];
  last.pressure_err[k] = r.err[0];
  last.pressure_refinements[k] = r.refinements;
  last.n_pressure++;
  if (!r.ok) {
    last_error = "[NS] Pressure Poisson solve failed";
    std::cerr << last_error << std::endl;
    return ST_PRESSURE_NOCONV;
  }
  // p += relax_p p'; Dirichlet outlet reset; normalize to reference node
  pressure_update(stream_, nv, d_p.p, d_pc.p, relax_p, d_bc.p, outlet_type == 0,
                  outlet_pressure);
  pressure_normalize(stream_, nv, d_p.p, reference_node, d_scalar.p);
  return ST_OK;
}

// A corrector after the first: solvePressurePoisson (NavierStokes.cpp:586-668)
// would rebuild the same right-hand side from the unchanged u* and solve the
// same L_bc system from zero, so p' (d_pc) is already its result; only the
// pressure update, outlet reset and normalization repeat.  correctVelocity
// (:702-771) would recompute u from the same u* and p', so u is already its
// result too.  The statistics repeat the first solve's.
void NSSolver::repeat_pressure_correction() {
  const int k = std::min(last.n_pressure, 7);
  last.pressure_iters[k] = last.pressure_iters[0];
  last.pressure_err[k] = last.pressure_err[0];
  last.pressure_refinements[k] = last.pressure_refinements[0];
  last.n_pressure++;
  pressure_update(stream_, nv, d_p.p, d_pc.p, relax_p, d_bc.p, outlet_type == 0,
                  outlet_pressure);
  pressure_normalize(stream_, nv, d_p.p, reference_node, d_scalar.p);
}

void NSSolver::correct_velocity() {
  spmv_grad3(stream_, P_, d_G3.p, d_pc.p, d_gp.p);
  cfsa::correct_velocity(stream_, nu, nv, d_bc.p, d_ginlet.p, d_ilm.p, d_gp.p, d_us.p, dt,
                         relax_u, d_u.p);
}

bool NSSolver::has_nans() {
  d_flag.zero(stream_);
  nan_check(stream_, nu * 3, d_u.p, d_flag.p);
  nan_check(stream_, nu * 3, d_us.p, d_flag.p);
  nan_check(stream_, nv, d_p.p, d_flag.p);
  nan_check(stream_, nv, d_pc.p, d_flag.p);
  int f = 0;
  CFSA_HIP(hipMemcpyAsync(&f, d_flag.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
  CFSA_HIP(hipStreamSynchronize(stream_));
  return f != 0;
}

int NSSolver::finish_step() {
  reset_velocity_bc(stream_, nu, d_bc.p, d_ginlet.p, d_u.p);
  time += dt;
  return has_nans() ? ST_NAN : ST_OK;
}

int NSSolver::momentum_predictor_step() {
  if (!initialized) return ST_NOT_INIT;
  prepare_operators();
  const int st = momentum_predictor();
  CFSA_HIP(hipStreamSynchronize(stream_));
  return st;
}

int NSSolver::pressure_poisson_step() {
  if (!initialized) return ST_NOT_INIT;
  prepare_operators();
  const int st = pressure_poisson();
  CFSA_HIP(hipStreamSynchronize(stream_));
  return st;
}

void NSSolver::correct_velocity_step() {
  if (!initialized) throw std::runtime_error("[NS] Error: solver not initialized");
  prepare_operators();
  correct_velocity();
  CFSA_HIP(hipStreamSynchronize(stream_));
}

void NSSolver::normalize_pressure_step() {
  if (!initialized) throw std::runtime_error("[NS] Error: solver not initialized");
  pressure_normalize(stream_, nv, d_p.p, reference_node, d_scalar.p);
  CFSA_HIP(hipStreamSynchronize(stream_));
}

SolveResult NSSolver::solve_poisson_bicgstab(const double *rhs, double *x) {
  if (!initialized) throw std::runtime_error("[NS] Error: solver not initialized");
  prepare_operators();
  d_bsb.upload(rhs, (size_t)nv);
  d_bsx.alloc((size_t)nv);
  d_bsinvd.alloc((size_t)nv);
  inv_diag(stream_, nv, d_Pdiag.p, d_L.p, d_bsinvd.p);
  const int saved = opt.stop_scaled;
  opt.stop_scaled = 0;  // Eigen's |r| <= tol |b|
  SolveResult r;
  try {
    r = bicgstab<double, 1>(P_, d_L.p, d_bsinvd.p, d_bsb.p, d_bsx.p, 1e-8, 1000, w1_);
  } catch (...) {
    opt.stop_scaled = saved;
    throw;
  }
  opt.stop_scaled = saved;
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_bsx.download(x, (size_t)nv);
  return r;
}

int NSSolver::piso_step() {
  if (!initialized) {
    last_error = "[NS] Error: solver not initialized";
    std::cerr << last_error << std::endl;
    return ST_NOT_INIT;
  }
  std::memset(&last, 0, sizeof(last));
  ev_used_ = 0;
  hipEvent_t ev[6] = {};
  const bool timing = opt.timing >= 1;
  if (timing) {
    for (auto &e : ev) CFSA_HIP(hipEventCreate(&e));
    CFSA_HIP(hipEventRecord(ev[0], stream_));
  }
  prepare_operators();
  int st = momentum_predictor();
  if (timing) CFSA_HIP(hipEventRecord(ev[1], stream_));
  if (st == ST_OK) {
    for (int k = 0; k < n_corrections && st == ST_OK; ++k) {
      if (k > 0 && opt.reuse_correction) {
        repeat_pressure_correction();
        continue;
      }
      st = pressure_poisson();
      if (st == ST_OK) correct_velocity();
    }
  }
  if (timing) CFSA_HIP(hipEventRecord(ev[2], stream_));
  if (st == ST_OK) st = finish_step();
  if (timing) {
    CFSA_HIP(hipEventRecord(ev[3], stream_));
    CFSA_HIP(hipEventSynchronize(ev[3]));
    float a = 0, b = 0, c = 0;
    CFSA_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
    CFSA_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
    CFSA_HIP(hipEventElapsedTime(&c, ev[0], ev[3]));
    last.ms_momentum = a;
    last.ms_pressure = b;
    last.ms_total = c;
    for (auto &e : ev) (void)hipEventDestroy(e);
  }
  if (opt.timing >= 2) {
    double tot = 0.0;
    for (int i = 0; i + 1 < ev_used_; i += 2) {
      float ms = 0;
      CFSA_HIP(hipEventElapsedTime(&ms, ev_pool_[(size_t)i], ev_pool_[(size_t)i + 1]));
      tot += ms;
    }
    last.spmv_launches = ev_used_ / 2;
    last.spmv_ms = tot;
  }
  return st;
}

// simpleStep (NavierStokes.cpp:968-1009): one pressure correction with an
// extra outlet-pressure reset before the velocity correction.
int NSSolver::simple_step() {
  if (!initialized) {
    last_error = "[NS-Simple] Error: solver not initialized";
    std::cerr << last_error << std::endl;
    return ST_NOT_INIT;
  }
  std::memset(&last, 0, sizeof(last));
  prepare_operators();
  int st = momentum_predictor();
  if (st != ST_OK) return st;
  st = pressure_poisson();
  if (st != ST_OK) return st;
  // reinforcePressureBCs (NavierStokes.cpp:1011-1024)
  if (outlet_type == 0) pressure_outlet_reset(stream_, nv, d_p.p, d_bc.p, outlet_pressure);
  correct_velocity();
  return finish_step();
}

// ---------------------------------------------------------------------------
// Field access and parity entry points
// ---------------------------------------------------------------------------
void NSSolver::get_velocity(double *out) {
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_u.download(out, (size_t)nu * 3);
}
// The snapshot is two device-to-device copies on the solver's stream (about
// 0.1 ms at 96^3), so later steps may overwrite u and p while the host copies
// run; the next snapshot waits for the previous host copies first.
void NSSolver::fields_async(double *u_out, double *p_out) {
  if (!copy_stream_) {
    CFSA_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
    CFSA_HIP(hipEventCreateWithFlags(&ev_snap_, hipEventDisableTiming));
    CFSA_HIP(hipEventCreateWithFlags(&ev_copied_, hipEventDisableTiming));
  }
  if (copy_pending_) CFSA_HIP(hipStreamWaitEvent(stream_, ev_copied_, 0));
  d_usnap.alloc((size_t)nu * 3);
  d_psnap.alloc((size_t)nv);
  CFSA_HIP(hipMemcpyAsync(d_usnap.p, d_u.p, (size_t)nu * 3 * sizeof(double),
                          hipMemcpyDeviceToDevice, stream_));
  CFSA_HIP(hipMemcpyAsync(d_psnap.p, d_p.p, (size_t)nv * sizeof(double),
                          hipMemcpyDeviceToDevice, stream_));
  CFSA_HIP(hipEventRecord(ev_snap_, stream_));
  CFSA_HIP(hipStreamWaitEvent(copy_stream_, ev_snap_, 0));
  if (u_out)
    CFSA_HIP(hipMemcpyAsync(u_out, d_usnap.p, (size_t)nu * 3 * sizeof(double),
                            hipMemcpyDeviceToHost, copy_stream_));
  if (p_out)
    CFSA_HIP(hipMemcpyAsync(p_out, d_psnap.p, (size_t)nv * sizeof(double),
                            hipMemcpyDeviceToHost, copy_stream_));
  CFSA_HIP(hipEventRecord(ev_copied_, copy_stream_));
  copy_pending_ = true;
}
void NSSolver::fields_wait() {
  if (!copy_pending_) return;
  CFSA_HIP(hipEventSynchronize(ev_copied_));
  copy_pending_ = false;
}
void NSSolver::get_velocity_at(int64_t node, double out[3]) {
  CFSA_HIP(hipStreamSynchronize(stream_));
  CFSA_HIP(hipMemcpy(out, d_u.p + (size_t)node * 3, 3 * sizeof(double), hipMemcpyDeviceToHost));
}
void NSSolver::get_pressure_at(int64_t node, double *out) {
  CFSA_HIP(hipStreamSynchronize(stream_));
  CFSA_HIP(hipMemcpy(out, d_p.p + (size_t)node, sizeof(double), hipMemcpyDeviceToHost));
}
void NSSolver::set_velocity(const double *in) {
  CFSA_HIP(hipStreamSynchronize(stream_));
  CFSA_HIP(hipMemcpy(d_u.p, in, (size_t)nu * 3 * sizeof(double), hipMemcpyHostToDevice));
}
void NSSolver::get_velocity_star(double *out) {
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_us.download(out, (size_t)nu * 3);
}
void NSSolver::get_pressure(double *out) {
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_p.download(out, (size_t)nv);
}
void NSSolver::set_pressure(const double *in) {
  CFSA_HIP(hipStreamSynchronize(stream_));
  CFSA_HIP(hipMemcpy(d_p.p, in, (size_t)nv * sizeof(double), hipMemcpyHostToDevice));
}
void NSSolver::get_pressure_correction(double *out) {
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_pc.download(out, (size_t)nv);
}
void NSSolver::get_inv_lumped_mass(double *out) {
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_ilm.download(out, (size_t)nu);
}
void NSSolver::get_colouring(int32_t *colour) const {
  for (int c = 0; c < ncolours; ++c)
    for (int q = h_colour_offsets[(size_t)c]; q < h_colour_offsets[(size_t)c + 1]; ++q)
      colour[h_order[(size_t)q]] = c;
}
void NSSolver::get_mesh_arrays(double *coords, int32_t *tets, uint8_t *p1_bc) const {
  if (coords) std::memcpy(coords, h_coords.data(), h_coords.size() * sizeof(double));
  if (tets) std::memcpy(tets, h_tets.data(), h_tets.size() * sizeof(int32_t));
  if (p1_bc) std::memcpy(p1_bc, h_bc.data(), (size_t)nv);
}

void NSSolver::get_incidence(int32_t *ptr, int32_t *tet) {
  d_incp.download(ptr, (size_t)nu + 1);
  d_inct.download(tet, (size_t)nt * 10);
}

// which: 0 M, 1 K, 2 momentum lhs (with penalties), 3 L, 4 L_bc, 5 Gx, 6 Gy, 7 Gz
void NSSolver::get_matrix(int which, int32_t *rowptr, int32_t *col, double *val) {
  prepare_operators();
  CFSA_HIP(hipStreamSynchronize(stream_));
  const bool p2 = which <= 2;
  const DBuf<int> &rp = p2 ? d_Arp : d_Prp;
  const DBuf<int> &cl = p2 ? d_Acol : d_Pcol;
  const int64_t nnz = p2 ? nnzA : nnzP;
  if (rowptr) rp.download(rowptr, rp.n);
  if (col) cl.download(col, (size_t)nnz);
  if (!val) return;
  if (which == 0) d_M.download(val, (size_t)nnz);
  else if (which == 1) d_K.download(val, (size_t)nnz);
  else if (which == 2) d_lhs.download(val, (size_t)nnz);
  else if (which == 3) d_L.download(val, (size_t)nnz);
  else if (which == 4) d_Lbc.download(val, (size_t)nnz);
  else if (which >= 5 && which <= 7) {
    std::vector<double> g((size_t)nnz * 3);
    d_G3.download(g.data(), g.size());
    for (int64_t k = 0; k < nnz; ++k) val[k] = g[(size_t)k * 3 + (which - 5)];
  } else
    throw std::invalid_argument("get_matrix: unknown matrix id");
}

void NSSolver::compute_advection(double *out) {
  conv_tet(stream_, nt, d_tets.p, d_grad.p, d_u.p, d_cT.p);
  conv_gather(stream_, nu, nv, d_incp.p, d_inct.p, d_vol.p, d_cT.p, d_conv.p);
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_conv.download(out, (size_t)nu * 3);
}

void NSSolver::compute_divergence(const double *u, double *out) {
  CFSA_HIP(hipMemcpy(d_xref.p, u, (size_t)nu * 3 * sizeof(double), hipMemcpyHostToDevice));
  div_tet(stream_, nt, d_tets.p, d_grad.p, d_vol.p, d_xref.p, d_wT.p);
  div_gather_prhs(stream_, nv, d_incp.p, d_inct.p, d_wT.p, d_bc.p, 0, -1, 1.0, 0.0, 0.0,
                  d_div.p, d_rhsp.p);
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_div.download(out, (size_t)nv);
}

// y = A x for matrix `which` (see get_matrix); nrhs 1 or 3 for P2 matrices,
// and for G (5..7) x has nv entries and y is nv x 3 (all three components).
void NSSolver::apply_matrix(int which, const double *x, double *y, int nrhs) {
  if (which == 8) {
    // the FP32 momentum operator of the mixed-precision inner solves (CSR copy
    // or element patches), x and y converted on the device
    if (nrhs != 3) throw std::invalid_argument("apply_matrix: the FP32 operator takes 3 rhs");
    const int saved = opt.precision;
    if (opt.precision == 0) opt.precision = 2;
    try {
      prepare_operators();
    } catch (...) {
      opt.precision = saved;
      throw;
    }
    opt.precision = saved;
    const size_t m = (size_t)nu * 3;
    DBuf<double> dx, dy;
    DBuf<float> fx, fy;
    dx.upload(x, m);
    dy.alloc(m);
    fx.alloc(m);
    fy.alloc(m);
    const DevCsr &A32 = momA32();
    if (A32.elem) {
      // patch order in, node order out
      DBuf<int> one;
      const int ones[3] = {1, 1, 1};
      one.upload(ones, 3);
      hipLaunchKernelGGL(k_demote_perm, dim3(2048), dim3(kBlock), 0, stream_, nu, dx.p, one.p,
                         d_erows.p, fx.p);
      CFSA_LAUNCH_CHECK();
      spmv3<float>(stream_, A32, d_lhs32.p, fx.p, fy.p, nullptr);
      CFSA_HIP(hipMemsetAsync(dy.p, 0, m * sizeof(double), stream_));
      hipLaunchKernelGGL(k_accumulate_perm, dim3(2048), dim3(kBlock), 0, stream_, nu, fy.p,
                         d_erows.p, dy.p);
      CFSA_LAUNCH_CHECK();
    } else {
      convert<float, double>(stream_, (int64_t)m, dx.p, fx.p);
      spmv3<float>(stream_, A32, d_lhs32.p, fx.p, fy.p, nullptr);
      convert<double, float>(stream_, (int64_t)m, fy.p, dy.p);
    }
    CFSA_HIP(hipStreamSynchronize(stream_));
    dy.download(y, m);
    return;
  }
  prepare_operators();
  const bool p2 = which <= 2;
  const int n = p2 ? nu : nv;
  const double *val = which == 0 ? d_M.p : which == 1 ? d_K.p : which == 2 ? d_lhs.p
                      : which == 3 ? d_L.p : d_Lbc.p;
  DBuf<double> dx, dy;
  if (which >= 5) {
    dx.upload(x, (size_t)nv);
    dy.alloc((size_t)nv * 3);
    spmv_grad3(stream_, P_, d_G3.p, dx.p, dy.p);
    CFSA_HIP(hipStreamSynchronize(stream_));
    dy.download(y, (size_t)nv * 3);
    return;
  }
  dx.upload(x, (size_t)n * nrhs);
  dy.alloc((size_t)n * nrhs);
  // the Krylov operators run through their sliced-ELL copies
  const DevCsr &A = which == 2 ? momA() : which == 4 ? preA() : p2 ? A_ : P_;
  if (which == 2) val = momV();
  if (which == 4) val = preV();
  if (nrhs == 3)
    spmv3<double>(stream_, A, val, dx.p, dy.p, nullptr);
  else
    spmv1<double>(stream_, A, val, dx.p, dy.p);
  CFSA_HIP(hipStreamSynchronize(stream_));
  dy.download(y, (size_t)n * nrhs);
}

// Solve matrix `which` (2 momentum lhs, 4 L_bc) with method 0 BiCGSTAB / 1 PCG /
// 2 AMG-PCG (L_bc) / 3 and 4 polynomial PCG of degree 1 and 2 (momentum, 3 rhs)
// in FP64 (precision 0) or FP32-inner refinement (precision 1).
SolveResult NSSolver::krylov_solve(int which, int method, int nrhs, const double *b,
                                   double *x, double tol, int maxit, int precision) {
  if (which != 2 && which != 4) throw std::invalid_argument("krylov_solve: matrix 2 or 4");
  if (nrhs != 1 && nrhs != 3) throw std::invalid_argument("krylov_solve: nrhs 1 or 3");
  if (which == 4 && nrhs != 1) throw std::invalid_argument("krylov_solve: L_bc takes 1 rhs");
  const int saved = opt.precision;
  opt.precision = precision;
  prepare_operators();
  opt.precision = saved;
  const DevCsr &A = which == 2 ? momA() : preA();
  const int n = A.n;
  DBuf<double> db, dx;
  db.upload(b, (size_t)n * nrhs);
  dx.alloc((size_t)n * nrhs);
  SolveResult r;
  if (method == 3 || method == 4) {
    if (which != 2 || nrhs != 3)
      throw std::invalid_argument("krylov_solve: methods 3 and 4 (polynomial PCG) solve the "
                                  "momentum matrix with 3 rhs");
    if (precision == 1)
      r = refine<3>(A, momV(), d_invdA.p, d_lhs32.p, d_invdA32.p, db.p, dx.p, tol, maxit,
                    method - 1);
    else
      r = pcg_poly<double, 3>(A, momV(), d_invdA.p, db.p, dx.p, tol, maxit, w3_, method - 2);
  } else if (method == 2) {
    if (which != 4) throw std::invalid_argument("krylov_solve: method 2 (AMG-PCG) solves L_bc");
    r = precision == 1 ? pcg_amg<float>(db.p, dx.p, tol, maxit)
                       : pcg_amg<double>(db.p, dx.p, tol, maxit);
  } else if (precision == 1) {
    if (nrhs == 3)
      r = refine<3>(A, momV(), d_invdA.p, d_lhs32.p, d_invdA32.p, db.p, dx.p, tol, maxit,
                    method);
    else if (which == 2)
      throw std::invalid_argument("krylov_solve: mixed momentum needs 3 rhs");
    else
      r = refine<1>(A, preV(), d_invdL.p, d_Lbc32.p, d_invdL32.p, db.p, dx.p, tol, maxit,
                    method);
  } else if (method == 1 && nrhs == 3) {
    r = pcg<double, 3>(A, momV(), d_invdA.p, db.p, dx.p, tol, maxit, w3_);
  } else if (method == 1) {
    r = pcg<double, 1>(A, which == 2 ? momV() : preV(), which == 2 ? d_invdA.p : d_invdL.p,
                       db.p, dx.p, tol, maxit, which == 2 ? w3_ : w1_);
  } else if (nrhs == 3) {
    r = bicgstab<double, 3>(A, momV(), d_invdA.p, db.p, dx.p, tol, maxit, w3_);
  } else {
    r = bicgstab<double, 1>(A, which == 2 ? momV() : preV(),
                            which == 2 ? d_invdA.p : d_invdL.p, db.p, dx.p, tol, maxit,
                            which == 2 ? w3_ : w1_);
  }
  CFSA_HIP(hipStreamSynchronize(stream_));
  dx.download(x, (size_t)n * nrhs);
  return r;
}

double NSSolver::time_momentum_spmv(int reps, int precision) {
  prepare_operators();
  if (precision == 1) {
    ensure_mixed_buffers();
    prepare_operators();
  }
  hipEvent_t a, b;
  CFSA_HIP(hipEventCreate(&a));
  CFSA_HIP(hipEventCreate(&b));
  CFSA_HIP(hipEventRecord(a, stream_));
  for (int i = 0; i < reps; ++i) {
    if (precision == 1)
      spmv3<float>(stream_, momA32(), d_lhs32.p, f3_.y.p, f3_.v.p, nullptr);
    else
      spmv3<double>(stream_, momA(), momV(), w3_.y.p, w3_.v.p, nullptr);
  }
  CFSA_HIP(hipEventRecord(b, stream_));
  CFSA_HIP(hipEventSynchronize(b));
  float ms = 0;
  CFSA_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

}  // namespace cfsa
