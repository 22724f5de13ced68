// AMG hierarchy of the pressure operator on the device (pressure_solver = 2):
// host set-up from the current L_bc values (amg.cpp), upload of every level
// as CSR-stream operators, and the parity entry points.
#include <chrono>
#include <cstring>

#include "amg.hpp"
#include "pattern.hpp"
#include "solver.hpp"

namespace cfsa {

namespace {

// uploads a host CSR as a CSR-stream operator (windows of up to 2048 nonzeros
// and kStreamRows rows, one right-hand side)
void upload_stream(const HostCsr &h, DBuf<int> &rp, DBuf<int> &ci, DBuf<int> &win,
                   DBuf<double> &val, DevCsr &d) {
  CsrPattern cp;
  cp.n = h.n;
  cp.rowptr = h.rp;
  const std::vector<int32_t> w = stream_windows(cp, kStreamB<double, 1>, kStreamRows);
  rp.upload(h.rp.data(), h.rp.size());
  ci.upload(h.ci.data(), h.ci.size());
  val.upload(h.v.data(), h.v.size());
  win.upload(w.data(), w.size());
  d = DevCsr{};
  d.n = (int)h.n;
  d.nnz = h.nnz();
  d.rowptr = rp.p;
  d.col = ci.p;
  d.win1 = win.p;
  d.nwin1 = (int)w.size() - 1;
}

}  // namespace

// Builds (or keeps) the hierarchy for the current L_bc.  A hierarchy is rebuilt
// when the values differ from those it was built from, unless amg_reuse allows
// keeping it for that many changes; identical values (a re-assembly of an
// unchanged mesh) never rebuild.
void NSSolver::ensure_amg() {
  if (amg_built_ && !amg_lbc_changed_) return;
  std::vector<double> v((size_t)nnzP);
  CFSA_HIP(hipStreamSynchronize(stream_));
  d_Lbc.download(v.data(), (size_t)nnzP);
  amg_lbc_changed_ = false;
  if (amg_built_) {
    if (v == amg_vals_) return;
    if (amg_stale_ < opt.amg_reuse) {
      ++amg_stale_;
      return;
    }
  }
  const auto t0 = std::chrono::steady_clock::now();
  HostCsr A0;
  A0.n = A0.m = nv;
  A0.rp.resize((size_t)nv + 1);
  A0.ci.resize((size_t)nnzP);
  d_Prp.download(A0.rp.data(), (size_t)nv + 1);
  d_Pcol.download(A0.ci.data(), (size_t)nnzP);
  A0.v = v;
  const AmgHierarchy H = amg_build(A0);
  const int L = (int)H.levels.size();
  clear_graphs();  // captured chunks hold the previous hierarchy's buffers
  amg_lv_.clear();
  amg_lv_.resize((size_t)L);
  amg_view_.assign((size_t)L, AmgLevelView{});
  amg_rows_.assign((size_t)L, 0);
  amg_nnz_.assign((size_t)L, 0);
  for (int l = 0; l < L; ++l) {
    const AmgLevel &h = H.levels[(size_t)l];
    AmgDevLevel &d = amg_lv_[(size_t)l];
    AmgLevelView &w = amg_view_[(size_t)l];
    w.n = (int)h.A.n;
    w.c0 = h.c0;
    w.c1 = h.c1;
    w.c2 = h.c2;
    amg_rows_[(size_t)l] = h.A.n;
    amg_nnz_[(size_t)l] = h.A.nnz();
    if (l > 0) {
      upload_stream(h.A, d.Arp, d.Aci, d.Awin, d.Av, w.A);
      w.Av = d.Av.p;
      d.dinv.upload(h.dinv.data(), h.dinv.size());
      w.dinv = d.dinv.p;
      d.b.alloc((size_t)h.A.n);
      w.b = d.b.p;
    }
    if (l < L - 1) {
      upload_stream(h.P, d.Prp, d.Pci, d.Pwin, d.Pv, w.P);
      w.Pv = d.Pv.p;
      upload_stream(h.R, d.Rrp, d.Rci, d.Rwin, d.Rv, w.R);
      w.Rv = d.Rv.p;
    }
    d.x.alloc((size_t)h.A.n);
    d.y.alloc((size_t)h.A.n);
    d.r.alloc((size_t)h.A.n);
    w.x = d.x.p;
    w.y = d.y.p;
    w.r = d.r.p;
  }
  amg_cjac_ = H.coarse_jacobi;
  amg_nc_ = (int)H.coarse_n;
  if (!amg_cjac_)
    amg_cinv_.upload(H.coarse_inv.data(), H.coarse_inv.size());
  else
    amg_cinv_.release();
  amg_vals_ = std::move(v);
  amg_stale_ = 0;
  amg_built_ = true;
  ++amg_builds_;
  amg_setup_ms_ =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void NSSolver::amg_apply(const double *r, double *z) {
  if (!initialized) throw std::runtime_error("amg_apply needs an initialized solver");
  prepare_operators();
  ensure_amg();
  AmgLevelView &top = amg_view_[0];
  top.A = preA();
  top.Av = preV();
  top.dinv = d_invdL.p;
  DBuf<double> db;
  db.upload(r, (size_t)nv);
  w1_.alloc((size_t)nv);
  top.b = w1_.r.p;
  int Ga = 0, Gc = 0;
  // r is copied into the CG residual buffer with the first smoothing step;
  // x goes to a scratch vector
  amg_cg_init(stream_, nv, db.p, d_invdL.p, top.c0, w1_.s.p, w1_.r.p, top.x, 0, d_part_a.p, &Ga);
  amg_vcycle(stream_, amg_view_.data(), (int)amg_view_.size(), amg_cinv_.p, amg_nc_, amg_cjac_,
             nullptr, d_part_c.p, &Gc);
  CFSA_HIP(hipStreamSynchronize(stream_));
  CFSA_HIP(hipMemcpy(z, top.y, (size_t)nv * sizeof(double), hipMemcpyDeviceToHost));
}

int NSSolver::amg_info(int cap, int64_t *rows, int64_t *nnz, double *setup_ms, int *builds) {
  const int L = amg_built_ ? (int)amg_rows_.size() : 0;
  for (int l = 0; l < L && l < cap; ++l) {
    if (rows) rows[l] = amg_rows_[(size_t)l];
    if (nnz) nnz[l] = amg_nnz_[(size_t)l];
  }
  if (setup_ms) *setup_ms = amg_setup_ms_;
  if (builds) *builds = amg_builds_;
  return L;
}

}  // namespace cfsa
