// AMG hierarchy of the pressure operator on the device (pressure_solver = 2):
// host set-up from the current L_bc values (amg.cpp), upload of every level's
// operators (CSR-stream, or sliced-ELL), and the parity entry points.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "amg.hpp"
#include "pattern.hpp"
#include "solver.hpp"

namespace cfsa {

namespace {

// uploads a host CSR as a CSR-stream operator (windows of up to 2048 nonzeros
// and kStreamRows rows, one right-hand side), or, when a row is longer than a
// window holds (coarse operators of unstructured meshes grow dense), as
// sliced-ELL (also when `sell` asks for it); the SpMV launchers pick the kernel
// from the layout
template <class M>
void upload_level_matrix(const HostCsr &h, M &m, DevCsr &d, bool sell) {
  CsrPattern cp;
  cp.n = h.n;
  cp.rowptr = h.rp;
  int32_t longest = 0;
  for (int64_t i = 0; i < h.n; ++i)
    longest = std::max(longest, h.rp[(size_t)i + 1] - h.rp[(size_t)i]);
  d = DevCsr{};
  d.n = (int)h.n;
  d.nnz = h.nnz();
  if (!sell && longest <= kStreamB<double, 1>) {
    const std::vector<int32_t> w = stream_windows(cp, kStreamB<double, 1>, kStreamRows);
    m.rp.upload(h.rp.data(), h.rp.size());
    m.ci.upload(h.ci.data(), h.ci.size());
    m.v.upload(h.v.data(), h.v.size());
    m.win.upload(w.data(), w.size());
    d.rowptr = m.rp.p;
    d.col = m.ci.p;
    d.win1 = m.win.p;
    d.nwin1 = (int)w.size() - 1;
    return;
  }
  cp.col = h.ci;
  const SellPattern S = build_sell(cp);
  std::vector<double> sv((size_t)S.stored());
  for (size_t e = 0; e < sv.size(); ++e) sv[e] = S.src[e] >= 0 ? h.v[(size_t)S.src[e]] : 0.0;
  m.sptr.upload(S.sptr.data(), S.sptr.size());
  m.rlen.upload(S.rlen.data(), S.rlen.size());
  m.perm.upload(S.perm.data(), S.perm.size());
  m.ci.upload(S.col.data(), S.col.size());
  m.v.upload(sv.data(), sv.size());
  d.col = m.ci.p;
  d.sptr = m.sptr.p;
  d.rlen = m.rlen.p;
  d.perm = m.perm.p;
  d.nslices = S.nslices;
}

}  // namespace

// Builds (or keeps) the hierarchy for the current L_bc.  A hierarchy is rebuilt
// when the values differ from the last ones seen (a 64-bit fingerprint taken on
// the device, so an unchanged operator costs no copy), unless amg_reuse allows
// keeping it for that many changes; identical values (a re-assembly of an
// unchanged mesh) never rebuild.
void NSSolver::ensure_amg() {
  if (amg_built_ && !amg_lbc_changed_) return;
  amg_lbc_changed_ = false;
  d_print_.alloc(1);
  fingerprint(stream_, nnzP, d_Lbc.p, d_print_.p);
  uint64_t print = 0;
  CFSA_HIP(hipMemcpyAsync(&print, d_print_.p, sizeof(uint64_t), hipMemcpyDeviceToHost,
                          stream_));
  CFSA_HIP(hipStreamSynchronize(stream_));
  if (amg_built_) {
    if (print == amg_print_) return;
    amg_print_ = print;
    if (amg_stale_ < opt.amg_reuse) {
      ++amg_stale_;
      amg_top_fresh_ = false;  // level 0 reads the current L_bc from here on
      amg32_built_ = false;
      return;
    }
  }
  amg_print_ = print;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<double> v((size_t)nnzP);
  d_Lbc.download(v.data(), (size_t)nnzP);
  HostCsr A0;
  A0.n = A0.m = nv;
  A0.rp.resize((size_t)nv + 1);
  A0.ci.resize((size_t)nnzP);
  d_Prp.download(A0.rp.data(), (size_t)nv + 1);
  d_Pcol.download(A0.ci.data(), (size_t)nnzP);
  A0.v = std::move(v);
  const AmgHierarchy H = amg_build(A0);
  const int L = (int)H.levels.size();
  // coarse operators follow spmv_layout 2 (sliced-ELL); otherwise CSR-stream
  // unless a row is too long for it
  const bool sell = opt.spmv_layout == 2;
  clear_graphs();  // captured chunks hold the previous hierarchy's buffers
  amg_lv_.clear();
  amg_lv_.resize((size_t)L);
  amg_view_.assign((size_t)L, AmgLevelView<double>{});
  amg32_built_ = false;
  amg_rows_.assign((size_t)L, 0);
  amg_nnz_.assign((size_t)L, 0);
  for (int l = 0; l < L; ++l) {
    const AmgLevel &h = H.levels[(size_t)l];
    AmgDevLevel &d = amg_lv_[(size_t)l];
    AmgLevelView<double> &w = amg_view_[(size_t)l];
    w.n = (int)h.A.n;
    w.c0 = h.c0;
    w.c1 = h.c1;
    w.c2 = h.c2;
    amg_rows_[(size_t)l] = h.A.n;
    amg_nnz_[(size_t)l] = h.A.nnz();
    if (l == 0) {
      // the zero-free fine operator, in the pressure Krylov operator's layout
      upload_level_matrix(h.A, d.A, amg_top_A_, preA().sptr != nullptr);
      // the CG's product with its fused inner product uses the Krylov
      // operator's window scratch (this operator has no more windows than it)
      amg_top_A_.wpart = preA().wpart;
      amg_top_A_.wpart_cap = preA().wpart_cap;
    } else {
      upload_level_matrix(h.A, d.A, w.A, sell);
      w.Av = d.A.v.p;
      d.dinv.upload(h.dinv.data(), h.dinv.size());
      w.dinv = d.dinv.p;
      d.b.alloc((size_t)h.A.n);
      w.b = d.b.p;
    }
    if (l < L - 1) {
      upload_level_matrix(h.P, d.P, w.P, sell);
      w.Pv = d.P.v.p;
      upload_level_matrix(h.R, d.R, w.R, sell);
      w.Rv = d.R.v.p;
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
  amg_stale_ = 0;
  amg_top_fresh_ = true;
  amg_built_ = true;
  ++amg_builds_;
  amg_setup_ms_ =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// FP32 copies of every level's operators, diagonal and coarse inverse (device
// conversions of the FP64 hierarchy) and FP32 work vectors; level 0 uses the
// FP32 pressure operator the mixed mode keeps in the Krylov layout
void NSSolver::ensure_amg32() {
  if (amg32_built_) return;
  const int L = (int)amg_lv_.size();
  amg_view32_.assign((size_t)L, AmgLevelView<float>{});
  auto conv = [&](const DBuf<double> &in, DBuf<float> &out) {
    out.alloc(in.n);
    if (in.n) convert<float, double>(stream_, (int64_t)in.n, in.p, out.p);
    return out.p;
  };
  for (int l = 0; l < L; ++l) {
    AmgDevLevel &d = amg_lv_[(size_t)l];
    const AmgLevelView<double> &w = amg_view_[(size_t)l];
    AmgLevelView<float> &f = amg_view32_[(size_t)l];
    f.n = w.n;
    f.c0 = w.c0;
    f.c1 = w.c1;
    f.c2 = w.c2;
    f.A = w.A;
    f.P = w.P;
    f.R = w.R;
    if (l > 0) {
      f.Av = conv(d.A.v, d.A.v32);
      f.dinv = conv(d.dinv, d.dinv32);
      d.b32.alloc((size_t)w.n);
      f.b = d.b32.p;
    } else {
      d.b32.alloc((size_t)w.n);  // FP32 copy of the CG residual
      f.b = d.b32.p;
      if (amg_top_fresh_) conv(d.A.v, d.A.v32);
    }
    if (l < L - 1) {
      f.Pv = conv(d.P.v, d.P.v32);
      f.Rv = conv(d.R.v, d.R.v32);
    }
    d.x32.alloc((size_t)w.n);
    d.y32.alloc((size_t)w.n);
    d.r32.alloc((size_t)w.n);
    f.x = d.x32.p;
    f.y = d.y32.p;
    f.r = d.r32.p;
  }
  if (!amg_cjac_) conv(amg_cinv_, amg_cinv32_);
  CFSA_HIP(hipStreamSynchronize(stream_));
  amg32_built_ = true;
}

template <>
AmgLevelView<double> *NSSolver::amg_views<double>() {
  AmgLevelView<double> &top = amg_view_[0];
  top.A = amg_fine_A();
  top.Av = amg_fine_v();
  top.dinv = d_invdL.p;
  top.b = w1_.r.p;  // the V-cycle reads the CG residual itself
  return amg_view_.data();
}
template <>
AmgLevelView<float> *NSSolver::amg_views<float>() {
  if (d_Lbc32.n == 0) throw std::logic_error("FP32 V-cycle without the FP32 pressure operator");
  ensure_amg32();
  AmgLevelView<float> &top = amg_view32_[0];
  top.A = amg_fine_A();
  top.Av = amg_top_fresh_ ? amg_lv_[0].A.v32.p : d_Lbc32.p;
  top.dinv = d_invdL32.p;
  return amg_view32_.data();
}
template <>
const double *NSSolver::amg_cinv<double>() {
  return amg_cinv_.p;
}
template <>
const float *NSSolver::amg_cinv<float>() {
  return amg_cinv32_.p;
}

void NSSolver::amg_apply(const double *r, double *z, int precision) {
  if (!initialized) throw std::runtime_error("amg_apply needs an initialized solver");
  if (precision != 0 && precision != 1) throw std::invalid_argument("amg_apply: precision 0 or 1");
  const int saved = opt.precision;
  opt.precision = precision == 1 ? 1 : saved;
  try {
    prepare_operators();
  } catch (...) {
    opt.precision = saved;
    throw;
  }
  opt.precision = saved;
  ensure_amg();
  DBuf<double> db;
  db.upload(r, (size_t)nv);
  w1_.alloc((size_t)nv);
  int Ga = 0, Gc = 0;
  const int L = (int)amg_view_.size();
  // the first smoothing step goes with copying r into the CG residual buffer;
  // x goes to a scratch vector
  if (precision == 1) {
    AmgLevelView<float> *lv = amg_views<float>();
    amg_cg_init<float>(stream_, nv, db.p, d_invdL.p, lv[0].dinv, lv[0].c0, w1_.s.p, w1_.r.p,
                       lv[0].b, lv[0].x, 0, d_part_a.p, &Ga);
    amg_vcycle<float>(stream_, lv, L, amg_cinv<float>(), amg_nc_, amg_cjac_, w1_.r.p, nullptr,
                      d_part_c.p, &Gc);
    amg_cg_p0<float>(stream_, nv, lv[0].y, w1_.t.p);
  } else {
    AmgLevelView<double> *lv = amg_views<double>();
    amg_cg_init<double>(stream_, nv, db.p, d_invdL.p, lv[0].dinv, lv[0].c0, w1_.s.p, w1_.r.p,
                        nullptr, lv[0].x, 0, d_part_a.p, &Ga);
    amg_vcycle<double>(stream_, lv, L, amg_cinv<double>(), amg_nc_, amg_cjac_, w1_.r.p, nullptr,
                       d_part_c.p, &Gc);
    CFSA_HIP(hipMemcpyAsync(w1_.t.p, lv[0].y, (size_t)nv * sizeof(double),
                            hipMemcpyDeviceToDevice, stream_));
  }
  CFSA_HIP(hipStreamSynchronize(stream_));
  CFSA_HIP(hipMemcpy(z, w1_.t.p, (size_t)nv * sizeof(double), hipMemcpyDeviceToHost));
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
