// Host-side launchers for the HIP kernels in kernels.hip.  All launches are
// asynchronous on the given stream.
#pragma once

#include "dev.hpp"

namespace cfsa {

// A device matrix pattern.  With sptr == nullptr it is CSR (rowptr, col) and
// value arrays are in CSR slot order.  With sptr set it is the sliced-ELL copy
// (col = per stored entry, sptr/rlen/perm as in SellPattern) and value arrays
// are in stored-entry order; the SpMV launchers pick the kernel from that.
// Element patches of the momentum operator on the device (elem_patch.hpp): an
// FP32 3-rhs SpMV on a DevCsr that carries them is applied tet by tet.
struct ElemPatchDev {
  int npatch = 0;
  const int *node_off = nullptr, *nodes = nullptr;
  const int *row_off = nullptr, *rows = nullptr;
  const unsigned short *starts = nullptr;
  const int *visit_off = nullptr;
  int64_t nvisit = 0, ntet = 0;
  const unsigned int *vstat = nullptr;  // 2 planes of 4 words per visit
  const float *vgeo = nullptr;          // 2 planes of 4 floats per tet
  const float *pen = nullptr;           // diagonal penalty of each owned row
};

// per-tet FP32 geometry and owned-row penalties of the element patches from
// the FP64 per-tet volumes and gradients (rho_dt = rho / dt, inv_mu = 1 / mu)
void elem_refresh(hipStream_t s, int64_t nt, int64_t nrows, const int *torder, const int *rows,
                  const double *vol, const double *grad, double rho_dt, double inv_mu,
                  const uint8_t *bc, double penalty, float *vgeo, float *pen);

struct DevCsr {
  int n = 0;
  int64_t nnz = 0;
  const int *rowptr = nullptr;
  const int *col = nullptr;
  const int *sptr = nullptr;
  const int *rlen = nullptr;
  const int *perm = nullptr;
  int nslices = 0;
  // CSR-stream windows (row starts, nwin + 1): 1 right-hand side, 3 (FP64),
  // 3 (FP32)
  const int *win1 = nullptr, *win3 = nullptr, *win3f = nullptr;
  int nwin1 = 0, nwin3 = 0, nwin3f = 0;
  // scratch of the CSR-stream inner products: pairs per (value, window)
  double *wpart = nullptr;
  int64_t wpart_cap = 0;  // values x windows the pair scratch holds
  const ElemPatchDev *elem = nullptr;  // host struct; set: the element-patch SpMV
};
// whether a three-rhs product with values of value_size bytes on A runs in the
// element-patch kernel (FP32 values on an operator that carries patches)
inline bool elem_product(const DevCsr &A, size_t value_size) {
  return value_size == 4 && A.elem != nullptr;
}

// Nonzeros staged per CSR-stream window (one block), measured best at 96^3
// with the fused inner products: 3 components 1024 (FP64, 24 KiB of LDS) or
// 1536 (FP32, 18 KiB); one component 2048.
template <class T, int KO>
constexpr int kStreamB = KO == 3 ? (sizeof(T) == 4 ? 1536 : 1024) : 2048;
// rows per CSR-stream window: the window's rowptr slice is staged in LDS, one
// entry per thread of the block
constexpr int kStreamRows = kBlock - 1;

// out[e] = src[e] >= 0 ? (T)in[src[e]] : 0 over the stored entries
template <class T>
void sell_fill(hipStream_t s, int64_t stored, const int *src, const double *in, T *out);

// ---- CSR-stream SpMV family (3 interleaved right-hand sides unless noted) ----
template <class T>
void spmv3(hipStream_t s, const DevCsr &A, const T *val, const T *x, T *y,
           const int *active);
// v = A y and per-rhs dot r0.v -> partials[3][G]
template <class T>
void spmv3_dot(hipStream_t s, const DevCsr &A, const T *val, const T *y, T *v,
               const T *r0, const int *active, double *partials, int *grid);
// t = A z and per-rhs dots t.t, t.s -> partials[3*2][G]
template <class T>
void spmv3_dot2(hipStream_t s, const DevCsr &A, const T *val, const T *z, T *t,
                const T *sv, const int *active, double *partials, int *grid);
// z = g0 w - g1 D^-1 (A w) with w = D^-1 r, pairs [3*2][G]: the stopping norm
// of r (|w|^2 when scaled, else |r|^2) and r.z; w == nullptr (element-patch
// products only): w is formed from r while x is staged
// wsep (degree 2's second product): the product's input is u, and w = D^-1 r
// comes from wv, or is formed from r when wv == nullptr (element-patch only)
template <class T>
void spmv3_polyz(hipStream_t s, const DevCsr &A, const T *val, const T *w, T *z, const T *r,
                 const T *invd, double g0, double g1, int scaled, const int *active,
                 double *partials, int *grid, int wsep = 0, const T *wv = nullptr);
// u = a1 w + a2 D^-1 (A w) with w = D^-1 r (the degree-2 preconditioner's first
// product); w == nullptr (element-patch only): w is formed from r while staged
template <class T>
void spmv3_polyu(hipStream_t s, const DevCsr &A, const T *val, const T *w, T *u, const T *r,
                 const T *invd, double a1, double a2, const int *active);
// r = b - A x (rhs of the active mask only), r0 = r, partials r.r [3][G]
template <class T>
void spmv3_residual(hipStream_t s, const DevCsr &A, const T *val, const T *x,
                    const T *b, T *r, T *r0, const T *sc, const int *active, double *partials,
                    int *grid);
// single right-hand side variants
template <class T>
void spmv1(hipStream_t s, const DevCsr &A, const T *val, const T *x, T *y);
template <class T>
void spmv1_dot(hipStream_t s, const DevCsr &A, const T *val, const T *y, T *v,
               const T *r0, const int *active, double *partials, int *grid);
template <class T>
void spmv1_dot2(hipStream_t s, const DevCsr &A, const T *val, const T *z, T *t,
                const T *sv, const int *active, double *partials, int *grid);
template <class T>
void spmv1_residual(hipStream_t s, const DevCsr &A, const T *val, const T *x,
                    const T *b, T *r, T *r0, const T *sc, const int *active, double *partials,
                    int *grid);
// y[i][c] = sum_j G_c[i,j] x[j]   (G values interleaved [nnz][3])
void spmv_grad3(hipStream_t s, const DevCsr &P, const double *gval, const double *x,
                double *y);
// momentum right-hand side, reference NavierStokes.cpp:430-495:
// rhs = (M u)/dt - [G p] (P1 rows) + conv + penalty * u_bc (Inlet/Wall rows)
void momentum_rhs(hipStream_t s, const DevCsr &A, const double *M, const double *u,
                  const double *gp, const double *conv, const uint8_t *bc,
                  const double *ginlet, double dt, int nv, double penalty,
                  double *rhs);

// adjoint building blocks (transposed G values gT on the symmetric P1 pattern)
// y = beta y + alpha sum_c G_c^T x_c   (x of stride xs, components 0..2)
void spmv_gradT_in3(hipStream_t s, const DevCsr &P, const double *gT, const double *x, int xs,
                    double *y, double alpha, double beta);
// y[i][c] += sum_j G_c[i,j] x[j] on the P1 rows of a [*][3] array
void spmv_grad3_acc(hipStream_t s, const DevCsr &P, const double *g, const double *x,
                    double *y);
// lu = (M lb)/dt; lg += pen lb on Inlet rows
void momentum_rhs_adjoint(hipStream_t s, const DevCsr &A, const double *M, const double *lb,
                          const uint8_t *bc, double dt, double pen, double *lu, double *lg);

// r = r0 = b (residual of x = 0) and the pairs of its stopping norm [K][G]
template <int K>
void residual_x0(hipStream_t s, int n, const double *b, double *r, double *r0, const double *sc,
                 double *partials, int *grid);

// ---- Krylov vector / scalar kernels ----
int vec_grid(int n);
template <class T, int K>
void bicg_init(hipStream_t s, int n, const T *b, T *x, T *r, T *r0, T *p, T *v,
               const T *stop_scale, double *partials, int *grid);
template <int K>
void bicg_s_init(hipStream_t s, KrylovState *st, const double *partials, int G);
template <int K>
void bicg_s_begin(hipStream_t s, KrylovState *st, const double *partials, int G);
template <int K>
void bicg_s_alpha(hipStream_t s, KrylovState *st, const double *partials, int G);
template <int K>
void bicg_s_omega(hipStream_t s, KrylovState *st, const double *partials, int G);
template <int K>
void bicg_s_restart(hipStream_t s, KrylovState *st, const double *partials, int G,
                    const int *mask);
template <class T, int K>
void bicg_v1(hipStream_t s, int n, const KrylovState *st, const T *r, T *p,
             const T *v, const T *invd, T *y);
template <class T, int K>
void bicg_v2(hipStream_t s, int n, const KrylovState *st, const T *r, const T *v,
             const T *invd, T *sv, T *z);
template <class T, int K>
void bicg_v3(hipStream_t s, int n, const KrylovState *st, const T *y, const T *z,
             const T *sv, const T *t, const T *r0, T *x, T *r, const T *stop_scale,
             double *partials, int *grid);
// Jacobi PCG (single right-hand side)
template <class T, int K>
void cg_init(hipStream_t s, int n, const T *b, const T *invd, T *x, T *r, T *p, int scaled,
             double *partials, int *grid);
template <int K>
void cg_s_init(hipStream_t s, KrylovState *st, const double *partials, int G);
// PCG iteration after the SpMV: r and the stopping / rho products (reducing
// the p.Ap pairs itself), then the convergence test, x and p updates
template <class T, int K>
void cg_xr(hipStream_t s, int n, KrylovState *st, const double *pq_partials, int Gq, int par,
           const T *q, const T *invd, T *r, int scaled, double *partials, int *grid);
template <class T, int K>
void cg_p(hipStream_t s, int n, KrylovState *st, const double *partials, int Gv, int par,
          const T *r, const T *invd, T *p, T *x, const T *zin = nullptr);
// CG with the polynomial preconditioner M^-1 = (g0 I - g1 D^-1 A) D^-1: x = 0,
// r = b, w = D^-1 b; then per iteration r -= alpha q, w = D^-1 r (cgp_xr) and
// z = M^-1 r with the stopping and r.z pairs (spmv3_polyz), p = z + beta p (cg_p
// with zin = z)
template <class T, int K>
void cgp_init(hipStream_t s, int n, const T *b, const T *invd, T *x, T *r, T *w);
template <class T, int K>
void cgp_xr(hipStream_t s, int n, KrylovState *st, const double *pq_partials, int Gq, int par,
            const T *q, const T *invd, T *r, T *w);  // w == nullptr: not stored
// power iteration on D^-1 A over three columns: x0, v = D^-1 y with |v|^2 pairs
// [3][G] (returns G), x = v * sc[column]
void power_init3(hipStream_t s, int n, double *x);
int power_v3(hipStream_t s, int n, const double *y, const double *invd, double *v,
             double *partials);
void power_scale3(hipStream_t s, int n, const double *v, const double *sc, double *x);

// ---- AMG-preconditioned CG (single right-hand side) ----
// One level of the smoothed-aggregation hierarchy on the device, in the V-cycle's
// value type T (double; float for the mixed precision = 1 mode).  Level 0's A
// is the pressure operator in its Krylov layout (sliced-ELL or CSR-stream), its
// b the CG residual (T = double) or its T copy, and its y the preconditioned
// residual z; coarser levels' A, and every level's P (n x nc) and R (nc x n),
// are CSR-stream (sliced-ELL for rows too long for a window).
template <class T>
struct AmgLevelView {
  int n = 0;
  DevCsr A, P, R;
  const T *Av = nullptr, *Pv = nullptr, *Rv = nullptr;
  const T *dinv = nullptr;
  T *b = nullptr, *x = nullptr, *y = nullptr, *r = nullptr;
  double c0 = 0.0, c1 = 0.0, c2 = 0.0;
};
// y0 = V(b0): pre-smoothing (x0 = c0 D^-1 b0 must be in lv[0].x), residual,
// restriction, coarser levels, prolongation, post-smoothing; the coarsest
// level is the dense inverse cinv (nc x nc) or, with coarse_jacobi, D^-1.
// Emits the pairs of rdot . y0 (rdot: the FP64 CG residual) into partials
// (*grid per value).  mask: the solve's active flag (nullptr: always run).
template <class T>
void amg_vcycle(hipStream_t s, const AmgLevelView<T> *lv, int nlev, const T *cinv, int nc,
                bool coarse_jacobi, const double *rdot, const int *mask, double *partials,
                int *grid);
// *out = order-independent 64-bit fingerprint of v[0..n) (not for use inside a
// captured graph: it clears *out with a memset)
void fingerprint(hipStream_t s, int64_t n, const double *v, uint64_t *out);
// x = 0, r = b, b0 = (T) b (skipped when b0 is null), x0 = c0 D0^-1 b and the
// stopping-norm pairs of b (invd: the FP64 Jacobi diagonal; dinv0: level 0's in T)
template <class T>
void amg_cg_init(hipStream_t s, int n, const double *b, const double *invd, const T *dinv0,
                 double c0, double *x, double *r, T *b0, T *x0, int scaled, double *partials,
                 int *grid);
// state from the stopping norm of b (pa) and rho = b . z0 (pc)
void amg_cg_s_init(hipStream_t s, KrylovState *st, const double *pa, int Ga, const double *pc,
                   int Gc);
// p = z0
template <class T>
void amg_cg_p0(hipStream_t s, int n, const T *z, double *p);
// alpha, r -= alpha q, b0 = (T) r, x0 = c0 D0^-1 r, stopping-norm pairs of r
template <class T>
void amg_cg_xr(hipStream_t s, int n, KrylovState *st, const double *pq, int Gq, int par,
               const double *q, const double *invd, const T *dinv0, double c0, double *r, T *b0,
               T *x0, int scaled, double *partials, int *grid);
// convergence test, x += alpha p, p = z + beta p
template <class T>
void amg_cg_p(hipStream_t s, int n, KrylovState *st, const double *pa, int Ga, const double *pc,
              int Gc, int par, const T *z, double *p, double *x);

// ---- FEM element / node kernels ----
void assemble_colour(hipStream_t s, int ntc, const int *order, const int *tets,
                     const int *ten, int nv, const double *vol, const double *grad,
                     double rho, double inv_mu, const DevCsr &A, double *M, double *K,
                     const DevCsr &P, double *L, double *G3);
void assemble_rows(hipStream_t s, int nu, const int *incp, const int *inct, const int *tets,
                   const int *ten, int nv, const double *vol, const double *grad, double rho,
                   double inv_mu, const DevCsr &A, double *M, double *K, const DevCsr &P,
                   double *L, double *G3);
// row gather through the slot map, one wavefront per row with LDS
// accumulators; also writes ilm.  Returns false when the longest row does not
// fit (maxA / maxP: longest P2 / P1 row).
bool assemble_waves(hipStream_t s, int nu, int nv, int maxA, int maxP, const int *incp,
                    const int *inct, const uint8_t *loc, const int *amap, const int *pmap,
                    const DevCsr &A, const DevCsr &P, const double *vol, const double *grad,
                    double rho, double inv_mu, double *M, double *K, double *L, double *G3,
                    double *ilm);
void build_slot_map(hipStream_t s, int nu, const int *incp, const int *inct, const int *tets,
                    const int *ten, int nv, const DevCsr &A, const DevCsr &P, uint8_t *loc,
                    int *amap, int *pmap);
void assemble_mapped(hipStream_t s, int nu, int nv, const int *incp, const int *inct,
                     const uint8_t *loc, const int *amap, const int *pmap, const double *vol,
                     const double *grad, double rho, double inv_mu, double *M, double *K,
                     double *L, double *G3);
void tet_geometry(hipStream_t s, int nt, const double *xyz, const int *tets, double *vol,
                  double *grad);
void lumped_mass_inverse(hipStream_t s, const DevCsr &A, const double *M, double *ilm);
void build_lhs(hipStream_t s, int64_t nnz, const double *M, const double *K, double dt,
               double *lhs);
void add_diag_penalty(hipStream_t s, int n, const int *diag, const uint8_t *bc,
                      int mode, int ref, double penalty, double *A);
void inv_diag(hipStream_t s, int n, const int *diag, const double *A, double *invd);
void conv_tet(hipStream_t s, int nt, const int *tets, const double *grad,
              const double *u, double *cT);
void conv_gather(hipStream_t s, int nu, int nv, const int *inc_ptr, const int *inc_tet,
                 const double *vol, const double *cT, double *conv);
void div_tet(hipStream_t s, int nt, const int *tets, const double *grad,
             const double *vol, const double *u, double *wT);
// div[i] = sum_T w_T;  rhs_p = -div/dt (+ penalty * p_out on Dirichlet rows)
void div_gather_prhs(hipStream_t s, int nv, const int *inc_ptr, const int *inc_tet,
                     const double *wT, const uint8_t *bc, int outlet_mode, int ref,
                     double dt, double p_out, double penalty, double *div,
                     double *rhs_p);
void pressure_update(hipStream_t s, int nv, double *p, const double *pc, double relax_p,
                     const uint8_t *bc, int dirichlet, double p_out);
void pressure_outlet_reset(hipStream_t s, int nv, double *p, const uint8_t *bc, double p_out);
void pressure_normalize(hipStream_t s, int nv, double *p, int ref, double *scratch);
void correct_velocity(hipStream_t s, int nu, int nv, const uint8_t *bc,
                      const double *ginlet, const double *ilm, const double *gp,
                      const double *us, double dt, double relax_u, double *u);
void reset_velocity_bc(hipStream_t s, int nu, const uint8_t *bc, const double *ginlet,
                       double *u);
void set_inlet_nodes(hipStream_t s, int nu, const uint8_t *bc, const double *ginlet,
                     double *u);
void fill3(hipStream_t s, int n, const double v[3], double *out);
void nan_check(hipStream_t s, int n, const double *a, int *flag);
template <class TO, class TI>
void convert(hipStream_t s, int64_t n, const TI *in, TO *out);

}  // namespace cfsa
