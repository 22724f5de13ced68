// HIP kernels of the PISO step (gfx950, wave64).
//
// CSR SpMV is "CSR-stream": a block owns a fixed run of rows; the block's
// nonzeros are streamed in coalesced chunks, every product a_ij * x_j is
// staged in LDS, then the thread owning (row, component) sums its row's
// products in ascending column order.  That sequential order equals the
// accumulation order of Eigen's column-major sparse * dense product
// (res(i) += A(i,j) * x(j) over ascending j), so with FMA contraction off the
// FP64 SpMV reproduces the reference bit for bit.
#include "ops.hpp"

#include <cfloat>

#include "dd.hpp"
#include "elem_patch.hpp"
#include "pattern.hpp"

namespace cfsa {

// ===========================================================================
// CSR-stream core
// ===========================================================================
// Rows are cut into windows (host-built, see stream_windows) of at most
// kStreamB<KO> nonzeros and kStreamRows rows, one block per window.  The block
// first loads the window's columns and values (clamped indices, so every load
// is issued before any is used) and its rowptr slice, then gathers x and
// stages the products in LDS, then sums them per (row, component), kBlock / KO
// rows at a time.  Nothing is read from global memory between the barrier and
// the sums except the epilogue operands of rows past the first pass.
//
// Inner products: each block reduces its rows' pairs and stores them in
// wpart[v][window]; k_window_reduce then merges each group of kStreamGroup
// consecutive windows in a fixed tree order into partials[v][group].  (A
// last-block-done merge inside this kernel would need a device-scope release
// per block, which writes back the XCD's whole L2 and costs far more.)
constexpr int kStreamGroup = 256;
// Window order: blocks are dispatched round-robin over the 8 XCDs (block b on
// XCD b % 8).  Block b takes window ((q / C) * 8 + b % 8) * C + q % C with
// q = b / 8 and C = kStreamChunk: every XCD runs chunks of C consecutive
// windows, whose x gathers overlap in its L2, and the 8 XCDs work on
// neighbouring chunks.  Measured against window = block at 96^3: FP32 momentum
// SpMV 396 -> 390 us, FP64 residual SpMV 793 -> 777 us; cylinder 70 -> 68 us.
// The grid is rounded up to whole rounds of 8 C blocks; blocks past the last
// window exit at once.
constexpr int kStreamChunk = 16;

template <class T, int KO, int ND, class Op>
__global__ __launch_bounds__(kBlock) void k_csr_stream(const int *__restrict__ rowptr,
                                                       const int *__restrict__ win, int nwin,
                                                       Op op, double *__restrict__ wpart) {
  constexpr int B = kStreamB<T, KO>;
  constexpr int R = kBlock / KO;   // rows summed per pass
  constexpr int PER = B / kBlock;  // staged entries per thread
  static_assert(PER * kBlock == B, "whole staging rows per thread");
  constexpr int NDs = ND > 0 ? ND : 1;
  constexpr int NV = KO * ND;
  // staged products; reused for the block's inner-product pairs at the end
  __shared__ T lds[B * KO];
  __shared__ int rps[kBlock];  // rowptr slice of the window, relative to k0
  static_assert(sizeof(T) * B * KO >= sizeof(double) * 2 * NDs * kBlock, "LDS reuse");
  if (!op.any()) return;
  DD dots[NDs];
#pragma unroll
  for (int d = 0; d < NDs; ++d) dd_zero(dots[d]);
  const int tid = threadIdx.x;
  const int myc = tid % KO;
  const bool act = op.active(myc) && tid < KO * R;
  const int q = (int)(blockIdx.x >> 3);
  const int w = ((q / kStreamChunk) * 8 + (int)(blockIdx.x & 7)) * kStreamChunk + q % kStreamChunk;
  if (w >= nwin) return;
  const int r0 = win[w], r1 = win[w + 1];  // r1 - r0 <= kStreamRows
  const int k0 = rowptr[r0];
  const int len = rowptr[r1] - k0;  // <= B by construction of the windows
  const int rpv = rowptr[min(r0 + tid, r1)];
  const int row0 = r0 + tid / KO;
  typename Op::Pre pre0{};
  if (act && row0 < r1) pre0 = op.pre(row0, myc);
  if (len > 0) {
    typename Op::Frag f[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) f[u] = op.fetch(k0 + min(tid + u * kBlock, len - 1));
    T p[PER][KO];
#pragma unroll
    for (int u = 0; u < PER; ++u) op.prod(f[u], p[u]);
    // every slot i < B is written (PER * kBlock == B); slots past len hold
    // copies of the last product and are never summed
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = tid + u * kBlock;
#pragma unroll
      for (int c = 0; c < KO; ++c) lds[i * KO + c] = p[u][c];
    }
  }
  rps[tid] = rpv - k0;
  __syncthreads();
  for (int row = row0; row < r1; row += R) {
    if (act) {
      const typename Op::Pre pr = row == row0 ? pre0 : op.pre(row, myc);
      const int ra = rps[row - r0], rb = rps[row - r0 + 1];
      // sequential sum in ascending column order; LDS reads are batched
      T acc = T(0);
      int k = ra;
      for (; k + 8 <= rb; k += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = lds[(k + u) * KO + myc];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u];
      }
      for (; k < rb; ++k) acc += lds[k * KO + myc];
      op.epi(row, myc, acc, pr, dots);
    }
  }
  if constexpr (ND > 0) {
    double *rh = reinterpret_cast<double *>(lds);  // [ND][kBlock]
    double *rl = rh + ND * kBlock;                 // [ND][kBlock]
    __syncthreads();
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      rh[d * kBlock + tid] = dots[d].hi;
      rl[d * kBlock + tid] = dots[d].lo;
    }
    __syncthreads();
    // value v = (component c, product d) owns 32 lanes (half a wavefront):
    // lane l merges the threads c + KO * (l + 32 q), then a fixed shuffle tree
    static_assert(NV * 32 <= kBlock, "lanes per value");
    if (tid < NV * 32) {
      const int v = tid / 32, l = tid % 32;
      const int c = v / ND, d = v % ND;
      DD s;
      dd_zero(s);
      for (int t = c + KO * l; t < kBlock; t += KO * 32)
        dd_merge(s, rh[d * kBlock + t], rl[d * kBlock + t]);
#pragma unroll
      for (int st = 16; st > 0; st >>= 1) {
        const double oh = __shfl_down(s.hi, st, 32);
        const double ol = __shfl_down(s.lo, st, 32);
        dd_merge(s, oh, ol);
      }
      if (l == 0) {
        wpart[2 * ((size_t)v * nwin + w)] = s.hi;
        wpart[2 * ((size_t)v * nwin + w) + 1] = s.lo;
      }
    }
  }
}

// partials[v][g] = merge of wpart[v][g * kStreamGroup + 0 .. 255] (tree order)
template <int NV>
__global__ __launch_bounds__(kBlock) void k_window_reduce(int nwin,
                                                          const double *__restrict__ wpart,
                                                          double *__restrict__ partials) {
  static_assert(kStreamGroup == kBlock, "one window per thread");
  DD v[NV];
  const int w = blockIdx.x * kStreamGroup + threadIdx.x;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    dd_zero(v[i]);
    if (w < nwin) {
      v[i].hi = wpart[2 * ((size_t)i * nwin + w)];
      v[i].lo = wpart[2 * ((size_t)i * nwin + w) + 1];
    }
  }
  block_partials<NV>(v, partials);
}

// ===========================================================================
// Sliced-ELL core (the Krylov operators lhs and L_bc): one wavefront per slice
// of 64 rows, one row per lane.  Matrix values and columns of entry j of the
// 64 rows are contiguous, so every load of them is a full 64-lane coalesced
// access, and no LDS staging or block barrier is needed; a lane sums its row
// in CSR order (bit-identical to the CSR-stream kernel and to the reference).
// ===========================================================================
constexpr int kSellGrid = 2048;  // max blocks; always a multiple of 8
constexpr int kSellUnroll = 8;   // entries in flight per lane

static int sell_grid(int nslices) {
  constexpr int wpb = kBlock / kSellSlice;
  const int share = div_up(nslices, 8);
  const int per_xcd = std::max(1, std::min(div_up(share, wpb), kSellGrid / 8));
  return 8 * per_xcd;
}

template <class T, int KO, int ND, int U, class Op>
__global__ __launch_bounds__(kBlock) void k_sell(int n, int nslices,
                                                 const int *__restrict__ sptr,
                                                 const int *__restrict__ rlen,
                                                 const int *__restrict__ perm, Op op,
                                                 double *__restrict__ partials) {
  constexpr int NDs = ND > 0 ? ND : 1;
  constexpr int wpb = kBlock / kSellSlice;
  if (!op.any()) return;
  DD dots[KO * NDs];
#pragma unroll
  for (int i = 0; i < KO * NDs; ++i) dd_zero(dots[i]);
  bool act[KO];
#pragma unroll
  for (int c = 0; c < KO; ++c) act[c] = op.active(c);
  const int lane = threadIdx.x % kSellSlice;
  const int wv = threadIdx.x / kSellSlice;
  // Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8):
  // each XCD gets one contiguous share of the slices, so the x entries its
  // rows gather stay in that XCD's L2.
  const int xcd = blockIdx.x & 7, bx = blockIdx.x >> 3, nbx = gridDim.x >> 3;
  const int share = (nslices + 7) >> 3;
  const int s0 = xcd * share, s1 = min(nslices, s0 + share);
  for (int s = s0 + bx * wpb + wv; s < s1; s += nbx * wpb) {
    const int p = s * kSellSlice + lane;
    const int len = rlen[p];
    const int base = sptr[s] + lane;
    T acc[KO];
#pragma unroll
    for (int c = 0; c < KO; ++c) acc[c] = T(0);
    int j = 0;
    for (; j + U <= len; j += U) {
      T pr[U][KO];
#pragma unroll
      for (int u = 0; u < U; ++u) op.prod(op.fetch(base + (j + u) * kSellSlice), pr[u]);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int c = 0; c < KO; ++c) acc[c] += pr[u][c];
    }
    for (; j < len; ++j) {
      T pr[KO];
      op.prod(op.fetch(base + j * kSellSlice), pr);
#pragma unroll
      for (int c = 0; c < KO; ++c) acc[c] += pr[c];
    }
    if (p < n) {
      const int row = perm[p];
#pragma unroll
      for (int c = 0; c < KO; ++c)
        if (act[c]) op.epi(row, c, acc[c], op.pre(row, c), &dots[c * NDs]);
    }
  }
  if constexpr (ND > 0) block_partials<KO * ND>(dots, partials);
}

template <class T>
__global__ void k_sell_fill(int64_t stored, const int *__restrict__ src,
                            const double *__restrict__ in, T *__restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < stored;
       e += (int64_t)gridDim.x * kBlock) {
    const int k = src[e];
    out[e] = k >= 0 ? (T)in[k] : T(0);
  }
}
template <class T>
void sell_fill(hipStream_t s, int64_t stored, const int *src, const double *in, T *out) {
  if (stored <= 0) return;
  hipLaunchKernelGGL((k_sell_fill<T>), dim3(4096), dim3(kBlock), 0, s, stored, src, in, out);
  CFSA_LAUNCH_CHECK();
}
template void sell_fill<double>(hipStream_t, int64_t, const int *, const double *, double *);
template void sell_fill<float>(hipStream_t, int64_t, const int *, const double *, float *);

// ===========================================================================
// Element-patch SpMV (FP32 momentum operator, 3 right-hand sides)
// ===========================================================================
// One block per patch (elem_patch.hpp), in the XCD-chunked order of the
// CSR-stream windows.  The patch's x values are staged in LDS first (one
// gather per node instead of one per node and visit).  Phase 1: each thread
// takes visits (tets) of the patch, reads x at the tet's 10 nodes and applies
//   lhs_e = (rho V / dt) C + B A A^T B^T,  A = sqrt(V / mu) Phi,
// writing the contribution to each owned node into that (row, tet) pair's LDS
// slot.  Phase 2: the thread owning (row, component) sums the row's slots in
// ascending tet order, adds the row's diagonal penalty and hands the sum to
// the operator's epilogue (the same functors as the CSR kernels), so the
// product and its fused inner products are deterministic.
constexpr int kElemBlock = 512;

// n < 32 bits at bit b of a 128-bit little-endian word (b, n compile-time
// after unrolling)
__device__ __forceinline__ unsigned int bits128(const uint4 &w, int b, int n) {
  const unsigned int v[4] = {w.x, w.y, w.z, w.w};
  const int k = b >> 5, o = b & 31;
  unsigned int f = v[k] >> o;
  if (o + n > 32) f |= v[k + 1] << (32 - o);
  return f & ((1u << n) - 1u);
}
// the visit's tet (patch numbering): bits 0-17 at bit 110 of word A, bits
// 18-25 at bit 120 of word B
__device__ __forceinline__ size_t visit_tet(const uint4 &a, const uint4 &b) {
  return (size_t)((a.w >> 14) | ((b.w >> 24) << 18));
}

template <int ND, class Op>
__global__ __launch_bounds__(kElemBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_elem_spmv(ElemPatchDev E, Op op,
                                                          double *__restrict__ wpart) {
  __shared__ float cnt[kElemSlots * 3];
  __shared__ float xs[kElemNodes * 3];  // x at the patch's nodes
  __shared__ unsigned short st[kElemRows + 1];
  constexpr int NDs = ND > 0 ? ND : 1;
  constexpr int NV = 3 * ND;
  static_assert(sizeof(float) * kElemSlots * 3 >= sizeof(double) * 2 * NDs * kElemBlock,
                "LDS reuse");
  if (!op.any()) return;
  const int q = (int)(blockIdx.x >> 3);
  const int p = ((q / kStreamChunk) * 8 + (int)(blockIdx.x & 7)) * kStreamChunk + q % kStreamChunk;
  if (p >= E.npatch) return;
  const int tid = threadIdx.x;
  const int n0 = E.node_off[p], nn = E.node_off[p + 1] - n0;
  const int r0 = E.row_off[p], nr = E.row_off[p + 1] - r0;
  const int v0 = E.visit_off[p], nvis = E.visit_off[p + 1] - v0;
  const float *__restrict__ x = op.x;
  // Everything a thread reads from global memory is issued up front, in one
  // branch-free sequence (indices clamped into the patch, which holds at least
  // one row, node and visit), so the block's latencies overlap and no load
  // waits for another it does not depend on: its first visit record, the
  // staging node ids, its slot starts, the epilogue operands of its rows, the
  // first visit's geometry, then the staging gathers.  Each later visit record
  // is loaded while the previous visit is computed.
  constexpr int RP = kElemBlock / 3;  // rows per pass of the row sums
  constexpr int PMAX = (kElemRows + RP - 1) / RP;
  static_assert(kElemNodes % kElemBlock == 0, "whole staging passes");
  constexpr int NPT = kElemNodes / kElemBlock;                    // staged nodes per thread
  constexpr int SPT = (kElemRows + 1 + kElemBlock - 1) / kElemBlock;  // slot starts per thread
  const int myc = tid % 3;
  // the component flags are read at uniform addresses (scalar loads)
  const bool act0 = op.active(0), act1 = op.active(1), act2 = op.active(2);
  const bool rowthr = tid < 3 * RP && (myc == 0 ? act0 : myc == 1 ? act1 : act2);
  // visit records are stored as two planes (one 16-byte word per visit each),
  // so every load of a wavefront reads one contiguous run; the geometry is
  // per tet (two planes of one float4 per tet), gathered by the record's tet
  const size_t nvt = (size_t)E.nvisit, ntt = (size_t)E.ntet;
  const uint4 *vs = reinterpret_cast<const uint4 *>(E.vstat);
  const float4 *vg = reinterpret_cast<const float4 *>(E.vgeo);
  // the records of the thread's first two visits (a patch has at most 2 x
  // kElemBlock visits on a Kuhn cube; later passes load their own)
  uint4 sa0, sb0, sa1, sb1;
  {
    const size_t va = (size_t)(v0 + min(tid, nvis - 1));
    const size_t vb = (size_t)(v0 + min(tid + kElemBlock, nvis - 1));
    sa0 = vs[va];
    sb0 = vs[nvt + va];
    sa1 = vs[vb];
    sb1 = vs[nvt + vb];
  }
  int jj[NPT];
#pragma unroll
  for (int u = 0; u < NPT; ++u) jj[u] = E.nodes[n0 + min(tid + u * kElemBlock, nn - 1)];
  unsigned short stv[SPT];
#pragma unroll
  for (int u = 0; u < SPT; ++u) stv[u] = E.starts[r0 + p + min(tid + u * kElemBlock, nr)];
  typename Op::PreE pre2[PMAX];
  float pen2[PMAX];
#pragma unroll
  for (int k = 0; k < PMAX; ++k) {
    const int i = min(tid / 3 + k * RP, nr - 1);
    pre2[k] = op.pre_e(r0 + i, myc);
    pen2[k] = E.pen[r0 + i];
  }
  const size_t ta = visit_tet(sa0, sb0), tb = visit_tet(sa1, sb1);
  const float4 ga0 = vg[ta], gb0 = vg[ntt + ta], ga1 = vg[tb], gb1 = vg[ntt + tb];
  // the patch's x values, one gather per node (a node serves ~14 visits);
  // an operator may stage x scaled per node (x = D^-1 r formed here)
  {
    float xv[NPT][3];
#pragma unroll
    for (int u = 0; u < NPT; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) xv[u][c] = x[(size_t)jj[u] * 3 + c];
    if constexpr (Op::kStageScale) {
      if (op.sscale) {
        float sv[NPT];
#pragma unroll
        for (int u = 0; u < NPT; ++u) sv[u] = op.sscale[jj[u]];
#pragma unroll
        for (int u = 0; u < NPT; ++u)
#pragma unroll
          for (int c = 0; c < 3; ++c) xv[u][c] = (float)((double)sv[u] * (double)xv[u][c]);
      }
    }
#pragma unroll
    for (int u = 0; u < NPT; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) xs[(tid + u * kElemBlock) * 3 + c] = xv[u][c];
  }
#pragma unroll
  for (int u = 0; u < SPT; ++u)
    if (tid + u * kElemBlock <= nr) st[tid + u * kElemBlock] = stv[u];
  __syncthreads();
  // One visit (a thread past the last visit computes a clamped copy and
  // writes nothing, so every thread runs the same passes).
  auto visit = [&](const uint4 &ca, const uint4 &cb, const float4 &gA, const float4 &gB,
                   bool valid) {
    // 11-bit local ids at bit 11 l of record word A, 12-bit LDS slots at bit
    // 12 l of record word B (elem_patch.hpp)
    int lid[10], slot[10];
#pragma unroll
    for (int l = 0; l < 10; ++l) {
      lid[l] = (int)bits128(ca, 11 * l, 11);
      slot[l] = (int)bits128(cb, 12 * l, 12);
    }
    float xe[10][3];
#pragma unroll
    for (int l = 0; l < 10; ++l) {
#pragma unroll
      for (int c = 0; c < 3; ++c) xe[l][c] = xs[lid[l] * 3 + c];
    }
    // rho V / dt and the Gram matrix G_kl = (V / mu) grad phi_k . grad phi_l (k, l = 1..3)
    const float m = gA.x;
    const float g11 = gA.y, g12 = gA.z, g13 = gA.w, g22 = gB.x, g23 = gB.y, g33 = gB.z;
    float ye[10][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      // FP32 operator of the inner solves: products and sums may contract to
      // FMAs here (the library builds with contraction off for the FP64 path)
#pragma clang fp contract(fast)
      const float x0 = xe[0][c], x1 = xe[1][c], x2 = xe[2][c], x3 = xe[3][c];
      const float e0 = xe[4][c], e1 = xe[5][c], e2 = xe[6][c], e3 = xe[7][c], e4 = xe[8][c],
                  e5 = xe[9][c];
      const float sv = (x0 + x1) + (x2 + x3);
      const float se = ((e0 + e1) + (e2 + e3)) + (e4 + e5);
      const float z0 = x0 + e0 + e1 + e2, z1 = x1 + e0 + e3 + e4, z2 = x2 + e1 + e3 + e5,
                  z3 = x3 + e2 + e4 + e5;
      const float d1 = z1 - z0, d2 = z2 - z0, d3 = z3 - z0;
      const float s1 = g11 * d1 + g12 * d2 + g13 * d3;
      const float s2 = g12 * d1 + g22 * d2 + g23 * d3;
      const float s3 = g13 * d1 + g23 * d2 + g33 * d3;
      const float s0 = -((s1 + s2) + s3);
      // vertex rows: m (x_a / 20 + Sv / 20 + Se / 30); edge rows: m (x_k / 20 + Se / 60 + Sv / 30)
      const float mv = m * (sv * (1.0f / 20.0f) + se * (1.0f / 30.0f));
      const float me = m * (se * (1.0f / 60.0f) + sv * (1.0f / 30.0f));
      const float m20 = m * (1.0f / 20.0f);
      ye[0][c] = m20 * x0 + mv + s0;
      ye[1][c] = m20 * x1 + mv + s1;
      ye[2][c] = m20 * x2 + mv + s2;
      ye[3][c] = m20 * x3 + mv + s3;
      ye[4][c] = m20 * e0 + me + (s0 + s1);
      ye[5][c] = m20 * e1 + me + (s0 + s2);
      ye[6][c] = m20 * e2 + me + (s0 + s3);
      ye[7][c] = m20 * e3 + me + (s1 + s2);
      ye[8][c] = m20 * e4 + me + (s1 + s3);
      ye[9][c] = m20 * e5 + me + (s2 + s3);
    }
#pragma unroll
    for (int l = 0; l < 10; ++l) {
      if (valid && lid[l] < nr) {
#pragma unroll
        for (int c = 0; c < 3; ++c) cnt[slot[l] * 3 + c] = ye[l][c];
      }
    }
  };
  const int npass = (nvis + kElemBlock - 1) / kElemBlock;  // block-uniform
  visit(sa0, sb0, ga0, gb0, tid < nvis);
  if (npass > 1) visit(sa1, sb1, ga1, gb1, tid + kElemBlock < nvis);
  for (int pass = 2; pass < npass; ++pass) {
    const int v = tid + pass * kElemBlock;
    const size_t vi = (size_t)(v0 + min(v, nvis - 1));
    const uint4 ca = vs[vi], cb = vs[nvt + vi];
    const size_t t = visit_tet(ca, cb);
    visit(ca, cb, vg[t], vg[ntt + t], v < nvis);
  }
  __syncthreads();
  // inner products in plain FP64: the operands are FP32 values, whose
  // products are exact in FP64, and the sums run in a fixed order
  double dots[NDs];
#pragma unroll
  for (int d = 0; d < NDs; ++d) dots[d] = 0.0;
  // all of the thread's row sums first, then their epilogues, so no LDS read
  // of a later row waits behind an earlier row's global stores; the row's own
  // x value comes from the staged copy (owned rows lead the node table)
  float accs[PMAX], xrow[PMAX];
#pragma unroll
  for (int k = 0; k < PMAX; ++k) {
    const int i = tid / 3 + k * RP;
    float acc = 0.0f, px = 0.0f;
    if (rowthr && i < nr) {
      // the row's slots in ascending tet order, LDS reads batched by 8 and 4
      int sl = st[i];
      const int se = st[i + 1];
      px = xs[i * 3 + myc];
      for (; sl + 8 <= se; sl += 8) {
        float c8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) c8[u] = cnt[(sl + u) * 3 + myc];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += c8[u];
      }
      if (sl + 4 <= se) {
        float c4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c4[u] = cnt[(sl + u) * 3 + myc];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += c4[u];
        sl += 4;
      }
      for (; sl < se; ++sl) acc += cnt[sl * 3 + myc];
      if (pen2[k] != 0.0f) acc += pen2[k] * px;
    }
    accs[k] = acc;
    xrow[k] = px;
  }
  if (rowthr) {
#pragma unroll
    for (int k = 0; k < PMAX; ++k) {
      const int i = tid / 3 + k * RP;
      // rows are numbered in patch order
      if (i < nr) op.epi_e(r0 + i, myc, accs[k], pre2[k], xrow[k], dots);
    }
  }
  if constexpr (ND > 0) {
    double *rh = reinterpret_cast<double *>(cnt);  // [ND][kElemBlock]
    __syncthreads();
#pragma unroll
    for (int d = 0; d < ND; ++d) rh[d * kElemBlock + tid] = dots[d];
    __syncthreads();
    // value v = (component c, product d) owns 32 lanes: lane l sums the
    // threads c + 3 (l + 32 q), then a fixed shuffle tree; the block's pair
    // is (sum, 0) in the double-double partial format
    if (tid < NV * 32) {
      const int v = tid / 32, l = tid % 32;
      const int c = v / ND, d = v % ND;
      double sum = 0.0;
      for (int t = c + 3 * l; t < kElemBlock; t += 3 * 32) sum += rh[d * kElemBlock + t];
#pragma unroll
      for (int sh = 16; sh > 0; sh >>= 1) sum += __shfl_down(sum, sh, 32);
      if (l == 0) {
        wpart[2 * ((size_t)v * E.npatch + p)] = sum;
        wpart[2 * ((size_t)v * E.npatch + p) + 1] = 0.0;
      }
    }
  }
}

// Per-tet FP32 geometry (rho V / dt and the Gram matrix (V / mu) grad phi_k .
// grad phi_l, k, l = 1..3, rounded from FP64 products of the FP64
// per-tet arrays; two planes of one float4 per tet) and each owned row's
// diagonal penalty (Inlet / Wall).
__global__ void k_elem_refresh(int64_t nt, int64_t nrows, const int *__restrict__ torder,
                               const int *__restrict__ rows, const double *__restrict__ vol,
                               const double *__restrict__ grad, double rho_dt, double inv_mu,
                               const uint8_t *__restrict__ bc, float penalty,
                               float4 *__restrict__ vgeo, float *__restrict__ pen) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < nt; j += stride) {
    const int64_t t = torder[j];  // tet j of the patch numbering
    const double V = vol[t];
    const double sc = V * inv_mu;
    const double *g = grad + (size_t)t * 12;
    auto G = [&](int k, int l) {
      const double *a = g + 3 * k, *b = g + 3 * l;
      return (float)(sc * (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]));
    };
    vgeo[j] = make_float4((float)(rho_dt * V), G(1, 1), G(1, 2), G(1, 3));
    vgeo[nt + j] = make_float4(G(2, 2), G(2, 3), G(3, 3), 0.0f);
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nrows; i += stride) {
    const uint8_t b = bc[rows[i]];
    pen[i] = (b == kInlet || b == kWall) ? penalty : 0.0f;
  }
}
void elem_refresh(hipStream_t s, int64_t nt, int64_t nrows, const int *torder, const int *rows,
                  const double *vol, const double *grad, double rho_dt, double inv_mu,
                  const uint8_t *bc, double penalty, float *vgeo, float *pen) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_elem_refresh, dim3(4096), dim3(kBlock), 0, s, nt, nrows, torder, rows, vol,
                     grad, rho_dt, inv_mu, bc, (float)penalty, reinterpret_cast<float4 *>(vgeo),
                     pen);
  CFSA_LAUNCH_CHECK();
}

// Launches the sliced-ELL kernel when A carries that layout, else CSR-stream
// (or the element-patch kernel for an FP32 3-rhs operator that carries
// patches); returns the grid size (the number of partial pairs per value).
template <class T, int KO, int ND, class Op>
static int launch_spmv(hipStream_t s, const DevCsr &A, const Op &op, double *partials) {
  if constexpr (sizeof(T) == 4 && KO == 3) {
    if (A.elem) {
      const ElemPatchDev &E = *A.elem;
      const int G = (E.npatch + kStreamGroup - 1) / kStreamGroup;
      if (ND > 0 && (!A.wpart || (int64_t)KO * ND * E.npatch > A.wpart_cap))
        throw std::logic_error("element-patch inner products need window scratch");
      if (G > kMaxPartials)
        throw std::runtime_error("element patches: too many for the partial buffers");
      constexpr int round = 8 * kStreamChunk;
      const int grid = (E.npatch + round - 1) / round * round;
      hipLaunchKernelGGL((k_elem_spmv<ND, Op>), dim3(grid), dim3(kElemBlock), 0, s, E, op,
                         A.wpart);
      CFSA_LAUNCH_CHECK();
      if constexpr (ND > 0) {
        hipLaunchKernelGGL((k_window_reduce<KO * ND>), dim3(G), dim3(kBlock), 0, s, E.npatch,
                           A.wpart, partials);
        CFSA_LAUNCH_CHECK();
      }
      return G;
    }
  }
  if (A.sptr) {
    const int G = sell_grid(A.nslices);
    hipLaunchKernelGGL((k_sell<T, KO, ND, kSellUnroll, Op>), dim3(G), dim3(kBlock), 0, s, A.n,
                       A.nslices, A.sptr, A.rlen, A.perm, op, partials);
    CFSA_LAUNCH_CHECK();
    return G;
  }
  const int *win = KO == 1 ? A.win1 : (sizeof(T) == 4 ? A.win3f : A.win3);
  const int nwin = KO == 1 ? A.nwin1 : (sizeof(T) == 4 ? A.nwin3f : A.nwin3);
  if (!win || nwin < 1) throw std::logic_error("CSR pattern without stream windows");
  const int G = (nwin + kStreamGroup - 1) / kStreamGroup;
  if (ND > 0 && (!A.wpart || (int64_t)KO * ND * nwin > A.wpart_cap))
    throw std::logic_error("CSR-stream inner products need window scratch");
  if (G > kMaxPartials)
    throw std::runtime_error("CSR-stream: matrix too large for the partial buffers");
  constexpr int round = 8 * kStreamChunk;
  const int grid = (nwin + round - 1) / round * round;
  hipLaunchKernelGGL((k_csr_stream<T, KO, ND, Op>), dim3(grid), dim3(kBlock), 0, s, A.rowptr,
                     win, nwin, op, A.wpart);
  CFSA_LAUNCH_CHECK();
  if constexpr (ND > 0) {
    // (a fully masked launch leaves stale pairs; its consumers skip them)
    hipLaunchKernelGGL((k_window_reduce<KO * ND>), dim3(G), dim3(kBlock), 0, s, nwin, A.wpart,
                       partials);
    CFSA_LAUNCH_CHECK();
  }
  return G;
}

// ---- operator functors -----------------------------------------------------
// Pre: epilogue operands of a (row, component), loaded before the row's sum
// so their latency overlaps the staging of the products.
struct MaskBase {
  struct Pre {};
  __device__ Pre pre(int, int) const { return {}; }
  const int *mask;
  __device__ bool active(int c) const { return mask == nullptr || mask[c] != 0; }
  __device__ bool any() const {
    return mask == nullptr || (mask[0] | mask[1] | mask[2]) != 0;
  }
};
struct Mask1Base {
  struct Pre {};
  __device__ Pre pre(int, int) const { return {}; }
  const int *mask;
  __device__ bool active(int) const { return mask == nullptr || mask[0] != 0; }
  __device__ bool any() const { return mask == nullptr || mask[0] != 0; }
};

template <class T, int KO, class Base>
struct OpAx : Base {
  const int *__restrict__ col;
  const T *__restrict__ val;
  const T *__restrict__ x;
  T *__restrict__ y;
  struct Frag {
    int j;
    T a;
  };
  __device__ Frag fetch(int k) const { return Frag{col[k], val[k]}; }
  __device__ void prod(const Frag &f, T *p) const {
#pragma unroll
    for (int c = 0; c < KO; ++c) p[c] = f.a * x[(size_t)f.j * KO + c];
  }
  __device__ void epi(int row, int c, T acc, const typename Base::Pre &, DD *) const {
    y[(size_t)row * KO + c] = acc;
  }
  // element-patch kernel: PreE / pre_e / epi_e take the row's x value from
  // the kernel's staged copy and accumulate inner products in plain FP64
  struct PreE {};
  __device__ PreE pre_e(int, int) const { return {}; }
  __device__ void epi_e(int row, int c, T acc, const PreE &, T, double *) const {
    y[(size_t)row * KO + c] = acc;
  }
  static constexpr bool kStageScale = false;
};

template <class T, int KO, class Base>
struct OpAxDot : OpAx<T, KO, Base> {
  const T *__restrict__ r0;
  struct Pre {
    T r0;
  };
  __device__ Pre pre(int row, int c) const { return Pre{r0[(size_t)row * KO + c]}; }
  __device__ void epi(int row, int c, T acc, const Pre &pr, DD *d) const {
    this->y[(size_t)row * KO + c] = acc;
    dd_fma(d[0], (double)pr.r0, (double)acc);
  }
  // (CG's p.Ap passes r0 == x: the operand is then not loaded again)
  using PreE = Pre;
  __device__ PreE pre_e(int row, int c) const {
    return PreE{r0 == this->x ? T(0) : r0[(size_t)row * KO + c]};
  }
  __device__ void epi_e(int row, int c, T acc, const PreE &pr, T xr, double *d) const {
    this->y[(size_t)row * KO + c] = acc;
    const T o = r0 == this->x ? xr : pr.r0;
    d[0] = fma((double)o, (double)acc, d[0]);
  }
};

template <class T, int KO, class Base>
struct OpAxDot2 : OpAx<T, KO, Base> {
  const T *__restrict__ sv;
  struct Pre {
    T sv;
  };
  __device__ Pre pre(int row, int c) const { return Pre{sv[(size_t)row * KO + c]}; }
  __device__ void epi(int row, int c, T acc, const Pre &pr, DD *d) const {
    this->y[(size_t)row * KO + c] = acc;
    const double t = (double)acc;
    dd_fma(d[0], t, t);
    dd_fma(d[1], t, (double)pr.sv);
  }
  using PreE = Pre;
  __device__ PreE pre_e(int row, int c) const { return PreE{sv[(size_t)row * KO + c]}; }
  __device__ void epi_e(int row, int c, T acc, const PreE &pr, T, double *d) const {
    this->y[(size_t)row * KO + c] = acc;
    const double t = (double)acc;
    d[0] = fma(t, t, d[0]);
    d[1] = fma(t, (double)pr.sv, d[1]);
  }
};

// The last product of the polynomial preconditioner of CG (momentum_solver 2
// and 3): acc = (A x)_row, output z = M^-1 r = g0 w - g1 D^-1 (A x) with
// w = D^-1 r; pairs: the stopping norm of r (|D^-1 r|^2 on the Jacobi-scaled
// rule, else |r|^2) and r.z.  Degree 1: x = w, so z = (g0 I - g1 D^-1 A) w.
// Degree 2 (wsep): x = u = (a1 I + a2 D^-1 A) w from OpPolyU, g0 = a0 and
// g1 = -1, so z = (a0 I + a1 D^-1 A + a2 (D^-1 A)^2) w; w is then read from wv
// (CSR) or formed from r (element kernel, the rounding of its staging).
template <class T, int KO, class Base>
struct OpPolyZ : OpAx<T, KO, Base> {
  const T *__restrict__ r;
  const T *__restrict__ invd;  // per node
  double g0, g1;
  int scaled;
  int wsep;
  const T *__restrict__ wv;
  struct Pre {
    T r, d, w;
  };
  __device__ Pre pre(int row, int c) const {
    const size_t i = (size_t)row * KO + c;
    return Pre{r[i], invd[row], wsep ? wv[i] : this->x[i]};
  }
  __device__ void epi(int row, int c, T acc, const Pre &pr, DD *d) const {
    const T z = (T)(g0 * (double)pr.w - g1 * ((double)pr.d * (double)acc));
    this->y[(size_t)row * KO + c] = z;
    const double rv = (double)pr.r;
    const double sn = scaled ? (double)pr.w : rv;
    dd_fma(d[0], sn, sn);
    dd_fma(d[1], rv, (double)z);
  }
  // w is the kernel's x: its row value comes from the staged copy.  With
  // sscale = D^-1 the element kernel is given x = r and stages w = D^-1 r
  // itself (the same rounding as a stored w), so w need not be stored.
  static constexpr bool kStageScale = true;
  const T *sscale;
  struct PreE {
    T r, d;
  };
  __device__ PreE pre_e(int row, int c) const {
    return PreE{r[(size_t)row * KO + c], invd[row]};
  }
  __device__ void epi_e(int row, int c, T acc, const PreE &pr, T xr, double *d) const {
    const T w = wsep ? (T)((double)pr.d * (double)pr.r) : xr;
    const T z = (T)(g0 * (double)w - g1 * ((double)pr.d * (double)acc));
    this->y[(size_t)row * KO + c] = z;
    const double rv = (double)pr.r;
    const double sn = scaled ? (double)w : rv;
    d[0] = fma(sn, sn, d[0]);
    d[1] = fma(rv, (double)z, d[1]);
  }
};

// The first product of the degree-2 polynomial preconditioner (momentum_solver
// 3): input w = D^-1 r (element kernel: r staged and scaled by D^-1), output
// u = a1 w + a2 D^-1 (A w); no inner products.
template <class T, int KO, class Base>
struct OpPolyU : OpAx<T, KO, Base> {
  const T *__restrict__ invd;  // per node
  double a1, a2;
  struct Pre {
    T d, w;
  };
  __device__ Pre pre(int row, int c) const {
    return Pre{invd[row], this->x[(size_t)row * KO + c]};
  }
  __device__ void epi(int row, int c, T acc, const Pre &pr, DD *) const {
    this->y[(size_t)row * KO + c] = (T)(a1 * (double)pr.w + a2 * ((double)pr.d * (double)acc));
  }
  static constexpr bool kStageScale = true;
  const T *sscale;
  struct PreE {
    T d;
  };
  __device__ PreE pre_e(int row, int) const { return PreE{invd[row]}; }
  __device__ void epi_e(int row, int c, T acc, const PreE &pr, T w, double *) const {
    this->y[(size_t)row * KO + c] = (T)(a1 * (double)w + a2 * ((double)pr.d * (double)acc));
  }
};

template <class T, int KO, class Base>
struct OpResidual : OpAx<T, KO, Base> {
  const T *__restrict__ b;
  T *__restrict__ r0;
  const T *__restrict__ sc;  // per-row scale of the norm (nullptr: |r|^2)
  struct Pre {
    T b, sc;
  };
  __device__ Pre pre(int row, int c) const {
    return Pre{b[(size_t)row * KO + c], sc ? sc[row] : T(1)};
  }
  __device__ void epi(int row, int c, T acc, const Pre &pr, DD *d) const {
    const size_t i = (size_t)row * KO + c;
    const T r = pr.b - acc;
    this->y[i] = r;
    r0[i] = r;
    const double rs = sc ? (double)pr.sc * (double)r : (double)r;
    dd_fma(d[0], rs, rs);
  }
  using PreE = Pre;
  __device__ PreE pre_e(int row, int c) const { return pre(row, c); }
  __device__ void epi_e(int row, int c, T acc, const PreE &pr, T, double *d) const {
    const size_t i = (size_t)row * KO + c;
    const T r = pr.b - acc;
    this->y[i] = r;
    r0[i] = r;
    const double rs = sc ? (double)pr.sc * (double)r : (double)r;
    d[0] = fma(rs, rs, d[0]);
  }
};

struct OpGrad3 : MaskBase {
  const int *__restrict__ col;
  const double *__restrict__ g;
  const double *__restrict__ x;
  double *__restrict__ y;
  struct Frag {
    int j;
    double g[3];
  };
  __device__ Frag fetch(int k) const {
    return Frag{col[k], {g[(size_t)k * 3], g[(size_t)k * 3 + 1], g[(size_t)k * 3 + 2]}};
  }
  __device__ void prod(const Frag &f, double *p) const {
    const double xj = x[f.j];
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = f.g[c] * xj;
  }
  __device__ void epi(int row, int c, double acc, const Pre &, DD *) const {
    y[(size_t)row * 3 + c] = acc;
  }
};

struct OpMomRhs : MaskBase {
  const int *__restrict__ col;
  const double *__restrict__ M;
  const double *__restrict__ u;
  const double *__restrict__ gp;
  const double *__restrict__ conv;
  const uint8_t *__restrict__ bc;
  const double *__restrict__ ginlet;
  double *__restrict__ rhs;
  double dt, penalty;
  int nv;
  struct Frag {
    int j;
    double a;
  };
  __device__ Frag fetch(int k) const { return Frag{col[k], M[k]}; }
  __device__ void prod(const Frag &f, double *p) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = f.a * u[(size_t)f.j * 3 + c];
  }
  __device__ void epi(int row, int c, double acc, const Pre &, DD *) const {
    const size_t i = (size_t)row * 3 + c;
    // rhs = 0; rhs += (M u)/dt; rhs(P1) -= G p; rhs += conv; rhs += pen*u_bc
    double v = 0.0 + acc / dt;
    if (row < nv) v = v - gp[i];
    v = v + conv[i];
    const uint8_t b = bc[row];
    if (b == kInlet)
      v = v + penalty * ginlet[i];
    else if (b == kWall)
      v = v + penalty * 0.0;
    rhs[i] = v;
  }
};

// ---- launchers ---------------------------------------------------------------
template <class T>
void spmv3(hipStream_t s, const DevCsr &A, const T *val, const T *x, T *y,
           const int *active) {
  OpAx<T, 3, MaskBase> op{};
  op.mask = active;
  op.col = A.col;
  op.val = val;
  op.x = x;
  op.y = y;
  (void)launch_spmv<T, 3, 0>(s, A, op, nullptr);
}

template <class T>
void spmv3_dot(hipStream_t s, const DevCsr &A, const T *val, const T *y, T *v,
               const T *r0, const int *active, double *partials, int *grid) {
  OpAxDot<T, 3, MaskBase> op{};
  op.mask = active;
  op.col = A.col;
  op.val = val;
  op.x = y;
  op.y = v;
  op.r0 = r0;
  const int G = launch_spmv<T, 3, 1>(s, A, op, partials);
  *grid = G;
}

template <class T>
void spmv3_dot2(hipStream_t s, const DevCsr &A, const T *val, const T *z, T *t,
                const T *sv, const int *active, double *partials, int *grid) {
  OpAxDot2<T, 3, MaskBase> op{};
  op.mask = active;
  op.col = A.col;
  op.val = val;
  op.x = z;
  op.y = t;
  op.sv = sv;
  const int G = launch_spmv<T, 3, 2>(s, A, op, partials);
  *grid = G;
}

template <class T>
void spmv3_polyz(hipStream_t s, const DevCsr &A, const T *val, const T *w, T *z, const T *r,
                 const T *invd, double g0, double g1, int scaled, const int *active,
                 double *partials, int *grid, int wsep, const T *wv) {
  OpPolyZ<T, 3, MaskBase> op{};
  op.wsep = wsep;
  op.wv = wv;
  if (wsep && !wv && !elem_product(A, sizeof(T)))
    throw std::logic_error("spmv3_polyz: w is formed from r by the element-patch kernel only");
  if (w == nullptr) {
    // w = D^-1 r formed in the element kernel's staging
    if (!elem_product(A, sizeof(T)))
      throw std::logic_error("spmv3_polyz: only the element-patch kernel forms w itself");
    w = r;
    op.sscale = invd;
  }
  op.mask = active;
  op.col = A.col;
  op.val = val;
  op.x = w;
  op.y = z;
  op.r = r;
  op.invd = invd;
  op.g0 = g0;
  op.g1 = g1;
  op.scaled = scaled;
  const int G = launch_spmv<T, 3, 2>(s, A, op, partials);
  *grid = G;
}

template <class T>
void spmv3_polyu(hipStream_t s, const DevCsr &A, const T *val, const T *w, T *u, const T *r,
                 const T *invd, double a1, double a2, const int *active) {
  OpPolyU<T, 3, MaskBase> op{};
  op.mask = active;
  op.col = A.col;
  op.val = val;
  op.x = w;
  op.y = u;
  op.invd = invd;
  op.a1 = a1;
  op.a2 = a2;
  if (w == nullptr) {
    // w = D^-1 r formed in the element kernel's staging
    if (!elem_product(A, sizeof(T)))
      throw std::logic_error("spmv3_polyu: only the element-patch kernel forms w itself");
    op.x = r;
    op.sscale = invd;
  }
  (void)launch_spmv<T, 3, 0>(s, A, op, nullptr);
}

template <class T>
void spmv3_residual(hipStream_t s, const DevCsr &A, const T *val, const T *x,
                    const T *b, T *r, T *r0, const T *sc, const int *active, double *partials,
                    int *grid) {
  OpResidual<T, 3, MaskBase> op{};
  op.mask = active;
  op.col = A.col;
  op.val = val;
  op.x = x;
  op.y = r;
  op.b = b;
  op.r0 = r0;
  op.sc = sc;
  const int G = launch_spmv<T, 3, 1>(s, A, op, partials);
  *grid = G;
}

template <class T>
void spmv1(hipStream_t s, const DevCsr &A, const T *val, const T *x, T *y) {
  OpAx<T, 1, Mask1Base> op{};
  op.mask = nullptr;
  op.col = A.col;
  op.val = val;
  op.x = x;
  op.y = y;
  (void)launch_spmv<T, 1, 0>(s, A, op, nullptr);
}

template <class T>
void spmv1_dot(hipStream_t s, const DevCsr &A, const T *val, const T *y, T *v,
               const T *r0, const int *active, double *partials, int *grid) {
  OpAxDot<T, 1, Mask1Base> op{};
  op.mask = active;
  op.col = A.col;
  op.val = val;
  op.x = y;
  op.y = v;
  op.r0 = r0;
  const int G = launch_spmv<T, 1, 1>(s, A, op, partials);
  *grid = G;
}

template <class T>
void spmv1_dot2(hipStream_t s, const DevCsr &A, const T *val, const T *z, T *t,
                const T *sv, const int *active, double *partials, int *grid) {
  OpAxDot2<T, 1, Mask1Base> op{};
  op.mask = active;
  op.col = A.col;
  op.val = val;
  op.x = z;
  op.y = t;
  op.sv = sv;
  const int G = launch_spmv<T, 1, 2>(s, A, op, partials);
  *grid = G;
}

template <class T>
void spmv1_residual(hipStream_t s, const DevCsr &A, const T *val, const T *x,
                    const T *b, T *r, T *r0, const T *sc, const int *active, double *partials,
                    int *grid) {
  OpResidual<T, 1, Mask1Base> op{};
  op.mask = active;
  op.col = A.col;
  op.val = val;
  op.x = x;
  op.y = r;
  op.b = b;
  op.r0 = r0;
  op.sc = sc;
  const int G = launch_spmv<T, 1, 1>(s, A, op, partials);
  *grid = G;
}

void spmv_grad3(hipStream_t s, const DevCsr &P, const double *gval, const double *x,
                double *y) {
  OpGrad3 op{};
  op.mask = nullptr;
  op.col = P.col;
  op.g = gval;
  op.x = x;
  op.y = y;
  (void)launch_spmv<double, 3, 0>(s, P, op, nullptr);
}

// y = beta y + alpha sum_c G_c^T x_c with transposed values gT and x of
// stride xs (the first 3 entries of each node record are the components)
struct OpGradT3 : Mask1Base {
  const int *__restrict__ col;
  const double *__restrict__ g;
  const double *__restrict__ x;
  double *__restrict__ y;
  int xs;
  double alpha, beta;
  struct Frag {
    int j;
    double g[3];
  };
  __device__ Frag fetch(int k) const {
    return Frag{col[k], {g[(size_t)k * 3], g[(size_t)k * 3 + 1], g[(size_t)k * 3 + 2]}};
  }
  __device__ void prod(const Frag &f, double *p) const {
    const double *xj = x + (size_t)f.j * xs;
    p[0] = f.g[0] * xj[0] + f.g[1] * xj[1] + f.g[2] * xj[2];
  }
  __device__ void epi(int row, int, double acc, const Pre &, DD *) const {
    y[row] = (beta != 0.0 ? beta * y[row] : 0.0) + alpha * acc;
  }
};

void spmv_gradT_in3(hipStream_t s, const DevCsr &P, const double *gT, const double *x, int xs,
                    double *y, double alpha, double beta) {
  OpGradT3 op{};
  op.mask = nullptr;
  op.col = P.col;
  op.g = gT;
  op.x = x;
  op.y = y;
  op.xs = xs;
  op.alpha = alpha;
  op.beta = beta;
  (void)launch_spmv<double, 1, 0>(s, P, op, nullptr);
}

// y[i][c] += sum_j G_c[i,j] x[j] over the P1 rows of a [*][3] array
struct OpGrad3Acc : OpGrad3 {
  __device__ void epi(int row, int c, double acc, const Pre &, DD *) const {
    y[(size_t)row * 3 + c] += acc;
  }
};
void spmv_grad3_acc(hipStream_t s, const DevCsr &P, const double *g, const double *x,
                    double *y) {
  OpGrad3Acc op{};
  op.mask = nullptr;
  op.col = P.col;
  op.g = g;
  op.x = x;
  op.y = y;
  (void)launch_spmv<double, 3, 0>(s, P, op, nullptr);
}

// adjoint of the momentum right-hand side: lu = (M lb)/dt, lg += pen lb on Inlet rows
struct OpMomAdj : MaskBase {
  const int *__restrict__ col;
  const double *__restrict__ M;
  const double *__restrict__ lb;
  const uint8_t *__restrict__ bc;
  double *__restrict__ lu;
  double *__restrict__ lg;
  double dt, pen;
  struct Frag {
    int j;
    double a;
  };
  __device__ Frag fetch(int k) const { return Frag{col[k], M[k]}; }
  __device__ void prod(const Frag &f, double *p) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = f.a * lb[(size_t)f.j * 3 + c];
  }
  __device__ void epi(int row, int c, double acc, const Pre &, DD *) const {
    const size_t i = (size_t)row * 3 + c;
    lu[i] = acc / dt;
    if (bc[row] == kInlet) lg[i] += pen * lb[i];
  }
};
void momentum_rhs_adjoint(hipStream_t s, const DevCsr &A, const double *M, const double *lb,
                          const uint8_t *bc, double dt, double pen, double *lu, double *lg) {
  OpMomAdj op{};
  op.mask = nullptr;
  op.col = A.col;
  op.M = M;
  op.lb = lb;
  op.bc = bc;
  op.lu = lu;
  op.lg = lg;
  op.dt = dt;
  op.pen = pen;
  (void)launch_spmv<double, 3, 0>(s, A, op, nullptr);
}

void momentum_rhs(hipStream_t s, const DevCsr &A, const double *M, const double *u,
                  const double *gp, const double *conv, const uint8_t *bc,
                  const double *ginlet, double dt, int nv, double penalty,
                  double *rhs) {
  OpMomRhs op{};
  op.mask = nullptr;
  op.col = A.col;
  op.M = M;
  op.u = u;
  op.gp = gp;
  op.conv = conv;
  op.bc = bc;
  op.ginlet = ginlet;
  op.rhs = rhs;
  op.dt = dt;
  op.penalty = penalty;
  op.nv = nv;
  (void)launch_spmv<double, 3, 0>(s, A, op, nullptr);
}

#define CFSA_INST_SPMV(T)                                                              \
  template void spmv3<T>(hipStream_t, const DevCsr &, const T *, const T *, T *,        \
                         const int *);                                                  \
  template void spmv3_dot<T>(hipStream_t, const DevCsr &, const T *, const T *, T *,    \
                             const T *, const int *, double *, int *);                  \
  template void spmv3_dot2<T>(hipStream_t, const DevCsr &, const T *, const T *, T *,   \
                              const T *, const int *, double *, int *);                 \
  template void spmv3_polyz<T>(hipStream_t, const DevCsr &, const T *, const T *, T *,  \
                               const T *, const T *, double, double, int, const int *,  \
                               double *, int *, int, const T *);                        \
  template void spmv3_polyu<T>(hipStream_t, const DevCsr &, const T *, const T *, T *,  \
                               const T *, const T *, double, double, const int *);      \
  template void spmv3_residual<T>(hipStream_t, const DevCsr &, const T *, const T *,    \
                                  const T *, T *, T *, const T *, const int *, double *, \
                                  int *);                                               \
  template void spmv1<T>(hipStream_t, const DevCsr &, const T *, const T *, T *);       \
  template void spmv1_dot<T>(hipStream_t, const DevCsr &, const T *, const T *, T *,    \
                             const T *, const int *, double *, int *);                  \
  template void spmv1_dot2<T>(hipStream_t, const DevCsr &, const T *, const T *, T *,   \
                              const T *, const int *, double *, int *);                 \
  template void spmv1_residual<T>(hipStream_t, const DevCsr &, const T *, const T *,    \
                                  const T *, T *, T *, const T *, const int *, double *, \
                                  int *);
CFSA_INST_SPMV(double)
CFSA_INST_SPMV(float)

// ===========================================================================
// Deterministic reductions
// ===========================================================================
int vec_grid(int n) { return std::max(1, std::min(div_up(n, kBlock), kVecGrid)); }

// Krylov vector kernels run over the flat element index i of a [n][K] array
// (one lane per component: consecutive lanes touch consecutive 4/8-byte words,
// three times fewer cache lines per load than one lane per node).  The grid is
// a multiple of K, so a lane always meets the same component c = i % K.
static int vec_grid_k(int n, int K) {
  int G = std::max(K, std::min(div_up((int64_t)n * K, kBlock), kVecGrid));
  return G - G % K;
}

// Block reduction of per-lane pairs binned by component: lane t of block b holds
// component (b * kBlock + t) % K.  32 lanes per (component, value) merge every
// K-th lane in a fixed order, then a fixed shuffle tree; lane 0 writes
// partials[(c * ND + d) * gridDim.x + b].
template <int K, int ND>
__device__ void block_partials_binned(DD (&v)[ND], double *__restrict__ partials) {
  __shared__ double rh[ND][kBlock], rl[ND][kBlock];
  const int tid = threadIdx.x;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    rh[d][tid] = v[d].hi;
    rl[d][tid] = v[d].lo;
  }
  __syncthreads();
  static_assert(K * ND * 32 <= kBlock, "lanes per value");
  if (tid < K * ND * 32) {
    const int q = tid / 32, l = tid % 32;
    const int c = q / ND, d = q % ND;
    const int c0 = (int)(((size_t)blockIdx.x * kBlock) % K);
    const int t0 = (c - c0 + K) % K;
    DD a;
    dd_zero(a);
    for (int t = t0 + K * l; t < kBlock; t += K * 32) dd_merge(a, rh[d][t], rl[d][t]);
#pragma unroll
    for (int st = 16; st > 0; st >>= 1) {
      const double oh = __shfl_down(a.hi, st, 32);
      const double ol = __shfl_down(a.lo, st, 32);
      dd_merge(a, oh, ol);
    }
    if (l == 0) {
      const size_t o = 2 * ((size_t)(c * ND + d) * gridDim.x + blockIdx.x);
      partials[o] = a.hi;
      partials[o + 1] = a.lo;
    }
  }
}

// ===========================================================================
// BiCGSTAB (Eigen's right-preconditioned van der Vorst form)
// ===========================================================================
template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_bicg_init(int n, const T *__restrict__ b,
                                                      T *__restrict__ x, T *__restrict__ r,
                                                      T *__restrict__ r0, T *__restrict__ p,
                                                      T *__restrict__ v,
                                                      const T *__restrict__ sc,
                                                      double *__restrict__ partials) {
  // partials: [2c] = stopping norm of b (|b|^2, or |D^-1 b|^2 when sc != null),
  //           [2c+1] = |b|^2 = r0.r of the first iteration
  DD acc[2];
  dd_zero(acc[0]);
  dd_zero(acc[1]);
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
    const T bi = b[i];
    x[i] = T(0);
    r[i] = bi;
    r0[i] = bi;
    p[i] = T(0);
    v[i] = T(0);
    const double ds = sc ? (double)sc[i / K] * (double)bi : (double)bi;
    dd_fma(acc[0], ds, ds);
    dd_fma(acc[1], (double)bi, (double)bi);
  }
  block_partials_binned<K, 2>(acc, partials);
}

template <int K>
__global__ void k_bicg_s_init(KrylovState *st, const double *partials, int G) {
  __shared__ double tot[2 * K];
  reduce_final<2 * K>(partials, G, tot);
  if (threadIdx.x != 0) return;
  int any = 0;
  for (int c = 0; c < K; ++c) {
    const double bb = tot[2 * c];
    st->rhs_sq[c] = bb;
    st->r0_sq[c] = tot[2 * c + 1];
    st->rr[c] = bb;
    st->rho[c] = 1.0;
    st->alpha[c] = 1.0;
    st->omega[c] = 1.0;
    st->beta[c] = 0.0;
    st->iters[c] = 0;
    st->restarts[c] = 0;
    st->active[c] = 0;
    st->tol2[c] = st->tol * st->tol * bb;
    if (bb == 0.0) {
      st->status[c] = KS_DONE;  // x = 0 and success
      st->err[c] = 0.0;
    } else {
      st->status[c] = KS_RUNNING;
      st->err[c] = 1.0;
      any = 1;
    }
  }
  for (int c = K; c < 3; ++c) {
    st->status[c] = KS_DONE;
    st->active[c] = 0;
  }
  st->any_active = any;
}

template <int K>
__global__ void k_bicg_s_begin(KrylovState *st, const double *partials, int G) {
  __shared__ double tot[2 * K];
  reduce_final<2 * K>(partials, G, tot);
  if (threadIdx.x != 0) return;
  const double eps = DBL_EPSILON;
  const double eps2 = eps * eps;
  int any = 0;
  for (int c = 0; c < K; ++c) {
    const int status = st->status[c];
    if (status == KS_RESUME) {
      st->beta[c] = (st->rho[c] / st->rho_old[c]) * (st->alpha[c] / st->omega[c]);
      st->status[c] = KS_RUNNING;
      st->active[c] = 1;
      any = 1;
      continue;
    }
    if (status != KS_RUNNING) {
      st->active[c] = 0;
      continue;
    }
    const double rr = tot[2 * c];
    st->rr[c] = rr;
    if (!(rr > st->tol2[c] && st->iters[c] < st->maxit)) {
      st->status[c] = KS_DONE;
      st->err[c] = sqrt(rr / st->rhs_sq[c]);
      st->active[c] = 0;
      continue;
    }
    st->rho_old[c] = st->rho[c];
    const double rho = tot[2 * c + 1];
    st->rho[c] = rho;
    if (fabs(rho) < st->restart_scale * eps2 * st->r0_sq[c]) {
      st->status[c] = KS_RESTART;
      st->active[c] = 0;
      continue;
    }
    st->beta[c] = (rho / st->rho_old[c]) * (st->alpha[c] / st->omega[c]);
    st->active[c] = 1;
    any = 1;
  }
  st->any_active = any;
}

template <int K>
__global__ void k_bicg_s_alpha(KrylovState *st, const double *partials, int G) {
  __shared__ double tot[K];
  reduce_final<K>(partials, G, tot);
  if (threadIdx.x != 0) return;
  for (int c = 0; c < K; ++c) {
    if (!st->active[c]) continue;
    st->r0v[c] = tot[c];
    st->alpha[c] = st->rho[c] / tot[c];
  }
}

template <int K>
__global__ void k_bicg_s_omega(KrylovState *st, const double *partials, int G) {
  __shared__ double tot[2 * K];
  reduce_final<2 * K>(partials, G, tot);
  if (threadIdx.x != 0) return;
  for (int c = 0; c < K; ++c) {
    if (!st->active[c]) continue;
    const double tt = tot[2 * c], ts = tot[2 * c + 1];
    st->tt[c] = tt;
    st->ts[c] = ts;
    st->omega[c] = (tt > 0.0) ? ts / tt : 0.0;
    st->iters[c] += 1;
  }
}

// After r = b - A x, r0 = r for the masked right-hand sides.
template <int K>
__global__ void k_bicg_s_restart(KrylovState *st, const double *partials, int G,
                                 const int *mask) {
  __shared__ double tot[K];
  reduce_final<K>(partials, G, tot);
  if (threadIdx.x != 0) return;
  for (int c = 0; c < K; ++c) {
    if (!mask[c]) continue;
    st->rho[c] = tot[c];
    st->r0_sq[c] = tot[c];
    if (st->restarts[c]++ == 0) st->iters[c] = 0;
    st->status[c] = KS_RESUME;
  }
}

// Accesses to 3-rhs Krylov vectors that a kernel touches once and the next
// SpMV does not gather (tens of MB each): nontemporal, so they do not displace
// the vector the SpMV gathers next from the caches.  Measured at 96^3: PCG
// (q in k_cg_xr, x in k_cg_p) SpMV 407 -> 401 us, k_cg_p 83 -> 82 us; BiCGSTAB
// FP64 (p, v in k_bicg_v1; r, v in k_bicg_v2; x, y, z, s, t, r in k_bicg_v3)
// v1 174 -> 153 us, v2 133 -> 126 us, v3 262 -> 247 us.  Single-rhs vectors
// stay cache resident and keep ordinary accesses.
template <int K, class T>
__device__ __forceinline__ T load_once(const T *p) {
  if constexpr (K == 3)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <int K, class T>
__device__ __forceinline__ void store_once(T v, T *p) {
  if constexpr (K == 3)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_bicg_v1(int n, const KrylovState *__restrict__ st,
                                                    const T *__restrict__ r, T *__restrict__ p,
                                                    const T *__restrict__ v,
                                                    const T *__restrict__ invd,
                                                    T *__restrict__ y) {
  if (!st->any_active) return;
  const size_t i0 = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const int c = (int)(i0 % K);
  if (!st->active[c]) return;
  const double beta = st->beta[c], om = st->omega[c];
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = i0; i < N; i += stride) {
    const double pi =
        (double)r[i] + beta * ((double)load_once<K>(&p[i]) - om * (double)load_once<K>(&v[i]));
    const T pt = (T)pi;
    store_once<K>(pt, &p[i]);
    y[i] = (T)((double)invd[i / K] * (double)pt);
  }
}

template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_bicg_v2(int n, const KrylovState *__restrict__ st,
                                                    const T *__restrict__ r,
                                                    const T *__restrict__ v,
                                                    const T *__restrict__ invd,
                                                    T *__restrict__ sv, T *__restrict__ z) {
  if (!st->any_active) return;
  const size_t i0 = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const int c = (int)(i0 % K);
  if (!st->active[c]) return;
  const double al = st->alpha[c];
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = i0; i < N; i += stride) {
    const T si = (T)((double)load_once<K>(&r[i]) - al * (double)load_once<K>(&v[i]));
    sv[i] = si;
    z[i] = (T)((double)invd[i / K] * (double)si);
  }
}

template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_bicg_v3(
    int n, const KrylovState *__restrict__ st, const T *__restrict__ y,
    const T *__restrict__ z, const T *__restrict__ sv, const T *__restrict__ t,
    const T *__restrict__ r0, T *__restrict__ x, T *__restrict__ r,
    const T *__restrict__ sc, double *__restrict__ partials) {
  if (!st->any_active) return;  // grid-uniform: the partials are not read then
  const size_t i0 = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const int c = (int)(i0 % K);
  const bool act = st->active[c] != 0;
  const double al = st->alpha[c], om = st->omega[c];
  DD acc[2];
  dd_zero(acc[0]);
  dd_zero(acc[1]);
  const size_t N = act ? (size_t)n * K : 0, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = i0; i < N; i += stride) {
    const double xi = (double)load_once<K>(&x[i]);
    store_once<K>(
        (T)(xi + (al * (double)load_once<K>(&y[i]) + om * (double)load_once<K>(&z[i]))), &x[i]);
    const T ri = (T)((double)load_once<K>(&sv[i]) - om * (double)load_once<K>(&t[i]));
    store_once<K>(ri, &r[i]);
    const double rs = sc ? (double)sc[i / K] * (double)ri : (double)ri;
    dd_fma(acc[0], rs, rs);
    dd_fma(acc[1], (double)r0[i], (double)ri);
  }
  block_partials_binned<K, 2>(acc, partials);
}

template <class T, int K>
void bicg_init(hipStream_t s, int n, const T *b, T *x, T *r, T *r0, T *p, T *v,
               const T *stop_scale, double *partials, int *grid) {
  const int G = vec_grid_k(n, K);
  *grid = G;
  hipLaunchKernelGGL((k_bicg_init<T, K>), dim3(G), dim3(kBlock), 0, s, n, b, x, r, r0, p,
                     v, stop_scale, partials);
  CFSA_LAUNCH_CHECK();
}
template <int K>
void bicg_s_init(hipStream_t s, KrylovState *st, const double *partials, int G) {
  hipLaunchKernelGGL((k_bicg_s_init<K>), dim3(1), dim3(kBlock), 0, s, st, partials, G);
  CFSA_LAUNCH_CHECK();
}
template <int K>
void bicg_s_begin(hipStream_t s, KrylovState *st, const double *partials, int G) {
  hipLaunchKernelGGL((k_bicg_s_begin<K>), dim3(1), dim3(kBlock), 0, s, st, partials, G);
  CFSA_LAUNCH_CHECK();
}
template <int K>
void bicg_s_alpha(hipStream_t s, KrylovState *st, const double *partials, int G) {
  hipLaunchKernelGGL((k_bicg_s_alpha<K>), dim3(1), dim3(kBlock), 0, s, st, partials, G);
  CFSA_LAUNCH_CHECK();
}
template <int K>
void bicg_s_omega(hipStream_t s, KrylovState *st, const double *partials, int G) {
  hipLaunchKernelGGL((k_bicg_s_omega<K>), dim3(1), dim3(kBlock), 0, s, st, partials, G);
  CFSA_LAUNCH_CHECK();
}
template <int K>
void bicg_s_restart(hipStream_t s, KrylovState *st, const double *partials, int G,
                    const int *mask) {
  hipLaunchKernelGGL((k_bicg_s_restart<K>), dim3(1), dim3(kBlock), 0, s, st, partials, G,
                     mask);
  CFSA_LAUNCH_CHECK();
}
template <class T, int K>
void bicg_v1(hipStream_t s, int n, const KrylovState *st, const T *r, T *p, const T *v,
             const T *invd, T *y) {
  hipLaunchKernelGGL((k_bicg_v1<T, K>), dim3(vec_grid_k(n, K)), dim3(kBlock), 0, s, n, st, r, p,
                     v, invd, y);
  CFSA_LAUNCH_CHECK();
}
template <class T, int K>
void bicg_v2(hipStream_t s, int n, const KrylovState *st, const T *r, const T *v,
             const T *invd, T *sv, T *z) {
  hipLaunchKernelGGL((k_bicg_v2<T, K>), dim3(vec_grid_k(n, K)), dim3(kBlock), 0, s, n, st, r, v,
                     invd, sv, z);
  CFSA_LAUNCH_CHECK();
}
template <class T, int K>
void bicg_v3(hipStream_t s, int n, const KrylovState *st, const T *y, const T *z,
             const T *sv, const T *t, const T *r0, T *x, T *r, const T *stop_scale,
             double *partials,
             int *grid) {
  const int G = vec_grid_k(n, K);
  *grid = G;
  hipLaunchKernelGGL((k_bicg_v3<T, K>), dim3(G), dim3(kBlock), 0, s, n, st, y, z, sv, t, r0,
                     x, r, stop_scale, partials);
  CFSA_LAUNCH_CHECK();
}

#define CFSA_INST_BICG(T, K)                                                                \
  template void bicg_init<T, K>(hipStream_t, int, const T *, T *, T *, T *, T *, T *, const T *, \
                                double *, int *);                                           \
  template void bicg_v1<T, K>(hipStream_t, int, const KrylovState *, const T *, T *,        \
                              const T *, const T *, T *);                                   \
  template void bicg_v2<T, K>(hipStream_t, int, const KrylovState *, const T *, const T *,  \
                              const T *, T *, T *);                                         \
  template void bicg_v3<T, K>(hipStream_t, int, const KrylovState *, const T *, const T *,  \
                              const T *, const T *, const T *, T *, T *, const T *, double *, \
                              int *);
CFSA_INST_BICG(double, 1)
CFSA_INST_BICG(double, 3)
CFSA_INST_BICG(float, 1)
CFSA_INST_BICG(float, 3)
#define CFSA_INST_BICG_S(K)                                                          \
  template void bicg_s_init<K>(hipStream_t, KrylovState *, const double *, int);     \
  template void bicg_s_begin<K>(hipStream_t, KrylovState *, const double *, int);    \
  template void bicg_s_alpha<K>(hipStream_t, KrylovState *, const double *, int);    \
  template void bicg_s_omega<K>(hipStream_t, KrylovState *, const double *, int);    \
  template void bicg_s_restart<K>(hipStream_t, KrylovState *, const double *, int,   \
                                  const int *);
CFSA_INST_BICG_S(1)
CFSA_INST_BICG_S(3)

// ===========================================================================
// Jacobi PCG (stopping rule of Eigen's conjugate_gradient), K interleaved
// right-hand sides solved independently with per-component scalars.  With
// K = 1 this is exactly the reference's pressure CG path.
// ===========================================================================
template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_cg_init(int n, const T *__restrict__ b,
                                                    const T *__restrict__ invd,
                                                    T *__restrict__ x, T *__restrict__ r,
                                                    T *__restrict__ p, int scaled,
                                                    double *__restrict__ partials) {
  // partials: [2c] = stopping norm of b, [2c+1] = b . M^-1 b
  DD acc[2];
  dd_zero(acc[0]);
  dd_zero(acc[1]);
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
    const T bi = b[i];
    x[i] = T(0);
    r[i] = bi;
    const T pi = (T)((double)invd[i / K] * (double)bi);
    p[i] = pi;
    const double sn = scaled ? (double)pi : (double)bi;
    dd_fma(acc[0], sn, sn);
    dd_fma(acc[1], (double)bi, (double)pi);
  }
  block_partials_binned<K, 2>(acc, partials);
}

template <int K>
__global__ void k_cg_s_init(KrylovState *st, const double *partials, int G) {
  __shared__ double tot[2 * K];
  reduce_final<2 * K>(partials, G, tot);
  if (threadIdx.x != 0) return;
  int any = 0;
  for (int c = 0; c < K; ++c) {
    const double bb = tot[2 * c];
    st->rhs_sq[c] = bb;
    st->iters[c] = 0;
    st->restarts[c] = 0;
    st->active[c] = 0;
    if (bb == 0.0) {
      st->status[c] = KS_DONE;
      st->err[c] = 0.0;
      continue;
    }
    const double thr = fmax(st->tol * st->tol * bb, DBL_MIN);
    st->tol2[c] = thr;
    st->rr[c] = bb;
    if (bb < thr) {
      st->status[c] = KS_DONE;
      st->err[c] = sqrt(bb / bb);
      continue;
    }
    st->rho[c] = tot[2 * c + 1];
    st->rho_s[0][c] = tot[2 * c + 1];
    st->it_s[0][c] = 0;
    st->status[c] = (st->maxit > 0) ? KS_RUNNING : KS_DONE;
    st->err[c] = 1.0;
    st->active[c] = st->maxit > 0;
    any |= st->active[c];
  }
  for (int c = K; c < 3; ++c) {
    st->status[c] = KS_DONE;
    st->active[c] = 0;
  }
  st->any_active = any;
}

// One PCG iteration is SpMV (q = A p, p.q), then k_cg_xr, then k_cg_p.  Each
// of the two vector kernels reduces the partial pairs it needs in every block
// (same fixed order, same result), so no single-block scalar kernel sits
// between them.  Iteration parity `par` selects the rho / count slot read; block
// 0 writes the other slot.  The one state field read and written within a
// kernel is `active` in k_cg_p: block 0 clears it only for a rhs that stops, so
// a block that reads the cleared flag decides the same; the x update there
// goes by the snapshot act_s[par] that k_cg_xr took.
template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_cg_xr(int n, KrylovState *__restrict__ st,
                                                  const double *__restrict__ pq_partials,
                                                  int Gq, int par, const T *__restrict__ q,
                                                  const T *__restrict__ invd,
                                                  T *__restrict__ r, int scaled,
                                                  double *__restrict__ partials) {
  // the snapshot k_cg_p updates x by, written even when nothing is active
  if (blockIdx.x == 0 && threadIdx.x < K) st->act_s[par][threadIdx.x] = st->active[threadIdx.x];
  if (!st->any_active) return;  // grid-uniform: the partials are not read then
  double pq[K];
  block_reduce_pairs<K>(pq_partials, Gq, pq);
  if (blockIdx.x == 0 && threadIdx.x < K && st->active[threadIdx.x]) {
    const int c = threadIdx.x;
    st->r0v[c] = pq[c];
    st->alpha[c] = st->rho_s[par][c] / pq[c];
  }
  const size_t i0 = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const int c = (int)(i0 % K);
  const bool act = st->active[c] != 0;
  // alpha = rho / p.Ap (Eigen conjugate_gradient); x += alpha p follows in k_cg_p
  const double al = act ? st->rho_s[par][c] / pq[c] : 0.0;
  DD acc[2];
  dd_zero(acc[0]);
  dd_zero(acc[1]);
  const size_t N = act ? (size_t)n * K : 0, stride = (size_t)gridDim.x * kBlock;
  // four values per pass, loads first; the pairs accumulate in the same order
  // as one value per pass
  size_t i = i0;
  for (; i + 3 * stride < N; i += 4 * stride) {
    T rv[4], qv[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t j = i + u * stride;
      rv[u] = r[j];
      qv[u] = load_once<K>(&q[j]);
      dv[u] = invd[j / K];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T ri = (T)((double)rv[u] - al * (double)qv[u]);
      r[i + u * stride] = ri;
      const T zi = (T)((double)dv[u] * (double)ri);  // z = M^-1 r, not stored
      const double sn = scaled ? (double)zi : (double)ri;
      dd_fma(acc[0], sn, sn);
      dd_fma(acc[1], (double)ri, (double)zi);
    }
  }
  for (; i < N; i += stride) {
    const T ri = (T)((double)r[i] - al * (double)load_once<K>(&q[i]));
    r[i] = ri;
    const T zi = (T)((double)invd[i / K] * (double)ri);  // z = M^-1 r, not stored
    const double sn = scaled ? (double)zi : (double)ri;
    dd_fma(acc[0], sn, sn);
    dd_fma(acc[1], (double)ri, (double)zi);
  }
  block_partials_binned<K, 2>(acc, partials);
}

// Convergence test of one PCG iteration for each rhs, on |r|^2 (or
// |M^-1 r|^2) = tot[2c], rho_new = tot[2c+1]: a rhs that converges or reaches
// maxit leaves the loop, the others get beta = rho_new / rho.  The state is
// written only when `w`; cont[c] and beta[c] return the decision.
template <int K>
__device__ void cg_decide(KrylovState *__restrict__ st, const double *tot, int par, bool w,
                          double *beta_out, int *cont_out) {
  int any = 0;
  for (int c = 0; c < K; ++c) {
    int cont = 0;
    double beta = 0.0;
    if (st->active[c]) {
      const double rn2 = tot[2 * c];
      if (rn2 < st->tol2[c]) {
        if (w) {
          st->rr[c] = rn2;
          st->status[c] = KS_DONE;
          st->err[c] = sqrt(rn2 / st->rhs_sq[c]);
          st->active[c] = 0;
        }
      } else {
        const double rz = tot[2 * c + 1];
        const int it = st->it_s[par][c] + 1;
        beta = rz / st->rho_s[par][c];
        if (w) {
          st->rr[c] = rn2;
          st->rho_old[c] = st->rho_s[par][c];
          st->rho[c] = rz;
          st->beta[c] = beta;
          st->rho_s[1 - par][c] = rz;
          st->it_s[1 - par][c] = it;
          st->iters[c] = it;
        }
        if (it >= st->maxit) {
          // Eigen updates the direction before leaving; it is unused after
          if (w) {
            st->status[c] = KS_DONE;
            st->err[c] = sqrt(rn2 / st->rhs_sq[c]);
            st->active[c] = 0;
          }
        } else {
          cont = 1;
          any = 1;
        }
      }
    }
    beta_out[c] = beta;
    cont_out[c] = cont;
  }
  if (w) st->any_active = any;
}

// Every block reduces k_cg_xr's pairs and decides (block 0 writes the state),
// then, per value: x += alpha p for the rhs k_cg_xr updated (Eigen's update of
// x, moved here because this kernel reads p anyway), and p = z + beta p with
// z = M^-1 r recomputed from r (same rounding as in k_cg_xr) for the rhs that
// continue, or read from zin (the polynomial preconditioner's product).
template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_cg_p(int n, KrylovState *__restrict__ st,
                                                 const double *__restrict__ partials, int Gv,
                                                 int par, const T *__restrict__ r,
                                                 const T *__restrict__ invd,
                                                 const T *__restrict__ zin,
                                                 T *__restrict__ p, T *__restrict__ x) {
  __shared__ double beta_s[K], alpha_s[K];
  __shared__ int cont_s[K], xact_s[K];
  // act_s[par] is stable during this kernel (block 0 rewrites only `active`),
  // so every block takes the same early exit
  int go = 0;
#pragma unroll
  for (int c = 0; c < K; ++c) go |= st->act_s[par][c];
  if (!go) return;
  double tot[2 * K];
  block_reduce_pairs<2 * K>(partials, Gv, tot);
  if (threadIdx.x == 0) {
    cg_decide<K>(st, tot, par, blockIdx.x == 0, beta_s, cont_s);
#pragma unroll
    for (int c = 0; c < K; ++c) {
      xact_s[c] = st->act_s[par][c];
      alpha_s[c] = xact_s[c] ? st->alpha[c] : 0.0;
    }
  }
  __syncthreads();
  const size_t i0 = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const int c = (int)(i0 % K);
  if (!xact_s[c]) return;
  const double be = beta_s[c], al = alpha_s[c];
  const bool cont = cont_s[c] != 0;
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  // four values per pass, loads first (stride is a multiple of K, so all four
  // belong to component c)
  size_t i = i0;
  for (; i + 3 * stride < N; i += 4 * stride) {
    T rv[4], pv[4], dv[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t j = i + u * stride;
      pv[u] = p[j];
      xv[u] = load_once<K>(&x[j]);
      if (cont) {
        rv[u] = zin ? zin[j] : r[j];
        dv[u] = zin ? T(1) : invd[j / K];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t j = i + u * stride;
      store_once<K>((T)((double)xv[u] + al * (double)pv[u]), &x[j]);
      if (cont) {
        const T zi = (T)((double)dv[u] * (double)rv[u]);
        p[j] = (T)((double)zi + be * (double)pv[u]);
      }
    }
  }
  for (; i < N; i += stride) {
    const T pi = p[i];
    store_once<K>((T)((double)load_once<K>(&x[i]) + al * (double)pi), &x[i]);
    if (cont) {
      const T zi = zin ? zin[i] : (T)((double)invd[i / K] * (double)r[i]);
      p[i] = (T)((double)zi + be * (double)pi);
    }
  }
}

// r = r0 = b, the residual of x = 0 (b - A 0 = b exactly), and the pairs of
// its stopping norm |r|^2 (or |sc r|^2 with a per-row scale sc), binned by
// component: what the residual SpMV returns for x = 0, without the product.
template <int K>
__global__ __launch_bounds__(kBlock) void k_residual_x0(int n, const double *__restrict__ b,
                                                        double *__restrict__ r,
                                                        double *__restrict__ r0,
                                                        const double *__restrict__ sc,
                                                        double *__restrict__ partials) {
  DD acc[1];
  dd_zero(acc[0]);
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
    const double bi = b[i];
    r[i] = bi;
    r0[i] = bi;
    const double rs = sc ? sc[i / K] * bi : bi;
    dd_fma(acc[0], rs, rs);
  }
  block_partials_binned<K, 1>(acc, partials);
}
template <int K>
void residual_x0(hipStream_t s, int n, const double *b, double *r, double *r0, const double *sc,
                 double *partials, int *grid) {
  const int G = vec_grid_k(n, K);
  *grid = G;
  hipLaunchKernelGGL((k_residual_x0<K>), dim3(G), dim3(kBlock), 0, s, n, b, r, r0, sc, partials);
  CFSA_LAUNCH_CHECK();
}
template void residual_x0<1>(hipStream_t, int, const double *, double *, double *,
                             const double *, double *, int *);
template void residual_x0<3>(hipStream_t, int, const double *, double *, double *,
                             const double *, double *, int *);

template <class T, int K>
void cg_init(hipStream_t s, int n, const T *b, const T *invd, T *x, T *r, T *p, int scaled,
             double *partials, int *grid) {
  const int G = vec_grid_k(n, K);
  *grid = G;
  hipLaunchKernelGGL((k_cg_init<T, K>), dim3(G), dim3(kBlock), 0, s, n, b, invd, x, r, p,
                     scaled, partials);
  CFSA_LAUNCH_CHECK();
}
template <int K>
void cg_s_init(hipStream_t s, KrylovState *st, const double *partials, int G) {
  hipLaunchKernelGGL((k_cg_s_init<K>), dim3(1), dim3(kBlock), 0, s, st, partials, G);
  CFSA_LAUNCH_CHECK();
}
template <class T, int K>
void cg_xr(hipStream_t s, int n, KrylovState *st, const double *pq_partials, int Gq, int par,
           const T *q, const T *invd, T *r, int scaled, double *partials, int *grid) {
  const int G = vec_grid_k(n, K);
  *grid = G;
  hipLaunchKernelGGL((k_cg_xr<T, K>), dim3(G), dim3(kBlock), 0, s, n, st, pq_partials, Gq, par,
                     q, invd, r, scaled, partials);
  CFSA_LAUNCH_CHECK();
}
template <class T, int K>
void cg_p(hipStream_t s, int n, KrylovState *st, const double *partials, int Gv, int par,
          const T *r, const T *invd, T *p, T *x, const T *zin) {
  hipLaunchKernelGGL((k_cg_p<T, K>), dim3(vec_grid_k(n, K)), dim3(kBlock), 0, s, n, st, partials,
                     Gv, par, r, invd, zin, p, x);
  CFSA_LAUNCH_CHECK();
}

// CG with the polynomial preconditioner: x = 0, r = b, w = D^-1 b (the
// preconditioner product's input; its SpMV emits the initial pairs)
template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_cgp_init(int n, const T *__restrict__ b,
                                                     const T *__restrict__ invd,
                                                     T *__restrict__ x, T *__restrict__ r,
                                                     T *__restrict__ w) {
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
    const T bi = b[i];
    x[i] = T(0);
    r[i] = bi;
    w[i] = (T)((double)invd[i / K] * (double)bi);
  }
}
template <class T, int K>
void cgp_init(hipStream_t s, int n, const T *b, const T *invd, T *x, T *r, T *w) {
  hipLaunchKernelGGL((k_cgp_init<T, K>), dim3(vec_grid_k(n, K)), dim3(kBlock), 0, s, n, b, invd,
                     x, r, w);
  CFSA_LAUNCH_CHECK();
}

// k_cg_xr for the polynomial preconditioner: alpha = rho / p.Ap, r -= alpha q
// and w = D^-1 r (the next product's input; not stored when w == nullptr,
// the element-patch product then forms it); the stopping and rho pairs come
// from that product
template <class T, int K>
__global__ __launch_bounds__(kBlock) void k_cgp_xr(int n, KrylovState *__restrict__ st,
                                                   const double *__restrict__ pq_partials,
                                                   int Gq, int par, const T *__restrict__ q,
                                                   const T *__restrict__ invd,
                                                   T *__restrict__ r, T *__restrict__ w) {
  if (blockIdx.x == 0 && threadIdx.x < K) st->act_s[par][threadIdx.x] = st->active[threadIdx.x];
  if (!st->any_active) return;
  double pq[K];
  block_reduce_pairs<K>(pq_partials, Gq, pq);
  if (blockIdx.x == 0 && threadIdx.x < K && st->active[threadIdx.x]) {
    const int c = threadIdx.x;
    st->r0v[c] = pq[c];
    st->alpha[c] = st->rho_s[par][c] / pq[c];
  }
  const size_t i0 = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const int c = (int)(i0 % K);
  const bool act = st->active[c] != 0;
  const double al = act ? st->rho_s[par][c] / pq[c] : 0.0;
  const size_t N = act ? (size_t)n * K : 0, stride = (size_t)gridDim.x * kBlock;
  size_t i = i0;
  for (; i + 3 * stride < N; i += 4 * stride) {
    T rv[4], qv[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t j = i + u * stride;
      rv[u] = r[j];
      qv[u] = load_once<K>(&q[j]);
      if (w) dv[u] = invd[j / K];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T ri = (T)((double)rv[u] - al * (double)qv[u]);
      r[i + u * stride] = ri;
      if (w) w[i + u * stride] = (T)((double)dv[u] * (double)ri);
    }
  }
  for (; i < N; i += stride) {
    const T ri = (T)((double)r[i] - al * (double)load_once<K>(&q[i]));
    r[i] = ri;
    if (w) w[i] = (T)((double)invd[i / K] * (double)ri);
  }
}
template <class T, int K>
void cgp_xr(hipStream_t s, int n, KrylovState *st, const double *pq_partials, int Gq, int par,
            const T *q, const T *invd, T *r, T *w) {
  hipLaunchKernelGGL((k_cgp_xr<T, K>), dim3(vec_grid_k(n, K)), dim3(kBlock), 0, s, n, st,
                     pq_partials, Gq, par, q, invd, r, w);
  CFSA_LAUNCH_CHECK();
}

// Power iteration for the spectral radius of D^-1 A (K columns at once):
// v = D^-1 y with the pairs of |v|^2 per column, then x = v / |v|
template <int K>
__global__ __launch_bounds__(kBlock) void k_power_v(int n, const double *__restrict__ y,
                                                    const double *__restrict__ invd,
                                                    double *__restrict__ v,
                                                    double *__restrict__ partials) {
  DD acc[1];
  dd_zero(acc[0]);
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
    const double vi = invd[i / K] * y[i];
    v[i] = vi;
    dd_fma(acc[0], vi, vi);
  }
  block_partials_binned<K, 1>(acc, partials);
}
template <int K>
__global__ __launch_bounds__(kBlock) void k_power_scale(int n, const double *__restrict__ v,
                                                        double s0, double s1, double s2,
                                                        double *__restrict__ x) {
  const double sc[3] = {s0, s1, s2};
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride)
    x[i] = v[i] * sc[i % K];
}
// x0: a fixed pseudo-random vector, one column per component
template <int K>
__global__ __launch_bounds__(kBlock) void k_power_init(int n, double *__restrict__ x) {
  const size_t N = (size_t)n * K, stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
    const uint32_t h = (uint32_t)(i * 2654435761u) ^ (uint32_t)(i >> 32);
    x[i] = (double)(h >> 8) / 16777216.0 - 0.5;
  }
}
void power_init3(hipStream_t s, int n, double *x) {
  hipLaunchKernelGGL((k_power_init<3>), dim3(vec_grid_k(n, 3)), dim3(kBlock), 0, s, n, x);
  CFSA_LAUNCH_CHECK();
}
int power_v3(hipStream_t s, int n, const double *y, const double *invd, double *v,
             double *partials) {
  const int G = vec_grid_k(n, 3);
  hipLaunchKernelGGL((k_power_v<3>), dim3(G), dim3(kBlock), 0, s, n, y, invd, v, partials);
  CFSA_LAUNCH_CHECK();
  return G;
}
void power_scale3(hipStream_t s, int n, const double *v, const double *sc, double *x) {
  hipLaunchKernelGGL((k_power_scale<3>), dim3(vec_grid_k(n, 3)), dim3(kBlock), 0, s, n, v,
                     sc[0], sc[1], sc[2], x);
  CFSA_LAUNCH_CHECK();
}
#define CFSA_INST_CG(T, K)                                                                   \
  template void cg_init<T, K>(hipStream_t, int, const T *, const T *, T *, T *, T *, int,      \
                              double *, int *);                                              \
  template void cg_xr<T, K>(hipStream_t, int, KrylovState *, const double *, int, int,       \
                            const T *, const T *, T *, int, double *, int *);                \
  template void cg_p<T, K>(hipStream_t, int, KrylovState *, const double *, int, int,        \
                           const T *, const T *, T *, T *, const T *);                        \
  template void cgp_init<T, K>(hipStream_t, int, const T *, const T *, T *, T *, T *);        \
  template void cgp_xr<T, K>(hipStream_t, int, KrylovState *, const double *, int, int,      \
                             const T *, const T *, T *, T *);
CFSA_INST_CG(double, 1)
CFSA_INST_CG(double, 3)
CFSA_INST_CG(float, 1)
CFSA_INST_CG(float, 3)
template void cg_s_init<1>(hipStream_t, KrylovState *, const double *, int);
template void cg_s_init<3>(hipStream_t, KrylovState *, const double *, int);

// ===========================================================================
// Element assembly into the fixed patterns: row gather (default) and
// graph-coloured, atomic-free element scatter
// ===========================================================================
__device__ __forceinline__ int find_slot(const int *__restrict__ rowptr,
                                         const int *__restrict__ col, int row, int c) {
  int lo = rowptr[row], hi = rowptr[row + 1] - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

constexpr int kEdgeV[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

// Row node[a] of tet t: adds the tet's entries (a, 0..9) of M and K and, for a
// vertex (a < 4), its entries (a, 0..3) of L and G_x, G_y, G_z.
// reference: buildMassMatrix (NavierStokes.cpp:227-310), buildStiffnessMatrix
// (:135-225), buildGradientMatrices (:70-133), buildPoissonMatrix (:316-356).
__device__ __forceinline__ void tet_row_entries(
    int t, int a, const int (&node)[10], const double *__restrict__ vol, const double *__restrict__ grad, double rho,
    double inv_mu, const int *__restrict__ Arp, const int *__restrict__ Acol,
    double *__restrict__ M, double *__restrict__ K, const int *__restrict__ Prp,
    const int *__restrict__ Pcol, double *__restrict__ L, double *__restrict__ G3) {
  double g[10][3];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) g[i][c] = grad[(size_t)t * 12 + i * 3 + c];
#pragma unroll
  for (int e = 0; e < 6; ++e)
#pragma unroll
    for (int c = 0; c < 3; ++c) g[4 + e][c] = g[kEdgeV[e][0]][c] + g[kEdgeV[e][1]][c];
  const double v = vol[t];
  const double mass = rho * v;
  const double muv = inv_mu * v;
  const int row = node[a];
  for (int b = 0; b < 10; ++b) {
    double m;
    if (a < 4 && b < 4)
      m = (a == b) ? mass / 10.0 : mass / 20.0;
    else if (a >= 4 && b >= 4)
      m = (a == b) ? mass / 15.0 : mass / 60.0;
    else
      m = mass / 30.0;
    // vertex-edge pairs use grad_edge . grad_vertex in the reference
    const double kv = (a >= 4 && b < 4) || (a < 4 && b >= 4)
                          ? muv * dot3(g[a >= 4 ? a : b], g[a >= 4 ? b : a])
                          : muv * dot3(g[a], g[b]);
    const int slot = find_slot(Arp, Acol, row, node[b]);
    M[slot] += m;
    K[slot] += kv;
  }
  if (a < 4) {
    const double contrib = v / 4.0;
    for (int b = 0; b < 4; ++b) {
      const int slot = find_slot(Prp, Pcol, row, node[b]);
      L[slot] += v * dot3(g[a], g[b]);
#pragma unroll
      for (int c = 0; c < 3; ++c) G3[(size_t)slot * 3 + c] += contrib * g[b][c];
    }
  }
}

__device__ __forceinline__ void tet_nodes(int t, const int *__restrict__ tets,
                                          const int *__restrict__ ten, int nv, int (&node)[10]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) node[i] = tets[(size_t)t * 4 + i];
#pragma unroll
  for (int e = 0; e < 6; ++e) node[4 + e] = nv + ten[(size_t)t * 6 + e];
}

// Row gather: one thread per velocity node r walks the tets containing r in
// ascending tet order, so every matrix entry is summed in the order Eigen's
// setFromTriplets sums the reference's duplicates (triplets are inserted tet
// by tet), and the assembled values equal the reference's bit for bit.
// Atomic-free: each thread writes only its own row's slots.
__global__ __launch_bounds__(kBlock) void k_assemble_rows(
    int nu, const int *__restrict__ incp, const int *__restrict__ inct,
    const int *__restrict__ tets, const int *__restrict__ ten, int nv,
    const double *__restrict__ vol, const double *__restrict__ grad, double rho,
    double inv_mu, const int *__restrict__ Arp, const int *__restrict__ Acol,
    double *__restrict__ M, double *__restrict__ K, const int *__restrict__ Prp,
    const int *__restrict__ Pcol, double *__restrict__ L, double *__restrict__ G3) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nu) return;
  for (int k = incp[r]; k < incp[r + 1]; ++k) {
    const int t = inct[k];
    int node[10];
    tet_nodes(t, tets, ten, nv, node);
    int a = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i)
      if (node[i] == r) a = i;
    tet_row_entries(t, a, node, vol, grad, rho, inv_mu, Arp, Acol, M, K, Prp, Pcol, L,
                    G3);
  }
}

// Slot map of the row gather (built once per pattern): for incidence entry k
// (row r, tet t = inct[k]), loc[k] = local index a of r in t, amap[10 k + b] =
// slot of (r, node b of t) in the P2 pattern and, for vertex rows (k <
// incp[nv]), pmap[4 k + b] = slot of (r, vertex b) in the P1 pattern.
__global__ __launch_bounds__(kBlock) void k_slot_map(
    int nu, const int *__restrict__ incp, const int *__restrict__ inct,
    const int *__restrict__ tets, const int *__restrict__ ten, int nv,
    const int *__restrict__ Arp, const int *__restrict__ Acol, const int *__restrict__ Prp,
    const int *__restrict__ Pcol, uint8_t *__restrict__ loc, int *__restrict__ amap,
    int *__restrict__ pmap) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nu) return;
  for (int k = incp[r]; k < incp[r + 1]; ++k) {
    int node[10];
    tet_nodes(inct[k], tets, ten, nv, node);
    int a = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i)
      if (node[i] == r) a = i;
    loc[k] = (uint8_t)a;
    for (int b = 0; b < 10; ++b) amap[(size_t)k * 10 + b] = find_slot(Arp, Acol, r, node[b]);
    if (r < nv)
      for (int b = 0; b < 4; ++b) pmap[(size_t)k * 4 + b] = find_slot(Prp, Pcol, r, node[b]);
  }
}

void build_slot_map(hipStream_t s, int nu, const int *incp, const int *inct, const int *tets,
                    const int *ten, int nv, const DevCsr &A, const DevCsr &P, uint8_t *loc,
                    int *amap, int *pmap) {
  if (nu <= 0) return;
  hipLaunchKernelGGL(k_slot_map, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, incp, inct,
                     tets, ten, nv, A.rowptr, A.col, P.rowptr, P.col, loc, amap, pmap);
  CFSA_LAUNCH_CHECK();
}

// Row gather with the slot map: the same contributions, in the same ascending
// tet order per row, without searching the rows for the slots.
__global__ __launch_bounds__(kBlock) void k_assemble_mapped(
    int nu, int nv, const int *__restrict__ incp, const int *__restrict__ inct,
    const uint8_t *__restrict__ loc, const int *__restrict__ amap,
    const int *__restrict__ pmap, const double *__restrict__ vol,
    const double *__restrict__ grad, double rho, double inv_mu, double *__restrict__ M,
    double *__restrict__ K, double *__restrict__ L, double *__restrict__ G3) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nu) return;
  for (int k = incp[r]; k < incp[r + 1]; ++k) {
    const int t = inct[k];
    const int a = loc[k];
    double g[10][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) g[i][c] = grad[(size_t)t * 12 + i * 3 + c];
#pragma unroll
    for (int e = 0; e < 6; ++e)
#pragma unroll
      for (int c = 0; c < 3; ++c) g[4 + e][c] = g[kEdgeV[e][0]][c] + g[kEdgeV[e][1]][c];
    const double v = vol[t];
    const double mass = rho * v;
    const double muv = inv_mu * v;
    int slot[10];
#pragma unroll
    for (int b = 0; b < 10; ++b) slot[b] = amap[(size_t)k * 10 + b];
#pragma unroll
    for (int b = 0; b < 10; ++b) {
      double m;
      if (a < 4 && b < 4)
        m = (a == b) ? mass / 10.0 : mass / 20.0;
      else if (a >= 4 && b >= 4)
        m = (a == b) ? mass / 15.0 : mass / 60.0;
      else
        m = mass / 30.0;
      // vertex-edge pairs use grad_edge . grad_vertex in the reference
      const double kv = (a >= 4 && b < 4) || (a < 4 && b >= 4)
                            ? muv * dot3(g[a >= 4 ? a : b], g[a >= 4 ? b : a])
                            : muv * dot3(g[a], g[b]);
      M[slot[b]] += m;
      K[slot[b]] += kv;
    }
    if (r < nv) {
      const double contrib = v / 4.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int ps = pmap[(size_t)k * 4 + b];
        L[ps] += v * dot3(g[a], g[b]);
#pragma unroll
        for (int c = 0; c < 3; ++c) G3[(size_t)ps * 3 + c] += contrib * g[b][c];
      }
    }
  }
}

// Row gather with six matrix rows per wavefront (a group of 10 lanes per row,
// lane b adding entry b of each incident tet's 10 x 10 block; the 10 slots of
// one tet are distinct), 24 rows per block.  Each row's accumulators live in
// LDS; a group visits its row's tets in ascending order (one tet ahead loaded),
// then writes the row once and its lane 0 forms the lumped-mass inverse in CSR
// order (k_lumped).  Same arithmetic and order as k_assemble_mapped, so the
// same bits.  Dynamic LDS: 24 x (2 maxA + 4 maxP) doubles.
constexpr int kAsmGroup = 10;
constexpr int kAsmGroupsPerWave = 6;
constexpr int kAsmRows = (kBlock / 64) * kAsmGroupsPerWave;

// Vertex gradient q of a tet (selects over the four, no indexed registers).
__device__ __forceinline__ void asm_vert(const double (&g4)[4][3], int q, double (&o)[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) o[c] = g4[0][c];
#pragma unroll
  for (int i = 1; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = q == i ? g4[i][c] : o[c];
}
// "Gradient" of local node q: the P1 gradient for a vertex, g[v0] + g[v1] for
// the edge (v0, v1) (NavierStokes.cpp:184-194), edges in kEdgeV order.
__device__ __forceinline__ void asm_node(const double (&g4)[4][3], int q, double (&o)[3]) {
  if (q < 4) {
    asm_vert(g4, q, o);
  } else {
    const int e = q - 4;  // 01 02 03 12 13 23
    const int v0 = e < 3 ? 0 : (e < 5 ? 1 : 2);
    const int v1 = e < 3 ? e + 1 : (e < 5 ? e - 1 : 3);
    double x0[3], x1[3];
    asm_vert(g4, v0, x0);
    asm_vert(g4, v1, x1);
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = x0[c] + x1[c];
  }
}

__global__ __launch_bounds__(kBlock) void k_assemble_waves(
    int nu, int nv, int maxA, int maxP, const int *__restrict__ incp,
    const int *__restrict__ inct, const uint8_t *__restrict__ loc,
    const int *__restrict__ amap, const int *__restrict__ pmap,
    const int *__restrict__ rowptrA, const int *__restrict__ rowptrP,
    const double *__restrict__ vol, const double *__restrict__ grad, double rho, double inv_mu,
    double *__restrict__ M, double *__restrict__ K, double *__restrict__ L,
    double *__restrict__ G3, double *__restrict__ ilm) {
  extern __shared__ double asm_lds[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int grp = lane / kAsmGroup, b = lane % kAsmGroup;  // grp 6: idle lanes 60..63
  const int base = grp * kAsmGroup;                        // first lane of the group
  const int r = (blockIdx.x * (kBlock / 64) + wv) * kAsmGroupsPerWave + grp;
  const bool live = grp < kAsmGroupsPerWave && r < nu;
  const int slot_grp = live ? wv * kAsmGroupsPerWave + grp : 0;
  double *accM = asm_lds + (size_t)slot_grp * (2 * maxA + 4 * maxP);
  double *accK = accM + maxA;
  double *accL = accK + maxA;
  double *accG = accL + maxP;
  int a0 = 0, la = 0, p0 = 0, lp = 0, kb = 0, ke = 0;
  const bool prow = live && r < nv;
  if (live) {
    a0 = rowptrA[r];
    la = rowptrA[r + 1] - a0;
    kb = incp[r];
    ke = incp[r + 1];
  }
  if (prow) {
    p0 = rowptrP[r];
    lp = rowptrP[r + 1] - p0;
  }
  for (int i = b; i < la; i += kAsmGroup) {
    accM[i] = 0.0;
    accK[i] = 0.0;
  }
  for (int i = b; i < lp; i += kAsmGroup) {
    accL[i] = 0.0;
    accG[3 * i] = accG[3 * i + 1] = accG[3 * i + 2] = 0.0;
  }
  __builtin_amdgcn_wave_barrier();
  const int cnt = ke - kb;  // incident tets of this lane's row (0 when idle)
  double gn[4][3] = {}, vn = 0.0;
  int sln = 0, psn = 0;
  int tl = 0, al = 0;
  for (int j = 0; __ballot(j < cnt) != 0; ++j) {
    // the group's lane q holds tet j0 + q of its row (j0 = j rounded down to 10)
    if (j % kAsmGroup == 0) {
      const int kk = kb + min(j + b, max(cnt - 1, 0));
      tl = live && cnt > 0 ? inct[kk] : 0;
      al = live && cnt > 0 ? loc[kk] : 0;
      if (j == 0) {
        const int t = __shfl(tl, base, 64);
        if (cnt > 0) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) gn[i][c] = grad[(size_t)t * 12 + i * 3 + c];
          vn = vol[t];
          sln = amap[(size_t)kb * 10 + b] - a0;
          if (prow && b < 4) psn = pmap[(size_t)kb * 4 + b] - p0;
        }
      }
    }
    const bool act = j < cnt;
    double g4[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) g4[i][c] = gn[i][c];
    const double v = vn;
    const int sl = sln, ps = psn;
    const int a = __shfl(al, base + j % kAsmGroup, 64);
    // one tet ahead: its id is in the current group of 10, or the next one
    const int jn = j + 1;
    int tn;
    if (jn % kAsmGroup == 0) {
      const int kk = kb + min(jn, max(cnt - 1, 0));
      tn = live && jn < cnt ? inct[kk] : 0;
    } else {
      tn = __shfl(tl, base + jn % kAsmGroup, 64);
    }
    if (jn < cnt) {
      const size_t kn = (size_t)kb + jn;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) gn[i][c] = grad[(size_t)tn * 12 + i * 3 + c];
      vn = vol[tn];
      sln = amap[kn * 10 + b] - a0;
      if (prow && b < 4) psn = pmap[kn * 4 + b] - p0;
    }
    if (act) {
      double ga[3], gb[3];
      asm_node(g4, a, ga);
      asm_node(g4, b, gb);
      const double mass = rho * v;
      const double muv = inv_mu * v;
      double m;
      if (a < 4 && b < 4)
        m = (a == b) ? mass / 10.0 : mass / 20.0;
      else if (a >= 4 && b >= 4)
        m = (a == b) ? mass / 15.0 : mass / 60.0;
      else
        m = mass / 30.0;
      // vertex-edge pairs use grad_edge . grad_vertex in the reference
      const double kv = (a < 4 && b >= 4) ? muv * dot3(gb, ga) : muv * dot3(ga, gb);
      accM[sl] += m;
      accK[sl] += kv;
      if (prow && b < 4) {
        accL[ps] += v * dot3(ga, gb);
        const double contrib = v / 4.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) accG[3 * ps + c] += contrib * gb[c];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  for (int i = b; i < la; i += kAsmGroup) {
    M[(size_t)a0 + i] = accM[i];
    K[(size_t)a0 + i] = accK[i];
  }
  for (int i = b; i < lp; i += kAsmGroup) {
    L[(size_t)p0 + i] = accL[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) G3[((size_t)p0 + i) * 3 + c] = accG[3 * i + c];
  }
  if (live && b == 0) {
    // inv_lumped_mass_ = 1 / (M * ones), NavierStokes.cpp:290-306
    double sum = 0.0;
    for (int i = 0; i < la; ++i) sum += accM[i] * 1.0;
    ilm[r] = (fabs(sum) > 1e-14) ? 1.0 / sum : 1.0;
  }
}

// Launches k_assemble_waves; false (nothing launched) when the rows'
// accumulators do not fit in LDS, for the caller to fall back.
bool assemble_waves(hipStream_t s, int nu, int nv, int maxA, int maxP, const int *incp,
                    const int *inct, const uint8_t *loc, const int *amap, const int *pmap,
                    const DevCsr &A, const DevCsr &P, const double *vol, const double *grad,
                    double rho, double inv_mu, double *M, double *K, double *L, double *G3,
                    double *ilm) {
  const size_t lds = (size_t)kAsmRows * (2 * maxA + 4 * maxP) * sizeof(double);
  if (lds > 64 * 1024) return false;
  if (nu <= 0) return true;
  hipLaunchKernelGGL(k_assemble_waves, dim3(div_up(nu, kAsmRows)), dim3(kBlock), lds, s, nu, nv,
                     maxA, maxP, incp, inct, loc, amap, pmap, A.rowptr, P.rowptr, vol, grad,
                     rho, inv_mu, M, K, L, G3, ilm);
  CFSA_LAUNCH_CHECK();
  return true;
}

void assemble_mapped(hipStream_t s, int nu, int nv, const int *incp, const int *inct,
                     const uint8_t *loc, const int *amap, const int *pmap, const double *vol,
                     const double *grad, double rho, double inv_mu, double *M, double *K,
                     double *L, double *G3) {
  if (nu <= 0) return;
  hipLaunchKernelGGL(k_assemble_mapped, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, nv,
                     incp, inct, loc, amap, pmap, vol, grad, rho, inv_mu, M, K, L, G3);
  CFSA_LAUNCH_CHECK();
}

void assemble_rows(hipStream_t s, int nu, const int *incp, const int *inct, const int *tets,
                   const int *ten, int nv, const double *vol, const double *grad, double rho,
                   double inv_mu, const DevCsr &A, double *M, double *K, const DevCsr &P,
                   double *L, double *G3) {
  if (nu <= 0) return;
  hipLaunchKernelGGL(k_assemble_rows, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, incp,
                     inct, tets, ten, nv, vol, grad, rho, inv_mu, A.rowptr, A.col, M, K,
                     P.rowptr, P.col, L, G3);
  CFSA_LAUNCH_CHECK();
}

// Coloured scatter: one thread per (tet, local node a) of one colour; tets of
// one colour share no vertex, so no two threads touch the same row.  Entries
// are summed in colour order, which differs from the reference's tet order in
// the last bits.
__global__ __launch_bounds__(kBlock) void k_assemble(
    int ntc, const int *__restrict__ order, const int *__restrict__ tets,
    const int *__restrict__ ten, int nv, const double *__restrict__ vol,
    const double *__restrict__ grad, double rho, double inv_mu,
    const int *__restrict__ Arp, const int *__restrict__ Acol, double *__restrict__ M,
    double *__restrict__ K, const int *__restrict__ Prp, const int *__restrict__ Pcol,
    double *__restrict__ L, double *__restrict__ G3) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  const int tl = idx / 10, a = idx % 10;
  if (tl >= ntc) return;
  const int t = order[tl];
  int node[10];
  tet_nodes(t, tets, ten, nv, node);
  tet_row_entries(t, a, node, vol, grad, rho, inv_mu, Arp, Acol, M, K, Prp, Pcol, L, G3);
}

void assemble_colour(hipStream_t s, int ntc, const int *order, const int *tets,
                     const int *ten, int nv, const double *vol, const double *grad,
                     double rho, double inv_mu, const DevCsr &A, double *M, double *K,
                     const DevCsr &P, double *L, double *G3) {
  if (ntc <= 0) return;
  hipLaunchKernelGGL(k_assemble, dim3(div_up((int64_t)ntc * 10, kBlock)), dim3(kBlock), 0,
                     s, ntc, order, tets, ten, nv, vol, grad, rho, inv_mu, A.rowptr, A.col,
                     M, K, P.rowptr, P.col, L, G3);
  CFSA_LAUNCH_CHECK();
}

// Tet volumes and P1 gradients from the vertex coordinates, the same
// expressions in the same order as the host mesh layer (Mesh::tetGeometry;
// reference Mesh.cpp:455 and computeShapeFunctionGradients, :590-614), so a
// moved mesh is re-measured on the GPU with identical bits.
struct DV3 {
  double x, y, z;
};
__device__ __forceinline__ DV3 dsub(DV3 a, DV3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ DV3 dcross(DV3 l, DV3 r) {
  return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x};
}
__device__ __forceinline__ double ddot(DV3 a, DV3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ DV3 dscale(double s, DV3 a) { return {s * a.x, s * a.y, s * a.z}; }

__global__ __launch_bounds__(kBlock) void k_tet_geometry(int nt, const double *__restrict__ xyz,
                                                         const int *__restrict__ tets,
                                                         double *__restrict__ vol,
                                                         double *__restrict__ grad) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  DV3 p[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int v = tets[(size_t)t * 4 + i];
    p[i] = {xyz[(size_t)v * 3], xyz[(size_t)v * 3 + 1], xyz[(size_t)v * 3 + 2]};
  }
  vol[t] = fabs(ddot(dcross(dsub(p[1], p[0]), dsub(p[2], p[0])), dsub(p[3], p[0])) / 6.0);
  const DV3 e1 = dsub(p[1], p[0]), e2 = dsub(p[2], p[0]), e3 = dsub(p[3], p[0]);
  const double sixv = ddot(e1, dcross(e2, e3));
  const double inv = 1.0 / sixv;
  const DV3 g[4] = {dscale(inv, dcross(dsub(p[3], p[1]), dsub(p[2], p[1]))),
                    dscale(inv, dcross(dsub(p[2], p[0]), dsub(p[3], p[0]))),
                    dscale(inv, dcross(dsub(p[3], p[0]), dsub(p[1], p[0]))),
                    dscale(inv, dcross(dsub(p[1], p[0]), dsub(p[2], p[0])))};
  double *o = grad + (size_t)t * 12;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    o[i * 3] = g[i].x;
    o[i * 3 + 1] = g[i].y;
    o[i * 3 + 2] = g[i].z;
  }
}
void tet_geometry(hipStream_t s, int nt, const double *xyz, const int *tets, double *vol,
                  double *grad) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_tet_geometry, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, s, nt, xyz,
                     tets, vol, grad);
  CFSA_LAUNCH_CHECK();
}

// inv_lumped_mass_ = 1 / (M * ones), reference NavierStokes.cpp:290-306
__global__ void k_lumped(int n, const int *__restrict__ rp, const double *__restrict__ M,
                         double *__restrict__ ilm) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; ++k) s += M[k] * 1.0;
  ilm[i] = (fabs(s) > 1e-14) ? 1.0 / s : 1.0;
}
void lumped_mass_inverse(hipStream_t s, const DevCsr &A, const double *M, double *ilm) {
  hipLaunchKernelGGL(k_lumped, dim3(div_up(A.n, kBlock)), dim3(kBlock), 0, s, A.n, A.rowptr,
                     M, ilm);
  CFSA_LAUNCH_CHECK();
}

// lhs = M/dt + K (reference NavierStokes.cpp:460)
__global__ void k_lhs(int64_t nnz, const double *__restrict__ M,
                      const double *__restrict__ K, double dt, double *__restrict__ out) {
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * kBlock)
    out[k] = M[k] / dt + K[k];
}
void build_lhs(hipStream_t s, int64_t nnz, const double *M, const double *K, double dt,
               double *lhs) {
  hipLaunchKernelGGL(k_lhs, dim3(4096), dim3(kBlock), 0, s, nnz, M, K, dt, lhs);
  CFSA_LAUNCH_CHECK();
}

// mode 0: Inlet/Wall rows (momentum); 1: Outlet rows (Dirichlet pressure);
// 2: only row `ref` (Neumann pressure reference node)
__global__ void k_diag_penalty(int n, const int *__restrict__ diag,
                               const uint8_t *__restrict__ bc, int mode, int ref,
                               double pen, double *__restrict__ A) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  bool on;
  if (mode == 0)
    on = bc[i] == kInlet || bc[i] == kWall;
  else if (mode == 1)
    on = bc[i] == kOutlet;
  else
    on = (i == ref);
  if (on) A[diag[i]] += pen;
}
void add_diag_penalty(hipStream_t s, int n, const int *diag, const uint8_t *bc, int mode,
                      int ref, double penalty, double *A) {
  hipLaunchKernelGGL(k_diag_penalty, dim3(div_up(n, kBlock)), dim3(kBlock), 0, s, n, diag,
                     bc, mode, ref, penalty, A);
  CFSA_LAUNCH_CHECK();
}

// Eigen DiagonalPreconditioner: 1/A_jj when the entry exists and is nonzero.
__global__ void k_invdiag(int n, const int *__restrict__ diag, const double *__restrict__ A,
                          double *__restrict__ invd) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int d = diag[i];
  invd[i] = (d >= 0 && A[d] != 0.0) ? 1.0 / A[d] : 1.0;
}
void inv_diag(hipStream_t s, int n, const int *diag, const double *A, double *invd) {
  hipLaunchKernelGGL(k_invdiag, dim3(div_up(n, kBlock)), dim3(kBlock), 0, s, n, diag, A,
                     invd);
  CFSA_LAUNCH_CHECK();
}

// ===========================================================================
// Convection / divergence: per-tet value, then a node gather over the
// node->tet incidence in ascending tet order (the reference's scatter order).
// ===========================================================================
// grad_u(a,b) = sum_j u_j(a) * gphi_j(b), accumulated j = 0..3 from zero.
__device__ __forceinline__ void tet_grad_u(const double (&u)[4][3], const double (&g)[4][3],
                                           double (&G)[3][3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) G[a][b] = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) G[a][b] = G[a][b] + u[j][a] * g[j][b];
}

// reference: computeAdvectionRHS per-tet part (NavierStokes.cpp:550-567)
__global__ __launch_bounds__(kBlock) void k_conv_tet(int nt, const int *__restrict__ tets,
                                                     const double *__restrict__ grad,
                                                     const double *__restrict__ uvel,
                                                     double *__restrict__ cT) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  double u[4][3], g[4][3];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int v = tets[(size_t)t * 4 + j];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      u[j][c] = uvel[(size_t)v * 3 + c];
      g[j][c] = grad[(size_t)t * 12 + j * 3 + c];
    }
  }
  double ue[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) ue[c] = ue[c] + 0.25 * u[j][c];
  double G[3][3];
  tet_grad_u(u, g, G);
#pragma unroll
  for (int a = 0; a < 3; ++a)
    cT[(size_t)t * 3 + a] = (G[a][0] * ue[0] + G[a][1] * ue[1]) + G[a][2] * ue[2];
}
void conv_tet(hipStream_t s, int nt, const int *tets, const double *grad, const double *u,
              double *cT) {
  hipLaunchKernelGGL(k_conv_tet, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, s, nt, tets,
                     grad, u, cT);
  CFSA_LAUNCH_CHECK();
}

// conv[node] -= (vol/4) c_T for vertices, (vol/24) c_T for edge nodes
// (NavierStokes.cpp:571-580)
__global__ __launch_bounds__(kBlock) void k_conv_gather(int nu, int nv,
                                                        const int *__restrict__ ip,
                                                        const int *__restrict__ it,
                                                        const double *__restrict__ vol,
                                                        const double *__restrict__ cT,
                                                        double *__restrict__ conv) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nu) return;
  const double div = (i < nv) ? 4.0 : 24.0;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int q = ip[i]; q < ip[i + 1]; ++q) {
    const int t = it[q];
    const double w = vol[t] / div;
    a0 = a0 - w * cT[(size_t)t * 3 + 0];
    a1 = a1 - w * cT[(size_t)t * 3 + 1];
    a2 = a2 - w * cT[(size_t)t * 3 + 2];
  }
  conv[(size_t)i * 3 + 0] = a0;
  conv[(size_t)i * 3 + 1] = a1;
  conv[(size_t)i * 3 + 2] = a2;
}
void conv_gather(hipStream_t s, int nu, int nv, const int *inc_ptr, const int *inc_tet,
                 const double *vol, const double *cT, double *conv) {
  hipLaunchKernelGGL(k_conv_gather, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, nv,
                     inc_ptr, inc_tet, vol, cT, conv);
  CFSA_LAUNCH_CHECK();
}

// reference: computeDivergence per-tet part (NavierStokes.cpp:684-697)
__global__ __launch_bounds__(kBlock) void k_div_tet(int nt, const int *__restrict__ tets,
                                                    const double *__restrict__ grad,
                                                    const double *__restrict__ vol,
                                                    const double *__restrict__ uvel,
                                                    double *__restrict__ wT) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  double u[4][3], g[4][3];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int v = tets[(size_t)t * 4 + j];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      u[j][c] = uvel[(size_t)v * 3 + c];
      g[j][c] = grad[(size_t)t * 12 + j * 3 + c];
    }
  }
  double G[3][3];
  tet_grad_u(u, g, G);
  const double d = (G[0][0] + G[1][1]) + G[2][2];
  wT[t] = (vol[t] / 4.0) * d;
}
void div_tet(hipStream_t s, int nt, const int *tets, const double *grad, const double *vol,
             const double *u, double *wT) {
  hipLaunchKernelGGL(k_div_tet, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, s, nt, tets, grad,
                     vol, u, wT);
  CFSA_LAUNCH_CHECK();
}

// pressure rhs, reference NavierStokes.cpp:598-626
__global__ __launch_bounds__(kBlock) void k_div_gather(
    int nv, const int *__restrict__ ip, const int *__restrict__ it,
    const double *__restrict__ wT, const uint8_t *__restrict__ bc, int mode, int ref,
    double dt, double p_out, double pen, double *__restrict__ div,
    double *__restrict__ rhs) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nv) return;
  double a = 0.0;
  for (int q = ip[i]; q < ip[i + 1]; ++q) a = a + wT[it[q]];
  div[i] = a;
  double r = -a / dt;
  if (mode == 1 && bc[i] == kOutlet) r = r + pen * p_out;
  if (mode == 2 && i == ref) r = r + pen * 0.0;
  rhs[i] = r;
}
void div_gather_prhs(hipStream_t s, int nv, const int *inc_ptr, const int *inc_tet,
                     const double *wT, const uint8_t *bc, int outlet_mode, int ref,
                     double dt, double p_out, double penalty, double *div,
                     double *rhs_p) {
  hipLaunchKernelGGL(k_div_gather, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, s, nv,
                     inc_ptr, inc_tet, wT, bc, outlet_mode, ref, dt, p_out, penalty, div,
                     rhs_p);
  CFSA_LAUNCH_CHECK();
}

// p += relax_p p'; then Dirichlet outlet reset (NavierStokes.cpp:657-662,1011-1024)
__global__ void k_p_update(int nv, double *__restrict__ p, const double *__restrict__ pc,
                           double relax, const uint8_t *__restrict__ bc, int dirichlet,
                           double p_out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nv) return;
  double v = p[i] + relax * pc[i];
  if (dirichlet && bc[i] == kOutlet) v = p_out;
  p[i] = v;
}
void pressure_update(hipStream_t s, int nv, double *p, const double *pc, double relax_p,
                     const uint8_t *bc, int dirichlet, double p_out) {
  hipLaunchKernelGGL(k_p_update, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, s, nv, p, pc,
                     relax_p, bc, dirichlet, p_out);
  CFSA_LAUNCH_CHECK();
}

__global__ void k_p_reset(int nv, double *__restrict__ p, const uint8_t *__restrict__ bc,
                          double p_out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nv && bc[i] == kOutlet) p[i] = p_out;
}
void pressure_outlet_reset(hipStream_t s, int nv, double *p, const uint8_t *bc, double p_out) {
  hipLaunchKernelGGL(k_p_reset, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, s, nv, p, bc, p_out);
  CFSA_LAUNCH_CHECK();
}

__global__ void k_copy1(const double *src, double *dst) { *dst = *src; }
__global__ void k_sub_scalar(int n, double *__restrict__ p, const double *__restrict__ s) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  p[i] = p[i] - *s;
}
// normalizePressire (NavierStokes.cpp:804-813)
void pressure_normalize(hipStream_t s, int nv, double *p, int ref, double *scratch) {
  if (ref < 0 || ref >= nv) return;
  hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, s, p + ref, scratch);
  CFSA_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sub_scalar, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, s, nv, p,
                     scratch);
  CFSA_LAUNCH_CHECK();
}

// correctVelocity (NavierStokes.cpp:725-768)
__global__ __launch_bounds__(kBlock) void k_correct(
    int nu, int nv, const uint8_t *__restrict__ bc, const double *__restrict__ g,
    const double *__restrict__ ilm, const double *__restrict__ gp,
    const double *__restrict__ us, double dt, double relax, double *__restrict__ u) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nu) return;
  const uint8_t b = bc[i];
  const size_t o = (size_t)i * 3;
  if (b == kInlet) {
    u[o] = g[o];
    u[o + 1] = g[o + 1];
    u[o + 2] = g[o + 2];
  } else if (b == kWall) {
    u[o] = 0.0;
    u[o + 1] = 0.0;
    u[o + 2] = 0.0;
  } else if (i < nv && ilm[i] > 1e-14) {
    const double sc = -dt * ilm[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) u[o + c] = us[o + c] + relax * (sc * gp[o + c]);
  } else {
    u[o] = us[o];
    u[o + 1] = us[o + 1];
    u[o + 2] = us[o + 2];
  }
}
void correct_velocity(hipStream_t s, int nu, int nv, const uint8_t *bc, const double *ginlet,
                      const double *ilm, const double *gp, const double *us, double dt,
                      double relax_u, double *u) {
  hipLaunchKernelGGL(k_correct, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, nv, bc,
                     ginlet, ilm, gp, us, dt, relax_u, u);
  CFSA_LAUNCH_CHECK();
}

// reenforceVelocityBCs (NavierStokes.cpp:815-839)
__global__ void k_bc_reset(int nu, const uint8_t *__restrict__ bc,
                           const double *__restrict__ g, double *__restrict__ u,
                           int inlet_only) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nu) return;
  const uint8_t b = bc[i];
  const size_t o = (size_t)i * 3;
  if (b == kInlet) {
    u[o] = g[o];
    u[o + 1] = g[o + 1];
    u[o + 2] = g[o + 2];
  } else if (b == kWall && !inlet_only) {
    u[o] = 0.0;
    u[o + 1] = 0.0;
    u[o + 2] = 0.0;
  }
}
void reset_velocity_bc(hipStream_t s, int nu, const uint8_t *bc, const double *ginlet,
                       double *u) {
  hipLaunchKernelGGL(k_bc_reset, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, bc,
                     ginlet, u, 0);
  CFSA_LAUNCH_CHECK();
}
// initialize(): inlet nodes take the inlet value (NavierStokes.cpp:33-46)
void set_inlet_nodes(hipStream_t s, int nu, const uint8_t *bc, const double *ginlet,
                     double *u) {
  hipLaunchKernelGGL(k_bc_reset, dim3(div_up(nu, kBlock)), dim3(kBlock), 0, s, nu, bc,
                     ginlet, u, 1);
  CFSA_LAUNCH_CHECK();
}

__global__ void k_fill3(int n, double a, double b, double c, double *__restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  out[(size_t)i * 3] = a;
  out[(size_t)i * 3 + 1] = b;
  out[(size_t)i * 3 + 2] = c;
}
void fill3(hipStream_t s, int n, const double v[3], double *out) {
  hipLaunchKernelGGL(k_fill3, dim3(div_up(n, kBlock)), dim3(kBlock), 0, s, n, v[0], v[1],
                     v[2], out);
  CFSA_LAUNCH_CHECK();
}

__global__ void k_nan(int n, const double *__restrict__ a, int *__restrict__ flag) {
  int f = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    f |= isnan(a[i]) ? 1 : 0;
  if (f) *flag = 1;
}
void nan_check(hipStream_t s, int n, const double *a, int *flag) {
  hipLaunchKernelGGL(k_nan, dim3(vec_grid(n)), dim3(kBlock), 0, s, n, a, flag);
  CFSA_LAUNCH_CHECK();
}

template <class TO, class TI>
__global__ void k_convert(int64_t n, const TI *__restrict__ in, TO *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    out[i] = (TO)in[i];
}
template <class TO, class TI>
void convert(hipStream_t s, int64_t n, const TI *in, TO *out) {
  hipLaunchKernelGGL((k_convert<TO, TI>), dim3(4096), dim3(kBlock), 0, s, n, in, out);
  CFSA_LAUNCH_CHECK();
}
template void convert<float, double>(hipStream_t, int64_t, const double *, float *);
template void convert<double, float>(hipStream_t, int64_t, const float *, double *);

// ===========================================================================
// AMG-preconditioned CG for the pressure system (pressure_solver = 2): the
// V-cycle of a smoothed-aggregation hierarchy (set up on the host, amg.cpp)
// with degree-2 Chebyshev smoothing, and the CG iteration around it.  Every
// smoothing step is one SpMV whose epilogue applies the polynomial update, so a
// level costs four operator products plus one restriction and one
// prolongation.  The V-cycle runs in T (double, or float for the mixed
// precision = 1 mode: values, vectors and SpMV accumulation); the CG around it
// is FP64 (iterate, residual, scalars, stopping test).  All kernels take the
// solve's `active` flag as their mask and exit at once after the solve stopped.
// ===========================================================================
template <class T>
struct OpAmgBase : Mask1Base {
  const int *__restrict__ col;
  const T *__restrict__ val;
  const T *__restrict__ x;  // the gathered vector
  struct Frag {
    int j;
    T a;
  };
  __device__ Frag fetch(int k) const { return Frag{col[k], val[k]}; }
  __device__ void prod(const Frag &f, T *p) const { p[0] = f.a * x[f.j]; }
};

// pre-smoothing from x0 = 0, second step: Y = X + c1 X + c2 D^-1 (b - A X)
// (the first step X = c0 D^-1 b is fused into the kernel that produced b)
template <class T>
struct OpChebPre : OpAmgBase<T> {
  const T *__restrict__ b;
  const T *__restrict__ dinv;
  T *__restrict__ y;
  T c1, c2;
  struct Pre {
    T b, x, d;
  };
  __device__ Pre pre(int row, int) const { return Pre{b[row], this->x[row], dinv[row]}; }
  __device__ void epi(int row, int, T acc, const Pre &pr, DD *) const {
    y[row] = (pr.x + c1 * pr.x) + c2 * (pr.d * (pr.b - acc));
  }
};

// r = b - A Y
template <class T>
struct OpAmgResidual : OpAmgBase<T> {
  const T *__restrict__ b;
  T *__restrict__ r;
  struct Pre {
    T b;
  };
  __device__ Pre pre(int row, int) const { return Pre{b[row]}; }
  __device__ void epi(int row, int, T acc, const Pre &pr, DD *) const { r[row] = pr.b - acc; }
};

// b_c = R r; X_c = c0 D_c^-1 b_c (first Chebyshev step of the coarse level)
// unless the coarse level is solved directly (xc == nullptr)
template <class T>
struct OpRestrict : OpAmgBase<T> {
  const T *__restrict__ dinvc;
  T *__restrict__ bc;
  T *__restrict__ xc;
  T c0;
  struct Pre {
    T d;
  };
  __device__ Pre pre(int row, int) const { return Pre{xc ? dinvc[row] : T(0)}; }
  __device__ void epi(int row, int, T acc, const Pre &pr, DD *) const {
    bc[row] = acc;
    if (xc) xc[row] = c0 * (pr.d * acc);
  }
};

// Y += P e_c
template <class T>
struct OpProlong : OpAmgBase<T> {
  T *__restrict__ y;
  struct Pre {
    T y;
  };
  __device__ Pre pre(int row, int) const { return Pre{y[row]}; }
  __device__ void epi(int row, int, T acc, const Pre &pr, DD *) const { y[row] = pr.y + acc; }
};

// post-smoothing from x0 = Y, first step: X = Y + c0 D^-1 (b - A Y)
template <class T>
struct OpChebPost1 : OpAmgBase<T> {
  const T *__restrict__ b;
  const T *__restrict__ dinv;
  T *__restrict__ xo;
  T c0;
  struct Pre {
    T b, y, d;
  };
  __device__ Pre pre(int row, int) const { return Pre{b[row], this->x[row], dinv[row]}; }
  __device__ void epi(int row, int, T acc, const Pre &pr, DD *) const {
    xo[row] = pr.y + c0 * (pr.d * (pr.b - acc));
  }
};

// second step: Y = X + c1 (X - Y) + c2 D^-1 (b - A X); on the finest level the
// pairs of r . Y (the CG's r . z, with the FP64 residual rdot) are emitted
template <class T>
struct OpChebPost2 : OpAmgBase<T> {
  const T *__restrict__ b;
  const T *__restrict__ dinv;
  const double *__restrict__ rdot;
  T *__restrict__ y;
  T c1, c2;
  struct Pre {
    T b, x, y, d;
    double r;
  };
  __device__ Pre pre(int row, int) const {
    return Pre{b[row], this->x[row], y[row], dinv[row], rdot ? rdot[row] : 0.0};
  }
  __device__ void epi(int row, int, T acc, const Pre &pr, DD *d) const {
    const T z = (pr.x + c1 * (pr.x - pr.y)) + c2 * (pr.d * (pr.b - acc));
    y[row] = z;
    dd_fma(d[0], pr.r, (double)z);
  }
};

// coarsest level: Y = C b with the dense (pseudo-)inverse C, one wavefront per
// row, lane-strided products and a fixed shuffle tree
template <class T>
__global__ __launch_bounds__(kBlock) void k_amg_dense(int nc, const T *__restrict__ cinv,
                                                      const T *__restrict__ b,
                                                      T *__restrict__ y,
                                                      const int *__restrict__ mask) {
  if (mask && !mask[0]) return;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (row >= nc) return;
  T s = T(0);
  for (int j = lane; j < nc; j += 64) s += cinv[(size_t)row * nc + j] * b[j];
#pragma unroll
  for (int st = 32; st > 0; st >>= 1) s += __shfl_down(s, st, 64);
  if (lane == 0) y[row] = s;
}

// coarsest level without a dense inverse: one Jacobi step Y = D^-1 b
template <class T>
__global__ __launch_bounds__(kBlock) void k_amg_jacobi(int n, const T *__restrict__ dinv,
                                                       const T *__restrict__ b,
                                                       T *__restrict__ y,
                                                       const int *__restrict__ mask) {
  if (mask && !mask[0]) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    y[i] = dinv[i] * b[i];
}

// pairs of r . y (a one-level hierarchy has no smoothing SpMV to carry them)
template <class T>
__global__ __launch_bounds__(kBlock) void k_amg_dot(int n, const double *__restrict__ r,
                                                    const T *__restrict__ y,
                                                    const int *__restrict__ mask,
                                                    double *__restrict__ partials) {
  if (mask && !mask[0]) return;
  DD acc[1];
  dd_zero(acc[0]);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    dd_fma(acc[0], r[i], (double)y[i]);
  block_partials_binned<1, 1>(acc, partials);
}

template <class Op, class T>
static Op amg_op(const DevCsr &A, const T *val, const T *x, const int *mask) {
  Op op{};
  op.mask = mask;
  op.col = A.col;
  op.val = val;
  op.x = x;
  return op;
}

template <class T>
static void amg_level(hipStream_t s, const AmgLevelView<T> *lv, int nlev, int l, const T *cinv,
                      int nc, bool cjac, const double *rdot, const int *mask, double *partials,
                      int *grid) {
  const AmgLevelView<T> &L = lv[l];
  const bool top = l == 0;
  if (l == nlev - 1) {
    if (cjac)
      hipLaunchKernelGGL((k_amg_jacobi<T>), dim3(vec_grid(L.n)), dim3(kBlock), 0, s, L.n, L.dinv,
                         L.b, L.y, mask);
    else
      hipLaunchKernelGGL((k_amg_dense<T>), dim3(div_up(nc, kBlock / 64)), dim3(kBlock), 0, s, nc,
                         cinv, L.b, L.y, mask);
    CFSA_LAUNCH_CHECK();
    if (top) {
      const int G = vec_grid(L.n);
      *grid = G;
      hipLaunchKernelGGL((k_amg_dot<T>), dim3(G), dim3(kBlock), 0, s, L.n, rdot, L.y, mask,
                         partials);
      CFSA_LAUNCH_CHECK();
    }
    return;
  }
  const AmgLevelView<T> &C = lv[l + 1];
  {
    auto op = amg_op<OpChebPre<T>>(L.A, L.Av, L.x, mask);
    op.b = L.b;
    op.dinv = L.dinv;
    op.y = L.y;
    op.c1 = (T)L.c1;
    op.c2 = (T)L.c2;
    launch_spmv<T, 1, 0>(s, L.A, op, nullptr);
  }
  {
    auto op = amg_op<OpAmgResidual<T>>(L.A, L.Av, L.y, mask);
    op.b = L.b;
    op.r = L.r;
    launch_spmv<T, 1, 0>(s, L.A, op, nullptr);
  }
  {
    const bool cdirect = l + 1 == nlev - 1;
    auto op = amg_op<OpRestrict<T>>(L.R, L.Rv, L.r, mask);
    op.dinvc = C.dinv;
    op.bc = C.b;
    op.xc = cdirect ? nullptr : C.x;
    op.c0 = (T)C.c0;
    launch_spmv<T, 1, 0>(s, L.R, op, nullptr);
  }
  amg_level(s, lv, nlev, l + 1, cinv, nc, cjac, rdot, mask, partials, grid);
  {
    auto op = amg_op<OpProlong<T>>(L.P, L.Pv, C.y, mask);
    op.y = L.y;
    launch_spmv<T, 1, 0>(s, L.P, op, nullptr);
  }
  {
    auto op = amg_op<OpChebPost1<T>>(L.A, L.Av, L.y, mask);
    op.b = L.b;
    op.dinv = L.dinv;
    op.xo = L.x;
    op.c0 = (T)L.c0;
    launch_spmv<T, 1, 0>(s, L.A, op, nullptr);
  }
  {
    auto op = amg_op<OpChebPost2<T>>(L.A, L.Av, L.x, mask);
    op.b = L.b;
    op.dinv = L.dinv;
    op.rdot = top ? rdot : nullptr;
    op.y = L.y;
    op.c1 = (T)L.c1;
    op.c2 = (T)L.c2;
    if (top)
      *grid = launch_spmv<T, 1, 1>(s, L.A, op, partials);
    else
      launch_spmv<T, 1, 0>(s, L.A, op, nullptr);
  }
}

template <class T>
void amg_vcycle(hipStream_t s, const AmgLevelView<T> *lv, int nlev, const T *cinv, int nc,
                bool coarse_jacobi, const double *rdot, const int *mask, double *partials,
                int *grid) {
  if (nlev < 1) throw std::logic_error("AMG V-cycle without levels");
  if (!coarse_jacobi && nc != lv[nlev - 1].n)
    throw std::logic_error("AMG coarse inverse does not match the coarsest level");
  if (!rdot) throw std::logic_error("AMG V-cycle without the residual of its dot");
  amg_level(s, lv, nlev, 0, cinv, nc, coarse_jacobi, rdot, mask, partials, grid);
}
template void amg_vcycle<double>(hipStream_t, const AmgLevelView<double> *, int, const double *,
                                 int, bool, const double *, const int *, double *, int *);
template void amg_vcycle<float>(hipStream_t, const AmgLevelView<float> *, int, const float *,
                                int, bool, const double *, const int *, double *, int *);

// 64-bit fingerprint of an array of doubles: the sum (mod 2^64) of a mixed
// hash of every (index, bit pattern), so it does not depend on the order the
// atomic adds land in; used to tell whether an operator's values changed
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__global__ __launch_bounds__(kBlock) void k_fingerprint(int64_t n, const double *__restrict__ v,
                                                        unsigned long long *out) {
  uint64_t h = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    h += mix64((uint64_t)__double_as_longlong(v[i]) + 0x9e3779b97f4a7c15ull * (uint64_t)i);
  for (int st = 32; st > 0; st >>= 1) h += __shfl_down(h, st, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)h);
}
void fingerprint(hipStream_t s, int64_t n, const double *v, uint64_t *out) {
  CFSA_HIP(hipMemsetAsync(out, 0, sizeof(uint64_t), s));
  hipLaunchKernelGGL(k_fingerprint, dim3(1024), dim3(kBlock), 0, s, n, v,
                     reinterpret_cast<unsigned long long *>(out));
  CFSA_LAUNCH_CHECK();
}

// x = 0, r = b, and for the V-cycle its finest right-hand side b0 = (T) b (not
// written for T = double: the V-cycle reads r) and X0 = c0 D^-1 b; the pairs of
// the stopping norm |b|^2 (or |D^-1 b|^2)
template <class T>
__global__ __launch_bounds__(kBlock) void k_amg_cg_init(int n, const double *__restrict__ b,
                                                        const double *__restrict__ invd,
                                                        const T *__restrict__ dinv0, double c0,
                                                        double *__restrict__ x,
                                                        double *__restrict__ r,
                                                        T *__restrict__ b0, T *__restrict__ x0,
                                                        int scaled,
                                                        double *__restrict__ partials) {
  DD acc[1];
  dd_zero(acc[0]);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double bi = b[i], d = invd[i];
    x[i] = 0.0;
    r[i] = bi;
    const T bt = (T)bi;
    if (b0) b0[i] = bt;
    x0[i] = (T)c0 * (dinv0[i] * bt);
    const double sn = scaled ? d * bi : bi;
    dd_fma(acc[0], sn, sn);
  }
  block_partials_binned<1, 1>(acc, partials);
}

// state of the solve from |b|^2 (pa) and rho = b . z0 (pc); p = z0 is set by
// k_amg_cg_p0
__global__ void k_amg_cg_s_init(KrylovState *st, const double *pa, int Ga, const double *pc,
                                int Gc) {
  __shared__ double bb[1], rz[1];
  reduce_final<1>(pa, Ga, bb);
  reduce_final<1>(pc, Gc, rz);
  if (threadIdx.x != 0) return;
  const double b2 = bb[0];
  st->rhs_sq[0] = b2;
  st->iters[0] = 0;
  st->restarts[0] = 0;
  st->active[0] = 0;
  st->any_active = 0;
  for (int c = 1; c < 3; ++c) {
    st->status[c] = KS_DONE;
    st->active[c] = 0;
  }
  if (b2 == 0.0) {
    st->status[0] = KS_DONE;
    st->err[0] = 0.0;
    return;
  }
  const double thr = fmax(st->tol * st->tol * b2, DBL_MIN);
  st->tol2[0] = thr;
  st->rr[0] = b2;
  if (b2 < thr) {
    st->status[0] = KS_DONE;
    st->err[0] = 1.0;
    return;
  }
  st->rho[0] = rz[0];
  st->rho_s[0][0] = rz[0];
  st->it_s[0][0] = 0;
  st->status[0] = st->maxit > 0 ? KS_RUNNING : KS_DONE;
  st->err[0] = 1.0;
  st->active[0] = st->maxit > 0;
  st->any_active = st->active[0];
}

// p = z0 (in FP64)
template <class T>
__global__ __launch_bounds__(kBlock) void k_amg_cg_p0(int n, const T *__restrict__ z,
                                                      double *__restrict__ p) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    p[i] = (double)z[i];
}

// alpha = rho / p.Ap; r -= alpha q; b0 = (T) r and X0 = c0 D^-1 r for the
// V-cycle; pairs of the stopping norm.  The snapshot of `active` is what
// k_amg_cg_p updates x by.
template <class T>
__global__ __launch_bounds__(kBlock) void k_amg_cg_xr(int n, KrylovState *__restrict__ st,
                                                      const double *__restrict__ pq_partials,
                                                      int Gq, int par,
                                                      const double *__restrict__ q,
                                                      const double *__restrict__ invd,
                                                      const T *__restrict__ dinv0, double c0,
                                                      double *__restrict__ r, T *__restrict__ b0,
                                                      T *__restrict__ x0, int scaled,
                                                      double *__restrict__ partials) {
  if (blockIdx.x == 0 && threadIdx.x == 0) st->act_s[par][0] = st->active[0];
  if (!st->active[0]) return;  // grid-uniform; the pairs are not read then
  double pq[1];
  block_reduce_pairs<1>(pq_partials, Gq, pq);
  const double al = st->rho_s[par][0] / pq[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->r0v[0] = pq[0];
    st->alpha[0] = al;
  }
  DD acc[1];
  dd_zero(acc[0]);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double ri = r[i] - al * q[i];
    const double d = invd[i];
    r[i] = ri;
    const T rt = (T)ri;
    if (b0) b0[i] = rt;
    x0[i] = (T)c0 * (dinv0[i] * rt);
    const double sn = scaled ? d * ri : ri;
    dd_fma(acc[0], sn, sn);
  }
  block_partials_binned<1, 1>(acc, partials);
}

// Every block reduces the stopping pairs (pa) and r . z (pc) and decides as
// Eigen's conjugate_gradient does (block 0 writes the state); then
// x += alpha p and, if the solve continues, p = z + beta p.
template <class T>
__global__ __launch_bounds__(kBlock) void k_amg_cg_p(int n, KrylovState *__restrict__ st,
                                                     const double *__restrict__ pa, int Ga,
                                                     const double *__restrict__ pc, int Gc,
                                                     int par, const T *__restrict__ z,
                                                     double *__restrict__ p,
                                                     double *__restrict__ x) {
  __shared__ double beta_s[1];
  __shared__ int cont_s[1];
  if (!st->act_s[par][0]) return;
  double rn[1], rz[1];
  block_reduce_pairs<1>(pa, Ga, rn);
  block_reduce_pairs<1>(pc, Gc, rz);
  if (threadIdx.x == 0) {
    const double tot[2] = {rn[0], rz[0]};
    cg_decide<1>(st, tot, par, blockIdx.x == 0, beta_s, cont_s);
  }
  __syncthreads();
  const double al = st->alpha[0], be = beta_s[0];
  const bool cont = cont_s[0] != 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double pi = p[i];
    x[i] = x[i] + al * pi;
    if (cont) p[i] = (double)z[i] + be * pi;
  }
}

template <class T>
void amg_cg_init(hipStream_t s, int n, const double *b, const double *invd, const T *dinv0,
                 double c0, double *x, double *r, T *b0, T *x0, int scaled, double *partials,
                 int *grid) {
  const int G = vec_grid(n);
  *grid = G;
  hipLaunchKernelGGL((k_amg_cg_init<T>), dim3(G), dim3(kBlock), 0, s, n, b, invd, dinv0, c0, x,
                     r, b0, x0, scaled, partials);
  CFSA_LAUNCH_CHECK();
}
void amg_cg_s_init(hipStream_t s, KrylovState *st, const double *pa, int Ga, const double *pc,
                   int Gc) {
  hipLaunchKernelGGL(k_amg_cg_s_init, dim3(1), dim3(kBlock), 0, s, st, pa, Ga, pc, Gc);
  CFSA_LAUNCH_CHECK();
}
template <class T>
void amg_cg_p0(hipStream_t s, int n, const T *z, double *p) {
  hipLaunchKernelGGL((k_amg_cg_p0<T>), dim3(vec_grid(n)), dim3(kBlock), 0, s, n, z, p);
  CFSA_LAUNCH_CHECK();
}
template <class T>
void amg_cg_xr(hipStream_t s, int n, KrylovState *st, const double *pq, int Gq, int par,
               const double *q, const double *invd, const T *dinv0, double c0, double *r, T *b0,
               T *x0, int scaled, double *partials, int *grid) {
  const int G = vec_grid(n);
  *grid = G;
  hipLaunchKernelGGL((k_amg_cg_xr<T>), dim3(G), dim3(kBlock), 0, s, n, st, pq, Gq, par, q, invd,
                     dinv0, c0, r, b0, x0, scaled, partials);
  CFSA_LAUNCH_CHECK();
}
template <class T>
void amg_cg_p(hipStream_t s, int n, KrylovState *st, const double *pa, int Ga, const double *pc,
              int Gc, int par, const T *z, double *p, double *x) {
  hipLaunchKernelGGL((k_amg_cg_p<T>), dim3(vec_grid(n)), dim3(kBlock), 0, s, n, st, pa, Ga, pc,
                     Gc, par, z, p, x);
  CFSA_LAUNCH_CHECK();
}
#define CFSA_INST_AMG(T)                                                                       \
  template void amg_cg_init<T>(hipStream_t, int, const double *, const double *, const T *,     \
                               double, double *, double *, T *, T *, int, double *, int *);     \
  template void amg_cg_p0<T>(hipStream_t, int, const T *, double *);                            \
  template void amg_cg_xr<T>(hipStream_t, int, KrylovState *, const double *, int, int,         \
                             const double *, const double *, const T *, double, double *, T *,  \
                             T *, int, double *, int *);                                        \
  template void amg_cg_p<T>(hipStream_t, int, KrylovState *, const double *, int,               \
                            const double *, int, int, const T *, double *, double *);
CFSA_INST_AMG(double)
CFSA_INST_AMG(float)

}  // namespace cfsa
