// Order-robust inner products for the Krylov solvers.
//
// Every inner product and squared norm of the solvers is accumulated as an
// unevaluated pair (hi, lo): each product is split exactly (x*y = h + r with
// r = fma(x, y, -h)) and each addition is error-free (TwoSum), the scheme of
// Ogita, Rump & Oishi, "Accurate Sum and Dot Product", SIAM J. Sci. Comput.
// 26(6), 2005 (Algorithm Dot2).  hi + lo carries the exact sum with an error of
// order n*u^2*sum|x_i y_i|, so its rounding to double is the same whichever
// order the terms were added in, unless the exact value lies within that error
// of a rounding boundary.  The GPU (per-lane, per-block, per-grid tree) and the
// CPU restatement (sequential loop) therefore produce the same scalars, and
// the FP64 Krylov iterates follow the CPU's bit for bit.
//
// Partial results of a reduction are stored as pairs: partials[2*(v*G+g)]
// = hi, partials[2*(v*G+g)+1] = lo for value v of block g out of G.
#pragma once

#include "dev.hpp"

namespace cfsa {

struct DD {
  double hi, lo;
};

__device__ __forceinline__ void dd_zero(DD &a) {
  a.hi = 0.0;
  a.lo = 0.0;
}

// a += x * y
__device__ __forceinline__ void dd_fma(DD &a, double x, double y) {
  const double h = x * y;
  const double r = fma(x, y, -h);
  const double s = a.hi + h;
  const double z = s - a.hi;
  const double q = (a.hi - (s - z)) + (h - z);
  a.hi = s;
  a.lo = a.lo + (q + r);
}

// a += b
__device__ __forceinline__ void dd_merge(DD &a, double bhi, double blo) {
  const double s = a.hi + bhi;
  const double z = s - a.hi;
  const double q = (a.hi - (s - z)) + (bhi - z);
  a.hi = s;
  a.lo = (a.lo + blo) + q;
}

__device__ __forceinline__ double dd_value(const DD &a) { return a.hi + a.lo; }

// Fixed-shape tree over one block of kBlock threads; writes the block's pairs.
template <int NV>
__device__ void block_partials(DD (&v)[NV], double *__restrict__ partials) {
  __shared__ double rh[NV][kBlock], rl[NV][kBlock];
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    rh[i][tid] = v[i].hi;
    rl[i][tid] = v[i].lo;
  }
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (tid < s) {
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        DD a{rh[i][tid], rl[i][tid]};
        dd_merge(a, rh[i][tid + s], rl[i][tid + s]);
        rh[i][tid] = a.hi;
        rl[i][tid] = a.lo;
      }
    }
    __syncthreads();
  }
  if (tid < NV) {
    const size_t o = 2 * ((size_t)tid * gridDim.x + blockIdx.x);
    partials[o] = rh[tid][0];
    partials[o + 1] = rl[tid][0];
  }
}

// Single-block final reduction of NV pair arrays of length G into out[NV]
// (visible to every thread of the block on return).
template <int NV>
__device__ void reduce_final(const double *__restrict__ partials, int G, double *out) {
  __shared__ double rh[NV][kBlock], rl[NV][kBlock];
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    DD a;
    dd_zero(a);
    for (int g = tid; g < G; g += kBlock) {
      const size_t o = 2 * ((size_t)i * G + g);
      dd_merge(a, partials[o], partials[o + 1]);
    }
    rh[i][tid] = a.hi;
    rl[i][tid] = a.lo;
  }
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (tid < s) {
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        DD a{rh[i][tid], rl[i][tid]};
        dd_merge(a, rh[i][tid + s], rl[i][tid + s]);
        rh[i][tid] = a.hi;
        rl[i][tid] = a.lo;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) out[i] = rh[i][0] + rl[i][0];
  }
  __syncthreads();
}

// Every block of a grid reduces the same NV pair arrays of length G in the
// same fixed order (lane-strided merge, wavefront shuffle tree, then the four
// wavefronts in order); out[] is identical in every block and thread.
template <int NV>
__device__ void block_reduce_pairs(const double *__restrict__ partials, int G, double (&out)[NV]) {
  constexpr int W = kBlock / 64;
  __shared__ double wh[NV][W], wl[NV][W];
  __shared__ double res[NV];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    DD a;
    dd_zero(a);
    for (int g = tid; g < G; g += kBlock) {
      const size_t o = 2 * ((size_t)v * G + g);
      dd_merge(a, partials[o], partials[o + 1]);
    }
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) {
      const double oh = __shfl_down(a.hi, st, 64);
      const double ol = __shfl_down(a.lo, st, 64);
      dd_merge(a, oh, ol);
    }
    if (lane == 0) {
      wh[v][wv] = a.hi;
      wl[v][wv] = a.lo;
    }
  }
  __syncthreads();
  if (tid < NV) {
    DD a{wh[tid][0], wl[tid][0]};
    for (int w = 1; w < W; ++w) dd_merge(a, wh[tid][w], wl[tid][w]);
    res[tid] = a.hi + a.lo;
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NV; ++v) out[v] = res[v];
}

// out[0..NV) = reduced values of NV pair arrays (one block of kBlock threads).
template <int NV>
__global__ void k_reduce_pairs(const double *__restrict__ partials, int G, double *out) {
  __shared__ double tot[NV];
  reduce_final<NV>(partials, G, tot);
  if (threadIdx.x < NV) out[threadIdx.x] = tot[threadIdx.x];
}

}  // namespace cfsa
