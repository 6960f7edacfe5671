// Device building blocks shared by the stream-launched and the graph-replayed eigensolver
// kernels, so both paths perform the same operations in the same order.
#pragma once

#include "mpsd_internal.h"

namespace mpsd {

constexpr int kVecThreads = 256;
constexpr int kDotMaxBlocks = 128;

// Partial-sum blocks of a dot product over n elements: a function of n only (one block per
// 2048 elements, 1..128), so a dot's reduction order never depends on anything else.
__host__ __device__ inline int dot_blocks(int64_t n) {
  const int64_t b = (n + 2047) / 2048;
  return b < 1 ? 1 : (b > kDotMaxBlocks ? kDotMaxBlocks : (int)b);
}

// Block `blk` of `nb` partial sums of conj(v) . w; the result is valid in thread 0.
__device__ inline double2 dot_block(const double2* __restrict__ v, const double2* __restrict__ w, int64_t n, int blk,
                                    int nb) {
  const int64_t chunk = (n + nb - 1) / nb;
  const int64_t lo = blk * chunk;
  const int64_t hi = min(n, lo + chunk);
  double2 acc{0.0, 0.0};
  // the thread's elements in ascending order, loads issued 8 at a time ahead of their sums
  for (int64_t t0 = lo + threadIdx.x; t0 < hi; t0 += 8 * kVecThreads) {
    double2 a[8], b[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t t = t0 + (int64_t)k * kVecThreads;
      if (t < hi) {
        a[k] = v[t];
        b[k] = w[t];
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (t0 + (int64_t)k * kVecThreads < hi) {
        acc.x = fma(a[k].x, b[k].x, fma(a[k].y, b[k].y, acc.x));
        acc.y = fma(a[k].x, b[k].y, fma(-a[k].y, b[k].x, acc.y));
      }
  }
  __shared__ double2 red[kVecThreads];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kVecThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[threadIdx.x].x += red[threadIdx.x + s].x;
      red[threadIdx.x].y += red[threadIdx.x + s].y;
    }
    __syncthreads();
  }
  return red[0];
}

// dot_final's result computed by one thread, for nb <= kDotFoldMax: the same halving tree over
// kDotMaxBlocks slots.  Slots at and beyond nb hold zeros, so the levels wider than P (the
// next power of two >= nb) only add +0.0 to each live slot, which is done once (it is
// idempotent), and the levels below P are the tree itself.  Bit for bit dot_final's value.
constexpr int kDotFoldMax = 16;
__device__ inline double2 dot_final_one(const double2* __restrict__ partial, int nb) {
  double2 r[kDotFoldMax];
  int P = 1;
  while (P < nb) P <<= 1;
#pragma unroll
  for (int i = 0; i < kDotFoldMax; ++i) r[i] = i < nb ? partial[i] : double2{0.0, 0.0};
  if (P < kDotMaxBlocks) {
#pragma unroll
    for (int i = 0; i < kDotFoldMax; ++i) {
      r[i].x += 0.0;
      r[i].y += 0.0;
    }
  }
#pragma unroll
  for (int s = kDotFoldMax / 2; s > 0; s >>= 1)
    if (s < P) {
#pragma unroll
      for (int i = 0; i < s; ++i) {
        r[i].x += r[i + s].x;
        r[i].y += r[i + s].y;
      }
    }
  return r[0];
}

// Sum of the nb partials of one dot (launched with kDotMaxBlocks threads); valid in thread 0.
__device__ inline double2 dot_final(const double2* __restrict__ partial, int nb) {
  __shared__ double2 red[kDotMaxBlocks];
  red[threadIdx.x] = threadIdx.x < nb ? partial[threadIdx.x] : double2{0.0, 0.0};
  __syncthreads();
  for (int s = kDotMaxBlocks / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[threadIdx.x].x += red[threadIdx.x + s].x;
      red[threadIdx.x].y += red[threadIdx.x + s].y;
    }
    __syncthreads();
  }
  return red[0];
}

// w[t] += sign * sum_{i<m} V_i[t] c[i] for one element t.
__device__ inline void gemv_elem(const double2* __restrict__ V, int64_t ldv, int m, const double2* __restrict__ c,
                                 double sign, double2* __restrict__ w, int64_t t) {
  double2 acc{0.0, 0.0};
  for (int i = 0; i < m; ++i) {
    const double2 v = V[(int64_t)i * ldv + t];
    const double2 ci = c[i];
    acc.x = fma(v.x, ci.x, fma(-v.y, ci.y, acc.x));
    acc.y = fma(v.x, ci.y, fma(v.y, ci.x, acc.y));
  }
  const double2 wi = w[t];
  w[t] = double2{fma(sign, acc.x, wi.x), fma(sign, acc.y, wi.y)};
}

// Lanczos coefficients of step j from the two Gram-Schmidt passes.
__device__ inline void lanczos_scalars(const double2* __restrict__ h1, const double2* __restrict__ h2, int j,
                                       double& alpha, double& beta, double& inv) {
  alpha = h1[j].x + h2[j].x;
  double corr = 0.0;
  for (int i = 0; i <= j; ++i) corr = fma(h2[i].x, h2[i].x, fma(h2[i].y, h2[i].y, corr));
  beta = sqrt(fmax(h2[j + 1].x - corr, 0.0));
  inv = beta > 1e-300 ? 1.0 / beta : 0.0;
}

}  // namespace mpsd
