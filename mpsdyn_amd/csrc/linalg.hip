// Device vector primitives for the FP64 eigensolver and the gauge steps of the sweep.
//
// Reductions use a fixed grid and a fixed summation order that depend only on the vector
// length, so a result is bit-identical whatever else runs on the device (a solve gives the
// same numbers alone and inside a mix of concurrent solves).
//
// The QR/LQ gauge moves keep the MPS in HBM: rocSOLVER's Householder QR/LQ runs on the
// site buffer in place and the triangular factor is multiplied into the neighbour with the
// BRGEMM kernel.

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mpsd_internal.h"
#include "vec_kernels.h"
#include "wave_xor.h"

namespace mpsd {
namespace {

constexpr int kT = 256;
constexpr int kDotBlocks = 128;  // partial sums per vector (fixed => fixed order)

// partial[j * kDotBlocks + blk] = block blk of nb partial sums of conj(V_j) . w
__global__ void __launch_bounds__(kT) dots_partial_kernel(const double2* __restrict__ V, int64_t ldv,
                                                          const double2* __restrict__ w, int64_t n, int nb,
                                                          double2* __restrict__ partial) {
  const double2 r = dot_block(V + (int64_t)blockIdx.y * ldv, w, n, blockIdx.x, nb);
  if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * kDotBlocks + blockIdx.x] = r;
}

__global__ void __launch_bounds__(kDotBlocks) dots_final_kernel(const double2* __restrict__ partial, int nb,
                                                               double2* __restrict__ out) {
  const double2 r = dot_final(partial + (int64_t)blockIdx.x * kDotBlocks, nb);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// w[i] += sign * sum_j V_j[i] * c[j]   (c on the device, m <= 1024)
__global__ void __launch_bounds__(kT) gemv_update_kernel(const double2* __restrict__ V, int64_t ldv, int m,
                                                         const double2* __restrict__ c, double sign,
                                                         double2* __restrict__ w, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) gemv_elem(V, ldv, m, c, sign, w, i);
}

// y[i] -= sum_j w_j h_j Phi[j*ldp + i]: the projector terms of TensorMultiplierProjMulti with
// h_j = <phi_j|x> already in device memory.
constexpr int kProjMax = 32;
struct ProjWeights {
  double w[kProjMax];
};
__global__ void __launch_bounds__(kT) proj_update_kernel(const double2* __restrict__ Phi, int64_t ldp, int m,
                                                         const double2* __restrict__ h, ProjWeights pw,
                                                         double2* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  double2 acc = y[i];
  for (int j = 0; j < m; ++j) {
    const double2 hj = h[j], p = Phi[(int64_t)j * ldp + i];
    const double2 c{pw.w[j] * hj.x, pw.w[j] * hj.y};
    acc.x -= c.x * p.x - c.y * p.y;
    acc.y -= c.x * p.y + c.y * p.x;
  }
  y[i] = acc;
}

// w = (w - V_{0..j} h2) / beta_j: the second update and lanczos_coef_kernel in one pass, with
// the same operations (gemv_elem's, then the scaling); block 0 stores alpha_j, beta_j.
__global__ void __launch_bounds__(kT) coef_update_kernel(const double2* __restrict__ V, int64_t ldv,
                                                         const double2* __restrict__ h1,
                                                         const double2* __restrict__ h2, int j,
                                                         double* __restrict__ ab, double2* __restrict__ w,
                                                         int64_t n) {
  double alpha, beta, inv;
  lanczos_scalars(h1, h2, j, alpha, beta, inv);
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n; t += (int64_t)gridDim.x * kT) {
    gemv_elem(V, ldv, j + 1, h2, -1.0, w, t);
    const double2 v = w[t];
    w[t] = double2{v.x * inv, v.y * inv};
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ab[2 * j] = alpha;
    ab[2 * j + 1] = beta;
  }
}

// The two Gram-Schmidt passes of a Lanczos step with the dots' last reduction done by the
// consumer (nb <= kDotFoldMax partial sums per dot): every block reduces the partials with
// dot_final_one, which returns dot_final's bits, so the step is the same arithmetic as the
// separate dots_final launches, one launch shorter per pass.  Block 0 stores the dots.
constexpr int kFoldM = 1024;
__global__ void __launch_bounds__(kT) gemv_fold_kernel(const double2* __restrict__ V, int64_t ldv, int m,
                                                       const double2* __restrict__ partial, int nb, double sign,
                                                       double2* __restrict__ w, int64_t n, double2* __restrict__ h) {
  __shared__ double2 c[kFoldM];
  for (int i = threadIdx.x; i < m; i += kT) c[i] = dot_final_one(partial + (int64_t)i * kDotBlocks, nb);
  __syncthreads();
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < m; i += kT) h[i] = c[i];
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t < n) gemv_elem(V, ldv, m, c, sign, w, t);
}

__global__ void __launch_bounds__(kT) coef_fold_kernel(const double2* __restrict__ V, int64_t ldv,
                                                       const double2* __restrict__ h1,
                                                       const double2* __restrict__ partial, int nb, int j,
                                                       double* __restrict__ ab, double2* __restrict__ w,
                                                       int64_t n, double2* __restrict__ h2) {
  __shared__ double2 c[kFoldM];
  for (int i = threadIdx.x; i < j + 2; i += kT) c[i] = dot_final_one(partial + (int64_t)i * kDotBlocks, nb);
  __syncthreads();
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < j + 2; i += kT) h2[i] = c[i];
  double alpha, beta, inv;
  lanczos_scalars(h1, c, j, alpha, beta, inv);
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n; t += (int64_t)gridDim.x * kT) {
    gemv_elem(V, ldv, j + 1, c, -1.0, w, t);
    const double2 v = w[t];
    w[t] = double2{v.x * inv, v.y * inv};
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ab[2 * j] = alpha;
    ab[2 * j + 1] = beta;
  }
}

// ---- Lanczos step with the Gram-Schmidt passes fused ---------------------------------------
// Classical Gram-Schmidt applied twice to w against V_0..V_j (m1 = j + 1 vectors):
//   h1 = V^H w,  w1 = w - V h1,  h2 = V^H w1 and |w1|^2,  w = (w1 - V h2) / beta_j.
// The unfused form reads V four times and writes w twice.  Here w1 is never stored:
//   dots_partial_kernel : block partials of h1 (dot_block's blocks);
//   gs_mid_kernel       : h1 folded from those partials, one wave per coefficient; per element w1 in
//                         registers and its terms of h2 and |w1|^2, reduced per block (nb2
//                         blocks; the One instance, with no loop, when a block has at most one
//                         element per thread);
//   gs_fold_kernel      : h2 from the block partials (one launch, or folded by a wave inside
//                         gs_update_kernel when nb2 <= kGsFoldMax);
//   gs_update_kernel    : w1 again, by the same operations (same bits), then w1 - V h2, scaled
//                         (one element per thread).
// Three passes over V, one store of w.  Every sum has an order fixed by n and j alone, so a
// step's bits do not depend on what else runs on the device.
// The basis columns of one element stay in registers, so the kernels are instantiated for
// M >= m1 columns (8, 16, 24, 32).
constexpr int kGsCols = 32;       // most basis vectors of a fused step
constexpr int kGsMaxBlocks = 512; // partial-sum blocks of the second pass, and their stride
constexpr int kGsFoldMax = 128;  // gs_fold_waves' limit

// one block per 256 elements (one per thread) up to kGsMaxBlocks
inline int gs_blocks(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return b < 1 ? 1 : (b > kGsMaxBlocks ? kGsMaxBlocks : (int)b);
}

// Columns V_i[t] (i < m1) into v and w[t] into wt.
template <int M>
__device__ __forceinline__ void gs_load(const double2* __restrict__ V, int64_t ldv, int m1,
                                        const double2* __restrict__ w, int64_t t, double2 (&v)[M], double2& wt) {
#pragma unroll
  for (int i = 0; i < M; ++i)
    if (i < m1) v[i] = V[(int64_t)i * ldv + t];
  wt = w[t];
}

// w1 = wi - sum_{i<m1} v[i] c[i], gemv_elem's operations on the loaded columns.
template <int M>
__device__ __forceinline__ double2 gs_w1(int m1, const double2* c, double2 wi, const double2 (&v)[M]) {
  double2 acc{0.0, 0.0};
#pragma unroll
  for (int i = 0; i < M; ++i)
    if (i < m1) {
      const double2 ci = c[i];
      acc.x = fma(v[i].x, ci.x, fma(-v[i].y, ci.y, acc.x));
      acc.y = fma(v[i].x, ci.y, fma(v[i].y, ci.x, acc.y));
    }
  return double2{fma(-1.0, acc.x, wi.x), fma(-1.0, acc.y, wi.y)};
}

// Sums of nb <= 128 partials, coefficient i at p + i * ld for i = wave, wave + 4, ... < nc
// (nc <= C): lane b adds p[b] and p[b + 64], then the wave's butterfly; lane 0 of the wave
// passes (i, sum) to put.  All of a wave's loads are issued before the first butterfly.
template <int C, class Put>
__device__ __forceinline__ void gs_fold_waves(const double2* __restrict__ p, int64_t ld, int nb, int nc,
                                              Put put) {
  constexpr int kIt = (C + 3) / 4;
  const int wave = threadIdx.x >> 6, b = (int)__lane_id();
  double2 s[kIt];
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int i = wave + 4 * it;
    s[it] = double2{0.0, 0.0};
    if (i < nc) {
      const double2* q = p + (int64_t)i * ld;
      if (b < nb) s[it] = q[b];
      if (b + 64 < nb) {
        const double2 u = q[b + 64];
        s[it].x += u.x;
        s[it].y += u.y;
      }
    }
  }
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int i = wave + 4 * it;
    if (i < nc) {
      const double2 r{wave_sum_x(s[it].x), wave_sum_x(s[it].y)};
      if (b == 0) put(i, r);
    }
  }
}

// Sum of acc[0..M) over the wave, scattered: returns the total of coefficient gs_lane_coef<M>()
// (lanes l and l ^ 1 hold the same one, for M = 32; for smaller M, more lanes share one).
// Halving steps over the lane bits from 32 down: at offset O a lane keeps the half of its
// values selected by its bit O and adds its partner's copy of that half; the remaining lane
// bits are summed as a butterfly.  A fixed order: the same inputs give the same bits.
template <int M>
__device__ __forceinline__ int gs_lane_coef() {
  const int l = (int)__lane_id();
  constexpr int steps = M == 32 ? 5 : (M == 16 ? 4 : 3);
  int idx = 0;
#pragma unroll
  for (int k = 0; k < steps; ++k) idx = 2 * idx + ((l >> (5 - k)) & 1);
  return idx;
}

template <int K, int O>
__device__ __forceinline__ void gs_halve(double2* a) {
  const bool up = (__lane_id() & O) != 0;
#pragma unroll
  for (int k = 0; k < K / 2; ++k) {
    const double2 send = up ? a[k] : a[k + K / 2];
    const double2 keep = up ? a[k + K / 2] : a[k];
    const double2 recv = xor_lane<O>(send);
    a[k] = double2{keep.x + recv.x, keep.y + recv.y};
  }
}

template <int M>
__device__ __forceinline__ double2 gs_wave_scatter(double2 (&acc)[M]) {
  gs_halve<M, 32>(acc);
  gs_halve<M / 2, 16>(acc);
  gs_halve<M / 4, 8>(acc);
  if constexpr (M >= 16) gs_halve<M / 8, 4>(acc);
  if constexpr (M >= 32) gs_halve<M / 16, 2>(acc);
  double2 r = acc[0];
  if constexpr (M <= 8) r = double2{r.x + xor_lane<4>(r).x, r.y + xor_lane<4>(r).y};
  if constexpr (M <= 16) r = double2{r.x + xor_lane<2>(r).x, r.y + xor_lane<2>(r).y};
  r = double2{r.x + xor_lane<1>(r).x, r.y + xor_lane<1>(r).y};
  return r;
}

template <int M, bool One>
__global__ void __launch_bounds__(kT) gs_mid_kernel(const double2* __restrict__ V, int64_t ldv, int m1,
                                                    const double2* __restrict__ p1, int nb1,
                                                    const double2* __restrict__ w, int64_t n, int nb2,
                                                    double2* __restrict__ h1, double2* __restrict__ p2) {
  __shared__ double2 c[M];
  __shared__ double2 red[kT / 64][M + 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t chunk = (n + nb2 - 1) / nb2;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = min(n, lo + chunk);
  gs_fold_waves<M>(p1, kDotBlocks, nb1, m1, [&](int i, double2 v) {
    c[i] = v;
    if (blockIdx.x == 0) h1[i] = v;
  });
  __syncthreads();
  constexpr int S = M <= 8 ? 8 : (M <= 16 ? 16 : 32);  // accumulators: a power of two for the scatter
  double2 acc[S];
#pragma unroll
  for (int i = 0; i < S; ++i) acc[i] = double2{0.0, 0.0};
  double nrm = 0.0;
  // (the columns are loaded after the fold: loaded before it, they would stay live across the
  // fold's registers, and at 16 to 32 columns the kernel spills and runs slower)
  if constexpr (One) {
    // at most one element per thread (chunk <= kT): straight-line, so each column's register
    // is free once its product is accumulated
    const int64_t t = lo + threadIdx.x;
    if (t < hi) {
      double2 v[M], wt;
      gs_load<M>(V, ldv, m1, w, t, v, wt);
      const double2 w1 = gs_w1<M>(m1, c, wt, v);
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < m1) {  // dot_block's operations, from zero
          acc[i].x = fma(v[i].x, w1.x, fma(v[i].y, w1.y, acc[i].x));
          acc[i].y = fma(v[i].x, w1.y, fma(-v[i].y, w1.x, acc[i].y));
        }
      nrm = fma(w1.x, w1.x, fma(w1.y, w1.y, nrm));
    }
  } else {
  for (int64_t t = lo + threadIdx.x; t < hi; t += kT) {
    // h1 is read from LDS at its use: kept in registers across the loop it would cost 4 M
    // VGPRs on top of the columns and the accumulators
    __asm__ volatile("" ::: "memory");
    double2 v[M], wt;
    gs_load<M>(V, ldv, m1, w, t, v, wt);
    const double2 w1 = gs_w1<M>(m1, c, wt, v);
#pragma unroll
    for (int i = 0; i < M; ++i)
      if (i < m1) {  // dot_block's operations
        acc[i].x = fma(v[i].x, w1.x, fma(v[i].y, w1.y, acc[i].x));
        acc[i].y = fma(v[i].x, w1.y, fma(-v[i].y, w1.x, acc[i].y));
      }
    nrm = fma(w1.x, w1.x, fma(w1.y, w1.y, nrm));
  }
  }
  const double2 r = gs_wave_scatter<S>(acc);
  constexpr int share = 64 / S;  // lanes holding one coefficient
  const int coef = gs_lane_coef<S>();
  if (lane % share == 0 && coef < M) red[wave][coef] = r;
  nrm = wave_sum_x(nrm);
  __syncthreads();  // red[.][m1] is written after every coefficient slot has been
  if (lane == 0) red[wave][m1] = double2{nrm, 0.0};
  __syncthreads();
  if ((int)threadIdx.x <= m1) {
    double2 s = red[0][threadIdx.x];
    for (int k = 1; k < kT / 64; ++k) {
      s.x += red[k][threadIdx.x].x;
      s.y += red[k][threadIdx.x].y;
    }
    p2[(int64_t)threadIdx.x * kGsMaxBlocks + blockIdx.x] = s;
  }
}

// h2 from the nb2 block partials of gs_mid_kernel (slot i at p2 + i * kGsMaxBlocks), i <= m1
// (slot m1 is |w1|^2): p[t] + p[t + 256] per thread, each wave's butterfly, then the four
// waves in order.  One block per slot; used when nb2 > kGsFoldMax.
__global__ void __launch_bounds__(kT) gs_fold_kernel(const double2* __restrict__ p2, int nb2,
                                                     double2* __restrict__ h2) {
  __shared__ double2 red[kT / 64];
  const double2* p = p2 + (int64_t)blockIdx.x * kGsMaxBlocks;
  const int t = threadIdx.x;
  double2 s = t < nb2 ? p[t] : double2{0.0, 0.0};
  if (t + kT < nb2) {
    s.x += p[t + kT].x;
    s.y += p[t + kT].y;
  }
  s.x = wave_sum_x(s.x);
  s.y = wave_sum_x(s.y);
  if ((t & 63) == 0) red[t >> 6] = s;
  __syncthreads();
  if (t == 0) {
    double2 r = red[0];
    for (int k = 1; k < kT / 64; ++k) {
      r.x += red[k].x;
      r.y += red[k].y;
    }
    h2[blockIdx.x] = r;
  }
}

// One element per thread (grid ceil(n / kT)), so the coefficients are read from LDS once.
template <int M>
__global__ void __launch_bounds__(kT) gs_update_kernel(const double2* __restrict__ V, int64_t ldv, int m1,
                                                       const double2* __restrict__ h1,
                                                       const double2* __restrict__ p2, int nb2, int fold,
                                                       double2* __restrict__ h2, double* __restrict__ ab,
                                                       double2* __restrict__ w, int64_t n) {
  __shared__ double2 c1[M];
  __shared__ double2 c2[M + 1];
  // the element's columns and w first: their latency overlaps the coefficient fold's
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  double2 v[M], wt{0.0, 0.0};
  if (t < n) gs_load<M>(V, ldv, m1, w, t, v, wt);
  if ((int)threadIdx.x < m1) c1[threadIdx.x] = h1[threadIdx.x];
  if (fold) {
    gs_fold_waves<M + 1>(p2, kGsMaxBlocks, nb2, m1 + 1, [&](int i, double2 v) {
      c2[i] = v;
      if (blockIdx.x == 0) h2[i] = v;
    });
  } else if ((int)threadIdx.x <= m1) {
    c2[threadIdx.x] = h2[threadIdx.x];
  }
  __syncthreads();
  const int j = m1 - 1;
  if (t < n) {
    const double2 w1 = gs_w1<M>(m1, c1, wt, v);
    double2 acc{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < M; ++i)
      if (i < m1) {
        const double2 ci = c2[i];
        acc.x = fma(v[i].x, ci.x, fma(-v[i].y, ci.y, acc.x));
        acc.y = fma(v[i].x, ci.y, fma(v[i].y, ci.x, acc.y));
      }
    const double2 w2{fma(-1.0, acc.x, w1.x), fma(-1.0, acc.y, w1.y)};
    double alpha, beta, inv;
    lanczos_scalars(c1, c2, j, alpha, beta, inv);
    w[t] = double2{w2.x * inv, w2.y * inv};
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double alpha, beta, inv;
    lanczos_scalars(c1, c2, j, alpha, beta, inv);
    ab[2 * j] = alpha;
    ab[2 * j + 1] = beta;
  }
}

// Lanczos coefficients of step j from the two Gram-Schmidt passes (h1 = V^H w before the
// first pass, h2 = [V_0..V_j, w]^H w after it): alpha_j = Re(h1[j] + h2[j]),
// beta_j = sqrt(|w|^2 - sum_i |h2[i]|^2) (the norm after the second pass), then
// w /= beta_j (w = 0 when beta_j underflows: the Krylov space is exhausted).  Every block
// recomputes the scalars in the same order; block 0 stores them.
__global__ void __launch_bounds__(kT) lanczos_coef_kernel(const double2* __restrict__ h1,
                                                          const double2* __restrict__ h2, int j,
                                                          double* __restrict__ ab, double2* __restrict__ w,
                                                          int64_t n) {
  double alpha, beta, inv;
  lanczos_scalars(h1, h2, j, alpha, beta, inv);
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double2 v = w[i];
    w[i] = double2{v.x * inv, v.y * inv};
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ab[2 * j] = alpha;
    ab[2 * j + 1] = beta;
  }
}


// y = a x + b y
__global__ void __launch_bounds__(kT) axpby_kernel(double2 a, const double2* __restrict__ x, double2 b,
                                                   double2* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double2 xv = x[i], yv = y[i];
  y[i] = double2{a.x * xv.x - a.y * xv.y + b.x * yv.x - b.y * yv.y,
                 a.x * xv.y + a.y * xv.x + b.x * yv.y + b.y * yv.x};
}

__global__ void __launch_bounds__(kT) fill_kernel(double2* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) x[i] = double2{0.0, 0.0};
}

// y = x / sqrt(nrm2->x) (y = x when the norm is zero or not finite: the caller checks it)
__global__ void __launch_bounds__(kT) scale_invsqrt_kernel(const double2* __restrict__ x, double2* __restrict__ y,
                                                           int64_t n, const double2* __restrict__ nrm2) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double q = nrm2->x;
  const double inv = (q > 0.0 && q < 1e300) ? 1.0 / sqrt(q) : 1.0;
  const double2 v = x[i];
  y[i] = double2{v.x * inv, v.y * inv};
}

// Upper-trapezoidal R (k x n) out of a geqrf result (lda = m), zeros below the diagonal.
__global__ void __launch_bounds__(kT) extract_r_kernel(const double2* __restrict__ A, int m, int k, int n,
                                                       double2* __restrict__ R) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= (int64_t)k * n) return;
  const int i = (int)(t % k), j = (int)(t / k);
  R[t] = (i <= j) ? A[i + (int64_t)m * j] : double2{0.0, 0.0};
}

// Lower-trapezoidal L (m x k) out of a gelqf result (lda = m), zeros above the diagonal.
__global__ void __launch_bounds__(kT) extract_l_kernel(const double2* __restrict__ A, int m, int k,
                                                       double2* __restrict__ L) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= (int64_t)m * k) return;
  const int i = (int)(t % m), j = (int)(t / m);
  L[t] = (i >= j) ? A[i + (int64_t)m * j] : double2{0.0, 0.0};
}

// out[a, s, c] = in[s, a, c] for in = [d, Dl, Dr]; ldo = leading dim (rows) of out.
__global__ void __launch_bounds__(kT) permute_sad_kernel(const double2* __restrict__ in, int d, int Dl, int Dr,
                                                         double2* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t total = (int64_t)d * Dl * Dr;
  if (t >= total) return;
  const int s = (int)(t % d);
  const int a = (int)((t / d) % Dl);
  const int64_t c = t / ((int64_t)d * Dl);
  out[a + (int64_t)Dl * s + (int64_t)Dl * d * c] = in[t];
}

// site[s, i, c] = Q[i, s + d*c] with Q stored with leading dimension ldq (k rows used).
__global__ void __launch_bounds__(kT) unpermute_q_kernel(const double2* __restrict__ Q, int ldq, int k, int d,
                                                         int Dr, double2* __restrict__ site) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t total = (int64_t)d * k * Dr;
  if (t >= total) return;
  const int s = (int)(t % d);
  const int i = (int)((t / d) % k);
  const int64_t c = t / ((int64_t)d * k);
  site[t] = Q[i + (int64_t)ldq * (s + (int64_t)d * c)];
}

// R[i + X j] = S[i] V[i + ldv j] for i < X, j < n (the S V^H factor of a truncated SVD).
__global__ void __launch_bounds__(kT) scale_rows_kernel(const double2* __restrict__ V, int ldv, const double* __restrict__ S,
                                                        int X, int n, double2* __restrict__ R) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= (int64_t)X * n) return;
  const int i = (int)(t % X);
  const int64_t j = t / X;
  const double2 v = V[i + (int64_t)ldv * j];
  R[t] = double2{S[i] * v.x, S[i] * v.y};
}

// L[a + m i] = U[a + m i] S[i] for i < X (the U S factor of a truncated SVD).
__global__ void __launch_bounds__(kT) scale_cols_kernel(const double2* __restrict__ U, int m, const double* __restrict__ S,
                                                        int X, double2* __restrict__ L) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= (int64_t)m * X) return;
  const int i = (int)(t / m);
  const double2 u = U[t];
  L[t] = double2{S[i] * u.x, S[i] * u.y};
}

// B[s + d c, a] = conj(site[s, a, c]): the adjoint of the LQ matrix M[a, s + d c] (B: n x Dl).
__global__ void __launch_bounds__(kT) site_to_bh_kernel(const double2* __restrict__ site, int d, int Dl, int Dr,
                                                        double2* __restrict__ B) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t n = (int64_t)d * Dr;
  if (t >= n * Dl) return;
  const int s = (int)(t % d);
  const int a = (int)((t / d) % Dl);
  const int64_t c = t / ((int64_t)d * Dl);
  const double2 v = site[t];
  B[(s + (int64_t)d * c) + n * a] = double2{v.x, -v.y};
}

// out[i, j] = conj(in[j, i]) for in (r x c), out (c x r)
__global__ void __launch_bounds__(kT) conj_transpose_kernel(const double2* __restrict__ in, int r, int c,
                                                            double2* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= (int64_t)r * c) return;
  const int i = (int)(t % c), j = (int)(t / c);
  const double2 v = in[j + (int64_t)r * i];
  out[t] = double2{v.x, -v.y};
}

// site[s, i, c] = conj(Q'[s + d c, i]) for Q' (n x k), n = d * Dr
__global__ void __launch_bounds__(kT) qh_to_site_kernel(const double2* __restrict__ Qp, int n, int k, int d, int Dr,
                                                        double2* __restrict__ site) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= (int64_t)d * k * Dr) return;
  const int s = (int)(t % d);
  const int i = (int)((t / d) % k);
  const int64_t c = t / ((int64_t)d * k);
  const double2 v = Qp[(s + (int64_t)d * c) + (int64_t)n * i];
  site[t] = double2{v.x, -v.y};
}

// Edge phase of Site::gaugeR/gaugeL (Site.cpp:202-207, 232-237): the 1x1 factor f becomes
// |f| and its phase f/|f| is multiplied into the site tensor.
__global__ void __launch_bounds__(kT) edge_phase_kernel(double2* __restrict__ site, int64_t n,
                                                        double2* __restrict__ f) {
  const double2 v = *f;
  const double a = sqrt(v.x * v.x + v.y * v.y);
  const double2 ph = a > 0.0 ? double2{v.x / a, v.y / a} : double2{1.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double2 x = site[i];
    site[i] = double2{x.x * ph.x - x.y * ph.y, x.x * ph.y + x.y * ph.x};
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) *f = double2{a, 0.0};
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

// Per-context rocBLAS handle bound to the context's stream (created on first use).
static int handle_of(mpsd_ctx* ctx, rocblas_handle* h) {
  if (!ctx->blas) {
    rocblas_handle b = nullptr;
    if (rocblas_create_handle(&b) != rocblas_status_success)
      return set_err(ctx, MPSD_ERR_HIP, "rocblas_create_handle failed");
    if (rocblas_set_stream(b, ctx->stream) != rocblas_status_success) {
      rocblas_destroy_handle(b);
      return set_err(ctx, MPSD_ERR_HIP, "rocblas_set_stream failed");
    }
    ctx->blas = b;
  }
  *h = (rocblas_handle)ctx->blas;
  return MPSD_OK;
}

int blas_handle(mpsd_ctx* ctx, void** out) {
  rocblas_handle h = nullptr;
  const int rc = handle_of(ctx, &h);
  *out = h;
  return rc;
}

void destroy_blas(mpsd_ctx* ctx) {
  if (ctx->blas) rocblas_destroy_handle((rocblas_handle)ctx->blas);
  ctx->blas = nullptr;
}

static int solver_err(mpsd_ctx* ctx, rocblas_status st, const char* what) {
  if (st == rocblas_status_success) return MPSD_OK;
  return set_err(ctx, MPSD_ERR_HIP, std::string(what) + ": " + rocblas_status_to_string(st));
}

}  // namespace mpsd

namespace mpsd {
static int ensure_dots(mpsd_ctx* ctx, int m) {
  const size_t need = sizeof(double2) * ((size_t)m * kDotBlocks + m);
  if (ctx->dots_bytes >= need) return MPSD_OK;
  if (ctx->dots) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->dots);
  }
  ctx->dots = nullptr;
  ctx->dots_bytes = 0;
  const size_t want = std::max(need, sizeof(double2) * (64 * kDotBlocks + 64));
  hipError_t e = hipMalloc(&ctx->dots, want);
  if (e != hipSuccess) return hip_err(ctx, e, "dots alloc");
  ctx->dots_bytes = want;
  return MPSD_OK;
}

}  // namespace mpsd

using namespace mpsd;

extern "C" {

int mpsd_vec_dots(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, const double* w, int64_t n,
                  double* out_host) {
  if (!ctx || !V || !w || !out_host || m < 0 || n < 0 || (m > 1 && ldv < n)) return MPSD_ERR_ARG;
  if (m == 0) return MPSD_OK;
  int rc = ensure_dots(ctx, m);
  if (rc) return rc;
  double2* partial = (double2*)ctx->dots;
  double2* out = partial + (size_t)m * kDotBlocks;
  const int nb = dot_blocks(n);
  hipLaunchKernelGGL(dots_partial_kernel, dim3(nb, m), dim3(kT), 0, ctx->stream, (const double2*)V, ldv,
                     (const double2*)w, n, nb, partial);
  hipLaunchKernelGGL(dots_final_kernel, dim3(m), dim3(kDotBlocks), 0, ctx->stream, partial, nb, out);
  hipError_t e = hipMemcpyAsync(out_host, out, sizeof(double2) * m, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return hip_err(ctx, e, "vec_dots");
}

int mpsd_vec_gemv(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, const double* c_host, double* w,
                  int64_t n) {
  if (!ctx || !V || !w || !c_host || m < 0 || m > 1024 || n < 0) return MPSD_ERR_ARG;
  if (m == 0 || n == 0) return MPSD_OK;
  // Coefficients go through a pinned host ring and a device ring of 8 slots; a slot is
  // reused only after the event of its previous copy has completed.
  if (!ctx->coef) {
    hipError_t e = hipMalloc(&ctx->coef, sizeof(double2) * 1024 * 8);
    if (e == hipSuccess) e = hipHostMalloc(&ctx->coef_host, sizeof(double2) * 1024 * 8, hipHostMallocDefault);
    for (int i = 0; e == hipSuccess && i < 8; ++i) e = hipEventCreateWithFlags(&ctx->coef_ev[i], hipEventDisableTiming);
    if (e != hipSuccess) return hip_err(ctx, e, "coef alloc");
  }
  const unsigned slot = ctx->coef_slot++ & 7;
  hipError_t e = hipEventSynchronize(ctx->coef_ev[slot]);
  if (e != hipSuccess) return hip_err(ctx, e, "coef slot wait");
  double2* ch = (double2*)ctx->coef_host + 1024 * slot;
  double2* c = (double2*)ctx->coef + 1024 * slot;
  memcpy(ch, c_host, sizeof(double2) * m);
  e = hipMemcpyAsync(c, ch, sizeof(double2) * m, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return hip_err(ctx, e, "vec_gemv H2D");
  hipLaunchKernelGGL(gemv_update_kernel, dim3(nblk(n)), dim3(kT), 0, ctx->stream, (const double2*)V, ldv, m,
                     c, 1.0, (double2*)w, n);
  e = hipEventRecord(ctx->coef_ev[slot], ctx->stream);
  return hip_err(ctx, e, "vec_gemv");
}

int mpsd_vec_dots_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, const double* w, int64_t n,
                      double* out_dev) {
  if (!ctx || !V || !w || !out_dev || m < 0 || n < 0 || (m > 1 && ldv < n)) return MPSD_ERR_ARG;
  if (m == 0) return MPSD_OK;
  int rc = ensure_dots(ctx, m);
  if (rc) return rc;
  double2* partial = (double2*)ctx->dots;
  const int nb = dot_blocks(n);
  hipLaunchKernelGGL(dots_partial_kernel, dim3(nb, m), dim3(kT), 0, ctx->stream, (const double2*)V, ldv,
                     (const double2*)w, n, nb, partial);
  hipLaunchKernelGGL(dots_final_kernel, dim3(m), dim3(kDotBlocks), 0, ctx->stream, partial, nb,
                     (double2*)out_dev);
  return hip_err(ctx, hipGetLastError(), "vec_dots_dev");
}

int mpsd_vec_gemv_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, const double* c_dev, double sign,
                      double* w, int64_t n) {
  if (!ctx || !V || !w || !c_dev || m < 0 || n < 0) return MPSD_ERR_ARG;
  if (m == 0 || n == 0) return MPSD_OK;
  hipLaunchKernelGGL(gemv_update_kernel, dim3(nblk(n)), dim3(kT), 0, ctx->stream, (const double2*)V, ldv, m,
                     (const double2*)c_dev, sign, (double2*)w, n);
  return hip_err(ctx, hipGetLastError(), "vec_gemv_dev");
}

int mpsd_vec_proj_update(mpsd_ctx* ctx, const double* Phi, int64_t ldphi, int m, const double* weights,
                         const double* x, double* y, int64_t n) {
  if (!ctx || (m > 0 && (!Phi || !weights)) || !x || !y || m < 0 || n < 0 || (m > 1 && ldphi < n))
    return MPSD_ERR_ARG;
  if (m > kProjMax) return set_err(ctx, MPSD_ERR_ARG, "vec_proj_update: at most 32 projected vectors");
  if (m == 0 || n == 0) return MPSD_OK;
  if (x == y) return set_err(ctx, MPSD_ERR_ARG, "vec_proj_update: x and y must differ");
  int rc = ensure_dots(ctx, m);
  if (rc) return rc;
  double2* partial = (double2*)ctx->dots;
  double2* h = partial + (size_t)m * kDotBlocks;
  const int nb = dot_blocks(n);
  hipLaunchKernelGGL(dots_partial_kernel, dim3(nb, m), dim3(kT), 0, ctx->stream, (const double2*)Phi, ldphi,
                     (const double2*)x, n, nb, partial);
  hipLaunchKernelGGL(dots_final_kernel, dim3(m), dim3(kDotBlocks), 0, ctx->stream, partial, nb, h);
  ProjWeights pw{};
  for (int j = 0; j < m; ++j) pw.w[j] = weights[j];
  hipLaunchKernelGGL(proj_update_kernel, dim3(nblk(n)), dim3(kT), 0, ctx->stream, (const double2*)Phi, ldphi, m,
                     (const double2*)h, pw, (double2*)y, n);
  return hip_err(ctx, hipGetLastError(), "vec_proj_update");
}

int mpsd_lanczos_coef(mpsd_ctx* ctx, const double* h1_dev, const double* h2_dev, int j, double* ab_dev, double* w,
                      int64_t n) {
  if (!ctx || !h1_dev || !h2_dev || !ab_dev || !w || j < 0 || n < 0) return MPSD_ERR_ARG;
  const unsigned blocks = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (n + kT - 1) / kT), 1024);
  hipLaunchKernelGGL(lanczos_coef_kernel, dim3(blocks), dim3(kT), 0, ctx->stream, (const double2*)h1_dev,
                     (const double2*)h2_dev, j, ab_dev, (double2*)w, n);
  return hip_err(ctx, hipGetLastError(), "lanczos_coef");
}

int mpsd_lanczos_coef_update(mpsd_ctx* ctx, const double* V, int64_t ldv, const double* h1_dev,
                             const double* h2_dev, int j, double* ab_dev, double* w, int64_t n) {
  if (!ctx || !V || !h1_dev || !h2_dev || !ab_dev || !w || j < 0 || n < 0 || (j > 0 && ldv < n))
    return MPSD_ERR_ARG;
  const unsigned blocks = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (n + kT - 1) / kT), 1024);
  hipLaunchKernelGGL(coef_update_kernel, dim3(blocks), dim3(kT), 0, ctx->stream, (const double2*)V, ldv,
                     (const double2*)h1_dev, (const double2*)h2_dev, j, ab_dev, (double2*)w, n);
  return hip_err(ctx, hipGetLastError(), "lanczos_coef_update");
}

int mpsd_lanczos_gs1_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, double* w, int64_t n, double* h_dev) {
  if (!ctx || !V || !w || !h_dev || m < 1 || n < 0 || (m > 1 && ldv < n)) return MPSD_ERR_ARG;
  const int nb = dot_blocks(n);
  if (nb > kDotFoldMax || m > kFoldM || n == 0) {
    int rc = mpsd_vec_dots_dev(ctx, V, ldv, m, w, n, h_dev);
    return rc ? rc : mpsd_vec_gemv_dev(ctx, V, ldv, m, h_dev, -1.0, w, n);
  }
  int rc = ensure_dots(ctx, m);
  if (rc) return rc;
  double2* partial = (double2*)ctx->dots;
  hipLaunchKernelGGL(dots_partial_kernel, dim3(nb, m), dim3(kT), 0, ctx->stream, (const double2*)V, ldv,
                     (const double2*)w, n, nb, partial);
  hipLaunchKernelGGL(gemv_fold_kernel, dim3(nblk(n)), dim3(kT), 0, ctx->stream, (const double2*)V, ldv, m,
                     (const double2*)partial, nb, -1.0, (double2*)w, n, (double2*)h_dev);
  return hip_err(ctx, hipGetLastError(), "lanczos_gs1");
}

int mpsd_lanczos_gs2_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int j, const double* h1_dev, double* h2_dev,
                         double* ab_dev, double* w, int64_t n) {
  if (!ctx || !V || !w || !h1_dev || !h2_dev || !ab_dev || j < 0 || n < 0 || ldv < n) return MPSD_ERR_ARG;
  const int nb = dot_blocks(n);
  if (nb > kDotFoldMax || j + 2 > kFoldM || n == 0) {
    int rc = mpsd_vec_dots_dev(ctx, V, ldv, j + 2, w, n, h2_dev);
    return rc ? rc : mpsd_lanczos_coef_update(ctx, V, ldv, h1_dev, h2_dev, j, ab_dev, w, n);
  }
  int rc = ensure_dots(ctx, j + 2);
  if (rc) return rc;
  double2* partial = (double2*)ctx->dots;
  hipLaunchKernelGGL(dots_partial_kernel, dim3(nb, j + 2), dim3(kT), 0, ctx->stream, (const double2*)V, ldv,
                     (const double2*)w, n, nb, partial);
  const unsigned blocks = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (n + kT - 1) / kT), 1024);
  hipLaunchKernelGGL(coef_fold_kernel, dim3(blocks), dim3(kT), 0, ctx->stream, (const double2*)V, ldv,
                     (const double2*)h1_dev, (const double2*)partial, nb, j, ab_dev, (double2*)w, n,
                     (double2*)h2_dev);
  return hip_err(ctx, hipGetLastError(), "lanczos_gs2");
}

int mpsd_lanczos_gs_step_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int j, double* w, int64_t n,
                             double* h1_dev, double* h2_dev, double* ab_dev) {
  if (!ctx || !V || !w || !h1_dev || !h2_dev || !ab_dev || j < 0 || n <= 0 || (j > 0 && ldv < n))
    return MPSD_ERR_ARG;
  const int m1 = j + 1;
  if (m1 > kGsCols) return set_err(ctx, MPSD_ERR_UNSUPPORTED, "lanczos_gs_step: at most 32 basis vectors");
  int rc = ensure_dots(ctx, m1);
  if (rc) return rc;
  if (!ctx->gs_partial) {
    const size_t bytes = sizeof(double2) * (size_t)(kGsCols + 1) * kGsMaxBlocks;
    hipError_t e = hipMalloc(&ctx->gs_partial, bytes);
    if (e != hipSuccess) {
      ctx->gs_partial = nullptr;
      return hip_err(ctx, e, "gs partial alloc");
    }
    ctx->gs_partial_bytes = bytes;
  }
  hipStream_t s = ctx->stream;
  const double2* Vc = (const double2*)V;
  double2* p1 = (double2*)ctx->dots;
  double2* p2 = (double2*)ctx->gs_partial;
  double2* h1 = (double2*)h1_dev;
  double2* h2 = (double2*)h2_dev;
  const int nb1 = dot_blocks(n), nb2 = gs_blocks(n);
  const int fold = nb2 <= kGsFoldMax ? 1 : 0;
  hipLaunchKernelGGL(dots_partial_kernel, dim3(nb1, m1), dim3(kT), 0, s, Vc, ldv, (const double2*)w, n, nb1, p1);
  const bool one = (n + nb2 - 1) / nb2 <= kT;  // gs_mid_kernel's chunk fits one pass
  auto steps = [&](auto mid1, auto mid, auto upd) {
    hipLaunchKernelGGL(one ? mid1 : mid, dim3(nb2), dim3(kT), 0, s, Vc, ldv, m1, (const double2*)p1, nb1, (const double2*)w, n,
                       nb2, h1, p2);
    if (!fold) hipLaunchKernelGGL(gs_fold_kernel, dim3(m1 + 1), dim3(kT), 0, s, (const double2*)p2, nb2, h2);
    hipLaunchKernelGGL(upd, dim3(nblk(n)), dim3(kT), 0, s, Vc, ldv, m1, (const double2*)h1, (const double2*)p2,
                       nb2, fold, h2, ab_dev, (double2*)w, n);
  };
  if (m1 <= 8)
    steps(gs_mid_kernel<8, true>, gs_mid_kernel<8, false>, gs_update_kernel<8>);
  else if (m1 <= 16)
    steps(gs_mid_kernel<16, true>, gs_mid_kernel<16, false>, gs_update_kernel<16>);
  else if (m1 <= 24)
    steps(gs_mid_kernel<24, true>, gs_mid_kernel<24, false>, gs_update_kernel<24>);
  else
    steps(gs_mid_kernel<32, true>, gs_mid_kernel<32, false>, gs_update_kernel<32>);
  return hip_err(ctx, hipGetLastError(), "lanczos_gs_step");
}

int mpsd_vec_normalize_dev(mpsd_ctx* ctx, const double* x, double* y, int64_t n, double* nrm2_dev) {
  if (!ctx || !x || !y || !nrm2_dev || n <= 0) return MPSD_ERR_ARG;
  int rc = mpsd_vec_dots_dev(ctx, x, n, 1, x, n, nrm2_dev);  // nrm2_dev[0..1] = (|x|^2, 0)
  if (rc) return rc;
  hipLaunchKernelGGL(scale_invsqrt_kernel, dim3(nblk(n)), dim3(kT), 0, ctx->stream, (const double2*)x, (double2*)y,
                     n, (const double2*)nrm2_dev);
  return hip_err(ctx, hipGetLastError(), "vec_normalize_dev");
}

int mpsd_vec_axpby(mpsd_ctx* ctx, const double* a, const double* x, const double* b, double* y, int64_t n) {
  if (!ctx || !a || !x || !b || !y || n < 0) return MPSD_ERR_ARG;
  if (n == 0) return MPSD_OK;
  hipLaunchKernelGGL(axpby_kernel, dim3(nblk(n)), dim3(kT), 0, ctx->stream, double2{a[0], a[1]},
                     (const double2*)x, double2{b[0], b[1]}, (double2*)y, n);
  return hip_err(ctx, hipGetLastError(), "vec_axpby");
}

int mpsd_fetch_async(mpsd_ctx* ctx, const void* dev, int64_t bytes, int slot) {
  if (!ctx || !dev || bytes < 0 || bytes > mpsd_ctx::kFetchBytes || slot < 0 || slot > 1) return MPSD_ERR_ARG;
  if (!ctx->fetch_host) {
    hipError_t e = hipHostMalloc(&ctx->fetch_host, 2 * mpsd_ctx::kFetchBytes, hipHostMallocDefault);
    for (int i = 0; e == hipSuccess && i < 2; ++i) e = hipEventCreateWithFlags(&ctx->fetch_ev[i], hipEventDisableTiming);
    if (e != hipSuccess) return hip_err(ctx, e, "fetch alloc");
  }
  // the slot's previous copy has been waited for by its reader (mpsd_fetch_wait)
  char* h = (char*)ctx->fetch_host + slot * mpsd_ctx::kFetchBytes;
  hipError_t e = bytes ? hipMemcpyAsync(h, dev, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream) : hipSuccess;
  if (e == hipSuccess) e = hipEventRecord(ctx->fetch_ev[slot], ctx->stream);
  return hip_err(ctx, e, "fetch_async");
}

int mpsd_fetch_wait(mpsd_ctx* ctx, int slot, void* host, int64_t bytes) {
  if (!ctx || !host || bytes < 0 || bytes > mpsd_ctx::kFetchBytes || slot < 0 || slot > 1 || !ctx->fetch_host)
    return MPSD_ERR_ARG;
  hipError_t e = hipEventSynchronize(ctx->fetch_ev[slot]);
  if (e != hipSuccess) return hip_err(ctx, e, "fetch_wait");
  memcpy(host, (const char*)ctx->fetch_host + slot * mpsd_ctx::kFetchBytes, (size_t)bytes);
  return MPSD_OK;
}

int mpsd_copy_to_host(mpsd_ctx* ctx, const void* dev, void* host, int64_t bytes) {
  if (!ctx || !dev || !host || bytes < 0) return MPSD_ERR_ARG;
  if (bytes == 0) return MPSD_OK;
  hipError_t e = hipMemcpyAsync(host, dev, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return hip_err(ctx, e, "copy_to_host");
}

int mpsd_vec_copy(mpsd_ctx* ctx, const double* x, double* y, int64_t n) {
  if (!ctx || !x || !y || n < 0) return MPSD_ERR_ARG;
  if (n == 0 || x == y) return MPSD_OK;
  return hip_err(ctx, hipMemcpyAsync(y, x, sizeof(double2) * n, hipMemcpyDeviceToDevice, ctx->stream), "vec_copy");
}

int mpsd_vec_zero(mpsd_ctx* ctx, double* x, int64_t n) {
  if (!ctx || !x || n < 0) return MPSD_ERR_ARG;
  if (n == 0) return MPSD_OK;
  hipLaunchKernelGGL(fill_kernel, dim3(nblk(n)), dim3(kT), 0, ctx->stream, (double2*)x, n);
  return hip_err(ctx, hipGetLastError(), "vec_zero");
}

int mpsd_vec_norm2(mpsd_ctx* ctx, const double* x, int64_t n, double* out) {
  if (!out) return MPSD_ERR_ARG;
  double h[2] = {0.0, 0.0};
  int rc = mpsd_vec_dots(ctx, x, n, 1, x, n, h);
  *out = h[0];
  return rc;
}

int mpsd_site_gauge_r(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* R_out) {
  if (!ctx || !site || !R_out || site == R_out) return MPSD_ERR_ARG;
  const int d = site->dims[0], Dl = site->rank > 1 ? site->dims[1] : 1, Dr = site->rank > 2 ? site->dims[2] : 1;
  const int m = d * Dl, n = Dr, k = m < n ? m : n;
  if (ctx->gauge_qr == 0 && m <= house_qr_max_rows()) {
    // library Householder QR: site (m x n) = Q R, Q (m x k) back into the site (written there
    // directly by the single-workgroup form)
    const bool small = house_qr_small(m, n);
    int rc = small ? 0 : ensure_scratch(ctx, 3, sizeof(double2) * (size_t)m * k);
    const int rd[2] = {k, n};
    if (!rc) rc = mpsd_tensor_reshape(R_out, 2, rd);
    if (!rc) rc = house_qr(ctx, site->ptr, m, n, small ? site->ptr : (double2*)ctx->scratch[3], R_out->ptr);
    if (rc) return rc;
    if (!small) {
      hipError_t e = hipMemcpyAsync(site->ptr, ctx->scratch[3], sizeof(double2) * (size_t)m * k,
                                    hipMemcpyDeviceToDevice, ctx->stream);
      if (e != hipSuccess) return hip_err(ctx, e, "gauge_r copy");
    }
    const int sd[3] = {d, Dl, k};
    rc = mpsd_tensor_reshape(site, 3, sd);
    if (rc) return rc;
    if (k == 1 && n == 1)
      hipLaunchKernelGGL(edge_phase_kernel, dim3(1), dim3(kT), 0, ctx->stream, site->ptr, site->n, R_out->ptr);
    ++site->version;
    ++R_out->version;
    return hip_err(ctx, hipGetLastError(), "gauge_r");
  }
  rocblas_handle h;
  int rc = handle_of(ctx, &h);
  if (rc) return rc;
  rc = ensure_scratch(ctx, 2, sizeof(double2) * (size_t)k);
  if (rc) return rc;
  rocblas_double_complex* A = (rocblas_double_complex*)site->ptr;
  rocblas_double_complex* tau = (rocblas_double_complex*)ctx->scratch[2];
  rc = solver_err(ctx, rocsolver_zgeqrf(h, m, n, A, m, tau), "zgeqrf");
  if (rc) return rc;
  const int rd[2] = {k, n};
  rc = mpsd_tensor_reshape(R_out, 2, rd);
  if (rc) return rc;
  hipLaunchKernelGGL(extract_r_kernel, dim3(nblk((int64_t)k * n)), dim3(kT), 0, ctx->stream, site->ptr, m, k, n,
                     R_out->ptr);
  rc = solver_err(ctx, rocsolver_zungqr(h, m, k, k, A, m, tau), "zungqr");
  if (rc) return rc;
  const int sd[3] = {d, Dl, k};
  rc = mpsd_tensor_reshape(site, 3, sd);
  if (rc) return rc;
  if (k == 1 && n == 1)
    hipLaunchKernelGGL(edge_phase_kernel, dim3(1), dim3(kT), 0, ctx->stream, site->ptr, site->n, R_out->ptr);
  ++site->version;
  ++R_out->version;
  return hip_err(ctx, hipGetLastError(), "gauge_r");
}

int mpsd_site_gauge_l(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* L_out) {
  if (!ctx || !site || !L_out || site == L_out) return MPSD_ERR_ARG;
  const int d = site->dims[0], Dl = site->rank > 1 ? site->dims[1] : 1, Dr = site->rank > 2 ? site->dims[2] : 1;
  const int m = Dl, n = d * Dr, k = m < n ? m : n;
  if (ctx->gauge_qr == 0 && house_qr_small(n, m)) {
    // the same LQ as below with the adjoint taken inside the single-workgroup QR
    const int ld[2] = {m, k};
    int rc = mpsd_tensor_reshape(L_out, 2, ld);
    if (!rc) rc = house_lq_site_small(ctx, site->ptr, d, Dl, Dr, site->ptr, L_out->ptr);
    const int sd[3] = {d, k, Dr};
    if (!rc) rc = mpsd_tensor_reshape(site, 3, sd);
    if (rc) return rc;
    if (k == 1 && m == 1)
      hipLaunchKernelGGL(edge_phase_kernel, dim3(1), dim3(kT), 0, ctx->stream, site->ptr, site->n, L_out->ptr);
    ++site->version;
    ++L_out->version;
    return hip_err(ctx, hipGetLastError(), "gauge_l");
  }
  if (ctx->gauge_qr == 0 && n <= house_qr_max_rows()) {
    // LQ of M[a, s + d c] = site[s, a, c] as the QR of B = M^H (n x m): M = R'^H Q'^H, so
    // L = R'^H (m x k) and the site becomes Q'^H reshaped to [d, k, Dr]
    int rc = ensure_scratch(ctx, 3, sizeof(double2) * (size_t)n * m);
    if (!rc) rc = ensure_scratch(ctx, 7, sizeof(double2) * (size_t)(n * k + k * m));
    if (rc) return rc;
    double2* B = (double2*)ctx->scratch[3];
    double2* Qp = (double2*)ctx->scratch[7];
    double2* Rp = Qp + (size_t)n * k;
    hipLaunchKernelGGL(site_to_bh_kernel, dim3(nblk((int64_t)m * n)), dim3(kT), 0, ctx->stream, site->ptr, d, Dl,
                       Dr, B);
    rc = house_qr(ctx, B, n, m, Qp, Rp);
    const int ld[2] = {m, k};
    if (!rc) rc = mpsd_tensor_reshape(L_out, 2, ld);
    const int sd[3] = {d, k, Dr};
    if (!rc) rc = mpsd_tensor_reshape(site, 3, sd);
    if (rc) return rc;
    hipLaunchKernelGGL(conj_transpose_kernel, dim3(nblk((int64_t)m * k)), dim3(kT), 0, ctx->stream, Rp, k, m,
                       L_out->ptr);
    hipLaunchKernelGGL(qh_to_site_kernel, dim3(nblk((int64_t)d * k * Dr)), dim3(kT), 0, ctx->stream, Qp, n, k, d, Dr,
                       site->ptr);
    if (k == 1 && m == 1)
      hipLaunchKernelGGL(edge_phase_kernel, dim3(1), dim3(kT), 0, ctx->stream, site->ptr, site->n, L_out->ptr);
    ++site->version;
    ++L_out->version;
    return hip_err(ctx, hipGetLastError(), "gauge_l");
  }
  rocblas_handle h;
  int rc = handle_of(ctx, &h);
  if (rc) return rc;
  rc = ensure_scratch(ctx, 2, sizeof(double2) * (size_t)k);
  if (!rc) rc = ensure_scratch(ctx, 3, sizeof(double2) * (size_t)m * n);
  if (rc) return rc;
  double2* T = (double2*)ctx->scratch[3];
  hipLaunchKernelGGL(permute_sad_kernel, dim3(nblk((int64_t)m * n)), dim3(kT), 0, ctx->stream, site->ptr, d, Dl,
                     Dr, T);
  rocblas_double_complex* tau = (rocblas_double_complex*)ctx->scratch[2];
  rc = solver_err(ctx, rocsolver_zgelqf(h, m, n, (rocblas_double_complex*)T, m, tau), "zgelqf");
  if (rc) return rc;
  const int ld[2] = {m, k};
  rc = mpsd_tensor_reshape(L_out, 2, ld);
  if (rc) return rc;
  hipLaunchKernelGGL(extract_l_kernel, dim3(nblk((int64_t)m * k)), dim3(kT), 0, ctx->stream, T, m, k, L_out->ptr);
  rc = solver_err(ctx, rocsolver_zunglq(h, k, n, k, (rocblas_double_complex*)T, m, tau), "zunglq");
  if (rc) return rc;
  const int sd[3] = {d, k, Dr};
  rc = mpsd_tensor_reshape(site, 3, sd);
  if (rc) return rc;
  hipLaunchKernelGGL(unpermute_q_kernel, dim3(nblk((int64_t)d * k * Dr)), dim3(kT), 0, ctx->stream, T, m, k, d,
                     Dr, site->ptr);
  if (k == 1 && m == 1)
    hipLaunchKernelGGL(edge_phase_kernel, dim3(1), dim3(kT), 0, ctx->stream, site->ptr, site->n, L_out->ptr);
  ++site->version;
  ++L_out->version;
  return hip_err(ctx, hipGetLastError(), "gauge_l");
}

}  // extern "C"

namespace mpsd {
// Thin SVD of the m x n matrix at A (destroyed): U (m x k) and V^H (k x n, leading dimension k)
// into U, V, singular values into S[0..k); returns the number kept under the rule of
// wrapper::svd with tol = 0 and Xi = cutD (src/wrapper/svd.cpp:23-118): the leading values
// above zero among the first min(cutD, k).
static int svd_keep(mpsd_ctx* ctx, double2* A, int m, int n, double2* U, double2* V, double* S, int cutD, int* kept) {
  rocblas_handle h;
  int rc = handle_of(ctx, &h);
  if (rc) return rc;
  const int k = m < n ? m : n;
  double* E = S + k;
  rocblas_int* info = (rocblas_int*)(S + 2 * k + 1);
  rc = solver_err(ctx,
                  rocsolver_zgesvd(h, rocblas_svect_singular, rocblas_svect_singular, m, n, (rocblas_double_complex*)A,
                                   m, S, (rocblas_double_complex*)U, m, (rocblas_double_complex*)V, k, E,
                                   rocblas_outofplace, info),
                  "zgesvd");
  if (rc) return rc;
  std::vector<double> hs(2 * k + 2);
  hipError_t e = hipMemcpyAsync(hs.data(), S, sizeof(double) * (2 * k + 2), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_err(ctx, e, "svd values");
  rocblas_int hinfo = 0;
  memcpy(&hinfo, &hs[2 * k + 1], sizeof(hinfo));
  if (hinfo != 0) return set_err(ctx, MPSD_ERR_NUMERIC, "zgesvd did not converge");
  const int lim = cutD < k ? cutD : k;
  int X = 0;
  while (X < lim && hs[X] > 0.0) ++X;
  if (X == 0) return set_err(ctx, MPSD_ERR_NUMERIC, "svd: no singular value above zero");
  *kept = X;
  return MPSD_OK;
}
}  // namespace mpsd

extern "C" {

int mpsd_site_gauge_r_svd(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* R_out, int cutD) {
  if (!ctx || !site || !R_out || site == R_out || cutD < 1) return MPSD_ERR_ARG;
  const int d = site->dims[0], Dl = site->rank > 1 ? site->dims[1] : 1, Dr = site->rank > 2 ? site->dims[2] : 1;
  const int m = d * Dl, n = Dr, k = m < n ? m : n;
  int rc = ensure_scratch(ctx, 3, sizeof(double2) * (size_t)m * k);
  if (!rc) rc = ensure_scratch(ctx, 7, sizeof(double2) * (size_t)k * n);
  if (!rc) rc = ensure_scratch(ctx, 2, sizeof(double) * (size_t)(2 * k + 2));
  if (rc) return rc;
  double2* U = (double2*)ctx->scratch[3];
  double2* V = (double2*)ctx->scratch[7];
  double* S = (double*)ctx->scratch[2];
  int X = 0;
  rc = svd_keep(ctx, site->ptr, m, n, U, V, S, cutD, &X);
  const int rd[2] = {X, n};
  if (!rc) rc = mpsd_tensor_reshape(R_out, 2, rd);
  const int sd[3] = {d, Dl, X};
  if (!rc) rc = mpsd_tensor_reshape(site, 3, sd);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(site->ptr, U, sizeof(double2) * (size_t)m * X, hipMemcpyDeviceToDevice, ctx->stream);
  if (e != hipSuccess) return hip_err(ctx, e, "gauge_r_svd copy");
  hipLaunchKernelGGL(scale_rows_kernel, dim3(nblk((int64_t)X * n)), dim3(kT), 0, ctx->stream, V, k, S, X, n,
                     R_out->ptr);
  if (X == 1 && n == 1)
    hipLaunchKernelGGL(edge_phase_kernel, dim3(1), dim3(kT), 0, ctx->stream, site->ptr, site->n, R_out->ptr);
  ++site->version;
  ++R_out->version;
  return hip_err(ctx, hipGetLastError(), "gauge_r_svd");
}

int mpsd_site_gauge_l_svd(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* L_out, int cutD) {
  if (!ctx || !site || !L_out || site == L_out || cutD < 1) return MPSD_ERR_ARG;
  const int d = site->dims[0], Dl = site->rank > 1 ? site->dims[1] : 1, Dr = site->rank > 2 ? site->dims[2] : 1;
  const int m = Dl, n = d * Dr, k = m < n ? m : n;
  int rc = ensure_scratch(ctx, 3, sizeof(double2) * (size_t)m * n);
  if (!rc) rc = ensure_scratch(ctx, 6, sizeof(double2) * (size_t)m * k);
  if (!rc) rc = ensure_scratch(ctx, 7, sizeof(double2) * (size_t)k * n);
  if (!rc) rc = ensure_scratch(ctx, 2, sizeof(double) * (size_t)(2 * k + 2));
  if (rc) return rc;
  double2* T = (double2*)ctx->scratch[3];
  double2* U = (double2*)ctx->scratch[6];
  double2* V = (double2*)ctx->scratch[7];
  double* S = (double*)ctx->scratch[2];
  // M[a, s + d c] = site[s, a, c]; M = U S V^H, leftterm = U S, site = V^H reshaped
  hipLaunchKernelGGL(permute_sad_kernel, dim3(nblk((int64_t)m * n)), dim3(kT), 0, ctx->stream, site->ptr, d, Dl, Dr, T);
  int X = 0;
  rc = svd_keep(ctx, T, m, n, U, V, S, cutD, &X);
  const int ld[2] = {m, X};
  if (!rc) rc = mpsd_tensor_reshape(L_out, 2, ld);
  const int sd[3] = {d, X, Dr};
  if (!rc) rc = mpsd_tensor_reshape(site, 3, sd);
  if (rc) return rc;
  hipLaunchKernelGGL(scale_cols_kernel, dim3(nblk((int64_t)m * X)), dim3(kT), 0, ctx->stream, U, m, S, X, L_out->ptr);
  hipLaunchKernelGGL(unpermute_q_kernel, dim3(nblk((int64_t)d * X * Dr)), dim3(kT), 0, ctx->stream, V, k, X, d, Dr,
                     site->ptr);
  if (X == 1 && m == 1)
    hipLaunchKernelGGL(edge_phase_kernel, dim3(1), dim3(kT), 0, ctx->stream, site->ptr, site->n, L_out->ptr);
  ++site->version;
  ++L_out->version;
  return hip_err(ctx, hipGetLastError(), "gauge_l_svd");
}

// The site's result goes to a fresh pooled buffer of n elements with shape [d, Dl, Dr], and the
// old buffer (returned in *old / *oldcap) is read by the launch that fills the new one, then
// pooled after it is enqueued: reuse is stream-ordered, so no copy of the old contents is needed.
static int take_new_buffer(mpsd_ctx* ctx, mpsd_tensor* site, int d, int Dl, int Dr, double2** old,
                           int64_t* oldcap) {
  const int64_t n = (int64_t)d * Dl * Dr;
  double2* nb = nullptr;
  int64_t ncap = 0;
  hipError_t e = pool_alloc(ctx, n, &nb, &ncap);
  if (e != hipSuccess) return hip_err(ctx, e, "site buffer");
  *old = site->ptr;
  *oldcap = site->cap;
  ctx->tensor_bytes += (int64_t)sizeof(double2) * (ncap - site->cap);
  site->ptr = nb;
  site->cap = ncap;
  site->rank = 3;
  for (int i = 0; i < 8; ++i) site->dims[i] = 0;
  site->dims[0] = d;
  site->dims[1] = Dl;
  site->dims[2] = Dr;
  site->n = n;
  return MPSD_OK;
}

int mpsd_site_reduce_dim_r(mpsd_ctx* ctx, mpsd_tensor* site, const mpsd_tensor* R) {
  // site[s, i, c] = sum_j R[i, j] site_old[s, j, c]   (Site::reduceDimR, Site.cpp:240-261)
  if (!ctx || !site || !R || site == R) return MPSD_ERR_ARG;
  const int d = site->dims[0], Dl = site->rank > 1 ? site->dims[1] : 1, Dr = site->rank > 2 ? site->dims[2] : 1;
  const int Rm = R->dims[0], Rn = R->rank > 1 ? R->dims[1] : 1;
  if (Rn != Dl) return set_err(ctx, MPSD_ERR_SHAPE, "reduce_dim_r: factor does not match the left bond");
  double2* old = nullptr;
  int64_t oldcap = 0;
  int rc = take_new_buffer(ctx, site, d, Rm, Dr, &old, &oldcap);
  if (rc) return rc;
  ZgemmDesc g;
  g.A = R->ptr;
  g.B = old;
  g.C = site->ptr;
  g.M = Rm;
  g.N = Dr;
  g.K = Dl;
  g.NB = d;
  g.sAm = 1;
  g.sAk = Rm;
  g.sAb = 0;
  g.sBk = d;
  g.sBn = (int64_t)d * Dl;
  g.sBb = 1;
  g.sCm = d;
  g.sCn = (int64_t)d * Rm;
  g.sCb = 1;
  g.a_mfast = 1;
  g.b_kfast = 1;
  ++site->version;
  const hipError_t e = launch_zbrgemm_f64(g, TileCfg{32, 32}, ctx->stream);
  pool_free(ctx, old, oldcap);
  return hip_err(ctx, e, "reduce_dim_r");
}

int mpsd_site_gauge_r_absorb(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* R_out, mpsd_tensor* next) {
  // gauge_r(site, R) then reduce_dim_r(next, R) (MPS::gaugeCond, MPS.cpp:322-337, 392-401); one
  // launch when the single-workgroup QR takes the site
  if (!ctx || !site || !R_out || !next || site == R_out || next == R_out || next == site) return MPSD_ERR_ARG;
  const int d = site->dims[0], Dl = site->rank > 1 ? site->dims[1] : 1, Dr = site->rank > 2 ? site->dims[2] : 1;
  const int m = d * Dl, n = Dr, k = m < n ? m : n;
  const int nd = next->dims[0], nDl = next->rank > 1 ? next->dims[1] : 1, nDr = next->rank > 2 ? next->dims[2] : 1;
  if (!(ctx->gauge_qr == 0 && house_qr_small(m, n) && !(k == 1 && n == 1) && nDl == n && next->rank <= 3)) {
    const int rc = mpsd_site_gauge_r(ctx, site, R_out);
    return rc ? rc : mpsd_site_reduce_dim_r(ctx, next, R_out);
  }
  const int rd[2] = {k, n};
  int rc = mpsd_tensor_reshape(R_out, 2, rd);
  if (rc) return rc;
  double2* old = nullptr;
  int64_t oldcap = 0;
  rc = take_new_buffer(ctx, next, nd, k, nDr, &old, &oldcap);
  if (rc) return rc;
  Absorb ab;
  ab.old = old;
  ab.out = next->ptr;
  ab.nd = nd;
  ab.nD = nDr;
  rc = house_qr_small_absorb(ctx, site->ptr, m, n, site->ptr, R_out->ptr, ab);
  pool_free(ctx, old, oldcap);
  if (rc) return rc;
  const int sd[3] = {d, Dl, k};
  rc = mpsd_tensor_reshape(site, 3, sd);
  if (rc) return rc;
  ++site->version;
  ++R_out->version;
  ++next->version;
  return MPSD_OK;
}

int mpsd_site_gauge_l_absorb(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* L_out, mpsd_tensor* prev) {
  // gauge_l(site, L) then reduce_dim_l(prev, L) (MPS::gaugeCond, MPS.cpp:310-320, 370-379); one
  // launch when the single-workgroup QR takes the site
  if (!ctx || !site || !L_out || !prev || site == L_out || prev == L_out || prev == site) return MPSD_ERR_ARG;
  const int d = site->dims[0], Dl = site->rank > 1 ? site->dims[1] : 1, Dr = site->rank > 2 ? site->dims[2] : 1;
  const int m = Dl, n = d * Dr, k = m < n ? m : n;
  const int pd = prev->dims[0], pDl = prev->rank > 1 ? prev->dims[1] : 1, pDr = prev->rank > 2 ? prev->dims[2] : 1;
  if (!(ctx->gauge_qr == 0 && house_qr_small(n, m) && !(k == 1 && m == 1) && pDr == m && prev->rank <= 3)) {
    const int rc = mpsd_site_gauge_l(ctx, site, L_out);
    return rc ? rc : mpsd_site_reduce_dim_l(ctx, prev, L_out);
  }
  const int ld[2] = {m, k};
  int rc = mpsd_tensor_reshape(L_out, 2, ld);
  if (rc) return rc;
  double2* old = nullptr;
  int64_t oldcap = 0;
  rc = take_new_buffer(ctx, prev, pd, pDl, k, &old, &oldcap);
  if (rc) return rc;
  Absorb ab;
  ab.old = old;
  ab.out = prev->ptr;
  ab.nd = pd;
  ab.nD = pDl;
  rc = house_lq_site_small(ctx, site->ptr, d, Dl, Dr, site->ptr, L_out->ptr, ab);
  pool_free(ctx, old, oldcap);
  if (rc) return rc;
  const int sd[3] = {d, k, Dr};
  rc = mpsd_tensor_reshape(site, 3, sd);
  if (rc) return rc;
  ++site->version;
  ++L_out->version;
  ++prev->version;
  return MPSD_OK;
}

int mpsd_site_reduce_dim_l(mpsd_ctx* ctx, mpsd_tensor* site, const mpsd_tensor* L) {
  // site[s, a, j] = sum_c site_old[s, a, c] L[c, j]   (Site::reduceDimL, Site.cpp:263-283)
  if (!ctx || !site || !L || site == L) return MPSD_ERR_ARG;
  const int d = site->dims[0], Dl = site->rank > 1 ? site->dims[1] : 1, Dr = site->rank > 2 ? site->dims[2] : 1;
  const int Lm = L->dims[0], Ln = L->rank > 1 ? L->dims[1] : 1;
  if (Lm != Dr) return set_err(ctx, MPSD_ERR_SHAPE, "reduce_dim_l: factor does not match the right bond");
  double2* old = nullptr;
  int64_t oldcap = 0;
  int rc = take_new_buffer(ctx, site, d, Dl, Ln, &old, &oldcap);
  if (rc) return rc;
  ZgemmDesc g;
  g.A = old;
  g.B = L->ptr;
  g.C = site->ptr;
  g.M = d * Dl;
  g.N = Ln;
  g.K = Dr;
  g.NB = 1;
  g.sAm = 1;
  g.sAk = (int64_t)d * Dl;
  g.sBk = 1;
  g.sBn = Lm;
  g.sCm = 1;
  g.sCn = (int64_t)d * Dl;
  g.a_mfast = 1;
  g.b_kfast = 1;
  ++site->version;
  const hipError_t e = launch_zbrgemm_f64(g, TileCfg{32, 32}, ctx->stream);
  pool_free(ctx, old, oldcap);
  return hip_err(ctx, e, "reduce_dim_l");
}

}  // extern "C"
