// Whole restarted-Lanczos eigensolve of a small effective Hamiltonian in one workgroup.
//
// For local problems of a thousand unknowns or so (the D <= 24 sites of a sweep) every
// kernel of a Lanczos step runs for a few microseconds and the launch latency dominates.  Here one
// workgroup of 1024 threads does everything a site's eigensolve needs, in one launch:
//
// * the matvec y = (L W R) x (+ offset x) in three stages through global scratch T1, T2
//   (the contraction order of TensorMultiplier::product, src/wrapper/multipliers/
//   TensorMultiplier.cpp:61-93);
// * two-pass classical Gram-Schmidt against the Krylov basis (one dot per wave, fixed order);
// * alpha_j, beta_j as the stream and graph solvers form them (beta from the second pass);
// * the tridiagonal eigenproblem every third step: the lowest eigenvalue by multisection
//   (256 Sturm counts per round), its eigenvector by inverse iteration on one thread;
// * restarts from the Ritz vector, and the final normalised Ritz vector in x.
//
// The host reads one small status block per eigensolve.

#include <cfloat>

#include "mpsd_internal.h"

namespace mpsd {
namespace {

constexpr int kFT = 1024;
constexpr int kFW = kFT / 64;
constexpr int kFM = 64;  // largest Krylov basis (m + 1)
constexpr int kMS = 256; // points per multisection round

struct FusedArgs {
  const double2* L;
  const double2* W;
  const double2* R;
  double offset;
  int d, Db, wl, Dk, Dbr, wr, Dkr;
  double2* x;
  double2* V;
  double2* T1;
  double2* T2;
  int m, max_restarts;
  double tol;
  double* out;  // theta, residual, matvecs, restarts, converged, status
};

__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {  // c + a b
  return double2{fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}
__device__ __forceinline__ double2 cjfma(double2 a, double2 b, double2 c) {  // c + conj(a) b
  return double2{fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y))};
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// y = H x on the whole workgroup (x and y distinct, both n long).
__device__ void fused_matvec(const FusedArgs& a, const double2* __restrict__ x, double2* __restrict__ y) {
  const int tid = threadIdx.x;
  const int d = a.d, Db = a.Db, wl = a.wl, Dk = a.Dk, Dbr = a.Dbr, wr = a.wr, Dkr = a.Dkr;
  // stage 1: T1[b, al, s, ar] = sum_a L[b, al, a] x[s, a, ar]
  const int64_t n1 = (int64_t)Db * wl * d * Dkr;
  for (int64_t t = tid; t < n1; t += kFT) {
    int64_t q = t;
    const int b = (int)(q % Db);
    q /= Db;
    const int al = (int)(q % wl);
    q /= wl;
    const int s = (int)(q % d);
    const int ar = (int)(q / d);
    const double2* Lp = a.L + b + (int64_t)Db * al;
    const double2* xp = x + s + (int64_t)d * Dk * ar;
    double2 acc{0.0, 0.0};
    for (int k = 0; k < Dk; ++k) acc = cfma(Lp[(int64_t)Db * wl * k], xp[(int64_t)d * k], acc);
    a.T1[t] = acc;
  }
  __syncthreads();
  // stage 2: T2[b, s', be, ar] = sum_{al, s} W[s', al, s, be] T1[b, al, s, ar]
  const int64_t n2 = (int64_t)Db * d * wr * Dkr;
  for (int64_t t = tid; t < n2; t += kFT) {
    int64_t q = t;
    const int b = (int)(q % Db);
    q /= Db;
    const int sp = (int)(q % d);
    q /= d;
    const int be = (int)(q % wr);
    const int ar = (int)(q / wr);
    double2 acc{0.0, 0.0};
    for (int s = 0; s < d; ++s)
      for (int al = 0; al < wl; ++al)
        acc = cfma(a.W[sp + (int64_t)d * (al + (int64_t)wl * (s + (int64_t)d * be))],
                   a.T1[b + (int64_t)Db * (al + (int64_t)wl * (s + (int64_t)d * ar))], acc);
    a.T2[t] = acc;
  }
  __syncthreads();
  // stage 3: y[s', b, c] = sum_{be, ar} T2[b, s', be, ar] R[c, be, ar] (+ offset x)
  const int64_t n3 = (int64_t)d * Db * Dbr;
  for (int64_t t = tid; t < n3; t += kFT) {
    int64_t q = t;
    const int sp = (int)(q % d);
    q /= d;
    const int b = (int)(q % Db);
    const int c = (int)(q / Db);
    double2 acc{0.0, 0.0};
    for (int ar = 0; ar < Dkr; ++ar)
      for (int be = 0; be < wr; ++be)
        acc = cfma(a.T2[b + (int64_t)Db * (sp + (int64_t)d * (be + (int64_t)wr * ar))],
                   a.R[c + (int64_t)Dbr * (be + (int64_t)wr * ar)], acc);
    if (a.offset != 0.0) {
      const double2 xi = x[t];
      acc.x = fma(a.offset, xi.x, acc.x);
      acc.y = fma(a.offset, xi.y, acc.y);
    }
    y[t] = acc;
  }
  __syncthreads();
}

// h[i] = V_i^H w for i < cnt (one dot per wave, fixed order); h is shared, valid after the call.
__device__ void multi_dot(const double2* __restrict__ V, int64_t n, int cnt, const double2* __restrict__ w,
                          double2* h) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = wave; i < cnt; i += kFW) {
    const double2* v = V + (int64_t)i * n;
    double2 acc{0.0, 0.0};
    for (int64_t t = lane; t < n; t += 64) acc = cjfma(v[t], w[t], acc);
    const double sx = wave_sum(acc.x), sy = wave_sum(acc.y);
    if (lane == 0) h[i] = double2{sx, sy};
  }
  __syncthreads();
}

// w -= sum_{i < cnt} V_i h[i]
__device__ void gemv_sub(const double2* __restrict__ V, int64_t n, int cnt, const double2* h, double2* __restrict__ w) {
  for (int64_t t = threadIdx.x; t < n; t += kFT) {
    double2 acc{0.0, 0.0};
    for (int i = 0; i < cnt; ++i) acc = cfma(V[(int64_t)i * n + t], h[i], acc);
    const double2 wi = w[t];
    w[t] = double2{wi.x - acc.x, wi.y - acc.y};
  }
  __syncthreads();
}

// Number of eigenvalues of the tridiagonal (a, b) below x (Sturm sequence).
__device__ int sturm(const double* a, const double* b, int m, double x) {
  const double tiny = 1e-300;
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0) ++cnt;
  for (int i = 1; i < m; ++i) {
    if (fabs(q) < tiny) q = q < 0 ? -tiny : tiny;
    q = a[i] - x - b[i - 1] * b[i - 1] / q;
    if (q < 0) ++cnt;
  }
  return cnt;
}

// (T - sigma) y = r in place on one thread (partial pivoting); dd, u, v, l: shared work
// arrays of length >= m (m <= kFM).
__device__ void tri_solve(const double* a, const double* b, int m, double sigma, double* r, double* dd, double* u,
                          double* v, double* l) {
  if (m == 1) {
    double p = a[0] - sigma;
    if (fabs(p) < 1e-300) p = 1e-300;
    r[0] /= p;
    return;
  }
  for (int i = 0; i < m; ++i) {
    dd[i] = a[i] - sigma;
    u[i] = v[i] = l[i] = 0.0;
  }
  for (int i = 0; i + 1 < m; ++i) {
    u[i] = b[i];
    l[i + 1] = b[i];
  }
  for (int i = 0; i + 1 < m; ++i) {
    if (fabs(l[i + 1]) > fabs(dd[i])) {
      double t;
      t = dd[i]; dd[i] = l[i + 1]; l[i + 1] = t;
      t = u[i]; u[i] = dd[i + 1]; dd[i + 1] = t;
      t = v[i]; v[i] = u[i + 1]; u[i + 1] = t;
      t = r[i]; r[i] = r[i + 1]; r[i + 1] = t;
    }
    double piv = dd[i];
    if (fabs(piv) < 1e-300) piv = dd[i] = 1e-300;
    const double f = l[i + 1] / piv;
    dd[i + 1] -= f * u[i];
    u[i + 1] -= f * v[i];
    r[i + 1] -= f * r[i];
  }
  if (fabs(dd[m - 1]) < 1e-300) dd[m - 1] = 1e-300;
  r[m - 1] /= dd[m - 1];
  r[m - 2] = (r[m - 2] - u[m - 2] * r[m - 1]) / dd[m - 2];
  for (int i = m - 3; i >= 0; --i) r[i] = (r[i] - u[i] * r[i + 1] - v[i] * r[i + 2]) / dd[i];
}

__global__ void __launch_bounds__(kFT) lanczos_fused_kernel(FusedArgs a) {
  __shared__ double2 h1[kFM + 1], h2[kFM + 1];
  __shared__ double al[kFM], be[kFM], s[kFM];
  __shared__ double wk[4][kFM];  // inverse-iteration work arrays (thread 0)
  __shared__ int cnt[kFT];
  __shared__ double lo_s, hi_s, inv_s, theta_s, res_s;
  __shared__ int stop_s, status_s;
  const int tid = threadIdx.x;
  const int64_t n = (int64_t)a.d * a.Dk * a.Dkr;
  const int m = a.m;
  double2* x = a.x;
  int matvecs = 0, restarts = 0, converged = 0;
  if (tid == 0) status_s = 0;
  for (int restart = 0; restart <= a.max_restarts; ++restart) {
    restarts = restart;
    // V_0 = x / |x|
    multi_dot(x, n, 1, x, h1);
    const double nx = h1[0].x;
    if (!(nx > 0.0) || !isfinite(nx)) {
      if (tid == 0) status_s = 2;  // the host reseeds
      break;
    }
    const double ix = 1.0 / sqrt(nx);
    for (int64_t t = tid; t < n; t += kFT) a.V[t] = double2{x[t].x * ix, x[t].y * ix};
    if (tid == 0) stop_s = 0;
    __syncthreads();
    int K = 0;
    for (int j = 0; j < m; ++j) {
      double2* w = a.V + (int64_t)(j + 1) * n;
      fused_matvec(a, a.V + (int64_t)j * n, w);
      ++matvecs;
      multi_dot(a.V, n, j + 1, w, h1);
      gemv_sub(a.V, n, j + 1, h1, w);
      multi_dot(a.V, n, j + 2, w, h2);  // h2[j + 1] = |w|^2
      gemv_sub(a.V, n, j + 1, h2, w);
      if (tid == 0) {
        const double alpha = h1[j].x + h2[j].x;
        double corr = 0.0;
        for (int i = 0; i <= j; ++i) corr = fma(h2[i].x, h2[i].x, fma(h2[i].y, h2[i].y, corr));
        const double beta = sqrt(fmax(h2[j + 1].x - corr, 0.0));
        al[j] = alpha;
        be[j] = beta;
        inv_s = beta > 1e-300 ? 1.0 / beta : 0.0;
        if (!isfinite(alpha) || !isfinite(beta)) status_s = 3;
        // Gershgorin interval of T_j
        double lo = al[0], hi = al[0];
        for (int i = 0; i <= j; ++i) {
          const double rad = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i < j ? fabs(be[i]) : 0.0);
          lo = fmin(lo, al[i] - rad);
          hi = fmax(hi, al[i] + rad);
        }
        lo_s = lo;
        hi_s = hi;
      }
      __syncthreads();
      if (status_s == 3) break;
      const double inv = inv_s;
      for (int64_t t = tid; t < n; t += kFT) w[t] = double2{w[t].x * inv, w[t].y * inv};
      __syncthreads();  // V_{j+1} complete before the next matvec reads it
      K = j + 1;
      // the tridiagonal problem is solved every third step and at the end of the basis
      if ((j + 1) % 3 != 0 && j + 1 < m && (int64_t)(j + 1) < n) continue;
      // lowest eigenvalue of T_j by multisection: kMS threads count at their own points
      for (int round = 0; round < 10; ++round) {
        const double lo = lo_s, hi = hi_s;
        if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi))) break;
        const double step = (hi - lo) / (kMS + 1);
        const double xk = lo + step * (tid + 1);
        if (tid < kMS) cnt[tid] = (xk > lo && xk < hi) ? (sturm(al, be, j + 1, xk) >= 1) : (xk >= hi ? 1 : 0);
        __syncthreads();
        // the first point whose count reaches one bounds the eigenvalue from above
        if (tid < kMS && cnt[tid] && (tid == 0 || !cnt[tid - 1])) {
          lo_s = tid == 0 ? lo : lo + step * tid;
          hi_s = xk;
        }
        __syncthreads();
        if (tid == 0 && !cnt[kMS - 1]) lo_s = lo + step * kMS;  // above the last point
        __syncthreads();
      }
      if (tid == 0) {
        const double theta = 0.5 * (lo_s + hi_s);
        const int mm = j + 1;
        for (int i = 0; i < mm; ++i) s[i] = 1.0;
        const double shift = theta - 4.0 * DBL_EPSILON * fmax(1.0, fabs(theta));
        for (int it = 0; it < 3; ++it) {
          tri_solve(al, be, mm, shift, s, wk[0], wk[1], wk[2], wk[3]);
          double nrm = 0.0;
          for (int i = 0; i < mm; ++i) nrm += s[i] * s[i];
          nrm = sqrt(nrm);
          if (!(nrm > 0.0) || !isfinite(nrm)) {
            for (int i = 0; i < mm; ++i) s[i] = i == 0 ? 1.0 : 0.0;
            break;
          }
          for (int i = 0; i < mm; ++i) s[i] /= nrm;
        }
        const double res = be[j] * fabs(s[j]);
        const double scale = fmax(1.0, fabs(theta));
        theta_s = theta;
        res_s = res;
        stop_s = (res <= a.tol * scale || be[j] <= 1e-14 * scale || (int64_t)(j + 1) >= n) ? 1 : 0;
      }
      __syncthreads();
      if (stop_s) break;
    }
    if (status_s == 3) break;
    // Ritz vector x = V s / |V s|
    for (int64_t t = tid; t < n; t += kFT) {
      double2 acc{0.0, 0.0};
      for (int i = 0; i < K; ++i) {
        const double2 v = a.V[(int64_t)i * n + t];
        acc.x = fma(v.x, s[i], acc.x);
        acc.y = fma(v.y, s[i], acc.y);
      }
      x[t] = acc;
    }
    __syncthreads();
    multi_dot(x, n, 1, x, h1);
    const double ir = h1[0].x > 0.0 ? 1.0 / sqrt(h1[0].x) : 0.0;
    for (int64_t t = tid; t < n; t += kFT) x[t] = double2{x[t].x * ir, x[t].y * ir};
    __syncthreads();
    if (stop_s) {
      converged = 1;
      break;
    }
  }
  if (tid == 0) {
    a.out[0] = theta_s;
    a.out[1] = res_s;
    a.out[2] = matvecs;
    a.out[3] = restarts;
    a.out[4] = converged;
    a.out[5] = status_s;
  }
}

}  // namespace

bool lanczos_fused_fits(const mpsd_heff* h, int m) {
  const int64_t n = (int64_t)h->d * h->Dk * h->Dkr;
  const double F = 8.0 * ((double)h->d * h->wl * h->Db * h->Dk * h->Dkr +
                          (double)h->d * h->d * h->wl * h->wr * h->Db * h->Dkr +
                          (double)h->d * h->wr * h->Db * h->Dbr * h->Dkr);
  // one CU's FP64 rate (~0.25 TFLOP/s) against the many-CU path's per-step launch cost
  // (~60 us): the one-workgroup solver wins up to ~1.5 MFLOP per matvec (D <= 24 at w = 3)
  return m + 1 <= kFM && n <= 2048 && F <= 1.5e6 && (int64_t)h->d * h->Db * h->Dbr == n &&
         h->precision == MPSD_PREC_FP64 && !h->herm;
}

int lanczos_fused(mpsd_heff* h, double* x, double tol, int m, int max_restarts, double* out6) {
  mpsd_ctx* ctx = h->ctx;
  const int64_t n = (int64_t)h->d * h->Dk * h->Dkr;
  const int64_t t1 = (int64_t)h->Db * h->wl * h->d * h->Dkr, t2 = (int64_t)h->Db * h->d * h->wr * h->Dkr;
  int rc = ensure_scratch(ctx, 0, sizeof(double2) * (size_t)t1);
  if (!rc) rc = ensure_scratch(ctx, 1, sizeof(double2) * (size_t)t2);
  if (!rc) rc = ensure_scratch(ctx, 8, sizeof(double2) * (size_t)((m + 1) * n) + 8 * sizeof(double));
  if (rc) return rc;
  FusedArgs a;
  a.L = h->L->ptr;
  a.W = h->W->ptr;
  a.R = h->R->ptr;
  a.offset = h->offset;
  a.d = h->d;
  a.Db = h->Db;
  a.wl = h->wl;
  a.Dk = h->Dk;
  a.Dbr = h->Dbr;
  a.wr = h->wr;
  a.Dkr = h->Dkr;
  a.x = (double2*)x;
  a.V = (double2*)ctx->scratch[8];
  a.T1 = (double2*)ctx->scratch[0];
  a.T2 = (double2*)ctx->scratch[1];
  a.m = m;
  a.max_restarts = max_restarts;
  a.tol = tol;
  a.out = (double*)(a.V + (m + 1) * n);
  hipLaunchKernelGGL(lanczos_fused_kernel, dim3(1), dim3(kFT), 0, ctx->stream, a);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out6, a.out, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return hip_err(ctx, e, "lanczos_fused");
}

}  // namespace mpsd

using namespace mpsd;

extern "C" int mpsd_heff_eigs_fused(mpsd_heff* h, double* x, double tol, int max_krylov, int max_restarts,
                                    double* out6) {
  if (!h || !x || !out6 || max_krylov < 1 || max_restarts < 0) return MPSD_ERR_ARG;
  if (!lanczos_fused_fits(h, max_krylov))
    return set_err(h->ctx, MPSD_ERR_UNSUPPORTED, "eigs_fused: problem too large for the one-workgroup solver");
  return lanczos_fused(h, x, tol > 0.0 ? tol : 1e-10, max_krylov, max_restarts, out6);
}
