// Whole restarted-Lanczos eigensolve of a small effective Hamiltonian in one workgroup.
//
// For local problems of a thousand unknowns or so (the D <= 24 sites of a sweep) every
// kernel of a Lanczos step runs for a few microseconds and the launch latency dominates.  Here one
// workgroup of 1024 threads does everything a site's eigensolve needs, in one launch:
//
// * the matvec y = (L W R) x (+ offset x) in three stages through T1, T2 (the contraction
//   order of TensorMultiplier::product, src/wrapper/multipliers/TensorMultiplier.cpp:61-93),
//   the two GEMM stages on FP64 MFMA one 16 x 16 block per wave, the MPO mix on the vector
//   units.
//   L, R, W, the current basis vector and the new vector live in LDS (the operands loaded
//   once per eigensolve), and T1, T2 too when they fit: the inner products of the three
//   stages read LDS instead of waiting on L2 in dependent chains;
// * two-pass classical Gram-Schmidt of the new vector (in LDS) against the Krylov basis (in
//   global memory; one dot per wave, fixed order);
// * alpha_j, beta_j as the stream and graph solvers form them (beta from the second pass),
//   reduced across one wave;
// * the tridiagonal eigenproblem every third step: the lowest eigenvalue by multisection
//   (256 Sturm counts per round), its eigenvector by two steps of inverse iteration on one
//   wave, row i of the tridiagonal in lane i's registers (taken by readlane, no LDS round
//   trip in the elimination's dependent chain); both divide through a refined hardware
//   reciprocal;
// * restarts from the Ritz vector, and the final normalised Ritz vector in x.
//
// The host reads one small status block per eigensolve.

#include <cfloat>

#include "mpsd_internal.h"
#include "wave_xor.h"

namespace mpsd {
namespace {

constexpr int kFT = 1024;
constexpr int kFW = kFT / 64;
constexpr int kFM = 64;       // largest Krylov basis (m + 1)
constexpr int kMS = 256;      // points per multisection round
constexpr double kGeoLog2 = 18.0 * 3.321928094887362 / kMS;  // log2(1/g), g^kMS = 1e-18
constexpr int kPool = 9472;   // complex elements of LDS for operands and vectors (148 KiB)

// LDS footprint of a problem, in complex elements: the operands with the current and the new
// basis vector, and the two intermediates.
struct FusedLds {
  int64_t ops, inter;
};
inline FusedLds fused_lds(const mpsd_heff* h) {
  const int64_t n = (int64_t)h->d * h->Dk * h->Dkr;
  FusedLds f;
  f.ops = (int64_t)h->Db * h->wl * h->Dk + (int64_t)h->Dbr * h->wr * h->Dkr +
          (int64_t)h->d * h->wl * h->d * h->wr + 2 * n;
  f.inter = (int64_t)h->Db * h->wl * h->d * h->Dkr + (int64_t)h->Db * h->d * h->wr * h->Dkr;
  return f;
}

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
__device__ __forceinline__ double wave_sum(double v) { return wave_sum_x(v); }
__device__ __forceinline__ double wave_min(double v) { return wave_min_x(v); }
__device__ __forceinline__ double wave_max(double v) { return wave_max_x(v); }

typedef double fd4 __attribute__((ext_vector_type(4)));

// 1 / v: the hardware reciprocal estimate refined by two Newton-Raphson steps
// (r <- r + r (1 - v r)), accurate to about an ulp; much shorter than an IEEE division in the
// serial recurrences of the tridiagonal solver.
__device__ __forceinline__ double recip(double v) {
  double r = __builtin_amdgcn_rcp(v);
  r = fma(r, fma(-v, r, 1.0), r);
  r = fma(r, fma(-v, r, 1.0), r);
  return r;
}

// One 16 x 16 block C(m0.., n0..) = sum_k A(m, k) B(k, n) of a small complex GEMM on one wave,
// with v_mfma_f64_16x16x4f64 in the three-product form (P1 = Ar Br, P2 = Ai Bi,
// P3 = (Ar + Ai)(Br + Bi); re = P1 - P2, im = P3 - P1 - P2).  A(m, k) and B(k, n) are read
// through the accessors (out-of-range indices read as zero); the result element (m0 + q + 4 r,
// n0 + (lane & 15)), q = lane >> 4, is handed to store(m, n, value) for r = 0..3.
template <class FA, class FB, class FS>
__device__ __forceinline__ void wave_zgemm16(int M, int N, int K, int m0, int n0, FA A, FB B, FS store) {
  const int lane = threadIdx.x & 63;
  const int lc = lane & 15, q = lane >> 4;
  fd4 p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
  const int m = m0 + lc, n = n0 + lc;
  // unrolled so that the LDS reads of later k-steps are issued ahead of this step's MFMAs
#pragma unroll 4
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + q;
    const double2 av = (m < M && k < K) ? A(m, k) : double2{0.0, 0.0};
    const double2 bv = (n < N && k < K) ? B(k, n) : double2{0.0, 0.0};
    p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, p1, 0, 0, 0);
    p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.y, p2, 0, 0, 0);
    p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x + av.y, bv.x + bv.y, p3, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int mo = m0 + q + 4 * r;
    if (mo < M && n < N) store(mo, n, double2{p1[r] - p2[r], p3[r] - (p1[r] + p2[r])});
  }
}

// Where the matvec's operands live: L, R, W, the input X and the output Y in LDS; T1, T2 in
// LDS or in global scratch.
struct FusedOps {
  const double2* L;
  const double2* R;
  const double2* W;
  double2* X;
  double2* Y;
  double2* T1;
  double2* T2;
};

// t / D as (t mgc) >> 32 with mgc = ceil(2^32 / D) (33 bits for D = 1): exact for D <= 2^10
// and t < 2^20, since t (mgc D - 2^32) < 2^32 keeps the error term below 1 / D
struct FastDiv {
  unsigned long long mgc;
  __device__ explicit FastDiv(unsigned d) : mgc((0x100000000ull + d - 1) / d) {}
  __device__ __forceinline__ unsigned div(unsigned t) const { return (unsigned)(((unsigned long long)t * mgc) >> 32); }
};

// Y = H X on the whole workgroup.
__device__ void fused_matvec(const FusedArgs& a, const FusedOps& o) {
  const int tid = threadIdx.x;
  const int d = a.d, Db = a.Db, wl = a.wl, Dk = a.Dk, Dbr = a.Dbr, wr = a.wr, Dkr = a.Dkr;
  const double2* x = o.X;
  const int wave = tid >> 6;
  // stage 1: T1[(b, al), (s, ar)] = sum_a L[(b, al), a] x[s, a, ar], one 16 x 16 block per wave
  {
    const int M = Db * wl, N = d * Dkr;
    const int gm = (M + 15) / 16, gn = (N + 15) / 16;
    const double2* Lm = o.L;
    double2* T1 = o.T1;
    for (int blk = wave; blk < gm * gn; blk += kFW)
      wave_zgemm16(
          M, N, Dk, 16 * (blk % gm), 16 * (blk / gm), [&](int mm, int k) { return Lm[mm + M * k]; },
          [&](int k, int nn) { return x[(nn % d) + d * k + d * Dk * (nn / d)]; },
          [&](int mm, int nn, double2 v) { T1[mm + M * nn] = v; });
  }
  __syncthreads();
  // stage 2: T2[b, s', be, ar] = sum_{al, s} W[s', al, s, be] T1[b, al, s, ar]
  const int n2 = Db * d * wr * Dkr;
  const FastDiv fDb(Db), fd(d), fwr(wr);
  for (int t = tid; t < n2; t += kFT) {
    unsigned q = (unsigned)t;
    unsigned q1 = fDb.div(q);
    const int b = (int)(q - q1 * (unsigned)Db);
    q = q1;
    q1 = fd.div(q);
    const int sp = (int)(q - q1 * (unsigned)d);
    q = q1;
    q1 = fwr.div(q);
    const int be = (int)(q - q1 * (unsigned)wr);
    const int ar = (int)q1;
    double2 acc{0.0, 0.0};
    for (int s = 0; s < d; ++s)
      for (int al = 0; al < wl; ++al)
        acc = cfma(o.W[sp + d * (al + wl * (s + d * be))], o.T1[b + Db * (al + wl * (s + d * ar))], acc);
    o.T2[t] = acc;
  }
  __syncthreads();
  // stage 3: y[s', b, c] = sum_{(be, ar)} T2[(b, s'), (be, ar)] R[c, (be, ar)] (+ offset x)
  {
    const int M = Db * d, N = Dbr, K = wr * Dkr;
    const int gm = (M + 15) / 16, gn = (N + 15) / 16;
    const double2* T2 = o.T2;
    const double2* Rm = o.R;
    double2* Y = o.Y;
    const double off = a.offset;
    for (int blk = wave; blk < gm * gn; blk += kFW)
      wave_zgemm16(
          M, N, K, 16 * (blk % gm), 16 * (blk / gm), [&](int mm, int k) { return T2[mm + M * k]; },
          [&](int k, int nn) { return Rm[nn + Dbr * k]; },
          [&](int mm, int nn, double2 v) {
            const int t = (mm / Db) + d * (mm % Db) + d * Db * nn;
            if (off != 0.0) {
              const double2 xi = x[t];
              v.x = fma(off, xi.x, v.x);
              v.y = fma(off, xi.y, v.y);
            }
            Y[t] = v;
          });
  }
  __syncthreads();
}

// h[i] = V_i^H w for i < cnt, and h[cnt] = w^H w when self is set (one dot per wave, fixed
// order); V in global memory, w in LDS; h is shared, valid after the call.
__device__ void multi_dot(const double2* __restrict__ V, int64_t n, int cnt, bool self, const double2* w,
                          double2* h) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tot = cnt + (self ? 1 : 0);
  for (int i = wave; i < tot; i += kFW) {
    double2 acc{0.0, 0.0};
    if (i < cnt) {
      const double2* v = V + (int64_t)i * n;
#pragma unroll 4
      for (int64_t t = lane; t < n; t += 64) acc = cjfma(v[t], w[t], acc);
    } else {
#pragma unroll 4
      for (int64_t t = lane; t < n; t += 64) acc = cjfma(w[t], w[t], acc);
    }
    const double sx = wave_sum(acc.x), sy = wave_sum(acc.y);
    if (lane == 0) h[i] = double2{sx, sy};
  }
  __syncthreads();
}

// w -= sum_{i < cnt} V_i h[i]  (V global, w in LDS)
__device__ void gemv_sub(const double2* __restrict__ V, int64_t n, int cnt, const double2* h, double2* w) {
  for (int64_t t = threadIdx.x; t < n; t += kFT) {
    double2 acc{0.0, 0.0};
#pragma unroll 8
    for (int i = 0; i < cnt; ++i) acc = cfma(V[(int64_t)i * n + t], h[i], acc);
    const double2 wi = w[t];
    w[t] = double2{wi.x - acc.x, wi.y - acc.y};
  }
  __syncthreads();
}

// Number of eigenvalues of the tridiagonal (diagonal a, squared off-diagonal b2) below x
// (Sturm sequence of the LDL^T pivots).
__device__ int sturm(const double* a, const double* b2, int m, double x) {
  const double tiny = 1e-300;
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0) ++cnt;
  for (int i = 1; i < m; ++i) {
    if (fabs(q) < tiny) q = q < 0 ? -tiny : tiny;
    q = a[i] - x - b2[i - 1] * recip(q);
    if (q < 0) ++cnt;
  }
  return cnt;
}

// lane l's value of x (l uniform), and x with lane l's value replaced by v (v uniform)
__device__ __forceinline__ double rl(double x, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double wl(double v, int l, double x) { return (int)(threadIdx.x & 63) == l ? v : x; }

// tri_solve on one whole wave with row i in lane i (a_l, b_l, r_l: the diagonal, the
// off-diagonal and the right-hand side; m <= 64): the same elimination and back substitution,
// each row's operands taken by readlane and its pivot row left in lane i by a select.  Returns the
// solution, lane i holding y[i].
__device__ double tri_solve_wave(double a_l, double b_l, int m, double sigma, double r_l) {
  const double tiny = 1e-300;
  if (m == 1) {
    double p = rl(a_l, 0) - sigma;
    if (fabs(p) < tiny) p = tiny;
    return wl(rl(r_l, 0) / p, 0, r_l);
  }
  double dd_l = 0.0, u_l = 0.0, v_l = 0.0;
  double c0 = rl(a_l, 0) - sigma, c1 = rl(b_l, 0), c2 = 0.0, cr = rl(r_l, 0);
  for (int i = 0; i + 1 < m; ++i) {
    const double n0 = rl(b_l, i), n1 = rl(a_l, i + 1) - sigma, n2 = (i + 2 < m) ? rl(b_l, i + 1) : 0.0,
                 nr = rl(r_l, i + 1);
    double p0, p1, p2, pr, o0, o1, o2, orr;
    if (fabs(n0) > fabs(c0)) {
      p0 = n0, p1 = n1, p2 = n2, pr = nr;
      o0 = c0, o1 = c1, o2 = c2, orr = cr;
    } else {
      p0 = c0, p1 = c1, p2 = c2, pr = cr;
      o0 = n0, o1 = n1, o2 = n2, orr = nr;
    }
    if (fabs(p0) < tiny) p0 = tiny;
    const double ip = recip(p0);
    const double f = o0 * ip;
    dd_l = wl(ip, i, dd_l);
    u_l = wl(p1, i, u_l);
    v_l = wl(p2, i, v_l);
    r_l = wl(pr, i, r_l);
    c0 = o1 - f * p1;
    c1 = o2 - f * p2;
    c2 = 0.0;
    cr = orr - f * pr;
  }
  if (fabs(c0) < tiny) c0 = tiny;
  double x1 = cr * recip(c0);
  r_l = wl(x1, m - 1, r_l);
  double x2 = 0.0;
  for (int i = m - 2; i >= 0; --i) {
    const double xi = (rl(r_l, i) - rl(u_l, i) * x1 - rl(v_l, i) * x2) * rl(dd_l, i);
    r_l = wl(xi, i, r_l);
    x2 = x1;
    x1 = xi;
  }
  return r_l;
}

template <bool TL>
__global__ void __launch_bounds__(kFT) lanczos_fused_kernel(FusedArgs a) {
  __shared__ double2 pool[kPool];
  __shared__ double2 h1[kFM + 1], h2[kFM + 1];
  __shared__ double al[kFM], be[kFM], be2[kFM], s[kFM];
  __shared__ int cnt[kFT];
  __shared__ double lo_s, hi_s, inv_s, theta_s, res_s;
  __shared__ int stop_s, status_s;
  __shared__ double warm_hi_s;     // upper end of the last bracket of this cycle
  __shared__ int warm_s, geo_s;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int64_t n = (int64_t)a.d * a.Dk * a.Dkr;
  const int m = a.m;
  double2* x = a.x;
  // operands into LDS once for the whole eigensolve
  const int64_t nL = (int64_t)a.Db * a.wl * a.Dk, nR = (int64_t)a.Dbr * a.wr * a.Dkr,
                nW = (int64_t)a.d * a.wl * a.d * a.wr, n1 = (int64_t)a.Db * a.wl * a.d * a.Dkr;
  FusedOps o;
  double2* Ls = pool;
  double2* Rs = Ls + nL;
  double2* Ws = Rs + nR;
  double2* X = Ws + nW;  // current basis vector V_j
  double2* Y = X + n;    // new vector
  o.L = Ls;
  o.R = Rs;
  o.W = Ws;
  o.X = X;
  o.Y = Y;
  if constexpr (TL) {
    o.T1 = Y + n;
    o.T2 = o.T1 + n1;
  } else {
    o.T1 = a.T1;
    o.T2 = a.T2;
  }
  for (int64_t t = tid; t < nL; t += kFT) Ls[t] = a.L[t];
  for (int64_t t = tid; t < nR; t += kFT) Rs[t] = a.R[t];
  for (int64_t t = tid; t < nW; t += kFT) Ws[t] = a.W[t];
  int matvecs = 0, restarts = 0, converged = 0;
  if (tid == 0) status_s = 0;
  __syncthreads();
  for (int restart = 0; restart <= a.max_restarts; ++restart) {
    restarts = restart;
    // V_0 = x / |x|
    for (int64_t t = tid; t < n; t += kFT) Y[t] = x[t];
    __syncthreads();
    multi_dot(nullptr, n, 0, true, Y, h1);
    const double nx = h1[0].x;
    if (!(nx > 0.0) || !isfinite(nx)) {
      if (tid == 0) status_s = 2;  // the host reseeds
      break;
    }
    const double ix = 1.0 / sqrt(nx);
    for (int64_t t = tid; t < n; t += kFT) {
      const double2 v = double2{Y[t].x * ix, Y[t].y * ix};
      a.V[t] = v;
      X[t] = v;
    }
    if (tid == 0) {
      stop_s = 0;
      warm_s = 0;
    }
    __syncthreads();
    int K = 0;
    for (int j = 0; j < m; ++j) {
      fused_matvec(a, o);  // Y = H V_j
      ++matvecs;
      multi_dot(a.V, n, j + 1, false, Y, h1);
      gemv_sub(a.V, n, j + 1, h1, Y);
      multi_dot(a.V, n, j + 1, true, Y, h2);  // h2[j + 1] = |w|^2
      gemv_sub(a.V, n, j + 1, h2, Y);
      if (tid < 64) {
        // alpha_j, beta_j and the Gershgorin interval of T_j across one wave
        double corr = 0.0;
        for (int i = lane; i <= j; i += 64) corr = fma(h2[i].x, h2[i].x, fma(h2[i].y, h2[i].y, corr));
        corr = wave_sum(corr);
        const double alpha = h1[j].x + h2[j].x;
        const double beta = sqrt(fmax(h2[j + 1].x - corr, 0.0));
        if (lane == 0) {
          al[j] = alpha;
          be[j] = beta;
          be2[j] = beta * beta;
        }
        double lo = DBL_MAX, hi = -DBL_MAX;
        for (int i = lane; i <= j; i += 64) {
          const double ai = i == j ? alpha : al[i], bi = i == j ? beta : be[i];
          const double rad = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i < j ? fabs(bi) : 0.0);
          lo = fmin(lo, ai - rad);
          hi = fmax(hi, ai + rad);
        }
        lo = wave_min(lo);
        hi = wave_max(hi);
        if (lane == 0) {
          inv_s = beta > 1e-300 ? 1.0 / beta : 0.0;
          if (!isfinite(alpha) || !isfinite(beta)) status_s = 3;
          // Within a cycle T_j grows by rows, so by interlacing its lowest eigenvalue lies at or
          // below the one bracketed at the last check: that bracket's upper end caps the search.
          const bool g = warm_s && warm_hi_s > lo && warm_hi_s < hi;
          lo_s = lo;
          hi_s = g ? warm_hi_s : hi;
          geo_s = g;
        }
      }
      __syncthreads();
      if (status_s == 3) break;
      const double inv = inv_s;
      // V_{j+1}: to the basis in global memory and, as the next matvec's input, to X
      for (int64_t t = tid; t < n; t += kFT) {
        const double2 v = double2{Y[t].x * inv, Y[t].y * inv};
        a.V[(int64_t)(j + 1) * n + t] = v;
        X[t] = v;
      }
      __syncthreads();
      K = j + 1;
      // the tridiagonal problem is solved every third step and at the end of the basis
      if ((j + 1) % 3 != 0 && j + 1 < m && (int64_t)(j + 1) < n) continue;
      // lowest eigenvalue of T_j by multisection: kMS threads count at their own points
      // With a capped bracket the first round's points close in on the cap geometrically,
      // hi - (hi - lo) g^k with g^kMS = 1e-18, so an eigenvalue that moved little since the
      // last check is bracketed to a small fraction of that move in one round.
      for (int round = 0; round < 10; ++round) {
        const double lo = lo_s, hi = hi_s;
        if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi))) break;
        const bool geo = round == 0 && geo_s;
        const double step = (hi - lo) / (kMS + 1);
        auto pt = [&](int k) {
          return geo ? (k == 0 ? lo : hi - (hi - lo) * exp2(-kGeoLog2 * k)) : lo + step * k;
        };
        const double xk = pt(tid + 1);
        if (tid < kMS) cnt[tid] = (xk > lo && xk < hi) ? (sturm(al, be2, j + 1, xk) >= 1) : (xk >= hi ? 1 : 0);
        __syncthreads();
        // the first point whose count reaches one bounds the eigenvalue from above
        if (tid < kMS && cnt[tid] && (tid == 0 || !cnt[tid - 1])) {
          lo_s = pt(tid);
          hi_s = xk;
        }
        __syncthreads();
        if (tid == 0 && !cnt[kMS - 1]) lo_s = pt(kMS);  // above the last point
        __syncthreads();
      }
      if (tid == 0) {
        warm_hi_s = hi_s;
        warm_s = 1;
      }
      if (tid < 64) {
        // the eigenvector by inverse iteration on wave 0, row i of T_j in lane i
        const double theta = 0.5 * (lo_s + hi_s);
        const int mm = j + 1;
        const double a_l = lane < mm ? al[lane] : 0.0, b_l = lane < mm ? be[lane] : 0.0;
        double s_l = lane < mm ? 1.0 : 0.0;
        const double shift = theta - 4.0 * DBL_EPSILON * fmax(1.0, fabs(theta));
        for (int it = 0; it < 2; ++it) {
          s_l = tri_solve_wave(a_l, b_l, mm, shift, s_l);
          if (lane >= mm) s_l = 0.0;
          const double nrm = sqrt(wave_sum(s_l * s_l));
          if (!(nrm > 0.0) || !isfinite(nrm)) {
            s_l = lane == 0 ? 1.0 : 0.0;
            break;
          }
          s_l *= 1.0 / nrm;
        }
        if (lane < mm) s[lane] = s_l;
        const double res = rl(b_l, j) * fabs(rl(s_l, j));
        const double scale = fmax(1.0, fabs(theta));
        if (lane == 0) {
          theta_s = theta;
          res_s = res;
          stop_s = (res <= a.tol * scale || be[j] <= 1e-14 * scale || (int64_t)(j + 1) >= n) ? 1 : 0;
        }
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
      Y[t] = acc;
    }
    __syncthreads();
    multi_dot(nullptr, n, 0, true, Y, h1);
    const double ir = h1[0].x > 0.0 ? 1.0 / sqrt(h1[0].x) : 0.0;
    for (int64_t t = tid; t < n; t += kFT) x[t] = double2{Y[t].x * ir, Y[t].y * ir};
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
         h->precision == MPSD_PREC_FP64 && !h->herm && fused_lds(h).ops <= kPool;
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
  const FusedLds f = fused_lds(h);
  if (f.ops + f.inter <= kPool)
    hipLaunchKernelGGL(lanczos_fused_kernel<true>, dim3(1), dim3(kFT), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL(lanczos_fused_kernel<false>, dim3(1), dim3(kFT), 0, ctx->stream, a);
  hipError_t e = hipGetLastError();
  // the status block through pinned memory: an asynchronous copy, then one wait
  if (e == hipSuccess && !ctx->status_host) e = hipHostMalloc((void**)&ctx->status_host, 64 * sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(ctx->status_host, a.out, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess)
    for (int i = 0; i < 6; ++i) out6[i] = ctx->status_host[i];
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
