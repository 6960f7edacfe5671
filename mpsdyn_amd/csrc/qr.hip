// Blocked Householder QR for the gauge steps (the role of wrapper::qr / lq, src/wrapper/qr.cpp:33-188,
// behind Site::gaugeR / gaugeL, src/core/cpp/Site.cpp:179-238).
//
// A (m x n, column-major) is factorised in panels of PB columns: A = Q R,
// Q = H_0 H_1 ... H_{k-1}, H_i = I - tau_i v_i v_i^H with LAPACK's zgeqrf / zlarfg conventions
// (v_i(i) = 1, real R diagonal beta = -sign(Re alpha) * ||(alpha, x)||).
//
// * One workgroup factorises a panel with the panel held in registers (a thread owns up to
//   RPT rows); each column costs two workgroup reductions.  It also forms the panel's
//   compact-WY factor T (zlarft, forward / columnwise) column by column from the Gram products
//   it reduces alongside, and writes V and -T.
// * The trailing columns take C <- (I - V T^H V^H) C = C + V ((-T)^H (V^H C)), and Q is
//   accumulated backwards, Q <- (I - V T V^H) Q = Q + V ((-T) (V^H Q)).  V^H C (16 rows, long K)
//   is a chunked product (one launch of chunk partials); a second launch sums the chunks in
//   order, applies the 16 x 16 factor and adds the rank-16 product to C, tile by tile.
//
// A panel step is a handful of launches where rocSOLVER's blocked QR issues several per
// column, which is what the small and medium gauge steps of a sweep are bound by.

#include "mpsd_internal.h"
#include "wave_xor.h"

namespace mpsd {
namespace {

constexpr int kQT = 256;  // threads of the small-panel kernel (one wave per SIMD: a large register budget)
constexpr int kPB = 16;   // panel width
constexpr int kQWmax = 8; // waves of the largest panel workgroup (512 threads)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// conj(a) * b
__device__ __forceinline__ double2 cjmul(double2 a, double2 b) { return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }

// Workgroup sum of one real value per thread (a squared norm), fixed order: out[0].x, with
// out[0].y = 0, visible to all threads on return.
template <int NW>
__device__ __forceinline__ void block_sum_real(double v, double2 (*red)[1], double2* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double x = wave_sum_x(v);
  if (lane == 0) red[wave][0] = double2{x, 0.0};
  __syncthreads();
  if (threadIdx.x == 0) {
    double sm = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) sm += red[w][0].x;
    out[0] = double2{sm, 0.0};
  }
  __syncthreads();
}

// Workgroup sum of kPB (16) complex values per thread, fixed order, into out[0..16) for all
// threads.  Within a wave the values are reduced and scattered at once over lane bits 5..2
// (8 + 4 + 2 + 1 exchanges instead of 16 x 4), then over bits 1..0; lane 4u ends with the
// wave's sum of value u.
template <int O>
__device__ __forceinline__ double2 shfl2(double2 v) { return xor_lane<O>(v); }
__device__ __forceinline__ double2 add2(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }

template <int NW>
__device__ __forceinline__ void block_sum16(const double2 (&v)[kPB], double2 (*red)[kPB], double2* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
  double2 w8[8], w4[4], w2[2];
#pragma unroll
  for (int u = 0; u < 8; ++u) w8[u] = add2(b5 ? v[u + 8] : v[u], shfl2<32>(b5 ? v[u] : v[u + 8]));
#pragma unroll
  for (int u = 0; u < 4; ++u) w4[u] = add2(b4 ? w8[u + 4] : w8[u], shfl2<16>(b4 ? w8[u] : w8[u + 4]));
#pragma unroll
  for (int u = 0; u < 2; ++u) w2[u] = add2(b3 ? w4[u + 2] : w4[u], shfl2<8>(b3 ? w4[u] : w4[u + 2]));
  double2 w1 = add2(b2 ? w2[1] : w2[0], shfl2<4>(b2 ? w2[0] : w2[1]));
  w1 = add2(w1, shfl2<2>(w1));
  w1 = add2(w1, shfl2<1>(w1));
  // this lane now holds value u = 8 b5 + 4 b4 + 2 b3 + b2
  if ((lane & 3) == 0) red[wave][lane >> 2] = w1;
  __syncthreads();
  if (threadIdx.x < kPB) {
    double2 sm{0.0, 0.0};
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) sm = add2(sm, red[ww][threadIdx.x]);
    out[threadIdx.x] = sm;
  }
  __syncthreads();
}

// Workgroup maximum of one non-negative value per thread; the result is returned to all threads.
template <int NW>
__device__ __forceinline__ double block_max(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_max_x(v);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) r = fmax(r, red[w]);
  __syncthreads();
  return r;
}

// The Householder vector of the panel column whose diagonal row is c, at row r (a: the stored value).
__device__ __forceinline__ double2 vval(int r, int c, double2 a) {
  return r > c ? a : (r == c ? double2{1.0, 0.0} : double2{0.0, 0.0});
}

// Factorise columns [j0, j0+pb) of A (rows j0..m-1), pb <= kPB, m - j0 <= RPT * NT.
// Writes R / v in place, V[:, j0+u] (ld = m, rows >= j0) and -T (kPB x kPB, row-major) to Tg.
template <int NT, int RPT>
__global__ void __launch_bounds__(NT) qr_panel_kernel(double2* __restrict__ A, int m, int j0, int pb,
                                                       double2* __restrict__ V, double2* __restrict__ Tg) {
  constexpr int NW = NT / 64;
  __shared__ double2 red[kQWmax][kPB];
  __shared__ double2 sums[kPB];
  __shared__ double2 Tn[kPB][kPB];  // -T, upper triangular
  __shared__ double2 tau_s[kPB];
  __shared__ double2 scal_s;
  const int tid = threadIdx.x;

  double2 a[RPT][kPB];
  int row[RPT];
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    row[i] = j0 + tid + i * NT;
#pragma unroll
    for (int u = 0; u < kPB; ++u)
      a[i][u] = (row[i] < m && u < pb) ? A[row[i] + (int64_t)m * (j0 + u)] : double2{0.0, 0.0};
  }

  // t is unrolled so every register index below is a compile-time constant
#pragma clang loop unroll(full)
  for (int t = 0; t < kPB; ++t) {
    if (t < pb) {
      const int c = j0 + t;
      // squared norm of the column below the diagonal
      double s1 = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; ++i)
        if (row[i] > c) s1 += a[i][t].x * a[i][t].x + a[i][t].y * a[i][t].y;
      block_sum_real<NW>(s1, reinterpret_cast<double2(*)[1]>(&red[0][0]), sums);
      // A column whose squares underflow or overflow is rescaled by its largest entry first (the
      // role of zlarfg's safmin loop and of dznrm2's scaling); the test reads shared memory, so
      // every thread takes the branch.
      double xs2 = sums[0].x, S = 1.0;
      if (xs2 < 1e-280 || !(xs2 < 1e280)) {
        double mx = 0.0;
#pragma unroll
        for (int i = 0; i < RPT; ++i)
          if (row[i] >= c) mx = fmax(mx, fmax(fabs(a[i][t].x), fabs(a[i][t].y)));
        S = block_max<NW>(mx, &red[0][0].x);
        double s2 = 0.0;
        if (S > 0.0) {
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (row[i] > c) {
              const double xr = a[i][t].x / S, xi = a[i][t].y / S;
              s2 += xr * xr + xi * xi;
            }
        }
        block_sum_real<NW>(s2, reinterpret_cast<double2(*)[1]>(&red[0][0]), sums);
        xs2 = sums[0].x;
        if (S == 0.0) S = 1.0;
      }
      // zlarfg on the thread that owns row c (c - j0 = t < kPB <= NT: thread t, slot 0)
      if (tid == t) {
        const double2 alpha = a[0][t];
        double2 tau{0.0, 0.0}, scal{0.0, 0.0};
        if (!(xs2 == 0.0 && alpha.y == 0.0)) {
          const double ar = alpha.x / S, ai = alpha.y / S;
          const double nrm = S * sqrt(ar * ar + ai * ai + xs2);
          const double beta = alpha.x >= 0.0 ? -nrm : nrm;
          tau = double2{(beta - alpha.x) / beta, -alpha.y / beta};
          // scal = 1 / (alpha - beta), scaled so |alpha - beta|^2 cannot underflow
          const double2 dd{alpha.x - beta, alpha.y};
          const double sd = fmax(fabs(dd.x), fabs(dd.y));
          const double2 ds{dd.x / sd, dd.y / sd};
          const double den = (ds.x * ds.x + ds.y * ds.y) * sd;
          scal = double2{ds.x / den, -ds.y / den};
          a[0][t] = double2{beta, 0.0};
        }
        tau_s[t] = tau;
        scal_s = scal;
      }
      __syncthreads();
      const double2 tau = tau_s[t], scal = scal_s;
      const bool reflect = !(tau.x == 0.0 && tau.y == 0.0);
#pragma unroll
      for (int i = 0; i < RPT; ++i)
        if (row[i] > c) a[i][t] = reflect ? cmul(a[i][t], scal) : double2{0.0, 0.0};
      // one reduction: w_u = v^H A[:, u] (u > t) and G[u][t] = V[:, u]^H v (u < t)
      double2 part[kPB];
#pragma unroll
      for (int u = 0; u < kPB; ++u) part[u] = double2{0.0, 0.0};
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const double2 vr = vval(row[i], j0 + t, a[i][t]);
#pragma unroll
        for (int u = 0; u < kPB; ++u) {
          if (u > t) {
            const double2 p = cjmul(vr, a[i][u]);
            part[u].x += p.x;
            part[u].y += p.y;
          } else if (u < t) {
            const double2 p = cjmul(vval(row[i], j0 + u, a[i][u]), vr);
            part[u].x += p.x;
            part[u].y += p.y;
          }
        }
      }
      block_sum16<NW>(part, red, sums);
      // column t of -T (zlarft, forward columnwise): Tn[t][t] = -tau_t and, with
      // G[j][t] = V[:, j]^H v = sums[j] (j < t), Tn[r][t] = -tau_t sum_{r <= j < t} Tn[r][j] G[j][t]
      if (tid < kPB) {
        const int r = tid;
        double2 v{0.0, 0.0};
        if (r == t) {
          v = double2{-tau.x, -tau.y};
        } else if (r < t) {
          double2 acc{0.0, 0.0};
          for (int j = r; j < t; ++j) {
            const double2 p = cmul(Tn[r][j], sums[j]);
            acc.x += p.x;
            acc.y += p.y;
          }
          const double2 q = cmul(tau, acc);
          v = double2{-q.x, -q.y};
        }
        Tn[r][t] = v;
      }
      // H^H on the remaining panel columns: a_u -= conj(tau) v w_u
      const double2 ctau{tau.x, -tau.y};
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        if (row[i] < c) continue;
        const double2 cv = cmul(ctau, vval(row[i], j0 + t, a[i][t]));
#pragma unroll
        for (int u = 0; u < kPB; ++u) {
          if (u > t && u < pb) {
            const double2 p = cmul(cv, sums[u]);
            a[i][u].x -= p.x;
            a[i][u].y -= p.y;
          }
        }
      }
      __syncthreads();  // sums / red are rewritten by the next column
    }
  }

  // columns t >= pb of -T are zero
  for (int e = tid; e < kPB * kPB; e += NT) {
    const int r = e / kPB, t = e % kPB;
    Tg[e] = t < pb ? Tn[r][t] : double2{0.0, 0.0};
  }
  // outputs: A (R / v) and V
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    if (row[i] >= m) continue;
#pragma unroll
    for (int u = 0; u < kPB; ++u) {
      if (u >= pb) continue;
      const int64_t col = (int64_t)m * (j0 + u);
      A[row[i] + col] = a[i][u];
      V[row[i] + col] = vval(row[i], j0 + u, a[i][u]);
    }
  }
}

// ---- whole QR of a small matrix in one workgroup ----------------------------------------------
// A (m x n, m <= 64, n <= 64) is factorised by one workgroup in the zlarfg / zung2r conventions
// of the panel path, Q (m x k) is formed, and Q and R are written out: one launch where
// the panel path takes a dozen for a gauge step of a small site.
//   Layout: lane u of every wave holds column u, wave w the rows [w RPT, (w + 1) RPT), in
// registers, nw = ceil(m / RPT) <= 16 waves; the matrix is mirrored in LDS (X, leading dimension
// m | 1).  A column step is one wave-uniform instruction stream, two barriers apart: every thread
// evaluates zlarfg from the column's partial squared norms and reads its rows x_r of column c
// from X as broadcasts; the reflector v = scal x below the diagonal (and 1 = scal (alpha - beta)
// on it, so the diagonal row takes x_c = alpha - beta and needs no branch) is never formed: a
// wave's partial of v^H a_u is conj(scal) sum conj(x_r) a_r, and the update is a_r -= x_r
// (scal g).  Every lane then writes its updated column and its partial squared norm below the
// next diagonal back to LDS.  Q is accumulated in the same k steps: the columns of I_m take
// every reflection with the matrix and end as Q_full^H, whose first k rows give Q.  They are
// never a reflector column, so they stay in registers: in the spare lanes n .. n + m - 1 when
// n + m <= 64 (at no extra cost per step), else in a second register slot of lane u with its
// own partial sums.  A is read whole (through LDS, coalesced) before anything is written, so Q
// may be A itself.
//   With lqd = d > 0 the kernel does the LQ of a site instead (Site::gaugeL): A is the site
// tensor [d, n, m / d] and the matrix factorised is B[s + d c, a] = conj(site[s, a, c])
// (m x n); Rg receives L = R^H (n x k) and Qg the new site [d, k, m / d] = Q^H reshaped.
constexpr int kSQW = 16;            // most waves (the partial sums below are 16-slot trees)
constexpr int kSQStage = 65 * 64;   // X: leading dimension m | 1

struct Reflector {
  double2 tau, scal, dd;  // dd = alpha - beta = 1 / scal
  double beta;
  bool reflect;
};

// 1 / x from the hardware reciprocal and two Newton steps (within an ulp or two of the quotient)
__device__ __forceinline__ double rcp2(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// sum of the kSQW slots of a column of partials, as a tree
template <typename T>
__device__ __forceinline__ T tree_sum(T (*v)[64], int u);
template <>
__device__ __forceinline__ double tree_sum<double>(double (*v)[64], int u) {
  double t[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) t[q] = v[q][u];
#pragma unroll
  for (int h = 8; h > 0; h >>= 1)
#pragma unroll
    for (int q = 0; q < h; ++q) t[q] += t[q + h];
  return t[0];
}
template <>
__device__ __forceinline__ double2 tree_sum<double2>(double2 (*v)[64], int u) {
  double2 t[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) t[q] = v[q][u];
#pragma unroll
  for (int h = 8; h > 0; h >>= 1)
#pragma unroll
    for (int q = 0; q < h; ++q) {
      t[q].x += t[q + h].x;
      t[q].y += t[q + h].y;
    }
  return t[0];
}

// zlarfg on alpha and the squared norm xs2 of the rest of the column.  Unless the column's squares
// underflow or overflow (scaled: xs2 is then the norm of the column over S, its largest entry),
// 1e-280 <= xs2 <= |alpha - beta|^2 < 1e281 and two plain reciprocals suffice; otherwise every step
// is scaled.
__device__ __forceinline__ Reflector zlarfg_of(double2 alpha, double xs2, double S, bool scaled) {
  Reflector f{{0.0, 0.0}, {0.0, 0.0}, {1.0, 0.0}, alpha.x, false};
  if (xs2 == 0.0 && alpha.y == 0.0) return f;
  f.reflect = true;
  if (!scaled) {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xs2);
    f.beta = alpha.x >= 0.0 ? -nrm : nrm;
    const double rb = rcp2(f.beta);
    f.tau = double2{(f.beta - alpha.x) * rb, -alpha.y * rb};
    f.dd = double2{alpha.x - f.beta, alpha.y};
    const double rd = rcp2(f.dd.x * f.dd.x + f.dd.y * f.dd.y);
    f.scal = double2{f.dd.x * rd, -f.dd.y * rd};
    return f;
  }
  const double ar = alpha.x / S, ai = alpha.y / S;
  const double nrm = S * sqrt(ar * ar + ai * ai + xs2);
  f.beta = alpha.x >= 0.0 ? -nrm : nrm;
  f.tau = double2{(f.beta - alpha.x) / f.beta, -alpha.y / f.beta};
  f.dd = double2{alpha.x - f.beta, alpha.y};
  const double sd = fmax(fabs(f.dd.x), fabs(f.dd.y));
  const double2 ds{f.dd.x / sd, f.dd.y / sd};
  const double den = (ds.x * ds.x + ds.y * ds.y) * sd;
  f.scal = double2{ds.x / den, -ds.y / den};
  return f;
}

// acc += conj(x) a and a -= x g, in fused multiply-adds
__device__ __forceinline__ void cj_fma(double2& acc, double2 x, double2 a) {
  acc.x = fma(x.x, a.x, acc.x);
  acc.x = fma(x.y, a.y, acc.x);
  acc.y = fma(x.x, a.y, acc.y);
  acc.y = fma(-x.y, a.x, acc.y);
}
__device__ __forceinline__ void sub_mul(double2& a, double2 x, double2 g) {
  a.x = fma(-x.x, g.x, a.x);
  a.x = fma(x.y, g.y, a.x);
  a.y = fma(-x.x, g.y, a.y);
  a.y = fma(-x.y, g.x, a.y);
}

template <int RPT, bool Q2>
__global__ void __launch_bounds__(64 * kSQW) qr_small_kernel(const double2* Ag, int m, int n, double2* Qg,
                                                             double2* __restrict__ Rg, int lqd, Absorb ab) {
  __shared__ double2 X[kSQStage];
  __shared__ double2 part[kSQW][64], partB[kSQW][64];
  __shared__ double nrm_s[kSQW][64];
  __shared__ double beta_s[64];
  const int tid = threadIdx.x, u = tid & 63, nt = blockDim.x, nw = nt >> 6;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), r0 = w * RPT;
  const int k = min(m, n), ld = m | 1;
  if (lqd > 0) {
    // site[s, a, c] (s + d (a + n c), read in order) -> B[s + d c, a] = conj
    for (int e = tid; e < m * n; e += nt) {
      const int s = e % lqd, a = (e / lqd) % n, c = e / (lqd * n);
      const double2 v = Ag[e];
      X[s + lqd * c + ld * a] = double2{v.x, -v.y};
    }
  } else {
    for (int e = tid; e < m * n; e += nt) X[e % m + ld * (e / m)] = Ag[e];
  }
  // slots of absent waves read as zeros
  for (int e = nw * 64 + tid; e < kSQW * 64; e += nt) {
    (&part[0][0])[e] = double2{0.0, 0.0};
    if constexpr (Q2) (&partB[0][0])[e] = double2{0.0, 0.0};
    (&nrm_s[0][0])[e] = 0.0;
  }
  __syncthreads();
  // With n + m <= 64 the lanes n .. n + m - 1 hold the columns of I_m and take every reflection
  // too, so they end as Q_full^H = H_{k-1}^H ... H_0^H and Q needs no backward pass.
  // Otherwise a second register slot per lane, b, holds column u of I_m (u < m) and takes the
  // same reflections, with its own partial sums.
  // (Q2 is set exactly when n + m > 64.)
  constexpr bool q2 = Q2, fq = !Q2;
  double2 a[RPT], xr[RPT], b[RPT];
#pragma unroll
  for (int i = 0; i < RPT; ++i) b[i] = double2{(q2 && r0 + i == u && u < m) ? 1.0 : 0.0, 0.0};
  {
    double s2 = 0.0;
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = r0 + i;
      a[i] = (u < n && r < m) ? X[r + ld * u] : double2{0.0, 0.0};
      if (r > 0) s2 = fma(a[i].x, a[i].x, fma(a[i].y, a[i].y, s2));
      if (fq && u >= n && r < m && r == u - n) a[i] = double2{1.0, 0.0};
    }
    nrm_s[w][u] = s2;
  }
  for (int c = 0; c < k; ++c) {
    __syncthreads();
    double xs2 = tree_sum<double>(nrm_s, c);
    const double2* xc = X + ld * c;
    const double2 alpha = xc[c];
    double S = 1.0;
    const bool scaled = xs2 < 1e-280 || !(xs2 < 1e280);
    if (scaled) {
      // squares that underflow or overflow: rescale by the column's largest entry (zlarfg's
      // safmin loop, dznrm2's scaling)
      double mx = 0.0;
      for (int r = c; r < m; ++r) mx = fmax(mx, fmax(fabs(xc[r].x), fabs(xc[r].y)));
      double s2 = 0.0;
      if (mx > 0.0)
        for (int r = c + 1; r < m; ++r) {
          const double p = xc[r].x / mx, q = xc[r].y / mx;
          s2 += p * p + q * q;
        }
      xs2 = s2;
      S = mx > 0.0 ? mx : 1.0;
    }
    const Reflector f = zlarfg_of(alpha, xs2, S, scaled);
    if (tid == 0) {
      beta_s[c] = f.beta;
    }
    // (H_c = I when !reflect: tau = scal = 0 and the step below only refreshes the partial norms)
    // v = scal x: for a rescaled column the step works on x / S and S scal, so that x^H a neither
    // overflows nor underflows (the same numbers as unscaled columns, where S = 1).
    const double2 sv = scaled ? double2{f.scal.x * S, f.scal.y * S} : f.scal;
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = r0 + i;
      const double2 x = r > c && r < m ? xc[r] : (r == c ? f.dd : double2{0.0, 0.0});
      xr[i] = scaled ? double2{x.x / S, x.y / S} : x;
    }
    // this wave's partial of v^H a_u
    const bool mine = (u > c && u < n) || (fq && u >= n && u < n + m);
    const bool mineB = q2 && u < m;
    {
      double2 acc0{0.0, 0.0}, acc1{0.0, 0.0};
#pragma unroll
      for (int i = 0; i < RPT; ++i) cj_fma(i & 1 ? acc1 : acc0, xr[i], a[i]);
      if (mine) part[w][u] = cjmul(sv, double2{acc0.x + acc1.x, acc0.y + acc1.y});
    }
    if constexpr (q2) {
      double2 acc0{0.0, 0.0}, acc1{0.0, 0.0};
#pragma unroll
      for (int i = 0; i < RPT; ++i) cj_fma(i & 1 ? acc1 : acc0, xr[i], b[i]);
      if (mineB) partB[w][u] = cjmul(sv, double2{acc0.x + acc1.x, acc0.y + acc1.y});
    }
    __syncthreads();
    if (mineB) {
      const double2 g2 = cmul(sv, cjmul(f.tau, tree_sum<double2>(partB, u)));
#pragma unroll
      for (int i = 0; i < RPT; ++i) sub_mul(b[i], xr[i], g2);
    }
    // H_c^H on the trailing columns: a_u -= conj(tau) v (v^H a_u), then back to LDS
    if (mine) {
      const double2 g2 = cmul(sv, cjmul(f.tau, tree_sum<double2>(part, u)));
      double s2 = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = r0 + i;
        sub_mul(a[i], xr[i], g2);
        if (u < n && r < m) X[r + ld * u] = a[i];
        if (r > c + 1) s2 = fma(a[i].x, a[i].x, fma(a[i].y, a[i].y, s2));
      }
      if (u < n) nrm_s[w][u] = s2;
    }
  }
  __syncthreads();
  // R (k x n): rows of X above the diagonal, beta on it; or L = R^H (n x k)
  for (int e = tid; e < k * n; e += nt) {
    const int i = lqd > 0 ? e / n : e % k, j = lqd > 0 ? e % n : e / k;
    const double2 r = i < j ? X[i + ld * j] : (i == j ? double2{beta_s[i], 0.0} : double2{0.0, 0.0});
    Rg[e] = lqd > 0 ? double2{r.x, -r.y} : r;
  }
  if (ab.out) {
    // the factor into the neighbour (Site::reduceDimR / reduceDimL) from R in LDS, with
    // R[i][j] = X[i + ld j] above the diagonal, beta_i on it and zero below:
    //   QR: next[s, i, c] = sum_{j >= i} R[i][j] old[s, j, c]        (old [nd, n, nD] -> [nd, k, nD])
    //   LQ: prev[s, a, j] = sum_{c >= j} old[s, a, c] conj(R[j][c])  (old [nd, nD, n] -> [nd, nD, k])
    const int tot = ab.nd * k * ab.nD;
    for (int e = tid; e < tot; e += nt) {
      const int s = e % ab.nd;
      double2 acc{0.0, 0.0};
      if (lqd == 0) {
        const int i = (e / ab.nd) % k, c = e / (ab.nd * k);
        const double2* o = ab.old + s + (int64_t)ab.nd * n * c;
        for (int j = i; j < n; ++j) {
          const double2 r = j == i ? double2{beta_s[i], 0.0} : X[i + ld * j];
          const double2 v = o[(int64_t)ab.nd * j];
          acc.x = fma(r.x, v.x, fma(-r.y, v.y, acc.x));
          acc.y = fma(r.x, v.y, fma(r.y, v.x, acc.y));
        }
      } else {
        const int a = (e / ab.nd) % ab.nD, j = e / (ab.nd * ab.nD);
        const double2* o = ab.old + s + (int64_t)ab.nd * a;
        for (int c = j; c < n; ++c) {
          const double2 r = c == j ? double2{beta_s[j], 0.0} : X[j + ld * c];
          const double2 v = o[(int64_t)ab.nd * ab.nD * c];
          acc.x = fma(v.x, r.x, fma(v.y, r.y, acc.x));
          acc.y = fma(v.y, r.x, fma(-v.x, r.y, acc.y));
        }
      }
      ab.out[e] = acc;
    }
  }
  {
    // Q[q, j] = conj(Q^H[j, q]): row j of lane n + q (fq) or of slot b of lane q, for j < k
    const int q = fq ? u - n : u;
    if (q >= 0 && q < m) {
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int j = r0 + i;
        if (j >= k) continue;
        const double2 h = fq ? a[i] : b[i];
        if (lqd > 0)
          Qg[q % lqd + lqd * (j + k * (q / lqd))] = h;  // new site[s, j, c] = conj(Q[s + d c, j])
        else
          Qg[q + m * j] = double2{h.x, -h.y};
      }
    }
  }
}

template <int RPT>
void launch_qr_small(hipStream_t s, const double2* A, int m, int n, double2* Q, double2* R, int lqd, Absorb ab) {
  const int nw = (m + RPT - 1) / RPT;
  if (n + m <= 64)
    hipLaunchKernelGGL((qr_small_kernel<RPT, false>), dim3(1), dim3(64 * nw), 0, s, A, m, n, Q, R, lqd, ab);
  else
    hipLaunchKernelGGL((qr_small_kernel<RPT, true>), dim3(1), dim3(64 * nw), 0, s, A, m, n, Q, R, lqd, ab);
}

// rows per wave: up to four waves of four rows for m <= 16 (fewer waves per SIMD beat shorter
// row loops there), else up to 16 waves of 2 or 4 rows
void launch_qr_small_any(hipStream_t s, const double2* A, int m, int n, double2* Q, double2* R, int lqd,
                         Absorb ab = Absorb{}) {
  if (m <= 16)
    launch_qr_small<4>(s, A, m, n, Q, R, lqd, ab);
  else if (m <= 32)
    launch_qr_small<2>(s, A, m, n, Q, R, lqd, ab);
  else
    launch_qr_small<4>(s, A, m, n, Q, R, lqd, ab);
}

// Q = I (m x k)
__global__ void __launch_bounds__(256) eye_kernel(double2* __restrict__ Q, int m, int k) {
  const int64_t total = (int64_t)m * k;
  for (int64_t t = blockIdx.x * 256LL + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int64_t i = t % m, j = t / m;
    Q[t] = double2{i == j ? 1.0 : 0.0, 0.0};
  }
}

// R (k x n) = upper trapezoid of the factorised A (ld m)
__global__ void __launch_bounds__(256) upper_kernel(const double2* __restrict__ A, int m, int k, int n,
                                                    double2* __restrict__ R) {
  const int64_t total = (int64_t)k * n;
  for (int64_t t = blockIdx.x * 256LL + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int64_t i = t % k, j = t / k;
    R[t] = i <= j ? A[i + (int64_t)m * j] : double2{0.0, 0.0};
  }
}

unsigned grid1(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, (n + 255) / 256)); }

// W = V^H C for a panel (pb <= 16 rows of W) over a long K: the rows are cut into chunks of
// kVC, each workgroup forms the 16 x 16 partial product of one chunk and one block of 16
// columns from LDS, and a second kernel sums the chunks in order (deterministic, no atomics).
constexpr int kVC = 64;
constexpr int kVJ = 16;

__global__ void __launch_bounds__(256) vh_partial_kernel(const double2* __restrict__ V, int ldv, int rows, int pb,
                                                         const double2* __restrict__ C, int ldc, int cols,
                                                         double2* __restrict__ P) {
  __shared__ double2 Vs[kVC][kPB + 1];
  __shared__ double2 Cs[kVC][kVJ + 1];
  const int tid = threadIdx.x;
  const int jb = blockIdx.x * kVJ, r0 = blockIdx.y * kVC;
  for (int e = tid; e < kVC * kPB; e += 256) {
    const int r = e % kVC, a = e / kVC;
    Vs[r][a] = (r0 + r < rows && a < pb) ? V[(r0 + r) + (int64_t)ldv * a] : double2{0.0, 0.0};
  }
  for (int e = tid; e < kVC * kVJ; e += 256) {
    const int r = e % kVC, j = e / kVC;
    Cs[r][j] = (r0 + r < rows && jb + j < cols) ? C[(r0 + r) + (int64_t)ldc * (jb + j)] : double2{0.0, 0.0};
  }
  __syncthreads();
  const int a = tid % kPB, j = tid / kPB;
  double2 acc{0.0, 0.0};
#pragma unroll 8
  for (int r = 0; r < kVC; ++r) {
    const double2 p = cjmul(Vs[r][a], Cs[r][j]);
    acc.x += p.x;
    acc.y += p.y;
  }
  if (a < pb && jb + j < cols) P[((int64_t)blockIdx.y * cols + jb + j) * kPB + a] = acc;
}

// C += V (op(Tn) (V^H C)) from the chunk partials of V^H C, in one launch after vh_partial_kernel
// (in place of vh_reduce_kernel and the rank-16 BRGEMM): workgroup (column block, 64-row tile)
// first forms its 16 x 16 block of W exactly as vh_reduce_kernel does (chunk sums in order, then
// the factor), then updates its 64 x 16 tile of C with the rank-pb product, summed over the
// panel's columns in order and added to C.  Every row tile repeats its column block's W (a few
// KB from L2).
constexpr int kVR = 64;  // rows of C per workgroup
__global__ void __launch_bounds__(256) vh_apply_kernel(const double2* __restrict__ V, int ldv, int rows, int pb,
                                                       double2* __restrict__ C, int ldc, int cols,
                                                       const double2* __restrict__ P, int nch,
                                                       const double2* __restrict__ Tg, int conjT) {
  __shared__ double2 Ws[kPB][kVJ + 1];
  __shared__ double2 Wf[kPB][kVJ + 1];
  __shared__ double2 Ts[kPB][kPB + 1];
  __shared__ double2 Vs[kVR][kPB + 1];
  const int tid = threadIdx.x, a = tid % kPB, jj = tid / kPB;
  const int jb = blockIdx.x * kVJ, j = jb + jj, r0 = blockIdx.y * kVR;
  Ts[tid / kPB][tid % kPB] = Tg[tid];
  double2 sm{0.0, 0.0};
  if (a < pb && j < cols)
    for (int ch = 0; ch < nch; ++ch) {
      const double2 v = P[((int64_t)ch * cols + j) * kPB + a];
      sm.x += v.x;
      sm.y += v.y;
    }
  Ws[a][jj] = sm;
  for (int e = tid; e < kVR * kPB; e += 256) {
    const int r = e % kVR, c = e / kVR;
    Vs[r][c] = (r0 + r < rows && c < pb) ? V[(r0 + r) + (int64_t)ldv * c] : double2{0.0, 0.0};
  }
  // this thread's elements of C, loaded now so their latency overlaps the factor product
  constexpr int kPer = kVR * kVJ / 256;
  double2 cv[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int e = tid + 256 * q, r = e % kVR, c = e / kVR;
    cv[q] = (r0 + r < rows && jb + c < cols) ? C[(r0 + r) + (int64_t)ldc * (jb + c)] : double2{0.0, 0.0};
  }
  __syncthreads();
  {
    double2 o{0.0, 0.0};
    if (a < pb && j < cols)
      for (int b = 0; b < pb; ++b) {
        const double2 p = conjT ? cjmul(Ts[b][a], Ws[b][jj]) : cmul(Ts[a][b], Ws[b][jj]);
        o.x += p.x;
        o.y += p.y;
      }
    Wf[a][jj] = o;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int e = tid + 256 * q, r = e % kVR, c = e / kVR;
    if (r0 + r >= rows || jb + c >= cols) continue;
    double2 t{0.0, 0.0};
    for (int b = 0; b < pb; ++b) {
      const double2 p = cmul(Vs[r][b], Wf[b][c]);
      t.x += p.x;
      t.y += p.y;
    }
    C[(r0 + r) + (int64_t)ldc * (jb + c)] = double2{cv[q].x + t.x, cv[q].y + t.y};
  }
}

// C += V op(Tn) V^H C for a panel: chunk partials, then vh_apply_kernel.  P: partials.
hipError_t block_reflect(const double2* V, int ldv, int rows, int pb, double2* C, int ldc, int cols,
                         const double2* Tn, int conjT, double2* P, hipStream_t s) {
  const int nch = (rows + kVC - 1) / kVC, ncb = (cols + kVJ - 1) / kVJ;
  hipLaunchKernelGGL(vh_partial_kernel, dim3(ncb, nch), dim3(256), 0, s, V, ldv, rows, pb, C, ldc, cols, P);
  hipLaunchKernelGGL(vh_apply_kernel, dim3(ncb, (rows + kVR - 1) / kVR), dim3(256), 0, s, V, ldv, rows, pb, C, ldc,
                     cols, (const double2*)P, nch, Tn, conjT);
  return hipGetLastError();
}

}  // namespace

int house_qr_max_rows() { return 1024; }

// m <= 64: beyond that (d = 2, D = 64) the panel path is as fast; the kernel's step is two barriers
// and a zlarfg whatever the size, so it pays where the panels' dozen launches dominate
bool house_qr_small(int m, int n) { return m <= 4 * kSQW && n <= 64; }

int house_lq_site_small(mpsd_ctx* ctx, const double2* site, int d, int Dl, int Dr, double2* site_out, double2* L,
                        Absorb ab) {
  if (d < 1 || Dl < 1 || Dr < 1 || !house_qr_small(d * Dr, Dl))
    return set_err(ctx, MPSD_ERR_ARG, "house_lq_site_small: unsupported size");
  launch_qr_small_any(ctx->stream, site, d * Dr, Dl, site_out, L, d, ab);
  return hip_err(ctx, hipGetLastError(), "house_lq_site_small");
}

int house_qr_small_absorb(mpsd_ctx* ctx, const double2* A, int m, int n, double2* Q, double2* R, Absorb ab) {
  if (!house_qr_small(m, n)) return set_err(ctx, MPSD_ERR_ARG, "house_qr_small_absorb: unsupported size");
  launch_qr_small_any(ctx->stream, A, m, n, Q, R, 0, ab);
  return hip_err(ctx, hipGetLastError(), "house_qr_small_absorb");
}

int house_qr(mpsd_ctx* ctx, double2* A, int m, int n, double2* Q, double2* R) {
  if (m < 1 || n < 1 || m > house_qr_max_rows()) return set_err(ctx, MPSD_ERR_ARG, "house_qr: unsupported size");
  if (house_qr_small(m, n)) {
    launch_qr_small_any(ctx->stream, A, m, n, Q, R, 0);
    return hip_err(ctx, hipGetLastError(), "house_qr small");
  }
  if (Q == A) return set_err(ctx, MPSD_ERR_ARG, "house_qr: Q may alias A only for small matrices");
  const int k = std::min(m, n);
  const int64_t mk = (int64_t)m * k;
  const int64_t wn = std::max(n, k), nch = (m + kVC - 1) / kVC, npan = (k + kPB - 1) / kPB;
  const size_t need = sizeof(double2) * (size_t)(mk + npan * kPB * kPB + kPB * wn * nch);
  int rc = ensure_scratch(ctx, 6, need);
  if (rc) return rc;
  double2* V = (double2*)ctx->scratch[6];
  double2* T = V + mk;                  // -T of panel p at T + p kPB^2
  double2* P = T + npan * kPB * kPB;    // chunk partials of V^H C
  hipStream_t s = ctx->stream;
  hipError_t e = hipSuccess;
  for (int j0 = 0; j0 < k && e == hipSuccess; j0 += kPB) {
    const int pb = std::min(kPB, k - j0), rows = m - j0;
    // up to 256 rows: one wave per SIMD; beyond, 512 threads (two waves per SIMD) with one or
    // two rows per thread
    double2* Tp = T + (int64_t)(j0 / kPB) * kPB * kPB;
    const double2* Vp = V + j0 + (int64_t)m * j0;
    if (rows <= kQT)
      hipLaunchKernelGGL((qr_panel_kernel<kQT, 1>), dim3(1), dim3(kQT), 0, s, A, m, j0, pb, V, Tp);
    else if (rows <= 512)
      hipLaunchKernelGGL((qr_panel_kernel<512, 1>), dim3(1), dim3(512), 0, s, A, m, j0, pb, V, Tp);
    else
      hipLaunchKernelGGL((qr_panel_kernel<512, 2>), dim3(1), dim3(512), 0, s, A, m, j0, pb, V, Tp);
    e = hipGetLastError();
    const int cols = n - j0 - pb;
    if (e == hipSuccess && cols > 0) {
      double2* C = A + j0 + (int64_t)m * (j0 + pb);
      e = block_reflect(Vp, m, rows, pb, C, m, cols, Tp, 1, P, s);
    }
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(upper_kernel, dim3(grid1((int64_t)k * n)), dim3(256), 0, s, A, m, k, n, R);
    hipLaunchKernelGGL(eye_kernel, dim3(grid1(mk)), dim3(256), 0, s, Q, m, k);
    e = hipGetLastError();
  }
  // Q = H_0 ... H_{k-1} I, accumulated from the last panel back
  const int last = ((k - 1) / kPB) * kPB;
  for (int j0 = last; j0 >= 0 && e == hipSuccess; j0 -= kPB) {
    const int pb = std::min(kPB, k - j0), rows = m - j0, cols = k - j0;
    double2* Qs = Q + j0 + (int64_t)m * j0;
    const double2* Vp = V + j0 + (int64_t)m * j0;
    const double2* Tp = T + (int64_t)(j0 / kPB) * kPB * kPB;
    e = block_reflect(Vp, m, rows, pb, Qs, m, cols, Tp, 0, P, s);
  }
  return hip_err(ctx, e, "house_qr");
}

}  // namespace mpsd
