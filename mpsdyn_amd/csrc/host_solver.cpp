// Device matvec multiplier and the FP64 Lanczos eigensolver of the sweep.

#include <algorithm>
#include <cmath>
#include <limits>

#include "mpsdyn/mpsdyn.h"

namespace mpsdyn {

// ------------------------------------------------------------------------------ Multiplier
void Multiplier::product(const mwArray& x, mwArray& y) {
  mpsd_ctx* c = ctx();
  const int64_t n = getSize();
  if (x.getNrComponents() != n) throw Error(MPSD_ERR_SHAPE, "Multiplier::product: wrong input size");
  DevTensor dx(c, x), dy(c, std::vector<int>{(int)n});
  product_device(dx.ptr(), dy.ptr());
  y = dy.download();
  y.reshape({(int)n, 1});
}

TensorMultiplierOffset::TensorMultiplierOffset(mpsd_ctx* ctx, const mpsd_tensor* L, const mpsd_tensor* W,
                                               const mpsd_tensor* R, double offset, int precision)
    : ctx_(ctx) {
  check(ctx_, mpsd_heff_create(ctx_, L, W, R, offset, precision, &h_), "heff_create");
}

TensorMultiplierOffset::~TensorMultiplierOffset() {
  if (h_) mpsd_heff_destroy(h_);
}

TensorMultiplierHermitianOffset::TensorMultiplierHermitianOffset(mpsd_ctx* ctx, const mpsd_tensor* L,
                                                                 const mpsd_tensor* W, const mpsd_tensor* R,
                                                                 double offset, int precision)
    : TensorMultiplierOffset(ctx, L, W, R, offset, precision) {
  check(ctx, mpsd_heff_set_hermitian(handle(), 1), "heff_set_hermitian");
}

int64_t TensorMultiplierOffset::getSize() const { return mpsd_heff_size(h_); }

void TensorMultiplierOffset::product_device(const double* x, double* y) {
  const int64_t n = getSize();
  check(ctx_, mpsd_heff_apply(h_, x, y, 1, n, n, 1), "heff_apply");
}

void TensorMultiplierOffset::getFullTensor(mwArray& result) {
  const int n = (int)getSize();
  result = mwArray(std::vector<int>{n, n});
  check(ctx_, mpsd_heff_full(h_, (double*)result.data()), "heff_full");
}

void TensorMultiplierOffset::setPrecision(int p) { check(ctx_, mpsd_heff_set_precision(h_, p), "set_precision"); }

// ------------------------------------------------------------------------------ tridiagonal
// Number of eigenvalues of T below x (Sturm count via the LDL^T pivots).
static int sturm_count(const std::vector<double>& a, const std::vector<double>& b, double x) {
  const double tiny = std::numeric_limits<double>::min() * 1e10;
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0) ++cnt;
  for (size_t i = 1; i < a.size(); ++i) {
    if (std::fabs(q) < tiny) q = q < 0 ? -tiny : tiny;
    q = a[i] - x - b[i - 1] * b[i - 1] / q;
    if (q < 0) ++cnt;
  }
  return cnt;
}

// Solve (T - sigma I) y = r by Gaussian elimination with partial pivoting (tridiagonal).
static void tridiag_solve(const std::vector<double>& a, const std::vector<double>& b, double sigma,
                          std::vector<double>& r) {
  const int m = (int)a.size();
  if (m == 1) {
    double p = a[0] - sigma;
    if (std::fabs(p) < 1e-300) p = 1e-300;
    r[0] /= p;
    return;
  }
  // rows: lower l[i] (i>=1), diag d[i], upper u[i], second upper v[i] (from pivoting)
  std::vector<double> d(m), u(m, 0.0), v(m, 0.0), l(m, 0.0);
  for (int i = 0; i < m; ++i) d[i] = a[i] - sigma;
  for (int i = 0; i + 1 < m; ++i) {
    u[i] = b[i];
    l[i + 1] = b[i];
  }
  std::vector<double> rr = r;
  for (int i = 0; i + 1 < m; ++i) {
    if (std::fabs(l[i + 1]) > std::fabs(d[i])) {  // swap rows i and i+1
      std::swap(d[i], l[i + 1]);
      std::swap(u[i], d[i + 1]);
      std::swap(v[i], u[i + 1]);
      std::swap(rr[i], rr[i + 1]);
    }
    double piv = d[i];
    if (std::fabs(piv) < 1e-300) piv = d[i] = 1e-300;
    const double f = l[i + 1] / piv;
    d[i + 1] -= f * u[i];
    u[i + 1] -= f * v[i];
    rr[i + 1] -= f * rr[i];
  }
  if (std::fabs(d[m - 1]) < 1e-300) d[m - 1] = 1e-300;
  rr[m - 1] /= d[m - 1];
  rr[m - 2] = (rr[m - 2] - u[m - 2] * rr[m - 1]) / d[m - 2];
  for (int i = m - 3; i >= 0; --i) rr[i] = (rr[i] - u[i] * rr[i + 1] - v[i] * rr[i + 2]) / d[i];
  r = rr;
}

double tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, std::vector<double>& s) {
  const int m = (int)a.size();
  double lo = a[0], hi = a[0];
  for (int i = 0; i < m; ++i) {
    const double rad = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - rad);
    hi = std::max(hi, a[i] + rad);
  }
  const double eps = std::numeric_limits<double>::epsilon();
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (hi - lo <= 2 * eps * std::max(std::fabs(lo), std::fabs(hi)) || mid == lo || mid == hi) break;
    if (sturm_count(a, b, mid) >= 1)
      hi = mid;
    else
      lo = mid;
  }
  const double theta = 0.5 * (lo + hi);
  s.assign(m, 1.0);
  const double shift = theta - 4 * eps * std::max(1.0, std::fabs(theta));
  for (int it = 0; it < 3; ++it) {
    tridiag_solve(a, b, shift, s);
    double nrm = 0;
    for (double v : s) nrm += v * v;
    nrm = std::sqrt(nrm);
    if (!(nrm > 0) || !std::isfinite(nrm)) {
      s.assign(m, 0.0);
      s[0] = 1.0;
      break;
    }
    for (double& v : s) v /= nrm;
  }
  return theta;
}

// ------------------------------------------------------------------------------ Lanczos
// Steps per chunk between two host checks: short chunks first (late sweeps converge within a
// few steps), then 4.  Shared by both launch paths so they build the same basis.
static int lanczos_chunk(int j) { return j < 6 ? 3 : 4; }

// Largest basis (j + 1 vectors) whose Gram-Schmidt step runs fused.
static constexpr int kFusedGsCols = 32;

LanczosSolver::~LanczosSolver() {
  if (lz_) mpsd_lanczos_destroy(lz_);
}

EigStats LanczosSolver::dense(Multiplier& H, double* x) {
  mpsd_heff* h = H.heff();
  if (!h) throw Error(MPSD_ERR_UNSUPPORTED, "dense eigensolver: the multiplier has no device handle");
  EigStats st;
  check(ctx_, mpsd_heff_eigs_dense(h, x, &st.value), "heff_eigs_dense");
  st.residual = 0.0;
  st.converged = true;
  return st;
}

EigStats LanczosSolver::lowest_graph(Multiplier& H, mpsd_heff* h, double* x, double tol, int m, int max_restarts) {
  // Device-driven steps replayed as graphs; the start and Ritz vectors are normalised on the
  // device, so a site costs one host round trip per chunk of steps.
  EigStats st;
  const int64_t n = H.getSize();
  if (!lz_ || lz_m_ != m) {
    if (lz_) mpsd_lanczos_destroy(lz_);
    lz_ = nullptr;
    check(ctx_, mpsd_lanczos_create(ctx_, m, &lz_), "lanczos_create");
    lz_m_ = m;
  }
  std::vector<double> abh(2 * m + 4), s;
  const int64_t rep0 = mpsd_lanczos_replays(lz_);
  bool seeded = false;
  for (int restart = 0; restart <= max_restarts; ++restart) {
    st.restarts = restart;
    check(ctx_, mpsd_lanczos_begin(lz_, h, x), "lanczos_begin");
    int K = 0;
    bool stop = false;
    double theta = 0.0, res = 0.0;
    for (int j = 0; j < m && !stop;) {
      const int k = lanczos_chunk(j);
      if (j + k > m) break;  // the tail shorter than a chunk ends this restart
      int jend = 0;
      check(ctx_, mpsd_lanczos_steps(lz_, k, abh.data(), &jend), "lanczos_steps");
      st.matvecs += k;
      if (j == 0) {
        const double nx = abh[0];
        if (!(nx > 0.0) || !std::isfinite(nx)) {
          if (seeded) throw Error(MPSD_ERR_NUMERIC, "Lanczos: start vector is zero or not finite");
          DevTensor g(ctx_, complex_gaussian(12345, {(int)n}));  // reseed and start over
          check(ctx_, mpsd_vec_copy(ctx_, g.ptr(), x, n), "seed");
          seeded = true;
          break;
        }
      }
      for (int kk = j; kk < jend; ++kk) {
        const double a = abh[2 + 2 * kk], b = abh[3 + 2 * kk];
        if (!std::isfinite(a) || !std::isfinite(b))
          throw Error(MPSD_ERR_NUMERIC, "Lanczos: non-finite value (NaN/Inf in the operator or vector)");
        std::vector<double> al(kk + 1), be(kk);
        for (int i = 0; i <= kk; ++i) al[i] = abh[2 + 2 * i];
        for (int i = 0; i < kk; ++i) be[i] = abh[3 + 2 * i];
        theta = tridiag_lowest(al, be, s);
        res = b * std::fabs(s[kk]);
        const double scale = std::max(1.0, std::fabs(theta));
        if (res <= tol * scale || b <= 1e-14 * scale || kk + 1 >= n) {
          stop = true;
          K = kk + 1;
          break;
        }
      }
      j = jend;
      if (!stop) K = j;
    }
    if (K == 0) {
      if (seeded && restart < max_restarts) continue;
      throw Error(MPSD_ERR_NUMERIC, "Lanczos: empty basis");
    }
    check(ctx_, mpsd_lanczos_ritz(lz_, s.data(), K, x), "ritz");
    st.value = theta;
    st.residual = res;
    if (stop) {
      st.converged = true;
      break;
    }
  }
  st.graph_replays = mpsd_lanczos_replays(lz_) - rep0;
  return st;
}

EigStats LanczosSolver::lowest(Multiplier& H, double* x, double tol, int max_krylov, int max_restarts) {
  // Restarted Lanczos with two-pass full reorthogonalisation.  The iterations run on the
  // device in chunks (lanczos_chunk): Gram-Schmidt coefficients, alpha and beta stay in device
  // memory, and the host reads the chunk's alphas/betas once to solve the tridiagonal
  // problem.  The stopping step is the first converged one inside the chunk.  Up to 32 basis
  // vectors a step's two passes run fused (mpsd_lanczos_gs_step_dev: three reads of the basis,
  // three launches below n = 32768 and four above); larger bases take the two separate passes.
  // The graph-replayed form (lowest_graph) keeps the separate passes, so the two forms agree to
  // rounding.
  EigStats st;
  const int64_t n = H.getSize();
  const int m = (int)std::max<int64_t>(1, std::min<int64_t>(max_krylov, n));
  mpsd_heff* hh = H.heff();
  mpsd_heff* hs = H.smallHeff();
  if (hs && mpsd_ctx_get_option(ctx_, MPSD_OPT_FUSED_LANCZOS) == 1) {
    // small local problems: the whole eigensolve in one workgroup, one launch
    double o[6];
    const int rc = mpsd_heff_eigs_fused(hs, x, tol, m, max_restarts, o);
    if (rc == MPSD_OK && o[5] == 3.0)
      throw Error(MPSD_ERR_NUMERIC, "Lanczos: non-finite value (NaN/Inf in the operator or vector)");
    if (rc == MPSD_OK && o[5] == 0.0) {
      st.value = o[0];
      st.residual = o[1];
      st.matvecs = (int)o[2];
      st.restarts = (int)o[3];
      st.converged = o[4] != 0.0;
      return st;
    }
    if (rc != MPSD_OK && rc != MPSD_ERR_UNSUPPORTED) check(ctx_, rc, "eigs_fused");
    // status 2 (zero / non-finite start vector) and problems that do not fit continue below
  }
  if (hh && m >= 3 && mpsd_ctx_get_option(ctx_, MPSD_OPT_GRAPH) == 1) {
    // the graph path needs d = 2 and the FP64 / FP32 forms; anything else runs below
    try {
      return lowest_graph(H, hh, x, tol, m, max_restarts);
    } catch (const Error& e) {
      if (e.code != MPSD_ERR_UNSUPPORTED) throw;
    }
  }
  if ((int64_t)(m + 1) * n > cap_) {
    V_ = DevTensor(ctx_, std::vector<int>{(int)n, m + 1});
    cap_ = (int64_t)(m + 1) * n;
  }
  if (coef_cap_ < m) {
    coef_ = DevTensor(ctx_, std::vector<int>{4 * (m + 2)});
    coef_cap_ = m;
  }
  double* V = V_.ptr();
  const int64_t ldv = n;
  auto vec = [&](int j) { return V + 2 * (int64_t)j * ldv; };
  double* h1 = coef_.ptr();       // m + 2 complex
  double* h2 = h1 + 2 * (m + 2);  // m + 2 complex
  double* ab = h2 + 2 * (m + 2);  // |x|^2, 0, then alpha_j, beta_j
  std::vector<double> abh(2 * m + 4), s;
  bool seeded = false;
  // steps jj0 .. jj1 - 1 of the basis, then the coefficients so far into fetch slot `slot`
  auto run_chunk = [&](int jj0, int jj1, int slot) {
    for (int jj = jj0; jj < jj1; ++jj) {
      double* w = vec(jj + 1);
      H.product_device(vec(jj), w);
      ++st.matvecs;
      if (jj + 1 <= kFusedGsCols) {
        check(ctx_, mpsd_lanczos_gs_step_dev(ctx_, V, ldv, jj, w, n, h1, h2, ab + 2), "Gram-Schmidt step");
      } else {
        check(ctx_, mpsd_lanczos_gs1_dev(ctx_, V, ldv, jj + 1, w, n, h1), "Gram-Schmidt pass 1");
        check(ctx_, mpsd_lanczos_gs2_dev(ctx_, V, ldv, jj, h1, h2, ab + 2, w, n), "Gram-Schmidt pass 2 + coef");
      }
    }
    check(ctx_, mpsd_fetch_async(ctx_, ab, (int64_t)sizeof(double) * (2 + 2 * jj1), slot), "coef fetch");
  };
  for (int restart = 0; restart <= max_restarts; ++restart) {
    st.restarts = restart;
    check(ctx_, mpsd_vec_normalize_dev(ctx_, x, vec(0), n, ab), "normalise");
    int K = 0;  // size of the Krylov basis used for the Ritz vector
    bool stop = false;
    double theta = 0.0, res = 0.0;
    int slot = 0;
    bool queued = false;  // the chunk starting at j was queued ahead, its coefficients in `slot`
    for (int j = 0; j < m && !stop;) {
      const int k = lanczos_chunk(j);
      if (j + k > m && j > 0) break;  // the tail shorter than a chunk ends this restart
      const int jend = std::min(m, j + k);
      if (!queued) run_chunk(j, jend, slot);
      // Queue the next chunk before reading this one's coefficients, so the device works through
      // the host's round trip.  Not from the first chunk, where late sweeps converge: an unused
      // chunk costs its matvecs.  Its steps only extend the basis past what this chunk's test
      // can keep, so the basis, the Ritz vector and the value are the same either way.
      queued = false;
      if (j > 0 && jend < m) {
        const int k2 = lanczos_chunk(jend);
        if (jend + k2 <= m) {
          run_chunk(jend, jend + k2, slot ^ 1);
          queued = true;
        }
      }
      check(ctx_, mpsd_fetch_wait(ctx_, slot, abh.data(), (int64_t)sizeof(double) * (2 + 2 * jend)), "coef read");
      slot ^= 1;
      if (j == 0) {
        const double nx = abh[0];
        if (!(nx > 0.0) || !std::isfinite(nx)) {
          if (seeded) throw Error(MPSD_ERR_NUMERIC, "Lanczos: start vector is zero or not finite");
          DevTensor g(ctx_, complex_gaussian(12345, {(int)n}));  // reseed and start over
          check(ctx_, mpsd_vec_copy(ctx_, g.ptr(), x, n), "seed");
          seeded = true;
          break;
        }
      }
      for (int kk = j; kk < jend; ++kk) {
        const double a = abh[2 + 2 * kk], b = abh[3 + 2 * kk];
        if (!std::isfinite(a) || !std::isfinite(b))
          throw Error(MPSD_ERR_NUMERIC, "Lanczos: non-finite value (NaN/Inf in the operator or vector)");
        std::vector<double> al(kk + 1), be(kk);
        for (int i = 0; i <= kk; ++i) al[i] = abh[2 + 2 * i];
        for (int i = 0; i < kk; ++i) be[i] = abh[3 + 2 * i];
        theta = tridiag_lowest(al, be, s);
        res = b * std::fabs(s[kk]);
        const double scale = std::max(1.0, std::fabs(theta));
        if (res <= tol * scale || b <= 1e-14 * scale || kk + 1 >= n) {
          stop = true;
          K = kk + 1;
          break;
        }
      }
      j = jend;
      if (!stop) K = j;
    }
    if (K == 0) {
      if (seeded && restart < max_restarts) continue;
      throw Error(MPSD_ERR_NUMERIC, "Lanczos: empty basis");
    }
    // Ritz vector x = V s / |V s|
    std::vector<double> c(2 * K, 0.0);
    for (int i = 0; i < K; ++i) c[2 * i] = s[i];
    check(ctx_, mpsd_vec_zero(ctx_, x, n), "zero");
    check(ctx_, mpsd_vec_gemv(ctx_, V, ldv, K, c.data(), x, n), "ritz");
    check(ctx_, mpsd_vec_normalize_dev(ctx_, x, x, n, h1), "normalise");
    st.value = theta;
    st.residual = res;
    if (stop) {
      st.converged = true;
      break;
    }
  }
  return st;
}

}  // namespace mpsdyn
