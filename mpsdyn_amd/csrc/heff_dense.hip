// Hermitian form of the matvec, the dense effective Hamiltonian and its dense lowest
// eigenpair.
//
// Hermitian form (TensorMultiplierHermitian(Offset), src/wrapper/multipliers/
// TensorMultiplierHermitian.cpp:8-66, TensorMultiplierHermitianOffset.cpp:13-16): the
// handle keeps conjugate-transposed copies L'[a,w,b] = conj(L[b,w,a]),
// W'[s,wl,s',wr] = conj(W[s',wl,s,wr]), R'[a,w,b] = conj(R[b,w,a]) and applies
// y = (H x + H' x) / 2 + offset x, H' being the same three-stage product over the copies.
//
// Dense form (TensorMultiplier(Offset)::getFullTensor, TensorMultiplier.cpp:102-117,
// TensorMultiplierOffset.cpp:20-24, TensorMultiplierHermitian(Offset).h) and the dense
// eigensolver the reference switches to for local problems below LOWEST_PRIMME and for
// the fulleig solver (Contractor.cpp:3087-3100, 3126-3133): the matrix is built on the
// device, averaged with its adjoint and diagonalised by rocSOLVER's zheevd.

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>

#include "mpsd_internal.h"

namespace mpsd {
namespace {

constexpr int kT = 256;

// out[i2, i1, i0, i3] = conj(in[i0, i1, i2, i3]) for a column-major [n0, n1, n2, n3] input
__global__ void __launch_bounds__(kT) conj_swap02_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                         int n0, int n1, int n2, int64_t total) {
  for (int64_t idx = blockIdx.x * (int64_t)kT + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * kT) {
    int64_t t = idx;
    const int64_t i0 = t % n0;
    t /= n0;
    const int64_t i1 = t % n1;
    t /= n1;
    const int64_t i2 = t % n2;
    const int64_t i3 = t / n2;
    const double2 v = in[idx];
    out[i2 + (int64_t)n2 * (i1 + (int64_t)n1 * (i0 + (int64_t)n0 * i3))] = double2{v.x, -v.y};
  }
}

// A <- (A + A^H) / 2 + offset I for a column-major n x n matrix; thread (i, j), i <= j,
// owns the pair A[i,j], A[j,i] (TensorMultiplierHermitian.h: .5*(result+Hconjugate(result))).
__global__ void __launch_bounds__(kT) herm_average_kernel(double2* __restrict__ A, int n, double offset) {
  const int64_t total = (int64_t)n * n;
  for (int64_t idx = blockIdx.x * (int64_t)kT + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * kT) {
    const int64_t i = idx % n, j = idx / n;
    if (i > j) continue;
    const double2 a = A[i + (int64_t)n * j];
    const double2 b = A[j + (int64_t)n * i];
    double2 s{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
    if (i == j) {
      s.y = 0.0;
      s.x += offset;
      A[i + (int64_t)n * j] = s;
    } else {
      A[i + (int64_t)n * j] = s;
      A[j + (int64_t)n * i] = double2{s.x, -s.y};
    }
  }
}

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (n + kT - 1) / kT)); }

int conj_swap(mpsd_ctx* ctx, const mpsd_tensor* in, int n3, mpsd_tensor** out) {
  const int n0 = in->dims[0], n1 = in->rank > 1 ? in->dims[1] : 1, n2 = in->rank > 2 ? in->dims[2] : 1;
  int dims[4] = {n2, n1, n0, n3};
  int rc = mpsd_tensor_create(ctx, n3 > 1 || in->rank == 4 ? 4 : 3, dims, out);
  if (rc) return rc;
  const int64_t total = (int64_t)n0 * n1 * n2 * n3;
  hipLaunchKernelGGL(conj_swap02_kernel, dim3(grid_of(total)), dim3(kT), 0, ctx->stream, in->ptr, (*out)->ptr, n0,
                     n1, n2, total);
  return hip_err(ctx, hipGetLastError(), "conj_swap");
}

}  // namespace

hipError_t heff_full_dev(mpsd_heff* h, double2* out) {
  const double offset = h->herm ? 0.0 : h->offset;
  hipError_t e = launch_heff_full_f64(h->L->ptr, h->W->ptr, h->R->ptr, out, h->d, h->Db, h->wl, h->Dk, h->Dbr,
                                      h->wr, h->Dkr, offset, h->ctx->stream);
  if (e != hipSuccess || !h->herm) return e;
  const int n = (int)((int64_t)h->d * h->Dk * h->Dkr);
  hipLaunchKernelGGL(herm_average_kernel, dim3(grid_of((int64_t)n * n)), dim3(kT), 0, h->ctx->stream, out, n,
                     h->offset);
  return hipGetLastError();
}

void heff_drop_hermitian(mpsd_heff* h) {
  if (h->hh) (void)mpsd_heff_destroy(h->hh);
  if (h->Lh) (void)mpsd_tensor_destroy(h->Lh);
  if (h->Wh) (void)mpsd_tensor_destroy(h->Wh);
  if (h->Rh) (void)mpsd_tensor_destroy(h->Rh);
  if (h->htmp) (void)hipFree(h->htmp);
  h->hh = nullptr;
  h->Lh = h->Wh = h->Rh = nullptr;
  h->htmp = nullptr;
  h->herm = 0;
}

}  // namespace mpsd

using namespace mpsd;

extern "C" {

int mpsd_heff_set_hermitian(mpsd_heff* h, int on) {
  if (!h) return MPSD_ERR_ARG;
  mpsd_ctx* ctx = h->ctx;
  heff_drop_hermitian(h);  // off, or a fresh copy of the current operands
  if (!on) return MPSD_OK;
  // TensorMultiplierHermitian CHECKDIMS (TensorMultiplierHermitian.cpp:14-20)
  if (h->Db != h->Dk || h->Dbr != h->Dkr || h->W->dims[2] != h->d)
    return set_err(ctx, MPSD_ERR_SHAPE, "set_hermitian: Heff is not square (bra and ket bonds / physical dims differ)");
  int rc = conj_swap(ctx, h->L, 1, &h->Lh);
  if (!rc) rc = conj_swap(ctx, h->W, h->wr, &h->Wh);
  if (!rc) rc = conj_swap(ctx, h->R, 1, &h->Rh);
  if (!rc) rc = mpsd_heff_create(ctx, h->Lh, h->Wh, h->Rh, 0.0, h->precision, &h->hh);
  if (!rc) {
    const int64_t n = (int64_t)h->d * h->Db * h->Dbr;
    hipError_t e = hipMalloc((void**)&h->htmp, sizeof(double2) * (size_t)std::max<int64_t>(n, 1));
    if (e != hipSuccess) rc = hip_err(ctx, e, "set_hermitian alloc");
  }
  if (rc) {
    heff_drop_hermitian(h);
    return rc;
  }
  h->herm = 1;
  return MPSD_OK;
}

int mpsd_heff_eigs_dense(mpsd_heff* h, double* x, double* value) {
  if (!h || !x || !value) return MPSD_ERR_ARG;
  mpsd_ctx* ctx = h->ctx;
  const int64_t n = mpsd_heff_size(h);
  if ((int64_t)h->d * h->Db * h->Dbr != n) return set_err(ctx, MPSD_ERR_SHAPE, "eigs_dense: Heff is not square");
  if (n > 16384) return set_err(ctx, MPSD_ERR_ARG, "eigs_dense: local dimension above 16384");
  void* bhv = nullptr;
  int rc = blas_handle(ctx, &bhv);
  rocblas_handle bh = (rocblas_handle)bhv;
  if (!rc) rc = ensure_scratch(ctx, 0, sizeof(double2) * (size_t)(n * n));
  if (!rc) rc = ensure_scratch(ctx, 1, sizeof(double) * (size_t)(2 * n + 2));
  if (rc) return rc;
  double2* A = (double2*)ctx->scratch[0];
  double* W = (double*)ctx->scratch[1];
  double* E = W + n;
  int* info = (int*)(E + n);
  hipError_t e = heff_full_dev(h, A);
  if (e != hipSuccess) return hip_err(ctx, e, "eigs_dense full");
  // the small-size switch averages with the adjoint (Contractor.cpp:3129); the Hermitian
  // form already is averaged, and averaging again leaves it unchanged
  hipLaunchKernelGGL(herm_average_kernel, dim3(grid_of(n * n)), dim3(kT), 0, ctx->stream, A, (int)n, 0.0);
  rocblas_status st = rocsolver_zheevd(bh, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n,
                                       (rocblas_double_complex*)A, (rocblas_int)n, W, E, info);
  if (st != rocblas_status_success)
    return set_err(ctx, MPSD_ERR_HIP, std::string("zheevd: ") + rocblas_status_to_string(st));
  int hinfo = 0;
  e = hipMemcpyAsync(x, A, sizeof(double2) * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream);  // lowest: column 0
  if (e == hipSuccess) e = hipMemcpyAsync(value, W, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_err(ctx, e, "eigs_dense");
  if (hinfo != 0) return set_err(ctx, MPSD_ERR_NUMERIC, "zheevd did not converge (info " + std::to_string(hinfo) + ")");
  return MPSD_OK;
}

}  // extern "C"
