/*
 * mpsd.h -- C-ABI of the MI355X device library for the DMRG effective-Hamiltonian hot path.
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types cross this boundary.
 * All tensors are complex128 stored as interleaved (re, im) double pairs in COLUMN-MAJOR
 * order (first index fastest), bit-compatible with the reference's mwArray buffers
 * (src/wrapper/mwArray.h:29-31, src/wrapper/mwArray.cpp:452-466):
 *   site tensor    A[d, Dl, Dr]                 (src/core/headers/Site.h:49)
 *   MPO tensor     W[d_up, wl, d_down, wr]      (src/core/headers/Operator.h:37-43)
 *   environment    E[D_bra, w, D_ket]           (TmpManager operL/operR, TmpManager.h:39-52)
 *
 * Every function returns MPSD_OK (0) or a negative MPSD_ERR_* code; the message of the
 * last failure on a context is available from mpsd_last_error().  The reference exits the
 * process on the same conditions (exit(212) for shape errors, Operator.cpp:283-289).
 *
 * Threading: a context owns one HIP stream and its scratch memory and is not re-entrant.
 * Concurrent solves use one context each; device tensors may be shared between contexts
 * of the same device when the caller orders the work (mpsd_ctx_sync).
 */
#ifndef MPSD_H
#define MPSD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPSD_OK 0
#define MPSD_ERR_ARG (-1)      /* null pointer, bad rank or size */
#define MPSD_ERR_SHAPE (-2)    /* operand dimensions do not chain */
#define MPSD_ERR_HIP (-3)      /* HIP runtime or device failure */
#define MPSD_ERR_NOMEM (-4)    /* device allocation failed */
#define MPSD_ERR_NUMERIC (-5)  /* NaN/Inf or breakdown detected */
#define MPSD_ERR_UNSUPPORTED (-6)

/* Arithmetic of a matvec.  FP64 is the drop-in; the others run the two GEMM stages in
 * lower precision with FP32 accumulation (the eigensolver around them stays FP64). */
#define MPSD_PREC_FP64 0
#define MPSD_PREC_FP32 1    /* complex FP32 MFMA, FP32 accumulate */
#define MPSD_PREC_BF16X3 2  /* split-BF16 (hi/lo, 3 products) MFMA, FP32 accumulate */

typedef struct mpsd_ctx mpsd_ctx;
typedef struct mpsd_tensor mpsd_tensor;
typedef struct mpsd_heff mpsd_heff;

/* ---- context --------------------------------------------------------------------------
 * Replaces the process-wide state of the reference (the Contractor singleton,
 * src/core/cpp/Contractor.cpp:49-53) with an explicit per-device, per-stream handle. */
int mpsd_ctx_create(int device, mpsd_ctx** out);
int mpsd_ctx_destroy(mpsd_ctx* ctx);
int mpsd_ctx_sync(mpsd_ctx* ctx);
const char* mpsd_last_error(const mpsd_ctx* ctx);
/* The HIP stream (hipStream_t) the context launches on, as an opaque pointer. */
void* mpsd_ctx_stream(mpsd_ctx* ctx);
/* Device memory currently held by the context's tensors and scratch, in bytes. */
int64_t mpsd_ctx_device_bytes(const mpsd_ctx* ctx);

/* ---- timing ---------------------------------------------------------------------------
 * HIP events around every kernel launch of the context, accumulated per kernel kind.  The
 * reference's equivalent is the -DTESTINGTIMES gettimeofday instrumentation of
 * TensorMultiplier::product (src/wrapper/multipliers/TensorMultiplier.cpp:57-99).
 * set_timing(on) also resets the totals; timing() synchronises the stream first. */
#define MPSD_KERNEL_STAGE1 0      /* environment x site GEMM (first stage, both sides) */
#define MPSD_KERNEL_MIX 1         /* MPO contraction of the middle stage */
#define MPSD_KERNEL_HEFF_STAGE3 2 /* last stage of the matvec */
#define MPSD_KERNEL_ENV_STAGE3 3  /* last stage of an environment update */
#define MPSD_KERNEL_PACK 4        /* packing of matvec operands (environments once per handle, x per call) */
#define MPSD_KERNEL_KINDS 5
int mpsd_ctx_set_timing(mpsd_ctx* ctx, int on);
int mpsd_ctx_timing(mpsd_ctx* ctx, int kind, double* total_ms, int64_t* launches);
/* Restrict timing to the kinds whose bit (1 << MPSD_KERNEL_*) is set in mask (default: all).
 * Each timed launch adds two event records to the stream, so timing only the kind being
 * measured keeps the others' schedule as it is untimed. */
int mpsd_ctx_set_timing_kinds(mpsd_ctx* ctx, unsigned mask);
/* Force the BRGEMM tile (tm, tn in {32, 64}, or 16, 16: one wave per workgroup); 0, 0 restores
 * the built-in choice (32 x 32 for FP64).  Every tile gives the same bits. */
int mpsd_ctx_set_tile(mpsd_ctx* ctx, int tm, int tn);
/* Other launch options of the context (used by the tile tuner). */
#define MPSD_OPT_TILE_K 1       /* K depth of a BRGEMM LDS stage: 16 (default) or 32 */
#define MPSD_OPT_SPLIT_STAGE3 2 /* matvec last stage split over the MPO bond: -1 auto, 0 off, 1 on */
#define MPSD_OPT_GRAPH 3        /* eigensolver steps as replayed hipGraphs: 0 (default: stream-launched
                                   steps, 4-14% faster per step for D = 32..512) or 1 */
#define MPSD_OPT_GAUGE_QR 5     /* gauge-step QR / LQ: 0 library Householder kernels (default), 1 rocSOLVER */
#define MPSD_OPT_FUSED_LANCZOS 6 /* small local problems on the one-workgroup eigensolver: 1 (default) or 0 */
#define MPSD_OPT_ENV_CHAIN 7    /* small environment updates as one single-workgroup launch: 1 (default)
                                   or 0 (three launches); the results are bit-identical */
#define MPSD_OPT_BATCH_GROUP 4  /* batches per BRGEMM workgroup: 0 built-in (default), 1, or 2 when an
                                   operand is shared by the batches; results are the same for all */
#define MPSD_OPT_KRYLOV 8       /* Krylov basis size of the ground-state sweep's local Lanczos solves:
                                   0 (default, 30) or 4..62 (more restarts, less Gram-Schmidt work) */
#define MPSD_OPT_LP_PACKED 9    /* precision forms: 1 (default) L and R packed once per handle into the
                                   form's arithmetic, x and the middle stage packed per matvec; 0 FP64
                                   operands converted while each tile is staged */
#define MPSD_OPT_LP_TILE 10     /* packed-operand BRGEMM tile: 0 automatic (default), 1 32x32, 2 64x32,
                                   3 32x64, 4 64x64 */
#define MPSD_OPT_PACKED_FP64 11 /* FP64 matvec: 1 (default) L and R packed once per handle, x per call,
                                   where the last stage is split over the MPO bond (same bits as 0);
                                   0 operands read in place */
#define MPSD_OPT_PACKED_FP64_TILE 12 /* its BRGEMM tile, as MPSD_OPT_LP_TILE */
#define MPSD_OPT_PACKED_FP64_X 13 /* its first stage on x: -1 automatic (default), 0 x read in place,
                                     1 x packed per call (same bits either way) */
int mpsd_ctx_set_option(mpsd_ctx* ctx, int option, int value);
int mpsd_ctx_get_option(const mpsd_ctx* ctx, int option);

/* ---- device tensors -------------------------------------------------------------------
 * HBM-resident counterpart of an mwArray (src/wrapper/mwArray.h:23-498): rank <= 8,
 * zero-initialised.  upload/download move the whole buffer (mwArray::getComponents()
 * layout, n complex numbers = 2n doubles). */
int mpsd_tensor_create(mpsd_ctx* ctx, int rank, const int* dims, mpsd_tensor** out);
int mpsd_tensor_destroy(mpsd_tensor* t);
int mpsd_tensor_rank(const mpsd_tensor* t);
int mpsd_tensor_dims(const mpsd_tensor* t, int* dims);
int64_t mpsd_tensor_size(const mpsd_tensor* t); /* number of complex elements */
/* Change the logical shape; reallocates only if the new size exceeds the capacity.
 * Replaces mwArray::reshape/resize of TmpManager slots (TmpManager.cpp:111-211). */
int mpsd_tensor_reshape(mpsd_tensor* t, int rank, const int* dims);
/* Change the dimensions keeping the contents by index: the element at each multi-index inside
 * both the old and the new dimensions is kept, new ones are zero, the rest are cut.  Same rank
 * (MPSD_ERR_SHAPE otherwise).  Replaces mwArray::resize with zero noise
 * (src/wrapper/mwArray.cpp:1604-1670), behind Site::changeDimensions (src/core/cpp/Site.cpp:
 * 525-559) as MPS::increaseBondDimension and MPS::adjustPhysDimensions use it.  The device
 * address changes (the copy goes to a new buffer). */
int mpsd_tensor_resize(mpsd_tensor* t, int rank, const int* dims);
/* Host data of the whole tensor to the device, stream-ordered: the host buffer may be reused
 * as soon as the call returns (the data goes through a pinned staging buffer of the context;
 * tensors larger than it are copied synchronously). */
int mpsd_tensor_upload(mpsd_tensor* t, const double* host, int64_t n_complex);
/* mpsd_tensor_create followed by mpsd_tensor_upload of n = prod(dims) complex numbers, without
 * the zero fill. */
int mpsd_tensor_create_from(mpsd_ctx* ctx, int rank, const int* dims, const double* host, mpsd_tensor** out);
/* mpsd_tensor_create without the zero fill: the contents are undefined until written (for
 * slots that a contraction or copy overwrites, like the environment cache's). */
int mpsd_tensor_create_empty(mpsd_ctx* ctx, int rank, const int* dims, mpsd_tensor** out);
int mpsd_tensor_download(const mpsd_tensor* t, double* host, int64_t n_complex);
int mpsd_tensor_copy(mpsd_tensor* dst, const mpsd_tensor* src); /* same size, stream-ordered */
/* Device address of the buffer (complex128 interleaved), for callers that keep vectors
 * on the device (mpsd_heff_apply with on_device = 1).  Valid until a call that changes the
 * tensor's shape (reshape past its capacity, gauge and reduce-dim steps). */
void* mpsd_tensor_devptr(mpsd_tensor* t);

/* ---- environment updates ---------------------------------------------------------------
 * out[Dr_bra, wr, Dr_ket] = sum conj(bra[s', b, c]) envL[b, wl, a] W[s', wl, s, wr] ket[s, a, a_r]
 * Replaces Operator::contractL -> contractleftop (src/core/cpp/Operator.cpp:273-334),
 * called from Contractor::calculateL (src/core/cpp/Contractor.cpp:723-742).
 * `out` is reshaped to [Dr_bra, wr, Dr_ket]; it must not alias an input. */
int mpsd_env_left(mpsd_ctx* ctx, const mpsd_tensor* envL, const mpsd_tensor* ket,
                  const mpsd_tensor* W, const mpsd_tensor* bra, mpsd_tensor* out);
/* out[Dl_bra, wl, Dl_ket] = sum conj(bra[s', b, c]) W[s', wl, s, wr] envR[c, wr, a] ket[s, a_l, a]
 * Replaces Operator::contractR -> contractrightop (src/core/cpp/Operator.cpp:336-398),
 * called from Contractor::calculateR (src/core/cpp/Contractor.cpp:744-764). */
int mpsd_env_right(mpsd_ctx* ctx, const mpsd_tensor* envR, const mpsd_tensor* ket,
                   const mpsd_tensor* W, const mpsd_tensor* bra, mpsd_tensor* out);

/* ---- effective Hamiltonian ---------------------------------------------------------------
 * Heff = (L . W . R) with y[s', b_l, b_r] = sum L[b_l,a,a_l] W[s',a,s,b] R[b_r,b,a_r] x[s,a_l,a_r]
 * (+ offset * x).  create replaces the TensorMultiplierOffset constructor
 * (src/wrapper/multipliers/TensorMultiplier.cpp:8-39, TensorMultiplierOffset.cpp:8-12) built
 * in Contractor::findGroundState (src/core/cpp/Contractor.cpp:1400-1406); the handle keeps
 * references to the three device tensors (no copies), which must outlive it. */
int mpsd_heff_create(mpsd_ctx* ctx, const mpsd_tensor* envL, const mpsd_tensor* W,
                     const mpsd_tensor* envR, double offset, int precision, mpsd_heff** out);
int mpsd_heff_destroy(mpsd_heff* h);
/* Multiplier::getSize (src/wrapper/multipliers/Multiplier.h:18-49): d * Dl_ket * Dr_ket. */
int64_t mpsd_heff_size(const mpsd_heff* h);
/* y = Heff x for nvec column vectors.  Replaces TensorMultiplierOffset::product
 * (TensorMultiplierOffset.cpp:14-17 -> TensorMultiplier.cpp:61-93) as called by the PRIMME
 * matvec callback matrixMatvec(x, ldx, y, ldy, blockSize, ...) (src/wrapper/eigs.cpp:296-302,
 * primme_headers_v3/primme_eigs.h:170-172) and by ARPACK reverse communication
 * (src/wrapper/eigs.cpp:125-138).  ldx/ldy are in complex elements (>= size).  With
 * on_device = 0, x and y are host buffers (copied over PCIe); with on_device = 1 they are
 * device addresses (e.g. mpsd_tensor_devptr).  The call is stream-ordered on the context
 * and, for host buffers, returns after y has been written. */
int mpsd_heff_apply(mpsd_heff* h, const double* x, double* y, int nvec, int64_t ldx,
                    int64_t ldy, int on_device);
/* Allocation-free matvec, for callers that cannot allocate device memory during the call (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It has no
 * counterpart in the reference.  y = Heff x + offset x in FP64 on device vectors x and y, with
 * Heff = (envL, W, envR) as in mpsd_heff_create, on ctx's stream; the result is bit-identical to
 * mpsd_heff_apply of a handle over the same tensors.  The MPO mix matrix and the intermediates
 * (T1 or the partial products, and T2: about 2 d w D^2 complex numbers) live in one static
 * 64 MiB device buffer per device, so no call allocates or frees (D = 512 with w = 3 fits;
 * larger shapes, e.g. D = 512 with w = 5, return MPSD_ERR_UNSUPPORTED).  Every call on a device
 * uses that device's one buffer: calls that overlap on one device (contexts on different streams)
 * race and must be ordered by the caller.  Nothing in the buffer needs clearing between calls;
 * each call overwrites what it reads. */
int mpsd_heff_apply_noalloc(mpsd_ctx* ctx, const mpsd_tensor* envL, const mpsd_tensor* W, const mpsd_tensor* envR,
                            double offset, const double* x, double* y);
/* Grouped form: y_i = Heff_i x_i for n effective Hamiltonians of any shapes (independent
 * solves), in one sequence of grouped persistent launches on ctx's stream.  x_i / y_i are
 * device vectors.  Each result is bit-identical to mpsd_heff_apply of the same handle
 * alone.  No counterpart in the reference (one solve at a time there). */
int mpsd_heff_apply_grouped(mpsd_ctx* ctx, int n, mpsd_heff* const* h, const double* const* x,
                            double* const* y);

/* ---- tile autotuner -------------------------------------------------------------------------
 * Times the BRGEMM launches of one operation under each candidate tile configuration (output
 * tile 32x32 / 64x32 / 32x64 / 64x64, K stage 16 / 32, one or two batches per workgroup) and
 * keeps the fastest per BRGEMM shape class on the context; later launches of that shape class
 * use it.  Every candidate accumulates each output element in the same order, so results are
 * bit-identical with and without tuning.  An explicit mpsd_ctx_set_tile / tile option takes
 * precedence over the table.  No counterpart in the reference (its GEMMs are BLAS calls).
 *   tune_heff: reps matvecs y = Heff x per candidate (x, y device vectors; y ends holding Heff x)
 *   tune_env : reps environment updates per candidate (side 0 = left, 1 = right; arguments as
 *              mpsd_env_left / mpsd_env_right)
 * out (optional) receives one row per shape class, *n_out the rows written. */
typedef struct mpsd_tune_result {
  int M, N, K, NB, arith, kind; /* shape class; kind = MPSD_KERNEL_* of the launch */
  int tm, tn, tk, nbw;          /* chosen configuration (tm = tn = 0: built-in) */
  double us_builtin, us_best;   /* average launch time, built-in choice and chosen */
} mpsd_tune_result;
int mpsd_tune_heff(mpsd_heff* h, const double* x, double* y, int reps, mpsd_tune_result* out, int max_out,
                   int* n_out);
int mpsd_tune_env(mpsd_ctx* ctx, int side, const mpsd_tensor* env, const mpsd_tensor* ket,
                  const mpsd_tensor* W, const mpsd_tensor* bra, mpsd_tensor* out_env, int reps,
                  mpsd_tune_result* out, int max_out, int* n_out);
/* Shape classes in the context's table; clear it; save / load it as text (load returns the
 * number of entries read and merges them into the table). */
int mpsd_tune_count(const mpsd_ctx* ctx);
int mpsd_tune_clear(mpsd_ctx* ctx);
int mpsd_tune_save(const mpsd_ctx* ctx, const char* path);
int mpsd_tune_load(mpsd_ctx* ctx, const char* path);
/* Dense Heff (size x size, column-major) into a host buffer.  Replaces
 * TensorMultiplier::getFullTensor (TensorMultiplier.cpp:102-117) / the Offset variant
 * (TensorMultiplierOffset.cpp:21-25), used by Contractor::solveEigs below 801
 * (src/core/cpp/Contractor.cpp:3126-3133). */
int mpsd_heff_full(mpsd_heff* h, double* host_out);
/* Hermitian form: on = 1 makes the handle apply (Heff + Heff^H) x / 2 + offset x, with Heff^H
 * the product over conjugate-transposed copies of L, W, R taken now (call again after changing
 * them); mpsd_heff_full then returns (H + H^H) / 2 + offset I.  Replaces
 * TensorMultiplierHermitian(Offset) (src/wrapper/multipliers/TensorMultiplierHermitian.cpp:8-66,
 * TensorMultiplierHermitianOffset.cpp:8-16, .h:20-24), which findGroundState builds for the
 * fulleig solver (src/core/cpp/Contractor.cpp:1403-1404).  Needs a square Heff (MPSD_ERR_SHAPE
 * otherwise, where the reference exits 212).  on = 0 releases the copies. */
int mpsd_heff_set_hermitian(mpsd_heff* h, int on);
/* Lowest eigenpair from the dense matrix: mpsd_heff_full's matrix on the device, averaged with
 * its adjoint, diagonalised by rocSOLVER zheevd; x (device, mpsd_heff_size elements) receives
 * the eigenvector, *value the eigenvalue (offset included).  Replaces the dense branch of
 * Contractor::solveEigs (src/core/cpp/Contractor.cpp:3087-3100 for fulleig, :3126-3133 below
 * LOWEST_PRIMME) with wrapper::eig (src/wrapper/eig.cpp:46-145).  Local size <= 16384. */
int mpsd_heff_eigs_dense(mpsd_heff* h, double* x, double* value);
/* Lowest eigenpair of a small FP64 Heff (local size <= 2048, matvec <= 1.5 MFLOP, Krylov basis
 * <= 63) by restarted Lanczos run entirely inside one workgroup: matvec, two-pass Gram-Schmidt,
 * tridiagonal eigenpair by multisection + inverse iteration, restarts and the Ritz vector, in
 * one launch.  x (device) holds the start vector and receives the eigenvector.  out6 (host):
 * eigenvalue, residual estimate, matvecs, restarts, converged (0/1), status (0 ok, 2 start
 * vector zero or not finite, 3 non-finite coefficients).  MPSD_ERR_UNSUPPORTED when the
 * problem does not fit.  Stands in for the PRIMME/ARPACK solve of small local problems
 * (src/core/cpp/Contractor.cpp:3067-3138, src/wrapper/eigs.cpp:325-422). */
int mpsd_heff_eigs_fused(mpsd_heff* h, double* x, double tol, int max_krylov, int max_restarts, double* out6);
/* Change the arithmetic of an existing handle (MPSD_PREC_*). */
int mpsd_heff_set_precision(mpsd_heff* h, int precision);

/* ---- device vectors (FP64 eigensolver support) ------------------------------------------
 * The orthogonalisation, Rayleigh-Ritz and residual arithmetic that PRIMME/ARPACK do on the
 * host (src/wrapper/eigs.cpp:71-422) run on device vectors here.  Vectors are complex128
 * device buffers of n complex elements; V holds m of them with leading dimension ldv.
 * Reductions have a fixed order that depends only on n (bit-reproducible). */
/* out_host[j] = sum_i conj(V[j*ldv + i]) * w[i], j < m (synchronises the stream). */
int mpsd_vec_dots(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, const double* w, int64_t n,
                  double* out_host);
/* w[i] += sum_j V[j*ldv + i] * c[j], j < m <= 1024, c a host array of m complex. */
int mpsd_vec_gemv(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, const double* c_host, double* w,
                  int64_t n);
/* Device-resident forms (no host synchronisation): out_dev = V^H w as m complex numbers in
 * device memory; w += sign * V c_dev. */
int mpsd_vec_dots_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, const double* w, int64_t n,
                      double* out_dev);
int mpsd_vec_gemv_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, const double* c_dev, double sign,
                      double* w, int64_t n);
/* y = x / |x| on the device (y may be x); nrm2_dev receives |x|^2 as a complex (re, 0).
 * A zero or non-finite norm leaves y = x. */
int mpsd_vec_normalize_dev(mpsd_ctx* ctx, const double* x, double* y, int64_t n, double* nrm2_dev);
/* Lanczos step j on the device: from h1 = V_{0..j}^H w (first Gram-Schmidt pass) and
 * h2 = [V_{0..j}, w]^H w (after it) store alpha_j = Re(h1[j] + h2[j]) and
 * beta_j = sqrt(|w|^2 - sum |h2[i]|^2) at ab_dev[2j], ab_dev[2j+1], and scale w by 1/beta_j
 * (w = 0 when beta_j underflows). */
int mpsd_lanczos_coef(mpsd_ctx* ctx, const double* h1_dev, const double* h2_dev, int j, double* ab_dev,
                      double* w, int64_t n);
/* mpsd_vec_gemv_dev(V, j + 1, h2, -1, w) followed by mpsd_lanczos_coef(h1, h2, j, ab, w) in
 * one launch, with the same arithmetic: w = (w - V_{0..j} h2) / beta_j, alpha_j and beta_j at
 * ab_dev[2j], ab_dev[2j+1]. */
int mpsd_lanczos_coef_update(mpsd_ctx* ctx, const double* V, int64_t ldv, const double* h1_dev,
                             const double* h2_dev, int j, double* ab_dev, double* w, int64_t n);
/* The two Gram-Schmidt passes of a stream-launched Lanczos step, each in two launches when the
 * vectors are short enough (n <= 32768): gs1 = mpsd_vec_dots_dev(V, m, w, h) then
 * mpsd_vec_gemv_dev(V, m, h, -1, w); gs2 = mpsd_vec_dots_dev(V, j + 2, w, h2) then
 * mpsd_lanczos_coef_update(V, h1, h2, j, ab, w).  The dots' last reduction runs inside the
 * consuming kernel, with the same bits. */
int mpsd_lanczos_gs1_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int m, double* w, int64_t n, double* h_dev);
int mpsd_lanczos_gs2_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int j, const double* h1_dev, double* h2_dev,
                         double* ab_dev, double* w, int64_t n);
/* Both Gram-Schmidt passes and the coefficients of Lanczos step j in one call, reading the
 * basis V_0..V_j three times instead of four: h1 = V^H w, w1 = w - V h1 (kept in registers,
 * never stored), h2[0..j] = V^H w1 and h2[j+1] = |w1|^2, then w = (w1 - V h2) / beta_j with
 * alpha_j, beta_j at ab_dev[2j], ab_dev[2j+1] (mpsd_lanczos_coef's formulas).  h1_dev holds
 * j + 1 and h2_dev j + 2 complex values on return.  The sums run in a fixed order of their own
 * (it depends on n and j only), so the results agree with gs1 + gs2 to rounding, not bit for
 * bit.  j + 1 <= 32 (MPSD_ERR_UNSUPPORTED above); w need not be a column of V.
 * Serves the eigensolver loop of Contractor::findGroundState (the PRIMME/ARPACK iterations of
 * src/wrapper/eigs.cpp run the same orthogonalisation inside the library). */
int mpsd_lanczos_gs_step_dev(mpsd_ctx* ctx, const double* V, int64_t ldv, int j, double* w, int64_t n,
                             double* h1_dev, double* h2_dev, double* ab_dev);
/* Projector terms of TensorMultiplierProjMulti::product (src/wrapper/multipliers/
 * TensorMultiplierProjMulti.cpp:87-96): y -= sum_j weights[j] <phi_j|x> phi_j, phi_j at
 * Phi + 2*j*ldphi, j < m <= 32; weights is a host array of m reals; x and y distinct device
 * vectors.  No host synchronisation. */
int mpsd_vec_proj_update(mpsd_ctx* ctx, const double* Phi, int64_t ldphi, int m, const double* weights,
                         const double* x, double* y, int64_t n);
/* y = a*x + b*y with complex scalars a, b given as {re, im}. */
int mpsd_vec_axpby(mpsd_ctx* ctx, const double* a, const double* x, const double* b, double* y, int64_t n);
/* y = x ; x = 0 (plain stores: a NaN in the old y does not survive). */
int mpsd_vec_copy(mpsd_ctx* ctx, const double* x, double* y, int64_t n);
int mpsd_vec_zero(mpsd_ctx* ctx, double* x, int64_t n);
/* Copy `bytes` from device memory to host after the context's queued work (synchronises). */
int mpsd_copy_to_host(mpsd_ctx* ctx, const void* dev, void* host, int64_t bytes);
/* The same in two halves: mpsd_fetch_async queues the copy of up to 64 KiB into pinned landing
 * slot 0 or 1 behind the context's queued work and returns; mpsd_fetch_wait waits for that copy
 * and moves the bytes to host.  Work queued after the fetch runs on meanwhile (the eigensolver
 * queues its next chunk of steps before it reads the coefficients of the last one). */
int mpsd_fetch_async(mpsd_ctx* ctx, const void* dev, int64_t bytes, int slot);
int mpsd_fetch_wait(mpsd_ctx* ctx, int slot, void* host, int64_t bytes);
/* out = sum |x_i|^2 (synchronises the stream). */
int mpsd_vec_norm2(mpsd_ctx* ctx, const double* x, int64_t n, double* out);

/* ---- gauge steps (kept in HBM) -----------------------------------------------------------
 * site [d, Dl, Dr] := Q of the QR of site viewed (d*Dl) x Dr; R_out := R (k x Dr),
 * k = min(d*Dl, Dr), and the site's right bond becomes k.  Replaces Site::gaugeR with
 * cutD = 0 (src/core/cpp/Site.cpp:179-208) over wrapper::qr (src/wrapper/qr.cpp:33-111),
 * including the edge phase convention for a 1x1 factor (Site.cpp:202-207). */
int mpsd_site_gauge_r(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* R_out);
/* site := Q of the LQ of site permuted to Dl x (d*Dr); L_out := L (Dl x k).  Replaces
 * Site::gaugeL (Site.cpp:210-238) over wrapper::lq (src/wrapper/qr.cpp:113-188). */
int mpsd_site_gauge_l(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* L_out);
/* Site::gaugeR / gaugeL with cutD > 0 (Site.cpp:178-238): an SVD (rocSOLVER zgesvd) in place of
 * the QR/LQ, truncated as wrapper::svd does with tol = 0 and Xi = cutD (src/wrapper/svd.cpp:
 * 23-118): the leading singular values above zero among the first min(cutD, k) are kept.
 * gauge_r_svd: site (d*Dl x Dr) = U S V^H; the site becomes U [d, Dl, X] and R_out = S V^H
 * (X x Dr).  gauge_l_svd: M[a, s + d c] = U S V^H; L_out = U S (Dl x X) and the site becomes
 * V^H [d, X, Dr].  The 1x1 edge factor's phase stays in the site, as with the QR forms.
 * Synchronises the stream (the kept count is decided on the host). */
int mpsd_site_gauge_r_svd(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* R_out, int cutD);
int mpsd_site_gauge_l_svd(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* L_out, int cutD);
/* site[s, :, c] := R . site[s, :, c]  -- Site::reduceDimR (Site.cpp:240-261). */
int mpsd_site_reduce_dim_r(mpsd_ctx* ctx, mpsd_tensor* site, const mpsd_tensor* R);
/* site[s, a, :] := site[s, a, :] . L  -- Site::reduceDimL (Site.cpp:263-283). */
int mpsd_site_reduce_dim_l(mpsd_ctx* ctx, mpsd_tensor* site, const mpsd_tensor* L);
/* A gauge step and the neighbour update that follows it, as MPS::gaugeCond runs them
 * (MPS.cpp:306-337, 366-410): gauge_r_absorb = mpsd_site_gauge_r(site, R_out) then
 * mpsd_site_reduce_dim_r(next, R_out); gauge_l_absorb = mpsd_site_gauge_l(site, L_out) then
 * mpsd_site_reduce_dim_l(prev, L_out).  Where the single-workgroup QR takes the site (the QR
 * path, d*Dl <= 64 and Dr <= 64 for gauge_r, d*Dr <= 64 and Dl <= 64 for gauge_l, not a 1x1
 * edge factor) it also applies the factor from its LDS: one launch instead of two.  The factor
 * is still written to R_out / L_out.  Same results as the two calls to rounding. */
int mpsd_site_gauge_r_absorb(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* R_out, mpsd_tensor* next);
int mpsd_site_gauge_l_absorb(mpsd_ctx* ctx, mpsd_tensor* site, mpsd_tensor* L_out, mpsd_tensor* prev);

/* ---- device-driven Lanczos steps (replayed hipGraph) ----------------------------------------
 * The Lanczos loop of the eigensolver with every step's kernels driven from a device
 * control block, so k steps are one captured hipGraph replayed for every chunk
 * of every site with the same launch shape.  Requires an Heff with d = 2 in the FP64 or
 * FP32 form.  Stands in for the
 * per-iteration host work of PRIMME/ARPACK (src/wrapper/eigs.cpp:71-422). */
typedef struct mpsd_lanczos mpsd_lanczos;
int mpsd_lanczos_create(mpsd_ctx* ctx, int max_krylov, mpsd_lanczos** out);
int mpsd_lanczos_destroy(mpsd_lanczos* lz);
/* Bind Heff h and start the basis with V_0 = x / |x| (x a device vector of the Heff size;
 * normalised on the device). */
int mpsd_lanczos_begin(mpsd_lanczos* lz, mpsd_heff* h, const double* x);
/* Run k (1..8) more steps as one graph replay (graphs are captured once per launch shape and
 * cached); ab_host receives |x|^2 of the start vector, 0, then alpha_i, beta_i (interleaved)
 * of all steps so far: 2 + 2 * (*j_end) doubles. */
int mpsd_lanczos_steps(mpsd_lanczos* lz, int k, double* ab_host, int* j_end);
/* x = sum_{i<K} s_host[i] V_i / |.| (Ritz vector from real coefficients, normalised on the
 * device). */
int mpsd_lanczos_ritz(mpsd_lanczos* lz, const double* s_host, int K, double* x);
/* Number of graph replays so far. */
int64_t mpsd_lanczos_replays(const mpsd_lanczos* lz);

#ifdef __cplusplus
}
#endif

#endif /* MPSD_H */
