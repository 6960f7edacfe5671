// Internal definitions shared by the device library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <atomic>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "mpsd.h"

namespace mpsd {

// One batch-reduce complex GEMM over strided operands (strides in complex elements):
//   C[b](m, n) = sum_{r < R} sum_{k < K} opA(A[b, r](m, k)) * B[b, r](k, n)  (+ beta * Dadd[b](m, n))
// opA conjugates when conjA is set.  Every permute of the reference's reshape/permute/ZGEMM
// chains becomes a choice of strides here, so nothing is materialised but the stage outputs.
struct ZgemmDesc {
  const double2* A = nullptr;
  const double2* B = nullptr;
  double2* C = nullptr;
  const double2* Dadd = nullptr;  // same strides as C; may be null
  double beta = 0.0;
  int M = 0, N = 0, K = 0, R = 1, NB = 1;
  int64_t sAm = 0, sAk = 0, sAr = 0, sAb = 0;
  int64_t sBk = 0, sBn = 0, sBr = 0, sBb = 0;
  int64_t sCm = 0, sCn = 0, sCb = 0;
  // Batches come in groups: batch = j + group * z, j < group.  With group = 0 the launcher
  // picks the group size and sets the z strides to group * s*b; with group > 0 (1 or 2)
  // the z strides below are used as given.
  int group = 0;
  int64_t sAz = 0, sBz = 0, sCz = 0;
  int conjA = 0;
  int arith = 0;    // MFMA arithmetic: 0 FP64, 1 FP32 (FP64 in and out), 2 split-BF16
  int a_mfast = 1;  // load order of the A tile: m fastest (1) or k fastest (0)
  int b_kfast = 1;  // load order of the B tile: k fastest (1) or n fastest (0)
};

// One problem of a grouped BRGEMM launch: a resolved descriptor (group strides set) and its
// tile range in the launch's concatenated tile list.
struct GroupedProblem {
  ZgemmDesc desc;
  int gm = 0, gn = 0, gz = 0;
  int ntiles = 0, tile_begin = 0;
  int pad_[3] = {0, 0, 0};
};

// One problem of a grouped elementwise launch (MPO mix or partial sum).
struct GroupedElem {
  const double2* in = nullptr;  // mix: T1; sum: partials P
  double2* out = nullptr;       // mix: T2; sum: y
  const double2* aux = nullptr; // mix: mix matrix; sum: x (for beta * x), may be null
  double beta = 0.0;
  int Bx = 0, nin = 0, nout = 0, np = 0;
  int64_t Ax = 0, stride = 0, n = 0;
  int64_t begin = 0;  // first element index of this problem in the launch
};

// Tile configuration of a BRGEMM launch (chosen by the tuner or by heuristics).
struct TileCfg {
  int tm = 0, tn = 0, tk = 16;  // tm = tn = 0: built-in choice
  int nbw = 0;                  // batches per workgroup: 0 built-in (choose_group), 1, 2
};

hipError_t launch_zbrgemm_f64(const ZgemmDesc& p, TileCfg cfg, hipStream_t s);

// Shape class of a BRGEMM launch: everything its fastest tile configuration depends on
// (sizes, operand layouts, sharing, arithmetic), none of the addresses.
struct GemmKey {
  int M = 0, N = 0, K = 0, R = 0, NB = 0, flags = 0;
  bool operator<(const GemmKey& o) const {
    return std::tie(M, N, K, R, NB, flags) < std::tie(o.M, o.N, o.K, o.R, o.NB, o.flags);
  }
};
GemmKey gemm_key(const ZgemmDesc& g);
hipError_t launch_zbrgemm_bf16x3(const ZgemmDesc& p, int nbw, hipStream_t s);

// Resolve a descriptor's batch group (sets group strides as launch_zbrgemm_f64 would) and
// return the batches per workgroup it uses.
int resolve_group(ZgemmDesc& p);
// The batch group every launch path uses for a descriptor with no explicit group (cfg_nbw:
// 0 built-in policy, 1 one batch per workgroup), then resolve_group.
int choose_group(ZgemmDesc& p, int cfg_nbw);

// Grouped persistent BRGEMM over nprob problems whose table lives in device memory; all
// problems share (tm, tn, tk, nbw, ASH/BSH kind, arith).  kind: 0 = none shared,
// 1 = A shared by the batch group, 2 = B shared.
// nprob_dev (optional): the problem count is read from device memory at run time instead.
hipError_t launch_zbrgemm_grouped(const GroupedProblem* dev_probs, int nprob, int grid, int tm, int tn, int tk,
                                  int nbw, int kind, int arith, hipStream_t s, const int* nprob_dev = nullptr);
// Grouped MPO mix / partial sums over problem tables in device memory.
hipError_t launch_zmix_grouped(const GroupedElem* dev_probs, int nprob, int grid, hipStream_t s,
                               const int* nprob_dev = nullptr);
hipError_t launch_zsum_grouped(const GroupedElem* dev_probs, int nprob, int grid, hipStream_t s,
                               const int* nprob_dev = nullptr);

// y[i] = sum_{p < np} P[p * stride + i] (+ beta * x[i]) in a fixed order (x may be null).
hipError_t launch_zsum_parts_f64(const double2* P, int np, int64_t stride, const double2* x, double beta, double2* y,
                                 int64_t n, hipStream_t s);

// Apply the MPO "mix" of the middle contraction stage:
//   out[bx + Bx*o + Bx*nout*ax] = sum_i mix[o + nout*i] * in[bx + Bx*i + Bx*nin*ax]
// A short chain of small FP64 contractions in one workgroup: up to three steps, each a BRGEMM
// (ZgemmDesc, batches b < NB at the s*b strides, no groups) or an MPO mix (the arguments of
// launch_zmix_f64), run in order with a barrier between them.  Every output element is
// accumulated exactly as launch_zbrgemm_f64 / launch_zmix_f64 accumulate it, so the results
// are bit-identical to the separate launches; only the launch count changes.
struct ChainStep {
  int kind = 0;  // 0: BRGEMM, 1: MPO mix
  ZgemmDesc g;
  const double2* in = nullptr;
  double2* out = nullptr;
  const double2* mix = nullptr;
  int Bx = 0, nin = 0, nout = 0;
  int64_t Ax = 0;
};
struct Chain {
  ChainStep step[3];
  int nsteps = 0;
};
hipError_t launch_zchain_f64(const Chain& c, hipStream_t s);
hipError_t launch_zmix_f64(const double2* in, double2* out, const double2* mix, int Bx, int nin,
                           int nout, int64_t Ax, hipStream_t s);

// Precision forms with packed low-precision operands (zgemm_lp.hip).  A packed operand holds,
// per batch, `rows` rows of Kp = K rounded up to 32 complex values (zeros past K), 8 bytes
// each (FP32 re/im or split-BF16 hi/lo planes, grouped 8 k at a time); offsets and strides
// below are in those 8-byte elements.
//   C[z](m, n) = sum_k A[z](m, k) B[z](n, k)  (FP32 out), with batch z split as
//   (z % zdiv, z / zdiv) for every operand's two batch strides.
struct LpDesc {
  const void* A = nullptr;
  const void* B = nullptr;
  float2* C = nullptr;    // FP32 and split-BF16 output
  double2* Cd = nullptr;  // FP64 output
  int M = 0, N = 0, Kp = 0, NB = 1, zdiv = 1;
  int prec = MPSD_PREC_FP32;
  int a_il = 0;  // A row-interleaved: 16-byte chunk c of group G (4 chunks) of row m at (G M + m) 4 + c
  // FP64 only: B read in place, row n and contraction index k at n sBn + k sBk (16-byte values),
  // zero for k >= Kb
  int b_ip = 0, Kb = 0;
  int64_t sBk = 0, sBn = 0;
  int64_t sA0 = 0, sA1 = 0, sB0 = 0, sB1 = 0;  // batch strides (packed elements)
  int64_t sCm = 0, sCn = 0, sC0 = 0, sC1 = 0;  // output strides (float2 elements)
  int group = 0;  // row-tiles per tile group of the workgroup order (0: by the launcher)
};
// Pack src(b, row, k) = src[b*sb + row*srow + k*sk] (k < K, FP64) into dst (nb x rows x Kp).
struct LpPackDesc {
  const double2* src = nullptr;
  void* dst = nullptr;
  int nb = 1, rows = 0, K = 0, Kp = 0;
  int64_t sb = 0, srow = 0, sk = 0;
};
// k per packed stage: 16 for FP64 (16-byte values), 32 for the low-precision forms (8 bytes)
inline int lp_stage(int prec) { return prec == MPSD_PREC_FP64 ? 16 : 32; }
inline int lp_pad(int k, int prec) { return (k + lp_stage(prec) - 1) / lp_stage(prec) * lp_stage(prec); }
// Packed FP64 (prec = MPSD_PREC_FP64): 16-byte values, within each 16-deep stage k = 16 t + 4 u + q
// in slot 4 q + u, the k grouping of zbrgemm_f64's MFMAs, so its products are the same bits.
hipError_t launch_pk64_mix(const double2* T1, double2* T2p, const double2* mix, int Db, int nin, int nout, int Dkr,
                           int Kp, hipStream_t s);
// tile: 1 = 32 x 32, 2 = 64 x 32, 3 = 32 x 64, 4 = 64 x 64, 0 = lp_auto_tile
int lp_auto_tile(const LpDesc& p);
// The tile of a packed launch: the tuning candidate, the user's option, the tuned table or 0.
// Call it where the launch is issued (inside timed), so a tuning pass records its kind.
GemmKey lp_key(const LpDesc& g);
hipError_t launch_lp_brgemm(const LpDesc& p, int tile, hipStream_t s);
hipError_t launch_lp_pack(const LpPackDesc& p, int prec, hipStream_t s);
// T2p[o] (rows b, k = a_r, packed row-interleaved) = sum_i mix[o + nout i] T1[b + Db i + Db nin a_r]
// (T1 in FP32)
hipError_t launch_lp_mix(const float2* T1, void* T2p, const double2* mix, int Db, int nin, int nout, int Dkr,
                         int Kp, int prec, hipStream_t s);
// y[i] = sum_{q < np} P[q stride + i] (FP32 partials, FP64 sum) + beta x[i] (x may be null)
hipError_t launch_lp_sum(const float2* P, int np, int64_t stride, const double2* x, double beta, double2* y,
                         int64_t n, hipStream_t s);

// Dense Heff (rows s',b_l,b_r ; cols s,a_l,a_r), column-major, plus offset on the diagonal.
hipError_t launch_heff_full_f64(const double2* L, const double2* W, const double2* R, double2* H,
                                int d, int Db, int wl, int Dk, int Dbr, int wr, int Dkr,
                                double offset, hipStream_t s);

}  // namespace mpsd

struct mpsd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // Scratch (grown on demand, never shrunk): 0, 1 = the two intermediates of a contraction;
  // 2 = Householder scalars of a gauge step; 3 = copy/permute buffer of a gauge step;
  // 4, 5 = intermediates (T1 + partials, T2) of a grouped matvec; 6 = Householder QR workspace
  // (V, T, chunk partials); 7 = second gauge buffer; 8 = Krylov basis of the one-workgroup eigensolver;
  // 9, 10, 11 = packed low-precision L, R and x of a precision-form matvec.
  static constexpr int kScratch = 12;
  void* scratch[kScratch] = {};
  size_t scratch_bytes[kScratch] = {};
  // Packed low-precision environments in slots 9 and 10: the stamp of the Heff handle (and
  // its precision) they were packed for, 0 when none.
  uint64_t lp_stamp = 0;
  int lp_prec = 0;
  int lp_packed = 1;  // precision forms on packed operands (1) or on FP64 operands converted per tile (0)
  int lp_tile = 0;    // packed BRGEMM tile: 0 automatic, 1 32x32, 2 64x32, 3 32x64, 4 64x64
  int pk64 = 1;       // FP64 matvec on packed operands where its last stage is split (same bits)
  int pk64_tile = 0;  // its tile, as lp_tile
  int pk64_x = -1;    // its stage 1 on x: -1 automatic, 0 read in place, 1 packed per call
  // Problem tables of grouped launches: device buffer + pinned host staging, reused after
  // the event of the previous upload has completed.
  void* gtab = nullptr;
  void* gtab_host = nullptr;
  size_t gtab_bytes = 0;
  hipEvent_t gtab_ev = nullptr;
  // Eigensolver support: rocBLAS handle (for rocSOLVER), reduction partials, coefficient ring.
  void* blas = nullptr;
  void* dots = nullptr;
  size_t dots_bytes = 0;
  void* gs_partial = nullptr;  // second-pass block partials of the fused Gram-Schmidt step
  std::atomic<int64_t> graph_launches{0};  // captured eigensolver-step graphs launched on the stream
  size_t gs_partial_bytes = 0;
  void* coef = nullptr;       // device ring, 8 slots x 1024 complex
  void* coef_host = nullptr;  // pinned host ring, same layout
  hipEvent_t coef_ev[8] = {nullptr};
  unsigned coef_slot = 0;
  // Two pinned landing slots for asynchronous reads of small device results (mpsd_fetch_async).
  static constexpr int64_t kFetchBytes = 64 << 10;
  void* fetch_host = nullptr;
  hipEvent_t fetch_ev[2] = {nullptr, nullptr};
  // Staging for host-buffer matvecs.
  void* stage[2] = {nullptr, nullptr};
  size_t stage_bytes[2] = {0, 0};
  int64_t tensor_bytes = 0;
  // Device buffers of destroyed tensors (and mix matrices), by capacity in complex elements,
  // for the next tensors of this context: a solve's setup creates and frees some hundred
  // tensors, and hipMalloc / hipFree cost more than the kernels that fill them.  All of a
  // context's work is on its stream, and a tensor is destroyed after that stream is drained,
  // so a pooled buffer is never still in use when it is handed out again.
  std::multimap<int64_t, double2*> pool;
  int64_t pool_bytes = 0;
  // Pinned staging of host uploads: data is copied in at up_off and DMA'd on the stream; the
  // stream is drained only when the buffer wraps.  status_host: pinned landing place of the
  // one-workgroup eigensolver's status block.
  char* up_host = nullptr;
  size_t up_off = 0;
  static constexpr size_t kUpBytes = size_t(8) << 20;
  double* status_host = nullptr;
  mpsd::TileCfg tile_override{0, 0, 16};
  bool tile_user = false;  // tile_override set away from the built-in choice
  // Autotuner: fastest configuration per BRGEMM shape class (mpsd_tune_*), and, while a
  // tuning run times one candidate, that candidate plus the (kind, shape) pairs it launched.
  std::map<mpsd::GemmKey, mpsd::TileCfg> tuned;
  const mpsd::TileCfg* tune_cand = nullptr;
  std::vector<std::pair<int, mpsd::GemmKey>> tune_seen;
  int cur_kind = -1;  // kind of the launch being issued (set by timed)
  int split3 = -1;  // last matvec stage split over the MPO bond: -1 auto, 0 off, 1 on
  bool graph = false;  // eigensolver steps as replayed hipGraphs where the Heff allows it
  int gauge_qr = 0;   // gauge-step QR / LQ: 0 the library's Householder kernels, 1 rocSOLVER
  int fused_lanczos = 1;  // small local problems: one-workgroup eigensolver (1) or the general one (0)
  int env_chain = 1;      // small environment updates: one single-workgroup launch (1) or three (0)
  int krylov = 0;         // Krylov basis size of a sweep's local eigensolves (0: the default, 30)
  // Launch timing (mpsd_ctx_set_timing): pending event pairs and per-kind totals.
  bool timing = false;
  unsigned timing_mask = ~0u;  // kernel kinds timed when timing is on
  struct Pending {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  double total_ms[MPSD_KERNEL_KINDS] = {0};
  int64_t launches[MPSD_KERNEL_KINDS] = {0};
};

// A device buffer of at least n complex elements from the context's pool (capacity in *cap),
// or a new one; and a buffer back to the pool (freed when the pool would exceed its bound).
namespace mpsd {
hipError_t pool_alloc(mpsd_ctx* ctx, int64_t n, double2** p, int64_t* cap);
void pool_free(mpsd_ctx* ctx, double2* p, int64_t cap);
}  // namespace mpsd

struct mpsd_tensor {
  mpsd_ctx* ctx = nullptr;
  int rank = 0;
  int dims[8] = {0};
  int64_t n = 0;    // complex elements
  int64_t cap = 0;  // capacity in complex elements
  double2* ptr = nullptr;
  // Contents version (bumped by every host upload / device copy into the tensor) and the
  // mix matrices derived from it when the tensor is used as an MPO tensor W.
  uint64_t version = 0;
  mutable double2* mix[2] = {nullptr, nullptr};
  mutable int64_t mix_cap[2] = {0, 0};  // capacity of each mix buffer, complex elements
  mutable uint64_t mix_version[2] = {0, 0};
};

struct mpsd_heff {
  mpsd_ctx* ctx = nullptr;
  const mpsd_tensor* L = nullptr;
  const mpsd_tensor* W = nullptr;
  const mpsd_tensor* R = nullptr;
  double offset = 0.0;
  int precision = MPSD_PREC_FP64;
  // Identifies the operands as this handle saw them when it was created or its precision was
  // last set: the packed low-precision copies of L and R are taken once per stamp.
  uint64_t lp_stamp = 0;
  int d = 0, Db = 0, wl = 0, Dk = 0, Dbr = 0, wr = 0, Dkr = 0;
  double2* mix = nullptr;  // device (d*wr) x (wl*d) mix matrix from W (owned by W's cache)
  // Hermitian form (mpsd_heff_set_hermitian): y = (H x + H' x) / 2 + offset x, where H' is the
  // product over conjugate-transposed copies of L, W, R taken when the form is switched on.
  int herm = 0;
  mpsd_tensor* Lh = nullptr;
  mpsd_tensor* Wh = nullptr;
  mpsd_tensor* Rh = nullptr;
  mpsd_heff* hh = nullptr;   // H' (offset 0), applied on this handle's context
  double2* htmp = nullptr;   // H' x, d * Db * Dbr elements
};

namespace mpsd {
// Release the Hermitian-form copies of a handle (no-op when the form is off).
void heff_drop_hermitian(mpsd_heff* h);
// Dense Heff (column-major, rows = output index) into device memory on h's stream: H + offset I,
// or (H + H^H) / 2 + offset I for the Hermitian form.
hipError_t heff_full_dev(mpsd_heff* h, double2* out);
// Householder QR of A (m x n, column-major, in place) on ctx's stream: Q (m x k) and R
// (k x n), k = min(m, n), in the LAPACK zgeqrf / zungqr conventions.  m <= house_qr_max_rows().
// Small matrices (house_qr_small) take one single-workgroup launch that leaves A as it was and
// lets Q be A itself; larger ones go through panels and overwrite A.
int house_qr(mpsd_ctx* ctx, double2* A, int m, int n, double2* Q, double2* R);
int house_qr_max_rows();
bool house_qr_small(int m, int n);
// The neighbour update that follows a gauge step (Site::reduceDimR / reduceDimL), done by the
// single-workgroup QR from R in its LDS: old [nd, n, nD] -> out [nd, k, nD] after a QR, old
// [nd, nD, n] -> out [nd, nD, k] after an LQ (n: the factorised matrix's column count).
// out == nullptr: no update.
struct Absorb {
  const double2* old = nullptr;
  double2* out = nullptr;
  int nd = 0, nD = 0;
};
// LQ of a site tensor [d, Dl, Dr] in one launch (house_qr_small(d Dr, Dl)): L (Dl x k) and the
// new site [d, k, Dr], k = min(Dl, d Dr), as Site::gaugeL defines them; site_out may be site;
// with ab.out set it also updates the left neighbour.
int house_lq_site_small(mpsd_ctx* ctx, const double2* site, int d, int Dl, int Dr, double2* site_out, double2* L,
                        Absorb ab = Absorb{});
// house_qr of a small matrix (house_qr_small(m, n)) that also applies R to a neighbour
int house_qr_small_absorb(mpsd_ctx* ctx, const double2* A, int m, int n, double2* Q, double2* R, Absorb ab);
// One-workgroup restarted Lanczos for small Heffs (lanczos_fused.hip): whether h fits, and
// the solve (out6: theta, residual, matvecs, restarts, converged, status 0 / 2 start vector
// zero or not finite / 3 non-finite coefficients).
bool lanczos_fused_fits(const mpsd_heff* h, int m);
int lanczos_fused(mpsd_heff* h, double* x, double tol, int m, int max_restarts, double* out6);
// The context's rocBLAS handle (created on first use, bound to the context's stream).
int blas_handle(mpsd_ctx* ctx, void** out);
int set_err(mpsd_ctx* ctx, int code, const std::string& msg);
int hip_err(mpsd_ctx* ctx, hipError_t e, const char* where);
int ensure_scratch(mpsd_ctx* ctx, int slot, size_t bytes);
// Whether the matvec's last stage is split over the MPO bond (a function of the shape and
// the context option only, so every launch path takes the same arithmetic).
bool stage3_split(const mpsd_heff* h, const mpsd_ctx* ctx);
// Tile configuration of a BRGEMM launch: the candidate of a running tuning pass, else an
// explicit context override, else the tuned entry of g's shape class, else built-in.  Every
// configuration gives bit-identical results, so the choice only moves time.
TileCfg tile_for(mpsd_ctx* ctx, const ZgemmDesc& g);
int lp_tile_for(mpsd_ctx* ctx, const LpDesc& g);
// Run `launch` on the context's stream, bracketed by HIP events when timing is on.
hipError_t timed(mpsd_ctx* ctx, int kind, const std::function<hipError_t()>& launch);
int flush_timing(mpsd_ctx* ctx);
void destroy_blas(mpsd_ctx* ctx);
}  // namespace mpsd
