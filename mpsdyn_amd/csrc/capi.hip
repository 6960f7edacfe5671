// C-ABI of the device library: contexts, HBM tensors, environment updates and the
// effective-Hamiltonian matvec.  Shapes and their chaining are checked on the host before
// any launch; every kernel below only ever sees operands whose sizes match its grid.

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "mpsd_internal.h"

namespace mpsd {

int set_err(mpsd_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_err(mpsd_ctx* ctx, hipError_t e, const char* where) {
  if (e == hipSuccess) return MPSD_OK;
  return set_err(ctx, e == hipErrorOutOfMemory ? MPSD_ERR_NOMEM : MPSD_ERR_HIP,
                 std::string(where) + ": " + hipGetErrorString(e));
}

int ensure_scratch(mpsd_ctx* ctx, int slot, size_t bytes) {
  if (ctx->scratch_bytes[slot] >= bytes) return MPSD_OK;
  if (ctx->scratch[slot]) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_err(ctx, e, "scratch sync");
    (void)hipFree(ctx->scratch[slot]);
    ctx->scratch[slot] = nullptr;
    ctx->scratch_bytes[slot] = 0;
  }
  size_t want = bytes + bytes / 4;  // headroom so a sweep's growing bonds reallocate rarely
  hipError_t e = hipMalloc(&ctx->scratch[slot], want);
  if (e != hipSuccess) return hip_err(ctx, e, "scratch alloc");
  ctx->scratch_bytes[slot] = want;
  return MPSD_OK;
}

static hipEvent_t pool_event(mpsd_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  // timing-only events: no system-scope fence, so recording one neither writes back nor
  // invalidates the caches between the kernels it brackets (elapsed times are read after a
  // stream synchronisation)
  hipEvent_t e = nullptr;
  (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
  return e;
}

hipError_t timed(mpsd_ctx* ctx, int kind, const std::function<hipError_t()>& launch) {
  ctx->cur_kind = kind;
  if (!ctx->timing || !((ctx->timing_mask >> kind) & 1u)) return launch();
  hipEvent_t a = pool_event(ctx), b = pool_event(ctx);
  (void)hipEventRecord(a, ctx->stream);
  hipError_t e = launch();
  (void)hipEventRecord(b, ctx->stream);
  ctx->pending.push_back({kind, a, b});
  return e;
}

int flush_timing(mpsd_ctx* ctx) {
  if (ctx->pending.empty()) return MPSD_OK;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_err(ctx, e, "timing sync");
  for (auto& p : ctx->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      ctx->total_ms[p.kind] += ms;
      ctx->launches[p.kind] += 1;
    }
    ctx->event_pool.push_back(p.a);
    ctx->event_pool.push_back(p.b);
  }
  ctx->pending.clear();
  return MPSD_OK;
}

static int ensure_stage(mpsd_ctx* ctx, int slot, size_t bytes) {
  if (ctx->stage_bytes[slot] >= bytes) return MPSD_OK;
  if (ctx->stage[slot]) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->stage[slot]);
    ctx->stage[slot] = nullptr;
    ctx->stage_bytes[slot] = 0;
  }
  hipError_t e = hipMalloc(&ctx->stage[slot], bytes);
  if (e != hipSuccess) return hip_err(ctx, e, "staging alloc");
  ctx->stage_bytes[slot] = bytes;
  return MPSD_OK;
}

static std::string dims_str(const mpsd_tensor* t) {
  std::string s = "[";
  for (int i = 0; i < t->rank; ++i) s += (i ? "," : "") + std::to_string(t->dims[i]);
  return s + "]";
}

// Dimension i of a tensor viewed with at least `r` indices (trailing ones), as mwArray and
// Operator pad a lower rank with unit dimensions (Operator.cpp:37-48).
static int dim(const mpsd_tensor* t, int i) { return i < t->rank ? t->dims[i] : 1; }

static bool rank_ok(const mpsd_tensor* t, int r) {
  for (int i = r; i < t->rank; ++i)
    if (t->dims[i] != 1) return false;
  return true;
}

// Stage 1 + 2 of a left-type contraction into T2 = scratch[1]:
//   T1[b, al, s, ar] = sum_a E[b, al, a] X[s, a, ar]           (BRGEMM, batch s)
//   T2[b, s', be, ar] = sum_{al,s} W[s', al, s, be] T1[b, al, s, ar]   (mix)
static int arith_of(int precision) {
  return precision == MPSD_PREC_FP32 ? 1 : precision == MPSD_PREC_BF16X3 ? 2 : 0;
}

// An environment update small enough for one workgroup (FP64 flops of its three stages; the
// one-workgroup eigensolver's bound): one launch instead of three, same bits.
static bool env_chain_fits(const mpsd_ctx* ctx, double flops) { return ctx->env_chain && flops <= 1.5e6; }

// s0, s1: where T1 and T2 go (the context's grow-only scratch slots 0 and 1 when null).
static int left_stages(mpsd_ctx* ctx, const double2* E, int Db, int wl, int Dk, const double2* X,
                       int d, int Dkr, const double2* mix, int wr, int arith = 0, Chain* chain = nullptr,
                       double2* s0 = nullptr, double2* s1 = nullptr) {
  if (!s0 || !s1) {
    const size_t t1 = sizeof(double2) * (size_t)Db * wl * d * Dkr;
    const size_t t2 = sizeof(double2) * (size_t)Db * d * wr * Dkr;
    int rc = ensure_scratch(ctx, 0, t1);
    if (rc) return rc;
    rc = ensure_scratch(ctx, 1, t2);
    if (rc) return rc;
    s0 = (double2*)ctx->scratch[0];
    s1 = (double2*)ctx->scratch[1];
  }
  ZgemmDesc g;
  g.A = E;
  g.B = X;
  g.C = s0;
  g.M = Db * wl;
  g.N = Dkr;
  g.K = Dk;
  g.NB = d;
  g.sAm = 1;
  g.sAk = (int64_t)Db * wl;
  g.sAb = 0;
  g.sBk = d;
  g.sBn = (int64_t)d * Dk;
  g.sBb = 1;
  g.sCm = 1;
  g.sCn = (int64_t)Db * wl * d;
  g.sCb = (int64_t)Db * wl;
  g.a_mfast = 1;
  g.b_kfast = 1;
  g.arith = arith;
  if (chain) {
    // the first two steps of a chain: stage 1 and the mix, as launched below
    chain->step[0].kind = 0;
    chain->step[0].g = g;
    ChainStep& mx = chain->step[1];
    mx.kind = 1;
    mx.in = s0;
    mx.out = s1;
    mx.mix = mix;
    mx.Bx = Db;
    mx.nin = wl * d;
    mx.nout = d * wr;
    mx.Ax = Dkr;
    chain->nsteps = 2;
    return MPSD_OK;
  }
  hipError_t e = timed(ctx, MPSD_KERNEL_STAGE1,
                       [&] { return launch_zbrgemm_f64(g, tile_for(ctx, g), ctx->stream); });
  if (e != hipSuccess) return hip_err(ctx, e, "stage1 launch");
  e = timed(ctx, MPSD_KERNEL_MIX, [&] {
    return launch_zmix_f64(s0, s1, mix, Db, wl * d, d * wr, Dkr, ctx->stream);
  });
  return hip_err(ctx, e, "mix launch");
}

// at most this many bytes of freed buffers are kept per context
constexpr int64_t kPoolMaxBytes = int64_t(1) << 31;

hipError_t pool_alloc(mpsd_ctx* ctx, int64_t n, double2** p, int64_t* cap) {
  // the smallest pooled buffer that holds n elements without wasting more than half of it
  auto it = ctx->pool.lower_bound(n);
  if (it != ctx->pool.end() && it->first <= 2 * n + 256) {
    *p = it->second;
    *cap = it->first;
    ctx->pool_bytes -= (int64_t)sizeof(double2) * it->first;
    ctx->pool.erase(it);
    return hipSuccess;
  }
  *cap = n;
  return hipMalloc((void**)p, sizeof(double2) * (size_t)std::max<int64_t>(n, 1));
}

void pool_free(mpsd_ctx* ctx, double2* p, int64_t cap) {
  if (!p) return;
  const int64_t bytes = (int64_t)sizeof(double2) * cap;
  if (ctx->pool_bytes + bytes > kPoolMaxBytes) {
    (void)hipFree(p);
    return;
  }
  ctx->pool.emplace(cap, p);
  ctx->pool_bytes += bytes;
}

namespace {
// The mix matrix of W, a permutation of its elements formed on the device (no read-back of W,
// no upload, nothing the host waits for):
//   left-type contraction (kind 0):  M[o = s' + d be, i = al + wl s] = W[s', al, s, be],
//   right-type contraction (kind 1): M[o = s' + d al, i = be + wr s] = W[s', al, s, be].
__global__ void __launch_bounds__(256) mix_perm_kernel(const double2* __restrict__ W, int d, int wl, int wr,
                                                       int kind, double2* __restrict__ m) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= d * wl * d * wr) return;
  const int sp = t % d, al = (t / d) % wl, s = (t / (d * wl)) % d, be = t / (d * wl * d);
  if (kind == 0)
    m[(sp + d * be) + (int64_t)d * wr * (al + wl * s)] = W[t];
  else
    m[(sp + d * al) + (int64_t)d * wl * (be + wr * s)] = W[t];
}

// Scratch of the allocation-free matvec (mpsd_heff_apply_noalloc): the MPO mix matrix, then T1
// and the partial products, then T2, each starting on a 256-byte boundary.  A file-scope device
// array, so each device holds its own copy for the life of the process and no call allocates.
// Host code never uses the name `scratch` as an address: the device address comes from
// hipGetSymbolAddress, once per device (noalloc_setup, called by mpsd_ctx_create).
constexpr size_t kNoAllocBytes = size_t(64) << 20;
__device__ __align__(256) double2 scratch[kNoAllocBytes / sizeof(double2)];
constexpr int kMaxDevices = 64;
std::mutex noalloc_mu;
void* noalloc_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be the current device.
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (noalloc_ptr[device]) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  noalloc_ptr[device] = p;
  return hipSuccess;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  return noalloc_ptr[device];
}

// Old and new dimensions of a resize (column-major, rank <= 8).
struct ResizeDims {
  int rank = 0;
  int od[8] = {0};
  int nd[8] = {0};
};

// mwArray::resize (src/wrapper/mwArray.cpp:1604-1670) with zero noise: element i of the new
// shape takes the old element at the same multi-index when every index is inside the old
// dimensions, else zero.  One thread per new element.
__global__ void __launch_bounds__(256) resize_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                     ResizeDims rd, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t rest = i, off = 0, ostride = 1;
    bool inside = true;
    for (int r = 0; r < rd.rank; ++r) {
      const int64_t idx = rest % rd.nd[r];
      rest /= rd.nd[r];
      inside = inside && idx < rd.od[r];
      off += idx * ostride;
      ostride *= rd.od[r];
    }
    out[i] = inside ? in[off] : double2{0.0, 0.0};
  }
}
}  // namespace

}  // namespace mpsd

using namespace mpsd;

namespace {
// The mix matrices derived from a W tensor are cached on the tensor and rebuilt only when
// its contents change (every host upload or device copy bumps its version).
int get_mix(mpsd_ctx* ctx, const mpsd_tensor* W, int kind, int d, int wl, int wr, double2** out) {
  if (W->mix[kind] && W->mix_version[kind] == W->version) {
    *out = W->mix[kind];
    return MPSD_OK;
  }
  const int64_t n = (int64_t)d * wl * d * wr;
  if (n != W->n) return set_err(ctx, MPSD_ERR_SHAPE, "mix: W size does not match its dimensions");
  if (W->mix[kind] && W->mix_cap[kind] < n) {
    (void)hipStreamSynchronize(ctx->stream);
    pool_free(ctx, W->mix[kind], W->mix_cap[kind]);
    W->mix[kind] = nullptr;
  }
  if (!W->mix[kind]) {
    hipError_t e = pool_alloc(ctx, n, &W->mix[kind], &W->mix_cap[kind]);
    if (e != hipSuccess) return hip_err(ctx, e, "mix alloc");
  }
  hipLaunchKernelGGL(mix_perm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, W->ptr, d, wl,
                     wr, kind, W->mix[kind]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_err(ctx, e, "mix build");
  W->mix_version[kind] = W->version;
  *out = W->mix[kind];
  return MPSD_OK;
}
}  // namespace

extern "C" {

int mpsd_ctx_create(int device, mpsd_ctx** out) {
  if (!out) return MPSD_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MPSD_ERR_HIP;
  if (device < 0 || device >= n) return MPSD_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return MPSD_ERR_HIP;
  mpsd_ctx* c = new mpsd_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MPSD_ERR_HIP;
  }
  // the allocation-free matvec's scratch address on this device (once per device)
  if (noalloc_setup(device) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MPSD_ERR_HIP;
  }
  *out = c;
  return MPSD_OK;
}

int mpsd_ctx_destroy(mpsd_ctx* ctx) {
  if (!ctx) return MPSD_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)flush_timing(ctx);
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  for (int i = 0; i < mpsd_ctx::kScratch; ++i)
    if (ctx->scratch[i]) (void)hipFree(ctx->scratch[i]);
  for (int i = 0; i < 2; ++i)
    if (ctx->stage[i]) (void)hipFree(ctx->stage[i]);
  if (ctx->dots) (void)hipFree(ctx->dots);
  if (ctx->gs_partial) (void)hipFree(ctx->gs_partial);
  if (ctx->gtab) (void)hipFree(ctx->gtab);
  if (ctx->gtab_host) (void)hipHostFree(ctx->gtab_host);
  if (ctx->gtab_ev) (void)hipEventDestroy(ctx->gtab_ev);
  if (ctx->coef) (void)hipFree(ctx->coef);
  if (ctx->coef_host) (void)hipHostFree(ctx->coef_host);
  if (ctx->fetch_host) (void)hipHostFree(ctx->fetch_host);
  for (hipEvent_t e : ctx->fetch_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& kv : ctx->pool) (void)hipFree(kv.second);
  ctx->pool.clear();
  if (ctx->up_host) (void)hipHostFree(ctx->up_host);
  if (ctx->status_host) (void)hipHostFree(ctx->status_host);
  for (hipEvent_t e : ctx->coef_ev)
    if (e) (void)hipEventDestroy(e);
  destroy_blas(ctx);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MPSD_OK;
}

int mpsd_ctx_sync(mpsd_ctx* ctx) {
  if (!ctx) return MPSD_ERR_ARG;
  return hip_err(ctx, hipStreamSynchronize(ctx->stream), "sync");
}

int mpsd_ctx_set_timing_kinds(mpsd_ctx* ctx, unsigned mask) {
  if (!ctx) return MPSD_ERR_ARG;
  ctx->timing_mask = mask;
  return MPSD_OK;
}

int mpsd_ctx_set_timing(mpsd_ctx* ctx, int on) {
  if (!ctx) return MPSD_ERR_ARG;
  int rc = flush_timing(ctx);
  for (int k = 0; k < MPSD_KERNEL_KINDS; ++k) {
    ctx->total_ms[k] = 0.0;
    ctx->launches[k] = 0;
  }
  ctx->timing = on != 0;
  return rc;
}

int mpsd_ctx_timing(mpsd_ctx* ctx, int kind, double* total_ms, int64_t* launches) {
  if (!ctx || kind < 0 || kind >= MPSD_KERNEL_KINDS || !total_ms || !launches) return MPSD_ERR_ARG;
  int rc = flush_timing(ctx);
  *total_ms = ctx->total_ms[kind];
  *launches = ctx->launches[kind];
  return rc;
}

static void update_tile_user(mpsd_ctx* ctx) {
  const mpsd::TileCfg& t = ctx->tile_override;
  ctx->tile_user = !(t.tm == 0 && t.tn == 0 && t.tk == 16 && t.nbw == 0);
}

int mpsd_ctx_set_tile(mpsd_ctx* ctx, int tm, int tn) {
  if (!ctx) return MPSD_ERR_ARG;
  const bool ok = (tm == 0 && tn == 0) || (tm == 16 && tn == 16) || ((tm == 32 || tm == 64) && (tn == 32 || tn == 64));
  if (!ok) return set_err(ctx, MPSD_ERR_ARG, "set_tile: tm, tn must be 32 or 64, both 16 or both 0");
  ctx->tile_override.tm = tm;
  ctx->tile_override.tn = tn;
  update_tile_user(ctx);
  return MPSD_OK;
}

int mpsd_ctx_get_option(const mpsd_ctx* ctx, int option) {
  if (!ctx) return MPSD_ERR_ARG;
  switch (option) {
    case MPSD_OPT_TILE_K: return ctx->tile_override.tk;
    case MPSD_OPT_SPLIT_STAGE3: return ctx->split3;
    case MPSD_OPT_GRAPH: return ctx->graph ? 1 : 0;
    case MPSD_OPT_BATCH_GROUP: return ctx->tile_override.nbw;
    case MPSD_OPT_GAUGE_QR: return ctx->gauge_qr;
    case MPSD_OPT_FUSED_LANCZOS: return ctx->fused_lanczos;
    case MPSD_OPT_ENV_CHAIN: return ctx->env_chain;
    case MPSD_OPT_KRYLOV: return ctx->krylov;
    case MPSD_OPT_LP_PACKED: return ctx->lp_packed;
    case MPSD_OPT_LP_TILE: return ctx->lp_tile;
    case MPSD_OPT_PACKED_FP64: return ctx->pk64;
    case MPSD_OPT_PACKED_FP64_TILE: return ctx->pk64_tile;
    case MPSD_OPT_PACKED_FP64_X: return ctx->pk64_x;
    default: return MPSD_ERR_ARG;
  }
}

int mpsd_ctx_set_option(mpsd_ctx* ctx, int option, int value) {
  if (!ctx) return MPSD_ERR_ARG;
  switch (option) {
    case MPSD_OPT_TILE_K:
      if (value != 16 && value != 32) return set_err(ctx, MPSD_ERR_ARG, "set_option: tile k must be 16 or 32");
      ctx->tile_override.tk = value;
      update_tile_user(ctx);
      return MPSD_OK;
    case MPSD_OPT_SPLIT_STAGE3:
      if (value < -1 || value > 1) return set_err(ctx, MPSD_ERR_ARG, "set_option: split must be -1, 0 or 1");
      ctx->split3 = value;
      return MPSD_OK;
    case MPSD_OPT_GRAPH:
      ctx->graph = value != 0;
      return MPSD_OK;
    case MPSD_OPT_GAUGE_QR:
      if (value != 0 && value != 1) return set_err(ctx, MPSD_ERR_ARG, "set_option: gauge QR must be 0 or 1");
      ctx->gauge_qr = value;
      return MPSD_OK;
    case MPSD_OPT_FUSED_LANCZOS:
      ctx->fused_lanczos = value != 0;
      return MPSD_OK;
    case MPSD_OPT_ENV_CHAIN:
      ctx->env_chain = value != 0;
      return MPSD_OK;
    case MPSD_OPT_KRYLOV:
      if (value != 0 && (value < 4 || value > 62))
        return set_err(ctx, MPSD_ERR_ARG, "set_option: Krylov basis size must be 0 (default) or 4..62");
      ctx->krylov = value;
      return MPSD_OK;
    case MPSD_OPT_LP_PACKED:
      ctx->lp_packed = value != 0;
      return MPSD_OK;
    case MPSD_OPT_LP_TILE:
      if (value < 0 || value > 4) return set_err(ctx, MPSD_ERR_ARG, "set_option: packed tile must be 0..4");
      ctx->lp_tile = value;
      return MPSD_OK;
    case MPSD_OPT_PACKED_FP64:
      ctx->pk64 = value != 0;
      return MPSD_OK;
    case MPSD_OPT_PACKED_FP64_TILE:
      if (value < 0 || value > 4) return set_err(ctx, MPSD_ERR_ARG, "set_option: packed tile must be 0..4");
      ctx->pk64_tile = value;
      return MPSD_OK;
    case MPSD_OPT_PACKED_FP64_X:
      if (value < -1 || value > 1) return set_err(ctx, MPSD_ERR_ARG, "set_option: packed x must be -1, 0 or 1");
      ctx->pk64_x = value;
      return MPSD_OK;
    case MPSD_OPT_BATCH_GROUP:
      if (value < 0 || value > 2) return set_err(ctx, MPSD_ERR_ARG, "set_option: batch group must be 0, 1 or 2");
      ctx->tile_override.nbw = value;
      update_tile_user(ctx);
      return MPSD_OK;
    default:
      return set_err(ctx, MPSD_ERR_ARG, "set_option: unknown option");
  }
}

const char* mpsd_last_error(const mpsd_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* mpsd_ctx_stream(mpsd_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int64_t mpsd_ctx_device_bytes(const mpsd_ctx* ctx) {
  if (!ctx) return 0;
  int64_t b = ctx->tensor_bytes + ctx->pool_bytes + (int64_t)ctx->stage_bytes[0] + (int64_t)ctx->stage_bytes[1] +
              (int64_t)ctx->dots_bytes + (int64_t)ctx->gs_partial_bytes;
  for (int i = 0; i < mpsd_ctx::kScratch; ++i) b += (int64_t)ctx->scratch_bytes[i];
  return b;
}

int mpsd_tensor_create(mpsd_ctx* ctx, int rank, const int* dims, mpsd_tensor** out) {
  if (!ctx || !out || rank < 0 || rank > 8 || (rank > 0 && !dims)) return set_err(ctx, MPSD_ERR_ARG, "tensor_create: bad arguments");
  int64_t n = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] <= 0) return set_err(ctx, MPSD_ERR_ARG, "tensor_create: non-positive dimension");
    n *= dims[i];
  }
  mpsd_tensor* t = new mpsd_tensor();
  t->ctx = ctx;
  t->rank = rank;
  for (int i = 0; i < rank; ++i) t->dims[i] = dims[i];
  t->n = n;
  hipError_t e = pool_alloc(ctx, n, &t->ptr, &t->cap);
  if (e == hipSuccess) e = hipMemsetAsync(t->ptr, 0, sizeof(double2) * (size_t)n, ctx->stream);
  if (e != hipSuccess) {
    if (t->ptr) pool_free(ctx, t->ptr, t->cap);
    delete t;
    return hip_err(ctx, e, "tensor_create");
  }
  ctx->tensor_bytes += (int64_t)sizeof(double2) * t->cap;
  *out = t;
  return MPSD_OK;
}

int mpsd_tensor_destroy(mpsd_tensor* t) {
  if (!t) return MPSD_ERR_ARG;
  (void)hipStreamSynchronize(t->ctx->stream);
  pool_free(t->ctx, t->ptr, t->cap);
  for (int k = 0; k < 2; ++k) pool_free(t->ctx, t->mix[k], t->mix_cap[k]);
  t->ctx->tensor_bytes -= (int64_t)sizeof(double2) * t->cap;
  delete t;
  return MPSD_OK;
}

int mpsd_tensor_rank(const mpsd_tensor* t) { return t ? t->rank : MPSD_ERR_ARG; }

int mpsd_tensor_dims(const mpsd_tensor* t, int* dims) {
  if (!t || !dims) return MPSD_ERR_ARG;
  for (int i = 0; i < t->rank; ++i) dims[i] = t->dims[i];
  return MPSD_OK;
}

int64_t mpsd_tensor_size(const mpsd_tensor* t) { return t ? t->n : MPSD_ERR_ARG; }

int mpsd_tensor_reshape(mpsd_tensor* t, int rank, const int* dims) {
  if (!t || rank < 0 || rank > 8 || (rank > 0 && !dims)) return MPSD_ERR_ARG;
  mpsd_ctx* ctx = t->ctx;
  int64_t n = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] <= 0) return set_err(ctx, MPSD_ERR_ARG, "tensor_reshape: non-positive dimension");
    n *= dims[i];
  }
  if (n > t->cap) {
    // the old buffer back to the pool and a larger one from it: both reuses are ordered on the
    // context's stream behind every launch that still reads the old one, so no wait
    hipError_t e = hipSuccess;
    pool_free(ctx, t->ptr, t->cap);
    ctx->tensor_bytes -= (int64_t)sizeof(double2) * t->cap;
    t->ptr = nullptr;
    t->cap = 0;
    e = pool_alloc(ctx, n, &t->ptr, &t->cap);
    if (e != hipSuccess) return hip_err(ctx, e, "reshape alloc");
    ctx->tensor_bytes += (int64_t)sizeof(double2) * t->cap;
  }
  t->rank = rank;
  for (int i = 0; i < 8; ++i) t->dims[i] = i < rank ? dims[i] : 0;
  t->n = n;
  return MPSD_OK;
}

int mpsd_tensor_resize(mpsd_tensor* t, int rank, const int* dims) {
  if (!t || rank < 1 || rank > 8 || !dims) return MPSD_ERR_ARG;
  mpsd_ctx* ctx = t->ctx;
  if (rank != t->rank)
    return set_err(ctx, MPSD_ERR_SHAPE, "tensor_resize: rank " + std::to_string(rank) + " for a tensor of rank " +
                                            std::to_string(t->rank));
  ResizeDims rd;
  rd.rank = rank;
  int64_t n = 1;
  bool same = true;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] <= 0) return set_err(ctx, MPSD_ERR_ARG, "tensor_resize: non-positive dimension");
    rd.nd[i] = dims[i];
    rd.od[i] = t->dims[i];
    same = same && dims[i] == t->dims[i];
    n *= dims[i];
  }
  if (same) return MPSD_OK;  // mwArray::resize: nothing to do
  double2* p = nullptr;
  int64_t cap = 0;
  hipError_t e = pool_alloc(ctx, n, &p, &cap);
  if (e != hipSuccess) return hip_err(ctx, e, "tensor_resize alloc");
  const unsigned blocks = (unsigned)std::min<int64_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(resize_kernel, dim3(blocks), dim3(256), 0, ctx->stream, t->ptr, p, rd, n);
  e = hipGetLastError();
  if (e != hipSuccess) {
    pool_free(ctx, p, cap);
    return hip_err(ctx, e, "tensor_resize");
  }
  // the old buffer goes back to the pool behind the copy on the same stream
  pool_free(ctx, t->ptr, t->cap);
  ctx->tensor_bytes += (int64_t)sizeof(double2) * (cap - t->cap);
  t->ptr = p;
  t->cap = cap;
  for (int i = 0; i < rank; ++i) t->dims[i] = dims[i];
  t->n = n;
  ++t->version;
  return MPSD_OK;
}

int mpsd_tensor_upload(mpsd_tensor* t, const double* host, int64_t n_complex) {
  if (!t || !host) return MPSD_ERR_ARG;
  if (n_complex != t->n) return set_err(t->ctx, MPSD_ERR_SHAPE, "tensor_upload: size mismatch");
  mpsd_ctx* ctx = t->ctx;
  const size_t bytes = sizeof(double2) * (size_t)t->n;
  hipError_t e = hipSuccess;
  if (bytes <= mpsd_ctx::kUpBytes / 4) {
    // through the pinned staging buffer: no wait for the copy (or anything before it)
    if (!ctx->up_host) e = hipHostMalloc((void**)&ctx->up_host, mpsd_ctx::kUpBytes, hipHostMallocDefault);
    if (e == hipSuccess && ctx->up_off + bytes > mpsd_ctx::kUpBytes) {
      e = hipStreamSynchronize(ctx->stream);  // every staged copy has landed: start over
      ctx->up_off = 0;
    }
    if (e == hipSuccess) {
      char* stage = ctx->up_host + ctx->up_off;
      std::memcpy(stage, host, bytes);
      ctx->up_off += (bytes + 255) & ~size_t(255);
      e = hipMemcpyAsync(t->ptr, stage, bytes, hipMemcpyHostToDevice, ctx->stream);
    }
  } else {
    e = hipMemcpyAsync(t->ptr, host, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  ++t->version;
  return hip_err(ctx, e, "tensor_upload");
}

int mpsd_tensor_create_empty(mpsd_ctx* ctx, int rank, const int* dims, mpsd_tensor** out) {
  if (!ctx || !out || rank < 0 || rank > 8 || (rank > 0 && !dims))
    return set_err(ctx, MPSD_ERR_ARG, "tensor_create_empty: bad arguments");
  int64_t n = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] <= 0) return set_err(ctx, MPSD_ERR_ARG, "tensor_create_empty: non-positive dimension");
    n *= dims[i];
  }
  mpsd_tensor* t = new mpsd_tensor();
  t->ctx = ctx;
  t->rank = rank;
  for (int i = 0; i < rank; ++i) t->dims[i] = dims[i];
  t->n = n;
  hipError_t e = pool_alloc(ctx, n, &t->ptr, &t->cap);
  if (e != hipSuccess) {
    delete t;
    return hip_err(ctx, e, "tensor_create_empty");
  }
  ctx->tensor_bytes += (int64_t)sizeof(double2) * t->cap;
  *out = t;
  return MPSD_OK;
}

int mpsd_tensor_create_from(mpsd_ctx* ctx, int rank, const int* dims, const double* host, mpsd_tensor** out) {
  if (!ctx || !out || !host || rank < 0 || rank > 8 || (rank > 0 && !dims))
    return set_err(ctx, MPSD_ERR_ARG, "tensor_create_from: bad arguments");
  int64_t n = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] <= 0) return set_err(ctx, MPSD_ERR_ARG, "tensor_create_from: non-positive dimension");
    n *= dims[i];
  }
  mpsd_tensor* t = new mpsd_tensor();
  t->ctx = ctx;
  t->rank = rank;
  for (int i = 0; i < rank; ++i) t->dims[i] = dims[i];
  t->n = n;
  hipError_t e = pool_alloc(ctx, n, &t->ptr, &t->cap);
  if (e != hipSuccess) {
    delete t;
    return hip_err(ctx, e, "tensor_create_from");
  }
  ctx->tensor_bytes += (int64_t)sizeof(double2) * t->cap;
  const int rc = mpsd_tensor_upload(t, host, n);
  if (rc) {
    (void)mpsd_tensor_destroy(t);
    return rc;
  }
  *out = t;
  return MPSD_OK;
}

int mpsd_tensor_download(const mpsd_tensor* t, double* host, int64_t n_complex) {
  if (!t || !host) return MPSD_ERR_ARG;
  if (n_complex != t->n) return set_err(t->ctx, MPSD_ERR_SHAPE, "tensor_download: size mismatch");
  hipError_t e = hipMemcpyAsync(host, t->ptr, sizeof(double2) * (size_t)t->n, hipMemcpyDeviceToHost,
                                t->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(t->ctx->stream);
  return hip_err(t->ctx, e, "tensor_download");
}

int mpsd_tensor_copy(mpsd_tensor* dst, const mpsd_tensor* src) {
  if (!dst || !src) return MPSD_ERR_ARG;
  if (dst->n != src->n) return set_err(dst->ctx, MPSD_ERR_SHAPE, "tensor_copy: size mismatch");
  ++dst->version;
  return hip_err(dst->ctx,
                 hipMemcpyAsync(dst->ptr, src->ptr, sizeof(double2) * (size_t)src->n,
                                hipMemcpyDeviceToDevice, dst->ctx->stream),
                 "tensor_copy");
}

void* mpsd_tensor_devptr(mpsd_tensor* t) { return t ? (void*)t->ptr : nullptr; }

int mpsd_env_left(mpsd_ctx* ctx, const mpsd_tensor* envL, const mpsd_tensor* ket,
                  const mpsd_tensor* W, const mpsd_tensor* bra, mpsd_tensor* out) {
  if (!ctx || !envL || !ket || !W || !bra || !out) return set_err(ctx, MPSD_ERR_ARG, "env_left: null argument");
  if (!rank_ok(envL, 3) || !rank_ok(ket, 3) || !rank_ok(W, 4) || !rank_ok(bra, 3))
    return set_err(ctx, MPSD_ERR_SHAPE, "env_left: rank too high");
  if (out == envL || out == ket || out == bra || out == W)
    return set_err(ctx, MPSD_ERR_ARG, "env_left: output aliases an input");
  const int Db = dim(envL, 0), wl = dim(envL, 1), Dk = dim(envL, 2);
  const int d = dim(bra, 0), Dbl = dim(bra, 1), Dbr = dim(bra, 2);
  const int dp = dim(ket, 0), Dkl = dim(ket, 1), Dkr = dim(ket, 2);
  const int du = dim(W, 0), al1 = dim(W, 1), dd = dim(W, 2), wr = dim(W, 3);
  // Operator::contractleftop dimension check (Operator.cpp:283-289)
  if (d != du || dp != dd || Dkl != Dk || Dbl != Db || al1 != wl || d != dp)
    return set_err(ctx, MPSD_ERR_SHAPE,
                   "env_left: wrong dimensions: bra " + dims_str(bra) + ", ket " + dims_str(ket) +
                       ", env " + dims_str(envL) + ", W " + dims_str(W));
  double2* mix = nullptr;
  int rc = get_mix(ctx, W, 0, d, wl, wr, &mix);
  if (rc) return rc;
  const int odims[3] = {Dbr, wr, Dkr};
  rc = mpsd_tensor_reshape(out, 3, odims);
  if (rc) return rc;
  const double flops = 8.0 * ((double)d * wl * Db * Dk * Dkr + (double)d * d * wl * wr * Db * Dkr +
                              (double)d * wr * Db * Dbr * Dkr);
  Chain chain;
  const bool one = env_chain_fits(ctx, flops);
  rc = left_stages(ctx, envL->ptr, Db, wl, Dk, ket->ptr, d, Dkr, mix, wr, 0, one ? &chain : nullptr);
  if (rc) return rc;
  // L'(c, n = be + wr*ar) = sum_{s'} sum_b conj(bra[s', b, c]) T2[b + Db*s' + Db*d*n]
  ZgemmDesc g;
  g.A = bra->ptr;
  g.B = (const double2*)ctx->scratch[1];
  g.C = out->ptr;
  g.M = Dbr;
  g.N = wr * Dkr;
  g.K = Db;
  g.R = d;
  g.NB = 1;
  g.sAm = (int64_t)d * Db;
  g.sAk = d;
  g.sAr = 1;
  g.sBk = 1;
  g.sBn = (int64_t)Db * d;
  g.sBr = Db;
  g.sCm = 1;
  g.sCn = Dbr;
  g.conjA = 1;
  g.a_mfast = 0;
  g.b_kfast = 1;
  if (one) {
    chain.step[2].kind = 0;
    chain.step[2].g = g;
    chain.nsteps = 3;
    return hip_err(ctx, timed(ctx, MPSD_KERNEL_ENV_STAGE3, [&] { return launch_zchain_f64(chain, ctx->stream); }),
                   "env_left chain");
  }
  return hip_err(ctx,
                 timed(ctx, MPSD_KERNEL_ENV_STAGE3,
                       [&] { return launch_zbrgemm_f64(g, tile_for(ctx, g), ctx->stream); }),
                 "env_left stage3");
}

int mpsd_env_right(mpsd_ctx* ctx, const mpsd_tensor* envR, const mpsd_tensor* ket,
                   const mpsd_tensor* W, const mpsd_tensor* bra, mpsd_tensor* out) {
  if (!ctx || !envR || !ket || !W || !bra || !out) return set_err(ctx, MPSD_ERR_ARG, "env_right: null argument");
  if (!rank_ok(envR, 3) || !rank_ok(ket, 3) || !rank_ok(W, 4) || !rank_ok(bra, 3))
    return set_err(ctx, MPSD_ERR_SHAPE, "env_right: rank too high");
  if (out == envR || out == ket || out == bra || out == W)
    return set_err(ctx, MPSD_ERR_ARG, "env_right: output aliases an input");
  const int Dbr = dim(envR, 0), wr = dim(envR, 1), Dkr = dim(envR, 2);
  const int d = dim(bra, 0), Dbl = dim(bra, 1), Dbr2 = dim(bra, 2);
  const int dp = dim(ket, 0), Dkl = dim(ket, 1), Dkr2 = dim(ket, 2);
  const int d1 = dim(W, 0), wl = dim(W, 1), d2 = dim(W, 2), al2 = dim(W, 3);
  // Operator::contractrightop dimension check (Operator.cpp:344-352)
  if (dp != d2 || d != d1 || Dkr2 != Dkr || Dbr2 != Dbr || al2 != wr || d != dp)
    return set_err(ctx, MPSD_ERR_SHAPE,
                   "env_right: wrong dimensions: bra " + dims_str(bra) + ", ket " + dims_str(ket) +
                       ", env " + dims_str(envR) + ", W " + dims_str(W));
  double2* mix = nullptr;
  int rc = get_mix(ctx, W, 1, d, wl, wr, &mix);
  if (rc) return rc;
  const int odims[3] = {Dbl, wl, Dkl};
  rc = mpsd_tensor_reshape(out, 3, odims);
  if (rc) return rc;
  const size_t u1 = sizeof(double2) * (size_t)Dbr * wr * d * Dkl;
  const size_t u2 = sizeof(double2) * (size_t)Dbr * d * wl * Dkl;
  rc = ensure_scratch(ctx, 0, u1);
  if (rc) return rc;
  rc = ensure_scratch(ctx, 1, u2);
  if (rc) return rc;
  // U1[c, be, s, al] = sum_ar R[c, be, ar] ket[s, al, ar]
  ZgemmDesc g;
  g.A = envR->ptr;
  g.B = ket->ptr;
  g.C = (double2*)ctx->scratch[0];
  g.M = Dbr * wr;
  g.N = Dkl;
  g.K = Dkr;
  g.NB = d;
  g.sAm = 1;
  g.sAk = (int64_t)Dbr * wr;
  g.sAb = 0;
  g.sBk = (int64_t)d * Dkl;
  g.sBn = d;
  g.sBb = 1;
  g.sCm = 1;
  g.sCn = (int64_t)Dbr * wr * d;
  g.sCb = (int64_t)Dbr * wr;
  g.a_mfast = 1;
  g.b_kfast = 0;
  const double flops = 8.0 * ((double)d * wr * Dbr * Dkr * Dkl + (double)d * d * wl * wr * Dbr * Dkl +
                              (double)d * wl * Dbr * Dbl * Dkl);
  const bool one = env_chain_fits(ctx, flops);
  Chain chain;
  if (one) {
    // stage 1 and the mix as the first two steps of one launch
    chain.step[0].kind = 0;
    chain.step[0].g = g;
    ChainStep& mx = chain.step[1];
    mx.kind = 1;
    mx.in = (const double2*)ctx->scratch[0];
    mx.out = (double2*)ctx->scratch[1];
    mx.mix = mix;
    mx.Bx = Dbr;
    mx.nin = wr * d;
    mx.nout = d * wl;
    mx.Ax = Dkl;
  } else {
    hipError_t e = timed(ctx, MPSD_KERNEL_STAGE1,
                         [&] { return launch_zbrgemm_f64(g, tile_for(ctx, g), ctx->stream); });
    if (e != hipSuccess) return hip_err(ctx, e, "env_right stage1");
    // U2[c, s', al, a] = sum_{be,s} W[s', al, s, be] U1[c, be, s, a]
    e = timed(ctx, MPSD_KERNEL_MIX, [&] {
      return launch_zmix_f64((const double2*)ctx->scratch[0], (double2*)ctx->scratch[1], mix, Dbr,
                             wr * d, d * wl, Dkl, ctx->stream);
    });
    if (e != hipSuccess) return hip_err(ctx, e, "env_right mix");
  }
  // R'(b, n = al + wl*a) = sum_{s'} sum_c conj(bra[s', b, c]) U2[c + Dbr*s' + Dbr*d*n]
  ZgemmDesc h;
  h.A = bra->ptr;
  h.B = (const double2*)ctx->scratch[1];
  h.C = out->ptr;
  h.M = Dbl;
  h.N = wl * Dkl;
  h.K = Dbr;
  h.R = d;
  h.NB = 1;
  h.sAm = d;
  h.sAk = (int64_t)d * Dbl;
  h.sAr = 1;
  h.sBk = 1;
  h.sBn = (int64_t)Dbr * d;
  h.sBr = Dbr;
  h.sCm = 1;
  h.sCn = Dbl;
  h.conjA = 1;
  h.a_mfast = 1;
  h.b_kfast = 1;
  if (one) {
    chain.step[2].kind = 0;
    chain.step[2].g = h;
    chain.nsteps = 3;
    return hip_err(ctx, timed(ctx, MPSD_KERNEL_ENV_STAGE3, [&] { return launch_zchain_f64(chain, ctx->stream); }),
                   "env_right chain");
  }
  return hip_err(ctx,
                 timed(ctx, MPSD_KERNEL_ENV_STAGE3,
                       [&] { return launch_zbrgemm_f64(h, tile_for(ctx, h), ctx->stream); }),
                 "env_right stage3");
}

// The packed FP64 matvec's first stage reads x in place up to D = 256 (Dk Dkr values per
// physical index): 5-11% faster per matvec there than packing x first, the same bits.  Above,
// the two are within 2% and the packed x keeps the first stage's loads contiguous.
constexpr int64_t kPackXMin = 256 * 256;

// Stamps of Heff handles (mpsd_heff::lp_stamp): unique over the process.
static std::atomic<uint64_t> g_lp_stamp{0};

int mpsd_heff_create(mpsd_ctx* ctx, const mpsd_tensor* envL, const mpsd_tensor* W,
                     const mpsd_tensor* envR, double offset, int precision, mpsd_heff** out) {
  if (!ctx || !envL || !W || !envR || !out) return set_err(ctx, MPSD_ERR_ARG, "heff_create: null argument");
  *out = nullptr;
  if (!rank_ok(envL, 3) || !rank_ok(envR, 3) || !rank_ok(W, 4))
    return set_err(ctx, MPSD_ERR_SHAPE, "heff_create: rank too high");
  if (precision < MPSD_PREC_FP64 || precision > MPSD_PREC_BF16X3)
    return set_err(ctx, MPSD_ERR_ARG, "heff_create: unknown precision");
  mpsd_heff* h = new mpsd_heff();
  h->ctx = ctx;
  h->L = envL;
  h->W = W;
  h->R = envR;
  h->offset = offset;
  h->precision = precision;
  h->lp_stamp = ++g_lp_stamp;
  h->Db = dim(envL, 0);
  h->wl = dim(envL, 1);
  h->Dk = dim(envL, 2);
  h->d = dim(W, 0);
  h->Dbr = dim(envR, 0);
  h->wr = dim(envR, 1);
  h->Dkr = dim(envR, 2);
  // TensorMultiplier CHECKDIMS (TensorMultiplier.cpp:22-28) plus square physical index
  if (dim(W, 1) != h->wl || dim(W, 3) != h->wr || dim(W, 2) != h->d) {
    delete h;
    return set_err(ctx, MPSD_ERR_SHAPE, "heff_create: wrong dimensions: L " + dims_str(envL) + ", W " +
                                            dims_str(W) + ", R " + dims_str(envR));
  }
  if (offset != 0.0 && (h->Db != h->Dk || h->Dbr != h->Dkr)) {
    delete h;
    return set_err(ctx, MPSD_ERR_SHAPE, "heff_create: offset needs equal bra and ket dimensions");
  }
  int rc = get_mix(ctx, W, 0, h->d, h->wl, h->wr, &h->mix);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return MPSD_OK;
}

int mpsd_heff_destroy(mpsd_heff* h) {
  if (!h) return MPSD_ERR_ARG;
  heff_drop_hermitian(h);  // the Hermitian form's copies (hipFree waits for their last use)
  delete h;  // otherwise owns no device memory (the mix matrix belongs to W's cache)
  return MPSD_OK;
}

int64_t mpsd_heff_size(const mpsd_heff* h) {
  return h ? (int64_t)h->d * h->Dk * h->Dkr : (int64_t)MPSD_ERR_ARG;
}

int mpsd_heff_set_precision(mpsd_heff* h, int precision) {
  if (!h) return MPSD_ERR_ARG;
  if (precision < MPSD_PREC_FP64 || precision > MPSD_PREC_BF16X3)
    return set_err(h->ctx, MPSD_ERR_ARG, "heff_set_precision: unknown precision");
  h->precision = precision;
  h->lp_stamp = ++g_lp_stamp;  // the packed operands are taken afresh
  return MPSD_OK;
}

// Matvec on packed operands: L and R packed once per handle stamp (scratch slots 9, 10), x
// packed per call (slot 11), T1 (slot 0), T2 packed row-interleaved by the mix (slot 1), the per-MPO-bond partial
// products (slot 0 again, T1 being dead by then), summed into y with the offset term.
//  - FP64 (taken only where the plain path splits its last stage, so the sums are the same):
//    packed FP64 rows in the k grouping of zbrgemm_f64's MFMAs, T1 and partials in FP64, the mix
//    and the part sum with zmix_f64's and zsum_parts_f64's arithmetic: bit-identical to the
//    plain path.
//  - FP32 / split-BF16: packed 8-byte values, T1 and partials in FP32, summed in FP64.
static int heff_apply_packed(mpsd_heff* h, const double2* x, double2* y, double offset) {
  mpsd_ctx* ctx = h->ctx;
  const int prec = h->precision;
  const bool f64 = prec == MPSD_PREC_FP64;
  const int d = h->d, Db = h->Db, wl = h->wl, Dk = h->Dk, Dbr = h->Dbr, wr = h->wr, Dkr = h->Dkr;
  const int M = Db * wl;
  const int Kp1 = lp_pad(Dk, prec), Kp3 = lp_pad(Dkr, prec);
  const int64_t ny = (int64_t)d * Db * Dbr;
  const size_t es = f64 ? 16 : 8;  // bytes per packed value, and per T1 / partial value
  const size_t bL = es * M * Kp1, bR = es * wr * Dbr * Kp3, bX = es * d * Dkr * Kp1;
  const size_t bT1 = es * M * d * Dkr, bP = es * wr * ny;
  const size_t bT2 = es * d * wr * Db * Kp3;
  if (ctx->scratch_bytes[9] < bL || ctx->scratch_bytes[10] < bR) ctx->lp_stamp = 0;  // slots reallocated
  int rc = ensure_scratch(ctx, 0, std::max(bT1, bP));
  if (!rc) rc = ensure_scratch(ctx, 1, bT2);
  if (!rc) rc = ensure_scratch(ctx, 9, bL);
  if (!rc) rc = ensure_scratch(ctx, 10, bR);
  if (!rc) rc = ensure_scratch(ctx, 11, bX);
  if (rc) return rc;
  hipError_t e = hipSuccess;
  if (ctx->lp_stamp != h->lp_stamp || ctx->lp_prec != prec) {
    // L as the first stage's A (rows b + Db al, k = a); R as the last stage's B per MPO bond
    // (rows c, k = a_r)
    LpPackDesc pl;
    pl.src = h->L->ptr;
    pl.dst = ctx->scratch[9];
    pl.rows = M;
    pl.K = Dk;
    pl.Kp = Kp1;
    pl.srow = 1;
    pl.sk = M;
    LpPackDesc pr;
    pr.src = h->R->ptr;
    pr.dst = ctx->scratch[10];
    pr.nb = wr;
    pr.rows = Dbr;
    pr.K = Dkr;
    pr.Kp = Kp3;
    pr.sb = Dbr;
    pr.srow = 1;
    pr.sk = (int64_t)Dbr * wr;
    e = timed(ctx, MPSD_KERNEL_PACK, [&] {
      hipError_t r = launch_lp_pack(pl, prec, ctx->stream);
      return r == hipSuccess ? launch_lp_pack(pr, prec, ctx->stream) : r;
    });
    if (e != hipSuccess) return hip_err(ctx, e, "heff pack environments");
    ctx->lp_stamp = h->lp_stamp;
    ctx->lp_prec = prec;
  }
  // x[s + d a + d Dk a_r] as the first stage's B per s (rows a_r, k = a)
  LpPackDesc px;
  px.src = x;
  px.dst = ctx->scratch[11];
  px.nb = d;
  px.rows = Dkr;
  px.K = Dk;
  px.Kp = Kp1;
  px.sb = 1;
  px.srow = (int64_t)d * Dk;
  px.sk = d;
  // T1[b + Db al + M s + M d a_r] = sum_a L[b, al, a] x[s, a, a_r]
  LpDesc g1;
  g1.A = ctx->scratch[9];
  g1.B = ctx->scratch[11];
  g1.C = (float2*)ctx->scratch[0];
  g1.Cd = (double2*)ctx->scratch[0];
  g1.M = M;
  g1.N = Dkr;
  g1.Kp = Kp1;
  g1.NB = d;
  g1.zdiv = d;
  g1.prec = prec;
  g1.sB0 = (int64_t)Dkr * Kp1;
  g1.sCm = 1;
  g1.sCn = (int64_t)M * d;
  g1.sC0 = M;
  // FP64: stage 1 reads x in place (same values, so the same bits) unless packing it pays
  const bool x_ip = f64 && (ctx->pk64_x == 0 || (ctx->pk64_x < 0 && (int64_t)Dk * Dkr <= kPackXMin));
  if (x_ip) {
    g1.B = x;
    g1.b_ip = 1;
    g1.Kb = Dk;
    g1.sB0 = 1;
    g1.sBk = d;
    g1.sBn = (int64_t)d * Dk;
  } else {
    e = timed(ctx, MPSD_KERNEL_PACK, [&] { return launch_lp_pack(px, prec, ctx->stream); });
    if (e != hipSuccess) return hip_err(ctx, e, "heff pack x");
  }
  e = timed(ctx, MPSD_KERNEL_STAGE1, [&] { return launch_lp_brgemm(g1, lp_tile_for(ctx, g1), ctx->stream); });
  if (e != hipSuccess) return hip_err(ctx, e, "heff packed stage1");
  // T2[o = s' + d be][b][a_r] = sum_{al, s} W[s', al, s, be] T1[b, al, s, a_r], packed
  e = timed(ctx, MPSD_KERNEL_MIX, [&] {
    return f64 ? launch_pk64_mix((const double2*)ctx->scratch[0], (double2*)ctx->scratch[1], h->mix, Db, wl * d,
                                 d * wr, Dkr, Kp3, ctx->stream)
               : launch_lp_mix((const float2*)ctx->scratch[0], ctx->scratch[1], h->mix, Db, wl * d, d * wr, Dkr, Kp3,
                               prec, ctx->stream);
  });
  if (e != hipSuccess) return hip_err(ctx, e, "heff packed mix");
  // P[be][s' + d b + d Db c] = sum_{a_r} T2[s' + d be][b][a_r] R[c, be, a_r]
  LpDesc g3;
  g3.A = ctx->scratch[1];
  g3.B = ctx->scratch[10];
  g3.C = (float2*)ctx->scratch[0];
  g3.Cd = (double2*)ctx->scratch[0];
  g3.M = Db;
  g3.N = Dbr;
  g3.Kp = Kp3;
  g3.NB = d * wr;
  g3.zdiv = d;
  g3.prec = prec;
  g3.sA0 = (int64_t)Db * Kp3;
  g3.sA1 = (int64_t)d * Db * Kp3;
  g3.sB1 = (int64_t)Dbr * Kp3;
  g3.sCm = d;
  g3.sCn = (int64_t)d * Db;
  g3.sC0 = 1;
  g3.sC1 = ny;
  g3.a_il = 1;  // as the mix writes it
  e = timed(ctx, MPSD_KERNEL_HEFF_STAGE3, [&] { return launch_lp_brgemm(g3, lp_tile_for(ctx, g3), ctx->stream); });
  if (e != hipSuccess) return hip_err(ctx, e, "heff packed stage3");
  e = timed(ctx, MPSD_KERNEL_MIX, [&] {
    return f64 ? launch_zsum_parts_f64((const double2*)ctx->scratch[0], wr, ny, offset != 0.0 ? x : nullptr, offset,
                                       y, ny, ctx->stream)
               : launch_lp_sum((const float2*)ctx->scratch[0], wr, ny, offset != 0.0 ? x : nullptr, offset, y, ny,
                               ctx->stream);
  });
  return hip_err(ctx, e, "heff packed sum");
}

// One device matvec y = Heff x (+ offset x) on device pointers.  With s0, s1 set, T1 and the
// partial products go to s0 and T2 to s1 (sized by the caller), and h->mix must already hold
// the mix matrix: nothing is allocated.  Otherwise the context's scratch slots and W's cached
// mix matrix are used (and grown or built on first use).
static int heff_apply_core(mpsd_heff* h, const double2* x, double2* y, double offset, double2* s0 = nullptr,
                           double2* s1 = nullptr) {
  mpsd_ctx* ctx = h->ctx;
  const bool fixed = s0 && s1;
  int rc = MPSD_OK;
  if (!fixed) rc = get_mix(ctx, h->W, 0, h->d, h->wl, h->wr, &h->mix);  // pointer check when cached
  if (rc) return rc;
  if (!fixed && h->precision != MPSD_PREC_FP64 && ctx->lp_packed) return heff_apply_packed(h, x, y, offset);
  if (!fixed && h->precision == MPSD_PREC_FP64 && ctx->pk64 && stage3_split(h, ctx))
    return heff_apply_packed(h, x, y, offset);
  const int arith = arith_of(h->precision);
  rc = left_stages(ctx, h->L->ptr, h->Db, h->wl, h->Dk, x, h->d, h->Dkr, h->mix, h->wr, arith, nullptr, s0, s1);
  if (rc) return rc;
  if (!fixed) {
    s0 = (double2*)ctx->scratch[0];
    s1 = (double2*)ctx->scratch[1];
  }
  const int64_t ny = (int64_t)h->d * h->Db * h->Dbr;
  // Split the last stage over the MPO bond when its plain grid would leave CUs idle: each
  // beta gives a partial product of the same shape as stage 1, summed in a fixed order.
  const bool split = stage3_split(h, ctx);
  if (split) {
    if (!fixed) {
      rc = ensure_scratch(ctx, 0, sizeof(double2) * (size_t)(ny * h->wr));
      if (rc) return rc;
      s0 = (double2*)ctx->scratch[0];
    }
    double2* P = s0;  // T1 is dead once the mix has run
    // P[be][s' + d*b + d*Db*c] = sum_ar T2[b + Db*s' + Db*d*be + Db*d*wr*ar] R[c + Dbr*be + Dbr*wr*ar]
    ZgemmDesc g;
    g.A = s1;
    g.B = h->R->ptr;
    g.C = P;
    g.M = h->Db;
    g.N = h->Dbr;
    g.K = h->Dkr;
    g.NB = h->d * h->wr;
    g.group = h->d;
    g.sAm = 1;
    g.sAk = (int64_t)h->Db * h->d * h->wr;
    g.sAb = h->Db;
    g.sAz = (int64_t)h->Db * h->d;
    g.sBk = (int64_t)h->Dbr * h->wr;
    g.sBn = 1;
    g.sBb = 0;
    g.sBz = h->Dbr;
    g.sCm = h->d;
    g.sCn = (int64_t)h->d * h->Db;
    g.sCb = 1;
    g.sCz = ny;
    g.a_mfast = 1;
    g.b_kfast = 0;
    g.arith = arith;
    hipError_t e = timed(ctx, MPSD_KERNEL_HEFF_STAGE3,
                         [&] { return launch_zbrgemm_f64(g, tile_for(ctx, g), ctx->stream); });
    if (e != hipSuccess) return hip_err(ctx, e, "heff stage3 split");
    e = timed(ctx, MPSD_KERNEL_MIX, [&] {
      return launch_zsum_parts_f64(P, h->wr, ny, offset != 0.0 ? x : nullptr, offset, y, ny, ctx->stream);
    });
    return hip_err(ctx, e, "heff stage3 sum");
  }
  // y[s', b, c] = sum_{k = be + wr*ar} T2[b + Db*s' + Db*d*k] R[c + Dbr*k]
  ZgemmDesc g;
  g.A = s1;
  g.B = h->R->ptr;
  g.C = y;
  g.M = h->Db;
  g.N = h->Dbr;
  g.K = h->wr * h->Dkr;
  g.NB = h->d;
  g.sAm = 1;
  g.sAk = (int64_t)h->Db * h->d;
  g.sAb = h->Db;
  g.sBk = h->Dbr;
  g.sBn = 1;
  g.sBb = 0;
  g.sCm = h->d;
  g.sCn = (int64_t)h->d * h->Db;
  g.sCb = 1;
  g.a_mfast = 1;
  g.b_kfast = 0;
  g.arith = arith;
  if (offset != 0.0) {
    g.Dadd = x;
    g.beta = offset;
  }
  return hip_err(ctx,
                 timed(ctx, MPSD_KERNEL_HEFF_STAGE3,
                       [&] { return launch_zbrgemm_f64(g, tile_for(ctx, g), ctx->stream); }),
                 "heff stage3");
}

__global__ void __launch_bounds__(256) herm_combine_kernel(double2* __restrict__ y, const double2* __restrict__ yh,
                                                           const double2* __restrict__ x, double offset, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    // TensorMultiplierHermitian(Offset)::product: result = .5 * (result + aux), then + offset * input
    double2 v = y[i];
    const double2 a = yh[i];
    v.x = 0.5 * (v.x + a.x);
    v.y = 0.5 * (v.y + a.y);
    if (offset != 0.0) {
      const double2 xi = x[i];
      v.x = fma(offset, xi.x, v.x);
      v.y = fma(offset, xi.y, v.y);
    }
    y[i] = v;
  }
}

// One device matvec y = Heff x (+ offset x) on device pointers; the Hermitian form adds the
// product over the conjugate-transposed operands and averages (TensorMultiplierHermitian.cpp:36-66).
static int heff_apply_dev(mpsd_heff* h, const double2* x, double2* y) {
  if (!h->herm) return heff_apply_core(h, x, y, h->offset);
  mpsd_ctx* ctx = h->ctx;
  h->hh->ctx = ctx;
  h->hh->precision = h->precision;
  int rc = heff_apply_core(h, x, y, 0.0);
  if (!rc) rc = heff_apply_core(h->hh, x, h->htmp, 0.0);
  if (rc) return rc;
  const int64_t n = (int64_t)h->d * h->Db * h->Dbr;
  const unsigned blocks = (unsigned)std::min<int64_t>(2048, (n + 255) / 256);
  hipLaunchKernelGGL(herm_combine_kernel, dim3(blocks), dim3(256), 0, ctx->stream, y, h->htmp, x, h->offset, n);
  return hip_err(ctx, hipGetLastError(), "heff hermitian combine");
}

int mpsd_heff_apply_noalloc(mpsd_ctx* ctx, const mpsd_tensor* envL, const mpsd_tensor* W, const mpsd_tensor* envR,
                            double offset, const double* x, double* y) {
  if (!ctx || !envL || !W || !envR || !x || !y) return set_err(ctx, MPSD_ERR_ARG, "heff_apply_noalloc: null argument");
  if (!rank_ok(envL, 3) || !rank_ok(envR, 3) || !rank_ok(W, 4))
    return set_err(ctx, MPSD_ERR_SHAPE, "heff_apply_noalloc: rank too high");
  // a handle on the stack: the shapes and operands of mpsd_heff_create, no device memory
  mpsd_heff h;
  h.ctx = ctx;
  h.L = envL;
  h.W = W;
  h.R = envR;
  h.offset = offset;
  h.precision = MPSD_PREC_FP64;
  h.Db = dim(envL, 0);
  h.wl = dim(envL, 1);
  h.Dk = dim(envL, 2);
  h.d = dim(W, 0);
  h.Dbr = dim(envR, 0);
  h.wr = dim(envR, 1);
  h.Dkr = dim(envR, 2);
  if (dim(W, 1) != h.wl || dim(W, 3) != h.wr || dim(W, 2) != h.d)
    return set_err(ctx, MPSD_ERR_SHAPE, "heff_apply_noalloc: wrong dimensions: L " + dims_str(envL) + ", W " +
                                            dims_str(W) + ", R " + dims_str(envR));
  if (offset != 0.0 && (h.Db != h.Dk || h.Dbr != h.Dkr))
    return set_err(ctx, MPSD_ERR_SHAPE, "heff_apply_noalloc: offset needs equal bra and ket dimensions");
  auto up = [](int64_t n) { return (n + 15) & ~int64_t(15); };  // complex elements, 256-byte steps
  const int64_t nmix = (int64_t)h.d * h.wl * h.d * h.wr;
  const int64_t ny = (int64_t)h.d * h.Db * h.Dbr;
  const int64_t t1 = (int64_t)h.Db * h.wl * h.d * h.Dkr;
  const int64_t t2 = (int64_t)h.Db * h.d * h.wr * h.Dkr;
  const bool split = stage3_split(&h, ctx);
  const int64_t s0n = std::max(t1, split ? ny * h.wr : int64_t(0));
  const int64_t need = (up(nmix) + up(s0n) + up(t2)) * (int64_t)sizeof(double2);
  if (need > (int64_t)kNoAllocBytes)
    return set_err(ctx, MPSD_ERR_UNSUPPORTED,
                   "heff_apply_noalloc: needs " + std::to_string(need) + " bytes of scratch, above the " +
                       std::to_string(kNoAllocBytes) + "-byte static buffer (e.g. D = 512 with w = 5): use "
                       "mpsd_heff_apply");
  double2* base = (double2*)noalloc_scratch(ctx->device);
  if (!base) return set_err(ctx, MPSD_ERR_HIP, "heff_apply_noalloc: scratch address not set up on this device");
  double2* mix = base;
  double2* s0 = base + up(nmix);
  double2* s1 = s0 + up(s0n);
  // the mix matrix into the scratch (kind 0, as get_mix builds it for a matvec)
  hipLaunchKernelGGL(mix_perm_kernel, dim3((unsigned)((nmix + 255) / 256)), dim3(256), 0, ctx->stream, W->ptr, h.d,
                     h.wl, h.wr, 0, mix);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_err(ctx, e, "heff_apply_noalloc mix");
  h.mix = mix;
  return heff_apply_core(&h, (const double2*)x, (double2*)y, offset, s0, s1);
}

int mpsd_heff_apply(mpsd_heff* h, const double* x, double* y, int nvec, int64_t ldx, int64_t ldy,
                    int on_device) {
  if (!h || !x || !y || nvec < 0) return MPSD_ERR_ARG;
  mpsd_ctx* ctx = h->ctx;
  const int64_t n = mpsd_heff_size(h);
  const int64_t ny = (int64_t)h->d * h->Db * h->Dbr;
  if (ldx < n || ldy < ny) return set_err(ctx, MPSD_ERR_ARG, "heff_apply: leading dimension too small");
  if (on_device) {
    for (int v = 0; v < nvec; ++v) {
      int rc = heff_apply_dev(h, (const double2*)x + v * ldx, (double2*)y + v * ldy);
      if (rc) return rc;
    }
    return MPSD_OK;
  }
  int rc = ensure_stage(ctx, 0, sizeof(double2) * (size_t)n);
  if (!rc) rc = ensure_stage(ctx, 1, sizeof(double2) * (size_t)ny);
  if (rc) return rc;
  for (int v = 0; v < nvec; ++v) {
    hipError_t e = hipMemcpyAsync(ctx->stage[0], (const double2*)x + v * ldx, sizeof(double2) * n,
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return hip_err(ctx, e, "heff_apply H2D");
    rc = heff_apply_dev(h, (const double2*)ctx->stage[0], (double2*)ctx->stage[1]);
    if (rc) return rc;
    e = hipMemcpyAsync((double2*)y + v * ldy, ctx->stage[1], sizeof(double2) * ny,
                       hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return hip_err(ctx, e, "heff_apply D2H");
  }
  return hip_err(ctx, hipStreamSynchronize(ctx->stream), "heff_apply sync");
}

}  // extern "C"

// Decide the last-stage split from the problem shape only (the same rule as heff_apply_dev),
// so a problem takes the same arithmetic path alone and inside a group.
namespace mpsd {
bool stage3_split(const mpsd_heff* h, const mpsd_ctx* ctx) {
  if (h->d != 2 || h->wr < 2) return false;
  if (ctx->split3 == 1) return true;
  if (ctx->split3 == 0) return false;
  const int64_t tiles = (int64_t)((h->Db + 31) / 32) * ((h->Dbr + 31) / 32);
  return tiles < 512;
}
}  // namespace mpsd

static int upload_tables(mpsd_ctx* ctx, const std::vector<char>& host, void** dev) {
  if (ctx->gtab_bytes < host.size()) {
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->gtab) (void)hipFree(ctx->gtab);
    if (ctx->gtab_host) (void)hipHostFree(ctx->gtab_host);
    ctx->gtab = ctx->gtab_host = nullptr;
    ctx->gtab_bytes = 0;
    const size_t want = host.size() * 2 + 4096;
    hipError_t e = hipMalloc(&ctx->gtab, want);
    if (e == hipSuccess) e = hipHostMalloc(&ctx->gtab_host, want, hipHostMallocDefault);
    if (e == hipSuccess && !ctx->gtab_ev) e = hipEventCreateWithFlags(&ctx->gtab_ev, hipEventDisableTiming);
    if (e != hipSuccess) return hip_err(ctx, e, "group table alloc");
    ctx->gtab_bytes = want;
  }
  hipError_t e = hipEventSynchronize(ctx->gtab_ev);  // previous upload done with the staging buffer
  if (e != hipSuccess) return hip_err(ctx, e, "group table wait");
  memcpy(ctx->gtab_host, host.data(), host.size());
  e = hipMemcpyAsync(ctx->gtab, ctx->gtab_host, host.size(), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ctx->gtab_ev, ctx->stream);
  *dev = ctx->gtab;
  return hip_err(ctx, e, "group table upload");
}

template <class T>
static size_t append(std::vector<char>& buf, const std::vector<T>& v) {
  const size_t off = (buf.size() + 255) & ~(size_t)255;
  buf.resize(off + sizeof(T) * v.size());
  if (!v.empty()) memcpy(buf.data() + off, v.data(), sizeof(T) * v.size());
  return off;
}

static int grouped_apply(mpsd_ctx* ctx, int n, mpsd_heff* const* hs, const double* const* xs, double* const* ys,
                         int arith) {
  // per-problem scratch: T1 (+ partials) in slot 4, T2 in slot 5
  std::vector<int64_t> o1(n), o2(n), op(n);
  int64_t t1 = 0, t2 = 0;
  for (int i = 0; i < n; ++i) {
    const mpsd_heff* h = hs[i];
    const int64_t ny = (int64_t)h->d * h->Db * h->Dbr;
    o1[i] = t1;
    t1 += (int64_t)h->Db * h->wl * h->d * h->Dkr;
    op[i] = t1;
    if (stage3_split(h, ctx)) t1 += ny * h->wr;
    o2[i] = t2;
    t2 += (int64_t)h->Db * h->d * h->wr * h->Dkr;
  }
  int rc = ensure_scratch(ctx, 4, sizeof(double2) * (size_t)std::max<int64_t>(t1, 1));
  if (!rc) rc = ensure_scratch(ctx, 5, sizeof(double2) * (size_t)std::max<int64_t>(t2, 1));
  if (rc) return rc;
  double2* S1 = (double2*)ctx->scratch[4];
  double2* S2 = (double2*)ctx->scratch[5];
  const int TM = 32, TN = 32;
  auto tiles = [&](GroupedProblem& g, int& begin) {
    g.gm = (g.desc.M + TM - 1) / TM;
    g.gn = (g.desc.N + TN - 1) / TN;
    g.gz = (g.desc.NB + g.desc.group - 1) / g.desc.group;
    g.ntiles = (g.desc.M > 0 && g.desc.N > 0) ? g.gm * g.gn * g.gz : 0;
    g.tile_begin = begin;
    begin += g.ntiles;
  };
  std::vector<GroupedProblem> st1, st3s, st3n;
  std::vector<GroupedElem> mx, sm;
  int b1 = 0, b3s = 0, b3n = 0;
  int64_t bm = 0, bs = 0;
  for (int i = 0; i < n; ++i) {
    mpsd_heff* h = hs[i];
    const double2* x = (const double2*)xs[i];
    double2* y = (double2*)ys[i];
    const int64_t ny = (int64_t)h->d * h->Db * h->Dbr;
    // stage 1, exactly as left_stages
    GroupedProblem g;
    ZgemmDesc& d = g.desc;
    d.A = h->L->ptr; d.B = x; d.C = S1 + o1[i];
    d.M = h->Db * h->wl; d.N = h->Dkr; d.K = h->Dk; d.NB = h->d;
    d.sAm = 1; d.sAk = (int64_t)h->Db * h->wl; d.sAb = 0;
    d.sBk = h->d; d.sBn = (int64_t)h->d * h->Dk; d.sBb = 1;
    d.sCm = 1; d.sCn = (int64_t)h->Db * h->wl * h->d; d.sCb = (int64_t)h->Db * h->wl;
    d.a_mfast = 1; d.b_kfast = 1; d.arith = arith;
    choose_group(d, ctx->tile_override.nbw);
    tiles(g, b1);
    st1.push_back(g);
    // mix
    GroupedElem e;
    e.in = S1 + o1[i]; e.out = S2 + o2[i]; e.aux = h->mix;
    e.Bx = h->Db; e.nin = h->wl * h->d; e.nout = h->d * h->wr; e.Ax = h->Dkr; e.begin = bm;
    bm += (int64_t)e.Bx * e.Ax;
    mx.push_back(e);
    // stage 3
    GroupedProblem g3;
    ZgemmDesc& t = g3.desc;
    t.A = S2 + o2[i]; t.B = h->R->ptr; t.a_mfast = 1; t.b_kfast = 0; t.arith = arith;
    t.M = h->Db; t.N = h->Dbr; t.sAm = 1; t.sBn = 1; t.sBb = 0;
    if (stage3_split(h, ctx)) {
      t.C = S1 + op[i];
      t.K = h->Dkr; t.NB = h->d * h->wr; t.group = h->d;
      t.sAk = (int64_t)h->Db * h->d * h->wr; t.sAb = h->Db; t.sAz = (int64_t)h->Db * h->d;
      t.sBk = (int64_t)h->Dbr * h->wr; t.sBz = h->Dbr;
      t.sCm = h->d; t.sCn = (int64_t)h->d * h->Db; t.sCb = 1; t.sCz = ny;
      resolve_group(t);
      tiles(g3, b3s);
      st3s.push_back(g3);
      GroupedElem su;
      su.in = S1 + op[i]; su.out = y; su.aux = h->offset != 0.0 ? x : nullptr; su.beta = h->offset;
      su.np = h->wr; su.stride = ny; su.n = ny; su.begin = bs;
      bs += ny;
      sm.push_back(su);
    } else {
      t.C = y;
      t.K = h->wr * h->Dkr; t.NB = h->d;
      t.sAk = (int64_t)h->Db * h->d; t.sAb = h->Db;
      t.sBk = h->Dbr;
      t.sCm = h->d; t.sCn = (int64_t)h->d * h->Db; t.sCb = 1;
      if (h->offset != 0.0) {
        t.Dadd = x;
        t.beta = h->offset;
      }
      choose_group(t, ctx->tile_override.nbw);
      tiles(g3, b3n);
      st3n.push_back(g3);
    }
  }
  std::vector<char> buf;
  const size_t f1 = append(buf, st1), fm = append(buf, mx), f3s = append(buf, st3s), f3n = append(buf, st3n),
               fs = append(buf, sm);
  void* dev = nullptr;
  rc = upload_tables(ctx, buf, &dev);
  if (rc) return rc;
  char* D = (char*)dev;
  // persistent grids: as many workgroups as fit at once (one-batch FP64 tiles: six per CU)
  const int persist = 256 * ((!st1.empty() && st1[0].desc.group == 1) ? 6 : 3);
  auto grid_of = [&](int64_t work, int per) { return (int)std::max<int64_t>(1, std::min<int64_t>(persist, (work + per - 1) / per)); };
  const int kind1 = st1.empty() ? 0 : (st1[0].desc.group == 2 ? 1 : 0);
  hipError_t e = timed(ctx, MPSD_KERNEL_STAGE1, [&] {
    return launch_zbrgemm_grouped((const GroupedProblem*)(D + f1), (int)st1.size(), grid_of(b1, 1), TM, TN, 16,
                                  st1.empty() ? 1 : st1[0].desc.group, kind1, arith, ctx->stream);
  });
  if (e == hipSuccess)
    e = timed(ctx, MPSD_KERNEL_MIX,
              [&] { return launch_zmix_grouped((const GroupedElem*)(D + fm), (int)mx.size(), grid_of(bm, 256), ctx->stream); });
  if (e == hipSuccess && !st3s.empty())
    e = timed(ctx, MPSD_KERNEL_HEFF_STAGE3, [&] {
      return launch_zbrgemm_grouped((const GroupedProblem*)(D + f3s), (int)st3s.size(), grid_of(b3s, 1), TM, TN, 16,
                                    2, 2, arith, ctx->stream);
    });
  if (e == hipSuccess && !st3n.empty())
    e = timed(ctx, MPSD_KERNEL_HEFF_STAGE3, [&] {
      return launch_zbrgemm_grouped((const GroupedProblem*)(D + f3n), (int)st3n.size(), grid_of(b3n, 1), TM, TN, 16,
                                    st3n[0].desc.group, st3n[0].desc.group == 2 ? 2 : 0, arith, ctx->stream);
    });
  if (e == hipSuccess && !sm.empty())
    e = timed(ctx, MPSD_KERNEL_MIX,
              [&] { return launch_zsum_grouped((const GroupedElem*)(D + fs), (int)sm.size(), grid_of(bs, 256), ctx->stream); });
  return hip_err(ctx, e, "grouped matvec");
}

extern "C" {

int mpsd_heff_apply_grouped(mpsd_ctx* ctx, int n, mpsd_heff* const* hs, const double* const* x, double* const* y) {
  if (!ctx || n < 0 || (n > 0 && (!hs || !x || !y))) return set_err(ctx, MPSD_ERR_ARG, "apply_grouped: bad arguments");
  if (n == 0) return MPSD_OK;
  std::vector<mpsd_heff*> g[2];
  std::vector<const double*> gx[2];
  std::vector<double*> gy[2];
  for (int i = 0; i < n; ++i) {
    mpsd_heff* h = hs[i];
    if (!h || !x[i] || !y[i]) return set_err(ctx, MPSD_ERR_ARG, "apply_grouped: null handle or vector");
    int rc = get_mix(ctx, h->W, 0, h->d, h->wl, h->wr, &h->mix);
    if (rc) return rc;
    if (h->precision == MPSD_PREC_BF16X3 || (h->precision != MPSD_PREC_FP64 && ctx->lp_packed) || h->d != 2 ||
        h->herm) {
      // the precision forms, d != 2 and the Hermitian form take the single-problem path on this
      // stream (grouped FP32 launches only on FP64 operands converted per tile)
      mpsd_ctx* own = h->ctx;
      h->ctx = ctx;
      rc = heff_apply_dev(h, (const double2*)x[i], (double2*)y[i]);
      h->ctx = own;
      if (rc) return rc;
      continue;
    }
    const int a = arith_of(h->precision);
    g[a].push_back(h);
    gx[a].push_back(x[i]);
    gy[a].push_back(y[i]);
  }
  for (int a = 0; a < 2; ++a) {
    if (g[a].empty()) continue;
    int rc = grouped_apply(ctx, (int)g[a].size(), g[a].data(), gx[a].data(), gy[a].data(), a);
    if (rc) return rc;
  }
  return MPSD_OK;
}

int mpsd_heff_full(mpsd_heff* h, double* host_out) {
  if (!h || !host_out) return MPSD_ERR_ARG;
  mpsd_ctx* ctx = h->ctx;
  const int64_t nr = (int64_t)h->d * h->Db * h->Dbr, nc = mpsd_heff_size(h);
  int rc = ensure_stage(ctx, 0, sizeof(double2) * (size_t)(nr * nc));
  if (rc) return rc;
  hipError_t e = heff_full_dev(h, (double2*)ctx->stage[0]);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host_out, ctx->stage[0], sizeof(double2) * (size_t)(nr * nc), hipMemcpyDeviceToHost,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return hip_err(ctx, e, "heff_full");
}

}  // extern "C"
