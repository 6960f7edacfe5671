// Device-driven Lanczos steps replayed as captured hipGraphs.
//
// Every kernel of a Lanczos step reads what changes from step to step -- the step index j,
// the basis columns V_j (matvec input) and V_{j+1} (output), the vector length -- from a
// small control block in device memory, and the matvec runs as grouped persistent launches
// over problem tables in device memory.  A one-thread "advance" kernel at the head of each
// step moves j forward and patches the tables' x/y pointers.  A graph of k consecutive steps
// therefore has fixed kernel arguments: it is captured once per solver and replayed for
// every chunk of every site, whatever the site's shapes; binding a site only rewrites the
// tables (one small host-to-device copy).  The matvec is the stream-launched path's (same
// tile function, same reduction order); the Gram-Schmidt passes run unfused here (dots, update,
// dots, update), where the stream path fuses them, so the two paths agree to rounding.

#include <algorithm>
#include <cstring>

#include <array>
#include <map>

#include "mpsd_internal.h"
#include "vec_kernels.h"

namespace mpsd {
namespace {

constexpr int kT = kVecThreads;
constexpr int kDotBlocks = kDotMaxBlocks;
constexpr int kPersist = 512;    // most workgroups of a persistent elementwise launch
constexpr int kPersistGemm = 768;  // most workgroups of a persistent BRGEMM launch: 3 per CU (x2 one-batch tiles)
constexpr int kVecBlocks = 512;  // most blocks of a grid-stride vector kernel

// Launch shape of one captured step: grids sized to the bound site's shapes.
typedef std::array<int, 10> GridKey;  // arith, k, st1, mix, st3, sum, vec, dot blocks, st1 group, st3 group

struct LanczosCtl {
  int j, j_next, split, has_offset;
  int n_sum, pad0, pad1, pad2;
  int64_t n, ldv;
  double2* V;
};

// Layout of the device block (and of its pinned host image).
struct Block {
  LanczosCtl ctl;
  alignas(256) GroupedProblem st1;
  alignas(256) GroupedProblem st3;
  alignas(256) GroupedElem mix;
  alignas(256) GroupedElem sum;
};

__global__ void advance_kernel(Block* b) {
  LanczosCtl& c = b->ctl;
  const int j = c.j_next;
  c.j = j;
  c.j_next = j + 1;
  double2* x = c.V + (int64_t)j * c.ldv;
  double2* y = x + c.ldv;
  b->st1.desc.B = x;
  if (c.split) {
    b->sum.out = y;
    b->sum.aux = c.has_offset ? x : nullptr;
  } else {
    b->st3.desc.C = y;
    b->st3.desc.Dadd = c.has_offset ? x : nullptr;
  }
}

// partial sums of conj(V_i) . w for i < j + add, w = V_{j+1} (same blocks as the stream path)
__global__ void __launch_bounds__(kT) dots_ctl_partial(const LanczosCtl* __restrict__ c, int add, int nb,
                                                       double2* __restrict__ partial) {
  const int i = blockIdx.y;
  if (i >= c->j + add) return;
  const double2 r =
      dot_block(c->V + (int64_t)i * c->ldv, c->V + (int64_t)(c->j + 1) * c->ldv, c->n, blockIdx.x, nb);
  if (threadIdx.x == 0) partial[(int64_t)i * kDotBlocks + blockIdx.x] = r;
}

__global__ void __launch_bounds__(kDotBlocks) dots_ctl_final(const LanczosCtl* __restrict__ c, int add, int nb,
                                                             const double2* __restrict__ partial,
                                                             double2* __restrict__ out) {
  const int i = blockIdx.x;
  if (i >= c->j + add) return;
  const double2 r = dot_final(partial + (int64_t)i * kDotBlocks, nb);
  if (threadIdx.x == 0) out[i] = r;
}

// w -= V_{0..j} coef
__global__ void __launch_bounds__(kT) gemv_ctl(const LanczosCtl* __restrict__ c, const double2* __restrict__ coef) {
  const int64_t n = c->n, ldv = c->ldv;
  double2* w = c->V + (int64_t)(c->j + 1) * ldv;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < n; t += (int64_t)gridDim.x * kT)
    gemv_elem(c->V, ldv, c->j + 1, coef, -1.0, w, t);
}

// alpha_j, beta_j and w /= beta_j
__global__ void __launch_bounds__(kT) coef_ctl(const LanczosCtl* __restrict__ c, const double2* __restrict__ h1,
                                               const double2* __restrict__ h2, double* __restrict__ ab) {
  const int j = c->j;
  const int64_t n = c->n;
  double2* w = c->V + (int64_t)(j + 1) * c->ldv;
  double alpha, beta, inv;
  lanczos_scalars(h1, h2, j, alpha, beta, inv);
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double2 v = w[i];
    w[i] = double2{v.x * inv, v.y * inv};
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ab[2 + 2 * j] = alpha;  // ab[0] holds |x|^2 of the start vector
    ab[3 + 2 * j] = beta;
  }
}

}  // namespace
}  // namespace mpsd

using namespace mpsd;

struct mpsd_lanczos {
  mpsd_ctx* ctx = nullptr;
  int mmax = 0;
  Block* dev = nullptr;   // device control block + tables
  Block* host = nullptr;  // pinned image
  hipEvent_t host_ev = nullptr;
  double2* partial = nullptr;
  double2* h1 = nullptr;
  double2* h2 = nullptr;
  double* ab = nullptr;
  double2* V = nullptr;
  int64_t vcap = 0;
  double2* S1 = nullptr;
  double2* S2 = nullptr;
  int64_t s1cap = 0, s2cap = 0;
  std::map<GridKey, hipGraphExec_t> exec;  // captured graphs by launch shape
  GridKey grids{};
  int arith = 0;
  int jn = 0;  // host mirror of the device step counter (the pinned image is never touched
               // after its upload is queued)
  int64_t n = 0;
  int64_t replays = 0;
  std::string err;
};

namespace {

int lz_err(mpsd_lanczos* lz, hipError_t e, const char* where) { return hip_err(lz->ctx, e, where); }

int grow(mpsd_lanczos* lz, double2** p, int64_t* cap, int64_t want) {
  if (*cap >= want) return MPSD_OK;
  hipError_t e = hipStreamSynchronize(lz->ctx->stream);
  if (e != hipSuccess) return lz_err(lz, e, "lanczos grow sync");
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  const int64_t n = want + want / 4;
  e = hipMalloc((void**)p, sizeof(double2) * (size_t)n);
  if (e != hipSuccess) return lz_err(lz, e, "lanczos alloc");
  *cap = n;
  return MPSD_OK;
}

// one Lanczos step as stream launches (captured into the graph)
hipError_t enqueue_step(mpsd_lanczos* lz, const GridKey& g) {
  hipStream_t s = lz->ctx->stream;
  Block* b = lz->dev;
  const int ma = lz->mmax + 2;
  const int arith = g[0], nb = g[7];
  hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, s, b);
  hipError_t e = launch_zbrgemm_grouped(&b->st1, 1, g[2], 32, 32, 16, g[8], g[8] == 2 ? 1 : 0, arith, s);
  if (e == hipSuccess) e = launch_zmix_grouped(&b->mix, 1, g[3], s);
  if (e == hipSuccess) e = launch_zbrgemm_grouped(&b->st3, 1, g[4], 32, 32, 16, g[9], g[9] == 2 ? 2 : 0, arith, s);
  if (e == hipSuccess) e = launch_zsum_grouped(&b->sum, 0, g[5], s, &b->ctl.n_sum);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(dots_ctl_partial, dim3(nb, ma), dim3(kT), 0, s, &b->ctl, 1, nb, lz->partial);
  hipLaunchKernelGGL(dots_ctl_final, dim3(ma), dim3(kDotBlocks), 0, s, &b->ctl, 1, nb, lz->partial, lz->h1);
  hipLaunchKernelGGL(gemv_ctl, dim3(g[6]), dim3(kT), 0, s, &b->ctl, lz->h1);
  hipLaunchKernelGGL(dots_ctl_partial, dim3(nb, ma), dim3(kT), 0, s, &b->ctl, 2, nb, lz->partial);
  hipLaunchKernelGGL(dots_ctl_final, dim3(ma), dim3(kDotBlocks), 0, s, &b->ctl, 2, nb, lz->partial, lz->h2);
  hipLaunchKernelGGL(gemv_ctl, dim3(g[6]), dim3(kT), 0, s, &b->ctl, lz->h2);
  hipLaunchKernelGGL(coef_ctl, dim3(g[6]), dim3(kT), 0, s, &b->ctl, lz->h1, lz->h2, lz->ab);
  return hipGetLastError();
}

int get_graph(mpsd_lanczos* lz, int k, hipGraphExec_t* out) {
  GridKey key = lz->grids;
  key[1] = k;
  hipGraphExec_t& ex = lz->exec[key];
  if (!ex) {
    hipStream_t s = lz->ctx->stream;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return lz_err(lz, e, "begin capture");
    hipError_t le = hipSuccess;
    for (int i = 0; i < k && le == hipSuccess; ++i) le = enqueue_step(lz, key);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (le != hipSuccess) e = le;
    if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      ex = nullptr;
      return lz_err(lz, e, "graph capture");
    }
  }
  *out = ex;
  return MPSD_OK;
}

int upload(mpsd_lanczos* lz) {
  hipError_t e = hipMemcpyAsync(lz->dev, lz->host, sizeof(Block), hipMemcpyHostToDevice, lz->ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(lz->host_ev, lz->ctx->stream);
  return lz_err(lz, e, "lanczos upload");
}

}  // namespace

extern "C" {

int mpsd_lanczos_create(mpsd_ctx* ctx, int max_krylov, mpsd_lanczos** out) {
  if (!ctx || !out || max_krylov < 1 || max_krylov > 512) return MPSD_ERR_ARG;
  *out = nullptr;
  mpsd_lanczos* lz = new mpsd_lanczos();
  lz->ctx = ctx;
  lz->mmax = max_krylov;
  const int ma = max_krylov + 2;
  hipError_t e = hipMalloc((void**)&lz->dev, sizeof(Block));
  if (e == hipSuccess) e = hipHostMalloc((void**)&lz->host, sizeof(Block), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&lz->host_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc((void**)&lz->partial, sizeof(double2) * kDotBlocks * ma);
  if (e == hipSuccess) e = hipMalloc((void**)&lz->h1, sizeof(double2) * ma);
  if (e == hipSuccess) e = hipMalloc((void**)&lz->h2, sizeof(double2) * ma);
  if (e == hipSuccess) e = hipMalloc((void**)&lz->ab, sizeof(double) * (2 * ma + 2));
  if (e != hipSuccess) {
    mpsd_lanczos_destroy(lz);
    return hip_err(ctx, e, "lanczos_create");
  }
  memset(lz->host, 0, sizeof(Block));
  *out = lz;
  return MPSD_OK;
}

int mpsd_lanczos_destroy(mpsd_lanczos* lz) {
  if (!lz) return MPSD_ERR_ARG;
  (void)hipStreamSynchronize(lz->ctx->stream);
  for (auto& kv : lz->exec)
    if (kv.second) (void)hipGraphExecDestroy(kv.second);
  for (void* p : {(void*)lz->dev, (void*)lz->partial, (void*)lz->h1, (void*)lz->h2, (void*)lz->ab, (void*)lz->V,
                  (void*)lz->S1, (void*)lz->S2})
    if (p) (void)hipFree(p);
  if (lz->host) (void)hipHostFree(lz->host);
  if (lz->host_ev) (void)hipEventDestroy(lz->host_ev);
  delete lz;
  return MPSD_OK;
}

int mpsd_lanczos_begin(mpsd_lanczos* lz, mpsd_heff* h, const double* x) {
  if (!lz || !h || !x) return MPSD_ERR_ARG;
  mpsd_ctx* ctx = lz->ctx;
  if (h->d != 2 || h->precision == MPSD_PREC_BF16X3)
    return set_err(ctx, MPSD_ERR_UNSUPPORTED, "lanczos graph: needs d = 2 and the FP64 or FP32 form");
  const int64_t n = (int64_t)h->d * h->Dk * h->Dkr;
  if ((int64_t)h->d * h->Db * h->Dbr != n) return set_err(ctx, MPSD_ERR_SHAPE, "lanczos graph: Heff is not square");
  int rc = grow(lz, &lz->V, &lz->vcap, (int64_t)(lz->mmax + 1) * n);
  // T1 + partials in S1, T2 in S2 (the layout of the grouped matvec)
  const int64_t t1 = (int64_t)h->Db * h->wl * h->d * h->Dkr, ny = n;
  const bool split = stage3_split(h, ctx);
  if (!rc) rc = grow(lz, &lz->S1, &lz->s1cap, t1 + (split ? ny * h->wr : 0));
  if (!rc) rc = grow(lz, &lz->S2, &lz->s2cap, (int64_t)h->Db * h->d * h->wr * h->Dkr);
  if (rc) return rc;
  // the mix matrix is cached on W by mpsd_heff_create
  if (!h->mix) return set_err(ctx, MPSD_ERR_ARG, "lanczos graph: Heff without mix matrix");
  hipError_t e = hipEventSynchronize(lz->host_ev);  // staging image free again
  if (e != hipSuccess) return lz_err(lz, e, "lanczos staging wait");
  Block& b = *lz->host;
  memset(&b, 0, sizeof(Block));
  const int arith = h->precision == MPSD_PREC_FP32 ? 1 : 0;
  b.ctl.j = 0;
  b.ctl.j_next = 0;
  b.ctl.split = split ? 1 : 0;
  b.ctl.has_offset = h->offset != 0.0 ? 1 : 0;
  b.ctl.n_sum = split ? 1 : 0;
  b.ctl.n = n;
  b.ctl.ldv = n;
  b.ctl.V = lz->V;
  {  // stage 1 (left_stages layout); B is patched per step
    ZgemmDesc& d = b.st1.desc;
    d = ZgemmDesc();
    d.A = h->L->ptr; d.C = lz->S1;
    d.M = h->Db * h->wl; d.N = h->Dkr; d.K = h->Dk; d.NB = h->d;
    d.sAm = 1; d.sAk = (int64_t)h->Db * h->wl; d.sAb = 0;
    d.sBk = h->d; d.sBn = (int64_t)h->d * h->Dk; d.sBb = 1;
    d.sCm = 1; d.sCn = (int64_t)h->Db * h->wl * h->d; d.sCb = (int64_t)h->Db * h->wl;
    d.a_mfast = 1; d.b_kfast = 1; d.arith = arith;
    choose_group(d, lz->ctx->tile_override.nbw);
    b.st1.gm = (d.M + 31) / 32; b.st1.gn = (d.N + 31) / 32; b.st1.gz = (d.NB + d.group - 1) / d.group;
    b.st1.ntiles = b.st1.gm * b.st1.gn * b.st1.gz; b.st1.tile_begin = 0;
  }
  {  // mix
    GroupedElem& m = b.mix;
    m.in = lz->S1; m.out = lz->S2; m.aux = h->mix;
    m.Bx = h->Db; m.nin = h->wl * h->d; m.nout = h->d * h->wr; m.Ax = h->Dkr; m.begin = 0;
  }
  {  // stage 3 (+ sum); C / Dadd or sum.out / sum.aux are patched per step
    ZgemmDesc& t = b.st3.desc;
    t = ZgemmDesc();
    t.A = lz->S2; t.B = h->R->ptr; t.a_mfast = 1; t.b_kfast = 0; t.arith = arith;
    t.M = h->Db; t.N = h->Dbr; t.sAm = 1; t.sBn = 1; t.sBb = 0;
    if (split) {
      double2* P = lz->S1 + t1;
      t.C = P;
      t.K = h->Dkr; t.NB = h->d * h->wr; t.group = h->d;
      t.sAk = (int64_t)h->Db * h->d * h->wr; t.sAb = h->Db; t.sAz = (int64_t)h->Db * h->d;
      t.sBk = (int64_t)h->Dbr * h->wr; t.sBz = h->Dbr;
      t.sCm = h->d; t.sCn = (int64_t)h->d * h->Db; t.sCb = 1; t.sCz = ny;
      GroupedElem& su = b.sum;
      su.in = P; su.beta = h->offset; su.np = h->wr; su.stride = ny; su.n = ny; su.begin = 0;
    } else {
      t.K = h->wr * h->Dkr; t.NB = h->d;
      t.sAk = (int64_t)h->Db * h->d; t.sAb = h->Db;
      t.sBk = h->Dbr;
      t.sCm = h->d; t.sCn = (int64_t)h->d * h->Db; t.sCb = 1;
      t.beta = h->offset;
    }
    choose_group(t, lz->ctx->tile_override.nbw);
    b.st3.gm = (t.M + 31) / 32; b.st3.gn = (t.N + 31) / 32; b.st3.gz = (t.NB + t.group - 1) / t.group;
    b.st3.ntiles = b.st3.gm * b.st3.gn * b.st3.gz; b.st3.tile_begin = 0;
  }
  lz->arith = arith;
  lz->n = n;
  lz->jn = 0;
  auto cap = [](int64_t v, int hi) { return (int)std::max<int64_t>(1, std::min<int64_t>(v, hi)); };
  const int p1 = b.st1.desc.group == 1 ? 2 * kPersistGemm : kPersistGemm;
  const int p3 = b.st3.desc.group == 1 ? 2 * kPersistGemm : kPersistGemm;
  lz->grids = GridKey{arith, 0, cap(b.st1.ntiles, p1), cap(((int64_t)b.mix.Bx * b.mix.Ax + kT - 1) / kT, kPersist),
                      cap(b.st3.ntiles, p3), split ? cap((ny + kT - 1) / kT, kPersist) : 1,
                      cap((n + kT - 1) / kT, kVecBlocks), dot_blocks(n), b.st1.desc.group, b.st3.desc.group};
  rc = upload(lz);
  if (rc) return rc;
  // V_0 = x / |x|, |x|^2 into ab[0..1]
  return mpsd_vec_normalize_dev(ctx, x, (double*)lz->V, n, lz->ab);
}

int mpsd_lanczos_steps(mpsd_lanczos* lz, int k, double* ab_host, int* j_end) {
  if (!lz || !ab_host || !j_end || k < 1 || k > 8) return MPSD_ERR_ARG;
  if (lz->jn + k > lz->mmax) return set_err(lz->ctx, MPSD_ERR_ARG, "lanczos_steps: past the basis size");
  hipGraphExec_t ex = nullptr;
  int rc = get_graph(lz, k, &ex);
  if (rc) return rc;
  hipError_t e = hipGraphLaunch(ex, lz->ctx->stream);
  if (e != hipSuccess) return lz_err(lz, e, "graph launch");
  ++lz->replays;
  ++lz->ctx->graph_launches;
  lz->jn += k;
  *j_end = lz->jn;
  return mpsd_copy_to_host(lz->ctx, lz->ab, ab_host, (int64_t)sizeof(double) * (2 + 2 * (*j_end)));
}

int mpsd_lanczos_ritz(mpsd_lanczos* lz, const double* s_host, int K, double* x) {
  if (!lz || !s_host || !x || K < 1 || K > lz->mmax) return MPSD_ERR_ARG;
  std::vector<double> c(2 * K, 0.0);
  for (int i = 0; i < K; ++i) c[2 * i] = s_host[i];
  int rc = mpsd_vec_zero(lz->ctx, x, lz->n);
  if (!rc) rc = mpsd_vec_gemv(lz->ctx, (const double*)lz->V, lz->n, K, c.data(), x, lz->n);
  if (!rc) rc = mpsd_vec_normalize_dev(lz->ctx, x, x, lz->n, (double*)lz->h1);  // h1 is free between chunks
  return rc;
}

int64_t mpsd_lanczos_replays(const mpsd_lanczos* lz) { return lz ? lz->replays : 0; }

}  // extern "C"
