// Tile autotuner of the FP64 / FP32 BRGEMM launches.
//
// A tuning pass runs one operation (a matvec or an environment update) once per candidate
// tile configuration, timing every launch with HIP events, and keeps, for each BRGEMM shape
// class the operation launched, the fastest candidate.  Later launches of the same shape
// class on the context take the tuned configuration (tile_for).  The candidates differ only
// in how the output is tiled, how deep a K stage is and how many batches a workgroup
// carries; the order in which every output element is accumulated is the same for all of
// them, so tuning never changes a result bit.  The grouped persistent launches of the
// serving loop and of the graph-replayed eigensolver keep their fixed 32x32x16 tiles.

#include <cstdio>
#include <map>

#include "mpsd_internal.h"

namespace mpsd {

GemmKey gemm_key(const ZgemmDesc& g) {
  GemmKey k;
  k.M = g.M;
  k.N = g.N;
  k.K = g.K;
  k.R = g.R;
  k.NB = g.NB;
  k.flags = (g.a_mfast ? 1 : 0) | (g.b_kfast ? 2 : 0) | (g.sAb == 0 ? 4 : 0) | (g.sBb == 0 ? 8 : 0) |
            ((g.arith & 3) << 4) | ((g.group & 7) << 6) | (g.conjA ? 512 : 0) | (g.Dadd ? 1024 : 0);
  return k;
}

TileCfg tile_for(mpsd_ctx* ctx, const ZgemmDesc& g) {
  if (ctx->tune_cand) {
    ctx->tune_seen.push_back({ctx->cur_kind, gemm_key(g)});
    return *ctx->tune_cand;
  }
  if (ctx->tile_user) return ctx->tile_override;
  if (!ctx->tuned.empty()) {
    auto it = ctx->tuned.find(gemm_key(g));
    if (it != ctx->tuned.end()) return it->second;
  }
  return TileCfg{};
}

// Packed-operand BRGEMMs (zgemm_lp.hip) share the tuner: their shape class carries bit 12 and
// their arithmetic, the candidate's tile maps onto the packed tiles (other candidates: the
// automatic choice), and every packed tile gives the same bits.
GemmKey lp_key(const LpDesc& g) {
  GemmKey k;
  k.M = g.M;
  k.N = g.N;
  k.K = g.Kp;
  k.R = 0;
  k.NB = g.NB;
  k.flags = 4096 | ((g.prec & 3) << 4) | (g.a_il ? 8192 : 0) | (g.b_ip ? 16384 : 0);
  return k;
}

// (a candidate with no packed counterpart, such as 16 x 16, times the automatic choice again)
static int lp_tile_of_cfg(const TileCfg& c) {
  if (c.tm == 32 && c.tn == 32) return 1;
  if (c.tm == 64 && c.tn == 32) return 2;
  if (c.tm == 32 && c.tn == 64) return 3;
  if (c.tm == 64 && c.tn == 64) return 4;
  return 0;
}

int lp_tile_for(mpsd_ctx* ctx, const LpDesc& g) {
  if (ctx->tune_cand) {
    ctx->tune_seen.push_back({ctx->cur_kind, lp_key(g)});
    return lp_tile_of_cfg(*ctx->tune_cand);
  }
  const int user = g.prec == MPSD_PREC_FP64 ? ctx->pk64_tile : ctx->lp_tile;
  if (user) return user;
  if (!ctx->tuned.empty()) {
    auto it = ctx->tuned.find(lp_key(g));
    if (it != ctx->tuned.end()) return lp_tile_of_cfg(it->second);
  }
  return 0;
}

namespace {

// Candidates: the built-in choice first (kept unless another is clearly faster).
const TileCfg kCandidates[] = {
    {0, 0, 16, 0},  {32, 32, 16, 1}, {32, 32, 32, 0}, {32, 32, 32, 1}, {64, 32, 16, 0},
    {64, 32, 16, 1}, {32, 64, 16, 0}, {32, 64, 16, 1}, {64, 64, 16, 0}, {16, 16, 16, 0},
};
constexpr int kNumCandidates = sizeof(kCandidates) / sizeof(kCandidates[0]);
constexpr double kKeepBuiltin = 0.97;  // a candidate must beat the built-in by 3%

struct Stat {
  int kind = -1;
  double us[kNumCandidates] = {0};
};

// One tuning pass of `op` (which issues its launches on ctx's stream).  Returns the number
// of shape classes added or updated; fills out[0..max_out) with one row per class.
int tune_op(mpsd_ctx* ctx, int reps, const std::function<int()>& op, mpsd_tune_result* out, int max_out,
            int* n_out) {
  if (n_out) *n_out = 0;
  if (reps < 1) reps = 1;
  if (ctx->tune_cand) return set_err(ctx, MPSD_ERR_ARG, "tune: a tuning pass is already running");
  int rc = mpsd_ctx_set_timing(ctx, 0);
  if (rc) return rc;
  std::map<GemmKey, Stat> stats;
  // Two round-robin passes over the candidates, keeping each one's faster pass, so clock
  // ramps and cache warm-up do not favour whichever candidate runs late.
  for (int pass = 0; pass < 2 * kNumCandidates; ++pass) {
    const int c = pass % kNumCandidates;
    ctx->tune_cand = &kCandidates[c];
    rc = pass < kNumCandidates ? op() : MPSD_OK;  // warm-up: code objects, caches
    if (!rc) rc = mpsd_ctx_set_timing(ctx, 1);
    ctx->tune_seen.clear();
    for (int r = 0; r < reps && !rc; ++r) rc = op();
    double ms[MPSD_KERNEL_KINDS] = {0};
    int64_t n[MPSD_KERNEL_KINDS] = {0};
    for (int k = 0; k < MPSD_KERNEL_KINDS && !rc; ++k) rc = mpsd_ctx_timing(ctx, k, &ms[k], &n[k]);
    ctx->tune_cand = nullptr;
    (void)mpsd_ctx_set_timing(ctx, 0);
    if (rc) {
      ctx->tune_seen.clear();
      return rc;
    }
    // time of a shape class = its kind's time per launch of that kind
    for (auto& kv : ctx->tune_seen) {
      const int kind = kv.first;
      if (kind < 0 || kind >= MPSD_KERNEL_KINDS || n[kind] == 0) continue;
      Stat& st = stats[kv.second];
      st.kind = kind;
      const double us = 1e3 * ms[kind] / (double)n[kind];
      st.us[c] = (st.us[c] > 0.0 && st.us[c] < us) ? st.us[c] : us;
    }
    ctx->tune_seen.clear();
  }
  int rows = 0;
  for (auto& kv : stats) {
    const Stat& st = kv.second;
    if (((kv.first.flags >> 4) & 3) == 2) continue;  // split-BF16 form: fixed tiles
    const bool packed = (kv.first.flags & 4096) != 0;
    int best = 0;
    for (int c = 1; c < kNumCandidates; ++c) {
      if (packed && lp_tile_of_cfg(kCandidates[c]) == 0) continue;  // the automatic tile again
      if (st.us[c] > 0.0 && st.us[c] < st.us[best] * (best == 0 ? kKeepBuiltin : 1.0)) best = c;
    }
    ctx->tuned[kv.first] = kCandidates[best];
    if (out && rows < max_out) {
      mpsd_tune_result& r = out[rows];
      r.M = kv.first.M;
      r.N = kv.first.N;
      r.K = kv.first.K;
      r.NB = kv.first.NB;
      r.arith = (kv.first.flags >> 4) & 3;
      r.kind = st.kind;
      r.tm = kCandidates[best].tm;
      r.tn = kCandidates[best].tn;
      r.tk = kCandidates[best].tk;
      r.nbw = kCandidates[best].nbw;
      r.us_builtin = st.us[0];
      r.us_best = st.us[best];
    }
    ++rows;
  }
  if (n_out) *n_out = rows < max_out ? rows : max_out;
  return MPSD_OK;
}

}  // namespace
}  // namespace mpsd

using namespace mpsd;

extern "C" {

int mpsd_tune_heff(mpsd_heff* h, const double* x, double* y, int reps, mpsd_tune_result* out, int max_out,
                   int* n_out) {
  if (!h || !x || !y) return MPSD_ERR_ARG;
  mpsd_ctx* ctx = h->ctx;
  const int64_t n = mpsd_heff_size(h);
  return tune_op(ctx, reps, [&] { return mpsd_heff_apply(h, x, y, 1, n, n, 1); }, out, max_out, n_out);
}

int mpsd_tune_env(mpsd_ctx* ctx, int side, const mpsd_tensor* env, const mpsd_tensor* ket, const mpsd_tensor* W,
                  const mpsd_tensor* bra, mpsd_tensor* outt, int reps, mpsd_tune_result* out, int max_out,
                  int* n_out) {
  if (!ctx || !env || !ket || !W || !bra || !outt || (side != 0 && side != 1)) return MPSD_ERR_ARG;
  return tune_op(
      ctx, reps,
      [&] { return side == 0 ? mpsd_env_left(ctx, env, ket, W, bra, outt) : mpsd_env_right(ctx, env, ket, W, bra, outt); },
      out, max_out, n_out);
}

int mpsd_tune_count(const mpsd_ctx* ctx) { return ctx ? (int)ctx->tuned.size() : MPSD_ERR_ARG; }

int mpsd_tune_clear(mpsd_ctx* ctx) {
  if (!ctx) return MPSD_ERR_ARG;
  ctx->tuned.clear();
  return MPSD_OK;
}

// Text table, one shape class per line: M N K R NB flags tm tn tk nbw
int mpsd_tune_save(const mpsd_ctx* ctx, const char* path) {
  if (!ctx || !path) return MPSD_ERR_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return set_err(const_cast<mpsd_ctx*>(ctx), MPSD_ERR_ARG, std::string("tune_save: cannot open ") + path);
  fprintf(f, "# mpsd tile table v1: M N K R NB flags tm tn tk nbw\n");
  for (auto& kv : ctx->tuned)
    fprintf(f, "%d %d %d %d %d %d %d %d %d %d\n", kv.first.M, kv.first.N, kv.first.K, kv.first.R, kv.first.NB,
            kv.first.flags, kv.second.tm, kv.second.tn, kv.second.tk, kv.second.nbw);
  const bool ok = fclose(f) == 0;
  return ok ? MPSD_OK : set_err(const_cast<mpsd_ctx*>(ctx), MPSD_ERR_ARG, "tune_save: write failed");
}

int mpsd_tune_load(mpsd_ctx* ctx, const char* path) {
  if (!ctx || !path) return MPSD_ERR_ARG;
  FILE* f = fopen(path, "r");
  if (!f) return set_err(ctx, MPSD_ERR_ARG, std::string("tune_load: cannot open ") + path);
  char line[256];
  int loaded = 0;
  std::map<GemmKey, TileCfg> add;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    GemmKey k;
    TileCfg c;
    if (sscanf(line, "%d %d %d %d %d %d %d %d %d %d", &k.M, &k.N, &k.K, &k.R, &k.NB, &k.flags, &c.tm, &c.tn, &c.tk,
               &c.nbw) != 10) {
      fclose(f);
      return set_err(ctx, MPSD_ERR_ARG, std::string("tune_load: malformed line: ") + line);
    }
    const bool tile_ok = (c.tm == 0 && c.tn == 0) || ((c.tm == 32 || c.tm == 64) && (c.tn == 32 || c.tn == 64));
    if (!tile_ok || (c.tk != 16 && c.tk != 32) || c.nbw < 0 || c.nbw > 2) {
      fclose(f);
      return set_err(ctx, MPSD_ERR_ARG, std::string("tune_load: invalid configuration: ") + line);
    }
    add[k] = c;
    ++loaded;
  }
  fclose(f);
  for (auto& kv : add) ctx->tuned[kv.first] = kv.second;
  return loaded;
}

}  // extern "C"
