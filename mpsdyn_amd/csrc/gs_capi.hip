// C-ABI over the C++ Contractor (include/mpsd_gs.h).  C++ exceptions stop here and become
// status codes plus the context's error message.

#include <chrono>
#include <cstring>

#include "mpsd_gs.h"
#include "mpsd_internal.h"
#include "mpsdyn/mpsdyn.h"

using namespace mpsdyn;

namespace {

std::unique_ptr<Hamiltonian> make_model(const mpsd_gs_config* c) {
  switch (c->model) {
    case MPSD_MODEL_ISING:
      return std::unique_ptr<Hamiltonian>(new IsingHamiltonian(c->L, 2, c->p[0], c->p[1], c->p[2], c->p[3]));
    case MPSD_MODEL_HEISENBERG:
      return std::unique_ptr<Hamiltonian>(new HeisenbergHamiltonian(c->L, c->p[0], c->p[1], c->p[2], c->p[3]));
    case MPSD_MODEL_SCHWINGER:
      return std::unique_ptr<Hamiltonian>(new SchwingerHamiltonian(c->L, c->p[0], c->p[1], c->p[2], c->p[3]));
    default:
      throw Error(MPSD_ERR_ARG, "unknown model");
  }
}

int fail(mpsd_ctx* ctx, const Error& e) {
  if (ctx) ctx->err = e.what();
  return e.code ? e.code : MPSD_ERR_ARG;
}

// Reads an MPO file (MPO::loadMPO / MPO::importMPOtext, src/core/cpp/MPO.cpp:250-316) and checks
// that it is a chain Hamiltonian the sweep can use: rank-4 [du, Dl, dd, Dr] operators with
// du == dd, matching bonds between neighbours and bond 1 at both ends.
void read_mpo(MPO& mpo, const char* path, int format) {
  if (!path) throw Error(MPSD_ERR_ARG, "MPO path is NULL");
  if (format == MPSD_MPO_BINARY)
    mpo.loadMPO(path);
  else if (format == MPSD_MPO_TEXT)
    mpo.importMPOtext(path);
  else
    throw Error(MPSD_ERR_ARG, "unknown MPO file format");
  const int L = mpo.getLength();
  if (L < 2) throw Error(MPSD_ERR_SHAPE, "MPO file: need at least two sites");
  for (int k = 0; k < L; ++k) {
    const Operator& W = mpo.getOp(k);
    if (W.getDimension(0) != W.getDimension(2))
      throw Error(MPSD_ERR_SHAPE, "MPO file: operator " + std::to_string(k) + " is not square in the physical index");
    const int left = k == 0 ? 1 : mpo.getOp(k - 1).getDimension(3);
    if (W.getDimension(1) != left)
      throw Error(MPSD_ERR_SHAPE, "MPO file: bond mismatch at operator " + std::to_string(k));
  }
  if (mpo.getOp(L - 1).getDimension(3) != 1) throw Error(MPSD_ERR_SHAPE, "MPO file: last bond is not 1");
}

// The Hamiltonian of a call: read from an MPO file when mpo_path is set, else cfg's model.
const MPO& load_hamiltonian(const mpsd_gs_config* cfg, const char* mpo_path, int format,
                            std::unique_ptr<Hamiltonian>& model, MPO& file) {
  if (mpo_path) {
    read_mpo(file, mpo_path, format);
    return file;
  }
  if (cfg->L < 2) throw Error(MPSD_ERR_ARG, "need L >= 2");
  model = make_model(cfg);
  return model->getHMPO();
}

void configure(Contractor& c, const mpsd_gs_config* cfg) {
  if (cfg->solver != MPSD_SOLVER_LANCZOS && cfg->solver != MPSD_SOLVER_FULLEIG)
    throw Error(MPSD_ERR_ARG, "unknown solver");
  c.setConvTol(cfg->convtol > 0 ? cfg->convtol : 1e-5);
  c.setEigTol(cfg->eigtol);
  c.setMaxHalfSweeps(cfg->max_halfsweeps);
  c.setPrecision(cfg->precision);
  if (cfg->solver == MPSD_SOLVER_FULLEIG) c.setEigenSolver(fulleig);
  c.setEigTolSchedule(cfg->tol_schedule != 0);
}

void fill_result(mpsd_ctx* ctx, const Contractor& c, double E, double seconds, mpsd_gs_result* out) {
  const SolveStats& st = c.lastStats();
  out->energy = E;
  out->halfsweeps = st.halfsweeps;
  out->matvecs = st.matvecs;
  out->env_updates = st.env_updates;
  out->residual = st.final_residual;
  out->seconds = seconds;
  out->device_bytes = mpsd_ctx_device_bytes(ctx);
  out->matvecs_lowp = st.matvecs_lowp;
  out->refine_from = st.refine_from;
  out->n_history = (int)std::min<size_t>(st.energies.size(), MPSD_GS_MAX_HISTORY);
  for (int i = 0; i < out->n_history; ++i) {
    out->history[i] = st.energies[i];
    out->hs_seconds[i] = i < (int)st.hs_seconds.size() ? st.hs_seconds[i] : 0.0;
    out->hs_tol[i] = i < (int)st.hs_tol.size() ? st.hs_tol[i] : 0.0;
    out->hs_matvecs[i] = i < (int)st.hs_matvecs.size() ? st.hs_matvecs[i] : 0;
  }
  out->residual_norm = -1.0;
  out->residual_floor = -1.0;
  out->matvec_flops = st.matvec_flops;
  out->env_flops = st.env_flops;
}

}  // namespace

namespace mpsdyn {
// The model of cfg with its energy shifted by -E, spread over the sites' local terms (the
// Ising constant K, the Heisenberg and Schwinger offsets): (H - E) as an MPO whose partial
// sums stay of the order of the fluctuations.
std::unique_ptr<Hamiltonian> shifted_model(const mpsd_gs_config* c, double E) {
  switch (c->model) {
    case MPSD_MODEL_ISING:
      return std::unique_ptr<Hamiltonian>(new IsingHamiltonian(c->L, 2, c->p[0], c->p[1], c->p[2], c->p[3] - E));
    case MPSD_MODEL_HEISENBERG:
      return std::unique_ptr<Hamiltonian>(new HeisenbergHamiltonian(c->L, c->p[0], c->p[1], c->p[2], c->p[3], 2, -E));
    case MPSD_MODEL_SCHWINGER:
      return std::unique_ptr<Hamiltonian>(new SchwingerHamiltonian(c->L, c->p[0], c->p[1], c->p[2], c->p[3], -E));
    default:
      throw Error(MPSD_ERR_ARG, "unknown model");
  }
}

// ||H psi - E psi|| for a normalised psi: sqrt(<psi|(H - E)^2|psi>) over the MPO of (H - E)^2,
// the product of the shifted model's MPO with itself (bond w^2), on the device.  *floor gets the
// rounding level of that variance: <(H - E - delta)^2> - delta^2 - <(H - E)^2> is zero in exact
// arithmetic (up to 2 delta (<H> - E), which is far smaller), so the difference of the two
// evaluations measures the rounding of either; a residual below sqrt of it is not resolved.
double residual_norm(Contractor& c, const mpsd_gs_config* cfg, MPS& mps, double E, double* floor) {
  auto Hs = shifted_model(cfg, E);
  const MPO sq = mpo_product(Hs->getHMPO(), Hs->getHMPO());
  const double v = c.contract(mps, sq, mps).real();
  const double delta = std::ldexp(1.0, -10);  // delta^2 exact
  auto Hd = shifted_model(cfg, E + delta);
  const MPO sqd = mpo_product(Hd->getHMPO(), Hd->getHMPO());
  const double vd = c.contract(mps, sqd, mps).real();
  *floor = std::sqrt(std::fabs(vd - delta * delta - v));
  return std::sqrt(std::max(v, 0.0));
}

int gs_solve_mpo_with(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const MPO& hmpo, const char* export_path,
                      const MultiplierFactory* factory, mpsd_gs_result* out, const char* start_path = nullptr,
                      bool model_residual = false);

// One solve on ctx with an optional multiplier factory (used by the serving loop).
int gs_solve_with(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* export_path, const MultiplierFactory* factory,
                  mpsd_gs_result* out) {
  if (!ctx || !cfg || !out) return MPSD_ERR_ARG;
  if (cfg->L < 2 || cfg->D < 1) return mpsd::set_err(ctx, MPSD_ERR_ARG, "gs_solve: need L >= 2 and D >= 1");
  try {
    auto H = make_model(cfg);
    return gs_solve_mpo_with(ctx, cfg, H->getHMPO(), export_path, factory, out, nullptr, true);
  } catch (const Error& e) {
    return fail(ctx, e);
  }
}

// The sweep itself, on any chain MPO; cfg supplies D, the tolerances, precision, seed and
// solver (its model, L and p are not read unless model_residual is set).  start_path: an
// exported MPS to start from instead of the seeded random state (findGroundState then fits it
// to the MPO and to D; D = 0 keeps its bond).
int gs_solve_mpo_with(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const MPO& hmpo, const char* export_path,
                      const MultiplierFactory* factory, mpsd_gs_result* out, const char* start_path,
                      bool model_residual) {
  if (!ctx || !cfg || !out) return MPSD_ERR_ARG;
  if (cfg->D < (start_path ? 0 : 1)) return mpsd::set_err(ctx, MPSD_ERR_ARG, "gs_solve: need D >= 1");
  if (cfg->solver != MPSD_SOLVER_LANCZOS && cfg->solver != MPSD_SOLVER_FULLEIG)
    return mpsd::set_err(ctx, MPSD_ERR_ARG, "gs_solve: unknown solver");
  if (cfg->want_residual && !model_residual)
    return mpsd::set_err(ctx, MPSD_ERR_UNSUPPORTED, "gs_solve: the residual norm needs a built-in model");
  memset(out, 0, sizeof(*out));
  try {
    MPS mps(hmpo.getLength(), std::max(cfg->D, 1), hmpo.getDimensions(), ctx);
    if (start_path) {
      mps.importMPS(start_path);
      if (mps.getLength() != hmpo.getLength())
        throw Error(MPSD_ERR_SHAPE, "gs_solve: start state and Hamiltonian lengths differ");
      // physical dimensions first, so that the start energy below can be evaluated (findGroundState
      // repeats this fit, a no-op by then)
      mps.adjustPhysDimensions(hmpo.getDimensions());
    } else {
      mps.setRandomState(cfg->seed);
    }
    Contractor c(ctx);
    if (factory) c.setMultiplierFactory(*factory);
    c.setConvTol(cfg->convtol > 0 ? cfg->convtol : 1e-5);
    c.setEigTol(cfg->eigtol);
    c.setMaxHalfSweeps(cfg->max_halfsweeps);
    c.setPrecision(cfg->precision);
    if (cfg->solver == MPSD_SOLVER_FULLEIG) c.setEigenSolver(fulleig);
    c.setBreakdown(cfg->breakdown != 0);
    c.setEigTolSchedule(cfg->tol_schedule != 0);
    out->initial_energy = c.contract(mps, hmpo, mps).real();
    double E = 0.0;
    auto t0 = std::chrono::steady_clock::now();
    c.findGroundState(hmpo, cfg->D, &E, mps, cfg->offset);
    check(ctx, mpsd_ctx_sync(ctx), "sync");
    auto t1 = std::chrono::steady_clock::now();
    const SolveStats& st = c.lastStats();
    out->energy = E;
    out->halfsweeps = st.halfsweeps;
    out->matvecs = st.matvecs;
    out->env_updates = st.env_updates;
    out->residual = st.final_residual;
    out->seconds = std::chrono::duration<double>(t1 - t0).count();
    out->device_bytes = mpsd_ctx_device_bytes(ctx);
    out->matvecs_lowp = st.matvecs_lowp;
    out->refine_from = st.refine_from;
    out->n_history = (int)std::min<size_t>(st.energies.size(), MPSD_GS_MAX_HISTORY);
    for (int i = 0; i < out->n_history; ++i) {
      out->history[i] = st.energies[i];
      out->hs_seconds[i] = i < (int)st.hs_seconds.size() ? st.hs_seconds[i] : 0.0;
      out->hs_tol[i] = i < (int)st.hs_tol.size() ? st.hs_tol[i] : 0.0;
      out->hs_matvecs[i] = i < (int)st.hs_matvecs.size() ? st.hs_matvecs[i] : 0;
    }
    out->t_env = st.t_env;
    out->t_eig = st.t_eig;
    out->t_gauge = st.t_gauge;
    out->t_matvec = st.t_matvec;
    out->matvec_flops = st.matvec_flops;
    out->env_flops = st.env_flops;
    out->residual_norm = -1.0;
    out->residual_floor = -1.0;
    if (cfg->want_residual) out->residual_norm = residual_norm(c, cfg, mps, E, &out->residual_floor);
    if (export_path) mps.exportMPS(export_path);
    return MPSD_OK;
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}
}  // namespace mpsdyn

namespace {
int solve(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* export_path, mpsd_gs_result* out) {
  return mpsdyn::gs_solve_with(ctx, cfg, export_path, nullptr, out);
}

// Non-owning Multiplier over a C-ABI Heff handle.
class HeffRef : public Multiplier {
 public:
  explicit HeffRef(mpsd_heff* h) : h_(h) {}
  int64_t getSize() const override { return mpsd_heff_size(h_); }
  void product_device(const double* x, double* y) override {
    const int64_t n = getSize();
    check(h_->ctx, mpsd_heff_apply(h_, x, y, 1, n, n, 1), "heff_apply");
  }
  void getFullTensor(mwArray& r) override {
    const int n = (int)getSize();
    r = mwArray(std::vector<int>{n, n});
    check(h_->ctx, mpsd_heff_full(h_, (double*)r.data()), "heff_full");
  }
  mpsd_ctx* ctx() const override { return h_->ctx; }
  mpsd_heff* heff() const override { return h_; }

 private:
  mpsd_heff* h_;
};

}  // namespace

extern "C" {

int mpsd_heff_eigs_lowest(mpsd_heff* h, double* x_dev, double tol, int max_krylov, double* value, double* residual,
                          int* matvecs) {
  if (!h || !x_dev || !value || !residual || !matvecs) return MPSD_ERR_ARG;
  try {
    HeffRef H(h);
    LanczosSolver solver(h->ctx);
    EigStats st = solver.lowest(H, x_dev, tol > 0 ? tol : 1e-10, max_krylov > 1 ? max_krylov : 30);
    *value = st.value;
    *residual = st.residual;
    *matvecs = st.matvecs;
    return st.converged ? MPSD_OK : mpsd::set_err(h->ctx, MPSD_ERR_NUMERIC, "eigensolver did not converge");
  } catch (const Error& e) {
    return fail(h->ctx, e);
  } catch (const std::exception& e) {
    h->ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

int mpsd_gs_solve(mpsd_ctx* ctx, const mpsd_gs_config* cfg, mpsd_gs_result* out) {
  return solve(ctx, cfg, nullptr, out);
}

int mpsd_gs_solve_export(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* export_path, mpsd_gs_result* out) {
  return solve(ctx, cfg, export_path, out);
}

int mpsd_mpo_save(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* path, int format) {
  if (!cfg || !path) return MPSD_ERR_ARG;
  if (cfg->L < 2) return ctx ? mpsd::set_err(ctx, MPSD_ERR_ARG, "mpo_save: need L >= 2") : MPSD_ERR_ARG;
  try {
    auto H = make_model(cfg);
    if (format == MPSD_MPO_BINARY)
      H->getHMPO().saveMPO(path);
    else if (format == MPSD_MPO_TEXT)
      H->getHMPO().exportMPOtext(path);
    else
      throw Error(MPSD_ERR_ARG, "unknown MPO file format");
    return MPSD_OK;
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

int mpsd_gs_solve_mpo(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mpo_path, int format,
                      const char* export_path, mpsd_gs_result* out) {
  if (!ctx || !cfg || !mpo_path || !out) return MPSD_ERR_ARG;
  try {
    MPO mpo;
    read_mpo(mpo, mpo_path, format);
    return mpsdyn::gs_solve_mpo_with(ctx, cfg, mpo, export_path, nullptr, out);
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

int mpsd_gs_solve_from(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mpo_path, int format, const char* mps_in,
                       const char* export_path, mpsd_gs_result* out) {
  if (!ctx || !cfg || !mps_in || !out) return MPSD_ERR_ARG;
  if (cfg->D < 0) return mpsd::set_err(ctx, MPSD_ERR_ARG, "gs_solve_from: need D >= 0");
  try {
    std::unique_ptr<Hamiltonian> model;
    MPO file;
    const MPO& H = load_hamiltonian(cfg, mpo_path, format, model, file);
    return mpsdyn::gs_solve_mpo_with(ctx, cfg, H, export_path, nullptr, out, mps_in, mpo_path == nullptr);
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

int mpsd_mpo_expectation(mpsd_ctx* ctx, const char* mpo_path, int format, const char* mps_path, double* re,
                         double* im) {
  if (!ctx || !mpo_path || !mps_path || !re || !im) return MPSD_ERR_ARG;
  try {
    MPO mpo;
    read_mpo(mpo, mpo_path, format);
    MPS mps(1, 1, 2, ctx);
    mps.importMPS(mps_path);
    if (mps.getLength() != mpo.getLength()) throw Error(MPSD_ERR_SHAPE, "mpo_expectation: MPS and MPO lengths differ");
    Contractor c(ctx);
    complex_t e = c.contract(mps, mpo, mps);
    *re = e.real();
    *im = e.imag();
    return MPSD_OK;
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

int mpsd_gs_solve_levels(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mpo_path, int format, int nlevels,
                         double exc_orth_tol, double* energies, double* overlaps, mpsd_gs_result* results) {
  if (!ctx || !cfg || !energies) return MPSD_ERR_ARG;
  if (nlevels < 1 || nlevels > 32) return mpsd::set_err(ctx, MPSD_ERR_ARG, "gs_solve_levels: need 1 <= nlevels <= 32");
  if (cfg->D < 1) return mpsd::set_err(ctx, MPSD_ERR_ARG, "gs_solve_levels: need D >= 1");
  try {
    std::unique_ptr<Hamiltonian> model;
    MPO file;
    const MPO& H = load_hamiltonian(cfg, mpo_path, format, model, file);
    Contractor c(ctx);
    configure(c, cfg);
    c.setExcOrthTol(exc_orth_tol);
    std::vector<std::unique_ptr<MPS>> levels;
    std::vector<MPS*> done;
    for (int k = 0; k < nlevels; ++k) {
      std::unique_ptr<MPS> m(new MPS(H.getLength(), cfg->D, H.getDimensions(), ctx));
      m->setRandomState(cfg->seed + 7919ull * (uint64_t)k);
      double E = 0.0;
      auto t0 = std::chrono::steady_clock::now();
      if (k == 0)
        c.findGroundState(H, cfg->D, &E, *m, cfg->offset);
      else
        c.findNextExcitedState(H, cfg->D, done, &E, *m, cfg->offset);
      check(ctx, mpsd_ctx_sync(ctx), "sync");
      auto t1 = std::chrono::steady_clock::now();
      energies[k] = E;
      if (results) {
        memset(&results[k], 0, sizeof(results[k]));
        fill_result(ctx, c, E, std::chrono::duration<double>(t1 - t0).count(), &results[k]);
      }
      levels.push_back(std::move(m));
      done.push_back(levels.back().get());
    }
    if (overlaps)
      for (int j = 0; j < nlevels; ++j)
        for (int i = 0; i < nlevels; ++i)
          overlaps[i + nlevels * j] = std::abs(c.contract(*levels[i], *levels[j]));
    return MPSD_OK;
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

int mpsd_gs_sweep_part(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mpo_path, int format, const char* mps_in,
                       int nproj, const char* const* proj_paths, const double* weights, int pos0, int lim,
                       const char* mps_out, double* energy) {
  if (!ctx || !cfg || !mps_in || nproj < 0 || (nproj > 0 && (!proj_paths || !weights))) return MPSD_ERR_ARG;
  try {
    std::unique_ptr<Hamiltonian> model;
    MPO file;
    const MPO& H = load_hamiltonian(cfg, mpo_path, format, model, file);
    MPS init(1, 1, 2, ctx);
    init.importMPS(mps_in);
    std::vector<std::unique_ptr<MPS>> proj;
    std::vector<MPS*> ptrs;
    for (int k = 0; k < nproj; ++k) {
      if (!proj_paths[k]) throw Error(MPSD_ERR_ARG, "gs_sweep_part: null projected-state path");
      proj.emplace_back(new MPS(1, 1, 2, ctx));
      proj.back()->importMPS(proj_paths[k]);
      ptrs.push_back(proj.back().get());
    }
    Contractor c(ctx);
    configure(c, cfg);
    if (nproj == 0)
      c.sweepPart(pos0, lim, H, cfg->D, init, cfg->offset);
    else
      c.sweepPartWithProjectorPenalty(pos0, lim, H, cfg->D, std::vector<double>(weights, weights + nproj), ptrs, init,
                                      cfg->offset);
    if (energy) *energy = c.contract(init, H, init).real() / c.contract(init, init).real();
    if (mps_out) init.exportMPS(mps_out);
    return MPSD_OK;
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

int mpsd_mps_gauge_site(mpsd_ctx* ctx, const char* mps_in, int pos, char dir, int next, const char* mps_out) {
  if (!ctx || !mps_in || !mps_out) return MPSD_ERR_ARG;
  try {
    MPS m(1, 1, 2, ctx);
    m.importMPS(mps_in);
    if (pos < 0 || pos >= m.getLength()) throw Error(MPSD_ERR_ARG, "mps_gauge_site: position out of range");
    if (dir != 'L' && dir != 'l' && dir != 'R' && dir != 'r')
      throw Error(MPSD_ERR_ARG, "mps_gauge_site: direction must be 'L' or 'R'");
    m.gaugeCond(pos, dir, next != 0);
    m.exportMPS(mps_out);
    return MPSD_OK;
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

int mpsd_mps_truncate(mpsd_ctx* ctx, const char* mps_in, int D, const char* mps_out) {
  if (!ctx || !mps_in || !mps_out || D < 1) return MPSD_ERR_ARG;
  try {
    MPS in(1, 1, 2, ctx);
    in.importMPS(mps_in);
    MPS out(in, D);
    out.exportMPS(mps_out);
    return MPSD_OK;
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

int mpsd_gs_expectation(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mps_path, double* re, double* im) {
  if (!ctx || !cfg || !mps_path || !re || !im) return MPSD_ERR_ARG;
  try {
    auto H = make_model(cfg);
    MPS mps(1, 1, 2, ctx);
    mps.importMPS(mps_path);
    Contractor c(ctx);
    complex_t e = c.contract(mps, H->getHMPO(), mps);
    *re = e.real();
    *im = e.imag();
    return MPSD_OK;
  } catch (const Error& e) {
    return fail(ctx, e);
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return MPSD_ERR_ARG;
  }
}

}  // extern "C"
