// Partial sweeps and excited states by projector penalties, over device-resident states:
// TensorMultiplierProjMulti and the Contractor's sweepPart / findNextExcitedState family.
// The projected levels enter each local problem through their effective vectors
// phi_k = (M_L (x) M_R) B_k, where M_L, M_R are the overlap environments of the state being
// optimised with level k; both are Heff-shaped contractions with the identity MPO, so they
// run on the same BRGEMM kernels as the matvec.

#include <chrono>
#include <cmath>
#include <cstdio>

#include "mpsdyn/mpsdyn.h"

namespace mpsdyn {

// ------------------------------------------------------------------------------ multipliers
TensorMultiplierProjMulti::TensorMultiplierProjMulti(mpsd_ctx* ctx, const mpsd_tensor* L, const mpsd_tensor* W,
                                                     const mpsd_tensor* R, const std::vector<double>& weights,
                                                     const std::vector<const double*>& phis, bool hermitian,
                                                     double offset)
    : ctx_(ctx), weights_(weights), phis_(phis) {
  if (weights.size() != phis.size())
    throw Error(MPSD_ERR_ARG, "TensorMultiplierProjMulti: one weight per projected vector");
  if (phis.size() > 32) throw Error(MPSD_ERR_ARG, "TensorMultiplierProjMulti: at most 32 projected vectors");
  for (const double* p : phis)
    if (!p) throw Error(MPSD_ERR_ARG, "TensorMultiplierProjMulti: null projected vector");
  if (hermitian)
    base_.reset(new TensorMultiplierHermitianOffset(ctx, L, W, R, offset));
  else
    base_.reset(new TensorMultiplierOffset(ctx, L, W, R, offset));
}

void TensorMultiplierProjMulti::product_device(const double* x, double* y) {
  base_->product_device(x, y);
  const int m = (int)phis_.size();
  if (m == 0) return;
  const int64_t n = getSize();
  // one launch sequence when the vectors are columns of one buffer (the sweep's layout)
  bool strided = true;
  for (int j = 1; j < m && strided; ++j) strided = phis_[j] == phis_[0] + 2 * n * j;
  if (strided) {
    check(ctx_, mpsd_vec_proj_update(ctx_, phis_[0], n, m, weights_.data(), x, y, n), "proj_update");
  } else {
    for (int j = 0; j < m; ++j)
      check(ctx_, mpsd_vec_proj_update(ctx_, phis_[j], n, 1, &weights_[j], x, y, n), "proj_update");
  }
}

void TensorMultiplierProjMulti::getFullTensor(mwArray& result) {
  base_->getFullTensor(result);
  const int n = (int)getSize();
  std::vector<complex_t> phi(n);
  for (size_t j = 0; j < phis_.size(); ++j) {
    check(ctx_, mpsd_copy_to_host(ctx_, phis_[j], phi.data(), (int64_t)sizeof(complex_t) * n), "proj_full");
    complex_t* F = result.data();
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < n; ++r) F[r + (int64_t)n * c] -= weights_[j] * phi[r] * std::conj(phi[c]);
  }
  // TensorMultiplierProjMulti.cpp:98: the result is averaged with its adjoint
  complex_t* F = result.data();
  for (int c = 0; c < n; ++c)
    for (int r = 0; r <= c; ++r) {
      const complex_t a = F[r + (int64_t)n * c], b = F[c + (int64_t)n * r];
      F[r + (int64_t)n * c] = 0.5 * (a + std::conj(b));
      F[c + (int64_t)n * r] = 0.5 * (b + std::conj(a));
    }
}

// ------------------------------------------------------------------------------ overlaps
const mpsd_tensor* Contractor::identityOp(int d) {
  for (auto& p : ident_)
    if (p.first == d) return p.second->device(ctx_);
  mwArray I(std::vector<int>{d, 1, d, 1});
  for (int s = 0; s < d; ++s) I.data()[s + d * s] = complex_t(1.0, 0.0);  // [s, 0, s, 0]
  ident_.emplace_back(d, std::unique_ptr<Operator>(new Operator(I)));
  return ident_.back().second->device(ctx_);
}

complex_t Contractor::contract(const MPS& ket, const MPS& bra) {
  const int L = ket.getLength();
  if (bra.getLength() != L) throw Error(MPSD_ERR_SHAPE, "contract(ket, bra): lengths differ");
  TmpManager mgr(L + 1, ctx_);
  for (int k = 1; k <= L; ++k)
    check(ctx_,
          mpsd_env_left(ctx_, mgr.operL(k - 1).t, ket.getA(k - 1).device(), identityOp(ket.getA(k - 1).getd()),
                        bra.getA(k - 1).device(), mgr.operL(k).t),
          "contract(ket, bra)");
  mwArray r = mgr.operL(L).download();
  return (ket.getNormFact() * bra.getNormFact()) * r.getElement(0);
}

void Contractor::checkLevels(const MPO& ops, const MPS& init, const std::vector<MPS*>& levels) const {
  const int L = ops.getLength();
  if (init.getLength() != L) throw Error(MPSD_ERR_SHAPE, "MPS and MPO lengths differ");
  if (init.ctx() != ctx_) throw Error(MPSD_ERR_ARG, "the MPS lives on another context");
  if (levels.size() > 32) throw Error(MPSD_ERR_ARG, "at most 32 projected levels");
  for (const MPS* m : levels) {
    if (!m) throw Error(MPSD_ERR_ARG, "null projected level");
    if (m == &init) throw Error(MPSD_ERR_ARG, "a projected level is the state being optimised");
    if (m->getLength() != L) throw Error(MPSD_ERR_SHAPE, "a projected level has another length");
    if (m->ctx() != ctx_) throw Error(MPSD_ERR_ARG, "a projected level lives on another context");
    for (int k = 0; k < L; ++k)
      if (m->getA(k).getd() != init.getA(k).getd())
        throw Error(MPSD_ERR_SHAPE, "a projected level has other physical dimensions");
  }
}

// Contractor::calculateLexck / calculateRexck (Contractor.cpp:2958-2989), iterative.
void Contractor::calculateLexck(int pos, TmpManagerExc& mgr, const MPO& ops, const MPS& ket,
                                const std::vector<MPS*>& levels) {
  int first = pos;
  while (first > 0 && !mgr.ops.isuptodateL(first)) --first;
  for (int k = first + 1; k <= pos; ++k) {
    const Site& a = ket.getA(k - 1);
    check(ctx_,
          mpsd_env_left(ctx_, mgr.ops.operL(k - 1).t, a.device(), ops.getOp(k - 1).device(ctx_), a.device(),
                        mgr.ops.operL(k).t),
          "env_left");
    for (size_t j = 0; j < levels.size(); ++j)
      check(ctx_,
            mpsd_env_left(ctx_, mgr.norms[j]->operL(k - 1).t, levels[j]->getA(k - 1).device(), identityOp(a.getd()),
                          a.device(), mgr.norms[j]->operL(k).t),
            "overlap_left");
    stats_.env_updates += 1 + (int64_t)levels.size();
    mgr.ops.storedL(k);
  }
}

void Contractor::calculateRexck(int pos, TmpManagerExc& mgr, const MPO& ops, const MPS& ket,
                                const std::vector<MPS*>& levels) {
  const int L = ket.getLength();
  int last = pos;
  while (last < L - 1 && !mgr.ops.isuptodateR(last)) ++last;
  for (int k = last - 1; k >= pos; --k) {
    const Site& a = ket.getA(k + 1);
    check(ctx_,
          mpsd_env_right(ctx_, mgr.ops.operR(k + 1).t, a.device(), ops.getOp(k + 1).device(ctx_), a.device(),
                         mgr.ops.operR(k).t),
          "env_right");
    for (size_t j = 0; j < levels.size(); ++j)
      check(ctx_,
            mpsd_env_right(ctx_, mgr.norms[j]->operR(k + 1).t, levels[j]->getA(k + 1).device(),
                           identityOp(a.getd()), a.device(), mgr.norms[j]->operR(k).t),
            "overlap_right");
    stats_.env_updates += 1 + (int64_t)levels.size();
    mgr.ops.storedR(k);
  }
}

// phi_j[s, a, c] = sum_{l, r} M_L[a, l] M_R[c, r] B_j[s, l, r]: the conjugate of the row that
// Contractor::contractMbraId (Contractor.cpp:880-903) builds, as an Heff product with the
// identity operator.
std::vector<const double*> Contractor::effectiveVectors(int pos, TmpManagerExc& mgr, const MPS& init,
                                                        const std::vector<MPS*>& levels) {
  const Site& a = init.getA(pos);
  const int d = a.getd();
  const int64_t n = (int64_t)d * a.getDl() * a.getDr();
  const int64_t need = n * (int64_t)levels.size();
  if (phis_.size() < need) phis_ = DevTensor(ctx_, std::vector<int>{(int)std::max<int64_t>(need, 1)});
  std::vector<const double*> out;
  for (size_t j = 0; j < levels.size(); ++j) {
    const Site& b = levels[j]->getA(pos);
    mpsd_heff* h = nullptr;
    check(ctx_,
          mpsd_heff_create(ctx_, mgr.norms[j]->operL(pos).t, identityOp(d), mgr.norms[j]->operR(pos).t, 0.0,
                           MPSD_PREC_FP64, &h),
          "effective vector");
    const int64_t nin = (int64_t)b.getd() * b.getDl() * b.getDr();
    double* dst = phis_.ptr() + 2 * n * (int64_t)j;
    const int rc = mpsd_heff_apply(h, (const double*)mpsd_tensor_devptr(b.device()), dst, 1, nin, n, 1);
    mpsd_heff_destroy(h);
    check(ctx_, rc, "effective vector");
    out.push_back(dst);
  }
  return out;
}

double Contractor::solvePenaltySite(int pos, TmpManagerExc& mgr, const MPO& ops, const std::vector<double>& weights,
                                    const std::vector<MPS*>& levels, MPS& init, double offset) {
  calculateLexck(pos, mgr, ops, init, levels);
  calculateRexck(pos, mgr, ops, init, levels);
  std::vector<const double*> phis = effectiveVectors(pos, mgr, init, levels);
  // forcing Hermiticity, as the reference does (Contractor.cpp:2637-2638, 2842-2843)
  TensorMultiplierProjMulti multi(ctx_, mgr.ops.operL(pos).t, ops.getOp(pos).device(ctx_), mgr.ops.operR(pos).t,
                                  weights, phis, true, offset);
  double* xv = (double*)mpsd_tensor_devptr(init.site(pos).device());
  EigStats es = eig_->lowest(multi, xv, eigtol_ > 0 ? eigtol_ : 1e-10);
  stats_.matvecs += es.matvecs;
  stats_.final_residual = es.residual;
  return es.value - offset;
}

// ------------------------------------------------------------------------------ sweeps
static void check_part_range(int pos0, int lim, int L) {
  if (pos0 < 0 || pos0 >= L || lim < -1 || lim > L) throw Error(MPSD_ERR_ARG, "sweepPart: positions out of range");
}

void Contractor::sweepPart(int pos0, int lim, const MPO& ops, int D, MPS& init, double offset, int knr) {
  (void)knr;
  const int L = ops.getLength();
  checkLevels(ops, init, {});
  adjustStart(ops, D, init);
  check_part_range(pos0, lim, L);
  stats_ = SolveStats();
  const bool right = lim >= pos0;
  const char dir = right ? 'R' : 'L';
  int pos = pos0;
  for (int k = 0; k < pos; ++k) init.gaugeCond(k, 'R', true);
  for (int k = L - 1; k > pos; --k) init.gaugeCond(k, 'L', true);
  TmpManager mgr(L, ctx_);
  const bool dense = solver_ == fulleig;
  const double tol = eigtol_ > 0 ? eigtol_ : 1e-10;
  double val = 0.0;
  bool done = false;
  while (!done) {
    calculateL(pos, mgr, ops, init, init);
    calculateR(pos, mgr, ops, init, init);
    double* xv = (double*)mpsd_tensor_devptr(init.site(pos).device());
    {
      std::unique_ptr<Multiplier> mp;
      if (dense)
        mp.reset(new TensorMultiplierHermitianOffset(ctx_, mgr.operL(pos).t, ops.getOp(pos).device(ctx_),
                                                     mgr.operR(pos).t, offset));
      else
        mp.reset(new TensorMultiplierOffset(ctx_, mgr.operL(pos).t, ops.getOp(pos).device(ctx_), mgr.operR(pos).t,
                                            offset));
      EigStats es = dense ? eig_->dense(*mp, xv) : eig_->lowest(*mp, xv, tol);
      stats_.matvecs += es.matvecs;
      stats_.final_residual = es.residual;
      val = es.value - offset;
    }
    // the gauge factor goes to the neighbour only at the last step (Contractor.cpp:1330-1333)
    const bool passOnX = (right && pos == lim - 1) || (!right && pos == lim + 1) || pos == lim;
    init.gaugeCond(pos, dir, passOnX);
    for (int l = pos + 1; l < L; ++l) mgr.outofdateL(l);
    for (int l = pos - 1; l >= 0; --l) mgr.outofdateR(l);
    pos += right ? 1 : -1;
    done = right ? pos >= lim : pos <= lim;
  }
  stats_.energies.push_back(val);
}

void Contractor::penaltySweepPart(int pos0, int lim, const MPO& ops, const std::vector<double>& weights,
                                  const std::vector<MPS*>& levels, MPS& init, double offset, bool passAlways) {
  const int L = ops.getLength();
  checkLevels(ops, init, levels);
  check_part_range(pos0, lim, L);
  if (weights.size() != levels.size()) throw Error(MPSD_ERR_ARG, "one weight per projected level");
  stats_ = SolveStats();
  const bool right = lim >= pos0;
  const char dir = right ? 'R' : 'L';
  int pos = pos0;
  for (int k = 0; k < pos; ++k) init.gaugeCond(k, 'R', true);
  for (int k = L - 1; k > pos; --k) init.gaugeCond(k, 'L', true);
  TmpManagerExc mgr(L, (int)levels.size(), ctx_);
  double val = 0.0;
  bool done = false;
  while (!done) {
    val = solvePenaltySite(pos, mgr, ops, weights, levels, init, offset);
    const bool passOnX = passAlways || (right && pos == lim - 1) || (!right && pos == lim + 1) || pos == lim;
    init.gaugeCond(pos, dir, passOnX);
    for (int l = pos + 1; l < L; ++l) mgr.outofdateL(l);
    for (int l = pos - 1; l >= 0; --l) mgr.outofdateR(l);
    pos += right ? 1 : -1;
    done = right ? pos >= lim : pos <= lim;
  }
  stats_.energies.push_back(val);
}

void Contractor::sweepPartWithProjectorPenalty(int pos0, int lim, const MPO& ops, int D,
                                               const std::vector<double>& weights, const std::vector<MPS*>& projected,
                                               MPS& init, double offset, int knr) {
  (void)knr;
  checkLevels(ops, init, projected);
  adjustStart(ops, D, init);
  penaltySweepPart(pos0, lim, ops, weights, projected, init, offset, false);
}

void Contractor::sweepPartNextExcitedState(int pos0, int lim, const MPO& ops, int D,
                                           const std::vector<MPS*>& computedLevels, MPS& init, double offset,
                                           int knr) {
  checkLevels(ops, init, computedLevels);
  std::vector<double> energies;
  for (MPS* m : computedLevels) energies.push_back(contract(*m, ops, *m).real() + offset);
  sweepPartWithProjectorPenalty(pos0, lim, ops, D, energies, computedLevels, init, offset, knr);
}

void Contractor::sweepPartWithProjPenalty(int pos0, int lim, const MPO& ops, int D, double mu, MPS& projected,
                                          MPS& init, double offset, int knr) {
  (void)knr;
  checkLevels(ops, init, {&projected});
  adjustStart(ops, D, init);
  penaltySweepPart(pos0, lim, ops, {-mu}, {&projected}, init, offset, true);
}

void Contractor::findGroundStateWithProjectorPenalty(const MPO& ops, int D, const std::vector<double>& penalties,
                                                     const std::vector<MPS*>& projected, double* lambdak, MPS& init,
                                                     double offset, int knr, const std::string& tmpfile, int maxTime,
                                                     bool useLM) {
  (void)knr;
  if (useLM) throw Error(MPSD_ERR_UNSUPPORTED, "findGroundStateWithProjectorPenalty: useLM is not part of this path");
  if (penalties.size() != projected.size())
    throw Error(MPSD_ERR_ARG, "findGroundStateWithProjectorPenalty: different number of projectors and penalties");
  checkLevels(ops, init, projected);
  adjustStart(ops, D, init);
  const int L = ops.getLength();
  stats_ = SolveStats();
  init.gaugeCond('R', true);
  init.gaugeCond('L', true);
  // energy with the penalty terms (Contractor.cpp:2779-2790); the multiplier subtracts its
  // weights, so they are minus the penalties
  double energy = contract(init, ops, init).real() + offset;
  std::vector<double> weights;
  for (size_t k = 0; k < projected.size(); ++k) {
    const complex_t ov = contract(init, *projected[k]);
    energy += penalties[k] * std::norm(ov);
    weights.push_back(-penalties[k]);
  }
  energy -= offset;
  TmpManagerExc mgr(L, (int)projected.size(), ctx_);
  auto lastSaved = std::chrono::steady_clock::now();
  int pos = 0, round = 0;
  bool right = true, done = false;
  double val = energy;
  while (!done) {
    val = solvePenaltySite(pos, mgr, ops, weights, projected, init, offset);
    // energy-increase check (Contractor.cpp:2851-2870): an error above 1e-6 relative
    const double inc = val - energy, scale = std::fabs(energy + offset);
    if (inc > 1e-8 && inc > scale * 1e-10 && inc > scale * 1e-6) {
      char buf[256];
      snprintf(buf, sizeof buf, "energy increasing in this step: old %.15g new %.15g (round %d, pos %d)", energy,
               val, round, pos);
      throw Error(MPSD_ERR_NUMERIC, buf);
    }
    init.gaugeCond(pos, right ? 'R' : 'L', true);
    for (int l = pos + 1; l < L; ++l) mgr.outofdateL(l);
    for (int l = pos - 1; l >= 0; --l) mgr.outofdateR(l);
    const bool endRight = right && pos == L - 1, endLeft = !right && pos == 0;
    if (!tmpfile.empty() && (endRight || endLeft)) {  // Contractor.cpp:2880-2897
      const auto now = std::chrono::steady_clock::now();
      if (maxTime <= 0 || std::chrono::duration<double>(now - lastSaved).count() >= maxTime) {
        init.exportMPS(tmpfile.c_str());
        lastSaved = now;
      }
    }
    if (endRight || endLeft) {
      // convergence (Contractor.cpp:2901-2951)
      done = std::fabs(energy - val) < convtol_ * std::fabs(energy + offset);
      if (!done && std::fabs(energy - val) < 1e-12) done = true;
      if (!done && round > 1000) done = true;
      energy = val;
      stats_.energies.push_back(val);
      ++round;
      if (max_halfsweeps_ > 0 && round >= max_halfsweeps_) done = true;
      if (!done) {
        if (endRight) {
          init.gaugeCond(pos, 'L', true);
          --pos;
        } else {
          init.gaugeCond(pos, 'R', true);
          ++pos;
        }
        right = !right;
      }
    } else {
      pos += right ? 1 : -1;
    }
  }
  stats_.halfsweeps = round;
  *lambdak = energy;
}

void Contractor::findNextExcitedState(const MPO& ops, int D, const std::vector<MPS*>& computedLevels,
                                      double* lambdak, MPS& init, double offset, int knr, const std::string& tmpfile,
                                      int maxTime, bool useLM) {
  if (useLM) throw Error(MPSD_ERR_UNSUPPORTED, "findNextExcitedState: useLM is not part of this path");
  if (computedLevels.empty()) throw Error(MPSD_ERR_ARG, "findNextExcitedState: no computed levels");
  checkLevels(ops, init, computedLevels);
  // penalty from the spread of the known levels (Contractor.cpp:2682-2707)
  const size_t K = computedLevels.size();
  double minEk = 0.0, maxEk = 0.0;
  for (size_t k = 0; k < K; ++k) {
    const double e = contract(*computedLevels[k], ops, *computedLevels[k]).real() + offset;
    if (k == 0 || e < minEk) minEk = e;
    if (k == 0 || e > maxEk) maxEk = e;
  }
  double penalty = std::max(20.0 * (maxEk - minEk), std::fabs(offset));
  if (K == 1 || penalty < 1e-6) penalty = computedLevels[0]->getLength() * 0.5;
  std::vector<double> penalties(K, penalty);
  int64_t matvecs = 0, envs = 0;
  for (int attempt = 0;; ++attempt) {
    findGroundStateWithProjectorPenalty(ops, D, penalties, computedLevels, lambdak, init, offset, knr, tmpfile,
                                        maxTime, false);
    matvecs += stats_.matvecs;
    envs += stats_.env_updates;
    bool done = true;
    if (excorthtol_ > 0) {  // Contractor.cpp:2712-2724
      for (size_t k = 0; k < K; ++k)
        if (std::abs(contract(init, *computedLevels[k])) > excorthtol_) {
          done = false;
          penalties[k] *= 2.0;
        }
    }
    if (done) break;
    if (attempt >= 40) throw Error(MPSD_ERR_NUMERIC, "findNextExcitedState: no orthogonal state after 40 penalty doublings");
  }
  stats_.matvecs = matvecs;
  stats_.env_updates = envs;
}

}  // namespace mpsdyn
