// Contractor: the single-site ground-state sweep over device-resident MPS and
// environments, and the model MPO builders of the benchmark configurations.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "mpsdyn/mpsdyn.h"

namespace mpsdyn {

Contractor::Contractor(mpsd_ctx* ctx) : ctx_(ctx ? ctx : default_ctx()), eig_(new LanczosSolver(ctx_)) {}

Contractor::~Contractor() {}

Contractor& Contractor::theContractor() {
  static Contractor c(nullptr);
  return c;
}

void Contractor::adjustStart(const MPO& ops, int D, MPS& init) {
  init.adjustPhysDimensions(ops.getDimensions());
  if (D > 0) {
    if (D > init.getBond())
      init.increaseBondDimension(D);
    else if (D < init.getBond())
      init = MPS(init, D);
  }
}

void Contractor::calculateL(int pos, TmpManager& mgr, const MPO& ops, const MPS& ket, const MPS& bra) {
  // Contractor::calculateL (Contractor.cpp:723-742), iterative instead of recursive.
  int first = pos;
  while (first > 0 && !mgr.isuptodateL(first)) --first;
  for (int k = first + 1; k <= pos; ++k) {
    check(ctx_,
          mpsd_env_left(ctx_, mgr.operL(k - 1).t, ket.getA(k - 1).device(), ops.getOp(k - 1).device(ctx_),
                        bra.getA(k - 1).device(), mgr.operL(k).t),
          "env_left");
    ++stats_.env_updates;
    const Site& s = ket.getA(k - 1);
    stats_.env_flops += contraction_flops(s.getd(), s.getDl(), s.getDr(), ops.getOp(k - 1).getDimension(1),
                                          ops.getOp(k - 1).getDimension(3));
    mgr.storedL(k);
  }
}

void Contractor::calculateR(int pos, TmpManager& mgr, const MPO& ops, const MPS& ket, const MPS& bra) {
  // Contractor::calculateR (Contractor.cpp:744-764), iterative instead of recursive.
  const int L = ket.getLength();
  int last = pos;
  while (last < L - 1 && !mgr.isuptodateR(last)) ++last;
  for (int k = last - 1; k >= pos; --k) {
    check(ctx_,
          mpsd_env_right(ctx_, mgr.operR(k + 1).t, ket.getA(k + 1).device(), ops.getOp(k + 1).device(ctx_),
                         bra.getA(k + 1).device(), mgr.operR(k).t),
          "env_right");
    ++stats_.env_updates;
    const Site& s = ket.getA(k + 1);
    stats_.env_flops += contraction_flops(s.getd(), s.getDl(), s.getDr(), ops.getOp(k + 1).getDimension(1),
                                          ops.getOp(k + 1).getDimension(3));
    mgr.storedR(k);
  }
}

complex_t Contractor::contract(const MPS& ket, const MPO& ops, const MPS& bra) {
  // left environments from the scalar 1 over every site, in two alternating slots
  const int L = ket.getLength();
  const complex_t one(1.0, 0.0);
  DevTensor e[2] = {DevTensor(ctx_, mwArray(std::vector<int>{1, 1, 1}, &one)),
                    DevTensor(ctx_, std::vector<int>{1, 1, 1}, false)};
  for (int k = 0; k < L; ++k)
    check(ctx_,
          mpsd_env_left(ctx_, e[k & 1].t, ket.getA(k).device(), ops.getOp(k).device(ctx_), bra.getA(k).device(),
                        e[(k + 1) & 1].t),
          "contract");
  mwArray r = e[L & 1].download();
  return r.getElement(0);
}

void Contractor::findGroundState(const MPO& ops, int D, double* lambda, MPS& init, double offset, int knr,
                                 const std::string& tmpfile, int freqSv, bool useLM) {
  (void)knr;
  if (useLM) throw Error(MPSD_ERR_UNSUPPORTED, "findGroundState: useLM is not part of this path");
  const int L = ops.getLength();
  if (init.getLength() != L) throw Error(MPSD_ERR_SHAPE, "findGroundState: MPS and MPO lengths differ");
  if (init.ctx() != ctx_) throw Error(MPSD_ERR_ARG, "findGroundState: MPS lives on another context");
  stats_ = SolveStats();
  adjustStart(ops, D, init);
  // Contractor.cpp:1371-1373: right then left gauge
  init.gaugeCond('R', true);
  init.gaugeCond('L', true);
  TmpManager mgr(L, ctx_);
  double energy = contract(init, ops, init).real();
  double val = energy;
  int pos = 0, round = 0;
  bool right = true, done = false;
  const double tol = eigtol_ > 0 ? eigtol_ : 1e-10;
  // fulleig: the Hermitian multiplier and the dense eigensolver, in FP64
  // (Contractor.cpp:1403-1404, 3087-3100)
  const bool dense = solver_ == fulleig;
  const bool lowp = precision_ != MPSD_PREC_FP64 && !dense;
  // Tolerance schedule (setEigTolSchedule, and always for the precision forms): the local solves
  // of a half-sweep stop at max(eigtol, min(1e-3, 1e-2 |dE| / |E| of the half-sweep before,
  // 1e-2 times that half-sweep's tolerance)), 1e-3 for the first ones.  While the environments
  // are far from converged a loose local solve loses nothing.  Each half-sweep tightens at least
  // 100-fold: one solved at the tolerance its sites already meet would change nothing, since
  // every local problem stops within its first chunk of steps.  The stopping rule
  // |dE| < convtol |E| is only tested between two half-sweeps solved at eigtol, since loose
  // solves can stall a sweep below convtol before the state has converged.
  const bool sched = (schedule_ || lowp) && !dense;
  double cur_tol = sched ? std::max(tol, 1e-3) : tol;
  int tight = 0;  // consecutive half-sweeps solved at eigtol
  // Precision form: a local solve at tolerance t uses the lowest arithmetic the form allows whose
  // matvec error (about 1e-5 relative for split-BF16, 1e-7 for complex FP32) stays two decades
  // below t; the eigensolver's vectors, the gauge steps and the stopping test stay FP64, and the
  // half-sweeps at eigtol run the FP64 matvec, so the final energy is an FP64 one.
  auto arith_for = [&](double t) {
    if (!lowp) return MPSD_PREC_FP64;
    if (precision_ == MPSD_PREC_BF16X3 && t >= 1e-3) return MPSD_PREC_BF16X3;
    if (t >= 1e-5) return MPSD_PREC_FP32;
    return MPSD_PREC_FP64;
  };
  bool base_fp64 = false;  // the energy-increase guard's baseline comes from an FP64 solve
  // phase breakdown: drain the stream at each phase boundary and charge the wall time since the
  // last boundary to the phase that ended
  auto last = std::chrono::steady_clock::now();
  auto mark = [&](double& acc) {
    if (!breakdown_) return;
    check(ctx_, mpsd_ctx_sync(ctx_), "breakdown sync");
    const auto now = std::chrono::steady_clock::now();
    acc += std::chrono::duration<double>(now - last).count();
    last = now;
  };
  auto hs_start = std::chrono::steady_clock::now();
  int64_t hs_mv0 = 0;
  const unsigned mv_kinds = (1u << MPSD_KERNEL_STAGE1) | (1u << MPSD_KERNEL_MIX) | (1u << MPSD_KERNEL_HEFF_STAGE3) |
                            (1u << MPSD_KERNEL_PACK);
  const int krylov_opt = mpsd_ctx_get_option(ctx_, MPSD_OPT_KRYLOV);
  const int krylov = krylov_opt > 0 ? krylov_opt : 30;
  while (!done) {
    mark(stats_.t_gauge);
    calculateL(pos, mgr, ops, init, init);
    calculateR(pos, mgr, ops, init, init);
    mark(stats_.t_env);
    if (breakdown_) {
      check(ctx_, mpsd_ctx_set_timing_kinds(ctx_, mv_kinds), "timing kinds");
      check(ctx_, mpsd_ctx_set_timing(ctx_, 1), "timing on");
    }
    Site& site = init.site(pos);
    double* xv = (double*)mpsd_tensor_devptr(site.device());
    const double site_flops = contraction_flops(site.getd(), site.getDl(), site.getDr(),
                                                ops.getOp(pos).getDimension(1), ops.getOp(pos).getDimension(3));
    const int arith = arith_for(cur_tol);
    {
      std::unique_ptr<Multiplier> mp;
      if (factory_)
        mp = factory_(ctx_, mgr.operL(pos).t, ops.getOp(pos).device(ctx_), mgr.operR(pos).t, offset, arith);
      else if (dense)
        mp.reset(new TensorMultiplierHermitianOffset(ctx_, mgr.operL(pos).t, ops.getOp(pos).device(ctx_),
                                                     mgr.operR(pos).t, offset, arith));
      else
        mp.reset(new TensorMultiplierOffset(ctx_, mgr.operL(pos).t, ops.getOp(pos).device(ctx_), mgr.operR(pos).t,
                                            offset, arith));
      EigStats es = dense ? eig_->dense(*mp, xv) : eig_->lowest(*mp, xv, cur_tol, krylov);
      stats_.matvecs += es.matvecs;
      if (arith != MPSD_PREC_FP64) stats_.matvecs_lowp += es.matvecs;
      stats_.final_residual = es.residual;
      stats_.matvec_flops += site_flops * (double)es.matvecs;
      val = es.value - offset;
    }
    if (breakdown_) {
      for (int k = 0; k < MPSD_KERNEL_KINDS; ++k) {
        double ms = 0.0;
        int64_t n = 0;
        if ((mv_kinds >> k) & 1u) check(ctx_, mpsd_ctx_timing(ctx_, k, &ms, &n), "timing");
        stats_.t_matvec += 1e-3 * ms;
      }
      check(ctx_, mpsd_ctx_set_timing(ctx_, 0), "timing off");
      check(ctx_, mpsd_ctx_set_timing_kinds(ctx_, ~0u), "timing kinds");
    }
    mark(stats_.t_eig);
    // energy-increase guard (Contractor.cpp:1417-1434), between FP64 energies only
    const bool fp64 = arith == MPSD_PREC_FP64;
    if (fp64 && base_fp64 && (val - energy > 1e-8) && (val - energy > std::fabs(energy + offset) * 1e-10) &&
        (val - energy > std::fabs(energy + offset) * convtol_)) {
      char buf[256];
      snprintf(buf, sizeof buf, "energy increasing in this step: old %.15g new %.15g (round %d, pos %d)", energy,
               val, round, pos);
      throw Error(MPSD_ERR_NUMERIC, buf);
    }
    if (fp64 && !base_fp64) {
      energy = val;  // first FP64 value: the guard's baseline from here on
      base_fp64 = true;
    }
    init.gaugeCond(pos, right ? 'R' : 'L', true);
    for (int l = pos + 1; l < L; ++l) mgr.outofdateL(l);
    for (int l = pos - 1; l >= 0; --l) mgr.outofdateR(l);
    const bool endRight = right && pos == L - 1;
    const bool endLeft = !right && pos == 0;
    if (endRight || endLeft) {
      const auto now = std::chrono::steady_clock::now();
      stats_.hs_seconds.push_back(std::chrono::duration<double>(now - hs_start).count());
      stats_.hs_matvecs.push_back(stats_.matvecs - hs_mv0);
      stats_.hs_tol.push_back(cur_tol);
      hs_start = now;
      hs_mv0 = stats_.matvecs;
      const double dE = std::fabs(energy - val);
      tight = cur_tol <= tol ? tight + 1 : 0;
      done = dE < convtol_ * std::fabs(energy + offset) && (!sched || tight >= 2);
      if (!done && std::fabs(val) < 1e-13) done = true;  // Contractor.cpp:1473-1478
      if (lowp && arith == MPSD_PREC_FP64 && stats_.refine_from < 0) stats_.refine_from = round;
      if (sched && !done) {
        const double rel = dE / std::max(std::fabs(val + offset), 1e-300);
        cur_tol = std::max(tol, std::min(std::min(1e-3, 1e-2 * rel), 1e-2 * cur_tol));
        if (dE < convtol_ * std::fabs(energy + offset)) cur_tol = tol;  // confirm at eigtol
      }
      energy = val;
      base_fp64 = fp64;
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
        if (!tmpfile.empty() && freqSv > 0 && round % freqSv == 0) init.exportMPS(tmpfile.c_str());
      }
    } else {
      pos += right ? 1 : -1;
    }
  }
  mark(stats_.t_gauge);
  stats_.halfsweeps = round;
  *lambda = energy;
}

// ------------------------------------------------------------------------------ models
static const complex_t I_c(0.0, 1.0);

// W[du, Dl, dd, Dr] = sum_o C[Dl, Dr, o] op_o[du, dd] -- the C*Z reshape and
// permute(3,1,4,2) of the reference's model builders (e.g. IsingHamiltonian.cpp:110-115).
static mwArray w_from_c(const std::vector<complex_t>& C, int Dl, int Dr, const std::vector<std::vector<complex_t>>& ops,
                        int d) {
  const int no = (int)ops.size();
  mwArray W(std::vector<int>{d, Dl, d, Dr});
  for (int l = 0; l < Dl; ++l)
    for (int r = 0; r < Dr; ++r)
      for (int o = 0; o < no; ++o) {
        const complex_t c = C[l + Dl * r + Dl * Dr * o];
        if (c == complex_t(0, 0)) continue;
        for (int u = 0; u < d; ++u)
          for (int dn = 0; dn < d; ++dn) {
            const complex_t v = ops[o][u + d * dn];  // column-major d x d
            if (v == complex_t(0, 0)) continue;
            W.data()[u + d * l + d * Dl * dn + d * Dl * d * r] += c * v;
          }
      }
  return W;
}

// column-major 2x2 operators
static const std::vector<complex_t> kI = {1, 0, 0, 1};
static const std::vector<complex_t> kX = {0, 1, 1, 0};
static const std::vector<complex_t> kY = {0, I_c, -I_c, 0};
static const std::vector<complex_t> kZ = {1, 0, 0, -1};

IsingHamiltonian::IsingHamiltonian(int N, int d, double J, double g, double h, double K) : hamil_(N) {
  if (d != 2) throw Error(MPSD_ERR_ARG, "IsingHamiltonian only defined for d=2");
  const std::vector<std::vector<complex_t>> ops = {kI, kZ, kX};
  const complex_t effJ = J < 0 ? std::sqrt(std::fabs(J)) * I_c : complex_t(std::sqrt(J), 0);
  std::shared_ptr<const Operator> bulk;
  for (int k = 0; k < N; ++k) {
    if (k > 1 && k < N - 1 && bulk) {  // IsingHamiltonian.cpp:74-75: bulk sites share one operator
      hamil_.setOpShared(k, bulk);
      continue;
    }
    const int Dl = k == 0 ? 1 : 3, Dr = k == N - 1 ? 1 : 3;
    std::vector<complex_t> C(Dl * Dr * 3, 0.0);
    auto at = [&](int l, int r, int o) -> complex_t& { return C[l + Dl * r + Dl * Dr * o]; };
    if (k != N - 1) at(0, 0, 0) = 1.0;
    if (k != 0) at(Dl - 1, Dr - 1, 0) = 1.0;
    at(0, Dr - 1, 2) = g;
    at(0, Dr - 1, 1) = h;
    at(0, Dr - 1, 0) = K / N;
    if (k != N - 1) at(0, 1, 1) = effJ;
    if (k != 0) at(1, Dr - 1, 1) = effJ;
    auto op = std::make_shared<const Operator>(w_from_c(C, Dl, Dr, ops, 2));
    hamil_.setOpShared(k, op);
    if (k == 1) bulk = op;
  }
}

HeisenbergHamiltonian::HeisenbergHamiltonian(int L, double Jx, double Jy, double Jz, double h, int d, double offset)
    : hamil_(L) {
  if (d != 2) throw Error(MPSD_ERR_ARG, "HeisenbergHamiltonian only supports d=2");
  init(std::vector<double>(std::max(L - 1, 1), Jx), std::vector<double>(std::max(L - 1, 1), Jy),
       std::vector<double>(std::max(L - 1, 1), Jz), std::vector<double>(L, h), offset);
}

HeisenbergHamiltonian::HeisenbergHamiltonian(int L, const std::vector<double>& Jx, const std::vector<double>& Jy,
                                             const std::vector<double>& Jz, const std::vector<double>& h, int d,
                                             double offset)
    : hamil_(L) {
  if (d != 2) throw Error(MPSD_ERR_ARG, "HeisenbergHamiltonian only supports d=2");
  if ((int)Jx.size() != L - 1 || (int)Jy.size() != L - 1 || (int)Jz.size() != L - 1 || (int)h.size() != L)
    throw Error(MPSD_ERR_ARG, "HeisenbergHamiltonian: coefficient vectors of the wrong length");
  init(Jx, Jy, Jz, h, offset);
}

void HeisenbergHamiltonian::init(const std::vector<double>& Jx, const std::vector<double>& Jy,
                                 const std::vector<double>& Jz, const std::vector<double>& h, double offset) {
  // HeisenbergHamiltonian::initZ/initHMPO (HeisenbergHamiltonian.cpp:266-331), no penalty term.
  const int L = hamil_.getLength();
  std::vector<complex_t> sx(4), sy(4), sz(4);
  for (int i = 0; i < 4; ++i) {
    sx[i] = 0.5 * kX[i];
    sy[i] = 0.5 * kY[i];
    sz[i] = 0.5 * kZ[i];
  }
  const std::vector<std::vector<complex_t>> ops = {kI, sx, sy, sz};
  for (int k = 0; k < L; ++k) {
    const int Dl = k == 0 ? 1 : 5, Dr = k == L - 1 ? 1 : 5;
    std::vector<complex_t> C(Dl * Dr * 4, 0.0);
    auto at = [&](int l, int r, int o) -> complex_t& { return C[l + Dl * r + Dl * Dr * o]; };
    if (k != L - 1) {
      at(0, 0, 0) = 1.0;
      at(0, 1, 1) = Jx[k];
      at(0, 2, 2) = Jy[k];
      at(0, 3, 3) = Jz[k];
    }
    if (k != 0) {
      at(Dl - 1, Dr - 1, 0) = 1.0;
      at(1, Dr - 1, 1) = 1.0;
      at(2, Dr - 1, 2) = 1.0;
      at(3, Dr - 1, 3) = 1.0;
    }
    at(0, Dr - 1, 3) = h[k];
    if (offset != 0) at(0, Dr - 1, 0) = offset / L;
    hamil_.setOpShared(k, std::make_shared<const Operator>(w_from_c(C, Dl, Dr, ops, 2)));
  }
}

SchwingerHamiltonian::SchwingerHamiltonian(int N, double mu, double x, double l0, double alpha, double offset,
                                           double nu, double y)
    : hamil_(N) {
  // SchwingerHamiltonian::initZ/initHMPO (SchwingerHamiltonian.cpp:23-121)
  const std::vector<std::vector<complex_t>> ops = {kI, kX, kY, kZ};
  const double a_ = alpha + l0;
  std::vector<double> alphas(N), sumA(N, 0.0);
  for (int k = 0; k < N; ++k) {
    const double sg = (k + 1) % 2 == 0 ? 1.0 : -1.0;
    alphas[k] = a_ - 0.25 * (1.0 - sg);
  }
  double tot = 0.0;
  for (int k = N - 2; k >= 0; --k) {
    tot += alphas[k];
    sumA[k] = tot;
  }
  for (int k = 0; k < N; ++k) {
    const int Dl = k == 0 ? 1 : 5, Dr = k == N - 1 ? 1 : 5;
    std::vector<complex_t> C(Dl * Dr * 4, 0.0);
    auto at = [&](int l, int r, int o) -> complex_t& { return C[l + Dl * r + Dl * Dr * o]; };
    if (k != N - 1) at(0, 0, 0) = 1.0;
    if (k != 0) at(4, Dr - 1, 0) = 1.0;
    if (k > 0 && k < N - 2) at(3, 3, 0) = 1.0;
    const double ak = alphas[k];
    const double sg = (k + 1) % 2 == 0 ? 1.0 : -1.0;
    if (k != N - 1) {
      at(0, Dr - 1, 0) = y * ak * ak + y * (k + 1) / 4.0 + mu / 2.0 + nu / 2.0 + offset / N;
      at(0, Dr - 1, 3) = sg * mu / 2.0 + nu / 2.0 + y * sumA[k];
    } else {
      at(0, Dr - 1, 0) = mu / 2.0 + nu / 2.0 + offset / N;
      at(0, Dr - 1, 3) = sg * mu / 2.0 + nu / 2.0;
    }
    if (k != N - 1) {
      at(0, 1, 1) = std::sqrt(x / 2);
      at(0, 2, 2) = std::sqrt(x / 2);
      if (k < N - 2) at(0, 3, 3) = y;
      if (k != 0) {
        at(1, 4, 1) = std::sqrt(x / 2);
        at(2, 4, 2) = std::sqrt(x / 2);
        at(3, 4, 3) = (N - k - 1) / 2.0;
      }
    } else {
      at(1, 0, 1) = std::sqrt(x / 2);
      at(2, 0, 2) = std::sqrt(x / 2);
    }
    hamil_.setOpShared(k, std::make_shared<const Operator>(w_from_c(C, Dl, Dr, ops, 2)));
  }
}

}  // namespace mpsdyn
