// mpsdyn.h -- C++ host API of the MI355X ground-state path.
//
// Same class and method names, argument meaning and index conventions as the reference
// library (mcbapo/mpsdyn), for the classes on the variational ground-state path:
//   mwArray   (src/wrapper/mwArray.h)     host dense complex tensor, column-major
//   Operator  (src/core/headers/Operator.h)   W[d_up, Dl, d_down, Dr]
//   MPO       (src/core/headers/MPO.h)
//   Site/MPS  (src/core/headers/Site.h, MPS.h)  -- site tensors live in HBM here
//   TmpManager (src/core/headers/TmpManager.h) -- environments live in HBM here
//   Multiplier / TensorMultiplierOffset (src/wrapper/multipliers/*.h)
//   Contractor (src/core/headers/Contractor.h): findGroundState, contract, setters
//   IsingHamiltonian, HeisenbergHamiltonian, SchwingerHamiltonian (src/models/headers/*.h)
// Every device operation goes through the C-ABI of include/mpsd.h.  Errors raise
// mpsdyn::Error (the reference calls exit(212)/exit(1) on the same conditions).
#pragma once

#include <complex>
#include <cstdint>
#include <fstream>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpsd.h"

namespace mpsdyn {

typedef std::complex<double> complex_t;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Throw Error(rc, mpsd_last_error(ctx)) when rc != 0.
void check(mpsd_ctx* ctx, int rc, const char* what);

// Process-wide default device context (device 0), created on first use.  The reference
// keeps its state in the Contractor singleton (Contractor.cpp:49-53); here every object
// may instead be bound to an explicit context (one per concurrent solve).
mpsd_ctx* default_ctx();

// ------------------------------------------------------------------------------ mwArray
class mwArray {
 public:
  mwArray() {}
  explicit mwArray(const std::vector<int>& dims);
  mwArray(const std::vector<int>& dims, const complex_t* data);
  explicit mwArray(complex_t scalar);
  int getRank() const { return (int)dims_.size(); }
  int getDimension(int k) const { return k < (int)dims_.size() ? dims_[k] : 1; }
  const std::vector<int>& getDimensions() const { return dims_; }
  int64_t getNrComponents() const { return (int64_t)data_.size(); }
  complex_t getElement(const std::vector<int>& idx) const;
  void setElement(complex_t v, const std::vector<int>& idx);
  complex_t getElement(int64_t k) const { return data_[k]; }
  void reshape(const std::vector<int>& dims);  // column-major, free
  void permute(const std::vector<int>& order, bool conj = false);  // 1-based order
  void conjugate();
  void fillWithZero();
  const complex_t* data() const { return data_.data(); }
  complex_t* data() { return data_.data(); }
  const double* getComponents() const { return reinterpret_cast<const double*>(data_.data()); }
  // binary layout of mwArray::save/load (src/wrapper/mwArray.cpp:382-443)
  void save(std::ofstream& out) const;
  void load(std::ifstream& in);
  // text layout of mwArray::savetext/loadtext (mwArray.cpp:1672-1717): "rank dims... re im ..."
  // separated by spaces, or only the real parts when real is set; written with 17 significant
  // digits so that a save/load round trip is exact
  void savetext(std::ostream& out, bool real = false) const;
  void loadtext(std::istream& in, bool real = false);

 private:
  std::vector<int> dims_;
  std::vector<complex_t> data_;
};

// A device tensor owned by a host object.
struct DevTensor {
  mpsd_tensor* t = nullptr;
  mpsd_ctx* ctx = nullptr;
  DevTensor() {}
  DevTensor(mpsd_ctx* c, const std::vector<int>& dims);
  // zero = false: contents undefined until written (mpsd_tensor_create_empty)
  DevTensor(mpsd_ctx* c, const std::vector<int>& dims, bool zero);
  DevTensor(mpsd_ctx* c, const mwArray& a);
  ~DevTensor();
  DevTensor(const DevTensor&) = delete;
  DevTensor& operator=(const DevTensor&) = delete;
  DevTensor(DevTensor&& o) noexcept { std::swap(t, o.t), std::swap(ctx, o.ctx); }
  DevTensor& operator=(DevTensor&& o) noexcept {
    std::swap(t, o.t);
    std::swap(ctx, o.ctx);
    return *this;
  }
  std::vector<int> dims() const;
  int64_t size() const { return t ? mpsd_tensor_size(t) : 0; }
  double* ptr() const { return (double*)mpsd_tensor_devptr(t); }
  mwArray download() const;
  void upload(const mwArray& a);
};

// ------------------------------------------------------------------------------ Operator / MPO
class Operator {
 public:
  explicit Operator(const mwArray& data);  // [du, Dl, dd, Dr]; lower ranks padded with 1
  const mwArray& getFullData() const { return data_; }
  int getDimension(int k) const { return data_.getDimension(k); }
  // The operator's tensor on a context's device (uploaded once per context).
  const mpsd_tensor* device(mpsd_ctx* ctx) const;

 private:
  mwArray data_;
  mutable std::vector<std::pair<mpsd_ctx*, std::shared_ptr<DevTensor>>> dev_;
};

class MPO {
 public:
  explicit MPO(int length = 0) : ops_(length) {}
  void initLength(int length) { ops_.assign(length, nullptr); }
  int getLength() const { return (int)ops_.size(); }
  // MPO::setOp (src/core/cpp/MPO.cpp:107-128): takes ownership when myop, else shares.
  void setOp(int k, Operator* op, bool myop = true);
  void setOp(int k, const Operator* shared) { ops_[k] = std::shared_ptr<const Operator>(shared, [](const Operator*) {}); }
  void setOpShared(int k, const std::shared_ptr<const Operator>& op) { ops_[k] = op; }
  const Operator& getOp(int k) const { return *ops_[k]; }
  std::vector<int> getDimensions() const;  // physical dimension per site
  // MPO::saveMPO / loadMPO (src/core/cpp/MPO.cpp:214-232, 293-316): int length, then each
  // operator's [du, Dl, dd, Dr] tensor in mwArray::save layout.
  void saveMPO(const char* filename) const;
  void loadMPO(const char* filename);
  // MPO::exportMPOtext / importMPOtext (MPO.cpp:234-271): "length " then mwArray::savetext of
  // each operator.
  void exportMPOtext(const char* filename) const;
  void importMPOtext(const char* filename);

 private:
  std::vector<std::shared_ptr<const Operator>> ops_;
};

// ------------------------------------------------------------------------------ Site / MPS
class Site {
 public:
  Site(mpsd_ctx* ctx, int pos, const std::vector<int>& dims);  // pos is 1-based like Site_t
  int getPos() const { return pos_; }
  int getd() const { return dims()[0]; }
  int getDl() const { return dims()[1]; }
  int getDr() const { return dims()[2]; }
  std::vector<int> dims() const;
  mwArray getA() const { return dev_.download(); }  // host copy
  void setA(const mwArray& a);
  mpsd_tensor* device() const { return dev_.t; }
  // Site::gaugeR / gaugeL (Site.cpp:178-238): QR / LQ, or a truncated SVD when cutD > 0
  void gaugeR(DevTensor& rightterm, int cutD = 0);
  void gaugeL(DevTensor& leftterm, int cutD = 0);
  void reduceDimR(const DevTensor& leftterm);   // Site.cpp:240-261
  void reduceDimL(const DevTensor& rightterm);  // Site.cpp:263-283
  // gaugeR then next.reduceDimR, gaugeL then prev.reduceDimL, as MPS::gaugeCond runs them
  // (MPS.cpp:306-337, 366-410): one launch for sites the single-workgroup QR takes
  void gaugeRInto(DevTensor& rightterm, Site& next);
  void gaugeLInto(DevTensor& leftterm, Site& prev);
  // Site::changeDimensions (Site.cpp:525-530): the tensor resized to [d, Dl, Dr] by index,
  // zero-padded where it grows and cut where it shrinks (mwArray::resize with zero noise)
  void changeDimensions(int d, int Dl, int Dr);
  void increaseBondDim(int Dl, int Dr) { changeDimensions(getd(), Dl, Dr); }  // Site.cpp:532-534
  void increasePhysDim(int d) {                                                // Site.cpp:545-552
    if (d > getd()) changeDimensions(d, getDl(), getDr());
  }
  void decreasePhysDim(int d) {  // Site.cpp:553-559
    if (d < getd()) changeDimensions(d, getDl(), getDr());
  }

 private:
  mpsd_ctx* ctx_;
  int pos_;
  DevTensor dev_;
};

enum Gauge { gNone, gLeft, gRight, gBoth };

class MPS {
 public:
  MPS() {}
  // MPS(len, bond, d) (src/core/cpp/MPS.cpp:15-35): sites [d,1,D], [d,D,D]..., [d,D,1].
  MPS(int length, int bond, int d, mpsd_ctx* ctx = nullptr);
  // MPS(len, bond, vector<int> d) (MPS.cpp:68-87): site k has physical dimension d[k].
  MPS(int length, int bond, const std::vector<int>& d, mpsd_ctx* ctx = nullptr);
  // MPS(const MPS&, D) (MPS.cpp:761-773): a copy truncated to bond D by gaugeCond('L', false, D).
  MPS(const MPS& other, int D);
  int getLength() const { return (int)A_.size(); }
  int getBond() const { return bond_; }
  const Site& getA(int pos) const { return *A_[pos]; }
  Site& site(int pos) { return *A_[pos]; }
  void setA(int pos, const mwArray& a) { A_[pos]->setA(a); }
  mpsd_ctx* ctx() const { return ctx_; }
  // Seeded complex-Gaussian state (SplitMix64 + Box-Muller, site k draws from seed*1000003+k)
  // at the canonical bond profile min(d^k, D, d^(L-k)), then right-gauged.  The reference
  // fills uniform [0,1) values from libc random() (MPS.cpp:282-294).
  void setRandomState(uint64_t seed = 1);
  // MPS.cpp:306-364: gauge every site; with cutD > 0 each bond is cut by a truncated SVD to at
  // most cutD; without normalisation the last factor's magnitude goes into the norm factor
  void gaugeCond(char dir, bool normal = false, int cutD = 0);
  void gaugeCond(int pos, char dir, bool next = false);  // MPS.cpp:366-419
  // MPS::increaseBondDimension (MPS.cpp:650-719, zero noise): every bond set to
  // min(D, d-product to its left, d-product to its right); sites zero-padded (or cut where an
  // old bond exceeds that profile); a no-op when D <= getBond()
  void increaseBondDimension(int D);
  // MPS::adjustPhysDimensions (MPS.cpp:722-756): site k resized to physical dimension d[k];
  // when any site changed, the bond profile is rebuilt by increaseBondDimension(getBond())
  void adjustPhysDimensions(const std::vector<int>& d);
  Gauge getGauge() const { return gauge_; }
  double getNormFact() const { return normfact_; }
  // Binary format of MPS::exportMPS/importMPS (MPS.cpp:442-509, Site.cpp:561-578).
  void exportMPS(const char* filename) const;
  void importMPS(const char* filename);

 private:
  DevTensor& factor();  // reusable device buffer for the QR/LQ factor of a gauge move
  mpsd_ctx* ctx_ = nullptr;
  int bond_ = 0;
  double normfact_ = 1.0;
  Gauge gauge_ = gNone;
  std::vector<std::unique_ptr<Site>> A_;
  std::unique_ptr<DevTensor> factor_;
};

// ------------------------------------------------------------------------------ environments
// TmpManager (src/core/cpp/TmpManager.cpp:22-211): operL/operR slots with up-to-date flags;
// edges are the scalar 1.  The slots are HBM tensors that keep their capacity.
class TmpManager {
 public:
  TmpManager(int length, mpsd_ctx* ctx);
  bool isuptodateL(int k) const { return upL_[k]; }
  bool isuptodateR(int k) const { return upR_[k]; }
  void outofdateL(int k) { upL_[k] = false; }
  void outofdateR(int k) { upR_[k] = false; }
  void storedL(int k) { upL_[k] = true; }
  void storedR(int k) { upR_[k] = true; }
  DevTensor& operL(int k) { return L_[k]; }
  DevTensor& operR(int k) { return R_[k]; }
  int64_t deviceBytes() const;

 private:
  std::vector<DevTensor> L_, R_;
  std::vector<bool> upL_, upR_;
};

// TmpManagerExc (src/core/headers/TmpManagerExc.h): the operator environments of the state
// being optimised plus, per projected level k, its overlap environments with that state.
// One set of up-to-date flags (the operator slots') covers all of them, as they are
// refreshed together.  Overlap slot k holds M[a, 1, l] = <state|level_k> contracted up to the
// bond (a: the state's bond, l: the level's), the conjugate transpose of the reference's
// normL/normR, so that the projected vector phi_k is a plain Heff product.
struct TmpManagerExc {
  TmpManagerExc(int length, int nlevels, mpsd_ctx* ctx) : ops(length, ctx) {
    for (int k = 0; k < nlevels; ++k) norms.emplace_back(new TmpManager(length, ctx));
  }
  void outofdateL(int k) { ops.outofdateL(k); }
  void outofdateR(int k) { ops.outofdateR(k); }
  TmpManager ops;
  std::vector<std::unique_ptr<TmpManager>> norms;
};

// ------------------------------------------------------------------------------ Multiplier
// Multiplier (src/wrapper/multipliers/Multiplier.h:18-49) with a device product: x and y
// are device vectors (complex128) of getSize() elements, ordered on the context's stream.
class Multiplier {
 public:
  virtual ~Multiplier() {}
  virtual int64_t getSize() const = 0;
  virtual void product_device(const double* x, double* y) = 0;
  // The reference's host form: product(const mwArray& x, mwArray& y).
  virtual void product(const mwArray& x, mwArray& y);
  virtual void getFullTensor(mwArray& result) = 0;
  virtual mpsd_ctx* ctx() const = 0;
  virtual void setPrecision(int) {}
  // The C-ABI handle when the product is a plain device Heff (enables the graph-driven
  // eigensolver path), else null.
  virtual mpsd_heff* heff() const { return nullptr; }
  // The C-ABI handle the one-workgroup eigensolver may use for a small local problem (it
  // then runs the whole eigensolve itself), else null.
  virtual mpsd_heff* smallHeff() const { return heff(); }
};

// TensorMultiplierOffset (TensorMultiplierOffset.cpp:8-25): Heff + offset, on the device.
class TensorMultiplierOffset : public Multiplier {
 public:
  TensorMultiplierOffset(mpsd_ctx* ctx, const mpsd_tensor* L, const mpsd_tensor* W, const mpsd_tensor* R,
                         double offset, int precision = MPSD_PREC_FP64);
  ~TensorMultiplierOffset();
  int64_t getSize() const override;
  void product_device(const double* x, double* y) override;
  void getFullTensor(mwArray& result) override;
  mpsd_ctx* ctx() const override { return ctx_; }
  void setPrecision(int p) override;
  mpsd_heff* handle() const { return h_; }
  mpsd_heff* heff() const override { return h_; }

 private:
  mpsd_ctx* ctx_;
  mpsd_heff* h_ = nullptr;
  std::vector<int> outdims_;
};

// TensorMultiplierHermitianOffset (src/wrapper/multipliers/TensorMultiplierHermitianOffset.h,
// TensorMultiplierHermitian.cpp:8-66): product = (Heff + Heff^H) x / 2 + offset x and
// getFullTensor = (H + H^H) / 2 + offset I.  The conjugate-transposed operand copies are taken
// at construction, as in the reference constructor.
class TensorMultiplierHermitianOffset : public TensorMultiplierOffset {
 public:
  TensorMultiplierHermitianOffset(mpsd_ctx* ctx, const mpsd_tensor* L, const mpsd_tensor* W, const mpsd_tensor* R,
                                  double offset, int precision = MPSD_PREC_FP64);
};

// TensorMultiplierProjMulti (src/wrapper/multipliers/TensorMultiplierProjMulti.cpp:13-100):
// product = B x - sum_k weights[k] <phi_k|x> phi_k + offset x, with B the plain Heff or, when
// hermitian, its Hermitian average (Heff + Heff^H) / 2.  getFullTensor is that matrix,
// averaged with its adjoint as in the reference.  The reference takes the projected vectors as
// bras v_k (rows, from Contractor::contractMbraId); here they are the kets phi_k = v_k^H as
// device vectors of getSize() complex numbers, which the caller keeps alive.
class TensorMultiplierProjMulti : public Multiplier {
 public:
  TensorMultiplierProjMulti(mpsd_ctx* ctx, const mpsd_tensor* L, const mpsd_tensor* W, const mpsd_tensor* R,
                            const std::vector<double>& weights, const std::vector<const double*>& phis,
                            bool hermitian = false, double offset = 0.0);
  int64_t getSize() const override { return base_->getSize(); }
  void product_device(const double* x, double* y) override;
  void getFullTensor(mwArray& result) override;
  mpsd_ctx* ctx() const override { return ctx_; }

 private:
  mpsd_ctx* ctx_;
  std::unique_ptr<TensorMultiplierOffset> base_;
  std::vector<double> weights_;
  std::vector<const double*> phis_;
};

// TensorMultiplierProj (TensorMultiplierProj.cpp:9-41): one projected vector, no offset:
// product = B x - weight <phi|x> phi.
class TensorMultiplierProj : public TensorMultiplierProjMulti {
 public:
  TensorMultiplierProj(mpsd_ctx* ctx, const mpsd_tensor* L, const mpsd_tensor* W, const mpsd_tensor* R, double weight,
                       const double* phi, bool hermitian = false)
      : TensorMultiplierProjMulti(ctx, L, W, R, {weight}, {phi}, hermitian, 0.0) {}
};

// ------------------------------------------------------------------------------ eigensolver
// Lowest eigenpair of a Hermitian Multiplier by restarted Lanczos with full (two-pass)
// reorthogonalisation on device vectors; the tridiagonal Rayleigh-Ritz step is FP64 on the
// host.  Stands in for PRIMME/ARPACK (src/wrapper/eigs.cpp:71-422), whose sources are not
// part of the reference.
struct EigStats {
  int64_t graph_replays = 0;
  double value = 0.0;
  double residual = 0.0;  // ||H x - value x|| (Lanczos estimate)
  int matvecs = 0;
  int restarts = 0;
  bool converged = false;
};

class LanczosSolver {
 public:
  explicit LanczosSolver(mpsd_ctx* ctx) : ctx_(ctx) {}
  ~LanczosSolver();
  // x: device vector, initial guess on entry (warm start), eigenvector on exit (unit norm).
  EigStats lowest(Multiplier& H, double* x, double tol, int max_krylov = 30, int max_restarts = 40);
  // Dense lowest eigenpair (mpsd_heff_eigs_dense): the fulleig solver and the reference's
  // small-size switch (Contractor.cpp:3087-3100, 3126-3133).
  EigStats dense(Multiplier& H, double* x);

 private:
  EigStats lowest_graph(Multiplier& H, mpsd_heff* h, double* x, double tol, int m, int max_restarts);
  mpsd_lanczos* lz_ = nullptr;
  int lz_m_ = 0;
  mpsd_ctx* ctx_;
  DevTensor V_, coef_;
  int64_t cap_ = 0;
  int coef_cap_ = 0;
};

// Lowest eigenpair of a symmetric tridiagonal matrix (diag a[0..m), offdiag b[0..m-1)):
// bisection on the Sturm count, then inverse iteration.  Returns theta; s gets the unit
// eigenvector.
double tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, std::vector<double>& s);

// ------------------------------------------------------------------------------ Contractor
enum EigSolver { primme, arpack, fulleig, lanczos };

struct SolveStats {
  int halfsweeps = 0;
  int64_t matvecs = 0;        // all matvecs
  int64_t matvecs_lowp = 0;   // of which in the low-precision form (precision != FP64)
  int refine_from = -1;       // first half-sweep with FP64 refinement (-1: FP64 throughout)
  int64_t env_updates = 0;
  double final_residual = 0.0;
  std::vector<double> energies;  // energy at the end of each half-sweep
  // per half-sweep: wall seconds, matvecs, and the local eigensolver tolerance it ran at
  std::vector<double> hs_seconds, hs_tol;
  std::vector<int64_t> hs_matvecs;
  // Wall-clock split of findGroundState's sweep in seconds, filled when the breakdown is on
  // (Contractor::setBreakdown): environment updates, local eigensolves (matvecs plus the
  // eigensolver's vector work), gauge steps; t_matvec is the matvec kernels' time inside the
  // eigensolves, from stream events.
  double t_env = 0.0, t_eig = 0.0, t_gauge = 0.0, t_matvec = 0.0;
  // algorithmic flops (complex MAC = 8) of the matvecs and of the environment updates done
  double matvec_flops = 0.0, env_flops = 0.0;
};

// Algorithmic real flops of one matvec or environment update over a site [d, Dl, Dr] with MPO
// bonds wl, wr (bra and ket bonds equal): 8 (d wl Dl^2 Dr + d^2 wl wr Dl Dr + d wr Dl Dr^2).
inline double contraction_flops(int d, int Dl, int Dr, int wl, int wr) {
  return 8.0 * ((double)d * wl * Dl * Dl * Dr + (double)d * d * wl * wr * Dl * Dr + (double)d * wr * Dl * Dr * Dr);
}

// Builds the Multiplier a sweep applies at one site (default: TensorMultiplierOffset on the
// Contractor's context).  A serving runtime substitutes one that batches matvecs of
// concurrent solves.
typedef std::function<std::unique_ptr<Multiplier>(mpsd_ctx*, const mpsd_tensor* L, const mpsd_tensor* W,
                                                  const mpsd_tensor* R, double offset, int precision)>
    MultiplierFactory;

class Contractor {
 public:
  explicit Contractor(mpsd_ctx* ctx = nullptr);
  ~Contractor();
  static Contractor& theContractor();  // Contractor.cpp:49-53
  void setConvTol(double tol) { convtol_ = tol; }
  void setEigTol(double tol) { eigtol_ = tol; }
  void setEigenSolver(EigSolver s) { solver_ = s; }
  double getConvTol() const { return convtol_; }
  // Limit on half-sweeps (the reference has none); <= 0 means unlimited.
  void setMaxHalfSweeps(int n) { max_halfsweeps_ = n; }
  void setPrecision(int p) { precision_ = p; }
  void setMultiplierFactory(MultiplierFactory f) { factory_ = std::move(f); }
  // Time the phases of findGroundState (SolveStats::t_*): the stream is drained at each phase
  // boundary, which costs a little time, so it is off by default.
  void setBreakdown(bool on) { breakdown_ = on; }
  // Local-tolerance schedule of findGroundState: off (default) solves every local problem to
  // the eigensolver tolerance, as the reference does; on, the local solves of a half-sweep stop
  // at max(eigtol, min(1e-3, 1e-2 |dE| / |E| of the previous half-sweep, 1e-2 times the previous
  // half-sweep's tolerance)), 1e-3 at the start.  The precision forms always run the schedule.
  void setEigTolSchedule(bool on) { schedule_ = on; }
  // Contractor::findGroundState (Contractor.h:255-257, Contractor.cpp:1352-1533).  knr, the number
  // of local eigenpairs the reference asks its eigensolver for, of which it keeps the lowest
  // (Contractor.cpp:1407-1416), is accepted: the device eigensolver computes that lowest pair.
  void findGroundState(const MPO& ops, int D, double* lambda, MPS& init, double offset = 0.,
                       int knr = 1, const std::string& tmpfile = "", int freqSv = 200, bool useLM = false);
  // <bra|ops|ket> via a full left-environment sweep (Contractor.cpp:528).
  complex_t contract(const MPS& ket, const MPO& ops, const MPS& bra);
  // <bra|ket> times both norm factors (Contractor::contract(ket, bra), Contractor.cpp:467-499).
  complex_t contract(const MPS& ket, const MPS& bra);
  const SolveStats& lastStats() const { return stats_; }

  // Part of a ground-state sweep, sites pos0 .. lim (excluded), moving right when lim >= pos0
  // (Contractor::sweepPart, Contractor.cpp:1260-1350).  Sites before pos0 are right-gauged
  // and sites after it left-gauged first; each site's gauge factor is dropped except at the
  // last step, as in the reference.  The start state is first fitted to ops and D as
  // findGroundState does.
  void sweepPart(int pos0, int lim, const MPO& ops, int D, MPS& init, double offset = 0., int knr = 1);

  // Excited states by projector penalties (Contractor.cpp:2561-2957).  The local problem at
  // each site is TensorMultiplierProjMulti in its Hermitian form with the projected levels'
  // effective vectors, solved by the device Lanczos solver whatever setEigenSolver says.
  //  * findGroundStateWithProjectorPenalty: ground state of H + sum_k penalties[k] |k><k|
  //    (the reference's `energies` argument, Contractor.cpp:2739-2956).
  //  * findNextExcitedState: the penalties are max(20 (maxE - minE), |offset|) over the given
  //    levels' energies, or L/2 for a single (or degenerate) level; with setExcOrthTol > 0 the
  //    solve repeats, doubling the penalty of each level the result overlaps by more than that
  //    (Contractor.cpp:2676-2727).
  //  * sweepPartWithProjectorPenalty: sweepPart on H - sum_k weights[k] |k><k| (the sign the
  //    reference's multiplier gives its `energies` there, Contractor.cpp:2575-2666);
  //    sweepPartNextExcitedState uses weights = the levels' energies + offset;
  //    sweepPartWithProjPenalty adds mu |projected><projected| (Contractor.cpp:3640-3734).
  // The largest-magnitude variants (useLM) are not part of this path.
  void findGroundStateWithProjectorPenalty(const MPO& ops, int D, const std::vector<double>& penalties,
                                           const std::vector<MPS*>& projected, double* lambdak, MPS& init,
                                           double offset = 0., int knr = 1, const std::string& tmpfile = "",
                                           int maxTime = 0, bool useLM = false);
  void findNextExcitedState(const MPO& ops, int D, const std::vector<MPS*>& computedLevels, double* lambdak,
                            MPS& init, double offset = 0., int knr = 1, const std::string& tmpfile = "",
                            int maxTime = 0, bool useLM = false);
  void sweepPartWithProjectorPenalty(int pos0, int lim, const MPO& ops, int D, const std::vector<double>& weights,
                                     const std::vector<MPS*>& projected, MPS& init, double offset = 0., int knr = 1);
  void sweepPartNextExcitedState(int pos0, int lim, const MPO& ops, int D, const std::vector<MPS*>& computedLevels,
                                 MPS& init, double offset = 0., int knr = 1);
  void sweepPartWithProjPenalty(int pos0, int lim, const MPO& ops, int D, double mu, MPS& projected, MPS& init,
                                double offset = 0., int knr = 1);
  void setExcOrthTol(double v) { excorthtol_ = v; }
  double getExcOrthTol() const { return excorthtol_; }

 private:
  void calculateL(int pos, TmpManager& mgr, const MPO& ops, const MPS& ket, const MPS& bra);
  void calculateR(int pos, TmpManager& mgr, const MPO& ops, const MPS& ket, const MPS& bra);
  void calculateLexck(int pos, TmpManagerExc& mgr, const MPO& ops, const MPS& ket, const std::vector<MPS*>& levels);
  void calculateRexck(int pos, TmpManagerExc& mgr, const MPO& ops, const MPS& ket, const std::vector<MPS*>& levels);
  // phi_k for every level at pos into phis_ (one column per level); returns the column pointers
  std::vector<const double*> effectiveVectors(int pos, TmpManagerExc& mgr, const MPS& init,
                                              const std::vector<MPS*>& levels);
  // the [d, 1, d, 1] identity operator on this context
  const mpsd_tensor* identityOp(int d);
  // one site of a projector-penalty sweep: solve, store, return the eigenvalue minus offset
  double solvePenaltySite(int pos, TmpManagerExc& mgr, const MPO& ops, const std::vector<double>& weights,
                          const std::vector<MPS*>& levels, MPS& init, double offset);
  void penaltySweepPart(int pos0, int lim, const MPO& ops, const std::vector<double>& weights,
                        const std::vector<MPS*>& levels, MPS& init, double offset, bool passAlways);
  void checkLevels(const MPO& ops, const MPS& init, const std::vector<MPS*>& levels) const;
  // The start state fitted to a call (Contractor.cpp:1360-1370, 1263-1279, 2580-2593, 2755-2766,
  // 3650-3660): physical dimensions from ops (MPS::adjustPhysDimensions), then, for D > 0, the
  // bond grown by zero padding (MPS::increaseBondDimension) or cut by the truncating copy
  // MPS(init, D)
  void adjustStart(const MPO& ops, int D, MPS& init);
  double excorthtol_ = 0.0;
  std::vector<std::pair<int, std::unique_ptr<Operator>>> ident_;
  DevTensor phis_;
  mpsd_ctx* ctx_;
  double convtol_ = 1e-5, eigtol_ = 0.0;
  EigSolver solver_ = primme;
  int max_halfsweeps_ = 0;
  int precision_ = MPSD_PREC_FP64;
  SolveStats stats_;
  std::unique_ptr<LanczosSolver> eig_;
  MultiplierFactory factory_;
  bool breakdown_ = false;
  bool schedule_ = false;
};

// ------------------------------------------------------------------------------ models
class Hamiltonian {
 public:
  virtual ~Hamiltonian() {}
  virtual const MPO& getHMPO() const = 0;
};

// H = sum J sz sz + g sx + h sz (+K) (IsingHamiltonian.cpp:17-121).
class IsingHamiltonian : public Hamiltonian {
 public:
  IsingHamiltonian(int N, int d, double J, double g, double h = 0., double K = 0.);
  const MPO& getHMPO() const override { return hamil_; }

 private:
  MPO hamil_;
};

// H = sum Jx SxSx + Jy SySy + Jz SzSz + h Sz, S = sigma/2 (HeisenbergHamiltonian.cpp:27-54,266-331).
class HeisenbergHamiltonian : public Hamiltonian {
 public:
  HeisenbergHamiltonian(int L, double Jx, double Jy, double Jz, double h, int d = 2, double offset = 0.);
  HeisenbergHamiltonian(int L, const std::vector<double>& Jx, const std::vector<double>& Jy,
                        const std::vector<double>& Jz, const std::vector<double>& h, int d = 2,
                        double offset = 0.);
  const MPO& getHMPO() const override { return hamil_; }

 private:
  void init(const std::vector<double>& Jx, const std::vector<double>& Jy, const std::vector<double>& Jz,
            const std::vector<double>& h, double offset);
  MPO hamil_;
};

// Lattice Schwinger model (SchwingerHamiltonian.cpp:13-121).
class SchwingerHamiltonian : public Hamiltonian {
 public:
  SchwingerHamiltonian(int N, double mu, double x, double l0, double alpha, double offset = 0,
                       double nu = 0., double y = 1.);
  const MPO& getHMPO() const override { return hamil_; }

 private:
  MPO hamil_;
};

// SplitMix64 + Box-Muller complex Gaussians, column-major over dims (same draws as the
// Python test generator).
mwArray complex_gaussian(uint64_t seed, const std::vector<int>& dims);

// The MPO of the operator product A B: site tensors contracted over the inner physical index
// (A's d_down with B's d_up), bonds merged as (a1, a2) -> a1 + Dl_A a2.
MPO mpo_product(const MPO& a, const MPO& b);

}  // namespace mpsdyn
