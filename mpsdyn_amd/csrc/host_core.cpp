// Host C++ data model of the ground-state path: mwArray, device tensors, Operator/MPO,
// Site/MPS (site tensors in HBM) and TmpManager (environments in HBM).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "mpsdyn/mpsdyn.h"

namespace mpsdyn {

void check(mpsd_ctx* ctx, int rc, const char* what) {
  if (rc != 0) throw Error(rc, std::string(what) + ": " + (ctx ? mpsd_last_error(ctx) : "error"));
}

mpsd_ctx* default_ctx() {
  static std::once_flag once;
  static mpsd_ctx* ctx = nullptr;
  std::call_once(once, [] {
    int rc = mpsd_ctx_create(0, &ctx);
    if (rc) throw Error(rc, "mpsd_ctx_create(0) failed: no HIP device");
  });
  return ctx;
}

// ------------------------------------------------------------------------------ mwArray
static int64_t prod(const std::vector<int>& d) {
  int64_t n = 1;
  for (int x : d) n *= x;
  return n;
}

mwArray::mwArray(const std::vector<int>& dims) : dims_(dims), data_(prod(dims)) {}

mwArray::mwArray(const std::vector<int>& dims, const complex_t* data) : dims_(dims), data_(data, data + prod(dims)) {}

mwArray::mwArray(complex_t scalar) : dims_{1}, data_{scalar} {}

static int64_t lin(const std::vector<int>& dims, const std::vector<int>& idx) {
  int64_t off = 0, stride = 1;
  for (size_t k = 0; k < dims.size(); ++k) {
    const int i = k < idx.size() ? idx[k] : 0;
    if (i < 0 || i >= dims[k]) throw Error(MPSD_ERR_ARG, "mwArray index out of range");
    off += stride * i;
    stride *= dims[k];
  }
  return off;
}

complex_t mwArray::getElement(const std::vector<int>& idx) const { return data_[lin(dims_, idx)]; }

void mwArray::setElement(complex_t v, const std::vector<int>& idx) { data_[lin(dims_, idx)] = v; }

void mwArray::reshape(const std::vector<int>& dims) {
  std::vector<int> nd = dims;
  int64_t known = 1;
  int neg = -1;
  for (size_t k = 0; k < nd.size(); ++k) {
    if (nd[k] < 0) neg = (int)k;
    else known *= nd[k];
  }
  if (neg >= 0) nd[neg] = (int)(data_.size() / std::max<int64_t>(known, 1));
  if (prod(nd) != (int64_t)data_.size()) throw Error(MPSD_ERR_SHAPE, "mwArray::reshape: size mismatch");
  dims_ = nd;
}

void mwArray::permute(const std::vector<int>& order, bool conj) {
  const int r = getRank();
  if ((int)order.size() != r) throw Error(MPSD_ERR_SHAPE, "mwArray::permute: wrong order length");
  std::vector<int> nd(r);
  std::vector<int64_t> ostride(r), nstride(r);
  int64_t s = 1;
  for (int k = 0; k < r; ++k) {
    ostride[k] = s;
    s *= dims_[k];
  }
  for (int k = 0; k < r; ++k) nd[k] = dims_[order[k] - 1];
  std::vector<complex_t> out(data_.size());
  std::vector<int> idx(r, 0);
  // iterate over new layout in column-major order
  for (int64_t t = 0; t < (int64_t)out.size(); ++t) {
    int64_t src = 0;
    for (int k = 0; k < r; ++k) src += idx[k] * ostride[order[k] - 1];
    out[t] = conj ? std::conj(data_[src]) : data_[src];
    for (int k = 0; k < r; ++k) {
      if (++idx[k] < nd[k]) break;
      idx[k] = 0;
    }
  }
  data_.swap(out);
  dims_ = nd;
}

void mwArray::conjugate() {
  for (auto& v : data_) v = std::conj(v);
}

void mwArray::fillWithZero() { std::fill(data_.begin(), data_.end(), complex_t(0.0, 0.0)); }

void mwArray::save(std::ofstream& out) const {
  int rank = getRank();
  out.write((const char*)&rank, sizeof(int));
  out.write((const char*)dims_.data(), sizeof(int) * rank);
  out.write((const char*)data_.data(), sizeof(complex_t) * data_.size());
  if (!out) throw Error(MPSD_ERR_ARG, "mwArray::save: write failed");
}

void mwArray::load(std::ifstream& in) {
  int rank = 0;
  in.read((char*)&rank, sizeof(int));
  if (!in || rank < 0 || rank > 64) throw Error(MPSD_ERR_ARG, "mwArray::load: bad rank");
  dims_.assign(rank, 0);
  in.read((char*)dims_.data(), sizeof(int) * rank);
  if (!in) throw Error(MPSD_ERR_ARG, "mwArray::load: bad dimensions");
  data_.assign(prod(dims_), complex_t(0, 0));
  in.read((char*)data_.data(), sizeof(complex_t) * data_.size());
  if (!in) throw Error(MPSD_ERR_ARG, "mwArray::load: truncated data");
}

void mwArray::savetext(std::ostream& out, bool real) const {
  const int rank = getRank();
  out << rank << " ";
  for (int k = 0; k < rank; ++k) out << dims_[k] << " ";
  const std::streamsize prec = out.precision(17);
  for (const complex_t& c : data_) {
    out << c.real() << " ";
    if (!real) out << c.imag() << " ";
  }
  out.precision(prec);
  if (!out) throw Error(MPSD_ERR_ARG, "mwArray::savetext: write failed");
}

void mwArray::loadtext(std::istream& in, bool real) {
  int rank = 0;
  in >> rank;
  if (!in || rank < 0 || rank > 64) throw Error(MPSD_ERR_ARG, "mwArray::loadtext: bad rank");
  dims_.assign(rank, 0);
  for (int k = 0; k < rank; ++k) {
    in >> dims_[k];
    if (!in || dims_[k] < 0) throw Error(MPSD_ERR_ARG, "mwArray::loadtext: bad dimensions");
  }
  data_.assign(prod(dims_), complex_t(0, 0));
  for (complex_t& c : data_) {
    double re = 0.0, im = 0.0;
    in >> re;
    if (!real) in >> im;
    if (!in) throw Error(MPSD_ERR_ARG, "mwArray::loadtext: truncated data");
    c = complex_t(re, im);
  }
}

// ------------------------------------------------------------------------------ DevTensor
DevTensor::DevTensor(mpsd_ctx* c, const std::vector<int>& dims) : ctx(c) {
  std::vector<int> d = dims.empty() ? std::vector<int>{1} : dims;
  check(ctx, mpsd_tensor_create(ctx, (int)d.size(), d.data(), &t), "tensor_create");
}

DevTensor::DevTensor(mpsd_ctx* c, const std::vector<int>& dims, bool zero) : ctx(c) {
  std::vector<int> d = dims.empty() ? std::vector<int>{1} : dims;
  if (zero)
    check(ctx, mpsd_tensor_create(ctx, (int)d.size(), d.data(), &t), "tensor_create");
  else
    check(ctx, mpsd_tensor_create_empty(ctx, (int)d.size(), d.data(), &t), "tensor_create_empty");
}

DevTensor::DevTensor(mpsd_ctx* c, const mwArray& a) : ctx(c) {
  const std::vector<int>& d = a.getDimensions();
  int64_t n = 1;
  for (int x : d) n *= x;
  if (d.empty() || n != a.getNrComponents()) {
    // shapes the upload rejects: the plain create and upload report it
    const std::vector<int> d1 = d.empty() ? std::vector<int>{1} : d;
    check(ctx, mpsd_tensor_create(ctx, (int)d1.size(), d1.data(), &t), "tensor_create");
    upload(a);
    return;
  }
  check(ctx, mpsd_tensor_create_from(ctx, (int)d.size(), d.data(), a.getComponents(), &t), "tensor_create_from");
}

DevTensor::~DevTensor() {
  if (t) mpsd_tensor_destroy(t);
}

std::vector<int> DevTensor::dims() const {
  int r = mpsd_tensor_rank(t);
  std::vector<int> d(std::max(r, 0));
  if (r > 0) mpsd_tensor_dims(t, d.data());
  return d;
}

mwArray DevTensor::download() const {
  mwArray a(dims());
  check(ctx, mpsd_tensor_download(t, (double*)a.data(), a.getNrComponents()), "tensor_download");
  return a;
}

void DevTensor::upload(const mwArray& a) {
  if (dims() != a.getDimensions()) {
    const auto& d = a.getDimensions();
    check(ctx, mpsd_tensor_reshape(t, (int)d.size(), d.data()), "tensor_reshape");
  }
  check(ctx, mpsd_tensor_upload(t, a.getComponents(), a.getNrComponents()), "tensor_upload");
}

// ------------------------------------------------------------------------------ Operator / MPO
Operator::Operator(const mwArray& data) : data_(data) {
  std::vector<int> d = data.getDimensions();
  if (d.size() > 4) throw Error(MPSD_ERR_SHAPE, "Operator: rank above 4");
  while (d.size() < 4) d.push_back(1);
  data_.reshape(d);
}

const mpsd_tensor* Operator::device(mpsd_ctx* ctx) const {
  for (auto& p : dev_)
    if (p.first == ctx) return p.second->t;
  dev_.emplace_back(ctx, std::make_shared<DevTensor>(ctx, data_));
  return dev_.back().second->t;
}

void MPO::setOp(int k, Operator* op, bool myop) {
  if (myop)
    ops_[k] = std::shared_ptr<const Operator>(op);
  else
    ops_[k] = std::shared_ptr<const Operator>(op, [](const Operator*) {});
}

void MPO::saveMPO(const char* filename) const {
  std::ofstream out(filename, std::ios::binary);
  if (!out.is_open()) throw Error(MPSD_ERR_ARG, std::string("saveMPO: cannot open ") + filename);
  const int length = getLength();
  out.write((const char*)&length, sizeof(int));
  for (int k = 0; k < length; ++k) getOp(k).getFullData().save(out);
  if (!out) throw Error(MPSD_ERR_ARG, "saveMPO: write failed");
}

void MPO::loadMPO(const char* filename) {
  std::ifstream in(filename, std::ios::binary);
  if (!in.is_open()) throw Error(MPSD_ERR_ARG, std::string("loadMPO: cannot open ") + filename);
  int length = 0;
  in.read((char*)&length, sizeof(int));
  if (!in || length < 0 || length > (1 << 20)) throw Error(MPSD_ERR_ARG, "loadMPO: bad length");
  initLength(length);
  for (int k = 0; k < length; ++k) {
    mwArray a;
    a.load(in);
    setOp(k, new Operator(a), true);
  }
}

void MPO::exportMPOtext(const char* filename) const {
  std::ofstream out(filename);
  if (!out.is_open()) throw Error(MPSD_ERR_ARG, std::string("exportMPOtext: cannot open ") + filename);
  const int length = getLength();
  out << length << " ";
  for (int k = 0; k < length; ++k) getOp(k).getFullData().savetext(out);
  if (!out) throw Error(MPSD_ERR_ARG, "exportMPOtext: write failed");
}

void MPO::importMPOtext(const char* filename) {
  std::ifstream in(filename);
  if (!in.is_open()) throw Error(MPSD_ERR_ARG, std::string("importMPOtext: cannot open ") + filename);
  int length = 0;
  in >> length;
  if (!in || length < 0 || length > (1 << 20)) throw Error(MPSD_ERR_ARG, "importMPOtext: bad length");
  initLength(length);
  for (int k = 0; k < length; ++k) {
    mwArray a;
    a.loadtext(in);
    setOp(k, new Operator(a), true);
  }
}

std::vector<int> MPO::getDimensions() const {
  std::vector<int> d;
  for (auto& o : ops_) d.push_back(o->getDimension(0));
  return d;
}

MPO mpo_product(const MPO& a, const MPO& b) {
  const int L = a.getLength();
  if (b.getLength() != L) throw Error(MPSD_ERR_SHAPE, "mpo_product: lengths differ");
  MPO out(L);
  for (int k = 0; k < L; ++k) {
    const mwArray& A = a.getOp(k).getFullData();
    const mwArray& B = b.getOp(k).getFullData();
    const int du = A.getDimension(0), al = A.getDimension(1), dm = A.getDimension(2), ar = A.getDimension(3);
    const int bl = B.getDimension(1), dd = B.getDimension(2), br = B.getDimension(3);
    if (B.getDimension(0) != dm) throw Error(MPSD_ERR_SHAPE, "mpo_product: inner physical dimensions differ");
    // P[u, a1 + al a2, v, b1 + ar b2] = sum_s A[u, a1, s, b1] B[s, a2, v, b2]
    const int pl = al * bl, pr = ar * br;
    mwArray P(std::vector<int>{du, pl, dd, pr});
    for (int b2 = 0; b2 < br; ++b2)
      for (int b1 = 0; b1 < ar; ++b1)
        for (int v = 0; v < dd; ++v)
          for (int a2 = 0; a2 < bl; ++a2)
            for (int a1 = 0; a1 < al; ++a1)
              for (int u = 0; u < du; ++u) {
                complex_t acc = 0.0;
                for (int s = 0; s < dm; ++s)
                  acc += A.data()[u + du * (a1 + al * (s + dm * b1))] * B.data()[s + dm * (a2 + bl * (v + dd * b2))];
                P.data()[u + du * ((a1 + al * a2) + pl * (v + dd * (b1 + ar * b2)))] = acc;
              }
    out.setOp(k, new Operator(P), true);
  }
  return out;
}

// ------------------------------------------------------------------------------ RNG
static inline uint64_t splitmix(uint64_t& state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

mwArray complex_gaussian(uint64_t seed, const std::vector<int>& dims) {
  mwArray a(dims);
  uint64_t st = seed;
  const double inv53 = 1.0 / 9007199254740992.0;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < a.getNrComponents(); ++i) {
    const double u1 = ((double)(splitmix(st) >> 11) + 1.0) * inv53;
    const double u2 = (double)(splitmix(st) >> 11) * inv53;
    const double r = std::sqrt(-2.0 * std::log(u1));
    a.data()[i] = complex_t(r * std::cos(two_pi * u2), r * std::sin(two_pi * u2));
  }
  return a;
}

// ------------------------------------------------------------------------------ Site
Site::Site(mpsd_ctx* ctx, int pos, const std::vector<int>& dims) : ctx_(ctx), pos_(pos), dev_(ctx, dims) {}

std::vector<int> Site::dims() const {
  std::vector<int> d = dev_.dims();
  while (d.size() < 3) d.push_back(1);
  return d;
}

void Site::setA(const mwArray& a) {
  mwArray b = a;
  std::vector<int> d = b.getDimensions();
  if (d.size() == 2 || (d.size() == 1 && d[0] == dev_.size())) b.reshape(dims());  // vector from the eigensolver
  dev_.upload(b);
}

void Site::gaugeR(DevTensor& rightterm, int cutD) {
  if (cutD > 0)
    check(ctx_, mpsd_site_gauge_r_svd(ctx_, dev_.t, rightterm.t, cutD), "gaugeR");
  else
    check(ctx_, mpsd_site_gauge_r(ctx_, dev_.t, rightterm.t), "gaugeR");
}

void Site::gaugeL(DevTensor& leftterm, int cutD) {
  if (cutD > 0)
    check(ctx_, mpsd_site_gauge_l_svd(ctx_, dev_.t, leftterm.t, cutD), "gaugeL");
  else
    check(ctx_, mpsd_site_gauge_l(ctx_, dev_.t, leftterm.t), "gaugeL");
}

void Site::reduceDimR(const DevTensor& leftterm) {
  check(ctx_, mpsd_site_reduce_dim_r(ctx_, dev_.t, leftterm.t), "reduceDimR");
}

void Site::reduceDimL(const DevTensor& rightterm) {
  check(ctx_, mpsd_site_reduce_dim_l(ctx_, dev_.t, rightterm.t), "reduceDimL");
}

void Site::gaugeRInto(DevTensor& rightterm, Site& next) {
  check(ctx_, mpsd_site_gauge_r_absorb(ctx_, dev_.t, rightterm.t, next.dev_.t), "gaugeR + reduceDimR");
}

void Site::gaugeLInto(DevTensor& leftterm, Site& prev) {
  check(ctx_, mpsd_site_gauge_l_absorb(ctx_, dev_.t, leftterm.t, prev.dev_.t), "gaugeL + reduceDimL");
}

void Site::changeDimensions(int d, int Dl, int Dr) {
  std::vector<int> cur = dev_.dims();
  if (cur.size() < 3) {
    // a lower-rank buffer (e.g. a vector set by the eigensolver) is a [d, Dl, Dr] site
    const std::vector<int> full = dims();
    check(ctx_, mpsd_tensor_reshape(dev_.t, 3, full.data()), "Site reshape");
  }
  const int nd[3] = {d, Dl, Dr};
  check(ctx_, mpsd_tensor_resize(dev_.t, 3, nd), "Site::changeDimensions");
}

// ------------------------------------------------------------------------------ MPS
MPS::MPS(int length, int bond, int d, mpsd_ctx* ctx) : MPS(length, bond, std::vector<int>(length, d), ctx) {}

MPS::MPS(int length, int bond, const std::vector<int>& d, mpsd_ctx* ctx) : ctx_(ctx ? ctx : default_ctx()), bond_(bond) {
  if ((int)d.size() < length) throw Error(MPSD_ERR_ARG, "MPS: fewer physical dimensions than sites");
  for (int k = 0; k < length; ++k) {
    if (d[k] < 1) throw Error(MPSD_ERR_ARG, "MPS: physical dimension below 1");
    const int Dl = (k == 0) ? 1 : bond;
    const int Dr = (k == length - 1) ? 1 : bond;
    A_.emplace_back(new Site(ctx_, k + 1, {d[k], Dl, Dr}));
  }
}

void MPS::setRandomState(uint64_t seed) {
  const int L = getLength();
  // bond k is capped by the state-space size on either side: prod d[0..k-1] and prod d[k..L-1]
  std::vector<int64_t> left(L + 1, 1), right(L + 1, 1);
  for (int k = 0; k < L; ++k) left[k + 1] = std::min<int64_t>(left[k] * A_[k]->getd(), bond_);
  for (int k = L - 1; k >= 0; --k) right[k] = std::min<int64_t>(right[k + 1] * A_[k]->getd(), bond_);
  std::vector<int> bd(L + 1);
  for (int k = 0; k <= L; ++k) bd[k] = (int)std::min<int64_t>(std::min<int64_t>(left[k], right[k]), bond_);
  for (int k = 0; k < L; ++k) {
    const int d = A_[k]->getd();
    // Each site is scaled by the power of two nearest 1 / sqrt(2 d Dr), so the unnormalised
    // state's norm stays near 1 along the chain: with unit-variance entries it would grow like
    // prod_k sqrt(2 d_k Dr_k), past the double range for long chains (e.g. 1e160 at L = 109,
    // D = 256), and the gauge factors absorbed below carry it.  A power of two scales exactly,
    // and the normalisation removes it, so the gauged state does not depend on it.
    mwArray a = complex_gaussian(seed * 1000003ull + (uint64_t)k, {d, bd[k], bd[k + 1]});
    const double sc = std::ldexp(1.0, -(int)std::lround(0.5 * std::log2(2.0 * d * bd[k + 1])));
    for (int64_t i = 0; i < a.getNrComponents(); ++i) a.data()[i] *= sc;
    A_[k]->setA(a);
  }
  gaugeCond('R', true);
}

DevTensor& MPS::factor() {
  if (!factor_) factor_.reset(new DevTensor(ctx_, std::vector<int>{1, 1}));
  return *factor_;
}

void MPS::gaugeCond(char dir, bool normal, int cutD) {
  DevTensor& tmp = factor();
  const int L = getLength();
  if (dir == 'L' || dir == 'l') {
    for (int pos = L - 1; pos >= 0; --pos) {
      if (pos > 0 && cutD <= 0) {
        A_[pos]->gaugeLInto(tmp, *A_[pos - 1]);
      } else {
        A_[pos]->gaugeL(tmp, cutD);
        if (pos > 0) A_[pos - 1]->reduceDimL(tmp);
      }
    }
    gauge_ = gLeft;
  } else if (dir == 'R' || dir == 'r') {
    for (int pos = 0; pos < L; ++pos) {
      if (pos < L - 1 && cutD <= 0) {
        A_[pos]->gaugeRInto(tmp, *A_[pos + 1]);
      } else {
        A_[pos]->gaugeR(tmp, cutD);
        if (pos < L - 1) A_[pos + 1]->reduceDimR(tmp);
      }
    }
    gauge_ = gRight;
  } else {
    throw Error(MPSD_ERR_ARG, "MPS::gaugeCond: unknown direction");
  }
  int maxb = 1;
  for (auto& s : A_) maxb = std::max(maxb, std::max(s->getDl(), s->getDr()));
  bond_ = std::min(bond_, maxb);
  if (normal) {
    normfact_ = 1.0;
  } else if (L > 0) {
    // the last factor is the 1x1 norm of the state (its phase stayed in the edge site)
    mwArray f = tmp.download();
    if (f.getNrComponents() != 1) throw Error(MPSD_ERR_SHAPE, "MPS::gaugeCond: last factor is not a scalar");
    normfact_ *= std::abs(f.getElement(0));
  }
}

MPS::MPS(const MPS& other, int D) : ctx_(other.ctx_), bond_(other.bond_), normfact_(other.normfact_), gauge_(other.gauge_) {
  for (int k = 0; k < other.getLength(); ++k) {
    const Site& s = other.getA(k);
    A_.emplace_back(new Site(ctx_, s.getPos(), s.dims()));
    check(ctx_,
          mpsd_vec_copy(ctx_, (const double*)mpsd_tensor_devptr(s.device()),
                        (double*)mpsd_tensor_devptr(A_.back()->device()), mpsd_tensor_size(s.device())),
          "MPS copy");
  }
  gaugeCond('L', false, D);
}

void MPS::gaugeCond(int pos, char dir, bool next) {
  DevTensor& tmp = factor();
  const int L = getLength();
  bool edge = false;  // the factor left over at the chain's end is the state's norm
  if (dir == 'L' || dir == 'l') {
    if (next && pos > 0) {
      A_[pos]->gaugeLInto(tmp, *A_[pos - 1]);
    } else {
      const int oldDl = A_[pos]->getDl();
      A_[pos]->gaugeL(tmp);
      if (A_[pos]->getDl() != oldDl && pos > 0) A_[pos - 1]->reduceDimL(tmp);
      edge = next && pos == 0;
    }
  } else {
    if (next && pos < L - 1) {
      A_[pos]->gaugeRInto(tmp, *A_[pos + 1]);
    } else {
      const int oldDr = A_[pos]->getDr();
      A_[pos]->gaugeR(tmp);
      if (pos < L - 1 && A_[pos]->getDr() != oldDr) A_[pos + 1]->reduceDimR(tmp);
      edge = next && pos == L - 1;
    }
  }
  if (edge) {
    // MPS.cpp:380-387, 402-409: the 1x1 factor's magnitude goes into the norm factor (its
    // phase stayed in the edge site)
    mwArray f = tmp.download();
    if (f.getNrComponents() != 1) throw Error(MPSD_ERR_SHAPE, "MPS::gaugeCond: edge factor is not a scalar");
    normfact_ *= std::abs(f.getElement(0));
  }
  gauge_ = gNone;
}

void MPS::increaseBondDimension(int D) {
  // MPS.cpp:650-719 with zero noise.  Bond k (0 < k < L) becomes min(D, prod_{j<k} d_j,
  // prod_{j>=k} d_j): the reference's running products cumd (left) and cumdr (right, D until
  // the first site from the right where it drops below D) give the same minimum.  The outer
  // bonds keep their dimensions.
  if (D <= bond_) return;
  const int L = getLength();
  if (L == 0) return;
  std::vector<int64_t> left(L + 1, 1), right(L + 1, 1);
  for (int k = 0; k < L; ++k) left[k + 1] = std::min<int64_t>(left[k] * A_[k]->getd(), D);
  for (int k = L - 1; k >= 0; --k) right[k] = std::min<int64_t>(right[k + 1] * A_[k]->getd(), D);
  std::vector<int> bd(L + 1);
  for (int k = 1; k < L; ++k) bd[k] = (int)std::min<int64_t>(std::min<int64_t>(left[k], right[k]), D);
  bd[0] = A_[0]->getDl();
  bd[L] = A_[L - 1]->getDr();
  int maxbond = 1;
  for (int k = 0; k < L; ++k) {
    A_[k]->increaseBondDim(bd[k], bd[k + 1]);
    maxbond = std::max(maxbond, std::max(bd[k], bd[k + 1]));
  }
  bond_ = maxbond;
  gauge_ = gNone;
}

void MPS::adjustPhysDimensions(const std::vector<int>& d) {
  const int L = getLength();
  if ((int)d.size() != L)
    throw Error(MPSD_ERR_SHAPE, "MPS::adjustPhysDimensions: " + std::to_string(d.size()) + " dimensions for " +
                                    std::to_string(L) + " sites");
  bool changed = false;
  for (int k = 0; k < L; ++k) {
    if (d[k] < 1) throw Error(MPSD_ERR_ARG, "MPS::adjustPhysDimensions: physical dimension below 1");
    if (d[k] > A_[k]->getd()) {
      A_[k]->increasePhysDim(d[k]);
      changed = true;
    } else if (d[k] < A_[k]->getd()) {
      A_[k]->decreasePhysDim(d[k]);
      changed = true;
    }
  }
  if (changed) {
    // MPS.cpp:749-754: the bond profile is rebuilt for the new dimensions
    const int D = bond_;
    bond_ = 1;
    gauge_ = gNone;
    increaseBondDimension(D);
  }
}

void MPS::exportMPS(const char* filename) const {
  std::ofstream out(filename, std::ios::binary);
  if (!out.is_open()) throw Error(MPSD_ERR_ARG, std::string("exportMPS: cannot open ") + filename);
  int length = getLength();
  int g = (int)gauge_;
  out.write((const char*)&length, sizeof(int));
  out.write((const char*)&bond_, sizeof(int));
  out.write((const char*)&normfact_, sizeof(double));
  out.write((const char*)&g, sizeof(int));
  for (auto& s : A_) {
    int pos = s->getPos();
    out.write((const char*)&pos, sizeof(int));
    s->getA().save(out);
  }
  if (!out) throw Error(MPSD_ERR_ARG, "exportMPS: write failed");
}

void MPS::importMPS(const char* filename) {
  std::ifstream in(filename, std::ios::binary);
  if (!in.is_open()) throw Error(MPSD_ERR_ARG, std::string("importMPS: cannot open ") + filename);
  if (!ctx_) ctx_ = default_ctx();
  int length = 0, g = 0;
  in.read((char*)&length, sizeof(int));
  in.read((char*)&bond_, sizeof(int));
  in.read((char*)&normfact_, sizeof(double));
  in.read((char*)&g, sizeof(int));
  if (!in || length < 0) throw Error(MPSD_ERR_ARG, "importMPS: bad header");
  gauge_ = (Gauge)g;
  A_.clear();
  for (int k = 0; k < length; ++k) {
    int pos = 0;
    in.read((char*)&pos, sizeof(int));
    mwArray a;
    a.load(in);
    std::vector<int> d = a.getDimensions();
    while (d.size() < 3) d.push_back(1);
    a.reshape(d);
    A_.emplace_back(new Site(ctx_, pos, d));
    A_.back()->setA(a);
  }
}

// ------------------------------------------------------------------------------ TmpManager
TmpManager::TmpManager(int length, mpsd_ctx* ctx) : upL_(length, false), upR_(length, false) {
  const complex_t one(1.0, 0.0);
  // every slot but the two edges is written by an environment update before it is read
  for (int k = 0; k < length; ++k) {
    L_.emplace_back(ctx, std::vector<int>{1, 1, 1}, false);
    R_.emplace_back(ctx, std::vector<int>{1, 1, 1}, false);
  }
  mwArray edge(std::vector<int>{1, 1, 1}, &one);
  L_[0].upload(edge);
  R_[length - 1].upload(edge);
  upL_[0] = true;
  upR_[length - 1] = true;
}

int64_t TmpManager::deviceBytes() const {
  int64_t b = 0;
  for (auto& t : L_) b += 16 * t.size();
  for (auto& t : R_) b += 16 * t.size();
  return b;
}

}  // namespace mpsdyn
