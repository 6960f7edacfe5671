// Serving loop: continuous admission of independent ground-state solves by HBM footprint,
// concurrent solves on their own streams, and a matvec batcher that merges the matvecs of
// concurrent solves into grouped persistent launches on a shared stream.
//
// Ordering between streams is carried by events only: a solve records "x ready" on its
// stream before handing a matvec to the batcher; the batcher makes its stream wait on those
// events, launches the grouped matvec, records "y ready" and the solve's stream waits on it.
// No host thread blocks on the GPU for a matvec.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "mpsd_internal.h"
#include "mpsd_serve.h"
#include "mpsdyn/mpsdyn.h"

namespace mpsdyn {
int gs_solve_with(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* export_path, const MultiplierFactory* factory,
                  mpsd_gs_result* out);
}

using namespace mpsdyn;

namespace {

typedef std::chrono::steady_clock Clock;

// ------------------------------------------------------------------------------ batcher
class Batcher {
 public:
  Batcher(mpsd_ctx* ctx, int window_us) : ctx_(ctx), window_us_(window_us) {
    th_ = std::thread([this] { run(); });
  }
  ~Batcher() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_in_.notify_all();
    th_.join();
    for (auto& b : pool_) (void)hipEventDestroy(b);
  }
  void solve_started() { ++active_; }
  void solve_finished() {
    --active_;
    cv_in_.notify_all();
  }
  // y = Heff x for a solve running on `wctx`; returns once the matvec is queued on the
  // device, with wctx's stream ordered after it.
  int apply(mpsd_ctx* wctx, mpsd_heff* h, const double* x, double* y, hipEvent_t ready) {
    hipStream_t ws = (hipStream_t)mpsd_ctx_stream(wctx);
    if (hipEventRecord(ready, ws) != hipSuccess) return MPSD_ERR_HIP;
    Item it{h, x, y, ready};
    {
      std::unique_lock<std::mutex> lk(mu_);
      pending_.push_back(&it);
      cv_in_.notify_all();
      cv_out_.wait(lk, [&] { return it.done; });
    }
    if (it.status) return it.status;
    if (hipStreamWaitEvent(ws, it.finished, 0) != hipSuccess) return MPSD_ERR_HIP;
    release(it.finished);
    return MPSD_OK;
  }
  mpsd_serve_stats stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return st_;
  }
  const std::string& error() const { return err_; }

 private:
  struct Item {
    mpsd_heff* h;
    const double* x;
    double* y;
    hipEvent_t ready;
    bool done = false;
    int status = 0;
    hipEvent_t finished = nullptr;
  };
  hipEvent_t acquire(int refs) {
    std::lock_guard<std::mutex> lk(ev_mu_);
    hipEvent_t e = nullptr;
    if (!free_.empty()) {
      e = free_.back();
      free_.pop_back();
    } else {
      (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
      pool_.push_back(e);
    }
    refs_[e] = refs;
    return e;
  }
  void release(hipEvent_t e) {
    std::lock_guard<std::mutex> lk(ev_mu_);
    if (--refs_[e] == 0) free_.push_back(e);
  }
  void run() {
    hipStream_t bs = (hipStream_t)mpsd_ctx_stream(ctx_);
    std::vector<Item*> batch;
    std::vector<mpsd_heff*> hs;
    std::vector<const double*> xs;
    std::vector<double*> ys;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_in_.wait(lk, [&] { return stop_ || !pending_.empty(); });
        if (stop_ && pending_.empty()) return;
        // gather: until every running solve has a matvec waiting, or the window ends
        const auto deadline = Clock::now() + std::chrono::microseconds(window_us_);
        while (!stop_ && (int)pending_.size() < active_.load() && Clock::now() < deadline)
          cv_in_.wait_until(lk, deadline);
        batch.swap(pending_);
        pending_.clear();
      }
      hs.clear();
      xs.clear();
      ys.clear();
      int status = MPSD_OK;
      for (Item* it : batch) {
        if (hipStreamWaitEvent(bs, it->ready, 0) != hipSuccess) status = MPSD_ERR_HIP;
        hs.push_back(it->h);
        xs.push_back(it->x);
        ys.push_back(it->y);
      }
      if (!status) status = mpsd_heff_apply_grouped(ctx_, (int)batch.size(), hs.data(), xs.data(), ys.data());
      if (status) err_ = mpsd_last_error(ctx_);
      hipEvent_t fin = acquire((int)batch.size());
      if (hipEventRecord(fin, bs) != hipSuccess && !status) status = MPSD_ERR_HIP;
      {
        std::lock_guard<std::mutex> lk(mu_);
        ++st_.batches;
        st_.batched_matvecs += (int64_t)batch.size();
        st_.max_batch = std::max<int64_t>(st_.max_batch, (int64_t)batch.size());
        for (Item* it : batch) {
          it->status = status;
          it->finished = fin;
          it->done = true;
        }
      }
      cv_out_.notify_all();
      batch.clear();
    }
  }
  mpsd_ctx* ctx_;
  int window_us_;
  std::thread th_;
  std::mutex mu_, ev_mu_;
  std::condition_variable cv_in_, cv_out_;
  std::vector<Item*> pending_;
  std::atomic<int> active_{0};
  bool stop_ = false;
  mpsd_serve_stats st_{};
  std::vector<hipEvent_t> pool_, free_;
  std::map<hipEvent_t, int> refs_;
  std::string err_;
};

// Multiplier whose matvecs go through the batcher.
class BatchedMultiplier : public Multiplier {
 public:
  BatchedMultiplier(Batcher* b, mpsd_ctx* ctx, const mpsd_tensor* L, const mpsd_tensor* W, const mpsd_tensor* R,
                    double offset, int precision, hipEvent_t ready)
      : b_(b), ctx_(ctx), ready_(ready) {
    check(ctx_, mpsd_heff_create(ctx_, L, W, R, offset, precision, &h_), "heff_create");
  }
  ~BatchedMultiplier() override {
    if (h_) mpsd_heff_destroy(h_);
  }
  int64_t getSize() const override { return mpsd_heff_size(h_); }
  void product_device(const double* x, double* y) override {
    int rc = b_->apply(ctx_, h_, x, y, ready_);
    if (rc) throw Error(rc, "batched matvec: " + b_->error());
  }
  void getFullTensor(mwArray& r) override {
    const int n = (int)getSize();
    r = mwArray(std::vector<int>{n, n});
    check(ctx_, mpsd_heff_full(h_, (double*)r.data()), "heff_full");
  }
  mpsd_ctx* ctx() const override { return ctx_; }
  void setPrecision(int p) override { check(ctx_, mpsd_heff_set_precision(h_, p), "set_precision"); }
  // small local problems are solved by the one-workgroup eigensolver, as they are alone
  mpsd_heff* smallHeff() const override { return h_; }

 private:
  Batcher* b_;
  mpsd_ctx* ctx_;
  hipEvent_t ready_;
  mpsd_heff* h_ = nullptr;
};

int64_t footprint(const mpsd_gs_config* c) {
  // complex128 elements: MPS + both environment sides + Lanczos basis (31 vectors) +
  // contraction scratch (T1, T2, partials) + gauge buffers; 10% slack for the allocator.
  const int L = c->L, D = c->D, d = 2;
  const int w = c->model == MPSD_MODEL_ISING ? 3 : 5;
  std::vector<double> bd(L + 1);
  for (int k = 0; k <= L; ++k) bd[k] = std::min<double>({(double)D, std::pow(2.0, std::min(k, 60)), std::pow(2.0, std::min(L - k, 60))});
  double mps = 0, env = 0, dmax = 0;
  for (int k = 0; k < L; ++k) {
    mps += d * bd[k] * bd[k + 1];
    env += 2.0 * w * bd[k] * bd[k];
    dmax = std::max(dmax, bd[k]);
  }
  const double n = d * dmax * dmax;
  const double elems = mps + env + 31.0 * n + 3.0 * w * n + 2.0 * n;
  return (int64_t)(1.1 * 16.0 * elems) + (64ll << 20);
}

// ------------------------------------------------------------------------------ server
struct Request {
  int id;
  mpsd_gs_config cfg;
  mpsd_serve_result res{};
  bool done = false;
};

}  // namespace

struct mpsd_server {
  mpsd_serve_config cfg{};
  Clock::time_point t0;
  mpsd_ctx* bctx = nullptr;
  std::unique_ptr<Batcher> batcher;
  std::vector<mpsd_ctx*> wctx;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::map<int, std::unique_ptr<Request>> reqs;
  int next_id = 0;
  int64_t admitted = 0;
  int running = 0;
  bool stop = false;
  mpsd_serve_stats st{};
  std::string err;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  void worker(int slot) {
    mpsd_ctx* ctx = wctx[slot];
    hipEvent_t ready = nullptr;
    (void)hipEventCreateWithFlags(&ready, hipEventDisableTiming);
    MultiplierFactory fac = [this, ready](mpsd_ctx* c, const mpsd_tensor* L, const mpsd_tensor* W,
                                          const mpsd_tensor* R, double off, int prec) {
      return std::unique_ptr<Multiplier>(new BatchedMultiplier(batcher.get(), c, L, W, R, off, prec, ready));
    };
    for (;;) {
      Request* r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        // FIFO admission: the head of the queue runs as soon as its footprint fits
        cv_work.wait(lk, [&] {
          return stop || (!queue.empty() && admitted + queue.front()->res.footprint_bytes <= cfg.hbm_budget);
        });
        if (stop && queue.empty()) break;
        if (stop) break;
        r = queue.front();
        queue.pop_front();
        admitted += r->res.footprint_bytes;
        ++running;
        st.peak_admitted_bytes = std::max(st.peak_admitted_bytes, admitted);
        st.peak_running = std::max(st.peak_running, running);
        r->res.admit_s = now();
      }
      cv_work.notify_all();
      if (batcher) batcher->solve_started();
      mpsd_gs_result g;
      const int rc = gs_solve_with(ctx, &r->cfg, nullptr, batcher ? &fac : nullptr, &g);
      if (batcher) batcher->solve_finished();
      {
        std::lock_guard<std::mutex> lk(mu);
        r->res.status = rc;
        if (rc) err = mpsd_last_error(ctx);
        r->res.energy = g.energy;
        r->res.halfsweeps = g.halfsweeps;
        r->res.matvecs = g.matvecs;
        r->res.matvecs_lowp = g.matvecs_lowp;
        r->res.matvec_flops = g.matvec_flops;
        r->res.finish_s = now();
        r->done = true;
        admitted -= r->res.footprint_bytes;
        --running;
        ++st.solves;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
    (void)hipEventDestroy(ready);
  }
};

extern "C" {

int64_t mpsd_gs_footprint(const mpsd_gs_config* cfg) { return cfg ? footprint(cfg) : MPSD_ERR_ARG; }

int mpsd_server_create(const mpsd_serve_config* cfg, mpsd_server** out) {
  if (!cfg || !out || cfg->slots < 1 || cfg->slots > 64) return MPSD_ERR_ARG;
  *out = nullptr;
  std::unique_ptr<mpsd_server> s(new mpsd_server());
  s->cfg = *cfg;
  s->t0 = Clock::now();
  if (s->cfg.hbm_budget <= 0) {
    size_t fr = 0, tot = 0;
    if (hipSetDevice(cfg->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) return MPSD_ERR_HIP;
    s->cfg.hbm_budget = (int64_t)(0.8 * (double)fr);
  }
  if (s->cfg.window_us <= 0) s->cfg.window_us = 200;
  for (int i = 0; i < cfg->slots; ++i) {
    mpsd_ctx* c = nullptr;
    int rc = mpsd_ctx_create(cfg->device, &c);
    if (rc) {
      for (auto* w : s->wctx) mpsd_ctx_destroy(w);
      return rc;
    }
    s->wctx.push_back(c);
    // eigensolver steps replayed from captured graphs (used where a solve's matvec is its own
    // launch sequence, i.e. without the batcher)
    if (cfg->graph && (rc = mpsd_ctx_set_option(c, MPSD_OPT_GRAPH, 1)) != 0) {
      for (auto* w : s->wctx) mpsd_ctx_destroy(w);
      return rc;
    }
  }
  if (cfg->batching) {
    int rc = mpsd_ctx_create(cfg->device, &s->bctx);
    if (rc) {
      for (auto* w : s->wctx) mpsd_ctx_destroy(w);
      return rc;
    }
    s->batcher.reset(new Batcher(s->bctx, s->cfg.window_us));
  }
  mpsd_server* raw = s.get();
  for (int i = 0; i < cfg->slots; ++i) raw->workers.emplace_back([raw, i] { raw->worker(i); });
  *out = s.release();
  return MPSD_OK;
}

int mpsd_server_destroy(mpsd_server* s) {
  if (!s) return MPSD_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  s->batcher.reset();
  for (auto* c : s->wctx) mpsd_ctx_destroy(c);
  if (s->bctx) mpsd_ctx_destroy(s->bctx);
  delete s;
  return MPSD_OK;
}

const char* mpsd_server_last_error(const mpsd_server* s) { return s ? s->err.c_str() : "null server"; }

int mpsd_server_submit(mpsd_server* s, const mpsd_gs_config* cfg, int* id) {
  if (!s || !cfg || !id) return MPSD_ERR_ARG;
  std::unique_ptr<Request> r(new Request());
  r->cfg = *cfg;
  r->res.footprint_bytes = footprint(cfg);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->res.footprint_bytes > s->cfg.hbm_budget) {
      s->err = "request does not fit in the HBM budget on its own";
      return MPSD_ERR_NOMEM;
    }
    r->id = s->next_id++;
    r->res.id = r->id;
    r->res.arrival_s = s->now();
    *id = r->id;
    s->queue.push_back(r.get());
    s->reqs[r->id] = std::move(r);
  }
  s->cv_work.notify_all();
  return MPSD_OK;
}

int mpsd_server_wait(mpsd_server* s, int id, mpsd_serve_result* out) {
  if (!s || !out) return MPSD_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return MPSD_ERR_ARG;
  Request* r = it->second.get();
  s->cv_done.wait(lk, [&] { return r->done; });
  *out = r->res;
  return MPSD_OK;
}

int mpsd_server_run_trace(mpsd_server* s, int n, const mpsd_gs_config* cfgs, const double* arrival_s,
                          mpsd_serve_result* out) {
  if (!s || n < 0 || (n > 0 && (!cfgs || !arrival_s || !out))) return MPSD_ERR_ARG;
  std::vector<int> ids(n);
  const auto start = Clock::now();
  for (int i = 0; i < n; ++i) {
    std::this_thread::sleep_until(start + std::chrono::duration_cast<Clock::duration>(
                                              std::chrono::duration<double>(arrival_s[i])));
    int rc = mpsd_server_submit(s, &cfgs[i], &ids[i]);
    if (rc) return rc;
  }
  int worst = MPSD_OK;
  for (int i = 0; i < n; ++i) {
    int rc = mpsd_server_wait(s, ids[i], &out[i]);
    if (rc) return rc;
    if (out[i].status && !worst) worst = out[i].status;
  }
  return worst;
}

int mpsd_server_stats(const mpsd_server* s, mpsd_serve_stats* out) {
  if (!s || !out) return MPSD_ERR_ARG;
  mpsd_server* m = const_cast<mpsd_server*>(s);
  std::lock_guard<std::mutex> lk(m->mu);
  *out = m->st;
  if (m->batcher) {
    mpsd_serve_stats b = m->batcher->stats();
    out->batches = b.batches;
    out->batched_matvecs = b.batched_matvecs;
    out->max_batch = b.max_batch;
  }
  out->graph_replays = 0;
  for (mpsd_ctx* c : m->wctx) out->graph_replays += c->graph_launches.load();
  return MPSD_OK;
}

}  // extern "C"
