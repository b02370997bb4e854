// rnpx_serve.cpp -- a least-squares request server on one GPU (no
// counterpart in the reference, which has no driver above QR::Factor).
//
// Requests (A: m x n, B: m x nrhs, host memory) are copied into the server on
// submit.  Worker threads, each with its own rnpx handle (HIP stream), pull
// work from the queue as they free up -- continuous batching:
//  * small requests (m <= 256) are drained together into one batched launch
//    (the persistent batched kernel, one problem per team);
//  * larger requests run one per worker; several run concurrently on the
//    workers' streams.  All panel kernels (they meet at grid barriers and
//    need their workgroups co-resident) go to one shared stream, so at most
//    one is in flight; the trailing updates of different requests overlap.
// Admission: a request is admitted only when its device footprint fits in
// what the HBM budget has left; a request larger than the whole budget is
// rejected.  The queue is served in arrival order, except that a request
// that does not fit may be overtaken by later ones that do, for at most
// kBypassLimit admissions.
// Precision: mode 0 solves in FP64 (QR + Q^H b + back substitution); mode 1
// uses the mixed-precision solver and falls back to FP64 for a request whose
// refinement does not reach the tolerance (ill-conditioned A).
// Each request's result depends only on its own data: the same bits alone or
// in any mix.
#include "rnpx_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdint>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace rnpx {

// dst (m x n, leading dimension m) = src (leading dimension lds): large
// matrices split over host threads by columns (one thread copies a few GB/s;
// a 8192 x 4096 request is 268 MB).
static void copy_columns(double* dst, const double* src, size_t m, size_t n, size_t lds) {
  const size_t bytes = sizeof(double) * m * n;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>({(size_t)8, (size_t)hw, n, std::max<size_t>(1, bytes >> 24)});  // 16 MB each
  auto part = [&](size_t j0, size_t j1) {
    for (size_t j = j0; j < j1; ++j) std::memcpy(dst + j * m, src + j * lds, sizeof(double) * m);
  };
  if (nt <= 1) {
    part(0, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (n + nt - 1) / nt;
  for (size_t q = 1; q < nt; ++q) th.emplace_back(part, std::min(n, q * per), std::min(n, (q + 1) * per));
  part(0, std::min(n, per));
  for (auto& x : th) x.join();
}

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  long long id = 0;
  int m = 0, n = 0, nrhs = 0;
  std::vector<double> A, B, X;  // column-major, leading dimension m (X: n)
  size_t bytes = 0;             // device footprint reserved while in flight
  rnpx_request_info info{};
  bool done = false;
};

struct Server {
  rnpx_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::deque<Request*> queue;
  std::map<long long, Request*> all;
  std::vector<std::thread> workers;
  std::vector<Handle*> handles;
  hipStream_t panel_stream = nullptr;
  size_t in_use = 0;
  long long next_id = 1;
  int bypassed = 0;
  bool stop = false;
  rnpx_server_stats stats{};
};

constexpr int kSmallM = 256;
constexpr int kBypassLimit = 8;
constexpr int kMaxBatch = 4096;

static size_t footprint(int m, int n, int nrhs, int mode) {
  const size_t e = sizeof(double);
  size_t b = e * ((size_t)m * n + (size_t)m * nrhs * 2 + (size_t)n * nrhs + n);
  if (m > kSmallM) {
    // factor copy (mixed), V/Gram/T double buffers, W/Z, split-K slices
    b += e * ((size_t)m * n * (mode == 1 ? 1 : 0) + (size_t)m * 64 * 2 + (size_t)64 * n * 2 + (size_t)32 * 64 * n);
  }
  return b;
}

// Dense request on one worker's handle.
static int solve_dense(Server* s, Handle* h, Request* r, Buffer& dA, Buffer& dA2, Buffer& dB, Buffer& dX) {
  const int m = r->m, n = r->n, q = r->nrhs;
  const size_t e = sizeof(double);
  hipStream_t st = h->stream;
  if (dA.ensure(e * (size_t)m * n, h) || dB.ensure(e * (size_t)m * q, h) || dX.ensure(e * ((size_t)n * q + n), h) ||
      dA2.ensure(e * (size_t)m * n, h))
    return RNPX_ERR_NOMEM;
  double* A = (double*)dA.p;
  double* B = (double*)dB.p;
  double* X = (double*)dX.p;
  double* tau = X + (size_t)n * q;
  if (hipMemcpyAsync(A, r->A.data(), e * (size_t)m * n, hipMemcpyHostToDevice, st) != hipSuccess) return RNPX_ERR_HIP;
  if (hipMemcpyAsync(B, r->B.data(), e * (size_t)m * q, hipMemcpyHostToDevice, st) != hipSuccess) return RNPX_ERR_HIP;
  int status = RNPX_OK;
  r->info.path = 0;
  if (s->cfg.mode == 1) {
    int iters = 0;
    double be = 0.0;
    status = gels_mixed_dev(h, m, n, q, A, m, B, m, X, n, s->cfg.max_iter, s->cfg.tol, &iters, &be);
    r->info.iters = iters;
    r->info.berr = be;
    r->info.path = 1;
    if (status == RNPX_ERR_NOTCONV) {
      r->info.path = 2;  // FP64 fallback
      status = RNPX_OK;
    }
  }
  if (r->info.path != 1) {
    double* A2 = (double*)dA2.p;
    hipMemcpyAsync(A2, A, e * (size_t)m * n, hipMemcpyDeviceToDevice, st);
    status = gels_dev<double>(h, m, n, q, A2, m, tau, B, m);
    if (status) return status;
    if (launch_copy<double>(n, q, B, m, X, n, st) != hipSuccess) return RNPX_ERR_HIP;
    // B was overwritten: restore it for the backward-error check
    hipMemcpyAsync(B, r->B.data(), e * (size_t)m * q, hipMemcpyHostToDevice, st);
    double be = 0.0;
    status = ls_berr_dev(h, m, n, q, A, m, B, m, X, n, &be);
    r->info.berr = be;
    if (status) return status;
  }
  if (status) return status;
  r->X.resize((size_t)n * q);
  if (hipMemcpyAsync(r->X.data(), X, e * (size_t)n * q, hipMemcpyDeviceToHost, st) != hipSuccess) return RNPX_ERR_HIP;
  return sync_handle(h);
}

// Many small requests in one batched launch.
static int solve_small_batch(Handle* h, std::vector<Request*>& rs, Buffer& dA, Buffer& dB, Buffer& dT) {
  const size_t e = sizeof(double);
  std::vector<rnpx_qr_desc> desc(rs.size());
  long long na = 0, nb = 0, nt = 0;
  for (size_t i = 0; i < rs.size(); ++i) {
    Request* r = rs[i];
    desc[i] = rnpx_qr_desc{r->m, r->n, r->m, r->m, r->nrhs, 0, na, nt, nb};
    na += (long long)r->m * r->n;
    nb += (long long)r->m * r->nrhs;
    nt += r->n;
  }
  if (dA.ensure(e * (size_t)std::max(na, 1LL), h) || dB.ensure(e * (size_t)std::max(nb, 1LL), h) ||
      dT.ensure(e * (size_t)std::max(nt, 1LL), h))
    return RNPX_ERR_NOMEM;
  std::vector<double> ha((size_t)na), hb((size_t)nb);
  for (size_t i = 0; i < rs.size(); ++i) {
    std::memcpy(ha.data() + desc[i].a_off, rs[i]->A.data(), e * rs[i]->A.size());
    std::memcpy(hb.data() + desc[i].b_off, rs[i]->B.data(), e * rs[i]->B.size());
  }
  hipStream_t st = h->stream;
  if (hipMemcpyAsync(dA.p, ha.data(), e * (size_t)na, hipMemcpyHostToDevice, st) != hipSuccess) return RNPX_ERR_HIP;
  if (hipMemcpyAsync(dB.p, hb.data(), e * (size_t)nb, hipMemcpyHostToDevice, st) != hipSuccess) return RNPX_ERR_HIP;
  int s = geqrf_batched<double>(h, (int)rs.size(), desc.data(), (double*)dA.p, (double*)dT.p, (double*)dB.p);
  if (s) return s;
  if (hipMemcpyAsync(hb.data(), dB.p, e * (size_t)nb, hipMemcpyDeviceToHost, st) != hipSuccess) return RNPX_ERR_HIP;
  s = sync_handle(h);
  if (s) return s;
  for (size_t i = 0; i < rs.size(); ++i) {
    Request* r = rs[i];
    r->X.resize((size_t)r->n * r->nrhs);
    for (int q = 0; q < r->nrhs; ++q)
      std::memcpy(r->X.data() + (size_t)q * r->n, hb.data() + desc[i].b_off + (size_t)q * r->m, e * r->n);
    r->info.path = 3;
  }
  return RNPX_OK;
}

static void finish(Server* s, std::vector<Request*>& rs, int status) {
  const double t = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  for (Request* r : rs) {
    r->info.status = status;
    r->info.t_done = t;
    r->done = true;
    s->in_use -= r->bytes;
    r->A.clear();
    r->A.shrink_to_fit();
    r->B.clear();
    r->B.shrink_to_fit();
    s->stats.completed += 1;
    if (status) s->stats.failed += 1;
  }
  s->cv_done.notify_all();
  s->cv.notify_all();
}

// Pick work for a free worker (called with the lock held).
static void take_work(Server* s, std::vector<Request*>& out) {
  out.clear();
  const size_t budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : (size_t)-1;
  // small requests: everything queued that fits, up to kMaxBatch
  if (!s->queue.empty() && s->queue.front()->m <= kSmallM) {
    for (auto it = s->queue.begin(); it != s->queue.end() && (int)out.size() < kMaxBatch;) {
      Request* r = *it;
      if (r->m <= kSmallM && s->in_use + r->bytes <= budget) {
        s->in_use += r->bytes;
        out.push_back(r);
        it = s->queue.erase(it);
      } else {
        ++it;
      }
    }
    if (!out.empty()) return;
  }
  for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
    Request* r = *it;
    const bool head = (it == s->queue.begin());
    if (!head && s->bypassed >= kBypassLimit) break;
    if (s->in_use + r->bytes <= budget) {
      s->in_use += r->bytes;
      s->bypassed = head ? 0 : s->bypassed + 1;
      out.push_back(r);
      s->queue.erase(it);
      return;
    }
  }
}

static void worker_main(Server* s, int w) {
  Handle* h = s->handles[w];
  hipSetDevice(h->device);
  Buffer dA, dA2, dB, dX, dT;
  std::vector<Request*> work;
  while (true) {
    work.clear();
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] {
        if (s->stop) return true;
        take_work(s, work);
        return !work.empty();
      });
      if (work.empty() && s->stop) break;
      const double t = now_s();
      for (Request* r : work) {
        r->info.t_admit = t;
        r->info.worker = w;
      }
      s->stats.admitted += (long long)work.size();
    }
    int status = RNPX_OK;
    if (work.size() > 1 || work[0]->m <= kSmallM) {
      status = solve_small_batch(h, work, dA, dB, dT);
      std::lock_guard<std::mutex> lk(s->mu);  // stats are read and written under the server lock
      s->stats.batches += 1;
    } else {
      status = solve_dense(s, h, work[0], dA, dA2, dB, dX);
    }
    finish(s, work, status);
  }
  hipStreamSynchronize(h->stream);
  for (Buffer* b : {&dA, &dA2, &dB, &dX, &dT}) b->release();
}

}  // namespace rnpx

using namespace rnpx;

struct rnpx_server_s {};

extern "C" {

int rnpx_server_create(const rnpx_server_config* cfg, rnpx_server* out) {
  if (!cfg) return -1;
  if (!out) return -2;
  if (cfg->workers < 1 || cfg->workers > 16) return -1;
  if (cfg->mode != 0 && cfg->mode != 1) return -1;
  int st = Handle::check_gpu();
  if (st) return st;
  Server* s = new Server();
  s->cfg = *cfg;
  if (s->cfg.max_iter <= 0) s->cfg.max_iter = 30;
  if (s->cfg.tol <= 0) s->cfg.tol = 1e-13;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (create_own_queue_stream(&s->panel_stream, cus > 0 ? cus : 256) != hipSuccess) {
    delete s;
    return RNPX_ERR_HIP;
  }
  for (int w = 0; w < cfg->workers; ++w) {
    Handle* h = nullptr;
    st = create_handle(&h, 0);
    if (st) break;
    h->ext_pstream = s->panel_stream;
    s->handles.push_back(h);
  }
  if (st) {
    for (Handle* h : s->handles) delete h;
    hipStreamDestroy(s->panel_stream);
    delete s;
    return st;
  }
  for (int w = 0; w < cfg->workers; ++w) s->workers.emplace_back(worker_main, s, w);
  *out = reinterpret_cast<rnpx_server>(s);
  return RNPX_OK;
}

int rnpx_server_submit(rnpx_server srv, size_t m, size_t n, size_t nrhs, const double* a, size_t lda,
                       const double* b, size_t ldb, long long* id) {
  Server* s = reinterpret_cast<Server*>(srv);
  if (!s) return -1;
  if (m < 1 || m > (size_t)INT32_MAX) return -2;
  if (n < 1 || n > m) return -3;
  if (nrhs < 1 || nrhs > 4096) return -4;
  if (!a) return -5;
  if (lda < m) return -6;
  if (!b) return -7;
  if (ldb < m) return -8;
  if (!id) return -9;
  const double t_in = now_s();  // the request's own copy of A and b below counts as its wait
  Request* r = new Request();
  r->m = (int)m;
  r->n = (int)n;
  r->nrhs = (int)nrhs;
  r->A.resize(m * n);
  r->B.resize(m * nrhs);
  copy_columns(r->A.data(), a, m, n, lda);
  copy_columns(r->B.data(), b, m, nrhs, ldb);
  r->bytes = footprint(r->m, r->n, r->nrhs, s->cfg.mode);
  r->info.t_submit = t_in;
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  *id = r->id;
  s->all[r->id] = r;
  s->stats.submitted += 1;
  if (s->cfg.hbm_budget && r->bytes > s->cfg.hbm_budget) {
    r->info.status = RNPX_ERR_NOMEM;
    r->info.t_done = r->info.t_submit;
    r->done = true;
    s->stats.rejected += 1;
    s->cv_done.notify_all();
    return RNPX_ERR_NOMEM;
  }
  s->queue.push_back(r);
  s->cv.notify_all();
  return RNPX_OK;
}

int rnpx_server_poll(rnpx_server srv, long long id, int wait, double* x, size_t ldx, rnpx_request_info* info) {
  Server* s = reinterpret_cast<Server*>(srv);
  if (!s) return -1;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return -2;
  Request* r = it->second;
  if (!r->done) {
    if (!wait) return 1;
    s->cv_done.wait(lk, [&] { return r->done; });
  }
  if (info) *info = r->info;
  if (x && r->info.status == RNPX_OK) {
    if (ldx < (size_t)r->n) return -5;
    for (int q = 0; q < r->nrhs; ++q) std::memcpy(x + (size_t)q * ldx, &r->X[(size_t)q * r->n], sizeof(double) * r->n);
  }
  return 0;
}

int rnpx_server_release(rnpx_server srv, long long id) {
  Server* s = reinterpret_cast<Server*>(srv);
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return -2;
  if (!it->second->done) return 1;
  delete it->second;
  s->all.erase(it);
  return 0;
}

int rnpx_server_stats_get(rnpx_server srv, rnpx_server_stats* out) {
  Server* s = reinterpret_cast<Server*>(srv);
  if (!s || !out) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->queued = (long long)s->queue.size();
  out->bytes_in_use = s->in_use;
  return 0;
}

int rnpx_server_destroy(rnpx_server srv) {
  Server* s = reinterpret_cast<Server*>(srv);
  if (!s) return -1;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
      if (!s->queue.empty()) return false;
      for (auto& kv : s->all)
        if (!kv.second->done) return false;
      return true;
    });
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto& t : s->workers) t.join();
  for (Handle* h : s->handles) delete h;
  hipStreamDestroy(s->panel_stream);
  for (auto& kv : s->all) delete kv.second;
  delete s;
  return 0;
}

}  // extern "C"
