// Serving form: a request loop with continuous batching over the batch engine
// (engine.hip).  Requests are admitted in FIFO order when their footprint
// estimate fits the HBM budget, join the engine at the next step boundary and
// advance level-synchronously with everything else in flight; a finished
// request's partition is copied back and its device memory returns to the
// pool at once, so new requests keep flowing in while others are on the GPU.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gkway.h"
#include "engine.hpp"
#include "runtime.hpp"

using namespace gkrt;

namespace {

struct Request {
  int64_t id = 0;
  gk_params prm{};
  int32_t n = 0, nnz = 0;
  std::vector<int32_t> x, a, w, v;  // host copy of the graph
  int64_t T = 0, foot = 0;
  double t_submit = 0, t_done = 0;
  std::vector<int32_t> part;
  gk_stats S{};
  int state = 0;  // 0 queued, 1 running, 2 finished
  int rc = GK_OK;
  std::string err;
  Buf<int32_t> dout;  // device partition while running
  bool claimed = false;  // a gk_serve_wait caller owns the result
  bool alone = false;    // re-run by itself after a step of several failed
};

}  // namespace

struct gk_server {
  gk_ctx* ctx = nullptr;
  gk_serve_config cfg{};
  int64_t budget = 0;
  std::thread th;
  std::mutex m;
  std::condition_variable cv_in, cv_out;
  std::deque<Request*> queue;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_id = 0;
  int waiters = 0;      // callers inside gk_serve_wait
  bool alone_running = false;  // an isolated re-run is on the device
  bool stop = false;
  bool have_shape = false;
  int32_t max_group = 0, max_moves = 0;
  int64_t submitted = 0, completed = 0, failed = 0, inflight = 0, edges_done = 0, steps = 0, isolated = 0;
  int64_t foot_now = 0, foot_peak = 0, device_peak = 0;
  std::vector<double> lat;
  double first_submit = -1, last_done = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

// Finish r under the server lock (result, statistics, wake-ups).
void finish_locked(gk_server* s, Request* r, int rc, const std::string& err) {
  r->state = 2;
  r->rc = rc;
  r->err = err;
  r->t_done = s->now();
  r->S.t_total_ms = r->t_done - r->t_submit;
  if (rc == GK_OK) {
    ++s->completed;
    s->edges_done += r->nnz / 2;
    s->lat.push_back(r->S.t_total_ms);
    s->last_done = std::max(s->last_done, r->t_done);
  } else {
    ++s->failed;
  }
  std::vector<int32_t>().swap(r->x);
  std::vector<int32_t>().swap(r->a);
  std::vector<int32_t>().swap(r->w);
  std::vector<int32_t>().swap(r->v);
  s->cv_out.notify_all();
}

int admit(gk_server* s, Engine* E, Request* r) {
  Pool& P = s->ctx->pool;
  hipStream_t st = s->ctx->stream;
  auto S0 = std::make_shared<Store>();
  S0->nj = 1;
  S0->n = r->n;
  S0->nnz = r->nnz;
  S0->v_off = {0, r->n};
  S0->e_off = {0, r->nnz};
  GK_TRY(E->upload(S0->dv_off, S0->v_off.data(), 2));
  GK_TRY(S0->bx.alloc(P, int64_t(r->n) + 1));
  GK_TRY(S0->ba.alloc(P, r->nnz));
  GK_TRY(S0->bw.alloc(P, r->nnz));
  GK_TRY(S0->bv.alloc(P, r->n));
  GK_TRY(r->dout.alloc(P, r->n));
  // The host copies stay alive until the request finishes.
  GK_HIP(hipMemcpyAsync(S0->bx.p, r->x.data(), sizeof(int32_t) * (size_t(r->n) + 1), hipMemcpyHostToDevice, st));
  if (r->nnz) {
    GK_HIP(hipMemcpyAsync(S0->ba.p, r->a.data(), sizeof(int32_t) * size_t(r->nnz), hipMemcpyHostToDevice, st));
    GK_HIP(hipMemcpyAsync(S0->bw.p, r->w.data(), sizeof(int32_t) * size_t(r->nnz), hipMemcpyHostToDevice, st));
  }
  GK_HIP(hipMemcpyAsync(S0->bv.p, r->v.data(), sizeof(int32_t) * size_t(r->n), hipMemcpyHostToDevice, st));
  S0->xadj = S0->bx.p;
  S0->adj = S0->ba.p;
  S0->w = S0->bw.p;
  S0->vw = S0->bv.p;
  auto j = std::make_unique<Job>();
  j->id = static_cast<int32_t>(r->id & 0x7fffffff);
  GK_TRY(E->job_setup(*j, r->prm, r->T, r->n));
  j->out = r->dout.p;
  j->user = r;
  JLevel l0;
  l0.n = r->n;
  l0.nnz = r->nnz;
  l0.store = std::move(S0);
  l0.slot = 0;
  j->L.push_back(std::move(l0));
  return E->add_job(std::move(j));
}

void serve_loop(gk_server* s) {
  (void)hipSetDevice(s->ctx->device);
  hipStream_t st = s->ctx->stream;
  gk_params wp{};
  wp.initpart_threads = s->cfg.workers;
  // Admitted and not finished (loop thread only).  A request leaves it under
  // the server lock, before finish_locked makes it visible to its waiter, who
  // may free it at once.
  std::vector<Request*> running;
  auto make_engine = [s, st, &wp, &running]() {
    auto E = std::make_unique<Engine>(s->ctx, st, batch_workers(wp));
    E->done_cb_ = [s, st, &running](Job& j) -> int {
      Request* r = static_cast<Request*>(j.user);
      r->part.resize(r->n);
      if (r->n)
        GK_HIP(hipMemcpyAsync(r->part.data(), r->dout.p, sizeof(int32_t) * size_t(r->n), hipMemcpyDeviceToHost, st));
      GK_HIP(hipStreamSynchronize(st));
      r->dout.release();
      r->S = j.S;
      std::lock_guard<std::mutex> lk(s->m);
      s->foot_now -= r->foot;
      --s->inflight;
      if (r->alone) s->alone_running = false;
      running.erase(std::remove(running.begin(), running.end(), r), running.end());
      finish_locked(s, r, j.rc, j.err);
      return GK_OK;
    };
    return E;
  };
  std::unique_ptr<Engine> E = make_engine();
  while (true) {
    std::vector<Request*> adm;
    {
      std::unique_lock<std::mutex> lk(s->m);
      if (E->live() == 0) s->cv_in.wait(lk, [s] { return s->stop || !s->queue.empty(); });
      if (s->stop && s->queue.empty() && E->live() == 0) break;
      while (!s->queue.empty() && !s->alone_running) {
        Request* r = s->queue.front();
        // An isolated re-run waits for an empty device and runs by itself.
        if (r->alone && s->inflight > 0) break;
        const bool fits = s->foot_now + r->foot <= s->budget || s->inflight == 0;
        const bool slot = s->cfg.max_inflight <= 0 || s->inflight < s->cfg.max_inflight;
        if (!fits || !slot) break;
        s->queue.pop_front();
        s->foot_now += r->foot;
        s->foot_peak = std::max(s->foot_peak, s->foot_now);
        ++s->inflight;
        r->state = 1;
        adm.push_back(r);
        if (r->alone) {
          s->alone_running = true;
          break;
        }
      }
    }
    int rc = GK_OK;
    for (Request* r : adm) {
      running.push_back(r);
      rc = admit(s, E.get(), r);
      if (rc != GK_OK) break;
    }
    if (rc == GK_OK && E->live() > 0) {
      bool worked = false;
      rc = E->step(&worked);
      E->drop_done();
    }
    {
      std::lock_guard<std::mutex> lk(s->m);
      ++s->steps;
      s->device_peak = std::max(s->device_peak, s->ctx->pool.peak());
    }
    if (rc == GK_OK) continue;
    // A failed admission or step.  The engine and its jobs are dropped (their
    // device buffers return to the pool) and a new engine takes over.  A
    // request that failed by itself is finished with the error; when several
    // were on the device, each is queued again, in order, to run alone, so
    // that a failure only reaches the request that causes it.
    const std::string msg = gk_last_error();
    (void)hipStreamSynchronize(st);
    (void)hipGetLastError();
    E.reset();
    E = make_engine();
    std::lock_guard<std::mutex> lk(s->m);
    const bool several = running.size() > 1;
    for (auto it = running.rbegin(); it != running.rend(); ++it) {
      Request* r = *it;
      r->dout.release();
      s->foot_now -= r->foot;
      --s->inflight;
      if (!several || r->alone) {
        finish_locked(s, r, rc, msg);
      } else {
        r->alone = true;
        r->state = 0;
        ++s->isolated;
        s->queue.push_front(r);
      }
    }
    running.clear();
    s->alone_running = false;
  }
  (void)hipStreamSynchronize(st);
}

int take_result(gk_server* s, std::unordered_map<int64_t, std::unique_ptr<Request>>::iterator it,
                int32_t* part_out, gk_stats* stats) {
  Request* r = it->second.get();
  const int rc = r->rc;
  if (rc == GK_OK) {
    if (part_out && r->n) std::memcpy(part_out, r->part.data(), sizeof(int32_t) * size_t(r->n));
    if (stats) *stats = r->S;
  } else {
    gkh::set_error("request " + std::to_string(r->id) + ": " + r->err);
  }
  s->reqs.erase(it);
  return rc;
}

}  // namespace

extern "C" {

int64_t gk_serve_footprint(int32_t n, int32_t nnz, const gk_params* p) {
  if (n <= 0) return 0;
  const int64_t csr = 4LL * (n + 1) + 8LL * nnz + 4LL * n;
  const int32_t fl = p ? gk_coarsen_floor(n, p->k, p->coarsen_policy, p->coarsen_to) : 0;
  // Levels: a level shrinks at least 3x while above the floor.
  int64_t levels = 1, m = n;
  while (m > fl && levels < GK_MAX_LEVELS) {
    m /= 3;
    ++levels;
  }
  return csr * levels + 96LL * n + 4096;
}

int gk_serve_open(gk_ctx* ctx, const gk_serve_config* cfg, gk_server** out) {
  if (!ctx || !out) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  *out = nullptr;
  gk_server* s = new gk_server;
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  s->budget = s->cfg.hbm_budget;
  if (s->budget <= 0) {
    (void)hipSetDevice(ctx->device);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      (void)hipGetLastError();
      delete s;
      gkh::set_error("hipMemGetInfo failed");
      return GK_ERR_HIP;
    }
    s->budget = static_cast<int64_t>(fr / 2);
  }
  ctx->pool.reset_peak();
  s->th = std::thread(serve_loop, s);
  *out = s;
  return GK_OK;
}

namespace {

// Validates one request and copies its host graph (no server state touched).
int make_request(const gk_csr* g, const gk_params* p, std::unique_ptr<Request>* out) {
  if (!g || !p) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  if (p->k < 1 || p->k > kMaxK || p->ubfactor < 0 || p->max_group < 1 || p->max_group > 64 ||
      p->max_moves < 1 || p->max_moves > 1024) {
    gkh::set_error("bad params");
    return GK_ERR_ARG;
  }
  GK_TRY(gk_graph_check(g, (p->flags & GK_FLAG_CHECK) != 0));
  auto r = std::make_unique<Request>();
  r->prm = *p;
  r->n = g->n;
  r->nnz = g->nnz;
  r->x.assign(g->xadj, g->xadj + g->n + 1);
  r->a.assign(g->adjncy, g->adjncy + g->nnz);
  if (g->adjwgt) r->w.assign(g->adjwgt, g->adjwgt + g->nnz); else r->w.assign(g->nnz, 1);
  if (g->vwgt) r->v.assign(g->vwgt, g->vwgt + g->n); else r->v.assign(g->n, 1);
  for (int32_t x : r->v) r->T += x;
  if (r->T >= (int64_t(1) << 31)) { gkh::set_error("total vertex weight >= 2^31"); return GK_ERR_LIMIT; }
  r->foot = gk_serve_footprint(g->n, g->nnz, p);
  *out = std::move(r);
  return GK_OK;
}

// Queues validated requests in order under one hold of the server lock, so
// that the loop sees all of them at its next admission.  All or none.
int enqueue(gk_server* s, std::vector<std::unique_ptr<Request>>& rs, int64_t* ids) {
  std::lock_guard<std::mutex> lk(s->m);
  if (s->stop) { gkh::set_error("server is closing"); return GK_ERR_ARG; }
  int32_t mg = s->have_shape ? s->max_group : rs[0]->prm.max_group;
  int32_t mm = s->have_shape ? s->max_moves : rs[0]->prm.max_moves;
  for (const auto& r : rs)
    if (r->prm.max_group != mg || r->prm.max_moves != mm) {
      gkh::set_error("max_group and max_moves must match the server's first request");
      return GK_ERR_ARG;
    }
  s->have_shape = true;
  s->max_group = mg;
  s->max_moves = mm;
  for (size_t i = 0; i < rs.size(); ++i) {
    auto& r = rs[i];
    r->id = s->next_id++;
    r->t_submit = s->now();
    if (s->first_submit < 0) s->first_submit = r->t_submit;
    ++s->submitted;
    ids[i] = r->id;
    Request* rp = r.get();
    s->reqs.emplace(r->id, std::move(r));
    if (rp->n == 0) finish_locked(s, rp, GK_OK, "");
    else s->queue.push_back(rp);
  }
  s->cv_in.notify_one();
  return GK_OK;
}

}  // namespace

int gk_serve_submit(gk_server* s, const gk_csr* g, const gk_params* p, int64_t* request) {
  return gk_serve_submit_many(s, 1, g, p, request);
}

int gk_serve_submit_many(gk_server* s, int32_t count, const gk_csr* graphs, const gk_params* params,
                         int64_t* requests) {
  if (!s || !graphs || !params || !requests || count < 1) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  std::vector<std::unique_ptr<Request>> rs(static_cast<size_t>(count));
  for (int32_t i = 0; i < count; ++i) GK_TRY(make_request(graphs + i, params + i, &rs[i]));
  return enqueue(s, rs, requests);
}

int gk_serve_poll(gk_server* s, int64_t request, int32_t* part_out, gk_stats* stats) {
  if (!s) { gkh::set_error("null server"); return GK_ERR_ARG; }
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->reqs.find(request);
  if (it == s->reqs.end()) { gkh::set_error("unknown request (or its result was already taken)"); return GK_ERR_ARG; }
  if (it->second->claimed) { gkh::set_error("request is being waited on by another caller"); return GK_ERR_ARG; }
  if (it->second->state != 2) return 0;
  const int rc = take_result(s, it, part_out, stats);
  return rc == GK_OK ? 1 : rc;
}

int gk_serve_wait(gk_server* s, int64_t request, int32_t* part_out, gk_stats* stats) {
  if (!s) { gkh::set_error("null server"); return GK_ERR_ARG; }
  std::unique_lock<std::mutex> lk(s->m);
  auto it = s->reqs.find(request);
  if (it == s->reqs.end()) { gkh::set_error("unknown request (or its result was already taken)"); return GK_ERR_ARG; }
  Request* r = it->second.get();
  if (r->claimed) { gkh::set_error("request is being waited on by another caller"); return GK_ERR_ARG; }
  // Claimed: no other caller can take (and free) the request while this one
  // sleeps; gk_serve_close waits for every waiter to leave.
  r->claimed = true;
  ++s->waiters;
  s->cv_out.wait(lk, [r] { return r->state == 2; });
  const int rc = take_result(s, s->reqs.find(request), part_out, stats);
  --s->waiters;
  s->cv_out.notify_all();
  return rc;
}

int gk_serve_query(gk_server* s, gk_serve_stats* out) {
  if (!s || !out) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  std::lock_guard<std::mutex> lk(s->m);
  std::memset(out, 0, sizeof *out);
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->failed = s->failed;
  out->queued = static_cast<int64_t>(s->queue.size());
  out->inflight = s->inflight;
  out->edges_done = s->edges_done;
  out->steps = s->steps;
  out->footprint_now = s->foot_now;
  out->footprint_peak = s->foot_peak;
  out->device_peak = s->device_peak;
  out->isolated = s->isolated;
  if (!s->lat.empty()) {
    std::vector<double> l = s->lat;
    std::sort(l.begin(), l.end());
    auto rank = [&l](double q) {  // nearest-rank percentile
      size_t i = static_cast<size_t>(std::ceil(q * double(l.size())));
      if (i < 1) i = 1;
      return l[std::min(i, l.size()) - 1];
    };
    out->p50_ms = rank(0.50);
    out->p99_ms = rank(0.99);
    out->max_ms = l.back();
  }
  const double span = s->last_done - s->first_submit;
  out->edges_per_s = span > 0 ? double(s->edges_done) / (span / 1e3) : 0.0;
  return GK_OK;
}

int gk_serve_close(gk_server* s) {
  if (!s) return GK_OK;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
  }
  s->cv_in.notify_all();
  if (s->th.joinable()) s->th.join();
  {
    // Every request has finished; callers still inside gk_serve_wait take
    // their results before the server goes away.
    std::unique_lock<std::mutex> lk(s->m);
    s->cv_out.wait(lk, [s] { return s->waiters == 0; });
  }
  delete s;
  return GK_OK;
}

}  // extern "C"
