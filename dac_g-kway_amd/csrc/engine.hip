// Batch engine: the multilevel V-cycle of many graphs ("jobs") advanced
// together, one step at a time.  A step
//   - coarsens every job still coarsening by one level: their current levels
//     are one packed CSR (global ids) and the single-graph coarsening stages
//     run once on the pack;
//   - starts the host initial partition of every job whose coarsening ended
//     (host worker pool, overlapping the device work of the other jobs);
//   - refines every job in the refinement phase by one level: fine graphs,
//     projection maps and coarse partitions are packed, and the refinement
//     rounds run on the pack with per-job caps, part weights and move order.
// Jobs never share an edge and every order-sensitive step is segmented by
// job, so each job's result equals gk_partition on that graph alone.
//
// Levels live in shared "stores": the output of one contraction holds the
// coarse graphs of all jobs of that step, and each job's level is a slice of
// it.  When the jobs coarsened next are exactly the slices of one store, in
// order, the store itself is the next pack (no copy).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <future>
#include <memory>
#include <mutex>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gkway.h"
#include "batch.hpp"
#include "engine.hpp"

#include <ctime>
#include "prim.hpp"
#include "runtime.hpp"

using gkd::DGraph;
using gkd::PackSrc;
using gkd::UnpackDst;

namespace gkrt {

// ---------------------------------------------------------------- thread pool
WorkerPool::WorkerPool(int n) {
  for (int i = 0; i < n; ++i)
    th_.emplace_back([this] {
      while (true) {
        std::function<void()> f;
        {
          std::unique_lock<std::mutex> lk(m_);
          cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
          if (stop_ && q_.empty()) return;
          f = std::move(q_.front());
          q_.pop();
        }
        f();
      }
    });
}

WorkerPool::~WorkerPool() {
  {
    std::lock_guard<std::mutex> lk(m_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto& t : th_) t.join();
}

std::future<int> WorkerPool::submit(std::function<int()> f) {
  auto task = std::make_shared<std::packaged_task<int()>>(std::move(f));
  std::future<int> fut = task->get_future();
  {
    std::lock_guard<std::mutex> lk(m_);
    q_.emplace([task] { (*task)(); });
  }
  cv_.notify_one();
  return fut;
}

// ---------------------------------------------------------------- pinned staging
Pinned::~Pinned() {
  if (base_) (void)hipHostFree(base_);
}

int Pinned::get(hipStream_t st, size_t bytes, void** out) {
  bytes = (bytes + 255) & ~size_t(255);
  if (used_ + bytes > cap_) {
    GK_HIP(hipStreamSynchronize(st));
    used_ = 0;
    if (bytes > cap_) {
      if (base_) GK_HIP(hipHostFree(base_));
      base_ = nullptr;
      cap_ = std::max<size_t>(bytes, std::max<size_t>(2 * cap_, size_t(1) << 20));
      if (hipHostMalloc(reinterpret_cast<void**>(&base_), cap_, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        base_ = nullptr;
        cap_ = 0;
        gkh::set_error("hipHostMalloc of staging memory failed");
        return GK_ERR_NOMEM;
      }
    }
  }
  *out = base_ + used_;
  used_ += bytes;
  return GK_OK;
}

// ---------------------------------------------------------------- engine
double Engine::now_ms() const {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
}

Engine::Engine(gk_ctx* ctx, hipStream_t st, int workers)
    : ctx_(ctx), st_(st), P_(ctx->pool), t0_(std::chrono::steady_clock::now()), up_(ctx->up), dl_(ctx->dl),
      pool_(workers) {}

Engine::~Engine() {
  // GK_ENGINE_PROFILE=1: one line per engine with the host wall time of its
  // step phases (diagnostics only).
  const char* e = std::getenv("GK_ENGINE_PROFILE");
  if (e && e[0] == '1' && steps_ > 0)
    std::fprintf(stderr,
                 "[gk engine] steps %lld rounds %lld | coarsen %.1f ms, init launch %.1f ms, init collect %.1f ms, "
                 "refine %.1f ms (rounds %.1f ms), idle wait %.1f ms, total %.1f ms; initial partitions %.1f ms "
                 "of worker CPU\n",
                 (long long)steps_, (long long)rounds_, prof_ms_[0], prof_ms_[1], prof_ms_[2], prof_ms_[3],
                 prof_ms_[5], prof_ms_[4], now_ms(), init_cpu_us_.load() / 1e3);
}

int Engine::sync() {
  GK_HIP(hipStreamSynchronize(st_));
  up_.drained();
  return GK_OK;
}

int Engine::download(std::initializer_list<Fetch> items) {
  size_t tot = 0;
  for (const Fetch& f : items) tot += (f.bytes + 255) & ~size_t(255);
  void* h = nullptr;
  GK_TRY(dl_.get(st_, tot, &h));
  char* c = static_cast<char*>(h);
  size_t off = 0;
  for (const Fetch& f : items) {
    if (f.bytes) GK_HIP(hipMemcpyAsync(c + off, f.dev, f.bytes, hipMemcpyDeviceToHost, st_));
    off += (f.bytes + 255) & ~size_t(255);
  }
  GK_TRY(sync());
  off = 0;
  for (const Fetch& f : items) {
    if (f.bytes) std::memcpy(f.host, c + off, f.bytes);
    off += (f.bytes + 255) & ~size_t(255);
  }
  dl_.drained();
  return GK_OK;
}

StoreP Engine::whole_store(const std::vector<const JLevel*>& lv) {
  if (lv.empty()) return nullptr;
  Store* s = lv[0]->store.get();
  if (static_cast<int32_t>(lv.size()) != s->nj) return nullptr;
  for (size_t i = 0; i < lv.size(); ++i)
    if (lv[i]->store.get() != s || lv[i]->slot != static_cast<int32_t>(i)) return nullptr;
  return lv[0]->store;
}

PackSrc Engine::src_of(const JLevel& lv) {
  const Store& s = *lv.store;
  const int32_t vo = s.v_off[lv.slot], eo = s.e_off[lv.slot];
  PackSrc p{};
  p.xadj = s.xadj + vo;
  p.adj = s.adj + eo;
  p.w = s.w + eo;
  p.vw = s.vw + vo;
  p.xbase = eo;
  p.abase = vo;
  if (lv.cmap) {
    p.cmap = lv.cmap->p + lv.cmap_off;
    p.cbase = lv.cbase;
  }
  return p;
}

int Engine::pack_copy(const std::vector<PackSrc>& src, const std::vector<int32_t>& nv,
                      const std::vector<int32_t>& ne, const std::vector<int32_t>* ncv, StoreP& out,
                      RefineSide* rs) {
  const int32_t J = static_cast<int32_t>(src.size());
  auto S = std::make_shared<Store>();
  S->nj = J;
  S->v_off.assign(J + 1, 0);
  S->e_off.assign(J + 1, 0);
  std::vector<int32_t> c_off(J + 1, 0);
  for (int32_t i = 0; i < J; ++i) {
    const int64_t a = int64_t(S->v_off[i]) + nv[i], b = int64_t(S->e_off[i]) + ne[i];
    const int64_t c = int64_t(c_off[i]) + (ncv ? (*ncv)[i] : 0);
    if (a >= INT32_MAX || b >= INT32_MAX || c >= INT32_MAX) {
      gkh::set_error("packed batch exceeds 2^31 vertices or entries");
      return GK_ERR_LIMIT;
    }
    S->v_off[i + 1] = int32_t(a);
    S->e_off[i + 1] = int32_t(b);
    c_off[i + 1] = int32_t(c);
  }
  S->n = S->v_off[J];
  S->nnz = S->e_off[J];
  const int32_t ctot = c_off[J];
  Buf<PackSrc> dsrc;
  Buf<int32_t> de_off, dc_off;
  GK_TRY(upload(dsrc, src.data(), src.size()));
  GK_TRY(upload(S->dv_off, S->v_off.data(), S->v_off.size()));
  GK_TRY(upload(de_off, S->e_off.data(), S->e_off.size()));
  GK_TRY(upload(dc_off, c_off.data(), c_off.size()));
  GK_TRY(S->bx.alloc(P_, int64_t(S->n) + 1));
  GK_TRY(S->ba.alloc(P_, S->nnz));
  GK_TRY(S->bw.alloc(P_, S->nnz));
  GK_TRY(S->bv.alloc(P_, S->n));
  if (rs) {
    GK_TRY(rs->vjob.alloc(P_, S->n));
    GK_TRY(rs->cmap.alloc(P_, S->n));
    GK_TRY(rs->cpart.alloc(P_, ctot));
    GK_TRY(rs->cbnd.alloc(P_, ctot));
    GK_TRY(rs->cgain.alloc(P_, ctot));
    GK_TRY(rs->cdest.alloc(P_, ctot));
    rs->c_off = c_off;
  }
  GK_STAGE(ctx_, GK_ST_PACK, 20LL * S->n + 16LL * S->nnz + (rs ? 16LL * S->n + 24LL * ctot : 0),
           gkd::launch_pack(J, dsrc.p, S->dv_off.p, de_off.p, dc_off.p, S->n, S->nnz, ctot, S->bx.p, S->ba.p,
                            S->bw.p, S->bv.p, rs ? rs->vjob.p : nullptr, rs ? rs->cmap.p : nullptr,
                            rs ? rs->cpart.p : nullptr, rs ? rs->cbnd.p : nullptr, rs ? rs->cgain.p : nullptr,
                            rs ? rs->cdest.p : nullptr, st_));
  S->xadj = S->bx.p;
  S->adj = S->ba.p;
  S->w = S->bw.p;
  S->vw = S->bv.p;
  out = std::move(S);
  return GK_OK;
}

void Engine::mark(Job& j, int phase) {
  const double t = now_ms();
  if (phase == INIT_READY) j.S.t_coarsen_ms = t - j.t_start;
  if (phase == REFINE || (phase == DONE && j.phase == INIT_RUNNING)) j.S.t_init_ms = t - j.t_start - j.S.t_coarsen_ms;
  if (phase == DONE) {
    j.S.t_total_ms = t - j.t_start;
    j.S.t_refine_ms = j.S.t_total_ms - j.S.t_coarsen_ms - j.S.t_init_ms;
    j.t_done = t;
    ++ndone_;
  }
  j.phase = phase;
}

// ------------------------------------------------------------ coarsening
int Engine::coarsen_step(bool* worked) {
  std::vector<Job*> ids;
  bool ready = false;
  for (auto& jp : jobs_) {
    Job& j = *jp;
    if (j.phase != COARSEN) continue;
    if (static_cast<int32_t>(j.L.size()) >= j.maxlv || !(j.L.back().n > j.floor_n)) {
      mark(j, INIT_READY);
      ready = true;
      continue;
    }
    ids.push_back(&j);
  }
  // Jobs whose coarsening ended go to the host workers before this step's
  // device work, not after it (measured on the 4096-graph batch: 866-872
  // against 877-882 ms per call, alternating on one box).
  if (ready) GK_TRY(init_step(worked));
  if (ids.empty()) return GK_OK;
  *worked = true;
  // Jobs of one coarsening pack share the stage parameters G (max_group).
  const int32_t G = ids[0]->prm.max_group;
  std::vector<const JLevel*> lv;
  for (Job* j : ids) lv.push_back(&j->L.back());
  StoreP pk = whole_store(lv);
  if (!pk) {
    std::vector<PackSrc> src;
    std::vector<int32_t> nv, ne;
    for (const JLevel* l : lv) {
      src.push_back(src_of(*l));
      nv.push_back(l->n);
      ne.push_back(l->nnz);
    }
    GK_TRY(pack_copy(src, nv, ne, nullptr, pk, nullptr));
  }
  const Store& F = *pk;
  const DGraph g = F.graph();
  const int32_t n = g.n, J = F.nj;
  Buf<int32_t> cand, label, dist, order, cptr, ce;
  auto cmap = std::make_shared<Buf<int32_t>>();
  GK_TRY(cand.alloc(P_, n));
  GK_TRY(label.alloc(P_, n));
  GK_TRY(cmap->alloc(P_, n));
  GK_STAGE(ctx_, GK_ST_MATCH, 16LL * g.nnz + 8LL * n + 4,
           match_stage(ctx_, st_, g, cand.p));
  int32_t rounds = 0;
  GK_TRY(group_stage(ctx_, st_, n, cand.p, label.p, dist, &rounds));
  int32_t nc = 0;
  GK_TRY(cluster_stage(ctx_, st_, n, label.p, dist.p, rounds + 1, G, order, cmap->p, cptr, &nc));
  auto C = std::make_shared<Store>();
  GK_TRY(C->dv_off.alloc(P_, int64_t(J) + 1));
  GK_TRY(gkd::launch_job_offsets(J, F.dv_off.p, order.p, cmap->p, nc, C->dv_off.p, st_));
  GK_TRY(C->bv.alloc(P_, nc));
  GK_STAGE(ctx_, GK_ST_CLUSTER, 8LL * n + 12LL * nc,
           gkd::launch_cvwgt(nc, cptr.p, order.p, g.vw, C->bv.p, st_));
  Level c;
  GK_TRY(contract_stage(ctx_, st_, g, nc, cmap->p, cptr.p, order.p, c));
  GK_TRY(ce.alloc(P_, int64_t(J) + 1));
  GK_TRY(gkd::launch_job_eoffsets(J, C->dv_off.p, c.xadj.p, ce.p, st_));
  C->v_off.resize(J + 1);
  C->e_off.resize(J + 1);
  GK_TRY(download({{C->v_off.data(), C->dv_off.p, sizeof(int32_t) * (J + 1)},
                   {C->e_off.data(), ce.p, sizeof(int32_t) * (J + 1)}}));
  C->nj = J;
  C->n = nc;
  C->nnz = c.g.nnz;
  C->bx = std::move(c.xadj);
  C->ba = std::move(c.adj);
  C->bw = std::move(c.w);
  C->xadj = C->bx.p;
  C->adj = C->ba.p;
  C->w = C->bw.p;
  C->vw = C->bv.p;
  for (int32_t i = 0; i < J; ++i) {
    Job& j = *ids[i];
    const int32_t nci = C->v_off[i + 1] - C->v_off[i];
    const int32_t nfi = F.v_off[i + 1] - F.v_off[i];
    bool keep = !(nci == nfi || nci < j.prm.k + 1);
    if (j.prm.coarsen_policy != GK_COARSEN_REF && nci < j.floor_n) keep = false;
    if (!keep) {
      mark(j, INIT_READY);
      continue;
    }
    JLevel& f = j.L.back();
    f.cmap = cmap;
    f.cmap_off = F.v_off[i];
    f.cbase = C->v_off[i];
    j.S.levels[j.L.size() - 1].sweeps = rounds;
    JLevel cl;
    cl.n = nci;
    cl.nnz = C->e_off[i + 1] - C->e_off[i];
    cl.store = C;
    cl.slot = i;
    j.L.push_back(std::move(cl));
    gk_level_stats& ls = j.S.levels[j.L.size() - 1];
    ls.n = nci;
    ls.nnz = j.L.back().nnz;
  }
  return GK_OK;
}

// ------------------------------------------------------------ initial partition
int Engine::init_step(bool* worked) {
  std::vector<Job*> ids;
  for (auto& jp : jobs_)
    if (jp->phase == INIT_READY) ids.push_back(jp.get());
  if (ids.empty()) return GK_OK;
  *worked = true;
  std::vector<const JLevel*> lv;
  for (Job* j : ids) lv.push_back(&j->L.back());
  StoreP pk = whole_store(lv);
  if (!pk) {
    std::vector<PackSrc> src;
    std::vector<int32_t> nv, ne;
    for (const JLevel* l : lv) {
      src.push_back(src_of(*l));
      nv.push_back(l->n);
      ne.push_back(l->nnz);
    }
    GK_TRY(pack_copy(src, nv, ne, nullptr, pk, nullptr));
  }
  const Store& Q = *pk;
  std::vector<int32_t> hx(int64_t(Q.n) + 1), ha(Q.nnz), hw(Q.nnz), hv(Q.n);
  GK_TRY(download({{hx.data(), Q.xadj, sizeof(int32_t) * (size_t(Q.n) + 1)},
                   {ha.data(), Q.adj, sizeof(int32_t) * size_t(Q.nnz)},
                   {hw.data(), Q.w, sizeof(int32_t) * size_t(Q.nnz)},
                   {hv.data(), Q.vw, sizeof(int32_t) * size_t(Q.n)}}));
  for (size_t i = 0; i < ids.size(); ++i) {
    Job& j = *ids[i];
    const int32_t vo = Q.v_off[i], eo = Q.e_off[i];
    const int32_t nt = Q.v_off[i + 1] - vo, et = Q.e_off[i + 1] - eo;
    j.hx.resize(int64_t(nt) + 1);
    for (int32_t v = 0; v <= nt; ++v) j.hx[v] = hx[int64_t(vo) + v] - eo;
    j.ha.resize(et);
    for (int32_t e = 0; e < et; ++e) j.ha[e] = ha[int64_t(eo) + e] - vo;
    j.hw.assign(hw.begin() + eo, hw.begin() + eo + et);
    j.hv.assign(hv.begin() + vo, hv.begin() + vo + nt);
    j.nl = static_cast<int32_t>(j.L.size());
    j.S.nlevels = j.nl;
    j.S.coarsest_n = nt;
    Job* jp = &j;
    j.fut = pool_.submit([this, jp]() -> int {
      Job& J2 = *jp;
      timespec c0{}, c1{};
      clock_gettime(CLOCK_THREAD_CPUTIME_ID, &c0);
      const int32_t nt2 = static_cast<int32_t>(J2.hv.size());
      gk_csr hc;
      hc.n = nt2;
      hc.nnz = static_cast<int32_t>(J2.ha.size());
      hc.xadj = J2.hx.data();
      hc.adjncy = J2.ha.data();
      hc.adjwgt = J2.hw.data();
      hc.vwgt = J2.hv.data();
      J2.hpart.assign(nt2, 0);
      J2.hbnd.assign(nt2, 0);
      int64_t cut0 = 0;
      const int rc = gk_initpart(&hc, J2.prm.k, J2.cap, J2.prm.seed, 1, J2.hpart.data(), &cut0);
      if (rc != GK_OK) {
        J2.err = gk_last_error();
        return rc;
      }
      J2.hpw.assign(J2.prm.k, 0);
      for (int32_t v = 0; v < nt2; ++v) {
        J2.hpw[J2.hpart[v]] += J2.hv[v];
        for (int32_t e = J2.hx[v]; e < J2.hx[v + 1]; ++e)
          if (J2.hpart[J2.ha[e]] != J2.hpart[v]) {
            J2.hbnd[v] = 1;
            break;
          }
      }
      J2.hcut = cut0;
      clock_gettime(CLOCK_THREAD_CPUTIME_ID, &c1);
      init_cpu_us_ += (c1.tv_sec - c0.tv_sec) * 1000000LL + (c1.tv_nsec - c0.tv_nsec) / 1000;
      return GK_OK;
    });
    mark(j, INIT_RUNNING);
  }
  return GK_OK;
}

int Engine::collect_inits(bool block, bool* worked) {
  std::vector<Job*> ready;
  Job* first = nullptr;
  for (auto& jp : jobs_) {
    if (jp->phase != INIT_RUNNING) continue;
    if (!first) first = jp.get();
    if (jp->fut.wait_for(std::chrono::seconds(0)) == std::future_status::ready) ready.push_back(jp.get());
  }
  if (ready.empty() && block && first) {
    first->fut.wait();
    ready.push_back(first);
  }
  if (ready.empty()) return GK_OK;
  *worked = true;
  int64_t tot = 0;
  std::vector<int32_t> off(ready.size() + 1, 0);
  for (size_t i = 0; i < ready.size(); ++i) {
    Job& j = *ready[i];
    const int rc = j.fut.get();
    if (rc != GK_OK) {
      gkh::set_error("initial partition of graph " + std::to_string(j.id) + ": " + j.err);
      return rc;
    }
    tot += static_cast<int64_t>(j.hpart.size());
    if (tot >= INT32_MAX) { gkh::set_error("coarsest graphs exceed 2^31 vertices"); return GK_ERR_LIMIT; }
    off[i + 1] = int32_t(tot);
  }
  auto pb = std::make_shared<Buf<int32_t>>();
  auto bb = std::make_shared<Buf<int32_t>>();
  {
    std::vector<int32_t> hp(tot), hb(tot);
    for (size_t i = 0; i < ready.size(); ++i) {
      std::copy(ready[i]->hpart.begin(), ready[i]->hpart.end(), hp.begin() + off[i]);
      std::copy(ready[i]->hbnd.begin(), ready[i]->hbnd.end(), hb.begin() + off[i]);
    }
    GK_TRY(upload(*pb, hp.data(), hp.size()));
    GK_TRY(upload(*bb, hb.data(), hb.size()));
  }
  // Jobs whose coarsest level is level 0 are finished by the initial partition.
  std::vector<UnpackDst> dst(ready.size());
  bool any_final = false;
  for (size_t i = 0; i < ready.size(); ++i) {
    Job& j = *ready[i];
    std::memset(&dst[i], 0, sizeof dst[i]);
    j.S.init_cut = j.hcut;
    j.S.levels[j.nl - 1].cut_after = j.hcut;
    j.cut = j.hcut;
    if (j.nl == 1) {
      dst[i].part = j.out;
      any_final = true;
      // Final metrics straight from the host copy of level 0.
      int64_t c = 0;
      const int32_t nt = static_cast<int32_t>(j.hv.size());
      for (int32_t v = 0; v < nt; ++v)
        for (int32_t e = j.hx[v]; e < j.hx[v + 1]; ++e)
          if (v < j.ha[e] && j.hpart[v] != j.hpart[j.ha[e]]) c += j.hw[e];
      finish_stats(j, c, j.hcut, std::vector<int64_t>(j.hpw.begin(), j.hpw.end()));
      mark(j, DONE);
    } else {
      j.cpart = pb;
      j.cbnd = bb;
      j.coff = off[i];
      j.cur = j.nl - 2;
      mark(j, REFINE);
    }
    j.release_host();
  }
  if (any_final) {
    Buf<UnpackDst> ddst;
    Buf<int32_t> doff;
    GK_TRY(upload(ddst, dst.data(), dst.size()));
    GK_TRY(upload(doff, off.data(), off.size()));
    GK_TRY(gkd::launch_unpack_part(static_cast<int32_t>(ready.size()), ddst.p, doff.p, int32_t(tot), pb->p,
                                   nullptr, st_));
    GK_TRY(sync());
    for (Job* j : ready)
      if (j->phase == DONE) GK_TRY(on_done(*j));
  }
  return GK_OK;
}

void Engine::finish_stats(Job& j, int64_t cut, int64_t tracked, const std::vector<int64_t>& pw) {
  j.S.cut = cut;
  j.S.cut_tracked = tracked;
  int64_t mx = 0;
  for (int64_t x : pw) mx = std::max(mx, x);
  j.S.max_part_wgt = mx;
  j.S.imbalance = j.T > 0 ? double(mx) * j.prm.k / double(j.T) : 1.0;
  if (cut != tracked && j.rc == GK_OK) {
    j.rc = GK_ERR_INTERNAL;
    j.err = "tracked cut " + std::to_string(tracked) + " != recomputed cut " + std::to_string(cut);
  }
}

// ------------------------------------------------------------ refinement
int Engine::refine_step(bool* worked) {
  std::vector<Job*> ids;
  for (auto& jp : jobs_)
    if (jp->phase == REFINE) ids.push_back(jp.get());
  if (ids.empty()) return GK_OK;
  *worked = true;
  const int32_t J = static_cast<int32_t>(ids.size());
  const int32_t M = ids[0]->prm.max_moves;
  std::vector<PackSrc> src(J);
  std::vector<int32_t> nv(J), ne(J), ncv(J), pw_off(J + 1, 0), kk(J);
  std::vector<long long> cap(J), cut(J);
  int32_t kmax = 1;
  for (int32_t i = 0; i < J; ++i) {
    Job& j = *ids[i];
    const JLevel& f = j.L[j.cur];
    src[i] = src_of(f);
    src[i].cpart = j.cpart->p + j.coff;
    src[i].cbnd = j.cbnd->p + j.coff;
    nv[i] = f.n;
    ne[i] = f.nnz;
    if (j.resume) {
      // The level continues from its own state (identity projection): the
      // partition, boundary flags, gains and destinations the uninterrupted
      // rounds would have.
      src[i].cmap = nullptr;
      src[i].cgain = j.cgain->p + j.coff;
      src[i].cdest = j.cdest->p + j.coff;
      ncv[i] = f.n;
    } else {
      ncv[i] = j.L[j.cur + 1].n;
    }
    kk[i] = j.prm.k;
    kmax = std::max(kmax, j.prm.k);
    pw_off[i + 1] = pw_off[i] + j.prm.k;
    cap[i] = j.cap;
    cut[i] = j.cut;
  }
  std::vector<int32_t> hpw(pw_off[J]);
  for (int32_t i = 0; i < J; ++i) std::copy(ids[i]->hpw.begin(), ids[i]->hpw.end(), hpw.begin() + pw_off[i]);
  RefineSide rs;
  StoreP pk;
  GK_TRY(pack_copy(src, nv, ne, &ncv, pk, &rs));
  const Store& F = *pk;
  const DGraph g = F.graph();
  const int32_t N = g.n;
  Buf<int32_t> dpw_off, dkk, dpw, upd, wl, heavy, list, hcnt, active, drounds, lastcnt, dmax, xlist, dthr, wpos,
      sscr;
  Buf<long long> dcap, dcut, dmoved, dpaths;
  Buf<uint64_t> cand, movable, selbits, defbits;
  Buf<uint32_t> hkey;
  auto part = std::make_shared<Buf<int32_t>>();
  auto bnd = std::make_shared<Buf<int32_t>>();
  auto gainp = std::make_shared<Buf<int32_t>>();
  auto destp = std::make_shared<Buf<int32_t>>();
  Buf<int32_t>& gain = *gainp;
  Buf<int32_t>& dest = *destp;
  GK_TRY(upload(dpw_off, pw_off.data(), pw_off.size()));
  GK_TRY(upload(dkk, kk.data(), kk.size()));
  GK_TRY(upload(dpw, hpw.data(), hpw.size()));
  GK_TRY(upload(dcap, cap.data(), cap.size()));
  GK_TRY(upload(dcut, cut.data(), cut.size()));
  GK_TRY(part->alloc(P_, N));
  GK_TRY(bnd->alloc(P_, N));
  GK_TRY(gain.alloc(P_, N));
  GK_TRY(dest.alloc(P_, N));
  GK_TRY(upd.alloc(P_, N));
  GK_TRY(wl.alloc(P_, N));
  GK_TRY(heavy.alloc(P_, N));
  GK_TRY(list.alloc(P_, N));
  GK_TRY(cand.alloc(P_, (int64_t(N) + 63) / 64 + 1));
  const int64_t nwords = (int64_t(N) + 63) / 64;
  GK_TRY(movable.alloc(P_, nwords + 1));
  GK_TRY(selbits.alloc(P_, nwords + 1));
  GK_TRY(defbits.alloc(P_, nwords + 1));
  GK_TRY(wpos.alloc(P_, nwords + 1));
  GK_TRY(sscr.alloc(P_, gkd::select_scratch_ints(N)));
  GK_TRY(xlist.alloc(P_, N));
  GK_TRY(dthr.alloc(P_, J));
  gkd::fill<int32_t>(dthr.p, 1, J, st_);
  gkd::fill<uint64_t>(selbits.p, 0ull, nwords + 1, st_);
  Buf<unsigned long long> selbytes;  // the selection kernels' algorithmic bytes, per block slot
  GK_TRY(selbytes.alloc(P_, gkd::kSelectByteSlots));
  gkd::fill<unsigned long long>(selbytes.p, 0ull, gkd::kSelectByteSlots, st_);
  const gkd::SelectBufs sb{movable.p, selbits.p, defbits.p, wpos.p, list.p, sscr.p, selbytes.p};
  GK_TRY(hkey.alloc(P_, N));
  GK_TRY(hcnt.alloc(P_, 2));
  GK_TRY(active.alloc(P_, 1));
  GK_TRY(drounds.alloc(P_, J));
  GK_TRY(lastcnt.alloc(P_, J));
  gkd::fill<int32_t>(lastcnt.p, 1, J, st_);
  GK_TRY(dmax.alloc(P_, 1));
  GK_TRY(dmoved.alloc(P_, J));
  gkd::fill<int32_t>(drounds.p, 0, J, st_);
  gkd::fill<long long>(dmoved.p, 0, J, st_);
  GK_TRY(dpaths.alloc(P_, 3LL * J));
  gkd::fill<long long>(dpaths.p, 0, 3LL * J, st_);
  Buf<long long> dab;  // the round kernel's algorithmic bytes per job
  GK_TRY(dab.alloc(P_, J));
  gkd::fill<long long>(dab.p, 0, J, st_);
  // Gain-pass statistics: a (vertices, entries) pair per block of the gain
  // launches (grids of at most kGainSlots blocks), summed after the rounds
  // into the stage's algorithmic bytes.
  constexpr int kGainSlots = gkd::kGainMaxBlocks;
  Buf<unsigned long long> gwork;
  GK_TRY(gwork.alloc(P_, 2 * kGainSlots));
  gkd::fill<unsigned long long>(gwork.p, 0ull, 2 * kGainSlots, st_);
  GK_STAGE(ctx_, GK_ST_PROJECT, 44LL * N,
           gkd::launch_project_b(N, rs.cmap.p, rs.cpart.p, rs.cbnd.p, rs.cgain.p, rs.cdest.p, part->p, bnd->p,
                                 gain.p, dest.p, upd.p, st_));
  GK_STAGE(ctx_, GK_ST_PROJECT, 8LL * N + N / 8, gkd::launch_cand_init(N, bnd->p, gain.p, cand.p, st_));
  const int lanes = gain_lanes(g);
  // Rows longer than the gain pass's lane count: the pass takes them itself
  // for k <= kGainInlineK; above it they go to the heavy-row kernel, whose
  // launch is skipped when the pack has none (which costs a read-back of the
  // largest row length).
  bool heavy_rows = false;
  if (gkd::gain_heavy_separate(kmax)) {
    int32_t maxdeg = 0;
    GK_TRY(gkd::launch_max_deg(g, dmax.p, st_));
    GK_TRY(fetch(ctx_, st_, {{&maxdeg, dmax.p, sizeof maxdeg}}));
    heavy_rows = maxdeg > lanes;
  }
  gkd::fill<int32_t>(hcnt.p, 0, 2, st_);
  GK_STAGE(ctx_, GK_ST_PROJECT, 4LL * N, gkd::launch_compact(N, upd.p, wl.p, hcnt.p, st_));
  GK_STAGE(ctx_, GK_ST_GAIN, 0,
           gkd::launch_gain_wl(g, kmax, lanes, wl.p, hcnt.p, heavy.p, part->p, upd.p, bnd->p, gain.p, dest.p,
                               gwork.p, N, cand.p, st_, steal_of(ctx_)));
  if (heavy_rows)
    GK_STAGE(ctx_, GK_ST_GAIN_HEAVY, 0,
             gkd::launch_gain_heavy(g, kmax, heavy.p, hcnt.p, part->p, upd.p, bnd->p, gain.p, dest.p, gwork.p,
                                    cand.p, st_));
  const int64_t avg_deg = N > 0 ? (int64_t(g.nnz) + N - 1) / N : 0;
  const int64_t wl_items = std::min<int64_t>(N, int64_t(J) * M * (1 + 2 * avg_deg));
  // Rounds run back to back on the device; the host reads the last round's
  // move count every kRoundsPerCheck rounds (every 8, 16, then 32 once the
  // step has run 32: a check costs a copy and a host sync).  A round in which no job selects
  // a move changes nothing, so rounds past convergence are harmless.  After
  // kMaxRounds rounds the step ends; jobs still moving resume their level in
  // the next step, so one slow job does not hold back the others.  The fixed
  // cost of a step (pack, projection, candidate set, a host sync) grows with
  // the pack, so bigger packs run more rounds per step: 16 per 16 Mi
  // vertices + entries, at most 256.
  constexpr int kRoundsPerCheck = 4;
  const int64_t pack_items = int64_t(N) + int64_t(g.nnz);
  const int kMaxRounds = static_cast<int>(std::min<int64_t>(256, 16 * std::max<int64_t>(1, pack_items >> 24)));
  auto one_round = [&]() -> int {
    GK_STAGE(ctx_, GK_ST_SELECT, 0,
             gkd::launch_select_list(g, rs.vjob.p, gain.p, dest.p, dpw.p, dpw_off.p, dcap.p, dthr.p, cand.p, sb,
                                     hcnt.p, active.p, st_));
    GK_STAGE(ctx_, GK_ST_BALANCE, 0,
             gkd::launch_round_b(J, kmax, F.dv_off.p, M, dkk.p, dcap.p, dpw_off.p, g, part->p, dest.p, gain.p,
                                 dpw.p, dcut.p, drounds.p, dmoved.p, sb, xlist.p, hkey.p, upd.p, wl.p, hcnt.p,
                                 lastcnt.p, active.p, cand.p, dthr.p, dpaths.p, nullptr, dab.p, st_));
    GK_STAGE(ctx_, GK_ST_GAIN, 0,
             gkd::launch_gain_wl(g, kmax, lanes, wl.p, hcnt.p, heavy.p, part->p, upd.p, bnd->p, gain.p, dest.p,
                                 gwork.p, wl_items, cand.p, st_, steal_of(ctx_)));
    if (heavy_rows)
      GK_STAGE(ctx_, GK_ST_GAIN_HEAVY, 0,
               gkd::launch_gain_heavy(g, kmax, heavy.p, hcnt.p, part->p, upd.p, bnd->p, gain.p, dest.p, gwork.p,
                                      cand.p, st_));
    return GK_OK;
  };
  // With GK_FLAG_GRAPH_ROUNDS (and no round stage event-timed), the
  // kRoundsPerCheck rounds are captured once and every check interval
  // replays that graph (intervals are multiples of it).
  RoundGraph rg;
  if (graph_rounds_ && !ctx_->clock.timed(GK_ST_SELECT) && !ctx_->clock.timed(GK_ST_BALANCE) &&
      !ctx_->clock.timed(GK_ST_GAIN) && !ctx_->clock.timed(GK_ST_GAIN_HEAVY))
    GK_TRY(rg.capture(ctx_, st_, [&]() -> int {
      for (int r = 0; r < kRoundsPerCheck; ++r) GK_TRY(one_round());
      return GK_OK;
    }));
  const double tr0 = now_ms();
  int done_rounds = 0, per_check = kRoundsPerCheck;
  while (done_rounds < kMaxRounds) {
    const int nr = std::min(per_check, kMaxRounds - done_rounds);
    rounds_ += nr;
    done_rounds += nr;
    if (rg.ready() && nr % kRoundsPerCheck == 0) {
      for (int r = 0; r < nr; r += kRoundsPerCheck) GK_TRY(rg.replay(ctx_, st_));
    } else {
      for (int r = 0; r < nr; ++r) GK_TRY(one_round());
    }
    int32_t live = 0;
    GK_TRY(fetch(ctx_, st_, {{&live, active.p, sizeof live}}));
    if (live == 0) break;
    if (done_rounds >= 32 && per_check < 32) per_check *= 2;
  }
  prof_ms_[5] += now_ms() - tr0;
  std::vector<long long> hcut(J), hmoved(J), hpaths(3 * size_t(J));
  std::vector<int32_t> hrounds(J), hlast(J);
  std::vector<unsigned long long> hwork(2 * kGainSlots);
  std::vector<long long> hab(J);
  std::vector<unsigned long long> hsel(gkd::kSelectByteSlots);
  GK_TRY(download({{hwork.data(), gwork.p, sizeof(unsigned long long) * hwork.size()},
                   {hsel.data(), selbytes.p, sizeof(unsigned long long) * hsel.size()},
                   {hab.data(), dab.p, sizeof(long long) * J},
                   {hpaths.data(), dpaths.p, sizeof(long long) * 3 * J},
                   {hcut.data(), dcut.p, sizeof(long long) * J},
                   {hpw.data(), dpw.p, sizeof(int32_t) * hpw.size()},
                   {hrounds.data(), drounds.p, sizeof(int32_t) * J},
                   {hmoved.data(), dmoved.p, sizeof(long long) * J},
                   {hlast.data(), lastcnt.p, sizeof(int32_t) * J}}));
  if (ctx_->cur) {
    unsigned long long wv = 0, we = 0;
    for (int i = 0; i < kGainSlots; ++i) {
      wv += hwork[2 * i];
      we += hwork[2 * i + 1];
    }
    ctx_->cur->stage_bytes[GK_ST_GAIN] += int64_t(28 * wv + 12 * we);
    for (int32_t i = 0; i < J; ++i) ctx_->cur->stage_bytes[GK_ST_BALANCE] += hab[i];
    for (unsigned long long b : hsel) ctx_->cur->stage_bytes[GK_ST_SELECT] += int64_t(b);
  }
  std::vector<UnpackDst> dst(J);
  std::vector<int32_t> done(J, 0);
  bool any_done = false;
  for (int32_t i = 0; i < J; ++i) {
    Job& j = *ids[i];
    gk_level_stats& ls = j.S.levels[j.cur];
    ls.rounds += hrounds[i];
    ls.moves += hmoved[i];
    for (int q = 0; q < 3; ++q) j.S.round_paths[q] += hpaths[3 * size_t(i) + q];
    ls.cut_after = hcut[i];
    j.cut = hcut[i];
    std::copy(hpw.begin() + pw_off[i], hpw.begin() + pw_off[i + 1], j.hpw.begin());
    std::memset(&dst[i], 0, sizeof dst[i]);
    if (hlast[i] > 0) {
      // Still moving after this step's rounds: the level resumes next step.
      j.resume = true;
      j.cpart = part;
      j.cbnd = bnd;
      j.cgain = gainp;
      j.cdest = destp;
      j.coff = F.v_off[i];
      continue;
    }
    j.resume = false;
    j.cgain.reset();
    j.cdest.reset();
    if (j.cur == 0) {
      dst[i].part = j.out;
      done[i] = 1;
      any_done = true;
    } else {
      j.cpart = part;
      j.cbnd = bnd;
      j.coff = F.v_off[i];
      j.L.pop_back();
      j.cur -= 1;
    }
  }
  if (any_done) {
    // Chunks of at most kCutChunk vertices of each finished job.
    constexpr int32_t kCutChunk = 2048;
    std::vector<int32_t> cj, cs, ce;
    for (int32_t i = 0; i < J; ++i) {
      if (!done[i]) continue;
      for (int32_t a = F.v_off[i]; a < F.v_off[i + 1]; a += kCutChunk) {
        cj.push_back(i);
        cs.push_back(a);
        ce.push_back(std::min(F.v_off[i + 1], a + kCutChunk));
      }
    }
    Buf<UnpackDst> ddst;
    Buf<int32_t> dcj, dcs, dce;
    Buf<unsigned long long> c64, pw64;
    GK_TRY(upload(ddst, dst.data(), dst.size()));
    GK_TRY(upload(dcj, cj.data(), cj.size()));
    GK_TRY(upload(dcs, cs.data(), cs.size()));
    GK_TRY(upload(dce, ce.data(), ce.size()));
    GK_TRY(c64.alloc(P_, J));
    GK_TRY(pw64.alloc(P_, pw_off[J]));
    gkd::fill<unsigned long long>(c64.p, 0ull, J, st_);
    gkd::fill<unsigned long long>(pw64.p, 0ull, pw_off[J], st_);
    GK_TRY(gkd::launch_unpack_part(J, ddst.p, F.dv_off.p, N, part->p, nullptr, st_));
    int64_t cut_bytes = 0;
    for (size_t c = 0; c < cj.size(); ++c)
      cut_bytes += 16LL * (ce[c] - cs[c]) + 12LL * (F.e_off[cj[c] + 1] - F.e_off[cj[c]]) * (ce[c] - cs[c]) /
                                               std::max<int64_t>(1, F.v_off[cj[c] + 1] - F.v_off[cj[c]]);
    GK_STAGE(ctx_, GK_ST_CUT, cut_bytes,
             gkd::launch_cut_chunks(static_cast<int32_t>(cj.size()), dcj.p, dcs.p, dce.p, kmax, dkk.p, g, part->p,
                                    dpw_off.p, c64.p, pw64.p, st_));
    std::vector<unsigned long long> hc64(J), hpw64(pw_off[J]);
    GK_TRY(download({{hc64.data(), c64.p, sizeof(unsigned long long) * J},
                     {hpw64.data(), pw64.p, sizeof(unsigned long long) * hpw64.size()}}));
    for (int32_t i = 0; i < J; ++i) {
      if (!done[i]) continue;
      Job& j = *ids[i];
      std::vector<int64_t> pw(hpw64.begin() + pw_off[i], hpw64.begin() + pw_off[i + 1]);
      finish_stats(j, static_cast<int64_t>(hc64[i]), j.cut, pw);
      j.L.clear();
      j.cpart.reset();
      j.cbnd.reset();
      mark(j, DONE);
      GK_TRY(on_done(j));
    }
  }
  return GK_OK;
}

int Engine::on_done(Job& j) {
  if (done_cb_) return done_cb_(j);
  return GK_OK;
}

int Engine::step(bool* worked) {
  *worked = false;
  double t = now_ms();
  auto lap = [&](int i) {
    const double u = now_ms();
    prof_ms_[i] += u - t;
    t = u;
  };
  GK_TRY(coarsen_step(worked));
  lap(0);
  GK_TRY(init_step(worked));
  lap(1);
  GK_TRY(collect_inits(false, worked));
  lap(2);
  GK_TRY(refine_step(worked));
  lap(3);
  if (!*worked) GK_TRY(collect_inits(true, worked));
  lap(4);
  ++steps_;
  return GK_OK;
}

int Engine::add_job(std::unique_ptr<Job> j) {
  j->t_start = now_ms();
  j->S.levels[0].n = j->L.empty() ? 0 : j->L[0].n;
  j->S.levels[0].nnz = j->L.empty() ? 0 : j->L[0].nnz;
  j->S.total_vwgt = j->T;
  j->S.cap = j->cap;
  if (j->L.empty() || j->L[0].n == 0) {
    j->phase = DONE;
    j->t_done = j->t_start;
    ++ndone_;
  }
  jobs_.push_back(std::move(j));
  return GK_OK;
}

void Engine::drop_done() {
  std::vector<std::unique_ptr<Job>> keep;
  for (auto& j : jobs_)
    if (j->phase != DONE) keep.push_back(std::move(j));
  ndone_ = 0;
  jobs_ = std::move(keep);
}

int Engine::job_setup(Job& j, const gk_params& prm, int64_t T, int32_t n0) {
  j.prm = prm;
  if (T >= (int64_t(1) << 31)) {
    gkh::set_error("total vertex weight >= 2^31 (graph " + std::to_string(j.id) + ")");
    return GK_ERR_LIMIT;
  }
  j.T = T;
  j.cap = gk_part_cap(T, prm.k, prm.ubfactor);
  j.floor_n = gk_coarsen_floor(n0, prm.k, prm.coarsen_policy, prm.coarsen_to);
  j.maxlv = prm.max_levels > 0 ? std::min(prm.max_levels, GK_MAX_LEVELS) : GK_MAX_LEVELS;
  return GK_OK;
}

int check_batch_params(const gk_params* p, int32_t count) {
  for (int32_t i = 0; i < count; ++i) {
    const gk_params& q = p[i];
    if (q.k < 1 || q.k > kMaxK) { gkh::set_error("k must be in [1, 8192]"); return GK_ERR_ARG; }
    if (q.ubfactor < 0) { gkh::set_error("ubfactor < 0"); return GK_ERR_ARG; }
    if (q.max_group < 1 || q.max_group > 64) { gkh::set_error("max_group must be in [1, 64]"); return GK_ERR_ARG; }
    if (q.max_moves < 1 || q.max_moves > 1024) { gkh::set_error("max_moves must be in [1, 1024]"); return GK_ERR_ARG; }
    if (q.max_group != p[0].max_group || q.max_moves != p[0].max_moves) {
      gkh::set_error("max_group and max_moves must be equal for all graphs of a batch");
      return GK_ERR_ARG;
    }
  }
  return GK_OK;
}

int batch_workers(const gk_params& p) {
  if (p.initpart_threads > 0) return p.initpart_threads;
  const unsigned hw = std::thread::hardware_concurrency();
  return static_cast<int>(hw ? std::min(hw, 16u) : 1u);
}

}  // namespace gkrt

using namespace gkrt;

namespace {

struct DeviceGuard {
  explicit DeviceGuard(int d) { (void)hipSetDevice(d); }
};

int run_packed(gk_ctx* ctx, hipStream_t st, const gk_packed_csr* pk, const gk_params* prm, int32_t* dpart,
               gk_stats* stats) {
  const int32_t G = pk->ngraphs;
  Pool& P = ctx->pool;
  gk_stats SB;
  std::memset(&SB, 0, sizeof SB);
  ctx->cur = &SB;
  ctx->clock.start((prm[0].flags & GK_FLAG_STAGE_TIME) != 0, st);
  struct Reset {
    gk_ctx* c;
    ~Reset() {
      c->cur = nullptr;
      c->clock.start(false, nullptr);
      gkrt::clear_launch_options(c);
    }
  } reset{ctx};
  GK_TRY(gkrt::set_launch_options(ctx, prm[0].flags));
  // Level 0 of every graph is a slice of the caller's pack.
  auto S0 = std::make_shared<Store>();
  S0->nj = G;
  S0->n = pk->n;
  S0->nnz = pk->nnz;
  S0->v_off.assign(pk->v_off, pk->v_off + G + 1);
  S0->e_off.assign(pk->e_off, pk->e_off + G + 1);
  S0->xadj = pk->xadj;
  S0->adj = pk->adjncy;
  S0->w = pk->adjwgt;
  S0->vw = pk->vwgt;
  if (!S0->w) {
    GK_TRY(S0->bw.alloc(P, pk->nnz));
    gkd::fill<int32_t>(S0->bw.p, 1, pk->nnz, st);
    S0->w = S0->bw.p;
  }
  if (!S0->vw) {
    GK_TRY(S0->bv.alloc(P, pk->n));
    gkd::fill<int32_t>(S0->bv.p, 1, pk->n, st);
    S0->vw = S0->bv.p;
  }
  Engine E(ctx, st, batch_workers(prm[0]));
  E.set_graph_rounds((prm[0].flags & GK_FLAG_GRAPH_ROUNDS) != 0);
  GK_TRY(E.upload(S0->dv_off, S0->v_off.data(), S0->v_off.size()));
  std::vector<long long> T(G);
  {
    Buf<long long> dT;
    GK_TRY(dT.alloc(P, G));
    GK_TRY(gkd::launch_job_sum(G, S0->dv_off.p, S0->vw, dT.p, st));
    GK_TRY(E.download({{T.data(), dT.p, sizeof(long long) * G}}));
  }
  for (int32_t g = 0; g < G; ++g) {
    auto j = std::make_unique<Job>();
    j->id = g;
    const int32_t n = S0->v_off[g + 1] - S0->v_off[g];
    GK_TRY(E.job_setup(*j, prm[g], T[g], n));
    j->out = dpart + S0->v_off[g];
    if (n > 0) {
      JLevel l0;
      l0.n = n;
      l0.nnz = S0->e_off[g + 1] - S0->e_off[g];
      l0.store = S0;
      l0.slot = g;
      j->L.push_back(std::move(l0));
    }
    GK_TRY(E.add_job(std::move(j)));
  }
  S0.reset();
  while (!E.all_done()) {
    bool worked = false;
    GK_TRY(E.step(&worked));
  }
  GK_TRY(E.sync());
  GK_TRY(ctx->clock.collect(SB));
  for (int i = 0; i < GK_ST_COUNT; ++i) SB.alg_bytes += SB.stage_bytes[i];
  int rc = GK_OK;
  for (int32_t g = 0; g < G; ++g) {
    Job& j = E.job(g);
    if (j.rc != GK_OK && rc == GK_OK) {
      rc = j.rc;
      gkh::set_error("graph " + std::to_string(g) + ": " + j.err);
    }
    if (stats) {
      gk_stats& s = stats[g];
      s = j.S;
      if (j.L.empty() && j.phase == DONE && j.S.nlevels == 0) std::memset(&s, 0, sizeof s);
      std::memcpy(s.stage_ms, SB.stage_ms, sizeof s.stage_ms);
      std::memcpy(s.stage_launches, SB.stage_launches, sizeof s.stage_launches);
      std::memcpy(s.stage_bytes, SB.stage_bytes, sizeof s.stage_bytes);
      s.alg_bytes = SB.alg_bytes;
    }
  }
  return rc;
}

}  // namespace

extern "C" {

int gk_partition_packed(gk_ctx* ctx, const gk_packed_csr* pk, const gk_params* prm, int32_t* dpart,
                        gk_stats* stats, void* stream) {
  if (!ctx || !pk || !prm || (!dpart && pk->n > 0)) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  if (pk->ngraphs < 0 || pk->n < 0 || pk->nnz < 0) { gkh::set_error("negative size"); return GK_ERR_ARG; }
  if (pk->ngraphs == 0) return GK_OK;
  if (!pk->v_off || !pk->e_off || (pk->n > 0 && !pk->xadj) || (pk->nnz > 0 && !pk->adjncy)) {
    gkh::set_error("bad packed graph description");
    return GK_ERR_ARG;
  }
  const int32_t G = pk->ngraphs;
  if (pk->v_off[0] != 0 || pk->e_off[0] != 0 || pk->v_off[G] != pk->n || pk->e_off[G] != pk->nnz) {
    gkh::set_error("v_off / e_off must start at 0 and end at n / nnz");
    return GK_ERR_ARG;
  }
  for (int32_t g = 0; g < G; ++g)
    if (pk->v_off[g + 1] < pk->v_off[g] || pk->e_off[g + 1] < pk->e_off[g]) {
      gkh::set_error("v_off / e_off must be nondecreasing");
      return GK_ERR_ARG;
    }
  GK_TRY(check_batch_params(prm, G));
  DeviceGuard guard(ctx->device);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return run_packed(ctx, st, pk, prm, dpart, stats);
}

int gk_partition_batch(gk_ctx* ctx, int32_t count, const gk_csr* graphs, const gk_params* prm,
                       int32_t* const* parts_out, gk_stats* stats) {
  if (!ctx || count < 0 || (count > 0 && (!graphs || !prm || !parts_out))) {
    gkh::set_error("null argument");
    return GK_ERR_ARG;
  }
  if (count == 0) return GK_OK;
  GK_TRY(check_batch_params(prm, count));
  std::vector<int32_t> v_off(count + 1, 0), e_off(count + 1, 0);
  bool any_w = false, any_v = false;
  for (int32_t g = 0; g < count; ++g) {
    GK_TRY(gk_graph_check(&graphs[g], (prm[g].flags & GK_FLAG_CHECK) != 0));
    const int64_t a = int64_t(v_off[g]) + graphs[g].n, b = int64_t(e_off[g]) + graphs[g].nnz;
    if (a >= INT32_MAX || b >= INT32_MAX) { gkh::set_error("batch exceeds 2^31 vertices or entries"); return GK_ERR_LIMIT; }
    v_off[g + 1] = int32_t(a);
    e_off[g + 1] = int32_t(b);
    any_w |= graphs[g].adjwgt != nullptr;
    any_v |= graphs[g].vwgt != nullptr;
    if (graphs[g].n > 0 && !parts_out[g]) { gkh::set_error("null partition output"); return GK_ERR_ARG; }
  }
  const int32_t N = v_off[count], E = e_off[count];
  // Host pack with global ids, then one copy per array.
  std::vector<int32_t> x(int64_t(N) + 1), a(E), w(any_w ? E : 0), v(any_v ? N : 0);
  for (int32_t g = 0; g < count; ++g) {
    const gk_csr& c = graphs[g];
    for (int32_t i = 0; i < c.n; ++i) x[v_off[g] + i] = e_off[g] + c.xadj[i];
    for (int32_t e = 0; e < c.nnz; ++e) a[e_off[g] + e] = v_off[g] + c.adjncy[e];
    if (any_w)
      for (int32_t e = 0; e < c.nnz; ++e) w[e_off[g] + e] = c.adjwgt ? c.adjwgt[e] : 1;
    if (any_v)
      for (int32_t i = 0; i < c.n; ++i) v[v_off[g] + i] = c.vwgt ? c.vwgt[i] : 1;
  }
  x[N] = E;
  DeviceGuard guard(ctx->device);
  hipStream_t st = ctx->stream;
  Buf<int32_t> dx, da, dw, dv, dp;
  GK_TRY(dx.alloc(ctx->pool, int64_t(N) + 1));
  GK_TRY(da.alloc(ctx->pool, E));
  GK_TRY(dp.alloc(ctx->pool, N));
  GK_HIP(hipMemcpyAsync(dx.p, x.data(), sizeof(int32_t) * (size_t(N) + 1), hipMemcpyHostToDevice, st));
  if (E) GK_HIP(hipMemcpyAsync(da.p, a.data(), sizeof(int32_t) * size_t(E), hipMemcpyHostToDevice, st));
  if (any_w) {
    GK_TRY(dw.alloc(ctx->pool, E));
    if (E) GK_HIP(hipMemcpyAsync(dw.p, w.data(), sizeof(int32_t) * size_t(E), hipMemcpyHostToDevice, st));
  }
  if (any_v) {
    GK_TRY(dv.alloc(ctx->pool, N));
    if (N) GK_HIP(hipMemcpyAsync(dv.p, v.data(), sizeof(int32_t) * size_t(N), hipMemcpyHostToDevice, st));
  }
  GK_HIP(hipStreamSynchronize(st));
  gk_packed_csr pk;
  pk.ngraphs = count;
  pk.n = N;
  pk.nnz = E;
  pk.xadj = dx.p;
  pk.adjncy = da.p;
  pk.adjwgt = any_w ? dw.p : nullptr;
  pk.vwgt = any_v ? dv.p : nullptr;
  pk.v_off = v_off.data();
  pk.e_off = e_off.data();
  GK_TRY(run_packed(ctx, st, &pk, prm, dp.p, stats));
  std::vector<int32_t> hp(N);
  if (N) GK_HIP(hipMemcpyAsync(hp.data(), dp.p, sizeof(int32_t) * size_t(N), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  for (int32_t g = 0; g < count; ++g)
    if (graphs[g].n > 0) std::copy(hp.begin() + v_off[g], hp.begin() + v_off[g + 1], parts_out[g]);
  return GK_OK;
}

}  // extern "C"
