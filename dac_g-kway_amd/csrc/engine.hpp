// Step machine shared by the batched form (gk_partition_packed /
// gk_partition_batch) and the serving loop: jobs, their level hierarchies as
// slices of shared packed stores, and the per-step coarsen / initial
// partition / refine passes over all jobs in flight (engine.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <functional>
#include <future>
#include <memory>
#include <atomic>
#include <mutex>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gkway.h"
#include "batch.hpp"
#include "runtime.hpp"

namespace gkrt {

// Host worker threads for the initial partitions.
class WorkerPool {
 public:
  explicit WorkerPool(int n);
  ~WorkerPool();
  std::future<int> submit(std::function<int()> f);

 private:
  std::vector<std::thread> th_;
  std::queue<std::function<void()>> q_;
  std::mutex m_;
  std::condition_variable cv_;
  bool stop_ = false;
};

// A packed CSR of several jobs' graphs with global ids: job slot i owns
// vertices [v_off[i], v_off[i+1]) and entries [e_off[i], e_off[i+1]).
struct Store {
  int32_t nj = 0, n = 0, nnz = 0;
  std::vector<int32_t> v_off, e_off;
  Buf<int32_t> dv_off;  // v_off on the device
  const int32_t *xadj = nullptr, *adj = nullptr, *w = nullptr, *vw = nullptr;
  Buf<int32_t> bx, ba, bw, bv;  // owned arrays (the caller's pack owns none)
  gkd::DGraph graph() const {
    gkd::DGraph g;
    g.n = n;
    g.nnz = nnz;
    g.xadj = xadj;
    g.adj = adj;
    g.w = w;
    g.vw = vw;
    return g;
  }
};
using StoreP = std::shared_ptr<Store>;
using IBufP = std::shared_ptr<Buf<int32_t>>;

// One level of a job: slot `slot` of `store`; cmap (once the next level
// exists) maps its vertices to the next level's: local coarse id =
// cmap->p[cmap_off + v] - cbase.
struct JLevel {
  int32_t n = 0, nnz = 0;
  StoreP store;
  int32_t slot = 0;
  IBufP cmap;
  int32_t cmap_off = 0, cbase = 0;
};

enum Phase { COARSEN = 0, INIT_READY = 1, INIT_RUNNING = 2, REFINE = 3, DONE = 4 };

struct Job {
  int32_t id = 0;
  gk_params prm{};
  int64_t T = 0, cap = 0;
  int32_t floor_n = 0, maxlv = 0;
  std::vector<JLevel> L;
  int32_t nl = 0;    // levels kept (set when coarsening ends)
  int phase = COARSEN;
  int32_t cur = -1;  // level refined by the next refinement step
  IBufP cpart, cbnd; // partition / boundary of level cur + 1 at offset coff, or,
  int32_t coff = 0;  // with resume, of level cur itself (a level not converged yet)
  bool resume = false;
  IBufP cgain, cdest;  // with resume: gains and destinations of level cur
  std::vector<int32_t> hpw;  // part weights
  long long cut = 0;         // tracked cut
  // initial partition (host worker)
  std::vector<int32_t> hx, ha, hw, hv, hpart, hbnd;
  int64_t hcut = 0;
  std::future<int> fut;
  std::string err;
  int rc = GK_OK;
  gk_stats S{};
  int32_t* out = nullptr;  // device partition output (local part ids)
  double t_start = 0, t_done = 0;
  void* user = nullptr;    // owner's tag (serving)
  void release_host() {
    std::vector<int32_t>().swap(hx);
    std::vector<int32_t>().swap(ha);
    std::vector<int32_t>().swap(hw);
    std::vector<int32_t>().swap(hv);
    std::vector<int32_t>().swap(hpart);
    std::vector<int32_t>().swap(hbnd);
  }
};

// Refinement pack extras: vertex -> job slot, packed projection map and the
// packed coarse partition / boundary.
struct RefineSide {
  Buf<int32_t> vjob, cmap, cpart, cbnd, cgain, cdest;
  std::vector<int32_t> c_off;
};

class Engine {
 public:
  Engine(gk_ctx* ctx, hipStream_t st, int workers);
  ~Engine();
  int job_setup(Job& j, const gk_params& prm, int64_t T, int32_t n0);
  int add_job(std::unique_ptr<Job> j);
  // One step over all jobs in flight; *worked = false when nothing could
  // advance (then the step has waited for an initial partition, if any).
  int step(bool* worked);
  bool all_done() const { return ndone_ == jobs_.size(); }
  size_t live() const { return jobs_.size() - ndone_; }
  Job& job(size_t i) { return *jobs_[i]; }
  size_t njobs() const { return jobs_.size(); }
  void drop_done();
  int sync();
  double now_ms() const;
  // Called for each job right after it is finished and its partition written.
  std::function<int(Job&)> done_cb_;

  template <class T>
  int upload(Buf<T>& dst, const T* src, size_t n) {
    GK_TRY(dst.alloc(P_, static_cast<int64_t>(n)));
    if (n == 0) return GK_OK;
    void* h = nullptr;
    GK_TRY(up_.get(st_, sizeof(T) * n, &h));
    std::memcpy(h, src, sizeof(T) * n);
    GK_HIP(hipMemcpyAsync(dst.p, h, sizeof(T) * n, hipMemcpyHostToDevice, st_));
    return GK_OK;
  }
  int download(std::initializer_list<Fetch> items);
  // Refinement rounds replayed from a captured hipGraph (GK_FLAG_GRAPH_ROUNDS).
  void set_graph_rounds(bool on) { graph_rounds_ = on; }

 private:
  static StoreP whole_store(const std::vector<const JLevel*>& lv);
  static gkd::PackSrc src_of(const JLevel& lv);
  int pack_copy(const std::vector<gkd::PackSrc>& src, const std::vector<int32_t>& nv,
                const std::vector<int32_t>& ne, const std::vector<int32_t>* ncv, StoreP& out, RefineSide* rs);
  void mark(Job& j, int phase);
  void finish_stats(Job& j, int64_t cut, int64_t tracked, const std::vector<int64_t>& pw);
  int coarsen_step(bool* worked);
  int init_step(bool* worked);
  int collect_inits(bool block, bool* worked);
  int refine_step(bool* worked);
  int on_done(Job& j);

  gk_ctx* ctx_;
  hipStream_t st_;
  Pool& P_;
  std::chrono::steady_clock::time_point t0_;
  Pinned& up_;  // the context's staging arenas
  Pinned& dl_;
  size_t ndone_ = 0;
  bool graph_rounds_ = false;
  int64_t steps_ = 0, rounds_ = 0;
  double prof_ms_[6] = {0, 0, 0, 0, 0, 0};  // host wall time per step phase
  std::atomic<long long> init_cpu_us_{0};   // worker CPU time in initial partitions
  std::vector<std::unique_ptr<Job>> jobs_;
  WorkerPool pool_;  // declared last: joined before the jobs it works on go away
};

int check_batch_params(const gk_params* p, int32_t count);
int batch_workers(const gk_params& p);

}  // namespace gkrt
