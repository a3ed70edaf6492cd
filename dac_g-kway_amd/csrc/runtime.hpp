// Host runtime shared by the plain driver (driver.hip) and the batch/serve
// engine (engine.hip): the caching device pool, buffers, per-stage HIP-event
// timing, the context, and small read-back helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gkway.h"
#include "dev_common.hpp"
#include "kernels.hpp"

namespace gkrt {

constexpr int32_t kMaxK = 8192;
constexpr int32_t kMaxJumpRounds = 40;

// ---------------------------------------------------------------- pool
// Caching device allocator: blocks return to a size-keyed free list and are
// reused by later requests of up to twice smaller size, so steady-state
// partitioning calls perform no hipMalloc.
class Pool {
 public:
  ~Pool() { release_all(); }
  int get(size_t bytes, void** out) {
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    if (limit_ > 0 && in_use_ + static_cast<int64_t>(bytes) > limit_) {
      gkh::set_error("device pool limit of " + std::to_string(limit_) + " bytes reached (" +
                     std::to_string(in_use_) + " in use, " + std::to_string(bytes) + " requested)");
      return GK_ERR_NOMEM;
    }
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= 2 * bytes && (limit_ <= 0 || in_use_ + int64_t(it->first) <= limit_)) {
      *out = it->second;
      note_out(it->first);
      free_.erase(it);
      return GK_OK;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      trim();
      if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        gkh::set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return GK_ERR_NOMEM;
      }
    }
    size_[p] = bytes;
    *out = p;
    note_out(bytes);
    return GK_OK;
  }
  void put(void* p) {
    if (!p) return;
    auto it = size_.find(p);
    if (it != size_.end()) {
      free_.emplace(it->second, p);
      in_use_ -= static_cast<int64_t>(it->second);
    }
  }
  int64_t in_use() const { return in_use_; }
  int64_t peak() const { return peak_; }
  // Bytes the pool may hand out at once (0 = no limit); a request beyond it
  // fails with GK_ERR_NOMEM.
  void set_limit(int64_t bytes) { limit_ = bytes > 0 ? bytes : 0; }
  void reset_peak() { peak_ = in_use_; }
  void trim() {
    for (auto& kv : free_) {
      (void)hipFree(kv.second);
      size_.erase(kv.second);
    }
    free_.clear();
  }
  void release_all() {
    for (auto& kv : size_) (void)hipFree(kv.first);
    size_.clear();
    free_.clear();
    in_use_ = 0;
  }

 private:
  void note_out(size_t b) {
    in_use_ += static_cast<int64_t>(b);
    if (in_use_ > peak_) peak_ = in_use_;
  }
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> size_;
  int64_t in_use_ = 0, peak_ = 0;  // bytes handed out now / at most
  int64_t limit_ = 0;
};

template <class T>
struct Buf {
  Pool* pool = nullptr;
  T* p = nullptr;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  Buf(Buf&& o) noexcept : pool(o.pool), p(o.p) { o.p = nullptr; }
  Buf& operator=(Buf&& o) noexcept {
    if (this != &o) { release(); pool = o.pool; p = o.p; o.p = nullptr; }
    return *this;
  }
  ~Buf() { release(); }
  int alloc(Pool& pl, int64_t n) {
    release();
    pool = &pl;
    void* q = nullptr;
    int rc = pl.get(sizeof(T) * static_cast<size_t>(n > 0 ? n : 1), &q);
    if (rc) return rc;
    p = static_cast<T*>(q);
    return GK_OK;
  }
  void release() {
    if (pool && p) pool->put(p);
    p = nullptr;
  }
};

inline int bits_for(int64_t x) {  // bits needed to hold values 0..x
  int b = 0;
  while (b < 63 && (int64_t(1) << b) <= x) ++b;
  return b < 1 ? 1 : b;
}

inline uint64_t fnv_host(const std::vector<int32_t>& a, uint64_t h = 0xcbf29ce484222325ull) {
  for (int32_t x : a) { h ^= uint32_t(x); h *= 0x100000001b3ull; }
  return h;
}

}  // namespace gkrt

// Per-stage device timing: each timed call is bracketed by two HIP events
// recorded on the work stream; elapsed times are summed at the end of a run.
namespace gkrt {
class StageClock {
 public:
  ~StageClock() {
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  }
  void start(bool on, hipStream_t st) {
    on_ = on;
    st_ = st;
    kclock() = nullptr;
    for (auto& v : recs_) v.clear();
    used_ = 0;
  }
  void set_mask(uint32_t m) { mask_ = m; }
  bool timed(int id) const { return on_ && ((mask_ >> id) & 1u); }
  // Stages whose launches go through gkd::gk_launch are timed kernel by
  // kernel (their time = the sum of their kernels' durations); the others
  // are bracketed by a pair of events on the stream.
  static bool per_kernel(int id) {
    return id == GK_ST_MATCH || id == GK_ST_CONTRACT || id == GK_ST_GAIN || id == GK_ST_GAIN_HEAVY ||
           id == GK_ST_SELECT || id == GK_ST_BALANCE;
  }
  void begin(int id) {
    if (!on_ || !((mask_ >> id) & 1u)) return;
    if (per_kernel(id)) {
      kclock() = this;
      kstage() = id;
      return;
    }
    cur_[id] = take();
    (void)hipEventRecord(cur_[id], st_);
  }
  void end(int id) {
    if (!on_ || !((mask_ >> id) & 1u)) return;
    if (per_kernel(id)) {
      kclock() = nullptr;
      return;
    }
    hipEvent_t e = take();
    (void)hipEventRecord(e, st_);
    recs_[id].emplace_back(cur_[id], e);
  }
  // An event pair for one launch of the stage being timed kernel by kernel
  // on this thread; false when there is none.
  static bool kernel_pair(hipEvent_t* a, hipEvent_t* b) {
    StageClock* c = kclock();
    if (!c) return false;
    *a = c->take();
    *b = c->take();
    if (!*a || !*b) return false;
    c->recs_[kstage()].emplace_back(*a, *b);
    return true;
  }
  int collect(gk_stats& S) {
    if (!on_) return GK_OK;
    GK_HIP(hipStreamSynchronize(st_));
    for (int id = 0; id < GK_ST_COUNT; ++id) {
      double ms = 0;
      for (auto& pr : recs_[id]) {
        float t = 0;
        GK_HIP(hipEventElapsedTime(&t, pr.first, pr.second));
        ms += t;
      }
      S.stage_ms[id] = ms;
    }
    return GK_OK;
  }

 private:
  hipEvent_t take() {
    if (used_ == pool_.size()) {
      hipEvent_t e;
      // Timing-only events: no system-scope fence (cache writeback and
      // invalidation) at each record, which would stretch the short
      // launches they bracket; the times are read after a stream sync.
      if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
      pool_.push_back(e);
    }
    return pool_[used_++];
  }
  static StageClock*& kclock() {
    thread_local StageClock* c = nullptr;
    return c;
  }
  static int& kstage() {
    thread_local int id = 0;
    return id;
  }
  bool on_ = false;
  uint32_t mask_ = ~0u;
  hipStream_t st_ = nullptr;
  std::vector<hipEvent_t> pool_;
  size_t used_ = 0;
  hipEvent_t cur_[GK_ST_COUNT] = {};
  std::vector<std::pair<hipEvent_t, hipEvent_t>> recs_[GK_ST_COUNT];
};
}  // namespace gkrt

namespace gkrt {
// Pinned host staging handed out by bumping a pointer; drained() after a
// stream synchronisation makes the whole arena free again.  When a request
// does not fit, the stream is synchronised and the arena starts over.
class Pinned {
 public:
  Pinned() = default;
  Pinned(const Pinned&) = delete;
  Pinned& operator=(const Pinned&) = delete;
  ~Pinned();
  int get(hipStream_t st, size_t bytes, void** out);
  void drained() { used_ = 0; }

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
};

}  // namespace gkrt

namespace gkrt {
// Launch shapes that do not change results (lanes per vertex or row, block
// size), chosen by timing.  A key names a kernel family and a level's degree
// profile: ceil(log2 n) and the average degree in half-octaves.  The first
// launch with a new key warms the device, then times every candidate shape on
// the real data, twice (each run writes the same output), and the fastest is
// kept for the life of the context.
class Tuner {
 public:
  struct Shape {
    int lanes;
    int block;
  };
  Tuner() = default;
  Tuner(const Tuner&) = delete;
  Tuner& operator=(const Tuner&) = delete;
  ~Tuner() {
    for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
  }
  static uint64_t key(int family, int64_t n, int64_t nnz) {
    int ln = 0;
    while (ln < 62 && (int64_t(1) << ln) < n) ++ln;
    const double avg = n > 0 ? double(nnz) / double(n) : 0.0;
    const int la = avg > 1.0 ? static_cast<int>(std::floor(2.0 * std::log2(avg))) : 0;
    return (uint64_t(family) << 32) | (uint64_t(ln) << 16) | uint64_t(la);
  }
  const Shape* find(uint64_t key) const {
    auto it = best_.find(key);
    return it == best_.end() ? nullptr : &it->second;
  }
  int tunings() const { return tunings_; }
  // launch(shape) enqueues one run on st and returns a status.  Ties go to
  // the earlier candidate.
  template <class F>
  int tune(uint64_t key, const std::vector<Shape>& cands, hipStream_t st, F&& launch, Shape* out) {
    const size_t m = cands.size();
    if (m == 0) {
      gkh::set_error("autotuner: no candidate launch shape");
      return GK_ERR_INTERNAL;
    }
    while (ev_.size() < 4 * m) {
      hipEvent_t e;
      GK_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
      ev_.push_back(e);
    }
    // Warm the device first (clock ramp after idle, caches, address
    // translation): the first candidate runs until two consecutive runs agree
    // within 5%, at most 16 times.
    float prev = -1.0f;
    for (int r = 0; r < 16; ++r) {
      GK_HIP(hipEventRecord(ev_[0], st));
      GK_TRY(launch(cands[0]));
      GK_HIP(hipEventRecord(ev_[1], st));
      GK_HIP(hipEventSynchronize(ev_[1]));
      float t = 0;
      GK_HIP(hipEventElapsedTime(&t, ev_[0], ev_[1]));
      if (prev > 0 && std::fabs(t - prev) <= 0.05f * prev) break;
      prev = t;
    }
    for (int pass = 0; pass < 2; ++pass)
      for (size_t i = 0; i < m; ++i) {
        GK_HIP(hipEventRecord(ev_[4 * i + 2 * pass], st));
        GK_TRY(launch(cands[i]));
        GK_HIP(hipEventRecord(ev_[4 * i + 2 * pass + 1], st));
      }
    GK_HIP(hipEventSynchronize(ev_[4 * (m - 1) + 3]));
    size_t bi = 0;
    float bt = 0;
    for (size_t i = 0; i < m; ++i) {
      float t0 = 0, t1 = 0;
      GK_HIP(hipEventElapsedTime(&t0, ev_[4 * i], ev_[4 * i + 1]));
      GK_HIP(hipEventElapsedTime(&t1, ev_[4 * i + 2], ev_[4 * i + 3]));
      const float t = std::min(t0, t1);
      if (i == 0 || t < bt) {
        bi = i;
        bt = t;
      }
    }
    // GK_TUNE_LOG=1: one line per tuning with every candidate's time
    // (diagnostics only).
    const char* lg = std::getenv("GK_TUNE_LOG");
    if (lg && lg[0] == '1') {
      std::fprintf(stderr, "[gk tune] family %u log2n %u avg2 %u:", unsigned(key >> 32), unsigned((key >> 16) & 0xffff),
                   unsigned(key & 0xffff));
      for (size_t i = 0; i < m; ++i) {
        float t0 = 0, t1 = 0;
        (void)hipEventElapsedTime(&t0, ev_[4 * i], ev_[4 * i + 1]);
        (void)hipEventElapsedTime(&t1, ev_[4 * i + 2], ev_[4 * i + 3]);
        std::fprintf(stderr, " %dx%d=%.1fus", cands[i].lanes, cands[i].block, 1e3 * std::min(t0, t1));
      }
      std::fprintf(stderr, " -> %dx%d\n", cands[bi].lanes, cands[bi].block);
    }
    best_[key] = cands[bi];
    ++tunings_;
    *out = cands[bi];
    return GK_OK;
  }

 private:
  std::unordered_map<uint64_t, Shape> best_;
  std::vector<hipEvent_t> ev_;
  int tunings_ = 0;
};
}  // namespace gkrt

struct gk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  gkrt::Pool pool;
  char* hpin = nullptr;  // pinned staging for small read-backs
  gkrt::StageClock clock;
  gk_stats* cur = nullptr;  // stats of the running partition (stage counters)
  gkrt::Pinned up, dl;      // pinned staging of the batched form (uploads / read-backs)
  gkrt::Tuner tuner;        // launch shapes chosen by timing (GK_FLAG_AUTOTUNE)
  bool autotune = false;    // set per call from GK_FLAG_AUTOTUNE
  bool persistent = false;  // set per call from GK_FLAG_PERSISTENT
  gkrt::Buf<unsigned long long> steal_ctr;  // queue counters of the work-stealing launches
  gkd::Steal steal;              // steal_ctr + the device's multiprocessor count
};

namespace gkrt {

// Per-call launch options from the flags.  GK_FLAG_PERSISTENT makes the
// matching, contraction-row and gain launches persistent work-stealing ones
// (the counters are allocated on first use and kept in the ctx).
inline int set_launch_options(gk_ctx* ctx, int32_t flags) {
  ctx->autotune = (flags & GK_FLAG_AUTOTUNE) != 0;
  ctx->persistent = (flags & GK_FLAG_PERSISTENT) != 0;
  if (ctx->persistent && !ctx->steal_ctr.p) {
    GK_TRY(ctx->steal_ctr.alloc(ctx->pool, gkd::kStealWords));
    int cus = 0;
    GK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    ctx->steal.ctr = ctx->steal_ctr.p;
    ctx->steal.cus = cus;
  }
  return GK_OK;
}

inline void clear_launch_options(gk_ctx* ctx) {
  ctx->autotune = false;
  ctx->persistent = false;
}

// The work-stealing handle for a launch, or null for a direct launch.
inline const gkd::Steal* steal_of(const gk_ctx* ctx) { return ctx->persistent ? &ctx->steal : nullptr; }

}  // namespace gkrt

namespace gkrt {

// Stage bracket: times `expr` when stage timing is on, counts the call and
// adds the stage's algorithmic bytes.
#define GK_STAGE(ctx, id, bytes, expr)                \
  do {                                                \
    (ctx)->clock.begin(id);                           \
    int rc_s_ = (expr);                               \
    (ctx)->clock.end(id);                             \
    if (rc_s_) return rc_s_;                          \
    count_stage(ctx, id, (int64_t)(bytes));           \
  } while (0)

inline void count_stage(gk_ctx* ctx, int id, int64_t bytes) {
  if (!ctx->cur) return;
  ctx->cur->stage_launches[id] += 1;
  ctx->cur->stage_bytes[id] += bytes;
}

// A fixed run of stream work -- the rounds of a refinement level between
// two host checks -- captured once into a hipGraph and replayed with one
// launch.  The captured calls do not run at capture; the stage counters that
// they advanced are advanced again at every replay instead.
class RoundGraph {
 public:
  RoundGraph() = default;
  RoundGraph(const RoundGraph&) = delete;
  RoundGraph& operator=(const RoundGraph&) = delete;
  ~RoundGraph() {
    if (!exec_) return;
    // Replays still queued (an error path) finish before the graph goes.
    if (st_) (void)hipStreamSynchronize(st_);
    (void)hipGraphExecDestroy(exec_);
  }
  bool ready() const { return exec_ != nullptr; }
  // body() enqueues the work on st and returns a status.
  template <class F>
  int capture(gk_ctx* ctx, hipStream_t st, F&& body) {
    gk_stats* S = ctx->cur;
    int32_t l0[GK_ST_COUNT] = {};
    int64_t b0[GK_ST_COUNT] = {};
    if (S)
      for (int i = 0; i < GK_ST_COUNT; ++i) {
        l0[i] = S->stage_launches[i];
        b0[i] = S->stage_bytes[i];
      }
    GK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &gr);
    if (S)
      for (int i = 0; i < GK_ST_COUNT; ++i) {
        launches_[i] = S->stage_launches[i] - l0[i];
        bytes_[i] = S->stage_bytes[i] - b0[i];
        S->stage_launches[i] = l0[i];
        S->stage_bytes[i] = b0[i];
      }
    if (rc != GK_OK || ec != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      if (rc != GK_OK) return rc;
      gkh::set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      return GK_ERR_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(&exec_, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ei != hipSuccess) {
      exec_ = nullptr;
      gkh::set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      return GK_ERR_HIP;
    }
    return GK_OK;
  }
  int replay(gk_ctx* ctx, hipStream_t st) {
    st_ = st;
    GK_HIP(hipGraphLaunch(exec_, st));
    if (ctx->cur)
      for (int i = 0; i < GK_ST_COUNT; ++i) {
        ctx->cur->stage_launches[i] += launches_[i];
        ctx->cur->stage_bytes[i] += bytes_[i];
      }
    return GK_OK;
  }

 private:
  hipGraphExec_t exec_ = nullptr;
  hipStream_t st_ = nullptr;  // stream of the last replay
  int32_t launches_[GK_ST_COUNT] = {};
  int64_t bytes_[GK_ST_COUNT] = {};
};

// Copy small device values to host variables through pinned staging, then
// synchronise the stream once.
struct Fetch {
  void* host;
  const void* dev;
  size_t bytes;
};
inline int fetch(gk_ctx* ctx, hipStream_t st, std::initializer_list<Fetch> items) {
  size_t off = 0;
  for (const Fetch& f : items) {
    GK_HIP(hipMemcpyAsync(ctx->hpin + off, f.dev, f.bytes, hipMemcpyDeviceToHost, st));
    off += (f.bytes + 7) & ~size_t(7);
  }
  GK_HIP(hipStreamSynchronize(st));
  off = 0;
  for (const Fetch& f : items) {
    std::memcpy(f.host, ctx->hpin + off, f.bytes);
    off += (f.bytes + 7) & ~size_t(7);
  }
  return GK_OK;
}

template <class T>
inline int to_host(hipStream_t st, std::vector<T>& out, const T* dev, int64_t n) {
  out.resize(static_cast<size_t>(n));
  if (n <= 0) return GK_OK;
  GK_HIP(hipMemcpyAsync(out.data(), dev, sizeof(T) * n, hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}

// Coarsening stages shared by the plain driver and the batch engine
// (defined in driver.hip).
struct Level;
int group_stage(gk_ctx* ctx, hipStream_t st, int32_t n, const int32_t* cand, int32_t* label,
                Buf<int32_t>& dist_out, int32_t* rounds);
int cluster_stage(gk_ctx* ctx, hipStream_t st, int32_t n, const int32_t* label, const int32_t* dist,
                  int32_t max_dist_bits, int32_t G, Buf<int32_t>& order, int32_t* cmap,
                  Buf<int32_t>& cptr, int32_t* nc);
int contract_stage(gk_ctx* ctx, hipStream_t st, const gkd::DGraph& g, int32_t nc, const int32_t* cmap,
                   const int32_t* cptr, const int32_t* order, Level& c);
// Lanes per vertex of the gain pass for this graph's average degree.
int gain_lanes(const gkd::DGraph& g);
// Candidate matching with the context's tuned launch shape (tuned on first
// use of a degree profile when ctx->autotune is set; the degree table
// otherwise).
int match_stage(gk_ctx* ctx, hipStream_t st, const gkd::DGraph& g, int32_t* cand);

struct Level {
  gkd::DGraph g;
  Buf<int32_t> xadj, adj, w, vw;  // owned storage (coarse levels, or level-0 fills)
  Buf<int32_t> cmap;              // fine -> coarse ids of the next level
};

}  // namespace gkrt
