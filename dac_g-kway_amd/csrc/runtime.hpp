// Host runtime shared by the plain driver (driver.hip) and the batch/serve
// engine (engine.hip): the caching device pool, buffers, per-stage HIP-event
// timing, the context, and small read-back helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
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
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= 2 * bytes) {
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
    for (auto& v : recs_) v.clear();
    used_ = 0;
  }
  void set_mask(uint32_t m) { mask_ = m; }
  void begin(int id) {
    if (!on_ || !((mask_ >> id) & 1u)) return;
    cur_[id] = take();
    (void)hipEventRecord(cur_[id], st_);
  }
  void end(int id) {
    if (!on_ || !((mask_ >> id) & 1u)) return;
    hipEvent_t e = take();
    (void)hipEventRecord(e, st_);
    recs_[id].emplace_back(cur_[id], e);
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
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool_.push_back(e);
    }
    return pool_[used_++];
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

struct gk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  gkrt::Pool pool;
  char* hpin = nullptr;  // pinned staging for small read-backs
  gkrt::StageClock clock;
  gk_stats* cur = nullptr;  // stats of the running partition (stage counters)
  gkrt::Pinned up, dl;      // pinned staging of the batched form (uploads / read-backs)
};

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

struct Level {
  gkd::DGraph g;
  Buf<int32_t> xadj, adj, w, vw;  // owned storage (coarse levels, or level-0 fills)
  Buf<int32_t> cmap;              // fine -> coarse ids of the next level
};

}  // namespace gkrt
