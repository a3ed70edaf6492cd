// Persistent work-stealing launches (GK_FLAG_PERSISTENT).
//
// A persistent launch is sized to the device -- its multiprocessor count
// times the workgroups of the kernel that fit on one -- instead of to the
// work.  Its waves take work from kStealQueues queues until every queue is
// empty.  The work is a run of "sweeps" (64 consecutive global thread slots,
// the unit one wave covers in the kernel's direct launch), cut into
// kStealQueues contiguous ranges; queue q hands out range q through a 64-bit
// counter in its own 128-byte line, a grab at a time.  A wave starts at
// queue blockIdx.x % kStealQueues (the dispatcher places consecutive
// workgroups on the eight XCDs in turn, so a range is mostly read through
// one XCD's L2) and moves on to the next queue when its own is empty; it
// reads a queue's counter before taking from it, so an empty queue costs a
// load, not an atomic.  Which wave takes which sweep
// depends on timing: the kernels that use this write results that do not
// depend on the schedule.
#pragma once
#include <mutex>
#include <unordered_map>

#include "dev_common.hpp"

namespace gkd {

constexpr int kStealQueues = 8;
constexpr int kStealLine = 16;  // 64-bit counters from one queue to the next (128 bytes)
constexpr int kStealWords = kStealQueues * kStealLine;

// Host-side handle: the counters (kStealWords 64-bit words of device memory,
// cleared by the launcher before every launch) and the device's
// multiprocessor count.
struct Steal {
  unsigned long long* ctr = nullptr;
  int cus = 0;
};

// Grab sizes (guided self-scheduling): a wave takes about 1 / div of the
// sweeps its queue has left, and at least min_chunk, so grabs are large while
// much remains and small at the end, where the last ones finish together.
// div = the queue's waves x kStealGuide; min_chunk = range / (div x
// kStealTail).  The grabs per queue stay near div x (1 + ln kStealTail).
// Same-address atomics serialise, and the grabs are what this schedule costs
// over a direct launch: measured on MI355X (2^22-point Delaunay, k = 64),
// coarsening took 5.8 ms with guide 1 / tail 16, 8.0 ms with guide 2, 9.8 ms
// with guide 4 and 5.3 ms with guide 1 / tail 4, against 3.4 ms for the
// direct launches.
constexpr int kStealGuide = 1;
constexpr int kStealTail = 4;

struct StealShape {
  int64_t div;        // remaining / div sweeps per grab
  int64_t min_chunk;  // sweeps per grab at least
};

__host__ __device__ inline StealShape steal_shape(int64_t sweeps, int64_t waves, int guide = kStealGuide,
                                                  int tail = kStealTail) {
  int64_t div = (waves / kStealQueues) * guide;
  if (div < 1) div = 1;
  const int64_t range = sweeps / kStealQueues + 1;
  int64_t mc = range / (div * tail);
  return {div, mc < 1 ? 1 : mc};
}

// Calls f(sw) for every sweep sw of [0, sweeps) that this wave takes.  The
// whole wave runs it together (wave-uniform control flow).
template <class F>
__device__ __forceinline__ void steal_sweeps(unsigned long long* ctr, int64_t sweeps, StealShape sh, F&& f) {
  const int home = static_cast<int>(blockIdx.x % kStealQueues);
  for (int i = 0; i < kStealQueues; ++i) {
    const int q = (home + i) % kStealQueues;
    const int64_t lo = sweeps * q / kStealQueues, hi = sweeps * (q + 1) / kStealQueues;
    const unsigned long long len = static_cast<unsigned long long>(hi - lo);
    unsigned long long* c = ctr + q * kStealLine;
    while (true) {
      unsigned long long start = ~0ull, chunk = 0;
      if (lane_id() == 0) {
        const unsigned long long seen = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (seen < len) {
          chunk = (len - seen) / static_cast<unsigned long long>(sh.div);
          if (chunk < static_cast<unsigned long long>(sh.min_chunk)) chunk = sh.min_chunk;
          start = atomicAdd(c, chunk);
        }
      }
      start = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(start)) |
              (static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(start >> 32))) << 32);
      chunk = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(chunk));
      if (start >= len) break;
      const int64_t s0 = lo + static_cast<int64_t>(start);
      const int64_t s1 = s0 + static_cast<int64_t>(chunk) < hi ? s0 + static_cast<int64_t>(chunk) : hi;
      for (int64_t sw = s0; sw < s1; ++sw) f(sw);
    }
  }
}

// Workgroups of `kern` (block size B, no dynamic LDS) that fit on one
// multiprocessor, from the occupancy calculator; remembered per kernel.
inline int steal_blocks_per_cu(const void* kern, int B) {
  static std::mutex mu;
  static std::unordered_map<const void*, int> memo;
  std::lock_guard<std::mutex> lk(mu);
  auto it = memo.find(kern);
  if (it != memo.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, B, 0) != hipSuccess || nb < 1) {
    (void)hipGetLastError();  // the query's failure is not the launch's
    nb = 1;
  }
  memo[kern] = nb;
  return nb;
}

// Grid of a persistent launch: every workgroup slot of the device, but no
// more workgroups than the direct launch would have.
inline unsigned steal_grid(const void* kern, int B, const Steal& ws, int64_t direct_grid, int64_t cap = 0) {
  int64_t g = int64_t(ws.cus > 0 ? ws.cus : 1) * steal_blocks_per_cu(kern, B);
  if (direct_grid < g) g = direct_grid;
  if (cap > 0 && cap < g) g = cap;
  return static_cast<unsigned>(g < 1 ? 1 : g);
}

}  // namespace gkd
