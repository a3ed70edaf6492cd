// Persistent work-stealing launches (GK_FLAG_PERSISTENT).
//
// A persistent launch is sized to the device -- its multiprocessor count
// times the workgroups of the kernel that fit on one -- instead of to the
// work, and its waves loop over the work.  The work is a run of "sweeps" (64
// consecutive global thread slots, the unit one wave covers in the kernel's
// direct launch).  Each wave first takes its static share of all but the
// last `tail` sweeps, with no shared state: the part is cut into eight
// contiguous ranges, one per XCD (workgroup b runs on XCD b % 8), and the
// waves of an XCD's workgroups stride over its range.  The tail -- one sweep per wave of the grid, when there
// are more than kStealStaticWaves sweeps per wave -- is handed out by
// kStealQueues queues, a workgroup at a time (one sweep per wave of the
// workgroup per grab): a workgroup whose static share ran long leaves the
// tail to the others.  The tail is cut into kStealQueues ranges, one per
// queue, each with a 64-bit counter in its own 128-byte line (cleared by the
// launcher before the launch); a workgroup drains its home queue
// (blockIdx.x % kStealQueues) and then steals from the next one (from all of
// them when the grid has fewer workgroups than queues, so that every range
// has a workgroup that drains it).
//
// Why so little stealing: same-address atomics serialise on MI355X, and the
// queue operations, not imbalance, set the cost of finer schedules.
// Measured on a 2^22-point Delaunay graph (k = 64; coarsening / refinement
// per V-cycle): direct launches 3.40 / 1.72 ms; this grid with the static
// schedule alone 3.55 / 1.58 ms; every wave grabbing 1/(its queue's waves)
// of what is left (guided self-scheduling) from all 8 queues 5.20 / 2.45 ms;
// fixed per-wave grabs from two queues 4.18 / 2.40 ms; this schedule 3.41 /
// 1.49 ms.
//
// Which wave takes which sweep depends on timing: the kernels that use this
// write results that do not depend on the schedule.
#pragma once
#include <mutex>
#include <unordered_map>

#include "dev_common.hpp"

namespace gkd {

constexpr int kStealQueues = 8;
constexpr int kStealLine = 16;  // 64-bit counters from one queue to the next (128 bytes)
constexpr int kStealWords = kStealQueues * kStealLine;
constexpr int kStealStaticWaves = 2;  // up to this many sweeps per wave: static schedule only

// Host-side handle: the counters (kStealWords 64-bit words of device memory)
// and the device's multiprocessor count.
struct Steal {
  unsigned long long* ctr = nullptr;
  int cus = 0;
};

// Sweeps of [0, sweeps) that go through the queues, for a grid of W waves.
__host__ __device__ inline int64_t steal_tail(int64_t sweeps, int64_t W) {
  return sweeps > int64_t(kStealStaticWaves) * W ? W : 0;
}

// Calls f(sw) for every sweep sw of [0, sweeps) that this wave takes.  Every
// thread of the workgroup calls it (the tail grabs are workgroup-wide); f is
// called by whole waves.  ctr == nullptr (counters not cleared for this
// launch): the static schedule takes everything.
template <class F>
__device__ __forceinline__ void steal_sweeps(unsigned long long* ctr, int64_t sweeps, F&& f) {
  __shared__ unsigned long long grab;
  const int wpb = static_cast<int>(blockDim.x >> 6);
  const int wave = static_cast<int>(threadIdx.x >> 6);
  const int64_t W = int64_t(gridDim.x) * wpb;
  const int64_t tail = ctr ? steal_tail(sweeps, W) : 0;
  const int64_t s0 = sweeps - tail;
  if (gridDim.x >= kStealQueues) {
    // By XCD: the workgroups of XCD x (blockIdx.x % 8 == x) stride over the
    // x-th eighth of the static part, so it is read through one L2.
    const uint32_t x = blockIdx.x % kStealQueues, i = blockIdx.x / kStealQueues;
    const int64_t gx = int64_t(gridDim.x / kStealQueues) + (x < gridDim.x % kStealQueues ? 1 : 0);
    const int64_t lo = s0 * x / kStealQueues, hi = s0 * (x + 1) / kStealQueues;
    for (int64_t sw = lo + int64_t(i) * wpb + wave; sw < hi; sw += gx * wpb) f(sw);
  } else {
    for (int64_t sw = int64_t(blockIdx.x) * wpb + wave; sw < s0; sw += W) f(sw);
  }
  if (tail == 0) return;
  const int home = static_cast<int>(blockIdx.x % kStealQueues);
  const int visits = gridDim.x < kStealQueues ? kStealQueues : 2;
  for (int i = 0; i < visits; ++i) {
    const int q = (home + i) % kStealQueues;
    const int64_t lo = s0 + tail * q / kStealQueues, hi = s0 + tail * (q + 1) / kStealQueues;
    const unsigned long long len = static_cast<unsigned long long>(hi - lo);
    unsigned long long* c = ctr + q * kStealLine;
    while (true) {
      __syncthreads();  // every wave has read the previous grab
      if (threadIdx.x == 0) grab = atomicAdd(c, static_cast<unsigned long long>(wpb));
      __syncthreads();
      const unsigned long long start = grab;
      if (start >= len) break;
      const int64_t sw = lo + static_cast<int64_t>(start) + wave;
      if (sw < hi) f(sw);
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
