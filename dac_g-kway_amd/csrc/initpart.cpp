// Coarsest-level initial partitioner (host C++).
//
// The reference hands its coarsest graph to METIS_PartGraphKway
// (gkway/metis_partition.hpp:46-47).  The METIS it uses exists only as a
// prebuilt binary, so this library supplies its own deterministic
// partitioner for that step:
//   multilevel recursive bisection
//     - coarsening by sequential heavy-edge matching in a seeded random order,
//     - initial bisection by greedy graph growing from several seeds,
//     - 2-way Fiduccia-Mattheyses refinement at every level,
//   followed by a k-way greedy boundary pass that repairs any part above the
//   cap and takes positive-gain moves.
// Every random choice comes from a SplitMix64 stream seeded by (seed, depth,
// first part id), so the output depends only on the input graph, k, cap and
// seed -- never on the number of threads.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/gkway.h"
#include "host_graph.hpp"

#ifndef GKIP_NCUTS0
#define GKIP_NCUTS0 4
#endif
#ifndef GKIP_NCUTS1
#define GKIP_NCUTS1 3
#endif
#ifndef GKIP_NCUTS2
#define GKIP_NCUTS2 2
#endif
#ifndef GKIP_FMPASS
#define GKIP_FMPASS 6
#endif
#ifndef GKIP_TRIALS
#define GKIP_TRIALS 8
#endif
#ifndef GKIP_SPAWN_MIN
#define GKIP_SPAWN_MIN 200  // smallest subgraph whose trials / halves become pool tasks
#endif

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  uint32_t below(uint32_t m) { return static_cast<uint32_t>((static_cast<uint64_t>(next() >> 32) * m) >> 32); }
};

struct G {
  int32_t n = 0;
  std::vector<int32_t> xadj, adj, w, vw;
  int64_t tw = 0;  // total vertex weight
};

// ------------------------------------------------------------------ heap
// Indexed binary max-heap of vertices ordered by (gain desc, vertex asc).
// Both go into one 64-bit key: the gain in the high 32 bits (a gain is
// bounded by a vertex's weighted degree, below 2^31) and the complemented
// vertex id in the low 32, so one integer compare orders two entries.  Keys
// are distinct, so the order in which vertices leave the heap depends on the
// keys alone, never on the heap's layout: building it in O(n) by sift-down
// gives the same sequence as n pushes.
class Heap {
 public:
  static int64_t pack(int64_t g, int32_t v) {
    return static_cast<int64_t>(static_cast<uint64_t>(g) << 32) | int64_t(0xFFFFFFFFu - static_cast<uint32_t>(v));
  }
  static int32_t vertex(int64_t key) { return static_cast<int32_t>(0xFFFFFFFFu - static_cast<uint32_t>(key)); }
  // Empties the heap over vertices [0, n); only the positions of vertices
  // still queued are cleared when the size is unchanged.
  void reset(int32_t n) {
    if (static_cast<int32_t>(pos_.size()) != n) {
      pos_.assign(n, -1);
    } else {
      for (int64_t k : h_) pos_[vertex(k)] = -1;
    }
    h_.clear();
  }
  // Replaces the contents (heap empty before) by `keys`, heapified in O(n).
  void build(const std::vector<int64_t>& keys) {
    h_.assign(keys.begin(), keys.end());
    const int32_t n = static_cast<int32_t>(h_.size());
    for (int32_t i = n / 2 - 1; i >= 0; --i) sift_down(i);
    for (int32_t i = 0; i < n; ++i) pos_[vertex(h_[i])] = i;
  }
  bool empty() const { return h_.empty(); }
  bool contains(int32_t v) const { return pos_[v] >= 0; }
  int32_t top() const { return vertex(h_[0]); }
  int64_t top_key() const { return h_[0] >> 32; }  // the gain of the top vertex
  void push(int32_t v, int64_t g) {
    const int32_t i = static_cast<int32_t>(h_.size());
    h_.push_back(pack(g, v));
    pos_[v] = i;
    up(i);
  }
  void update(int32_t v, int64_t g) {
    const int32_t i = pos_[v];
    const int64_t old = h_[i];
    h_[i] = pack(g, v);
    if (h_[i] > old) up(i); else down(i);
  }
  void remove(int32_t v) {
    const int32_t i = pos_[v];
    pos_[v] = -1;
    const int64_t last = h_.back();
    h_.pop_back();
    if (i == static_cast<int32_t>(h_.size())) return;
    h_[i] = last;
    pos_[vertex(last)] = i;
    up(i);
    down(pos_[vertex(last)]);
  }

 private:
  void up(int32_t i) {
    const int64_t k = h_[i];
    while (i > 0) {
      const int32_t p = (i - 1) / 2;
      if (h_[p] >= k) break;
      h_[i] = h_[p];
      pos_[vertex(h_[i])] = i;
      i = p;
    }
    h_[i] = k;
    pos_[vertex(k)] = i;
  }
  void down(int32_t i) {
    const int32_t n = static_cast<int32_t>(h_.size());
    const int64_t k = h_[i];
    while (true) {
      int32_t c = 2 * i + 1;
      if (c >= n) break;
      if (c + 1 < n && h_[c + 1] > h_[c]) ++c;
      if (h_[c] <= k) break;
      h_[i] = h_[c];
      pos_[vertex(h_[i])] = i;
      i = c;
    }
    h_[i] = k;
    pos_[vertex(k)] = i;
  }
  // down() without position updates, for build().
  void sift_down(int32_t i) {
    const int32_t n = static_cast<int32_t>(h_.size());
    const int64_t k = h_[i];
    while (true) {
      int32_t c = 2 * i + 1;
      if (c >= n) break;
      if (c + 1 < n && h_[c + 1] > h_[c]) ++c;
      if (h_[c] <= k) break;
      h_[i] = h_[c];
      i = c;
    }
    h_[i] = k;
  }
  std::vector<int64_t> h_;
  std::vector<int32_t> pos_;
};

// ------------------------------------------------------------------ fork-join
// Process-wide worker threads for the independent pieces of the recursive
// bisection (the trial bisections of a node and the two halves below it).
// A thread waiting for its own tasks runs queued tasks meanwhile, so nested
// fork-join never blocks a worker.  Results never depend on which thread runs
// a task.
class ForkJoin {
 public:
  struct Task {
    std::function<void()> fn;
    bool done = false;  // guarded by the pool mutex
  };
  static ForkJoin& get() {
    static ForkJoin pool;
    return pool;
  }
  int workers() const { return static_cast<int>(th_.size()); }
  std::shared_ptr<Task> spawn(std::function<void()> fn) {
    auto t = std::make_shared<Task>();
    t->fn = std::move(fn);
    {
      std::lock_guard<std::mutex> lk(m_);
      q_.push_back(t);
    }
    cv_.notify_all();
    return t;
  }
  // Runs queued tasks until t is done; sleeps when there is nothing to run.
  void wait(const std::shared_ptr<Task>& t) {
    while (true) {
      std::shared_ptr<Task> o;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return t->done || !q_.empty(); });
        if (t->done) return;
        o = q_.back();  // newest first: the waiter's own, most likely
        q_.pop_back();
      }
      run(*o);
    }
  }

 private:
  ForkJoin() {
    const unsigned hw = std::thread::hardware_concurrency();
    const int n = static_cast<int>(hw ? std::min(hw, 16u) : 1u) - 1;
    for (int i = 0; i < n; ++i)
      th_.emplace_back([this] {
        while (true) {
          std::shared_ptr<Task> t;
          {
            std::unique_lock<std::mutex> lk(m_);
            cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
            if (stop_ && q_.empty()) return;
            t = q_.front();
            q_.pop_front();
          }
          run(*t);
        }
      });
  }
  ~ForkJoin() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void run(Task& t) {
    t.fn();
    {
      std::lock_guard<std::mutex> lk(m_);
      t.done = true;
    }
    cv_.notify_all();
  }
  std::vector<std::thread> th_;
  std::deque<std::shared_ptr<Task>> q_;
  std::mutex m_;
  std::condition_variable cv_;
  bool stop_ = false;
};

// Runs fn(c) for c in [0, chunks): chunk 0 on the calling thread, the others
// as pool tasks (the caller may itself be a pool task).
template <class F>
void par_chunks(int chunks, const F& fn) {
  if (chunks <= 1) {
    if (chunks == 1) fn(0);
    return;
  }
  std::vector<std::shared_ptr<ForkJoin::Task>> ts;
  for (int c = 1; c < chunks; ++c) ts.push_back(ForkJoin::get().spawn([&fn, c] { fn(c); }));
  fn(0);
  for (auto& t : ts) ForkJoin::get().wait(t);
}

// ------------------------------------------------------------------ coarsening
// Heavy-edge matching in a seeded random visiting order.  Ties between equal
// edge weights go to the lighter neighbour, then the lower id.  cmap gets the
// coarse id of every vertex; returns the coarse graph.
G coarsen_hem(const G& g, Rng& rng, int64_t maxvw, std::vector<int32_t>& cmap, bool par) {
  const int32_t n = g.n;
  std::vector<int32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int32_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[rng.below(static_cast<uint32_t>(i + 1))]);
  std::vector<int32_t> match(n, -1);
  for (int32_t v : perm) {
    if (match[v] >= 0) continue;
    int32_t best = -1;
    int32_t bw = 0, bvw = 0;
    for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
      int32_t u = g.adj[j];
      if (match[u] >= 0) continue;
      if (int64_t(g.vw[v]) + g.vw[u] > maxvw) continue;
      int32_t w = g.w[j];
      if (best < 0 || w > bw || (w == bw && (g.vw[u] < bvw || (g.vw[u] == bvw && u < best)))) {
        best = u; bw = w; bvw = g.vw[u];
      }
    }
    if (best >= 0) { match[v] = best; match[best] = v; }
    else match[v] = v;
  }
  cmap.assign(n, -1);
  int32_t nc = 0;
  for (int32_t v = 0; v < n; ++v) {
    if (cmap[v] >= 0) continue;
    cmap[v] = nc;
    if (match[v] != v) cmap[match[v]] = nc;
    ++nc;
  }
  // Coarse rows in coarse-id order: coarse vertex cv is first reached at its
  // smaller member v (min(v, match[v])); its row lists the neighbours of v,
  // then of match[v], in order of first appearance, weights summed.
  G c;
  c.n = nc;
  c.vw.assign(nc, 0);
  c.xadj.assign(nc + 1, 0);
  c.tw = g.tw;
  const int chunks = par ? static_cast<int>(std::min<size_t>(8, g.adj.size() / 32768)) : 1;
  if (chunks > 1) {
    // Rows are independent: ranges of coarse ids are built on pool threads
    // into their own buffers and concatenated (the same arrays as below).
    std::vector<int32_t> firstv(nc);
    for (int32_t v = 0; v < n; ++v)
      if (match[v] >= v) firstv[cmap[v]] = v;
    struct Piece {
      std::vector<int32_t> adj, w;
    };
    std::vector<Piece> pieces(chunks);
    par_chunks(chunks, [&](int ci) {
      const int32_t lo = static_cast<int32_t>(int64_t(nc) * ci / chunks);
      const int32_t hi = static_cast<int32_t>(int64_t(nc) * (ci + 1) / chunks);
      Piece& pc = pieces[ci];
      pc.adj.reserve(g.adj.size() / chunks + 1024);
      pc.w.reserve(g.adj.size() / chunks + 1024);
      std::vector<int32_t> slot(nc, -1);
      for (int32_t cv = lo; cv < hi; ++cv) {
        const int32_t v = firstv[cv], m = match[v];
        const int32_t start = static_cast<int32_t>(pc.adj.size());
        const int32_t mem[2] = {v, m};
        for (int k = 0; k < (m != v ? 2 : 1); ++k) {
          const int32_t x = mem[k];
          c.vw[cv] += g.vw[x];
          for (int32_t j = g.xadj[x]; j < g.xadj[x + 1]; ++j) {
            const int32_t cu = cmap[g.adj[j]];
            if (cu == cv) continue;
            if (slot[cu] >= start) {
              pc.w[slot[cu]] += g.w[j];
            } else {
              slot[cu] = static_cast<int32_t>(pc.adj.size());
              pc.adj.push_back(cu);
              pc.w.push_back(g.w[j]);
            }
          }
        }
        c.xadj[cv + 1] = static_cast<int32_t>(pc.adj.size()) - start;  // row length for now
      }
    });
    for (int32_t cv = 0; cv < nc; ++cv) c.xadj[cv + 1] += c.xadj[cv];
    c.adj.resize(c.xadj[nc]);
    c.w.resize(c.xadj[nc]);
    int64_t off = 0;
    for (int ci = 0; ci < chunks; ++ci) {
      std::copy(pieces[ci].adj.begin(), pieces[ci].adj.end(), c.adj.begin() + off);
      std::copy(pieces[ci].w.begin(), pieces[ci].w.end(), c.w.begin() + off);
      off += static_cast<int64_t>(pieces[ci].adj.size());
    }
    return c;
  }
  // Rows are built in per-thread buffers that only ever grow (this path
  // never waits on the pool, so no other coarsening runs on the thread
  // meanwhile), then copied out at their exact size.
  thread_local std::vector<int32_t> tadj, tw, slot;
  if (tadj.size() < g.adj.size()) {
    tadj.resize(g.adj.size());
    tw.resize(g.adj.size());
  }
  slot.assign(nc, -1);
  const int32_t* xa = g.xadj.data();
  const int32_t* ad = g.adj.data();
  const int32_t* wt = g.w.data();
  const int32_t* cm = cmap.data();
  int32_t* ca = tadj.data();
  int32_t* cw = tw.data();
  int32_t* sl = slot.data();
  int32_t top = 0;
  for (int32_t v = 0; v < n; ++v) {
    const int32_t m = match[v];
    if (m < v) continue;  // second member: handled with the first
    const int32_t cv = cm[v];
    const int32_t start = top;
    const int32_t mem[2] = {v, m};
    for (int k = 0; k < (m != v ? 2 : 1); ++k) {
      const int32_t x = mem[k];
      c.vw[cv] += g.vw[x];
      for (int32_t j = xa[x]; j < xa[x + 1]; ++j) {
        const int32_t cu = cm[ad[j]];
        if (cu == cv) continue;
        if (sl[cu] >= start) {
          cw[sl[cu]] += wt[j];
        } else {
          sl[cu] = top;
          ca[top] = cu;
          cw[top] = wt[j];
          ++top;
        }
      }
    }
    c.xadj[cv + 1] = top;
  }
  c.adj.assign(ca, ca + top);
  c.w.assign(cw, cw + top);
  return c;
}

// ------------------------------------------------------------------ 2-way FM
struct Bal {
  int64_t max[2];
};

inline int64_t overload(const int64_t pw[2], const Bal& b) {
  return std::max<int64_t>(0, pw[0] - b.max[0]) + std::max<int64_t>(0, pw[1] - b.max[1]);
}

// Boundary FM refinement.  Moves that keep both sides within their maxima
// (or reduce the overload) are taken in max-gain order; the best prefix of a
// pass (lowest (overload, cut)) is kept.
// Buffers of fm2, one set per thread (fm2 never waits on the pool, so no
// other fm2 call can run on its thread while it holds them).
struct FmWork {
  std::vector<int64_t> id, ed, keys[2];
  std::vector<uint8_t> locked;
  std::vector<int32_t> moved;
  Heap q[2];
};

// conn0 / inc (optional): every vertex's edge weight to side 0 and in all,
// as a growing run leaves them; the degrees then come in O(n).
void fm2(const G& g, std::vector<uint8_t>& side, const Bal& bal, int passes, int64_t& cut,
         const int64_t* conn0 = nullptr, const int64_t* inc = nullptr) {
  const int32_t n = g.n;
  thread_local FmWork W;
  std::vector<int64_t>& id = W.id;
  std::vector<int64_t>& ed = W.ed;
  id.assign(n, 0);
  ed.assign(n, 0);
  int64_t pw[2] = {0, 0};
  for (int32_t v = 0; v < n; ++v) pw[side[v]] += g.vw[v];
  Heap* q = W.q;
  std::vector<uint8_t>& locked = W.locked;
  std::vector<int32_t>& moved = W.moved;
  locked.resize(n);
  const int32_t limit = std::max<int32_t>(25, std::min<int32_t>(n / 20, 300));
  // Internal / external degrees: computed once, then kept exact through the
  // moves of a pass and their rollback.
  if (conn0) {
    for (int32_t v = 0; v < n; ++v) {
      const int64_t a = side[v] == 0 ? conn0[v] : inc[v] - conn0[v];
      id[v] = a;
      ed[v] = inc[v] - a;
    }
  } else {
    for (int32_t v = 0; v < n; ++v) {
      const uint8_t sv = side[v];
      int64_t a = 0, t = 0;
      for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
        const int32_t wj = g.w[j];
        t += wj;
        a += side[g.adj[j]] == sv ? wj : 0;
      }
      id[v] = a;
      ed[v] = t - a;
    }
  }
  for (int pass = 0; pass < passes; ++pass) {
    q[0].reset(n);
    q[1].reset(n);
    std::fill(locked.begin(), locked.end(), 0);
    W.keys[0].clear();
    W.keys[1].clear();
    for (int32_t v = 0; v < n; ++v)
      if (ed[v] > 0) W.keys[side[v]].push_back(Heap::pack(ed[v] - id[v], v));
    q[0].build(W.keys[0]);
    q[1].build(W.keys[1]);
    moved.clear();
    int64_t cur_cut = cut, best_cut = cut;
    int64_t best_ol = overload(pw, bal);
    int32_t best_len = 0, since_best = 0;
    int64_t cur_pw[2] = {pw[0], pw[1]};
    while (true) {
      // Pick the source side: an overloaded side must shed weight; otherwise
      // take the better top among moves that keep the target within bounds.
      int from = -1;
      const bool ol0 = cur_pw[0] > bal.max[0], ol1 = cur_pw[1] > bal.max[1];
      if (ol0 && !q[0].empty()) from = 0;
      else if (ol1 && !q[1].empty()) from = 1;
      else {
        int64_t bg = INT64_MIN;
        for (int s = 0; s < 2; ++s) {
          if (q[s].empty()) continue;
          int32_t v = q[s].top();
          if (cur_pw[1 - s] + g.vw[v] > bal.max[1 - s]) continue;
          if (q[s].top_key() > bg) { bg = q[s].top_key(); from = s; }
        }
      }
      if (from < 0) break;
      int32_t v = q[from].top();
      q[from].remove(v);
      const int to = 1 - from;
      if (!(ol0 || ol1) && cur_pw[to] + g.vw[v] > bal.max[to]) continue;
      cur_cut -= ed[v] - id[v];
      cur_pw[from] -= g.vw[v];
      cur_pw[to] += g.vw[v];
      side[v] = static_cast<uint8_t>(to);
      locked[v] = 1;
      std::swap(id[v], ed[v]);
      moved.push_back(v);
      for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
        int32_t u = g.adj[j];
        int32_t w = g.w[j];
        if (side[u] == to) { id[u] += w; ed[u] -= w; }
        else { id[u] -= w; ed[u] += w; }
        if (locked[u]) continue;
        int s = side[u];
        if (q[s].contains(u)) {
          if (ed[u] > 0) q[s].update(u, ed[u] - id[u]); else q[s].remove(u);
        } else if (ed[u] > 0) {
          q[s].push(u, ed[u] - id[u]);
        }
      }
      int64_t ol = overload(cur_pw, bal);
      if (ol < best_ol || (ol == best_ol && cur_cut < best_cut)) {
        best_ol = ol;
        best_cut = cur_cut;
        best_len = static_cast<int32_t>(moved.size());
        since_best = 0;
      } else if (++since_best > limit) {
        break;
      }
    }
    // Roll back past the best prefix, newest first, undoing each move's
    // degree updates.
    for (int32_t i = static_cast<int32_t>(moved.size()) - 1; i >= best_len; --i) {
      const int32_t v = moved[i];
      const int to = side[v], from = 1 - to;
      for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
        const int32_t u = g.adj[j];
        const int32_t w = g.w[j];
        if (side[u] == to) { id[u] -= w; ed[u] += w; }
        else { id[u] += w; ed[u] -= w; }
      }
      std::swap(id[v], ed[v]);
      side[v] = static_cast<uint8_t>(from);
      cur_pw[to] -= g.vw[v];
      cur_pw[from] += g.vw[v];
    }
    pw[0] = cur_pw[0];
    pw[1] = cur_pw[1];
    bool improved = best_cut < cut || best_len > 0;
    cut = best_cut;
    if (!improved || best_len == 0) break;
  }
}

// Buffers of the growing trials of one bisection: the incident edge weight
// of every vertex (computed once), side-0 connectivity, room for one row's
// side-1 neighbours and the queue.
struct GrowWork {
  std::vector<int64_t> inc, conn0;
  std::vector<int32_t> front;  // side-1 neighbours of the vertex being added
  Heap q;
  explicit GrowWork(const G& g) : inc(g.n, 0) {
    int32_t maxdeg = 1;
    for (int32_t u = 0; u < g.n; ++u) {
      for (int32_t k = g.xadj[u]; k < g.xadj[u + 1]; ++k) inc[u] += g.w[k];
      maxdeg = std::max(maxdeg, g.xadj[u + 1] - g.xadj[u]);
    }
    front.resize(maxdeg);
  }
};

// Greedy graph growing from `seedv`: side 0 grows by max gain until it
// reaches its target weight.  Returns the cut of the result (each added
// vertex turns its edges to side 0 internal and the rest cut).  gw.conn0
// ends as every vertex's edge weight to side 0.
int64_t grow(const G& g, GrowWork& gw, int32_t seedv, int64_t target0, std::vector<uint8_t>& side) {
  const int32_t n = g.n;
  Heap& q = gw.q;
  side.assign(n, 1);
  std::vector<int64_t>& conn0 = gw.conn0;
  conn0.assign(n, 0);
  const std::vector<int64_t>& inc = gw.inc;
  int32_t* front = gw.front.data();
  q.reset(n);
  int64_t w0 = 0, cut = 0;
  int32_t next_unreached = 0;
  auto add = [&](int32_t v) {
    side[v] = 0;
    w0 += g.vw[v];
    cut += inc[v] - 2 * conn0[v];
    // connectivity of every neighbour; the side-1 ones are listed without a
    // branch and their queue keys raised after
    int32_t nf = 0;
    for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
      const int32_t u = g.adj[j];
      conn0[u] += g.w[j];
      front[nf] = u;
      nf += side[u] != 0;
    }
    for (int32_t i = 0; i < nf; ++i) {
      const int32_t u = front[i];
      // gain of moving u to side 0: 2*conn0 - total incident weight
      const int64_t gain = 2 * conn0[u] - inc[u];
      if (q.contains(u)) q.update(u, gain); else q.push(u, gain);
    }
  };
  add(seedv);
  while (w0 < target0) {
    int32_t v = -1;
    if (!q.empty()) { v = q.top(); q.remove(v); }
    else {
      while (next_unreached < n && side[next_unreached] == 0) ++next_unreached;
      if (next_unreached >= n) break;
      v = next_unreached;
    }
    if (side[v] == 0) continue;
    // Stop before overshooting by more than half the vertex weight.
    if (w0 + g.vw[v] > target0 && (w0 + g.vw[v] - target0) > (target0 - w0)) break;
    add(v);
  }
  return cut;
}


// Multilevel bisection of g.  target0 = desired weight of side 0.  Returns
// the cut of the result (tracked exactly through growth, FM and projection).
int64_t ml_bisect(const G& g, int64_t target0, double tol, uint64_t seed, std::vector<uint8_t>& side, bool par) {
  Rng rng(seed);
  const int64_t t1 = g.tw - target0;
  Bal bal;
  bal.max[0] = static_cast<int64_t>(std::floor(double(target0) * (1.0 + tol)));
  bal.max[1] = static_cast<int64_t>(std::floor(double(t1) * (1.0 + tol)));
  // Coarsen.
  std::vector<G> levels;
  std::vector<std::vector<int32_t>> cmaps;
  const G* cur = &g;
  const int64_t maxvw = std::max<int64_t>(1, static_cast<int64_t>(1.5 * double(g.tw) / 60.0));
  while (cur->n > 120) {
    std::vector<int32_t> cm;
    G c = coarsen_hem(*cur, rng, maxvw, cm, par);
    if (c.n > 0.92 * cur->n) break;
    cmaps.push_back(std::move(cm));
    levels.push_back(std::move(c));
    cur = &levels.back();
  }
  // Initial bisection: best of several grown regions.
  const G& cg = *cur;
  std::vector<uint8_t> best, trial;
  int64_t best_cut = INT64_MAX, best_ol = INT64_MAX;
  const int trials = cg.n <= 1 ? 1 : GKIP_TRIALS;
  GrowWork gw(cg);
  for (int t = 0; t < trials; ++t) {
    int32_t s = static_cast<int32_t>(rng.below(static_cast<uint32_t>(std::max(1, cg.n))));
    int64_t c = grow(cg, gw, s, target0, trial);
    fm2(cg, trial, bal, 4, c, gw.conn0.data(), gw.inc.data());
    int64_t pw[2] = {0, 0};
    for (int32_t v = 0; v < cg.n; ++v) pw[trial[v]] += cg.vw[v];
    int64_t ol = overload(pw, bal);
    if (ol < best_ol || (ol == best_ol && c < best_cut)) { best_ol = ol; best_cut = c; best = trial; }
  }
  // Uncoarsen with refinement.
  side = std::move(best);
  for (int32_t l = static_cast<int32_t>(levels.size()) - 1; l >= 0; --l) {
    const G& fine = (l == 0) ? g : levels[l - 1];
    const std::vector<int32_t>& cm = cmaps[l];
    std::vector<uint8_t> fs(fine.n);
    for (int32_t v = 0; v < fine.n; ++v) fs[v] = side[cm[v]];
    side = std::move(fs);
    int64_t c = best_cut;  // projection preserves the cut
    fm2(fine, side, bal, GKIP_FMPASS, c);
    best_cut = c;
  }
  return best_cut;
}

// Induced subgraph of the vertices with side[v] == s.
G induced(const G& g, const std::vector<uint8_t>& side, uint8_t s, std::vector<int32_t>& map_back) {
  std::vector<int32_t> loc(g.n, -1);
  map_back.clear();
  int64_t ub = 0;  // entries of the kept rows: room for the induced rows
  for (int32_t v = 0; v < g.n; ++v)
    if (side[v] == s) {
      loc[v] = static_cast<int32_t>(map_back.size());
      map_back.push_back(v);
      ub += g.xadj[v + 1] - g.xadj[v];
    }
  G h;
  h.n = static_cast<int32_t>(map_back.size());
  h.adj.reserve(ub);
  h.w.reserve(ub);
  h.xadj.assign(h.n + 1, 0);
  h.vw.resize(h.n);
  for (int32_t i = 0; i < h.n; ++i) {
    int32_t v = map_back[i];
    h.vw[i] = g.vw[v];
    h.tw += g.vw[v];
    for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
      int32_t u = loc[g.adj[j]];
      if (u < 0) continue;
      h.adj.push_back(u);
      h.w.push_back(g.w[j]);
    }
    h.xadj[i + 1] = static_cast<int32_t>(h.adj.size());
  }
  return h;
}

struct RBCtx {
  double tol;
  int ncuts0, ncuts1, ncuts2;
  uint64_t seed;
  bool parallel;  // use the fork-join pool
};

void rb(const G& g, int32_t k, int32_t part0, int depth, RBCtx& ctx, int32_t* out,
        const std::vector<int32_t>& ids) {
  if (k == 1 || g.n == 0) {
    for (int32_t i = 0; i < g.n; ++i) out[ids[i]] = part0;
    return;
  }
  if (g.n == 1) {
    out[ids[0]] = part0;
    return;
  }
  const int32_t k0 = k / 2, k1 = k - k0;
  const int64_t target0 = g.tw * k0 / k;
  std::vector<uint8_t> side;
  const uint64_t s = ctx.seed * 0x9E3779B97F4A7C15ull ^ (uint64_t(depth) << 40) ^ uint64_t(part0) * 0xD1B54A32D192ED03ull;
  // Several independent multilevel bisections (more near the root, where a
  // bisection decides the most); keep the lowest (overload, cut, trial).
  // Trials run on spare threads; the pick does not depend on scheduling.
  const int ncuts = depth == 0 ? ctx.ncuts0 : (depth < 3 ? ctx.ncuts1 : ctx.ncuts2);
  std::vector<std::vector<uint8_t>> trial(ncuts);
  std::vector<int64_t> key_ol(ncuts), key_cut(ncuts);
  auto run_trial = [&](int c) {
    key_cut[c] = ml_bisect(g, target0, ctx.tol, s + 0x51ED2701ull * uint64_t(c), trial[c], ctx.parallel);
    int64_t pw0 = 0;
    for (int32_t v = 0; v < g.n; ++v) if (trial[c][v] == 0) pw0 += g.vw[v];
    const int64_t t1 = g.tw - target0;
    const int64_t m0 = static_cast<int64_t>(std::floor(double(target0) * (1.0 + ctx.tol)));
    const int64_t m1 = static_cast<int64_t>(std::floor(double(t1) * (1.0 + ctx.tol)));
    key_ol[c] = std::max<int64_t>(0, pw0 - m0) + std::max<int64_t>(0, (g.tw - pw0) - m1);
  };
  if (ctx.parallel && g.n > GKIP_SPAWN_MIN) {
    std::vector<std::shared_ptr<ForkJoin::Task>> ts;
    for (int c = 1; c < ncuts; ++c) ts.push_back(ForkJoin::get().spawn([&, c] { run_trial(c); }));
    run_trial(0);
    for (auto& t : ts) ForkJoin::get().wait(t);
  } else {
    for (int c = 0; c < ncuts; ++c) run_trial(c);
  }
  int bc = 0;
  for (int c = 1; c < ncuts; ++c)
    if (key_ol[c] < key_ol[bc] || (key_ol[c] == key_ol[bc] && key_cut[c] < key_cut[bc])) bc = c;
  side.swap(trial[bc]);
  std::vector<int32_t> mb0, mb1;
  G g0, g1;
  if (ctx.parallel && g.n > GKIP_SPAWN_MIN) {
    auto t = ForkJoin::get().spawn([&] { g1 = induced(g, side, 1, mb1); });
    g0 = induced(g, side, 0, mb0);
    ForkJoin::get().wait(t);
  } else {
    g0 = induced(g, side, 0, mb0);
    g1 = induced(g, side, 1, mb1);
  }
  std::vector<int32_t> ids0(mb0.size()), ids1(mb1.size());
  for (size_t i = 0; i < mb0.size(); ++i) ids0[i] = ids[mb0[i]];
  for (size_t i = 0; i < mb1.size(); ++i) ids1[i] = ids[mb1[i]];
  // The two halves are independent: one may run on another thread.
  if (ctx.parallel && g.n > GKIP_SPAWN_MIN) {
    auto t = ForkJoin::get().spawn([&] { rb(g0, k0, part0, depth + 1, ctx, out, ids0); });
    rb(g1, k1, part0 + k0, depth + 1, ctx, out, ids1);
    ForkJoin::get().wait(t);
  } else {
    rb(g0, k0, part0, depth + 1, ctx, out, ids0);
    rb(g1, k1, part0 + k0, depth + 1, ctx, out, ids1);
  }
}

// k-way greedy boundary pass: repair parts above cap, then positive-gain
// moves that respect the cap.  Vertices are visited in id order each sweep.
void kway_polish(const G& g, int32_t k, int64_t cap, int32_t* part, int sweeps) {
  std::vector<int64_t> pw(k, 0);
  for (int32_t v = 0; v < g.n; ++v) pw[part[v]] += g.vw[v];
  std::vector<int64_t> conn(k, 0);
  std::vector<int32_t> touched;
  for (int s = 0; s < sweeps; ++s) {
    int64_t moves = 0;
    for (int32_t v = 0; v < g.n; ++v) {
      const int32_t p = part[v];
      touched.clear();
      bool boundary = false;
      for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
        int32_t q = part[g.adj[j]];
        if (conn[q] == 0) touched.push_back(q);
        conn[q] += g.w[j];
        if (q != p) boundary = true;
      }
      if (boundary) {
        const bool over = pw[p] > cap;
        int32_t bestq = -1;
        int64_t bestg = 0;
        for (int32_t q : touched) {
          if (q == p) continue;
          if (pw[q] + g.vw[v] > cap) continue;
          int64_t gain = conn[q] - conn[p];
          bool better;
          if (over) better = bestq < 0 || gain > bestg || (gain == bestg && (pw[q] < pw[bestq] || (pw[q] == pw[bestq] && q < bestq)));
          else better = gain > bestg || (gain == bestg && gain > 0 && q < bestq);
          if (better) { bestq = q; bestg = gain; }
        }
        if (bestq >= 0 && (over || bestg > 0)) {
          pw[p] -= g.vw[v];
          pw[bestq] += g.vw[v];
          part[v] = bestq;
          ++moves;
        }
      }
      for (int32_t q : touched) conn[q] = 0;
    }
    if (moves == 0) break;
  }
}

}  // namespace

extern "C" int64_t gk_part_cap(int64_t total_vwgt, int32_t k, int32_t ubfactor) {
  if (k < 1) return total_vwgt;
  int64_t a = total_vwgt * (1000 + int64_t(ubfactor)) / (1000 * int64_t(k));
  int64_t b = (total_vwgt + k - 1) / k;
  return std::max(a, b);
}

extern "C" int gk_initpart(const gk_csr* hg, int32_t k, int64_t cap, int32_t seed,
                           int32_t nthreads, int32_t* part, int64_t* cut) {
  if (!hg || !part || k < 1 || hg->n < 0) return GK_ERR_ARG;
  G g;
  g.n = hg->n;
  g.xadj.assign(hg->xadj, hg->xadj + hg->n + 1);
  const int32_t nnz = g.xadj[g.n];
  g.adj.assign(hg->adjncy, hg->adjncy + nnz);
  if (hg->adjwgt) g.w.assign(hg->adjwgt, hg->adjwgt + nnz); else g.w.assign(nnz, 1);
  if (hg->vwgt) g.vw.assign(hg->vwgt, hg->vwgt + g.n); else g.vw.assign(g.n, 1);
  g.tw = 0;
  for (int32_t v = 0; v < g.n; ++v) g.tw += g.vw[v];
  if (k == 1) {
    std::fill(part, part + g.n, 0);
    if (cut) *cut = 0;
    return GK_OK;
  }
  // Per-bisection tolerance: (cap / (T/k))^(1/levels) - 1, slightly tightened
  // so that rounding of the recursive targets stays inside the cap.
  const int levels = static_cast<int>(std::ceil(std::log2(double(k))));
  const double ratio = double(cap) * k / std::max<double>(1.0, double(g.tw));
  double tol = std::pow(std::max(1.0, ratio), 1.0 / std::max(1, levels)) - 1.0;
  tol = std::max(0.0, tol * 0.9);
  RBCtx ctx;
  ctx.tol = tol;
  ctx.ncuts0 = GKIP_NCUTS0;
  ctx.ncuts1 = GKIP_NCUTS1;
  ctx.ncuts2 = GKIP_NCUTS2;
  ctx.seed = static_cast<uint64_t>(seed) + 1;
  // nthreads == 1: everything on the calling thread; otherwise the
  // process-wide pool (up to 16 threads) helps.
  ctx.parallel = nthreads != 1 && ForkJoin::get().workers() > 0;
  std::vector<int32_t> ids(g.n);
  std::iota(ids.begin(), ids.end(), 0);
  rb(g, k, 0, 0, ctx, part, ids);
  kway_polish(g, k, cap, part, 4);
  if (cut) {
    int64_t c = 0;
    for (int32_t v = 0; v < g.n; ++v)
      for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j)
        if (v < g.adj[j] && part[v] != part[g.adj[j]]) c += g.w[j];
    *cut = c;
  }
  return GK_OK;
}
