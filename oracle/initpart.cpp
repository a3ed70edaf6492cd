// Oracle restatement of the coarsest-level initial partitioner (test
// infrastructure only; see oracle.h).
//
// The reference hands its coarsest graph to METIS_PartGraphKway
// (gkway/metis_partition.hpp:46-47), which exists there only as a prebuilt
// library that may not be loaded here.  The product therefore supplies its
// own partitioner for that step (libgkway.so, gk_initpart), and this file
// restates its rules one at a time, single-threaded and without the
// product's data structures (ordered sets instead of indexed heaps, no
// thread pool, no buffer reuse), so that the product's function is checked
// against a second formulation rather than against itself:
//
//   R1  cap-derived tolerance per bisection: t = (cap*k/T)^(1/ceil(log2 k)) - 1,
//       times 0.9, floored at 0;
//   R2  recursive bisection: k0 = floor(k/2) parts get target weight
//       floor(T*k0/k); several independent multilevel bisections per node
//       (4 at depth 0, 3 at depths 1-2, 2 below), the one with the lowest
//       (overload, cut, trial index) is kept; halves recurse in order;
//   R3  one multilevel bisection: heavy-edge matching in a random visiting
//       order (ties: lighter neighbour, then smaller id; pair weight <=
//       1.5*T/60) while the graph has > 120 vertices and shrinks below 92%;
//       coarse ids in order of the smaller member; coarse rows list the
//       neighbours of the first member then the second, in order of first
//       appearance, parallel weights summed, self loops dropped;
//   R4  initial bisection of the coarsest graph: best (overload, cut) of 8
//       greedy growths from random seeds, each followed by 4 FM passes;
//   R5  FM pass: moves by (gain desc, id asc) from the overloaded side, else
//       the better side whose move fits; best (overload, cut) prefix kept;
//       a pass stops max(25, min(n/20, 300)) moves after its last best;
//       passes stop when one keeps nothing; 6 passes per uncoarsening level;
//   R6  a final k-way sweep (at most 4) in id order: a vertex of an
//       overweight part goes to the fitting adjacent part of best (gain,
//       -weight, -id); otherwise to the fitting adjacent part of largest
//       positive gain (smallest id on ties).
// Random numbers: SplitMix64 (Steele, Lea, Flood 2014), one stream per
// bisection trial seeded from (seed, depth, first part id, trial).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <set>
#include <utility>
#include <vector>

#include "oracle.h"

namespace {

using i32 = int32_t;
using i64 = int64_t;
using u64 = uint64_t;

// SplitMix64 stream; below(m) maps the high 32 bits onto [0, m).
struct Stream {
  u64 state;
  u64 next() {
    state += 0x9E3779B97F4A7C15ull;
    u64 z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  i32 below(i32 m) { return i32(((next() >> 32) * u64(uint32_t(m))) >> 32); }
};

struct Sub {
  i32 n = 0;
  std::vector<i32> xadj{0}, adj, w, vw;
  i64 total = 0;
};

// Move queue of one side: ordered by (gain desc, vertex asc).
struct Queue {
  std::set<std::pair<i64, i32>> s;  // (-gain, v)
  std::vector<i64> key;             // current gain of a queued vertex
  std::vector<char> in;
  void clear(i32 n) {
    s.clear();
    key.assign(n, 0);
    in.assign(n, 0);
  }
  bool empty() const { return s.empty(); }
  i32 top_v() const { return s.begin()->second; }
  i64 top_gain() const { return -s.begin()->first; }
  void put(i32 v, i64 g) {
    if (in[v]) s.erase({-key[v], v});
    s.insert({-g, v});
    key[v] = g;
    in[v] = 1;
  }
  void drop(i32 v) {
    if (!in[v]) return;
    s.erase({-key[v], v});
    in[v] = 0;
  }
};

i64 excess(i64 w0, i64 w1, i64 m0, i64 m1) { return std::max<i64>(0, w0 - m0) + std::max<i64>(0, w1 - m1); }

i64 bisection_cut(const Sub& g, const std::vector<char>& side) {
  i64 c = 0;
  for (i32 v = 0; v < g.n; ++v)
    for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j)
      if (v < g.adj[j] && side[v] != side[g.adj[j]]) c += g.w[j];
  return c;
}

// R3: one coarsening step.  Returns the coarse graph; to[v] = coarse id.
Sub coarsen_step(const Sub& g, Stream& rs, i64 pair_max, std::vector<i32>& to) {
  std::vector<i32> order(g.n);
  for (i32 i = 0; i < g.n; ++i) order[i] = i;
  for (i32 i = g.n - 1; i >= 1; --i) std::swap(order[i], order[rs.below(i + 1)]);
  std::vector<i32> mate(g.n, -1);
  for (i32 v : order) {
    if (mate[v] != -1) continue;
    i32 pick = -1;
    for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
      const i32 u = g.adj[j];
      if (mate[u] != -1 || i64(g.vw[v]) + g.vw[u] > pair_max) continue;
      if (pick == -1) { pick = j; continue; }
      const i32 pu = g.adj[pick];
      const bool better = g.w[j] != g.w[pick] ? g.w[j] > g.w[pick]
                          : g.vw[u] != g.vw[pu] ? g.vw[u] < g.vw[pu] : u < pu;
      if (better) pick = j;
    }
    if (pick == -1) {
      mate[v] = v;
    } else {
      mate[v] = g.adj[pick];
      mate[g.adj[pick]] = v;
    }
  }
  to.assign(g.n, -1);
  Sub c;
  for (i32 v = 0; v < g.n; ++v) {
    if (to[v] != -1) continue;
    to[v] = c.n;
    to[mate[v]] = c.n;
    ++c.n;
  }
  c.vw.assign(c.n, 0);
  c.total = g.total;
  for (i32 v = 0; v < g.n; ++v) {
    if (mate[v] < v) continue;  // the second member of a pair
    const i32 cv = to[v];
    std::vector<i32> nb, nw;  // this coarse row, in order of first appearance
    std::vector<i32> members{v};
    if (mate[v] != v) members.push_back(mate[v]);
    for (i32 x : members) {
      c.vw[cv] += g.vw[x];
      for (i32 j = g.xadj[x]; j < g.xadj[x + 1]; ++j) {
        const i32 cu = to[g.adj[j]];
        if (cu == cv) continue;
        size_t at = std::find(nb.begin(), nb.end(), cu) - nb.begin();
        if (at == nb.size()) {
          nb.push_back(cu);
          nw.push_back(g.w[j]);
        } else {
          nw[at] += g.w[j];
        }
      }
    }
    c.adj.insert(c.adj.end(), nb.begin(), nb.end());
    c.w.insert(c.w.end(), nw.begin(), nw.end());
    c.xadj.push_back(i32(c.adj.size()));
  }
  return c;
}

// R5: FM passes on a bisection; `cut` is the bisection's cut on entry and
// the refined cut on return.
void fm_passes(const Sub& g, std::vector<char>& side, i64 m0, i64 m1, int passes, i64& cut) {
  const i64 maxw[2] = {m0, m1};
  // Connectivity of every vertex to its own side (in) and the other (out).
  std::vector<i64> in(g.n, 0), out(g.n, 0);
  for (i32 v = 0; v < g.n; ++v)
    for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j)
      (side[g.adj[j]] == side[v] ? in[v] : out[v]) += g.w[j];
  i64 wt[2] = {0, 0};
  for (i32 v = 0; v < g.n; ++v) wt[int(side[v])] += g.vw[v];
  const i32 patience = std::max<i32>(25, std::min<i32>(g.n / 20, 300));
  Queue q[2];
  for (int pass = 0; pass < passes; ++pass) {
    q[0].clear(g.n);
    q[1].clear(g.n);
    std::vector<char> done(g.n, 0);
    for (i32 v = 0; v < g.n; ++v)
      if (out[v] > 0) q[int(side[v])].put(v, out[v] - in[v]);
    std::vector<i32> seq;
    i64 now_cut = cut, keep_cut = cut;
    i64 keep_ex = excess(wt[0], wt[1], m0, m1);
    size_t keep_len = 0;
    i32 stale = 0;
    i64 cw[2] = {wt[0], wt[1]};
    for (;;) {
      const bool over0 = cw[0] > maxw[0], over1 = cw[1] > maxw[1];
      int src = -1;
      if (over0 && !q[0].empty()) {
        src = 0;
      } else if (over1 && !q[1].empty()) {
        src = 1;
      } else {
        for (int s = 0; s < 2; ++s) {
          if (q[s].empty()) continue;
          if (cw[1 - s] + g.vw[q[s].top_v()] > maxw[1 - s]) continue;
          if (src == -1 || q[s].top_gain() > q[src].top_gain()) src = s;
        }
      }
      if (src == -1) break;
      const i32 v = q[src].top_v();
      q[src].drop(v);
      const int dst = 1 - src;
      if (!over0 && !over1 && cw[dst] + g.vw[v] > maxw[dst]) continue;
      now_cut -= out[v] - in[v];
      cw[src] -= g.vw[v];
      cw[dst] += g.vw[v];
      side[v] = char(dst);
      done[v] = 1;
      std::swap(in[v], out[v]);
      seq.push_back(v);
      for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
        const i32 u = g.adj[j];
        if (side[u] == dst) { in[u] += g.w[j]; out[u] -= g.w[j]; }
        else { in[u] -= g.w[j]; out[u] += g.w[j]; }
        if (done[u]) continue;
        if (out[u] > 0) q[int(side[u])].put(u, out[u] - in[u]);
        else q[int(side[u])].drop(u);
      }
      const i64 ex = excess(cw[0], cw[1], m0, m1);
      if (ex < keep_ex || (ex == keep_ex && now_cut < keep_cut)) {
        keep_ex = ex;
        keep_cut = now_cut;
        keep_len = seq.size();
        stale = 0;
      } else if (++stale > patience) {
        break;
      }
    }
    // Undo the moves after the kept prefix, last first.
    while (seq.size() > keep_len) {
      const i32 v = seq.back();
      seq.pop_back();
      const int now = side[v], back = 1 - now;
      for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
        const i32 u = g.adj[j];
        if (side[u] == now) { in[u] -= g.w[j]; out[u] += g.w[j]; }
        else { in[u] += g.w[j]; out[u] -= g.w[j]; }
      }
      std::swap(in[v], out[v]);
      side[v] = char(back);
      cw[now] -= g.vw[v];
      cw[back] += g.vw[v];
    }
    wt[0] = cw[0];
    wt[1] = cw[1];
    cut = keep_cut;
    if (keep_len == 0) break;
  }
}

// R4: greedy growth of side 0 from `start` until it holds `want` weight.
// Returns the cut of the grown bisection.
i64 grow_from(const Sub& g, i32 start, i64 want, std::vector<char>& side) {
  side.assign(g.n, 1);
  std::vector<i64> deg(g.n, 0), to0(g.n, 0);
  for (i32 v = 0; v < g.n; ++v)
    for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j) deg[v] += g.w[j];
  Queue q;
  q.clear(g.n);
  i64 have = 0, cut = 0;
  i32 scan = 0;
  auto take = [&](i32 v) {
    side[v] = 0;
    have += g.vw[v];
    cut += deg[v] - 2 * to0[v];
    for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
      const i32 u = g.adj[j];
      if (side[u] == 0) continue;
      to0[u] += g.w[j];
      q.put(u, 2 * to0[u] - deg[u]);
    }
  };
  take(start);
  while (have < want) {
    i32 v;
    if (!q.empty()) {
      v = q.top_v();
      q.drop(v);
    } else {
      while (scan < g.n && side[scan] == 0) ++scan;
      if (scan == g.n) break;
      v = scan;
    }
    if (side[v] == 0) continue;
    const i64 after = have + g.vw[v];
    if (after > want && after - want > want - have) break;
    take(v);
  }
  return cut;
}

// R3 + R4 + R5: one multilevel bisection of g toward side-0 weight `want`.
std::vector<char> multilevel_bisect(const Sub& g, i64 want, double tol, u64 seed) {
  Stream rs{seed};
  const i64 m0 = i64(std::floor(double(want) * (1.0 + tol)));
  const i64 m1 = i64(std::floor(double(g.total - want) * (1.0 + tol)));
  const i64 pair_max = std::max<i64>(1, i64(1.5 * double(g.total) / 60.0));
  std::vector<Sub> hier{g};
  std::vector<std::vector<i32>> maps;
  while (hier.back().n > 120) {
    std::vector<i32> to;
    Sub c = coarsen_step(hier.back(), rs, pair_max, to);
    if (double(c.n) > 0.92 * double(hier.back().n)) break;
    maps.push_back(std::move(to));
    hier.push_back(std::move(c));
  }
  const Sub& top = hier.back();
  std::vector<char> best;
  i64 best_cut = 0, best_ex = 0;
  const int tries = top.n <= 1 ? 1 : 8;
  for (int t = 0; t < tries; ++t) {
    const i32 start = rs.below(std::max<i32>(1, top.n));
    std::vector<char> side;
    i64 c = grow_from(top, start, want, side);
    fm_passes(top, side, m0, m1, 4, c);
    i64 w0 = 0;
    for (i32 v = 0; v < top.n; ++v)
      if (side[v] == 0) w0 += top.vw[v];
    const i64 ex = excess(w0, top.total - w0, m0, m1);
    if (t == 0 || ex < best_ex || (ex == best_ex && c < best_cut)) {
      best = side;
      best_cut = c;
      best_ex = ex;
    }
  }
  for (i32 l = i32(maps.size()) - 1; l >= 0; --l) {
    const Sub& fine = hier[l];
    std::vector<char> side(fine.n);
    for (i32 v = 0; v < fine.n; ++v) side[v] = best[maps[l][v]];
    fm_passes(fine, side, m0, m1, 6, best_cut);
    best = std::move(side);
  }
  return best;
}

Sub half(const Sub& g, const std::vector<char>& side, char s, std::vector<i32>& ids) {
  std::vector<i32> at(g.n, -1);
  ids.clear();
  for (i32 v = 0; v < g.n; ++v)
    if (side[v] == s) {
      at[v] = i32(ids.size());
      ids.push_back(v);
    }
  Sub h;
  h.n = i32(ids.size());
  for (i32 v : ids) {
    h.vw.push_back(g.vw[v]);
    h.total += g.vw[v];
    for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j)
      if (at[g.adj[j]] != -1) {
        h.adj.push_back(at[g.adj[j]]);
        h.w.push_back(g.w[j]);
      }
    h.xadj.push_back(i32(h.adj.size()));
  }
  return h;
}

// R2: parts [first, first + k) for the vertices of g (global ids `ids`).
void bisect_parts(const Sub& g, i32 k, i32 first, int depth, double tol, u64 seed, const std::vector<i32>& ids,
                  i32* part) {
  if (k == 1 || g.n <= 1) {
    for (i32 i = 0; i < g.n; ++i) part[ids[i]] = first;
    return;
  }
  const i32 k0 = k / 2;
  const i64 want = g.total * k0 / k;
  const i64 m0 = i64(std::floor(double(want) * (1.0 + tol)));
  const i64 m1 = i64(std::floor(double(g.total - want) * (1.0 + tol)));
  const u64 node = (seed * 0x9E3779B97F4A7C15ull) ^ (u64(depth) << 40) ^ (u64(first) * 0xD1B54A32D192ED03ull);
  const int trials = depth == 0 ? 4 : depth < 3 ? 3 : 2;
  std::vector<char> side;
  i64 best_ex = 0, best_cut = 0;
  for (int t = 0; t < trials; ++t) {
    std::vector<char> s = multilevel_bisect(g, want, tol, node + 0x51ED2701ull * u64(t));
    i64 w0 = 0;
    for (i32 v = 0; v < g.n; ++v)
      if (s[v] == 0) w0 += g.vw[v];
    const i64 ex = excess(w0, g.total - w0, m0, m1);
    const i64 c = bisection_cut(g, s);
    if (t == 0 || ex < best_ex || (ex == best_ex && c < best_cut)) {
      side = std::move(s);
      best_ex = ex;
      best_cut = c;
    }
  }
  for (char s = 0; s < 2; ++s) {
    std::vector<i32> local;
    Sub h = half(g, side, s, local);
    std::vector<i32> global(local.size());
    for (size_t i = 0; i < local.size(); ++i) global[i] = ids[local[i]];
    if (s == 0) bisect_parts(h, k0, first, depth + 1, tol, seed, global, part);
    else bisect_parts(h, k - k0, first + k0, depth + 1, tol, seed, global, part);
  }
}

// R6.
void kway_sweeps(const Sub& g, i32 k, i64 cap, i32* part) {
  std::vector<i64> pw(k, 0);
  for (i32 v = 0; v < g.n; ++v) pw[part[v]] += g.vw[v];
  for (int sweep = 0; sweep < 4; ++sweep) {
    i64 moved = 0;
    for (i32 v = 0; v < g.n; ++v) {
      const i32 p = part[v];
      // Connectivity to each adjacent part, parts in order of first appearance.
      std::vector<std::pair<i32, i64>> con;
      bool edge_out = false;
      for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
        const i32 q = part[g.adj[j]];
        if (q != p) edge_out = true;
        auto it = std::find_if(con.begin(), con.end(), [q](const std::pair<i32, i64>& e) { return e.first == q; });
        if (it == con.end()) con.push_back({q, g.w[j]});
        else it->second += g.w[j];
      }
      if (!edge_out) continue;
      i64 own = 0;
      for (auto& e : con)
        if (e.first == p) own = e.second;
      const bool heavy = pw[p] > cap;
      i32 to = -1;
      i64 to_gain = 0;
      for (auto& e : con) {
        const i32 q = e.first;
        if (q == p || pw[q] + g.vw[v] > cap) continue;
        const i64 gain = e.second - own;
        bool take;
        if (heavy)
          take = to == -1 || gain > to_gain ||
                 (gain == to_gain && (pw[q] < pw[to] || (pw[q] == pw[to] && q < to)));
        else
          take = gain > to_gain || (gain == to_gain && gain > 0 && q < to);
        if (take) {
          to = q;
          to_gain = gain;
        }
      }
      if (to != -1 && (heavy || to_gain > 0)) {
        pw[p] -= g.vw[v];
        pw[to] += g.vw[v];
        part[v] = to;
        ++moved;
      }
    }
    if (moved == 0) break;
  }
}

}  // namespace

extern "C" int oracle_initpart(const gk_csr* hg, int32_t k, int64_t cap, int32_t seed, int32_t nthreads,
                               int32_t* part, int64_t* cut) {
  (void)nthreads;  // single-threaded by construction
  if (!hg || !part || k < 1 || hg->n < 0) return GK_ERR_ARG;
  Sub g;
  g.n = hg->n;
  g.xadj.assign(hg->xadj, hg->xadj + g.n + 1);
  const i32 nnz = g.xadj[g.n];
  for (i32 j = 0; j < nnz; ++j) {
    g.adj.push_back(hg->adjncy[j]);
    g.w.push_back(hg->adjwgt ? hg->adjwgt[j] : 1);
  }
  for (i32 v = 0; v < g.n; ++v) {
    g.vw.push_back(hg->vwgt ? hg->vwgt[v] : 1);
    g.total += g.vw.back();
  }
  if (k == 1) {
    for (i32 v = 0; v < g.n; ++v) part[v] = 0;
    if (cut) *cut = 0;
    return GK_OK;
  }
  // R1.
  const int depth = int(std::ceil(std::log2(double(k))));
  const double ratio = double(cap) * k / std::max(1.0, double(g.total));
  double tol = std::pow(std::max(1.0, ratio), 1.0 / std::max(1, depth)) - 1.0;
  tol = std::max(0.0, tol * 0.9);
  std::vector<i32> ids(g.n);
  for (i32 v = 0; v < g.n; ++v) ids[v] = v;
  bisect_parts(g, k, 0, 0, tol, u64(int64_t(seed)) + 1, ids, part);
  kway_sweeps(g, k, cap, part);
  if (cut) {
    i64 c = 0;
    for (i32 v = 0; v < g.n; ++v)
      for (i32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j)
        if (v < g.adj[j] && part[v] != part[g.adj[j]]) c += g.w[j];
    *cut = c;
  }
  return GK_OK;
}
