// CPU oracle of the multilevel k-way partitioning path (test infrastructure
// only; see oracle.h).  Single-threaded, literal restatement of each rule.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

using i32 = int32_t;
using i64 = int64_t;

uint64_t fnv(const i32* a, i64 n, uint64_t h) {
  for (i64 i = 0; i < n; ++i) { h ^= uint32_t(a[i]); h *= 0x100000001b3ull; }
  return h;
}
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull;

struct Level {
  i32 n = 0;
  std::vector<i32> xadj, adj, w, vw;
  std::vector<i32> cmap;  // to the next coarser level
  gk_csr view() const {
    gk_csr c;
    c.n = n;
    c.nnz = xadj.empty() ? 0 : xadj[n];
    c.xadj = xadj.data();
    c.adjncy = adj.data();
    c.adjwgt = w.data();
    c.vwgt = vw.data();
    return c;
  }
};

inline i32 W(const gk_csr* g, i32 j) { return g->adjwgt ? g->adjwgt[j] : 1; }
inline i32 VW(const gk_csr* g, i32 v) { return g->vwgt ? g->vwgt[v] : 1; }

}  // namespace

extern "C" {

int64_t oracle_part_cap(int64_t T, int32_t k, int32_t ub) {
  // Exact integer form of the reference's float cap
  // T/k*(1+IMBALANCE_RATIO)+1 (gkway/graph_partitioner.hpp:230): the larger of
  // floor(T*(1000+ub)/(1000k)) and ceil(T/k).
  if (k < 1) return T;
  int64_t a = T * (1000 + int64_t(ub)) / (1000 * int64_t(k));
  int64_t b = (T + k - 1) / k;
  return a > b ? a : b;
}

int32_t oracle_coarsen_floor(int32_t n0, int32_t k, int32_t policy, int32_t override_to) {
  // REF: the reference coarsens while n > 20k (gkway/graph_partitioner.hpp:22,66).
  if (policy == GK_COARSEN_REF) return 20 * k;
  if (override_to > 0) return override_to;
  int lg = 0;
  while ((int64_t(1) << lg) < k) ++lg;
  if (lg < 1) lg = 1;
  int64_t a = std::min<int64_t>(int64_t(n0) / (20 * int64_t(lg)), 240 * int64_t(k));
  int64_t b = 30 * int64_t(k);
  return int32_t(std::min<int64_t>(INT32_MAX, std::max(a, b)));
}

// matching_with_score (gkway/coarsen.hpp:39-78): heaviest incident edge,
// ties to the neighbour of smaller degree (:68-74), final ties to the smaller
// id (the reference keeps the first in row order; its coarse rows are sorted
// by id, coarsen.hpp:338, so the two agree there).  Isolated -> itself (:59).
void oracle_match(const gk_csr* g, int32_t* cand) {
  for (i32 v = 0; v < g->n; ++v) {
    i32 bu = v, bw = 0, bd = 0;
    for (i32 j = g->xadj[v]; j < g->xadj[v + 1]; ++j) {
      i32 u = g->adjncy[j], w = W(g, j);
      i32 d = g->xadj[u + 1] - g->xadj[u];
      if (w > bw || (w == bw && (d < bd || (d == bd && u < bu)))) {
        bu = u; bw = w; bd = d;
      }
    }
    cand[v] = bu;
  }
}

// init_group_id + HEC_union_find sweep loop (gkway/coarsen.hpp:351-359,
// 80-118, 461-469).  The groups are the connected components of the
// candidate graph {v, cand[v]}; label[v] = the largest id of v's component,
// the value the reference's max-propagation converges to (:103-112).  The
// reference's stamp is the sweep in which the label last changed, which
// depends on the thread schedule; here stamp[v] = the hop distance from v to
// its component's mutual pair (u, cand[u]) with cand[cand[u]] == u, following
// cand.  Every component of this graph is a tree hanging off exactly one such
// pair (an isolated vertex is its own pair).  Returns the number of
// pointer-jumping rounds that move a pointer: ceil(log2(max stamp)), 0 when
// the max stamp is <= 1; -1 if some cand chain never reaches a pair.
int32_t oracle_group(int32_t n, const int32_t* cand, int32_t* label, int32_t* stamp) {
  std::vector<i32> depth(n, -1), root(n, -1);
  std::vector<i32> path;
  for (i32 v = 0; v < n; ++v) {
    if (depth[v] >= 0) continue;
    path.clear();
    i32 u = v;
    i64 guard = 0;
    while (depth[u] < 0 && cand[cand[u]] != u) {
      path.push_back(u);
      u = cand[u];
      if (++guard > n) return -1;
    }
    if (depth[u] < 0) { depth[u] = 0; root[u] = u; }
    for (i32 i = i32(path.size()) - 1; i >= 0; --i) {
      const i32 x = path[i];
      depth[x] = depth[cand[x]] + 1;
      root[x] = root[cand[x]];
    }
  }
  std::vector<i32> compmax(n, -1);
  i32 maxd = 0;
  for (i32 v = 0; v < n; ++v) {
    const i32 r = root[v];
    const i32 key = std::max(r, cand[r]);
    compmax[key] = std::max(compmax[key], v);
    maxd = std::max(maxd, depth[v]);
  }
  for (i32 v = 0; v < n; ++v) {
    const i32 r = root[v];
    label[v] = compmax[std::max(r, cand[r])];
    stamp[v] = depth[v];
  }
  int32_t rounds = 0;
  while ((i64(1) << rounds) < maxd) ++rounds;
  return rounds;
}

// constraint_group_size + coarse numbering + cal_coarsen_vwgt
// (gkway/coarsen.hpp:361-421, 120-171, 472-483): stable order by
// (label, stamp, id); every group split into chunks of max_group in that
// order; coarse ids numbered by (label, chunk); coarse weights summed into a
// zeroed array (the reference never zeroes it, graph_partitioner.hpp:270).
int32_t oracle_cluster(int32_t n, const int32_t* label, const int32_t* stamp,
                       const int32_t* vwgt, int32_t G, int32_t* order, int32_t* cmap,
                       int32_t* cvwgt) {
  std::vector<i32> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](i32 a, i32 b) {
    if (label[a] != label[b]) return label[a] < label[b];
    return stamp[a] < stamp[b];
  });
  i32 nc = 0;
  i32 gstart = 0;
  for (i32 i = 0; i < n; ++i) {
    if (i == 0 || label[ord[i]] != label[ord[i - 1]]) gstart = i;
    if ((i - gstart) % G == 0) ++nc;
    cmap[ord[i]] = nc - 1;
    order[i] = ord[i];
  }
  for (i32 c = 0; c < nc; ++c) cvwgt[c] = 0;
  for (i32 v = 0; v < n; ++v) cvwgt[cmap[v]] += vwgt ? vwgt[v] : 1;
  return nc;
}

// contraction (gkway/coarsen.hpp:302-349): coarse row c = the neighbours of
// its members mapped through cmap, sorted by id (:338), self entries dropped
// (:276-278), parallel entries merged by summing weights (:280-295).
int32_t oracle_contract(const gk_csr* g, const int32_t* cmap, int32_t nc, int32_t* cxadj,
                        int32_t* cadj, int32_t* cw) {
  std::vector<std::vector<i32>> members(nc);
  for (i32 v = 0; v < g->n; ++v) members[cmap[v]].push_back(v);
  std::vector<std::pair<i32, i32>> row;
  i32 e = 0;
  cxadj[0] = 0;
  for (i32 c = 0; c < nc; ++c) {
    row.clear();
    for (i32 v : members[c])
      for (i32 j = g->xadj[v]; j < g->xadj[v + 1]; ++j) row.emplace_back(cmap[g->adjncy[j]], W(g, j));
    std::sort(row.begin(), row.end());
    for (size_t i = 0; i < row.size();) {
      size_t j = i;
      i64 s = 0;
      while (j < row.size() && row[j].first == row[i].first) s += row[j++].second;
      if (row[i].first != c) { cadj[e] = row[i].first; cw[e] = i32(s); ++e; }
      i = j;
    }
    cxadj[c + 1] = e;
  }
  return e;
}

// expand_partition (gkway/uncoarsen.hpp:325-356).
void oracle_project(int32_t n, const int32_t* cmap, const int32_t* cpart, const int32_t* cbnd,
                    int32_t* part, int32_t* bnd, int32_t* gain, int32_t* dest, int32_t* upd) {
  for (i32 v = 0; v < n; ++v) {
    i32 c = cmap[v];
    part[v] = cpart[c];
    bnd[v] = cbnd[c];
    gain[v] = 0;
    dest[v] = part[v];
    upd[v] = bnd[v] ? 1 : 0;
  }
}

// update_vertex_info + calculate_gain (gkway/uncoarsen.hpp:175-237, 82-109),
// with a signed comparison (the reference compares int gains against an
// unsigned running max, :222-229): gain = max(0, max_q conn_q - conn_p) over
// the parts q != p present in the row, lowest q on ties.
void oracle_gain(const gk_csr* g, int32_t k, const int32_t* part, const int32_t* upd,
                 int32_t* bnd, int32_t* gain, int32_t* dest) {
  std::vector<i64> conn(k, 0);
  std::vector<char> present(k, 0);
  for (i32 v = 0; v < g->n; ++v) {
    if (!upd[v]) continue;
    const i32 p = part[v];
    bool b = false;
    for (i32 j = g->xadj[v]; j < g->xadj[v + 1]; ++j) {
      i32 q = part[g->adjncy[j]];
      conn[q] += W(g, j);
      present[q] = 1;
      if (q != p) b = true;
    }
    bnd[v] = b ? 1 : 0;
    i64 best = 0;
    i32 bq = p;
    if (b) {
      for (i32 q = 0; q < k; ++q) {
        if (q == p || !present[q]) continue;
        i64 gq = conn[q] - conn[p];
        if (gq > best) { best = gq; bq = q; }
      }
    }
    gain[v] = i32(best);
    dest[v] = bq;
    for (i32 j = g->xadj[v]; j < g->xadj[v + 1]; ++j) {
      i32 q = part[g->adjncy[j]];
      conn[q] = 0;
      present[q] = 0;
    }
  }
}

// create_independent_move_buffer (gkway/uncoarsen.hpp:111-173) + mergesort
// by mvRequest::operator< (gkway/move_request.hpp:36-41) + truncation
// (gkway/uncoarsen.hpp:475).  Returns the number of moves kept; *nsel gets
// the number selected before truncation.
int32_t oracle_select(const gk_csr* g, const int32_t* part, const int32_t* bnd, const int32_t* gain,
                      const int32_t* dest, const int32_t* pw, int64_t cap, int32_t M,
                      int32_t* upd, int32_t* moves, int32_t* nsel) {
  (void)part;
  auto fits = [&](i32 u) { return i64(pw[dest[u]]) + VW(g, u) <= cap; };
  std::vector<i32> sel;
  for (i32 v = 0; v < g->n; ++v) {
    upd[v] = 0;
    if (!bnd[v] || gain[v] <= 0 || !fits(v)) continue;
    bool ok = true;
    for (i32 j = g->xadj[v]; j < g->xadj[v + 1]; ++j) {
      i32 u = g->adjncy[j];
      if (gain[u] > 0 && fits(u) && u < v) { ok = false; break; }
    }
    if (ok) sel.push_back(v);
  }
  std::sort(sel.begin(), sel.end(), [&](i32 a, i32 b) {
    if (gain[a] != gain[b]) return gain[a] > gain[b];
    return a < b;
  });
  *nsel = i32(sel.size());
  i32 B = std::min<i32>(i32(sel.size()), M);
  for (i32 i = 0; i < B; ++i) moves[i] = sel[i];
  return B;
}

// find_mv_wgt_sequence + scan + find_balance_sequence + the pw update of
// apply_sequence_move (gkway/uncoarsen.hpp:239-288, 37-80, 305-308): the
// largest i such that every part q with a positive running delta stays
// within the cap after moves 0..i.  pw is advanced by that prefix.
int32_t oracle_balance(const gk_csr* g, int32_t k, const int32_t* moves, int32_t B,
                       const int32_t* part, const int32_t* dest, int32_t* pw, int64_t cap) {
  if (B <= 0) return -1;
  std::vector<i64> delta(k, 0);
  i32 pos = -1;
  for (i32 i = 0; i < B; ++i) {
    i32 v = moves[i];
    delta[part[v]] -= VW(g, v);
    delta[dest[v]] += VW(g, v);
    bool ok = true;
    for (i32 q = 0; q < k; ++q)
      if (delta[q] > 0 && pw[q] + delta[q] > cap) { ok = false; break; }
    if (ok) pos = i;
  }
  std::fill(delta.begin(), delta.end(), 0);
  for (i32 i = 0; i <= pos; ++i) {
    i32 v = moves[i];
    delta[part[v]] -= VW(g, v);
    delta[dest[v]] += VW(g, v);
  }
  for (i32 q = 0; q < k; ++q) pw[q] = i32(pw[q] + delta[q]);
  return pos;
}

// apply_sequence_move (gkway/uncoarsen.hpp:290-323) minus the pw update.
int64_t oracle_apply(const gk_csr* g, const int32_t* moves, int32_t pos, const int32_t* gain,
                     const int32_t* dest, int32_t* part, int32_t* upd) {
  i64 s = 0;
  for (i32 i = 0; i <= pos; ++i) {
    i32 v = moves[i];
    part[v] = dest[v];
    upd[v] = 1;
    for (i32 j = g->xadj[v]; j < g->xadj[v + 1]; ++j) upd[g->adjncy[j]] = 1;
    s += gain[v];
  }
  return s;
}

// calculate_cutsize (gkway/uncoarsen.hpp:398-424) + calculate_partition_wgt
// (gkway/cuda_check.hpp:61-73), 64-bit.
void oracle_cut(const gk_csr* g, int32_t k, const int32_t* part, int64_t* cut, int64_t* pw) {
  i64 c = 0;
  for (i32 q = 0; q < k; ++q) pw[q] = 0;
  for (i32 v = 0; v < g->n; ++v) {
    pw[part[v]] += VW(g, v);
    for (i32 j = g->xadj[v]; j < g->xadj[v + 1]; ++j) {
      i32 u = g->adjncy[j];
      if (v < u && part[v] != part[u]) c += W(g, j);
    }
  }
  *cut = c;
}

// graph_partitioner (gkway/graph_partitioner.hpp:222-365): coarsening level
// loop (:14-123), initial partition at the coarsest level (:125-131 ->
// metis_partition.hpp), uncoarsening + refinement of every finer level
// (:133-220; the coarsest level itself is not refined, :174), final metrics
// (:194-202).
int oracle_partition(const gk_csr* g0, const gk_params* p, oracle_initpart_fn initpart,
                     int32_t* part_out, gk_stats* st) {
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  if (!g0 || !p || !initpart || !part_out || p->k < 1) return GK_ERR_ARG;
  const i32 k = p->k;
  gk_stats S;
  std::memset(&S, 0, sizeof S);
  const bool trace = (p->flags & GK_FLAG_TRACE) != 0;
  const i32 maxlv = p->max_levels > 0 ? std::min(p->max_levels, GK_MAX_LEVELS) : GK_MAX_LEVELS;

  std::vector<Level> L(1);
  {
    Level& a = L[0];
    a.n = g0->n;
    a.xadj.assign(g0->xadj, g0->xadj + g0->n + 1);
    i32 nnz = a.xadj[a.n];
    a.adj.assign(g0->adjncy, g0->adjncy + nnz);
    a.w.resize(nnz);
    for (i32 j = 0; j < nnz; ++j) a.w[j] = W(g0, j);
    a.vw.resize(a.n);
    for (i32 v = 0; v < a.n; ++v) a.vw[v] = VW(g0, v);
  }
  i64 T = 0;
  for (i32 v : L[0].vw) T += v;
  const i64 cap = oracle_part_cap(T, k, p->ubfactor);
  const i32 floor_n = oracle_coarsen_floor(g0->n, k, p->coarsen_policy, p->coarsen_to);
  S.total_vwgt = T;
  S.cap = cap;

  // ---- coarsening
  std::vector<i32> sweeps_of(1, 0);
  std::vector<uint64_t> hcand(1, 0), hcmap(1, 0), hcoarse(1, 0);
  while (i32(L.size()) < maxlv) {
    Level& f = L.back();
    if (!(f.n > floor_n)) break;
    gk_csr fv = f.view();
    std::vector<i32> cand(f.n), label(f.n), stamp(f.n), order(f.n), cmap(f.n), cvw(f.n);
    oracle_match(&fv, cand.data());
    i32 sw = oracle_group(f.n, cand.data(), label.data(), stamp.data());
    if (sw < 0) return GK_ERR_GRAPH;
    i32 nc = oracle_cluster(f.n, label.data(), stamp.data(), f.vw.data(), p->max_group,
                            order.data(), cmap.data(), cvw.data());
    if (nc == f.n || nc < k + 1) break;
    if (p->coarsen_policy != GK_COARSEN_REF && nc < floor_n) break;
    Level c;
    c.n = nc;
    c.xadj.resize(nc + 1);
    c.adj.resize(fv.nnz);
    c.w.resize(fv.nnz);
    i32 e = oracle_contract(&fv, cmap.data(), nc, c.xadj.data(), c.adj.data(), c.w.data());
    c.adj.resize(e);
    c.w.resize(e);
    c.vw.assign(cvw.begin(), cvw.begin() + nc);
    f.cmap = cmap;
    if (trace) {
      hcand.back() = fnv(cand.data(), f.n, kFnvBasis);
      hcmap.back() = fnv(cmap.data(), f.n, kFnvBasis);
      uint64_t h = fnv(c.xadj.data(), nc + 1, kFnvBasis);
      h = fnv(c.adj.data(), e, h);
      h = fnv(c.w.data(), e, h);
      hcoarse.back() = fnv(c.vw.data(), nc, h);
    }
    sweeps_of.back() = sw;
    L.push_back(std::move(c));
    sweeps_of.push_back(0);
    hcand.push_back(0);
    hcmap.push_back(0);
    hcoarse.push_back(0);
  }
  const i32 nl = i32(L.size());
  auto t1 = clk::now();

  // ---- initial partition of the coarsest level
  Level& top = L.back();
  gk_csr tv = top.view();
  std::vector<i32> part(top.n);
  i64 cut = 0;
  int rc = initpart(&tv, k, cap, p->seed, p->initpart_threads, part.data(), &cut);
  if (rc != GK_OK) return rc;
  S.init_cut = cut;
  std::vector<i32> bnd(top.n, 0);
  for (i32 v = 0; v < top.n; ++v)
    for (i32 j = top.xadj[v]; j < top.xadj[v + 1]; ++j)
      if (part[top.adj[j]] != part[v]) { bnd[v] = 1; break; }
  std::vector<i32> pw(k, 0);
  for (i32 v = 0; v < top.n; ++v) pw[part[v]] += top.vw[v];
  auto t2 = clk::now();

  std::vector<i32> rounds_of(nl, 0);
  std::vector<i64> moves_of(nl, 0), cut_of(nl, 0);
  std::vector<uint64_t> hpart(nl, 0);
  cut_of[nl - 1] = cut;
  if (trace) hpart[nl - 1] = fnv(part.data(), top.n, kFnvBasis);

  // ---- uncoarsening + refinement
  for (i32 l = nl - 2; l >= 0; --l) {
    Level& f = L[l];
    gk_csr fv = f.view();
    std::vector<i32> fpart(f.n), fbnd(f.n), gain(f.n), dest(f.n), upd(f.n);
    oracle_project(f.n, f.cmap.data(), part.data(), bnd.data(), fpart.data(), fbnd.data(),
                   gain.data(), dest.data(), upd.data());
    part.swap(fpart);
    bnd.swap(fbnd);
    oracle_gain(&fv, k, part.data(), upd.data(), bnd.data(), gain.data(), dest.data());
    std::vector<i32> moves(std::max(1, p->max_moves));
    while (true) {
      i32 nsel = 0;
      i32 B = oracle_select(&fv, part.data(), bnd.data(), gain.data(), dest.data(), pw.data(), cap,
                            p->max_moves, upd.data(), moves.data(), &nsel);
      if (B == 0) break;
      ++rounds_of[l];
      i32 pos = oracle_balance(&fv, k, moves.data(), B, part.data(), dest.data(), pw.data(), cap);
      i64 gs = oracle_apply(&fv, moves.data(), pos, gain.data(), dest.data(), part.data(), upd.data());
      cut -= gs;
      moves_of[l] += pos + 1;
      oracle_gain(&fv, k, part.data(), upd.data(), bnd.data(), gain.data(), dest.data());
    }
    cut_of[l] = cut;
    if (trace) hpart[l] = fnv(part.data(), f.n, kFnvBasis);
  }
  auto t3 = clk::now();

  // ---- final metrics
  gk_csr g0v = L[0].view();
  std::vector<i64> pw64(k);
  i64 final_cut = 0;
  oracle_cut(&g0v, k, part.data(), &final_cut, pw64.data());
  std::copy(part.begin(), part.end(), part_out);
  S.nlevels = nl;
  S.coarsest_n = top.n;
  S.cut = final_cut;
  S.cut_tracked = cut;
  i64 mx = 0;
  for (i64 x : pw64) mx = std::max(mx, x);
  S.max_part_wgt = mx;
  S.imbalance = T > 0 ? double(mx) * k / double(T) : 1.0;
  for (i32 l = 0; l < nl && l < GK_MAX_LEVELS; ++l) {
    gk_level_stats& ls = S.levels[l];
    ls.n = L[l].n;
    ls.nnz = L[l].xadj[L[l].n];
    ls.sweeps = sweeps_of[l];
    ls.rounds = rounds_of[l];
    ls.moves = moves_of[l];
    ls.cut_after = cut_of[l];
    ls.h_cand = hcand[l];
    ls.h_cmap = hcmap[l];
    ls.h_coarse = hcoarse[l];
    ls.h_part = hpart[l];
  }
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  S.t_coarsen_ms = ms(t0, t1);
  S.t_init_ms = ms(t1, t2);
  S.t_refine_ms = ms(t2, t3);
  S.t_total_ms = ms(t0, clk::now());
  if (st) *st = S;
  return GK_OK;
}

}  // extern "C"
