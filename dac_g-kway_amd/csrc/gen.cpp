// Seeded workload generators (host C++): Delaunay meshes, road-like graphs,
// layered circuit-like DAGs.  Exposed through include/gkgen.h.
//
// Delaunay: incremental Bowyer-Watson over a triangulation closed by one
// symbolic vertex "at infinity" (ghost triangles), so no bounding triangle
// perturbs the hull.  Points sit on a 2^24 integer grid; orient uses int64 and
// incircle __int128, both exact, so every conflict decision is exact.  Points
// are inserted in Hilbert-curve order and located by a visibility walk from
// the previous insertion, which keeps walks O(1) for uniform points.
#include "../../include/gkgen.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace {

// ---------------------------------------------------------------- PRNG
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, m)
  uint64_t below(uint64_t m) { return static_cast<uint64_t>((static_cast<unsigned __int128>(next()) * m) >> 64); }
  double unit() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

constexpr int kGridBits = 24;

// ---------------------------------------------------------------- points
void make_points(int32_t n, uint64_t seed, std::vector<int32_t>& x, std::vector<int32_t>& y) {
  SplitMix64 rng(seed * 0x2545F4914F6CDD1Dull + 0x1234567ull);
  x.resize(n);
  y.resize(n);
  for (int32_t i = 0; i < n; ++i) {
    x[i] = static_cast<int32_t>(rng.next() >> (64 - kGridBits));
    y[i] = static_cast<int32_t>(rng.next() >> (64 - kGridBits));
  }
  // Redraw duplicates (keep the first occurrence) until all points differ.
  while (true) {
    std::vector<std::pair<uint64_t, int32_t>> key(n);
    for (int32_t i = 0; i < n; ++i) key[i] = {(uint64_t(uint32_t(x[i])) << 32) | uint32_t(y[i]), i};
    std::sort(key.begin(), key.end());
    bool dup = false;
    for (int32_t i = 1; i < n; ++i) {
      if (key[i].first == key[i - 1].first) {
        int32_t j = key[i].second;  // later index (pairs sorted by key then index)
        x[j] = static_cast<int32_t>(rng.next() >> (64 - kGridBits));
        y[j] = static_cast<int32_t>(rng.next() >> (64 - kGridBits));
        dup = true;
      }
    }
    if (!dup) break;
  }
}

// ---------------------------------------------------------------- Hilbert key
uint64_t hilbert_key(uint32_t x, uint32_t y, int bits) {
  uint64_t d = 0;
  for (uint32_t s = 1u << (bits - 1); s > 0; s >>= 1) {
    uint32_t rx = (x & s) ? 1 : 0;
    uint32_t ry = (y & s) ? 1 : 0;
    d += uint64_t(s) * uint64_t(s) * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - (x & (s - 1)) + (x & ~(s - 1));
        y = s - 1 - (y & (s - 1)) + (y & ~(s - 1));
        x &= (s << 1) - 1;
        y &= (s << 1) - 1;
      }
      std::swap(x, y);
    }
  }
  return d;
}

// ---------------------------------------------------------------- Delaunay
constexpr int32_t GHOST = -1;

struct Tri {
  int32_t v[3];   // CCW; a ghost triangle keeps the ghost vertex in v[2]
  int32_t nb[3];  // nb[i] = triangle across the edge opposite v[i]
};

class Delaunay {
 public:
  Delaunay(const int32_t* x, const int32_t* y, int32_t n) : X(x), Y(y), n_(n) {}

  // Returns undirected edges (u < v).
  std::vector<std::pair<int32_t, int32_t>> run();

 private:
  const int32_t* X;
  const int32_t* Y;
  int32_t n_;
  std::vector<int32_t> LX, LY;
  std::vector<Tri> T;
  std::vector<int32_t> freelist;
  std::vector<uint32_t> mark;
  uint32_t epoch = 0;
  int32_t last = 0;

  int64_t orient(int32_t a, int32_t b, int32_t c) const {
    int64_t abx = int64_t(X[b]) - X[a], aby = int64_t(Y[b]) - Y[a];
    int64_t acx = int64_t(X[c]) - X[a], acy = int64_t(Y[c]) - Y[a];
    return abx * acy - aby * acx;
  }
  // > 0 when d lies strictly inside the circumcircle of CCW triangle (a,b,c).
  int incircle(int32_t a, int32_t b, int32_t c, int32_t d) const {
    __int128 adx = int64_t(X[a]) - X[d], ady = int64_t(Y[a]) - Y[d];
    __int128 bdx = int64_t(X[b]) - X[d], bdy = int64_t(Y[b]) - Y[d];
    __int128 cdx = int64_t(X[c]) - X[d], cdy = int64_t(Y[c]) - Y[d];
    __int128 ad = adx * adx + ady * ady;
    __int128 bd = bdx * bdx + bdy * bdy;
    __int128 cd = cdx * cdx + cdy * cdy;
    __int128 det = ad * (bdx * cdy - cdx * bdy) + bd * (cdx * ady - adx * cdy) + cd * (adx * bdy - bdx * ady);
    return det > 0 ? 1 : (det < 0 ? -1 : 0);
  }
  bool conflict(int32_t t, int32_t p) const {
    const Tri& r = T[t];
    if (r.v[2] != GHOST) return incircle(r.v[0], r.v[1], r.v[2], p) > 0;
    // Ghost (a, b, inf): the open half plane left of a->b, plus the open
    // segment ab itself.
    int32_t a = r.v[0], b = r.v[1];
    int64_t o = orient(a, b, p);
    if (o > 0) return true;
    if (o < 0) return false;
    int64_t dx = int64_t(X[b]) - X[a], dy = int64_t(Y[b]) - Y[a];
    int64_t t1 = (int64_t(X[p]) - X[a]) * dx + (int64_t(Y[p]) - Y[a]) * dy;
    int64_t t2 = dx * dx + dy * dy;
    return t1 > 0 && t1 < t2;
  }
  int32_t alloc_tri() {
    if (!freelist.empty()) {
      int32_t t = freelist.back();
      freelist.pop_back();
      return t;
    }
    T.push_back(Tri{});
    mark.push_back(0);
    return static_cast<int32_t>(T.size() - 1);
  }
  int32_t locate(int32_t p);
  void insert(int32_t p);
  // scratch
  std::vector<int32_t> cavity, stack;
  struct BEdge { int32_t a, b, out, newt; };
  std::vector<BEdge> bedges;
};

int32_t Delaunay::locate(int32_t p) {
  int32_t t = last;
  if (T[t].v[2] == GHOST) t = T[t].nb[2];
  // Visibility walk over finite triangles; stepping into a ghost triangle
  // means p lies outside the current hull beyond that edge.
  int64_t guard = 0;
  while (true) {
    const Tri& r = T[t];
    if (r.v[2] == GHOST) return t;
    bool moved = false;
    for (int i = 0; i < 3; ++i) {
      int32_t a = r.v[(i + 1) % 3], b = r.v[(i + 2) % 3];
      if (orient(a, b, p) < 0) {
        t = r.nb[i];
        moved = true;
        break;
      }
    }
    if (!moved) return t;
    if (++guard > 4 * int64_t(T.size()) + 16) throw std::runtime_error("delaunay: walk did not terminate");
  }
}

void Delaunay::insert(int32_t p) {
  int32_t t0 = locate(p);
  if (!conflict(t0, p)) {
    // p lies on an edge of a finite triangle that is also a hull edge and the
    // walk stopped in the finite triangle: it always conflicts; a ghost start
    // that does not conflict cannot happen for a proper walk.
    throw std::runtime_error("delaunay: start triangle not in conflict");
  }
  ++epoch;
  cavity.clear();
  bedges.clear();
  stack.clear();
  stack.push_back(t0);
  mark[t0] = epoch;
  while (!stack.empty()) {
    int32_t t = stack.back();
    stack.pop_back();
    cavity.push_back(t);
    for (int i = 0; i < 3; ++i) {
      int32_t u = T[t].nb[i];
      if (mark[u] == epoch) continue;
      if (conflict(u, p)) {
        mark[u] = epoch;
        stack.push_back(u);
      } else {
        bedges.push_back(BEdge{T[t].v[(i + 1) % 3], T[t].v[(i + 2) % 3], u, -1});
      }
    }
  }
  // New triangles (a, b, p) for each boundary edge a->b (cavity on its left).
  for (int32_t t : cavity) freelist.push_back(t);
  // Map: start vertex of boundary edge -> index in bedges.  The boundary is a
  // simple cycle around p, so every vertex (ghost included) starts at most one
  // edge.  Small linear maps are fine: the cavity has O(1) edges on average.
  const size_t nb = bedges.size();
  for (size_t i = 0; i < nb; ++i) bedges[i].newt = alloc_tri();
  for (size_t i = 0; i < nb; ++i) {
    BEdge& e = bedges[i];
    Tri& r = T[e.newt];
    // Find the edge starting at e.b (shares edge b->p) and the edge ending at
    // e.a (shares edge p->a).
    int32_t next = -1, prev = -1;
    for (size_t j = 0; j < nb; ++j) {
      if (bedges[j].a == e.b) next = bedges[j].newt;
      if (bedges[j].b == e.a) prev = bedges[j].newt;
    }
    if (next < 0 || prev < 0) throw std::runtime_error("delaunay: cavity boundary is not a cycle");
    // Triangle (a, b, p): nb opposite a = across (b, p) = next; opposite b =
    // across (p, a) = prev; opposite p = across (a, b) = out.
    int32_t va = e.a, vb = e.b;
    if (va == GHOST) {
      // (G, b, p) -> rotate to (b, p, G)
      r.v[0] = vb; r.v[1] = p; r.v[2] = GHOST;
      r.nb[0] = prev;   // opposite b: across (p, G) ... edge (p,a=G) -> prev
      r.nb[1] = e.out;  // opposite p: across (G, b)  -> wait: edge (a,b)=(G,b) -> out
      r.nb[2] = next;   // opposite G: across (b, p) -> next
    } else if (vb == GHOST) {
      // (a, G, p) -> rotate to (p, a, G)
      r.v[0] = p; r.v[1] = va; r.v[2] = GHOST;
      r.nb[0] = e.out;  // opposite p: across (a, G) -> out
      r.nb[1] = next;   // opposite a: across (G, p) -> next
      r.nb[2] = prev;   // opposite G: across (p, a) -> prev
    } else {
      r.v[0] = va; r.v[1] = vb; r.v[2] = p;
      r.nb[0] = next;
      r.nb[1] = prev;
      r.nb[2] = e.out;
    }
    // Re-point the outside neighbour at the new triangle.
    Tri& o = T[e.out];
    for (int k = 0; k < 3; ++k) {
      int32_t a2 = o.v[(k + 1) % 3], b2 = o.v[(k + 2) % 3];
      if (a2 == e.b && b2 == e.a) { o.nb[k] = e.newt; break; }
    }
    mark[e.newt] = 0;
    if (r.v[2] != GHOST) last = e.newt;
  }
}

std::vector<std::pair<int32_t, int32_t>> Delaunay::run() {
  std::vector<std::pair<int32_t, int32_t>> edges;
  if (n_ < 2) return edges;
  // Insertion order: Hilbert curve over the grid.
  std::vector<std::pair<uint64_t, int32_t>> ord(n_);
  for (int32_t i = 0; i < n_; ++i) ord[i] = {hilbert_key(uint32_t(X[i]), uint32_t(Y[i]), kGridBits), i};
  std::sort(ord.begin(), ord.end());
  // Work on coordinates stored in insertion order (cache locality of the
  // walk and the predicates); map ids back when edges are emitted.
  std::vector<int32_t> orig(n_), seq(n_);
  LX.resize(n_);
  LY.resize(n_);
  for (int32_t i = 0; i < n_; ++i) {
    orig[i] = ord[i].second;
    LX[i] = X[orig[i]];
    LY[i] = Y[orig[i]];
    seq[i] = i;
  }
  X = LX.data();
  Y = LY.data();
  // First triangle: seq[0], seq[1] and the first point not collinear with them.
  int32_t a = seq[0], b = seq[1], c = -1, ci = -1;
  for (int32_t i = 2; i < n_; ++i) {
    if (orient(a, b, seq[i]) != 0) { c = seq[i]; ci = i; break; }
  }
  if (c < 0) {
    // All points collinear: the triangulation is the sorted chain.
    std::vector<int32_t> s(seq);
    std::sort(s.begin(), s.end(), [&](int32_t p, int32_t q) {
      return X[p] != X[q] ? X[p] < X[q] : Y[p] < Y[q];
    });
    for (int32_t i = 1; i < n_; ++i) {
      int32_t p = orig[s[i - 1]], q = orig[s[i]];
      edges.emplace_back(std::min(p, q), std::max(p, q));
    }
    return edges;
  }
  if (orient(a, b, c) < 0) std::swap(a, b);
  T.reserve(size_t(n_) * 2 + 16);
  mark.reserve(size_t(n_) * 2 + 16);
  // t0 = (a,b,c); ghosts: across a->b: (b,a,G); across b->c: (c,b,G); across c->a: (a,c,G)
  for (int i = 0; i < 4; ++i) { T.push_back(Tri{}); mark.push_back(0); }
  T[0] = Tri{{a, b, c}, {2, 3, 1}};
  T[1] = Tri{{b, a, GHOST}, {3, 2, 0}};  // nb0: across (a,G) -> (a,c,G)=3? see below
  T[2] = Tri{{c, b, GHOST}, {1, 3, 0}};
  T[3] = Tri{{a, c, GHOST}, {2, 1, 0}};
  // Fix ghost adjacency by edge matching to stay independent of hand wiring.
  for (int t = 1; t <= 3; ++t) {
    for (int i = 0; i < 3; ++i) {
      int32_t ea = T[t].v[(i + 1) % 3], eb = T[t].v[(i + 2) % 3];
      for (int u = 0; u <= 3; ++u) {
        if (u == t) continue;
        for (int k = 0; k < 3; ++k) {
          if (T[u].v[(k + 1) % 3] == eb && T[u].v[(k + 2) % 3] == ea) T[t].nb[i] = u;
        }
      }
    }
  }
  for (int i = 0; i < 3; ++i) {
    int32_t ea = T[0].v[(i + 1) % 3], eb = T[0].v[(i + 2) % 3];
    for (int u = 1; u <= 3; ++u)
      for (int k = 0; k < 3; ++k)
        if (T[u].v[(k + 1) % 3] == eb && T[u].v[(k + 2) % 3] == ea) T[0].nb[i] = u;
  }
  last = 0;
  for (int32_t i = 2; i < n_; ++i) {
    if (i == ci) continue;
    insert(seq[i]);
  }
  // Collect every finite directed edge u->v with u < v; each undirected edge
  // appears exactly once this way (its two directions live in the two
  // triangles sharing it, ghost triangles included).
  std::vector<char> live(T.size(), 1);
  for (int32_t t : freelist) live[t] = 0;
  edges.reserve(size_t(n_) * 3);
  for (size_t t = 0; t < T.size(); ++t) {
    if (!live[t]) continue;
    for (int i = 0; i < 3; ++i) {
      int32_t u = T[t].v[i], v = T[t].v[(i + 1) % 3];
      if (u == GHOST || v == GHOST) continue;
      int32_t ou = orig[u], ov = orig[v];
      if (ou < ov) edges.emplace_back(ou, ov);
    }
  }
  return edges;
}

// ---------------------------------------------------------------- CSR build
int to_csr(int32_t n, std::vector<std::pair<int32_t, int32_t>>& edges,
           const std::vector<int32_t>* ew, const std::vector<int32_t>* vw, gkgen_graph* out) {
  int64_t m = static_cast<int64_t>(edges.size());
  if (2 * m >= (int64_t(1) << 31)) return -7;
  std::vector<int32_t> deg(n + 1, 0);
  for (auto& e : edges) { deg[e.first]++; deg[e.second]++; }
  int32_t* xadj = static_cast<int32_t*>(malloc(sizeof(int32_t) * (n + 1)));
  int32_t* adj = static_cast<int32_t*>(malloc(sizeof(int32_t) * std::max<int64_t>(1, 2 * m)));
  int32_t* aw = ew ? static_cast<int32_t*>(malloc(sizeof(int32_t) * std::max<int64_t>(1, 2 * m))) : nullptr;
  int32_t* vv = vw ? static_cast<int32_t*>(malloc(sizeof(int32_t) * std::max(1, n))) : nullptr;
  if (!xadj || !adj || (ew && !aw) || (vw && !vv)) {
    free(xadj); free(adj); free(aw); free(vv);
    return -6;
  }
  xadj[0] = 0;
  for (int32_t v = 0; v < n; ++v) xadj[v + 1] = xadj[v] + deg[v];
  std::vector<int32_t> fill(xadj, xadj + n);
  for (size_t i = 0; i < edges.size(); ++i) {
    int32_t u = edges[i].first, v = edges[i].second;
    int32_t w = ew ? (*ew)[i] : 1;
    adj[fill[u]] = v; if (aw) aw[fill[u]] = w; fill[u]++;
    adj[fill[v]] = u; if (aw) aw[fill[v]] = w; fill[v]++;
  }
  // Rows sorted by neighbour id (weights follow).
  std::vector<std::pair<int32_t, int32_t>> row;
  for (int32_t v = 0; v < n; ++v) {
    int32_t s = xadj[v], e = xadj[v + 1];
    row.clear();
    for (int32_t j = s; j < e; ++j) row.emplace_back(adj[j], aw ? aw[j] : 1);
    std::sort(row.begin(), row.end());
    for (int32_t j = s; j < e; ++j) { adj[j] = row[j - s].first; if (aw) aw[j] = row[j - s].second; }
  }
  if (vv) for (int32_t v = 0; v < n; ++v) vv[v] = (*vw)[v];
  out->n = n;
  out->nnz = static_cast<int32_t>(2 * m);
  out->xadj = xadj;
  out->adjncy = adj;
  out->adjwgt = aw;
  out->vwgt = vv;
  return 0;
}

}  // namespace

extern "C" {

int gkgen_points(int32_t n, uint64_t seed, int32_t* x, int32_t* y) {
  if (n < 0 || !x || !y) return -1;
  std::vector<int32_t> X, Y;
  make_points(n, seed, X, Y);
  std::copy(X.begin(), X.end(), x);
  std::copy(Y.begin(), Y.end(), y);
  return 0;
}

int gkgen_delaunay_points(int32_t n, const int32_t* x, const int32_t* y, gkgen_graph* out) {
  if (n < 0 || !out || (n > 0 && (!x || !y))) return -1;
  try {
    Delaunay d(x, y, n);
    auto edges = d.run();
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    return to_csr(n, edges, nullptr, nullptr, out);
  } catch (const std::exception&) {
    return -8;
  }
}

int gkgen_delaunay(int32_t n, uint64_t seed, gkgen_graph* out) {
  if (n < 0 || !out) return -1;
  std::vector<int32_t> X, Y;
  make_points(n, seed, X, Y);
  return gkgen_delaunay_points(n, X.data(), Y.data(), out);
}

int gkgen_road(int32_t n, uint64_t seed, int32_t avg_deg_x100, gkgen_graph* out) {
  if (n < 2 || !out || avg_deg_x100 < 200) return -1;
  std::vector<int32_t> X, Y;
  make_points(n, seed, X, Y);
  std::vector<std::pair<int32_t, int32_t>> edges;
  try {
    Delaunay d(X.data(), Y.data(), n);
    edges = d.run();
  } catch (const std::exception&) {
    return -8;
  }
  // Kruskal over Delaunay edges by (squared length, u, v): the EMST is a
  // subgraph of the Delaunay triangulation.
  struct E { int64_t len; int32_t u, v; };
  std::vector<E> es(edges.size());
  for (size_t i = 0; i < edges.size(); ++i) {
    int32_t u = edges[i].first, v = edges[i].second;
    int64_t dx = int64_t(X[u]) - X[v], dy = int64_t(Y[u]) - Y[v];
    es[i] = E{dx * dx + dy * dy, u, v};
  }
  std::sort(es.begin(), es.end(), [](const E& a, const E& b) {
    if (a.len != b.len) return a.len < b.len;
    if (a.u != b.u) return a.u < b.u;
    return a.v < b.v;
  });
  std::vector<int32_t> uf(n);
  std::iota(uf.begin(), uf.end(), 0);
  auto find = [&](int32_t a) {
    while (uf[a] != a) { uf[a] = uf[uf[a]]; a = uf[a]; }
    return a;
  };
  const int64_t target_m = (int64_t(n) * avg_deg_x100 + 100) / 200;  // round(n * deg / 2)
  std::vector<char> used(es.size(), 0);
  std::vector<std::pair<int32_t, int32_t>> out_e;
  out_e.reserve(target_m);
  for (size_t i = 0; i < es.size(); ++i) {
    int32_t a = find(es[i].u), b = find(es[i].v);
    if (a != b) { uf[a] = b; used[i] = 1; out_e.emplace_back(es[i].u, es[i].v); }
  }
  for (size_t i = 0; i < es.size() && int64_t(out_e.size()) < target_m; ++i)
    if (!used[i]) out_e.emplace_back(es[i].u, es[i].v);
  std::sort(out_e.begin(), out_e.end());
  return to_csr(n, out_e, nullptr, nullptr, out);
}

int gkgen_dag(int32_t n, int32_t layers, uint64_t seed, int32_t max_fanout, double exponent,
              int32_t window, int32_t wmax, int32_t vwmax, gkgen_graph* out) {
  if (n < 2 || layers < 2 || layers > n || max_fanout < 1 || window < 1 || !out) return -1;
  SplitMix64 rng(seed * 0x9E3779B97F4A7C15ull + 0xDA6ull);
  // Layer l holds vertices [lo[l], lo[l+1]).
  std::vector<int64_t> lo(layers + 1);
  for (int32_t l = 0; l <= layers; ++l) lo[l] = int64_t(n) * l / layers;
  // Fanout CDF, P(f) ~ f^-exponent.
  std::vector<double> cdf(max_fanout);
  double acc = 0;
  for (int32_t f = 1; f <= max_fanout; ++f) { acc += std::pow(double(f), -exponent); cdf[f - 1] = acc; }
  for (auto& c : cdf) c /= acc;
  std::vector<std::pair<int32_t, int32_t>> edges;
  edges.reserve(size_t(n) * 2);
  for (int32_t l = 0; l + 1 < layers; ++l) {
    for (int64_t v = lo[l]; v < lo[l + 1]; ++v) {
      double u = rng.unit();
      int32_t f = int32_t(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
      if (f > max_fanout) f = max_fanout;
      for (int32_t j = 0; j < f; ++j) {
        int32_t d = 1 + int32_t(rng.below(window));
        int32_t tl = std::min(l + d, layers - 1);
        int64_t sz = lo[tl + 1] - lo[tl];
        int32_t t = int32_t(lo[tl] + int64_t(rng.below(uint64_t(sz))));
        edges.emplace_back(int32_t(v), t);  // v < t since tl > l
      }
    }
  }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  std::vector<int32_t> ew, vw;
  if (wmax > 1) {
    ew.resize(edges.size());
    for (size_t i = 0; i < edges.size(); ++i) {
      uint64_t h = mix64(seed ^ mix64((uint64_t(uint32_t(edges[i].first)) << 32) | uint32_t(edges[i].second)));
      ew[i] = 1 + int32_t(h % uint64_t(wmax));
    }
  }
  if (vwmax > 1) {
    vw.resize(n);
    SplitMix64 vr(seed * 31 + 7);
    for (int32_t v = 0; v < n; ++v) vw[v] = 1 + int32_t(vr.below(uint64_t(vwmax)));
  }
  return to_csr(n, edges, wmax > 1 ? &ew : nullptr, vwmax > 1 ? &vw : nullptr, out);
}

void gkgen_free(gkgen_graph* g) {
  if (!g) return;
  free(g->xadj); free(g->adjncy); free(g->adjwgt); free(g->vwgt);
  g->xadj = g->adjncy = g->adjwgt = g->vwgt = nullptr;
  g->n = g->nnz = 0;
}

uint64_t gkgen_fnv1a_i32(const int32_t* a, int64_t n, uint64_t h) {
  for (int64_t i = 0; i < n; ++i) { h ^= uint32_t(a[i]); h *= 0x100000001b3ull; }
  return h;
}

}  // extern "C"
