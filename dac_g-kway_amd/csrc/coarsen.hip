// Coarsening kernels for gfx950: heavy-edge candidate matching, grouping of
// the candidate pseudo-forest by pointer jumping, chunking into coarse
// vertices, and CSR contraction (relabel, per-row sort, merge of parallel
// edges) -- the work of gkway/coarsen.hpp.  Every kernel is integer-exact and
// writes results that do not depend on the launch schedule: atomics are used
// only for commutative integer max/add or for slot reservation whose order is
// discarded by a later sort.
#include "kernels.hpp"
#include "prim.hpp"
#include "steal.hpp"
#include "../../include/gkway.h"

namespace gkd {

// About one lane per two to four entries: measured on MI355X, 4 lanes beat 8
// on meshes of average degree 4.5-6 (up to 24% at 2^24 vertices), 8 beat 16
// at degree 7-16 and 16 beat 32 at degree 16-22.
int lanes_for_degree(double avg) {
  if (avg <= 8.0) return 4;
  if (avg <= 16.0) return 8;
  if (avg <= 32.0) return 16;
  if (avg <= 64.0) return 32;
  return 64;
}

static inline int hip_status() { return hipGetLastError() == hipSuccess ? 0 : GK_ERR_HIP; }

// --------------------------------------------------------------- matching
// matching_with_score (gkway/coarsen.hpp:39-78) with S lanes per vertex: the
// lanes stride over the row, then reduce the packed preference
// (weight desc, neighbour degree asc, neighbour id asc) by xor-shuffles.
// Global thread slot gt serves vertex gt / S as lane gt % S of its group.
template <int S>
__device__ __forceinline__ void match_slot(int64_t gt, int32_t n, const int32_t* __restrict__ xadj,
                                           const int32_t* __restrict__ adj, const int32_t* __restrict__ w,
                                           int32_t* __restrict__ cand) {
  const int32_t v = static_cast<int32_t>(gt / S);
  const int sub = static_cast<int>(gt % S);
  if (v >= n) return;
  int bw = 0, bd = 0, bu = v;
  const int s = xadj[v], e = xadj[v + 1];
  for (int j = s + sub; j < e; j += S) {
    const int u = adj[j];
    const int wj = w[j];
    const int d = xadj[u + 1] - xadj[u];
    if (wj > bw || (wj == bw && (d < bd || (d == bd && u < bu)))) { bw = wj; bd = d; bu = u; }
  }
#pragma unroll
  for (int o = S / 2; o > 0; o >>= 1) {
    const int ow = __shfl_xor(bw, o, S), od = __shfl_xor(bd, o, S), ou = __shfl_xor(bu, o, S);
    if (ow > bw || (ow == bw && (od < bd || (od == bd && ou < bu)))) { bw = ow; bd = od; bu = ou; }
  }
  if (sub == 0) cand[v] = bu;
}

template <int S, int B>
__global__ void __launch_bounds__(B) k_match(int32_t n, const int32_t* __restrict__ xadj,
                                             const int32_t* __restrict__ adj,
                                             const int32_t* __restrict__ w, int32_t* __restrict__ cand) {
  match_slot<S>(int64_t(xcd_block()) * B + threadIdx.x, n, xadj, adj, w, cand);
}

// The same slots, taken from the work-stealing queues.
template <int S, int B>
__global__ void __launch_bounds__(B) k_match_ws(int32_t n, const int32_t* __restrict__ xadj,
                                                const int32_t* __restrict__ adj,
                                                const int32_t* __restrict__ w, int32_t* __restrict__ cand,
                                                unsigned long long* __restrict__ ctr) {
  const int64_t sweeps = (int64_t(n) * S + 63) / 64;
  steal_sweeps(ctr, sweeps, [&](int64_t sw) { match_slot<S>(sw * 64 + lane_id(), n, xadj, adj, w, cand); });
}

template <int S, int B>
static void match_go(const DGraph& g, int32_t* cand, const Steal* ws, hipStream_t st) {
  const unsigned grid = grid_for(int64_t(g.n) * S, B);
  if (!ws) {
    gk_launch(k_match<S, B>, dim3(grid), dim3(B), 0, st, g.n, g.xadj, g.adj, g.w, cand);
    return;
  }
  const unsigned pg = steal_grid(reinterpret_cast<const void*>(&k_match_ws<S, B>), B, *ws, grid);
  unsigned long long* ctr = nullptr;  // queues cleared and used only when the work has a tail
  if (steal_tail((int64_t(g.n) * S + 63) / 64, int64_t(pg) * (B / 64)) > 0) {
    fill<unsigned long long>(ws->ctr, 0ull, kStealWords, st);
    ctr = ws->ctr;
  }
  gk_launch(k_match_ws<S, B>, dim3(pg), dim3(B), 0, st, g.n, g.xadj, g.adj, g.w, cand, ctr);
}

template <int B>
static void match_block(const DGraph& g, int lanes, int32_t* cand, const Steal* ws, hipStream_t st) {
  switch (lanes) {
    case 1: match_go<1, B>(g, cand, ws, st); break;
    case 2: match_go<2, B>(g, cand, ws, st); break;
    case 4: match_go<4, B>(g, cand, ws, st); break;
    case 8: match_go<8, B>(g, cand, ws, st); break;
    case 16: match_go<16, B>(g, cand, ws, st); break;
    case 32: match_go<32, B>(g, cand, ws, st); break;
    default: match_go<64, B>(g, cand, ws, st); break;
  }
}

int launch_match(const DGraph& g, int lanes, int block, int32_t* cand, hipStream_t st, const Steal* ws) {
  if (g.n <= 0) return 0;
  switch (block) {
    case 128: match_block<128>(g, lanes, cand, ws, st); break;
    case 512: match_block<512>(g, lanes, cand, ws, st); break;
    default: match_block<256>(g, lanes, cand, ws, st); break;
  }
  return hip_status();
}

// --------------------------------------------------------------- grouping
// The candidate graph {v, cand[v]} is a pseudo-forest whose only cycles are
// mutual pairs (or a self loop for an isolated vertex).  Pointer jumping
// finds for every vertex its root pair and the hop distance to it.
__global__ void __launch_bounds__(256) k_jump_init(int32_t n, const int32_t* __restrict__ cand,
                                                   int32_t* __restrict__ next, int32_t* __restrict__ dist) {
  const int64_t v = int64_t(xcd_block()) * 256 + threadIdx.x;
  if (v >= n) return;
  const int32_t c = cand[v];
  const bool root = cand[c] == v;
  next[v] = root ? int32_t(v) : c;
  dist[v] = root ? 0 : 1;
}

__global__ void __launch_bounds__(256) k_jump_round(int32_t n, const int32_t* __restrict__ next_in,
                                                    const int32_t* __restrict__ dist_in,
                                                    int32_t* __restrict__ next_out,
                                                    int32_t* __restrict__ dist_out,
                                                    int32_t* __restrict__ flags, int32_t round) {
  if (round > 0 && flags[round - 1] == 0) return;  // converged in an earlier round
  const int64_t v = int64_t(xcd_block()) * 256 + threadIdx.x;
  bool moved = false;
  if (v < n) {
    const int32_t nx = next_in[v];
    const int32_t nn = next_in[nx];
    if (nn == nx) {
      next_out[v] = nx;
      dist_out[v] = dist_in[v];
    } else {
      next_out[v] = nn;
      dist_out[v] = dist_in[v] + dist_in[nx];
      moved = true;
    }
  }
  if (ballot(moved) != 0 && lane_id() == 0) flags[round] = 1;
}

int launch_jump_init(int32_t n, const int32_t* cand, int32_t* next, int32_t* dist, hipStream_t st) {
  if (n <= 0) return 0;
  k_jump_init<<<grid_for(n, 256), 256, 0, st>>>(n, cand, next, dist);
  return hip_status();
}

int launch_jump_round(int32_t n, const int32_t* cand, const int32_t* next_in, const int32_t* dist_in,
                      int32_t* next_out, int32_t* dist_out, int32_t* flags, int32_t round,
                      hipStream_t st) {
  (void)cand;
  if (n <= 0) return 0;
  k_jump_round<<<grid_for(n, 256), 256, 0, st>>>(n, next_in, dist_in, next_out, dist_out, flags, round);
  return hip_status();
}

__global__ void __launch_bounds__(256) k_label_max(int32_t n, const int32_t* __restrict__ cand,
                                                   const int32_t* __restrict__ next,
                                                   int32_t* __restrict__ compmax) {
  const int64_t v = int64_t(xcd_block()) * 256 + threadIdx.x;
  if (v >= n) return;
  const int32_t r = next[v];
  const int32_t key = max(r, cand[r]);
  atomicMax(&compmax[key], int32_t(v));
}

__global__ void __launch_bounds__(256) k_label_get(int32_t n, const int32_t* __restrict__ cand,
                                                   const int32_t* __restrict__ next,
                                                   const int32_t* __restrict__ compmax,
                                                   int32_t* __restrict__ label) {
  const int64_t v = int64_t(xcd_block()) * 256 + threadIdx.x;
  if (v >= n) return;
  const int32_t r = next[v];
  label[v] = compmax[max(r, cand[r])];
}

int launch_label(int32_t n, const int32_t* cand, const int32_t* next, int32_t* compmax,
                 int32_t* label, hipStream_t st) {
  if (n <= 0) return 0;
  fill<int32_t>(compmax, 0, n, st);
  k_label_max<<<grid_for(n, 256), 256, 0, st>>>(n, cand, next, compmax);
  k_label_get<<<grid_for(n, 256), 256, 0, st>>>(n, cand, next, compmax, label);
  return hip_status();
}

__global__ void __launch_bounds__(256) k_make_keys(int32_t n, const int32_t* __restrict__ label,
                                                   const int32_t* __restrict__ dist, int sbits,
                                                   uint64_t* __restrict__ keys, int32_t* __restrict__ ids) {
  const int64_t v = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (v >= n) return;
  keys[v] = (uint64_t(uint32_t(label[v])) << sbits) | uint64_t(uint32_t(dist[v]));
  ids[v] = int32_t(v);
}

int launch_make_keys(int32_t n, const int32_t* label, const int32_t* dist, int sbits,
                     uint64_t* keys, int32_t* ids, hipStream_t st) {
  if (n <= 0) return 0;
  k_make_keys<<<grid_for(n, 256), 256, 0, st>>>(n, label, dist, sbits, keys, ids);
  return hip_status();
}

// --------------------------------------------------------------- chunking
// constraint_group_size + coarse numbering (gkway/coarsen.hpp:361-421,
// 120-155): in the sorted order, an item starts a chunk when its rank in its
// group is a multiple of G; coarse ids are the running count of chunk starts.
__global__ void __launch_bounds__(256) k_group_start(int32_t n, const uint64_t* __restrict__ skeys, int sbits,
                                                     int32_t* __restrict__ gstart) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t lab = skeys[i] >> sbits;
  if (i == 0 || (skeys[i - 1] >> sbits) != lab) gstart[lab] = int32_t(i);
}

struct ChunkFlag {
  const uint64_t* skeys;
  const int32_t* gstart;
  int sbits;
  int32_t G;
  __device__ int operator()(int64_t i) const {
    const int32_t r = int32_t(i) - gstart[skeys[i] >> sbits];
    return (r % G) == 0 ? 1 : 0;
  }
};
struct StoreExcl {
  int32_t* out;
  __device__ void operator()(int64_t i, int excl, int) const { out[i] = excl; }
};

__global__ void __launch_bounds__(256) k_cmap(int32_t n, const uint64_t* __restrict__ skeys,
                                              const int32_t* __restrict__ sids, int sbits,
                                              const int32_t* __restrict__ gstart, int32_t G,
                                              const int32_t* __restrict__ excl,
                                              int32_t* __restrict__ cmap, int32_t* __restrict__ cptr) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int32_t r = int32_t(i) - gstart[skeys[i] >> sbits];
  const int f = (r % G) == 0;
  const int32_t c = f ? excl[i] : excl[i] - 1;
  cmap[sids[i]] = c;
  if (f) cptr[c] = int32_t(i);
  if (i == n - 1) cptr[c + 1] = n;
}

int launch_cluster(int32_t n, const uint64_t* skeys, const int32_t* sids, int sbits,
                   const int32_t* vwgt, int32_t G, int32_t* gstart, int32_t* flagscan,
                   int32_t* cmap, int32_t* cptr, int32_t* cvwgt, int32_t* nc_dev,
                   int32_t* scratch, hipStream_t st) {
  (void)vwgt;
  (void)cvwgt;
  if (n <= 0) return 0;
  k_group_start<<<grid_for(n, 256), 256, 0, st>>>(n, skeys, sbits, gstart);
  int rc = scan(n, ChunkFlag{skeys, gstart, sbits, G}, StoreExcl{flagscan}, nc_dev, scratch, st);
  if (rc) return GK_ERR_HIP;
  k_cmap<<<grid_for(n, 256), 256, 0, st>>>(n, skeys, sids, sbits, gstart, G, flagscan, cmap, cptr);
  return hip_status();
}

// cal_coarsen_vwgt (gkway/coarsen.hpp:157-171) as an ordered sum over the
// (at most G) members of each chunk; no atomics, destination fully written.
__global__ void __launch_bounds__(256) k_cvwgt(int32_t nc, const int32_t* __restrict__ cptr,
                                               const int32_t* __restrict__ order,
                                               const int32_t* __restrict__ vw, int32_t* __restrict__ cvw) {
  const int64_t c = int64_t(xcd_block()) * 256 + threadIdx.x;
  if (c >= nc) return;
  int32_t s = 0;
  for (int32_t j = cptr[c]; j < cptr[c + 1]; ++j) s += vw[order[j]];
  cvw[c] = s;
}

int launch_cvwgt(int32_t nc, const int32_t* cptr, const int32_t* order, const int32_t* vw,
                 int32_t* cvw, hipStream_t st) {
  if (nc <= 0) return 0;
  k_cvwgt<<<grid_for(nc, 256), 256, 0, st>>>(nc, cptr, order, vw, cvw);
  return hip_status();
}

// --------------------------------------------------------------- contraction
// Per coarse row: upper bound of entries (sum of member degrees) and its
// lane class; row ids appended to one list per class, each with its member
// range in order[] (desc: first member, member count).  Member m's row start
// and length are copied to mrow[m] on the way, so that the row kernels reach
// a member's entries with one load instead of order[] then xadj[].
__global__ void __launch_bounds__(256) k_row_bounds(int32_t nc, const int32_t* __restrict__ cptr,
                                                    const int32_t* __restrict__ order,
                                                    const int32_t* __restrict__ xadj,
                                                    int32_t* __restrict__ ub, int32_t* __restrict__ lists,
                                                    int2* __restrict__ desc, int2* __restrict__ mrow,
                                                    int32_t* __restrict__ cnt) {
  const int64_t c = int64_t(xcd_block()) * 256 + threadIdx.x;
  int cls = -1;
  int32_t m0 = 0, m1 = 0;
  if (c < nc) {
    m0 = cptr[c];
    m1 = cptr[c + 1];
    int32_t s = 0;
    for (int32_t j = m0; j < m1; ++j) {
      const int32_t v = order[j];
      const int32_t a = xadj[v], d = xadj[v + 1] - a;
      mrow[j] = make_int2(a, d);
      s += d;
    }
    ub[c] = s;
    cls = s <= 16 ? 0 : (s <= 32 ? 1 : (s <= 64 ? 2 : 3));
  }
  // Block-aggregated append: per class, wave counts in LDS, one global
  // atomic per block and class.
  __shared__ int wcnt[4][4];
  __shared__ int bbase[4];
  const int wave = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  uint64_t m[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    m[k] = ballot(cls == k);
    if (lane_id() == 0) wcnt[k][wave] = popc64(m[k]);
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = threadIdx.x;
    const int t = wcnt[k][0] + wcnt[k][1] + wcnt[k][2] + wcnt[k][3];
    bbase[k] = t ? atomicAdd(&cnt[k], t) : 0;
  }
  __syncthreads();
  if (cls >= 0) {
    int off = bbase[cls];
    for (int w = 0; w < wave; ++w) off += wcnt[cls][w];
    const uint64_t mine = cls == 0 ? m[0] : (cls == 1 ? m[1] : (cls == 2 ? m[2] : m[3]));
    const int64_t slot = int64_t(cls) * nc + off + popc64(mine & lt);
    lists[slot] = int32_t(c);
    desc[slot] = make_int2(m0, m1 - m0);
  }
}

struct UbLoad {
  const int32_t* ub;
  int32_t nc;
  __device__ int operator()(int64_t i) const { return i < nc ? ub[i] : 0; }
};

int launch_contract_bounds(const DGraph& g, int32_t nc, const int32_t* cptr, const int32_t* order,
                           int32_t* ub, int32_t* toff, int32_t* lists, int2* desc, int2* mrow,
                           int32_t* list_cnt, int32_t* scratch, hipStream_t st) {
  fill<int32_t>(list_cnt, 0, 4, st);
  if (nc <= 0) return 0;
  k_row_bounds<<<grid_for(nc, 256), 256, 0, st>>>(nc, cptr, order, g.xadj, ub, lists, desc, mrow, list_cnt);
  int rc = scan(int64_t(nc) + 1, UbLoad{ub, nc}, StoreExcl{toff}, nullptr, scratch, st);
  return rc ? GK_ERR_HIP : hip_status();
}

// Rows whose upper bound fits S lanes, any coarse id width: one entry per
// lane, bitonic sort of (coarse neighbour, weight) pairs across the S lanes
// by shuffles, segmented sum of equal neighbours, self entries dropped, run
// tails compacted with a ballot.  Used when the coarse ids do not leave room
// for the lane number beside them in 32 bits (k_contract_rows).
template <int S, int B>
__global__ void __launch_bounds__(B) k_contract_rows_wide(const int32_t* __restrict__ rows,
                                                       const int2* __restrict__ desc, int32_t nrows,
                                                       const int2* __restrict__ mrow,
                                                       const int32_t* __restrict__ adj,
                                                       const int32_t* __restrict__ w,
                                                       const int32_t* __restrict__ cmap,
                                                       const int32_t* __restrict__ toff,
                                                       int32_t* __restrict__ tadj, int32_t* __restrict__ tw,
                                                       int32_t* __restrict__ len) {
  const int64_t gt = int64_t(xcd_block()) * B + threadIdx.x;
  const int64_t ri = gt / S;
  const int sub = static_cast<int>(gt % S);
  if (ri >= nrows) return;
  const int32_t c = rows[ri];
  const int2 d = desc[ri];
  const int32_t m0 = d.x, nm = d.y;
  // Locate this lane's entry in the concatenation of the member rows: lane m
  // loads member m's row start and length, a shuffle scan gives the offsets,
  // and each lane finds its member among them.  More members than lanes
  // (max_group > S): member by member.
  int32_t j = -1;
  if (nm <= S) {
    int32_t ms = 0, md = 0;
    if (sub < nm) {
      const int2 r = mrow[m0 + sub];
      ms = r.x;
      md = r.y;
    }
    int32_t inc = md;
#pragma unroll
    for (int o = 1; o < S; o <<= 1) {
      const int32_t y = __shfl_up(inc, o, S);
      if (sub >= o) inc += y;
    }
    for (int m = 0; m < nm; ++m) {
      const int32_t im = __shfl(inc, m, S), sm = __shfl(ms, m, S), dm = __shfl(md, m, S);
      if (sub < im && sub >= im - dm) j = sm + (sub - (im - dm));
    }
  } else {
    int32_t pre = 0;
    for (int32_t m = m0; m < m0 + nm; ++m) {
      const int2 r = mrow[m];
      if (sub >= pre && sub < pre + r.y) j = r.x + (sub - pre);
      pre += r.y;
    }
  }
  int32_t key = INT32_MAX, val = 0;
  if (j >= 0) { key = cmap[adj[j]]; val = w[j]; }
#pragma unroll
  for (int k2 = 2; k2 <= S; k2 <<= 1) {
#pragma unroll
    for (int j2 = k2 >> 1; j2 > 0; j2 >>= 1) {
      const int32_t ok = __shfl_xor(key, j2, S);
      const int32_t ov = __shfl_xor(val, j2, S);
      const bool up = (sub & k2) == 0;
      const bool lower = (sub & j2) == 0;
      const bool take = lower ? (up ? ok < key : ok > key) : (up ? ok > key : ok < key);
      if (take) { key = ok; val = ov; }
    }
  }
  int32_t x = val;
#pragma unroll
  for (int o = 1; o < S; o <<= 1) {
    const int32_t y = __shfl_up(x, o, S);
    const int32_t ky = __shfl_up(key, o, S);
    if (sub >= o && ky == key) x += y;
  }
  const int32_t kn = __shfl_down(key, 1, S);
  const bool tail = (sub == S - 1) || kn != key;
  const bool out = tail && key != c && key != INT32_MAX;
  const uint64_t b = ballot(out);
  const int base_lane = lane_id() & ~(S - 1);
  const uint64_t seg = (S == 64) ? b : ((b >> base_lane) & ((1ull << S) - 1));
  const int pos = popc64(seg & ((1ull << sub) - 1));
  if (out) {
    const int32_t o = toff[c] + pos;
    tadj[o] = key;
    tw[o] = x;
  }
  if (sub == 0) len[c] = popc64(seg);
}

// Rows whose upper bound fits S lanes (S = 16, 32, 64; B / S rows per
// block), coarse ids below 2^(32 - log2 S) - 1.  Per row:
//   1. lane m < nm loads member m's row bounds (mrow); an add-scan of the
//      lengths gives each member's first lane, which the member marks in a
//      per-wave LDS slot array; lane i takes the entry of the last member
//      marked at or below it (a ballot of the marks);
//   2. its entry (coarse neighbour, weight); self entries become the
//      sentinel 0xffffffff, others the key neighbour << log2 S | lane;
//   3. an in-register bitonic sort of the keys (DPP partners), the weight
//      follows its key by one ds_bpermute;
//   4. runs of equal neighbours: a segmented add-scan minus the exclusive
//      sum at the run's head gives the run total at its tail; tails are
//      compacted with a ballot.
// The dependent loads per entry: the row's list slot (id and member range),
// the members' row bounds, the entry, its coarse id.
// Global thread slot gt serves row gt / S as lane gt % S of its segment; wm
// is the wave's 64-slot LDS mark array.  The whole wave calls it together.
template <int S>
__device__ __forceinline__ void contract_row_slot(int64_t gt, int32_t* wm, const int32_t* __restrict__ rows,
                                                  const int2* __restrict__ desc, int32_t nrows,
                                                  const int2* __restrict__ mrow, const int32_t* __restrict__ adj,
                                                  const int32_t* __restrict__ w, const int32_t* __restrict__ cmap,
                                                  const int32_t* __restrict__ toff, int32_t* __restrict__ tadj,
                                                  int32_t* __restrict__ tw, int32_t* __restrict__ len) {
  static_assert(S == 16 || S == 32 || S == 64, "segment of 16, 32 or 64 lanes");
  constexpr int LB = S == 16 ? 4 : (S == 32 ? 5 : 6);
  constexpr uint32_t kNone = 0xffffffffu;
  const int64_t ri = gt >> LB;
  const int lane = lane_id();
  const int sub = lane & (S - 1);
  const int base = lane - sub;
  // Segments past the last row stay active (masked stores) so that every
  // cross-lane step runs on whole segments.
  const bool live = ri < nrows;
  int32_t c = -1, m0 = 0, nm = 0, to = 0;
  if (live) {
    c = rows[ri];
    const int2 d = desc[ri];
    m0 = d.x;
    nm = d.y;
    to = toff[c];
  }
  int32_t j = -1;
  if (nm <= S) {
    int32_t ms = 0, md = 0;
    if (sub < nm) {
      const int2 r = mrow[m0 + sub];
      ms = r.x;
      md = r.y;
    }
    const int32_t inc = seg_incl_scan<S>(md);
    const int32_t e = inc - md;
    const int32_t tot = __shfl(inc, base + S - 1, 64);
    wm[lane] = INT32_MIN;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (md > 0) wm[base + e] = ms - e;  // distinct lanes: members with entries start apart
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint64_t starts = ballot(wm[lane] != INT32_MIN);
    if (sub < tot) j = wm[last_set_at_or_below(starts, lane)] + sub;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    // More members than lanes (members without entries): member by member.
    int32_t pre = 0;
    for (int32_t m = m0; m < m0 + nm; ++m) {
      const int2 r = mrow[m];
      if (sub >= pre && sub < pre + r.y) j = r.x + (sub - pre);
      pre += r.y;
    }
  }
  uint32_t key = kNone;
  int32_t val = 0;
  if (j >= 0) {
    const int32_t q = cmap[adj[j]];
    val = w[j];
    if (q != c) key = (uint32_t(q) << LB) | uint32_t(sub);
  }
  key = seg_sort_u32<S>(key);
  const bool valid = key != kNone;
  const int32_t sv = __shfl(val, base + int(key & (S - 1)), 64);
  const uint32_t nb = key >> LB;
  const uint32_t prev = dpp_mov<dpp::kWaveShr1>(key), next = dpp_mov<dpp::kWaveShl1>(key);
  const bool head = valid && (sub == 0 || (prev >> LB) != nb);
  const bool tail = valid && (sub == S - 1 || (next >> LB) != nb);
  const int32_t x = valid ? sv : 0;
  const int32_t p = seg_incl_scan<S>(x);
  const int h = last_set_at_or_below(ballot(head), lane);
  const int32_t eh = __shfl(p - x, h < 0 ? lane : h, 64);
  const uint64_t tails = ballot(tail);
  const uint64_t segm = S == 64 ? ~0ull : (((1ull << S) - 1) << base);
  if (live && tail) {
    const int32_t o = to + popc64(tails & segm & lanemask_lt());
    tadj[o] = int32_t(nb);
    tw[o] = p - eh;
  }
  if (live && sub == 0) len[c] = popc64(tails & segm);
}

template <int S, int B>
__global__ void __launch_bounds__(B) k_contract_rows(const int32_t* __restrict__ rows,
                                                       const int2* __restrict__ desc, int32_t nrows,
                                                       const int2* __restrict__ mrow,
                                                       const int32_t* __restrict__ adj,
                                                       const int32_t* __restrict__ w,
                                                       const int32_t* __restrict__ cmap,
                                                       const int32_t* __restrict__ toff,
                                                       int32_t* __restrict__ tadj, int32_t* __restrict__ tw,
                                                       int32_t* __restrict__ len) {
  static_assert(B % 64 == 0, "whole waves");
  __shared__ int32_t mark[B];
  contract_row_slot<S>(int64_t(xcd_block()) * B + threadIdx.x, mark + (threadIdx.x & ~63u), rows, desc, nrows, mrow,
                       adj, w, cmap, toff, tadj, tw, len);
}

// The same slots, taken from the work-stealing queues.
template <int S, int B>
__global__ void __launch_bounds__(B) k_contract_rows_ws(const int32_t* __restrict__ rows,
                                                          const int2* __restrict__ desc, int32_t nrows,
                                                          const int2* __restrict__ mrow,
                                                          const int32_t* __restrict__ adj,
                                                          const int32_t* __restrict__ w,
                                                          const int32_t* __restrict__ cmap,
                                                          const int32_t* __restrict__ toff,
                                                          int32_t* __restrict__ tadj, int32_t* __restrict__ tw,
                                                          int32_t* __restrict__ len,
                                                          unsigned long long* __restrict__ ctr) {
  static_assert(B % 64 == 0, "whole waves");
  __shared__ int32_t mark[B];
  int32_t* wm = mark + (threadIdx.x & ~63u);
  const int64_t sweeps = (int64_t(nrows) * S + 63) / 64;
  steal_sweeps(ctr, sweeps, [&](int64_t sw) {
    contract_row_slot<S>(sw * 64 + lane_id(), wm, rows, desc, nrows, mrow, adj, w, cmap, toff, tadj, tw, len);
  });
}

// Largest coarse id count for k_contract_rows<64, B> (ids below 2^26 - 1);
// the 16- and 32-lane kernels allow more, one limit keeps the choice simple.
constexpr int32_t kNarrowIdsMax = (1 << 26) - 1;

// Wide rows (coarse graphs of 2^26 - 1 vertices or more) always take the
// direct launch; the narrow ones may take the work-stealing form.
template <int B, bool Wide>
static void contract_rows_block(const DGraph& g, int lanes, const int32_t* rows, const int2* desc, int32_t nrows,
                                const int2* mrow, const int32_t* cmap, const int32_t* toff, int32_t* tadj,
                                int32_t* tw, int32_t* len, const Steal* ws, hipStream_t st) {
  const unsigned grid = grid_for(int64_t(nrows) * lanes, B);
  auto go = [&](auto kern) {
    gk_launch(kern, dim3(grid), dim3(B), 0, st, rows, desc, nrows, mrow, g.adj, g.w, cmap, toff, tadj, tw, len);
  };
  auto go_ws = [&](auto kern, int S) {
    const unsigned pg = steal_grid(reinterpret_cast<const void*>(kern), B, *ws, grid);
    unsigned long long* ctr = nullptr;  // queues cleared and used only when the work has a tail
    if (steal_tail((int64_t(nrows) * S + 63) / 64, int64_t(pg) * (B / 64)) > 0) {
      fill<unsigned long long>(ws->ctr, 0ull, kStealWords, st);
      ctr = ws->ctr;
    }
    gk_launch(kern, dim3(pg), dim3(B), 0, st, rows, desc, nrows, mrow, g.adj, g.w, cmap, toff, tadj, tw, len, ctr);
  };
  if (!Wide && ws) {
    switch (lanes) {
      case 16: go_ws(k_contract_rows_ws<16, B>, 16); break;
      case 32: go_ws(k_contract_rows_ws<32, B>, 32); break;
      default: go_ws(k_contract_rows_ws<64, B>, 64); break;
    }
    return;
  }
  switch (lanes) {
    case 16: Wide ? go(k_contract_rows_wide<16, B>) : go(k_contract_rows<16, B>); break;
    case 32: Wide ? go(k_contract_rows_wide<32, B>) : go(k_contract_rows<32, B>); break;
    default: Wide ? go(k_contract_rows_wide<64, B>) : go(k_contract_rows<64, B>); break;
  }
}

template <bool Wide>
static void contract_rows_any(const DGraph& g, int lanes, int block, const int32_t* rows, const int2* desc,
                              int32_t nrows, const int2* mrow, const int32_t* cmap, const int32_t* toff,
                              int32_t* tadj, int32_t* tw, int32_t* len, const Steal* ws, hipStream_t st) {
  switch (block) {
    case 64: contract_rows_block<64, Wide>(g, lanes, rows, desc, nrows, mrow, cmap, toff, tadj, tw, len, ws, st); break;
    case 128: contract_rows_block<128, Wide>(g, lanes, rows, desc, nrows, mrow, cmap, toff, tadj, tw, len, ws, st); break;
    case 512: contract_rows_block<512, Wide>(g, lanes, rows, desc, nrows, mrow, cmap, toff, tadj, tw, len, ws, st); break;
    default: contract_rows_block<256, Wide>(g, lanes, rows, desc, nrows, mrow, cmap, toff, tadj, tw, len, ws, st); break;
  }
}

int launch_contract_rows(const DGraph& g, int lanes, int block, const int32_t* rows, const int2* desc,
                         int32_t nrows, int32_t nc, const int2* mrow, const int32_t* cmap, const int32_t* toff,
                         int32_t* tadj, int32_t* tw, int32_t* len, bool force_wide, hipStream_t st,
                         const Steal* ws) {
  if (nrows <= 0) return 0;
  if (force_wide || nc > kNarrowIdsMax)
    contract_rows_any<true>(g, lanes, block, rows, desc, nrows, mrow, cmap, toff, tadj, tw, len, nullptr, st);
  else
    contract_rows_any<false>(g, lanes, block, rows, desc, nrows, mrow, cmap, toff, tadj, tw, len, ws, st);
  return hip_status();
}

struct BigUbLoad {
  const int32_t* rows;
  const int32_t* ub;
  int32_t nrows;
  __device__ int operator()(int64_t i) const { return i < nrows ? ub[rows[i]] : 0; }
};

int launch_big_offsets(const int32_t* rows, int32_t nrows, const int32_t* ub, int32_t* boff,
                       int32_t* total, int32_t* scratch, hipStream_t st) {
  int rc = scan(int64_t(nrows) + 1, BigUbLoad{rows, ub, nrows}, StoreExcl{boff}, total, scratch, st);
  return rc ? GK_ERR_HIP : hip_status();
}

// Big rows: every entry becomes the key (row rank << 32 | coarse neighbour);
// one block per row walks the members in order.
__global__ void __launch_bounds__(256) k_big_gather(const int32_t* __restrict__ rows,
                                                    const int32_t* __restrict__ boff,
                                                    const int32_t* __restrict__ cptr,
                                                    const int32_t* __restrict__ order,
                                                    const int32_t* __restrict__ xadj,
                                                    const int32_t* __restrict__ adj,
                                                    const int32_t* __restrict__ w,
                                                    const int32_t* __restrict__ cmap,
                                                    uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int32_t i = blockIdx.x;
  const int32_t c = rows[i];
  int32_t o = boff[i];
  for (int32_t m = cptr[c]; m < cptr[c + 1]; ++m) {
    const int32_t v = order[m];
    const int32_t s = xadj[v], d = xadj[v + 1] - s;
    for (int32_t t = threadIdx.x; t < d; t += 256) {
      keys[o + t] = (uint64_t(uint32_t(i)) << 32) | uint64_t(uint32_t(cmap[adj[s + t]]));
      vals[o + t] = w[s + t];
    }
    o += d;
  }
}

int launch_contract_big_gather(const DGraph& g, const int32_t* rows, int32_t nrows,
                               const int32_t* boff, const int32_t* cptr, const int32_t* order,
                               const int32_t* cmap, uint64_t* keys, int32_t* vals, hipStream_t st) {
  if (nrows <= 0) return 0;
  k_big_gather<<<nrows, 256, 0, st>>>(rows, boff, cptr, order, g.xadj, g.adj, g.w, cmap, keys, vals);
  return hip_status();
}

// One wave per big row walks its sorted entries in chunks of 64 with a
// segmented sum carried across chunks.
__global__ void __launch_bounds__(64) k_big_merge(const int32_t* __restrict__ rows,
                                                  const int32_t* __restrict__ boff,
                                                  const uint64_t* __restrict__ keys,
                                                  const int32_t* __restrict__ vals,
                                                  const int32_t* __restrict__ toff,
                                                  int32_t* __restrict__ tadj, int32_t* __restrict__ tw,
                                                  int32_t* __restrict__ len) {
  const int32_t i = blockIdx.x;
  const int32_t c = rows[i];
  const int32_t b0 = boff[i], b1 = boff[i + 1];
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();
  int32_t carry_key = -1, carry_sum = 0;
  int32_t count = 0;
  const int32_t obase = toff[c];
  for (int32_t base = b0; base < b1; base += 64) {
    const int32_t e = base + lane;
    const bool valid = e < b1;
    const int32_t key = valid ? int32_t(uint32_t(keys[e])) : INT32_MAX;
    int32_t x = valid ? vals[e] : 0;
    if (valid && key == carry_key) x += (lane == 0) ? carry_sum : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(x, o, 64);
      const int32_t ky = __shfl_up(key, o, 64);
      if (lane >= o && ky == key) x += y;
    }
    // Tail: the next element (possibly in the next chunk) has another key.
    int32_t kn = __shfl_down(key, 1, 64);
    if (lane == 63) kn = (e + 1 < b1) ? int32_t(uint32_t(keys[e + 1])) : INT32_MAX;
    const bool tail = valid && (e + 1 == b1 || kn != key);
    const bool out = tail && key != c;
    const uint64_t bm = ballot(out);
    if (out) {
      const int32_t o = obase + count + popc64(bm & lt);
      tadj[o] = key;
      tw[o] = x;
    }
    count += popc64(bm);
    // Carry the last run into the next chunk when it continues.
    const int32_t last_key = __shfl(key, 63, 64);
    const int32_t last_x = __shfl(x, 63, 64);
    const bool last_tail = __shfl(static_cast<int>(tail), 63, 64) != 0;
    carry_key = last_tail ? -1 : last_key;
    carry_sum = last_x;
  }
  if (lane == 0) len[c] = count;
}

int launch_contract_big_merge(const int32_t* rows, int32_t nrows, const int32_t* boff,
                              const uint64_t* keys, const int32_t* vals, const int32_t* toff,
                              int32_t* tadj, int32_t* tw, int32_t* len, hipStream_t st) {
  if (nrows <= 0) return 0;
  k_big_merge<<<nrows, 64, 0, st>>>(rows, boff, keys, vals, toff, tadj, tw, len);
  return hip_status();
}

struct LenLoad {
  const int32_t* len;
  int32_t nc;
  __device__ int operator()(int64_t i) const { return i < nc ? len[i] : 0; }
};

int launch_len_scan(int32_t nc, const int32_t* len, int32_t* cxadj, int32_t* cnnz,
                    int32_t* scratch, hipStream_t st) {
  int rc = scan(int64_t(nc) + 1, LenLoad{len, nc}, StoreExcl{cxadj}, cnnz, scratch, st);
  return rc ? GK_ERR_HIP : hip_status();
}

// Copy merged rows to the final CSR (8 lanes per row).
__global__ void __launch_bounds__(256) k_contract_copy(int32_t nc, const int32_t* __restrict__ toff,
                                                       const int32_t* __restrict__ len,
                                                       const int32_t* __restrict__ cxadj,
                                                       const int32_t* __restrict__ tadj,
                                                       const int32_t* __restrict__ tw,
                                                       int32_t* __restrict__ cadj, int32_t* __restrict__ cw) {
  const int64_t gt = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int64_t c = gt / 8;
  const int sub = static_cast<int>(gt % 8);
  if (c >= nc) return;
  const int32_t s = toff[c], d = cxadj[c], l = len[c];
  for (int32_t t = sub; t < l; t += 8) {
    cadj[d + t] = tadj[s + t];
    cw[d + t] = tw[s + t];
  }
}

int launch_contract_copy(int32_t nc, const int32_t* toff, const int32_t* len, const int32_t* cxadj,
                         const int32_t* tadj, const int32_t* tw, int32_t* cadj, int32_t* cw,
                         hipStream_t st) {
  if (nc <= 0) return 0;
  k_contract_copy<<<grid_for(int64_t(nc) * 8, 256), 256, 0, st>>>(nc, toff, len, cxadj, tadj, tw, cadj, cw);
  return hip_status();
}

}  // namespace gkd
