// Uncoarsening and refinement kernels for gfx950: projection, boundary/gain
// computation, independent-set move selection, balanced-prefix search,
// move application and the final cut/part weights -- the work of
// gkway/uncoarsen.hpp and gkway/cuda_check.hpp:61-73.  Integer-exact and
// schedule-independent: selected moves are ordered by a sort on the packed
// key (gain desc, id asc) before any order-sensitive step.
#include <algorithm>

#include "kernels.hpp"
#include "prim.hpp"
#include "balance.hpp"
#include "../../include/gkway.h"

namespace gkd {

static inline int hip_status() { return hipGetLastError() == hipSuccess ? 0 : GK_ERR_HIP; }

// expand_partition (gkway/uncoarsen.hpp:325-356), fused with the two
// device-to-device copies that follow it there (:381-382).
__global__ void __launch_bounds__(256) k_project(int32_t n, const int32_t* __restrict__ cmap,
                                                 const int32_t* __restrict__ cpart,
                                                 const int32_t* __restrict__ cbnd,
                                                 int32_t* __restrict__ part, int32_t* __restrict__ bnd,
                                                 int32_t* __restrict__ gain, int32_t* __restrict__ dest,
                                                 int32_t* __restrict__ upd) {
  const int64_t v = int64_t(xcd_block()) * 256 + threadIdx.x;
  if (v >= n) return;
  const int32_t c = cmap[v];
  const int32_t p = cpart[c];
  const int32_t b = cbnd[c] ? 1 : 0;
  part[v] = p;
  bnd[v] = b;
  gain[v] = 0;
  dest[v] = p;
  upd[v] = b;
}

int launch_project(int32_t n, const int32_t* cmap, const int32_t* cpart, const int32_t* cbnd,
                   int32_t* part, int32_t* bnd, int32_t* gain, int32_t* dest, int32_t* upd,
                   hipStream_t st) {
  if (n <= 0) return 0;
  k_project<<<grid_for(n, 256), 256, 0, st>>>(n, cmap, cpart, cbnd, part, bnd, gain, dest, upd);
  return hip_status();
}

// Worklist compaction: the indices v with flags[v] != 0, in any order (the
// gain pass is per vertex).  A block covers kCompactTile flags (16 per
// thread, wave-contiguous slots): it counts them, takes one global atomic,
// then writes -- few enough atomics that the counter is never contended.
constexpr int kCompactIPT = 16;
constexpr int kCompactTile = 256 * kCompactIPT;

__global__ void __launch_bounds__(256) k_compact(int32_t n, const int32_t* __restrict__ flags,
                                                 int32_t* __restrict__ list, int32_t* __restrict__ cnt) {
  __shared__ int wsum[4];
  __shared__ int base_s;
  const int wave = threadIdx.x >> 6;
  const int64_t wbase = int64_t(blockIdx.x) * kCompactTile + int64_t(wave) * 64 * kCompactIPT;
  uint64_t m[kCompactIPT];
  int tot = 0;
#pragma unroll
  for (int j = 0; j < kCompactIPT; ++j) {
    const int64_t v = wbase + int64_t(j) * 64 + lane_id();
    m[j] = ballot(v < n && flags[v] != 0);
    tot += popc64(m[j]);
  }
  if (lane_id() == 0) wsum[wave] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    base_s = t ? atomicAdd(cnt, t) : 0;
  }
  __syncthreads();
  int off = base_s;
  for (int w = 0; w < wave; ++w) off += wsum[w];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kCompactIPT; ++j) {
    if ((m[j] >> lane_id()) & 1ull) list[off + popc64(m[j] & lt)] = int32_t(wbase + int64_t(j) * 64 + lane_id());
    off += popc64(m[j]);
  }
}

int launch_compact(int32_t n, const int32_t* flags, int32_t* list, int32_t* cnt, hipStream_t st) {
  if (n <= 0) return 0;
  k_compact<<<grid_for(n, kCompactTile), 256, 0, st>>>(n, flags, list, cnt);
  return hip_status();
}

// Candidate bitmask of the batched form: bit v set iff v is a boundary vertex
// with positive gain (the vertices its selection looks at).
__device__ __forceinline__ void set_cand(uint64_t* cand, int32_t v, bool on) {
  const uint64_t b = 1ull << (v & 63);
  if (on) atomicOr(reinterpret_cast<unsigned long long*>(&cand[v >> 6]), (unsigned long long)b);
  else atomicAnd(reinterpret_cast<unsigned long long*>(&cand[v >> 6]), (unsigned long long)~b);
}

// update_vertex_info + calculate_gain (gkway/uncoarsen.hpp:175-237, 82-109)
// over a worklist of updated vertices (wlc[0] entries), S lanes per vertex,
// grid-stride.  Each lane holds one (part, weight) entry of the row and sums
// the weights of the lanes with the same part (an O(S) shuffle loop), which
// gives the connectivity to every adjacent part without a histogram.  Gain is
// the signed max over the adjacent parts of conn_q - conn_p, floored at 0,
// lowest q on ties.  Rows longer than S: with k <= kGainInlineK the whole
// wave takes them one at a time after each pass, with a per-wave LDS
// histogram of the k connectivities (as k_gain_heavy); with larger k they are
// appended to the heavy list (wlc[1] entries) for k_gain_heavy.  Processed
// vertices get upd[v] = 0.
//
// Direct form (Steal = false): group g of the grid (S lanes) takes worklist
// entries g, g + nsg, ...; the loop runs while any group of the wave has an
// entry, so the wave can take its long rows together.  Work-stealing form
// (Steal = true, k_gain_wl_ws): the wave's passes of 64 / S entries are
// taken from the queues at ctr.  Each entry is processed once
// either way and the statistics are sums, so both give the same results.
template <int S, bool Steal>
__device__ __forceinline__ void gain_wl_body(const int32_t* __restrict__ wl, int32_t* __restrict__ wlc, int32_t k,
                                             const int32_t* __restrict__ xadj, const int32_t* __restrict__ adj,
                                             const int32_t* __restrict__ w, const int32_t* __restrict__ part,
                                             int32_t* __restrict__ upd, int32_t* __restrict__ bnd,
                                             int32_t* __restrict__ gain, int32_t* __restrict__ dest,
                                             int32_t* __restrict__ heavy, unsigned long long* __restrict__ work,
                                             uint64_t* __restrict__ cand, unsigned long long* __restrict__ ctr,
                                             int32_t (*hh)[kGainInlineK], int (*ws)[4]) {
  const int32_t cnt = wlc[0];
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  const int sub = threadIdx.x & (S - 1);
  const bool inline_heavy = k <= kGainInlineK;
  const int grp = lane / S;
  int done_v = 0, done_e = 0;
  auto pass = [&](int64_t eb) {
    const int64_t e = eb + grp;
    int32_t v = 0, s = 0, d = 0;
    if (e < cnt) {
      v = wl[e];
      s = xadj[v];
      d = xadj[v + 1] - s;
    }
    const bool long_row = e < cnt && d > S;
    if (long_row && !inline_heavy && sub == 0) heavy[atomicAdd(&wlc[1], 1)] = v;
    if (e < cnt && d <= S) {
      const int32_t p = part[v];
      int32_t q = -1, wj = 0;
      if (sub < d) {
        q = part[adj[s + sub]];
        wj = w[s + sub];
      }
      int32_t conn = 0, id = 0;
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const int32_t qi = __shfl(q, i, S);
        const int32_t wi = __shfl(wj, i, S);
        conn += (qi == q) ? wi : 0;
        id += (qi == p) ? wi : 0;
      }
      const bool ext = q >= 0 && q != p;
      int32_t g = ext ? conn - id : INT32_MIN;
      int32_t bq = ext ? q : INT32_MAX;
      int32_t anyext = ext ? 1 : 0;
#pragma unroll
      for (int o = S / 2; o > 0; o >>= 1) {
        const int32_t og = __shfl_xor(g, o, S), oq = __shfl_xor(bq, o, S);
        if (og > g || (og == g && oq < bq)) { g = og; bq = oq; }
        anyext |= __shfl_xor(anyext, o, S);
      }
      if (sub == 0) {
        bnd[v] = anyext;
        if (anyext && g > 0) {
          gain[v] = g;
          dest[v] = bq;
        } else {
          gain[v] = 0;
          dest[v] = p;
        }
        if (cand) set_cand(cand, v, anyext && g > 0);
        upd[v] = 0;
        ++done_v;
        done_e += d;
      }
    }
    if (!inline_heavy) return;
    // The pass's long rows, one at a time, by the whole wave.
    int32_t* hist = hh[wv];
    uint64_t lm = ballot(long_row && sub == 0);
    while (lm) {
      const int src = __ffsll(static_cast<long long>(lm)) - 1;
      lm &= lm - 1;
      const int32_t hv = __shfl(v, src, 64), hs = __shfl(s, src, 64), he = hs + __shfl(d, src, 64);
      for (int32_t q = lane; q < k; q += 64) hist[q] = 0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const int32_t p = part[hv];
      bool ext = false;
      for (int32_t j = hs + lane; j < he; j += 64) {
        const int32_t q = part[adj[j]];
        atomicAdd(&hist[q], w[j]);
        ext |= (q != p);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const bool any = ballot(ext) != 0;
      const int32_t id = hist[p];
      int32_t g = INT32_MIN, bq = INT32_MAX;
      for (int32_t q = lane; q < k; q += 64) {
        const int32_t h = hist[q];
        if (q != p && h > 0) {
          const int32_t gq = h - id;
          if (gq > g || (gq == g && q < bq)) { g = gq; bq = q; }
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const int32_t og = __shfl_xor(g, o, 64), oq = __shfl_xor(bq, o, 64);
        if (og > g || (og == g && oq < bq)) { g = og; bq = oq; }
      }
      if (lane == 0) {
        bnd[hv] = any ? 1 : 0;
        if (any && g > 0) {
          gain[hv] = g;
          dest[hv] = bq;
        } else {
          gain[hv] = 0;
          dest[hv] = p;
        }
        if (cand) set_cand(cand, hv, any && g > 0);
        upd[hv] = 0;
        ++done_v;
        done_e += he - hs;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  };
  if constexpr (Steal) {
    const int64_t sweeps = (int64_t(cnt) + (64 / S) - 1) / (64 / S);
    steal_sweeps(ctr, sweeps, [&](int64_t sw) { pass(sw * (64 / S)); });
  } else {
    const int64_t nsg = int64_t(gridDim.x) * (256 / S);
    const int64_t wave_g0 = (int64_t(xcd_block()) * 4 + wv) * (64 / S);
    for (int64_t eb = wave_g0; eb < cnt; eb += nsg) pass(eb);
  }
  if (work) {
    // Statistics: one slot pair per block (no two blocks of a launch share
    // one, launches are stream-ordered), summed on the host.
    const int a2 = wave_sum(done_v), b2 = wave_sum(done_e);
    if (lane_id() == 0) {
      ws[0][threadIdx.x >> 6] = a2;
      ws[1][threadIdx.x >> 6] = b2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      work[2 * blockIdx.x] += (unsigned long long)(ws[0][0] + ws[0][1] + ws[0][2] + ws[0][3]);
      work[2 * blockIdx.x + 1] += (unsigned long long)(ws[1][0] + ws[1][1] + ws[1][2] + ws[1][3]);
    }
  }
}

template <int S>
__global__ void __launch_bounds__(256) k_gain_wl(const int32_t* __restrict__ wl, int32_t* __restrict__ wlc,
                                                 int32_t k, const int32_t* __restrict__ xadj,
                                                 const int32_t* __restrict__ adj,
                                                 const int32_t* __restrict__ w,
                                                 const int32_t* __restrict__ part,
                                                 int32_t* __restrict__ upd, int32_t* __restrict__ bnd,
                                                 int32_t* __restrict__ gain, int32_t* __restrict__ dest,
                                                 int32_t* __restrict__ heavy,
                                                 unsigned long long* __restrict__ work,
                                                 uint64_t* __restrict__ cand) {
  __shared__ int32_t hh[4][kGainInlineK];
  __shared__ int ws[2][4];
  gain_wl_body<S, false>(wl, wlc, k, xadj, adj, w, part, upd, bnd, gain, dest, heavy, work, cand, nullptr, hh, ws);
}

template <int S>
__global__ void __launch_bounds__(256) k_gain_wl_ws(const int32_t* __restrict__ wl, int32_t* __restrict__ wlc,
                                                    int32_t k, const int32_t* __restrict__ xadj,
                                                    const int32_t* __restrict__ adj,
                                                    const int32_t* __restrict__ w,
                                                    const int32_t* __restrict__ part,
                                                    int32_t* __restrict__ upd, int32_t* __restrict__ bnd,
                                                    int32_t* __restrict__ gain, int32_t* __restrict__ dest,
                                                    int32_t* __restrict__ heavy,
                                                    unsigned long long* __restrict__ work,
                                                    uint64_t* __restrict__ cand,
                                                    unsigned long long* __restrict__ ctr) {
  __shared__ int32_t hh[4][kGainInlineK];
  __shared__ int ws[2][4];
  gain_wl_body<S, true>(wl, wlc, k, xadj, adj, w, part, upd, bnd, gain, dest, heavy, work, cand, ctr, hh, ws);
}

// Rows above S entries: one wave per vertex with a dense LDS histogram of the
// k part connectivities (integer LDS atomics, order-independent).
__global__ void __launch_bounds__(64) k_gain_heavy(const int32_t* __restrict__ heavy,
                                                   const int32_t* __restrict__ wlc, int32_t k,
                                                   const int32_t* __restrict__ xadj,
                                                   const int32_t* __restrict__ adj,
                                                   const int32_t* __restrict__ w,
                                                   const int32_t* __restrict__ part,
                                                   int32_t* __restrict__ upd,
                                                   int32_t* __restrict__ bnd, int32_t* __restrict__ gain,
                                                   int32_t* __restrict__ dest,
                                                   unsigned long long* __restrict__ work,
                                                   uint64_t* __restrict__ cand) {
  extern __shared__ int32_t hist[];
  const int lane = lane_id();
  const int32_t cnt = wlc[1];
  unsigned long long done_v = 0, done_e = 0;
  for (int32_t hi = blockIdx.x; hi < cnt; hi += gridDim.x) {
    const int32_t v = heavy[hi];
    for (int32_t q = lane; q < k; q += 64) hist[q] = 0;
    __syncthreads();
    const int32_t p = part[v];
    const int32_t s = xadj[v], e = xadj[v + 1];
    bool ext = false;
    for (int32_t j = s + lane; j < e; j += 64) {
      const int32_t q = part[adj[j]];
      atomicAdd(&hist[q], w[j]);
      ext |= (q != p);
    }
    __syncthreads();
    const bool any = ballot(ext) != 0;
    const int32_t id = hist[p];
    int32_t g = INT32_MIN, bq = INT32_MAX;
    for (int32_t q = lane; q < k; q += 64) {
      const int32_t h = hist[q];
      if (q != p && h > 0) {
        const int32_t gq = h - id;
        if (gq > g || (gq == g && q < bq)) { g = gq; bq = q; }
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const int32_t og = __shfl_xor(g, o, 64), oq = __shfl_xor(bq, o, 64);
      if (og > g || (og == g && oq < bq)) { g = og; bq = oq; }
    }
    if (lane == 0) {
      bnd[v] = any ? 1 : 0;
      if (any && g > 0) { gain[v] = g; dest[v] = bq; }
      else { gain[v] = 0; dest[v] = p; }
      if (cand) set_cand(cand, v, any && g > 0);
      upd[v] = 0;
      done_v += 1;
      done_e += (unsigned long long)(e - s);
    }
    __syncthreads();
  }
  if (work && lane == 0 && done_v) {
    work[2 * blockIdx.x] += done_v;  // one slot pair per block, as in k_gain_wl
    work[2 * blockIdx.x + 1] += done_e;
  }
}

template <int S>
static void gain_go(const DGraph& g, int32_t k, const int32_t* wl, int32_t* wlc, int32_t* heavy, const int32_t* part,
                    int32_t* upd, int32_t* bnd, int32_t* gain, int32_t* dest, unsigned long long* work,
                    unsigned grid, int64_t items, uint64_t* cand, const Steal* ws, hipStream_t st) {
  if (!ws) {
    gk_launch(k_gain_wl<S>, dim3(grid), dim3(256), 0, st, wl, wlc, k, g.xadj, g.adj, g.w, part, upd, bnd, gain, dest,
              heavy, work, cand);
    return;
  }
  // Every workgroup slot of the device, at most kGainMaxBlocks (the
  // statistics slots).  The count is on the device; the queues are cleared
  // and offered only when the caller's expected count would reach the tail
  // (a round's worklist rarely does), else the static schedule takes all.
  const unsigned pg = steal_grid(reinterpret_cast<const void*>(&k_gain_wl_ws<S>), 256, *ws, kGainMaxBlocks);
  unsigned long long* ctr = nullptr;
  if (steal_tail((items + (64 / S) - 1) / (64 / S), int64_t(pg) * 4) > 0) {
    fill<unsigned long long>(ws->ctr, 0ull, kStealWords, st);
    ctr = ws->ctr;
  }
  gk_launch(k_gain_wl_ws<S>, dim3(pg), dim3(256), 0, st, wl, wlc, k, g.xadj, g.adj, g.w, part, upd, bnd, gain, dest,
            heavy, work, cand, ctr);
}

int launch_gain_wl(const DGraph& g, int32_t k, int lanes, const int32_t* wl, int32_t* wlc, int32_t* heavy,
                   const int32_t* part, int32_t* upd, int32_t* bnd, int32_t* gain, int32_t* dest,
                   unsigned long long* work, int64_t max_items, uint64_t* cand, hipStream_t st, const Steal* ws) {
  if (g.n <= 0) return 0;
  // The grid covers at most max_items worklist entries (a bound the caller
  // knows without reading the count); the kernel grid-strides beyond it.
  const int64_t items = std::min<int64_t>(g.n, std::max<int64_t>(max_items, 1));
  unsigned grid = grid_for(items * lanes, 256);
  if (grid > kGainMaxBlocks) grid = kGainMaxBlocks;
  switch (lanes) {
    case 4: gain_go<4>(g, k, wl, wlc, heavy, part, upd, bnd, gain, dest, work, grid, items, cand, ws, st); break;
    case 8: gain_go<8>(g, k, wl, wlc, heavy, part, upd, bnd, gain, dest, work, grid, items, cand, ws, st); break;
    case 16: gain_go<16>(g, k, wl, wlc, heavy, part, upd, bnd, gain, dest, work, grid, items, cand, ws, st); break;
    case 32: gain_go<32>(g, k, wl, wlc, heavy, part, upd, bnd, gain, dest, work, grid, items, cand, ws, st); break;
    default: gain_go<64>(g, k, wl, wlc, heavy, part, upd, bnd, gain, dest, work, grid, items, cand, ws, st); break;
  }
  return hip_status();
}

int launch_gain_heavy(const DGraph& g, int32_t k, const int32_t* heavy, const int32_t* wlc, const int32_t* part,
                      int32_t* upd, int32_t* bnd, int32_t* gain, int32_t* dest, unsigned long long* work,
                      uint64_t* cand, hipStream_t st) {
  if (g.n <= 0) return 0;
  // Grid-stride over the heavy list; more waves when k (the LDS histogram)
  // is small.
  const unsigned grid = k <= 1024 ? 2048 : 256;
  gk_launch(k_gain_heavy, dim3(grid), dim3(64), uint32_t(sizeof(int32_t) * k), st, heavy, wlc, k, g.xadj, g.adj, g.w,
            part, upd, bnd, gain, dest, work, cand);
  return hip_status();
}

// create_independent_move_buffer (gkway/uncoarsen.hpp:111-173): clears the
// update flags, keeps boundary vertices with positive gain whose move fits the
// cap and that have no smaller-id neighbour with the same property, and
// appends them with one atomic per wave (ballot + mbcnt instead of the
// reference's 32-lane __activemask/__popc).
__global__ void __launch_bounds__(256) k_select(int32_t n, const int32_t* __restrict__ xadj,
                                                const int32_t* __restrict__ adj,
                                                const int32_t* __restrict__ vw,
                                                const int32_t* __restrict__ bnd,
                                                const int32_t* __restrict__ gain,
                                                const int32_t* __restrict__ dest,
                                                const int32_t* __restrict__ pw, long long cap, int idbits,
                                                int32_t* __restrict__ upd, uint64_t* __restrict__ keys,
                                                int32_t* __restrict__ ctl,
                                                unsigned long long* __restrict__ work, int clear_upd,
                                                int32_t* __restrict__ reset) {
  const int64_t v = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (reset && v == 0) { reset[0] = 0; reset[1] = 0; }
  bool sel = false, cand = false, walked = false;
  int32_t g = 0, scanned = 0;
  if (v < n) {
    if (clear_upd) upd[v] = 0;
    g = gain[v];
    cand = bnd[v] && g > 0;
    if (cand && (long long)pw[dest[v]] + vw[v] <= cap) {
      sel = true;
      walked = true;
      for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j) {
        ++scanned;
        const int32_t u = adj[j];
        if (u < v && gain[u] > 0 && (long long)pw[dest[u]] + vw[u] <= cap) { sel = false; break; }
      }
    }
  }
  if (work) {
    const int c0 = popc64(ballot(cand)), c1 = popc64(ballot(walked));
    const int c2 = wave_sum(scanned);
    if (lane_id() == 0 && (c0 | c1 | c2)) {
      atomicAdd(&work[2], (unsigned long long)c0);
      atomicAdd(&work[3], (unsigned long long)c1);
      atomicAdd(&work[4], (unsigned long long)c2);
    }
  }
  const uint64_t m = ballot(sel);
  if (m == 0) return;
  const int leader = __ffsll(static_cast<long long>(m)) - 1;
  const int gmax = wave_max(sel ? g : 0);
  int base = 0;
  if (lane_id() == leader) {
    base = atomicAdd(&ctl[0], popc64(m));
    atomicMax(&ctl[1], gmax);
  }
  base = __shfl(base, leader, 64);
  if (sel)
    keys[base + popc64(m & lanemask_lt())] =
        (uint64_t(uint32_t(~g) & 0x7fffffffu) << idbits) | uint64_t(uint32_t(v));
}

int launch_select(const DGraph& g, const int32_t* bnd, const int32_t* gain, const int32_t* dest,
                  const int32_t* pw, int64_t cap, int idbits, int32_t* upd, uint64_t* keys,
                  int32_t* ctl, unsigned long long* work, int clear_upd, int32_t* reset, hipStream_t st) {
  fill<int32_t>(ctl, 0, 2, st);
  if (g.n <= 0) return 0;
  k_select<<<grid_for(g.n, 256), 256, 0, st>>>(g.n, g.xadj, g.adj, g.vw, bnd, gain, dest, pw, cap,
                                               idbits, upd, keys, ctl, work, clear_upd, reset);
  return hip_status();
}

// One workgroup: order the (at most 1024) first moves, then the largest
// balanced prefix (balance.hpp); ctl[2] receives its last index.
__global__ void __launch_bounds__(kBalThreads) k_balance(int32_t k, const uint64_t* __restrict__ keys,
                                                         int32_t count, int32_t M, int presorted,
                                                         int idbits, const int32_t* __restrict__ part,
                                                         const int32_t* __restrict__ dest,
                                                         const int32_t* __restrict__ gain,
                                                         const int32_t* __restrict__ vw,
                                                         int32_t* __restrict__ pw, long long cap,
                                                         int32_t* __restrict__ moves,
                                                         int32_t* __restrict__ ctl,
                                                         long long* __restrict__ cut) {
  __shared__ BalanceLds L;
  extern __shared__ int32_t dpw[];  // k entries
  const uint64_t idmask = (idbits >= 64) ? ~0ull : ((1ull << idbits) - 1);
  balance_block(L, dpw, k, keys, count, M, presorted, idmask, part, dest, gain, vw, pw, cap, moves, &ctl[2], cut);
}

int launch_balance(const DGraph& g, int32_t k, const uint64_t* keys, int32_t count, int32_t M,
                   bool presorted, int idbits, const int32_t* part, const int32_t* dest,
                   const int32_t* gain, int32_t* pw, int64_t cap, int32_t* moves, int32_t* ctl,
                   long long* cut, hipStream_t st) {
  if (M > kBalMax || (!presorted && count > kBalMax)) return GK_ERR_LIMIT;
  k_balance<<<1, kBalThreads, sizeof(int32_t) * k, st>>>(k, keys, count, M, presorted ? 1 : 0, idbits,
                                                         part, dest, gain, g.vw, pw, cap, moves, ctl, cut);
  return hip_status();
}

// apply_sequence_move (gkway/uncoarsen.hpp:290-323) without the pw/cut parts
// (done in k_balance): move, and mark the mover and its neighbours updated.
// A vertex whose flag goes 0 -> 1 is appended to the gain worklist, so each
// updated vertex enters it once.
__global__ void __launch_bounds__(256) k_apply(int32_t M, const int32_t* __restrict__ xadj,
                                               const int32_t* __restrict__ adj,
                                               const int32_t* __restrict__ moves,
                                               const int32_t* __restrict__ ctl,
                                               const int32_t* __restrict__ dest,
                                               int32_t* __restrict__ part, int32_t* __restrict__ upd,
                                               int32_t* __restrict__ wl, int32_t* __restrict__ wlc) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const bool mv = j < M && j <= ctl[2];
  int32_t v = 0, es = 0, deg = 0;
  if (mv) {
    v = moves[j];
    part[v] = dest[v];
    es = xadj[v];
    deg = xadj[v + 1] - es;
  }
  // The mover (e = -1), then its neighbours; appends are aggregated per wave.
  int32_t e = -1;
  bool act = mv;
  while (ballot(act) != 0) {
    int32_t u = 0;
    bool app = false;
    if (act) {
      u = e < 0 ? v : adj[es + e];
      app = atomicExch(&upd[u], 1) == 0;
    }
    const uint64_t am = ballot(app);
    if (am) {
      const int leader = __ffsll(static_cast<long long>(am)) - 1;
      int base = 0;
      if (lane_id() == leader) base = atomicAdd(&wlc[0], popc64(am));
      base = __shfl(base, leader, 64);
      if (app) wl[base + popc64(am & lanemask_lt())] = u;
    }
    if (act) {
      ++e;
      act = e < deg;
    }
  }
}

int launch_apply(const DGraph& g, int32_t M, const int32_t* moves, const int32_t* ctl,
                 const int32_t* dest, int32_t* part, int32_t* upd, int32_t* wl, int32_t* wlc,
                 hipStream_t st) {
  k_apply<<<grid_for(M, 256), 256, 0, st>>>(M, g.xadj, g.adj, moves, ctl, dest, part, upd, wl, wlc);
  return hip_status();
}

__global__ void __launch_bounds__(256) k_boundary(int32_t n, const int32_t* __restrict__ xadj,
                                                  const int32_t* __restrict__ adj,
                                                  const int32_t* __restrict__ part, int32_t* __restrict__ bnd) {
  const int64_t v = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (v >= n) return;
  const int32_t p = part[v];
  int32_t b = 0;
  for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j)
    if (part[adj[j]] != p) { b = 1; break; }
  bnd[v] = b;
}

int launch_boundary(const DGraph& g, const int32_t* part, int32_t* bnd, hipStream_t st) {
  if (g.n <= 0) return 0;
  k_boundary<<<grid_for(g.n, 256), 256, 0, st>>>(g.n, g.xadj, g.adj, part, bnd);
  return hip_status();
}

// calculate_cutsize + calculate_partition_wgt (gkway/uncoarsen.hpp:398-424,
// gkway/cuda_check.hpp:61-73): grid-stride, 8 lanes per vertex; part weights
// gathered in an LDS histogram per block, then one global atomic per block
// and part (a block's share of the total weight fits 32 bits, the total
// being below 2^31); the cut reduced per wave in 64 bits.
// Blocks of the cut kernel: at most 2048, and at most 4 Mi part-weight slots.
inline unsigned cut_grid(int64_t n, int32_t k) {
  unsigned grid = grid_for(n * 8, 256);
  const int64_t cap = std::max<int64_t>(1, (int64_t(1) << 22) / std::max<int32_t>(k, 1));
  if (grid > 2048) grid = 2048;
  if (grid > cap) grid = static_cast<unsigned>(cap);
  return grid;
}

int64_t cut_scratch(int64_t n, int32_t k) { return int64_t(cut_grid(n, k)) * (int64_t(k) + 1); }

// Each block writes its share of the cut and its part-weight histogram to its
// own slots (bcut[b], bpw[b * k + q]); k_cut_sum adds them up.  No counter is
// shared between blocks.
__global__ void __launch_bounds__(256) k_cut(int32_t n, int32_t k, const int32_t* __restrict__ xadj,
                                             const int32_t* __restrict__ adj,
                                             const int32_t* __restrict__ w,
                                             const int32_t* __restrict__ vw,
                                             const int32_t* __restrict__ part,
                                             unsigned long long* __restrict__ bcut,
                                             unsigned long long* __restrict__ bpw) {
  extern __shared__ unsigned int hist32[];
  __shared__ long long wc[4];
  for (int32_t q = threadIdx.x; q < k; q += 256) hist32[q] = 0;
  __syncthreads();
  long long c = 0;
  const int sub = threadIdx.x & 7;
  for (int64_t v = (int64_t(xcd_block()) * 256 + threadIdx.x) / 8; v < n; v += int64_t(gridDim.x) * 32) {
    const int32_t p = part[v];
    for (int32_t j = xadj[v] + sub; j < xadj[v + 1]; j += 8) {
      const int32_t u = adj[j];
      if (v < u && part[u] != p) c += w[j];
    }
    if (sub == 0) atomicAdd(&hist32[p], (unsigned int)vw[v]);
  }
  c = wave_sum64(c);
  if (lane_id() == 0) wc[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) bcut[blockIdx.x] = (unsigned long long)(wc[0] + wc[1] + wc[2] + wc[3]);
  for (int32_t q = threadIdx.x; q < k; q += 256) bpw[int64_t(blockIdx.x) * k + q] = hist32[q];
}

// Block q < k: pw64[q] = sum of the blocks' slots for part q; block k: the
// cut.
__global__ void __launch_bounds__(256) k_cut_sum(int32_t nb, int32_t k, const unsigned long long* __restrict__ bcut,
                                                 const unsigned long long* __restrict__ bpw,
                                                 unsigned long long* __restrict__ cut,
                                                 unsigned long long* __restrict__ pw64) {
  __shared__ unsigned long long ws[4];
  const int32_t q = blockIdx.x;
  unsigned long long s = 0;
  if (q < k) {
    for (int32_t b = threadIdx.x; b < nb; b += 256) s += bpw[int64_t(b) * k + q];
  } else {
    for (int32_t b = threadIdx.x; b < nb; b += 256) s += bcut[b];
  }
  s = (unsigned long long)wave_sum64((long long)s);
  if (lane_id() == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = ws[0] + ws[1] + ws[2] + ws[3];
    if (q < k) pw64[q] = t; else *cut = t;
  }
}

int launch_cut(const DGraph& g, int32_t k, const int32_t* part, unsigned long long* cut,
               unsigned long long* pw64, unsigned long long* scratch, hipStream_t st,
               int64_t max_slots) {
  if (g.n <= 0) {
    fill<unsigned long long>(cut, 0ull, 1, st);
    fill<unsigned long long>(pw64, 0ull, k, st);
    return hip_status();
  }
  unsigned grid = cut_grid(g.n, k);
  const int64_t fit = max_slots / (int64_t(k) + 1);  // blocks whose slots fit
  if (fit < 1) return GK_ERR_LIMIT;
  if (int64_t(grid) > fit) grid = static_cast<unsigned>(fit);
  unsigned long long* bcut = scratch;
  unsigned long long* bpw = scratch + grid;
  k_cut<<<grid, 256, sizeof(unsigned int) * k, st>>>(g.n, k, g.xadj, g.adj, g.w, g.vw, part, bcut, bpw);
  k_cut_sum<<<k + 1, 256, 0, st>>>(int32_t(grid), k, bcut, bpw, cut, pw64);
  return hip_status();
}

// Sum of an int32 array into a 64-bit accumulator (block reduce, one atomic
// per block).
__global__ void __launch_bounds__(256) k_sum(const int32_t* __restrict__ a, int64_t n,
                                             unsigned long long* __restrict__ out) {
  __shared__ long long part_s[4];
  long long s = 0;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) s += a[i];
  s = wave_sum64(s);
  if (lane_id() == 0) part_s[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)(part_s[0] + part_s[1] + part_s[2] + part_s[3]));
}

int launch_sum(const int32_t* a, int64_t n, unsigned long long* out, hipStream_t st) {
  fill<unsigned long long>(out, 0ull, 1, st);
  if (n <= 0) return 0;
  unsigned grid = grid_for(n, 256);
  if (grid > 1024) grid = 1024;
  k_sum<<<grid, 256, 0, st>>>(a, n, out);
  return hip_status();
}

}  // namespace gkd
