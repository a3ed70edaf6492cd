// Kernels of the batched form: many graphs (jobs) packed into one CSR with
// global vertex ids, processed by one launch per stage.  Jobs never share an
// edge, so the coarsening kernels of coarsen.hip run on the pack unchanged;
// this file adds the per-job glue:
//   pack / unpack    descriptor-driven copies from per-job slices (local ids)
//                    into a pack (global ids), and of partitions back out;
//   job offsets      first coarse id / first coarse entry of every job;
//   rounds           independent-set selection with per-job caps and part
//                    weights as a bitmask, then one workgroup per job for the
//                    ordered move list, the best moves by (gain desc, id asc),
//                    the balanced prefix (balance.hpp) and the apply;
//   cut              final cut and part weights with per-job offsets.
// Every per-job result equals the single-graph path's: ids keep their order
// inside a job and every order-sensitive step is segmented by job.
#include <algorithm>

#include "balance.hpp"
#include "batch.hpp"
#include "kernels.hpp"
#include "prim.hpp"
#include "../../include/gkway.h"

namespace gkd {

static inline int hip_status() { return hipGetLastError() == hipSuccess ? 0 : GK_ERR_HIP; }

// Index j in [lo, hi) with off[j] <= i < off[j+1] (off is nondecreasing;
// the caller knows off[lo] <= i < off[hi]).
__device__ __forceinline__ int32_t job_in(const int32_t* __restrict__ off, int32_t lo, int32_t hi, int64_t i) {
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int32_t job_of(const int32_t* __restrict__ off, int32_t nj, int64_t i) {
  return job_in(off, 0, nj, i);
}

// Descriptor-driven copies: a block covers kPackPer * 256 consecutive items
// (vertices or entries) and looks up the jobs of its first and last item
// once; items of a block inside one job (nearly all) then need no per-item
// search.
constexpr int kPackPer = 8;

__device__ __forceinline__ void block_jobs(const int32_t* __restrict__ off, int32_t nj, int64_t total, int64_t i0,
                                           int32_t* sj) {
  if (threadIdx.x == 0) {
    const int64_t last = (i0 + int64_t(kPackPer) * 256 < total ? i0 + int64_t(kPackPer) * 256 : total) - 1;
    sj[0] = job_of(off, nj, i0);
    sj[1] = job_in(off, sj[0], nj, last);
  }
  __syncthreads();
}

// ---------------------------------------------------------------- pack
__global__ void __launch_bounds__(256) k_pack_vert(int32_t nj, const PackSrc* __restrict__ src,
                                                   const int32_t* __restrict__ v_off,
                                                   const int32_t* __restrict__ e_off,
                                                   const int32_t* __restrict__ c_off, int32_t ntot,
                                                   int32_t* __restrict__ xadj, int32_t* __restrict__ vw,
                                                   int32_t* __restrict__ vjob, int32_t* __restrict__ cmap) {
  __shared__ int32_t sj[2];
  const int64_t i0 = int64_t(blockIdx.x) * kPackPer * 256;
  block_jobs(v_off, nj, ntot, i0, sj);
  for (int r = 0; r < kPackPer; ++r) {
    const int64_t i = i0 + r * 256 + threadIdx.x;
    if (i >= ntot) break;
    const int32_t j = sj[0] == sj[1] ? sj[0] : job_in(v_off, sj[0], sj[1] + 1, i);
    const PackSrc& s = src[j];
    const int32_t li = int32_t(i - v_off[j]);
    xadj[i] = e_off[j] + (s.xadj[li] - s.xbase);
    vw[i] = s.vw[li];
    if (vjob) vjob[i] = j;
    // No source map: the coarse side has the fine level's ids (identity).
    if (cmap) cmap[i] = c_off[j] + (s.cmap ? s.cmap[li] - s.cbase : li);
  }
}

__global__ void __launch_bounds__(256) k_pack_edge(int32_t nj, const PackSrc* __restrict__ src,
                                                   const int32_t* __restrict__ v_off,
                                                   const int32_t* __restrict__ e_off, int32_t etot,
                                                   int32_t* __restrict__ adj, int32_t* __restrict__ w) {
  __shared__ int32_t sj[2];
  const int64_t i0 = int64_t(blockIdx.x) * kPackPer * 256;
  block_jobs(e_off, nj, etot, i0, sj);
  if (sj[0] == sj[1]) {
    const int32_t j = sj[0];
    const int32_t* __restrict__ sa = src[j].adj;
    const int32_t* __restrict__ sw = src[j].w;
    const int32_t shift = v_off[j] - src[j].abase;
    const int32_t eo = e_off[j];
#pragma unroll
    for (int r = 0; r < kPackPer; ++r) {
      const int64_t e = i0 + r * 256 + threadIdx.x;
      if (e < etot) {
        adj[e] = sa[e - eo] + shift;
        w[e] = sw[e - eo];
      }
    }
    return;
  }
  for (int r = 0; r < kPackPer; ++r) {
    const int64_t e = i0 + r * 256 + threadIdx.x;
    if (e >= etot) break;
    const int32_t j = job_in(e_off, sj[0], sj[1] + 1, e);
    const PackSrc& s = src[j];
    const int32_t le = int32_t(e - e_off[j]);
    adj[e] = v_off[j] + (s.adj[le] - s.abase);
    w[e] = s.w[le];
  }
}

// Coarse-side arrays of a refinement pack (partition and boundary flags of
// the next coarser level, one range per job).
__global__ void __launch_bounds__(256) k_pack_coarse(int32_t nj, const PackSrc* __restrict__ src,
                                                     const int32_t* __restrict__ c_off, int32_t ctot,
                                                     int32_t* __restrict__ cpart, int32_t* __restrict__ cbnd,
                                                     int32_t* __restrict__ cgain, int32_t* __restrict__ cdest) {
  __shared__ int32_t sj[2];
  const int64_t i0 = int64_t(blockIdx.x) * kPackPer * 256;
  block_jobs(c_off, nj, ctot, i0, sj);
  for (int r = 0; r < kPackPer; ++r) {
    const int64_t i = i0 + r * 256 + threadIdx.x;
    if (i >= ctot) break;
    const int32_t j = sj[0] == sj[1] ? sj[0] : job_in(c_off, sj[0], sj[1] + 1, i);
    const int32_t li = int32_t(i - c_off[j]);
    const PackSrc& s = src[j];
    cpart[i] = s.cpart[li];
    cbnd[i] = s.cbnd[li];
    if (cgain) {
      cgain[i] = s.cgain ? s.cgain[li] : 0;
      cdest[i] = s.cgain ? s.cdest[li] : -1;
    }
  }
}

int launch_pack(int32_t nj, const PackSrc* src, const int32_t* v_off, const int32_t* e_off, const int32_t* c_off,
                int32_t ntot, int32_t etot, int32_t ctot, int32_t* xadj, int32_t* adj, int32_t* w, int32_t* vw,
                int32_t* vjob, int32_t* cmap, int32_t* cpart, int32_t* cbnd, int32_t* cgain, int32_t* cdest,
                hipStream_t st) {
  constexpr int kTile = kPackPer * 256;
  if (ntot > 0)
    k_pack_vert<<<grid_for(ntot, kTile), 256, 0, st>>>(nj, src, v_off, e_off, c_off, ntot, xadj, vw, vjob, cmap);
  fill<int32_t>(xadj + ntot, etot, 1, st);
  if (etot > 0) k_pack_edge<<<grid_for(etot, kTile), 256, 0, st>>>(nj, src, v_off, e_off, etot, adj, w);
  if (cpart && ctot > 0)
    k_pack_coarse<<<grid_for(ctot, kTile), 256, 0, st>>>(nj, src, c_off, ctot, cpart, cbnd, cgain, cdest);
  return hip_status();
}

// expand_partition (gkway/uncoarsen.hpp:325-356) on a refinement pack, with
// resumed levels taking their carried state.
__global__ void __launch_bounds__(256) k_project_b(int32_t n, const int32_t* __restrict__ cmap,
                                                   const int32_t* __restrict__ cpart,
                                                   const int32_t* __restrict__ cbnd,
                                                   const int32_t* __restrict__ cgain,
                                                   const int32_t* __restrict__ cdest, int32_t* __restrict__ part,
                                                   int32_t* __restrict__ bnd, int32_t* __restrict__ gain,
                                                   int32_t* __restrict__ dest, int32_t* __restrict__ upd) {
  const int64_t v = int64_t(xcd_block()) * 256 + threadIdx.x;
  if (v >= n) return;
  const int32_t c = cmap[v];
  const int32_t p = cpart[c];
  const int32_t b = cbnd[c] ? 1 : 0;
  const int32_t d = cdest[c];
  part[v] = p;
  bnd[v] = b;
  if (d >= 0) {
    gain[v] = cgain[c];
    dest[v] = d;
    upd[v] = 0;
  } else {
    gain[v] = 0;
    dest[v] = p;
    upd[v] = b;
  }
}

int launch_project_b(int32_t n, const int32_t* cmap, const int32_t* cpart, const int32_t* cbnd, const int32_t* cgain,
                     const int32_t* cdest, int32_t* part, int32_t* bnd, int32_t* gain, int32_t* dest, int32_t* upd,
                     hipStream_t st) {
  if (n > 0) k_project_b<<<grid_for(n, 256), 256, 0, st>>>(n, cmap, cpart, cbnd, cgain, cdest, part, bnd, gain, dest, upd);
  return hip_status();
}

// Candidate bitmask from boundary flags and gains, one wave per word.
__global__ void __launch_bounds__(256) k_cand_init(int32_t n, const int32_t* __restrict__ bnd,
                                                   const int32_t* __restrict__ gain, uint64_t* __restrict__ cand) {
  const int64_t v = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const bool c = v < n && bnd[v] && gain[v] > 0;
  const uint64_t m = ballot(c);
  if (lane_id() == 0 && (v >> 6) <= (int64_t(n) - 1) / 64) cand[v >> 6] = m;
}

int launch_cand_init(int32_t n, const int32_t* bnd, const int32_t* gain, uint64_t* cand, hipStream_t st) {
  if (n > 0) k_cand_init<<<grid_for(n, 256), 256, 0, st>>>(n, bnd, gain, cand);
  return hip_status();
}

// Total vertex weight of every job (one block per job).
__global__ void __launch_bounds__(256) k_job_sum(const int32_t* __restrict__ v_off, const int32_t* __restrict__ vw,
                                                 long long* __restrict__ out) {
  __shared__ long long ws[4];
  const int32_t j = blockIdx.x;
  long long s = 0;
  for (int64_t i = v_off[j] + threadIdx.x; i < v_off[j + 1]; i += 256) s += vw[i];
  s = wave_sum64(s);
  if (lane_id() == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[j] = ws[0] + ws[1] + ws[2] + ws[3];
}

int launch_job_sum(int32_t nj, const int32_t* v_off, const int32_t* vw, long long* out, hipStream_t st) {
  if (nj > 0) k_job_sum<<<nj, 256, 0, st>>>(v_off, vw, out);
  return hip_status();
}

// ---------------------------------------------------------------- job offsets
// After clustering the pack: job j's coarse ids start at the coarse id of its
// first vertex in sorted order (jobs occupy consecutive sorted ranges because
// labels stay inside a job's id range).
__global__ void k_job_coff(int32_t nj, const int32_t* __restrict__ v_off, const int32_t* __restrict__ order,
                           const int32_t* __restrict__ cmap, int32_t nc, int32_t* __restrict__ c_off) {
  const int32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j > nj) return;
  if (j == nj) { c_off[j] = nc; return; }
  c_off[j] = cmap[order[v_off[j]]];
}

__global__ void k_job_eoff(int32_t nj, const int32_t* __restrict__ c_off, const int32_t* __restrict__ cxadj,
                           int32_t* __restrict__ ce_off) {
  const int32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j > nj) return;
  ce_off[j] = cxadj[c_off[j]];
}

int launch_job_offsets(int32_t nj, const int32_t* v_off, const int32_t* order, const int32_t* cmap, int32_t nc,
                       int32_t* c_off, hipStream_t st) {
  k_job_coff<<<grid_for(int64_t(nj) + 1, 256), 256, 0, st>>>(nj, v_off, order, cmap, nc, c_off);
  return hip_status();
}

int launch_job_eoffsets(int32_t nj, const int32_t* c_off, const int32_t* cxadj, int32_t* ce_off, hipStream_t st) {
  k_job_eoff<<<grid_for(int64_t(nj) + 1, 256), 256, 0, st>>>(nj, c_off, cxadj, ce_off);
  return hip_status();
}

// ---------------------------------------------------------------- unpack
// Refined partition of the pack -> per-job arrays.
__global__ void __launch_bounds__(256) k_unpack_part(int32_t nj, const UnpackDst* __restrict__ dst,
                                                     const int32_t* __restrict__ v_off, int32_t n,
                                                     const int32_t* __restrict__ part,
                                                     const int32_t* __restrict__ bnd) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int32_t j = job_of(v_off, nj, i);
  const UnpackDst d = dst[j];
  const int32_t li = int32_t(i - v_off[j]);
  if (d.part) d.part[li] = part[i];
  if (d.bnd) d.bnd[li] = bnd[i];
}

int launch_unpack_part(int32_t nj, const UnpackDst* dst, const int32_t* v_off, int32_t n, const int32_t* part,
                       const int32_t* bnd, hipStream_t st) {
  if (n > 0) k_unpack_part<<<grid_for(n, 256), 256, 0, st>>>(nj, dst, v_off, n, part, bnd);
  return hip_status();
}

// ---------------------------------------------------------------- rounds
// Selection of a refinement round: create_independent_move_buffer
// (gkway/uncoarsen.hpp:111-173) with the cap and part weights of the vertex's
// job, over the candidate bitmask kept by the gain kernels (boundary vertices
// with positive gain), so vertices outside the mask cost one bit of a
// coalesced load.  k_select_fit turns the candidates into the movable set
// (move fits the cap); k_select_list keeps the movable vertices with no
// smaller-id movable neighbour, as a bitmask; a scan turns
// the bitmask into a list ordered by vertex id, in which every job's selected
// vertices form one slice.  No kernel of the selection shares a counter
// between waves (same-address atomics serialise in L2).  vjob == null: one
// job.
// Candidate bitmask words read per wave by k_select_fit / k_select_list, and
// per 256-thread block.
constexpr int kSelectWords = 8;
constexpr int kSelectBlockWords = 4 * kSelectWords;
// Rows longer than this are tested by the whole wave, 64 neighbours a step.
constexpr int32_t kSelectShortRow = 32;

// Position of the k-th (0-based) set bit of w (k < popcount(w)).
__device__ __forceinline__ int nth_set_bit(uint64_t w, int k) {
  int bit = 0;
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) {
    const int pc = popc64((w >> bit) & ((1ull << s) - 1));
    if (pc <= k) {
      k -= pc;
      bit += s;
    }
  }
  return bit;
}

// The wave's bit of rank r among the set bits of its kSelectWords words: lane
// l holds word l and the running count incl of words 0..l.  Every lane
// calls it (wave shuffles); the result is meaningful for r < the total.
__device__ __forceinline__ void deal_bit(int incl, uint64_t word, int r, int* wi_out, int* bit_out) {
  int wi = 0;
#pragma unroll
  for (int s = kSelectWords / 2; s >= 1; s >>= 1) {
    const int x = __shfl(incl, wi + s - 1, 64);
    if (x <= r) wi += s;
  }
  const int before = __shfl(incl, wi > 0 ? wi - 1 : 0, 64);
  const uint64_t wd = __shfl(word, wi < kSelectWords ? wi : 0, 64);
  *wi_out = wi;
  *bit_out = nth_set_bit(wd, r - (wi > 0 ? before : 0));
}

// Movable set of a round: the candidates (boundary, positive gain) whose move
// fits their job's cap, as a bitmask.  Each wave reads kSelectWords words of
// the candidate mask with one load and deals their set bits out to its lanes,
// 64 per pass; the movable bits are gathered in LDS and written back for
// every word.  Thread 0 also clears the worklist counters and the round's
// move counter, which only the round kernel adds to.
__global__ void __launch_bounds__(256) k_select_fit(int32_t n, const int32_t* __restrict__ vw,
                                                    const int32_t* __restrict__ vjob,
                                                    const int32_t* __restrict__ dest,
                                                    const int32_t* __restrict__ pw,
                                                    const int32_t* __restrict__ pw_off,
                                                    const long long* __restrict__ cap,
                                                    const uint64_t* __restrict__ cand,
                                                    uint64_t* __restrict__ movable, int32_t* __restrict__ reset,
                                                    int32_t* __restrict__ active,
                                                    unsigned long long* __restrict__ abytes) {
  __shared__ unsigned long long mw[4][kSelectWords];
  __shared__ int32_t wtot[4];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    reset[0] = 0;
    reset[1] = 0;
    active[0] = 0;
  }
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  if (lane < kSelectWords) mw[wv][lane] = 0;
  __syncthreads();
  const int64_t nwords = (int64_t(n) + 63) >> 6;
  const int64_t w0 = ((int64_t(blockIdx.x) * 256 + threadIdx.x) >> 6) * kSelectWords;
  uint64_t word = 0;
  if (lane < kSelectWords && w0 + lane < nwords) word = cand[w0 + lane];
  const int incl = wave_incl_scan(popc64(word));
  const int total = __shfl(incl, kSelectWords - 1, 64);
  for (int base = 0; base < total; base += 64) {
    const int r = base + lane;
    int wi, bit;
    deal_bit(incl, word, r, &wi, &bit);
    if (r < total) {
      const int32_t v = int32_t(((w0 + wi) << 6) + bit);
      const int32_t j = vjob ? vjob[v] : 0;
      if ((long long)pw[pw_off[j] + dest[v]] + vw[v] <= cap[j]) atomicOr(&mw[wv][wi], 1ull << bit);
    }
  }
  if (lane == 0) wtot[wv] = total;
  __syncthreads();
  if (lane < kSelectWords && w0 + lane < nwords) movable[w0 + lane] = mw[wv][lane];
  // Algorithmic bytes: the block's candidate and movable words, and per
  // candidate its job, destination, part weight and vertex weight.
  if (abytes && threadIdx.x == 0)
    atomicAdd(&abytes[blockIdx.x % kSelectByteSlots],
              (unsigned long long)(16 * kSelectBlockWords + 16 * (wtot[0] + wtot[1] + wtot[2] + wtot[3])));
}

// Selection proper: a movable vertex is selected when no smaller-id
// neighbour is movable (one bit per neighbour).  Same dealing of the movable
// mask's bits to lanes as k_select_fit.  Only vertices with gain >= thr[job]
// are tested; the others get a bit in defbits, and the round kernel tests
// them only when fewer than M were selected (every key above the threshold
// beats every key below it).  A lane tests a row of up to kSelectShortRow
// entries itself; longer rows of a dealing pass are then tested one after
// another by the whole wave, 64 neighbours per step.  Results are bitmask
// words (selbits, defbits), all written, so no counter is shared between
// waves.
__global__ void __launch_bounds__(256) k_select_list(int32_t n, const int32_t* __restrict__ xadj,
                                                     const int32_t* __restrict__ adj,
                                                     const int32_t* __restrict__ vjob,
                                                     const int32_t* __restrict__ gain,
                                                     const int32_t* __restrict__ thr,
                                                     const uint64_t* __restrict__ movable,
                                                     uint64_t* __restrict__ selbits, uint64_t* __restrict__ defbits,
                                                     int32_t* __restrict__ bsum,
                                                     unsigned long long* __restrict__ abytes) {
  __shared__ unsigned long long sw[4][2][kSelectWords];
  __shared__ int32_t wsum[4];
  __shared__ long long wbytes[4];
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  if (lane < kSelectWords) {
    sw[wv][0][lane] = 0;
    sw[wv][1][lane] = 0;
  }
  __syncthreads();
  const int64_t nwords = (int64_t(n) + 63) >> 6;
  const int64_t w0 = ((int64_t(blockIdx.x) * 256 + threadIdx.x) >> 6) * kSelectWords;
  uint64_t word = 0;
  if (lane < kSelectWords && w0 + lane < nwords) word = movable[w0 + lane];
  const int incl = wave_incl_scan(popc64(word));
  const int total = __shfl(incl, kSelectWords - 1, 64);
  long long tested = 0;  // row entries of the tested vertices
  for (int base = 0; base < total; base += 64) {
    const int r = base + lane;
    int wi, bit;
    deal_bit(incl, word, r, &wi, &bit);
    int32_t v = 0, s = 0, e = 0;
    int which = -1;  // 0 selected, 1 deferred
    bool long_row = false;
    if (r < total) {
      v = int32_t(((w0 + wi) << 6) + bit);
      const int32_t j = vjob ? vjob[v] : 0;
      s = xadj[v];
      e = xadj[v + 1];
      if (gain[v] < thr[j]) {
        which = 1;
      } else if (e - s > kSelectShortRow) {
        tested += e - s;
        long_row = true;
      } else {
        tested += e - s;
        // eight neighbours per step: their ids, then their movable bits
        bool sel = true;
        for (int32_t x = s; x < e && sel; x += 8) {
          int32_t u[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) u[i] = x + i < e ? adj[x + i] : v;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const uint64_t mwd = movable[u[i] >> 6];
            if (u[i] < v && ((mwd >> (u[i] & 63)) & 1ull)) sel = false;
          }
        }
        if (sel) which = 0;
      }
    }
    // The pass's long rows, one at a time, by the whole wave.
    uint64_t lm = ballot(long_row);
    while (lm) {
      const int src = __ffsll(static_cast<long long>(lm)) - 1;
      lm &= lm - 1;
      const int32_t lv = __shfl(v, src, 64), ls = __shfl(s, src, 64), le = __shfl(e, src, 64);
      bool blocked = false;
      for (int32_t x0 = ls; x0 < le && !blocked; x0 += 64) {
        const int32_t x = x0 + lane;
        bool b = false;
        if (x < le) {
          const int32_t u = adj[x];
          b = u < lv && ((movable[u >> 6] >> (u & 63)) & 1ull);
        }
        blocked = ballot(b) != 0;
      }
      if (lane == src && !blocked) which = 0;
    }
    if (which >= 0) atomicOr(&sw[wv][which][wi], 1ull << bit);
  }
  __syncthreads();
  int c = 0;
  if (lane < kSelectWords && w0 + lane < nwords) {
    selbits[w0 + lane] = sw[wv][0][lane];
    defbits[w0 + lane] = sw[wv][1][lane];
    c = popc64(sw[wv][0][lane]);
  }
  // selected vertices of the block's kSelectBlockWords words
  c = wave_sum(c);
  tested = wave_sum64(tested);
  if (lane == 0) {
    wsum[wv] = c;
    // movable words read, selected and deferred words written; per movable
    // vertex its row bounds, job and gain, per tested entry the neighbour id
    // and one movable bit
    wbytes[wv] = 8LL * total + 4LL * tested + tested / 8;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    bsum[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (abytes)
      atomicAdd(&abytes[blockIdx.x % kSelectByteSlots],
                (unsigned long long)(24LL * kSelectBlockWords + wbytes[0] + wbytes[1] + wbytes[2] + wbytes[3]));
  }
}

// Ordered list of the selected vertices: every word's first list slot
// (wpos) is the number of selected vertices before it.  Block b covers
// kScatterWords words; its offset is the sum of the selection blocks' counts
// (bsum) before its first word, then a block scan over its words.  One launch,
// no shared counter.  wpos[nwords] receives the total.
constexpr int kScatterWords = 2048;

__global__ void __launch_bounds__(256) k_sel_scatter(int64_t nwords, const uint64_t* __restrict__ selbits,
                                                     const int32_t* __restrict__ bsum, int32_t* __restrict__ wpos,
                                                     int32_t* __restrict__ list,
                                                     unsigned long long* __restrict__ abytes) {
  __shared__ int lds[8];
  const int64_t wb = int64_t(blockIdx.x) * kScatterWords;
  // offset: selection blocks [0, wb / kSelectBlockWords)
  const int64_t nb = wb / kSelectBlockWords;
  int off = 0;
  for (int64_t i = threadIdx.x; i < nb; i += 256) off += bsum[i];
  int tot;
  (void)block_excl_scan_256(off, &tot, lds);
  off = tot;
  constexpr int kPer = kScatterWords / 256;
  const int64_t w0 = wb + int64_t(threadIdx.x) * kPer;
  uint64_t b[kPer];
  int c = 0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    b[i] = w0 + i < nwords ? selbits[w0 + i] : 0;
    c += popc64(b[i]);
  }
  int r = off + block_excl_scan_256(c, &tot, lds);
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    if (w0 + i < nwords) wpos[w0 + i] = r;
    uint64_t x = b[i];
    while (x) {
      const int bit = __ffsll(static_cast<long long>(x)) - 1;
      x &= x - 1;
      list[r++] = int32_t(((w0 + i) << 6) + bit);
    }
  }
  if (wb + kScatterWords >= nwords && threadIdx.x == 255) wpos[nwords] = r;
  // Algorithmic bytes: the block's selected words read and word positions
  // written, and one list slot per selected vertex (tot: the block's count).
  if (abytes && threadIdx.x == 0)
    atomicAdd(&abytes[blockIdx.x % kSelectByteSlots], (unsigned long long)(12LL * kScatterWords + 4LL * tot));
}

// List slot of vertex x (the number of selected vertices before it).
__device__ __forceinline__ int32_t sel_slot(const int32_t* __restrict__ wpos, const uint64_t* __restrict__ selbits,
                                            int64_t x) {
  const int64_t w = x >> 6;
  const int b = int(x & 63);
  return wpos[w] + (b ? popc64(selbits[w] & ((1ull << b) - 1)) : 0);
}

int select_scratch_ints(int64_t n) {
  const int64_t nwords = (n + 63) / 64;
  return static_cast<int>((nwords + kSelectBlockWords - 1) / kSelectBlockWords + 1);
}

int launch_select_list(const DGraph& g, const int32_t* vjob, const int32_t* gain, const int32_t* dest,
                       const int32_t* pw, const int32_t* pw_off, const long long* cap, const int32_t* thr,
                       const uint64_t* cand, const SelectBufs& sb, int32_t* reset, int32_t* active,
                       hipStream_t st) {
  if (g.n <= 0) return 0;
  const int64_t nwords = (int64_t(g.n) + 63) / 64;
  const int64_t waves = (nwords + kSelectWords - 1) / kSelectWords;
  const unsigned grid = grid_for(waves * 64, 256);
  gk_launch(k_select_fit, dim3(grid), dim3(256), 0, st, g.n, g.vw, vjob, dest, pw, pw_off, cap, cand, sb.movable,
            reset, active, sb.abytes);
  gk_launch(k_select_list, dim3(grid), dim3(256), 0, st, g.n, g.xadj, g.adj, vjob, gain, thr, sb.movable, sb.selbits,
            sb.defbits, sb.scratch, sb.abytes);
  gk_launch(k_sel_scatter, dim3(grid_for(nwords, kScatterWords)), dim3(256), 0, st, nwords, sb.selbits, sb.scratch,
            sb.wpos, sb.list, sb.abytes);
  return hip_status();
}

struct RoundLds {
  BalanceLds b;  // b.sk doubles as the 2048-bin radix histogram
  int32_t mdst[kBalMax];  // each ordered move's destination
  int32_t ht[1024];  // the apply's table of first occurrences (cleared at the start)
  int32_t ws[40];
  int32_t misc[4];
  int32_t lz[4];  // deferred tests: selected count; next threshold's digit; next threshold
};

// Gain threshold of the next round's selection: aim at about
// kLazyFactor * M tested vertices per job.
constexpr int kLazyFactor = 2;

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
  return wave_reduce(x, [](uint32_t a, uint32_t b) { return b < a ? b : a; });
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
  return wave_reduce(x, [](uint32_t a, uint32_t b) { return b > a ? b : a; });
}

// Bits needed to hold values 0..x (at least 1).
__device__ __forceinline__ int bits_of(uint32_t x) { return x ? 32 - __clz(x) : 1; }

// Move key of a selected vertex: ascending order = gain descending, id
// ascending (the mvRequest order of gkway/move_request.hpp:36-41).
__device__ __forceinline__ uint64_t move_key(const int32_t* __restrict__ gain, int32_t v) {
  return (uint64_t(uint32_t(0x7fffffff - gain[v])) << 32) | uint64_t(uint32_t(v));
}

// One refinement round of one job per workgroup, after k_select_list:
//   1. the job's selected vertices: its slice of the selection list (plus,
//      when fewer than M and a gain threshold deferred some, those of the
//      deferred vertices that pass the test);
//   2. the min(count, M) best moves by (gain desc, id asc): all of them when
//      count <= M, else an exact radix select on a compressed key -- keys are
//      distinct, so exactly M are taken -- whose first digit also gives the
//      gain threshold of the job's next round;
//   3. LDS bitonic sort, balanced prefix (balance_prefix);
//   4. apply of the prefix: part, update flags and the gain worklist
//      (apply_sequence_move, gkway/uncoarsen.hpp:290-323).
// A job with nothing selected is left untouched, so extra rounds after the
// last productive one change nothing.
__global__ void __launch_bounds__(kBalThreads) k_round_b(
    const int32_t* __restrict__ v_off, int32_t M, int32_t kmax, const int32_t* __restrict__ kk,
    const long long* __restrict__ cap,
    const int32_t* __restrict__ pw_off, const int32_t* __restrict__ xadj, const int32_t* __restrict__ adj,
    int32_t* __restrict__ part, const int32_t* __restrict__ dest, const int32_t* __restrict__ gain,
    const int32_t* __restrict__ vw, int32_t* __restrict__ pw, long long* __restrict__ cut,
    int32_t* __restrict__ rounds, long long* __restrict__ moved, const SelectBufs sb, int32_t* __restrict__ xlist,
    uint32_t* __restrict__ hkey, int32_t* __restrict__ upd, int32_t* __restrict__ wl, int32_t* __restrict__ wlc,
    int32_t* __restrict__ lastcnt, int32_t* __restrict__ active, uint64_t* __restrict__ cand,
    int32_t* __restrict__ thr, long long* __restrict__ paths, int32_t* __restrict__ order,
    long long* __restrict__ abytes) {
  __shared__ RoundLds R;
  extern __shared__ int32_t dpw[];  // 2 * kmax ints: part weight changes, then the part weights
  int32_t* spw = dpw + kmax;
  const int t = threadIdx.x;
  const int32_t j = blockIdx.x;
  const int32_t vb = v_off[j], ve = v_off[j + 1];
  // Loads that the end of the round needs are issued now, off its critical
  // path: the job's first kBalThreads part weights (stored to LDS before the
  // balanced prefix) and its counters (thread 0).
  const int32_t kj = kk[j];
  const int32_t* pwj = pw + pw_off[j];
  const int32_t pw_t = t < kj ? pwj[t] : 0;
  R.ht[t] = -1;  // kBalThreads == table slots; the barriers before the apply order it
  long long cut0 = 0, moved0 = 0, ab0 = 0;
  int32_t rounds0 = 0;
  if (t == 0) {
    cut0 = cut[j];
    moved0 = moved[j];
    rounds0 = rounds[j];
    if (abytes) ab0 = abytes[j];
    R.misc[3] = sel_slot(sb.wpos, sb.selbits, vb);
    R.misc[1] = sel_slot(sb.wpos, sb.selbits, ve);
    R.misc[2] = lastcnt[j];
  }
  __syncthreads();
  const int32_t* seg = sb.list + R.misc[3];
  int32_t count = R.misc[1] - R.misc[3];
  const int32_t prev = R.misc[2];
  const bool deferred = count < M && thr[j] > 1;
  if (paths && deferred && t == 0) paths[3 * j + 2] += 1;
  if (deferred) {
    // The threshold left fewer than M selected: test the deferred (lower
    // gain) movable vertices of the job too, into the job's segment of xlist
    // after the selected ones, so that it holds every selected vertex.
    int32_t* xs = xlist + vb;
    for (int32_t i = t; i < count; i += kBalThreads) xs[i] = seg[i];
    if (t == 0) R.lz[0] = count;
    __syncthreads();
    for (int64_t w = (int64_t(vb) >> 6) + t; ve > vb && w <= ((int64_t(ve) - 1) >> 6); w += kBalThreads) {
      const int64_t lo = (vb > (w << 6) ? vb : (w << 6)) - (w << 6);
      const int64_t hi = (ve < (w << 6) + 64 ? ve : (w << 6) + 64) - (w << 6);
      uint64_t db = sb.defbits[w] & ((hi - lo == 64) ? ~0ull : (((1ull << (hi - lo)) - 1) << lo));
      while (db) {
        const int bit = __ffsll(static_cast<long long>(db)) - 1;
        db &= db - 1;
        const int32_t v = int32_t((w << 6) + bit);
        bool blocked = false;
        for (int32_t x = xadj[v]; x < xadj[v + 1] && !blocked; ++x) {
          const int32_t u = adj[x];
          blocked = u < v && ((sb.movable[u >> 6] >> (u & 63)) & 1ull);
        }
        if (!blocked) xs[atomicAdd(&R.lz[0], 1)] = v;
      }
    }
    __syncthreads();
    count = R.lz[0];
    seg = xs;
  }
  if (t == 0) {
    lastcnt[j] = count;
    if (count) atomicAdd(active, count);
    R.lz[2] = 1;  // next threshold: test everything unless set below
  }
  if (count == 0) {
    if (t == 0) thr[j] = 1;
    // Nothing selected: the job's state (partition, gains, part weights) is
    // unchanged, so every later round of this level selects nothing either.
    // The first time, its candidate bits are dropped so that the selection
    // stops re-checking them (with one job the rounds end at the next check
    // anyway).
    if (prev != 0 && gridDim.x > 1) {
      const int64_t a = v_off[j], b = v_off[j + 1];
      for (int64_t w = (a >> 6) + t; b > a && w <= ((b - 1) >> 6); w += kBalThreads) {
        const int64_t lo = (a > (w << 6) ? a : (w << 6)) - (w << 6);
        const int64_t hi = (b < (w << 6) + 64 ? b : (w << 6) + 64) - (w << 6);
        const uint64_t m = (hi - lo == 64) ? ~0ull : (((1ull << (hi - lo)) - 1) << lo);
        if (m == ~0ull) cand[w] = 0;
        else atomicAnd(reinterpret_cast<unsigned long long*>(&cand[w]), (unsigned long long)~m);
      }
    }
    return;
  }
  const int32_t B = count < M ? count : M;
  bool narrow = false;  // keys already sorted as 32-bit compressed keys
  // 2. the B best keys into L.sk[0, B)
  if (count <= M) {
    R.b.sk[t] = t < count ? move_key(gain, seg[t]) : ~0ull;
  } else {
    // Exact top-B by radix select on a compressed key: (gain key - its
    // minimum) above (id - first id of the job); the key is monotone in
    // (gain desc, id asc) and has only as many bits as the ranges need.
    // Pass 0 gathers the gain keys and finds their range; the passes then
    // use 11-bit digits from the top.  Up to kKeep keys per thread stay in
    // registers across the passes (with their vertices); more go through the
    // job's slice of hkey.
    constexpr int kKeep = 4;
    const bool inreg = count <= kKeep * kBalThreads;
    uint32_t* hk = hkey + vb;
    uint32_t hr[kKeep];
    int32_t vr[kKeep];
    uint32_t lo = 0xffffffffu, hi = 0;
    if (inreg) {
#pragma unroll
      for (int r = 0; r < kKeep; ++r) {
        const int32_t i = t + r * kBalThreads;
        vr[r] = i < count ? seg[i] : vb;
      }
#pragma unroll
      for (int r = 0; r < kKeep; ++r) {
        const uint32_t h = uint32_t(0x7fffffff - gain[vr[r]]);
        hr[r] = h;
        if (t + r * kBalThreads < count) {
          lo = h < lo ? h : lo;
          hi = h > hi ? h : hi;
        }
      }
    } else {
      for (int32_t i = t; i < count; i += kBalThreads) {
        const uint32_t h = uint32_t(0x7fffffff - gain[seg[i]]);
        hk[i] = h;
        lo = h < lo ? h : lo;
        hi = h > hi ? h : hi;
      }
    }
    // fn(gain key, vertex) for every selected vertex this thread holds
    auto each = [&](auto&& fn) {
      if (inreg) {
#pragma unroll
        for (int r = 0; r < kKeep; ++r)
          if (t + r * kBalThreads < count) fn(hr[r], vr[r]);
      } else {
        for (int32_t i = t; i < count; i += kBalThreads) fn(hk[i], seg[i]);
      }
    };
    lo = wave_min_u32(lo);
    hi = wave_max_u32(hi);
    if (lane_id() == 0) {
      R.ws[t >> 6] = int32_t(lo);
      R.ws[16 + (t >> 6)] = int32_t(hi);
    }
    __syncthreads();
    if (t == 0) {
      uint32_t a2 = 0xffffffffu, b2 = 0;
      for (int w = 0; w < kBalThreads / 64; ++w) {
        a2 = uint32_t(R.ws[w]) < a2 ? uint32_t(R.ws[w]) : a2;
        b2 = uint32_t(R.ws[16 + w]) > b2 ? uint32_t(R.ws[16 + w]) : b2;
      }
      R.ws[32] = int32_t(a2);
      R.ws[33] = int32_t(b2);
    }
    __syncthreads();
    const uint32_t hmin = uint32_t(R.ws[32]), hrange = uint32_t(R.ws[33]) - hmin;
    const int ib = bits_of(uint32_t(v_off[j + 1] - vb - 1));
    const int tb = bits_of(hrange) + ib;
    auto ckey = [&](uint32_t h, int32_t v) -> uint64_t {
      return (uint64_t(h - hmin) << ib) | uint64_t(uint32_t(v - vb));
    };
    constexpr int kDigit = 11;
    int32_t* hist = reinterpret_cast<int32_t*>(R.b.sk);  // 2048 ints, free until the keys are collected
    uint64_t prefix = 0, pmask = 0;
    int32_t need = B;
    for (int shift = ((tb + kDigit - 1) / kDigit - 1) * kDigit; shift >= 0; shift -= kDigit) {
      hist[2 * t] = 0;
      hist[2 * t + 1] = 0;
      __syncthreads();
      each([&](uint32_t h, int32_t v) {
        const uint64_t key = ckey(h, v);
        if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 2047], 1);
      });
      __syncthreads();
      // The digit where the running count reaches `need`: a block scan over
      // the 2048 bins, two per thread.
      const int32_t h0 = hist[2 * t], h1 = hist[2 * t + 1];
      int tot;
      const int32_t ex = block_excl_scan_1024(h0 + h1, &tot, R.ws);
      if (ex < need && need <= ex + h0) {
        R.misc[0] = 2 * t;
        R.misc[1] = ex;
        R.misc[3] = h0;
      } else if (ex + h0 < need && need <= ex + h0 + h1) {
        R.misc[0] = 2 * t + 1;
        R.misc[1] = ex + h0;
        R.misc[3] = h1;
      }
      // First digit: also the digit holding rank kLazyFactor * M, for the
      // next round's gain threshold.
      const bool lazy = pmask == 0 && count > kLazyFactor * M;
      const int32_t r2 = kLazyFactor * M;
      if (lazy && ex < r2 && r2 <= ex + h0) R.lz[1] = 2 * t;
      else if (lazy && ex + h0 < r2 && r2 <= ex + h0 + h1) R.lz[1] = 2 * t + 1;
      __syncthreads();
      if (lazy && t == 0) {
        // every key in digits <= lz[1] has a gain key <= hu, i.e. a gain >= 0x7fffffff - hu
        const uint64_t top = ((uint64_t(R.lz[1]) + 1) << shift) - 1;
        const uint64_t hu = uint64_t(hmin) + (top >> ib);
        R.lz[2] = hu >= 0x7ffffffeull ? 1 : int32_t(0x7fffffffull - hu);
      }
      need -= R.misc[1];
      prefix |= uint64_t(R.misc[0]) << shift;
      pmask |= 2047ull << shift;
      // Every key of the chosen bucket is needed: the B best are the keys
      // whose digits so far are <= prefix, and the lower digits can stay
      // unread.  (Only R.misc[2], R.b.sk and R.b.sv are written next; none
      // of them is read in this pass after its last barrier.)
      if (need == R.misc[3]) break;
      __syncthreads();
    }
    // (key & pmask) <= prefix holds for exactly the B best keys: prefix is
    // the B-th smallest key's leading digits and the keys are distinct.
    // Compressed keys of at most 32 bits are sorted as 32-bit words (half the
    // shuffle and LDS traffic of the 64-bit keys) and expanded afterwards.
    narrow = tb <= 32;
    uint32_t* nk = reinterpret_cast<uint32_t*>(R.b.sv);  // free until balance_prefix
    if (t == 0) R.misc[2] = 0;
    if (narrow) nk[t] = 0xffffffffu;
    else R.b.sk[t] = ~0ull;
    __syncthreads();
    each([&](uint32_t h, int32_t v) {
      const uint64_t ck = ckey(h, v);
      if ((ck & pmask) <= prefix) {
        const int slot = atomicAdd(&R.misc[2], 1);
        if (narrow) nk[slot] = uint32_t(ck);
        else R.b.sk[slot] = (uint64_t(h) << 32) | uint64_t(uint32_t(v));
      }
    });
    if (narrow) {
      __syncthreads();
      int P = 128;
      while (P < B) P <<= 1;
      bitonic_lds(nk, P);
      const uint32_t ck = nk[t];
      const uint32_t idm = uint32_t((1ull << ib) - 1);
      R.b.sk[t] = t < B ? (uint64_t(hmin + (ck >> ib)) << 32) | uint64_t(uint32_t(vb) + (ck & idm)) : ~0ull;
    }
  }
  __syncthreads();
  // 3. sort and balanced prefix
  if (!narrow) lds_bitonic(R.b, B);
  if (t < kj) spw[t] = pw_t;
  for (int32_t q = t + kBalThreads; q < kj; q += kBalThreads) spw[q] = pwj[q];
  BalanceOpt bo;
  bo.spw = spw;
  bo.mdst = R.mdst;
  bo.gain_in_key = true;
  bo.cut_given = true;
  bo.cut_now = cut0;
  const int pos = balance_prefix(R.b, dpw, kj, B, 0xffffffffull, part, dest, gain, vw, pw + pw_off[j], cap[j],
                                 nullptr, nullptr, &cut[j], bo);
  // Optional trace of the round: which ordering path ran (top-M radix
  // select; compressed keys wider than 32 bits) and the ordered moves with B
  // and pos.  R.b.sv keeps the ordered vertices through the apply below.
  if (paths && t == 0 && count > M) {
    paths[3 * j] += 1;
    if (!narrow) paths[3 * j + 1] += 1;
  }
  if (order) {
    int32_t* o = order + int64_t(j) * (M + 2);
    if (t < B) o[2 + t] = R.b.sv[t];
    if (t == 0) {
      o[0] = B;
      o[1] = pos;
    }
  }
  // 4. apply moves 0..pos (balance_prefix read part[] before its last barrier):
  // the movers and their neighbours are items [0, tot) -- mover m owns
  // [istart[m], istart[m] + deg + 1), first itself, then its row -- dealt out
  // in passes of kApplyPer items per thread.  An item's vertex joins the gain
  // worklist if its update flag was clear.  One workgroup (one job) owns the
  // worklist count, which the selection zeroed: it appends at a running
  // offset and stores the count at the end; several jobs append with one
  // global atomic per pass.
  constexpr int kApplyPer = 8;
  const bool sole = gridDim.x == 1;
  int32_t* istart = R.b.sw;     // free after balance_prefix
  int32_t* rstart = R.b.cover;  // likewise
  int tot = 0;
  // Items [a0, a0 + kApplyPer) of [0, tot): their vertices (-1 past the end)
  // and the bit mask of the first occurrences.  With a table ht (LDS, kHt
  // slots, all -1, more than twice the items): first occurrences within the
  // round by insertion into ht, the update flag then set by a plain store --
  // every flag is clear when a round starts (the gain pass of the previous
  // round cleared those it set).  Without: by the update flag's atomic swap.
  constexpr int kHt = sizeof(R.ht) / sizeof(int32_t);
  static_assert(kHt == 1024, "10-bit table index");
  auto take = [&](int a0, int32_t (&us)[kApplyPer], int32_t* ht) -> uint32_t {
    uint32_t win = 0;
#pragma unroll
    for (int i = 0; i < kApplyPer; ++i) us[i] = -1;
    if (a0 >= tot) return 0;
    int lo = 0, hi = pos;  // last mover m with istart[m] <= a0
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (istart[mid] <= a0) lo = mid; else hi = mid - 1;
    }
    int m = lo;
#pragma unroll
    for (int i = 0; i < kApplyPer; ++i) {
      const int e2 = a0 + i;
      if (e2 < tot) {
        while (m < pos && istart[m + 1] <= e2) ++m;
        const int off = e2 - istart[m];
        us[i] = off == 0 ? R.b.sv[m] : adj[rstart[m] + off - 1];
      }
    }
    if (ht) {
#pragma unroll
      for (int i = 0; i < kApplyPer; ++i) {
        if (us[i] < 0) continue;
        uint32_t sl = (uint32_t(us[i]) * 0x9E3779B1u) >> 22;  // 10 bits: kHt slots
        while (true) {
          const int32_t o = atomicCAS(&ht[sl], -1, us[i]);
          if (o == -1) {
            win |= 1u << i;
            upd[us[i]] = 1;
            break;
          }
          if (o == us[i]) break;
          sl = (sl + 1) & (kHt - 1);
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < kApplyPer; ++i)
        if (us[i] >= 0 && atomicExch(&upd[us[i]], 1) == 0) win |= 1u << i;
    }
    return win;
  };
  int32_t wrun = 0;
  if (pos < 64) {
    // At most 64 movers (the usual case once the parts sit at the cap): wave
    // 0 alone, with wave-level scans and no block barriers.
    if (t < 64) {
      const bool mv = t <= pos;
      int32_t es = 0, cnt = 0;
      if (mv) {
        const int32_t v = R.b.sv[t];
        part[v] = R.mdst[t];
        es = xadj[v];
        cnt = xadj[v + 1] - es + 1;
      }
      const int32_t incl = seg_incl_scan<64>(cnt);
      tot = __shfl(incl, 63, 64);
      if (mv) {
        istart[t] = incl - cnt;
        rstart[t] = es;
      }
      // first occurrences in LDS when the items fit the table
      int32_t* ht = tot <= kHt / 2 ? R.ht : nullptr;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (int pb = 0; pb < tot; pb += kApplyPer * 64) {
        int32_t us[kApplyPer];
        const uint32_t win = take(pb + kApplyPer * t, us, ht);
        const int c = __popc(win);
        const int wincl = seg_incl_scan<64>(c);
        const int wtot = __shfl(wincl, 63, 64);
        int base = wrun;
        if (!sole) {
          int b0 = 0;
          if (t == 0 && wtot > 0) b0 = atomicAdd(&wlc[0], wtot);
          base = __shfl(b0, 0, 64);
        }
        int r = base + wincl - c;
#pragma unroll
        for (int i = 0; i < kApplyPer; ++i)
          if ((win >> i) & 1u) wl[r++] = us[i];
        wrun += wtot;
      }
    }
  } else {
    const bool mv = t <= pos;
    int32_t es = 0, deg = 0;
    if (mv) {
      const int32_t v = R.b.sv[t];
      part[v] = R.mdst[t];
      es = xadj[v];
      deg = xadj[v + 1] - es;
    }
    const int ib = block_excl_scan_1024(mv ? deg + 1 : 0, &tot, R.ws);
    if (mv) {
      istart[t] = ib;
      rstart[t] = es;
    }
    __syncthreads();
    for (int pb = 0; pb < tot; pb += kApplyPer * kBalThreads) {
      int32_t us[kApplyPer];
      const uint32_t win = take(pb + kApplyPer * t, us, nullptr);
      int wtot;
      const int wex = block_excl_scan_1024(__popc(win), &wtot, R.ws);
      if (t == 0) {
        if (sole) {
          R.misc[0] = wrun;
          wrun += wtot;
        } else if (wtot > 0) {
          R.misc[0] = atomicAdd(&wlc[0], wtot);
        }
      }
      __syncthreads();
      const int wbase = R.misc[0] + wex;
      int r = 0;
#pragma unroll
      for (int i = 0; i < kApplyPer; ++i)
        if ((win >> i) & 1u) wl[wbase + r++] = us[i];
      __syncthreads();  // R.misc[0] is reused by the next pass
    }
  }
  if (t == 0) {
    if (sole) wlc[0] = wrun;
    thr[j] = R.lz[2];
    rounds[j] = rounds0 + 1;
    moved[j] = moved0 + pos + 1;
    // Algorithmic bytes of the round: each selected vertex's id and gain,
    // each ordered move's weight, gain, part and destination, the job's part
    // weights, and per applied item (a mover or a neighbour) its id, update
    // flag and worklist slot.
    if (abytes) abytes[j] = ab0 + 8LL * count + 16LL * B + 4LL * kj + 12LL * tot;
  }
}

int launch_round_b(int32_t nj, int32_t kmax, const int32_t* v_off, int32_t M, const int32_t* kk, const long long* cap,
                   const int32_t* pw_off, const DGraph& g, int32_t* part, const int32_t* dest, const int32_t* gain,
                   int32_t* pw, long long* cut, int32_t* rounds, long long* moved, const SelectBufs& sb,
                   int32_t* xlist, uint32_t* hkey, int32_t* upd, int32_t* wl, int32_t* wlc, int32_t* lastcnt,
                   int32_t* active, uint64_t* cand, int32_t* thr, long long* paths, int32_t* order,
                   long long* abytes, hipStream_t st) {
  if (nj <= 0) return 0;
  if (M > kBalMax) return GK_ERR_LIMIT;
  gk_launch(k_round_b, dim3(nj), dim3(kBalThreads), uint32_t(2 * sizeof(int32_t) * kmax), st, v_off, M, kmax, kk, cap,
            pw_off,
            g.xadj, g.adj, part, dest, gain, g.vw, pw, cut, rounds, moved, sb, xlist, hkey, upd, wl, wlc, lastcnt,
            active, cand, thr, paths, order, abytes);
  return hip_status();
}

// Largest row length of a graph (atomicMax per wave into *out, zeroed here).
// Grid-stride; one atomic per block.
__global__ void __launch_bounds__(256) k_max_deg(int32_t n, const int32_t* __restrict__ xadj, int32_t* __restrict__ out) {
  __shared__ int wm[4];
  int d = 0;
  for (int64_t v = int64_t(blockIdx.x) * 256 + threadIdx.x; v < n; v += int64_t(gridDim.x) * 256) {
    const int x = xadj[v + 1] - xadj[v];
    d = x > d ? x : d;
  }
  d = wave_max(d);
  if (lane_id() == 0) wm[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int m = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
    if (m > 0) atomicMax(out, m);
  }
}

int launch_max_deg(const DGraph& g, int32_t* out, hipStream_t st) {
  fill<int32_t>(out, 0, 1, st);
  unsigned grid = grid_for(g.n, 256);
  if (grid > 1024) grid = 1024;
  if (g.n > 0) k_max_deg<<<grid, 256, 0, st>>>(g.n, g.xadj, out);
  return hip_status();
}

// Final cut and part weights (calculate_cutsize + calculate_partition_wgt,
// gkway/uncoarsen.hpp:398-424, gkway/cuda_check.hpp:61-73) per chunk of one
// job's vertices: the block sums its vertex weights per part in an LDS
// histogram and its cut edges in registers, then adds them to the job's
// totals with one atomic per touched part and one for the cut.
__global__ void __launch_bounds__(256) k_cut_chunks(const int32_t* __restrict__ cj, const int32_t* __restrict__ cs,
                                                    const int32_t* __restrict__ ce, const int32_t* __restrict__ kk,
                                                    const int32_t* __restrict__ xadj,
                                                    const int32_t* __restrict__ adj, const int32_t* __restrict__ w,
                                                    const int32_t* __restrict__ vw,
                                                    const int32_t* __restrict__ part,
                                                    const int32_t* __restrict__ pw_off,
                                                    unsigned long long* __restrict__ cut,
                                                    unsigned long long* __restrict__ pw64) {
  extern __shared__ unsigned int hist[];
  __shared__ long long wsum[4];
  const int32_t c = static_cast<int32_t>(xcd_block());
  const int32_t j = cj[c], k = kk[j];
  for (int32_t q = threadIdx.x; q < k; q += 256) hist[q] = 0;
  __syncthreads();
  long long cc = 0;
  for (int32_t v = cs[c] + threadIdx.x; v < ce[c]; v += 256) {
    const int32_t p = part[v];
    atomicAdd(&hist[p], (unsigned int)vw[v]);
    for (int32_t e = xadj[v]; e < xadj[v + 1]; ++e) {
      const int32_t u = adj[e];
      if (v < u && part[u] != p) cc += w[e];
    }
  }
  cc = wave_sum64(cc);
  if (lane_id() == 0) wsum[threadIdx.x >> 6] = cc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const long long s = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (s) atomicAdd(&cut[j], (unsigned long long)s);
  }
  for (int32_t q = threadIdx.x; q < k; q += 256)
    if (hist[q]) atomicAdd(&pw64[pw_off[j] + q], (unsigned long long)hist[q]);
}

int launch_cut_chunks(int32_t nchunks, const int32_t* cj, const int32_t* cs, const int32_t* ce, int32_t kmax,
                      const int32_t* kk, const DGraph& g, const int32_t* part, const int32_t* pw_off,
                      unsigned long long* cut, unsigned long long* pw64, hipStream_t st) {
  if (nchunks <= 0) return 0;
  k_cut_chunks<<<nchunks, 256, sizeof(unsigned int) * kmax, st>>>(cj, cs, ce, kk, g.xadj, g.adj, g.w, g.vw, part,
                                                                  pw_off, cut, pw64);
  return hip_status();
}

}  // namespace gkd
