// Kernels of the batched form: many graphs (jobs) packed into one CSR with
// global vertex ids, processed by one launch per stage.  Jobs never share an
// edge, so the coarsening kernels of coarsen.hip run on the pack unchanged;
// this file adds the per-job glue:
//   pack / unpack    descriptor-driven copies between per-job storage (local
//                    ids) and the pack (global ids), one launch each;
//   job offsets      first coarse id / first coarse entry of every job;
//   select / keys    independent-set selection with per-job caps and part
//                    weights, then keys ordered by (job, gain desc, id asc);
//   segments         per-job ranges of the sorted keys;
//   balance          one workgroup per job (balance.hpp);
//   apply, cut       flat over the pack with per-job offsets.
// Every per-job result equals the single-graph path's: ids keep their order
// inside a job and every order-sensitive step is segmented by job.
#include <algorithm>

#include "balance.hpp"
#include "batch.hpp"
#include "prim.hpp"
#include "../../include/gkway.h"

namespace gkd {

static inline int hip_status() { return hipGetLastError() == hipSuccess ? 0 : GK_ERR_HIP; }

// Index j with off[j] <= i < off[j+1] (off is nondecreasing, off[0] = 0).
__device__ __forceinline__ int32_t job_of(const int32_t* __restrict__ off, int32_t nj, int64_t i) {
  int32_t lo = 0, hi = nj;  // answer in [lo, hi)
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------- pack
__global__ void __launch_bounds__(256) k_pack_vert(int32_t nj, const PackSrc* __restrict__ src,
                                                   const int32_t* __restrict__ v_off,
                                                   const int32_t* __restrict__ e_off,
                                                   const int32_t* __restrict__ c_off, int32_t ntot,
                                                   int32_t* __restrict__ xadj, int32_t* __restrict__ vw,
                                                   int32_t* __restrict__ vjob, int32_t* __restrict__ cmap) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= ntot) return;
  const int32_t j = job_of(v_off, nj, i);
  const PackSrc s = src[j];
  const int32_t li = int32_t(i - v_off[j]);
  xadj[i] = e_off[j] + (s.xadj[li] - s.xbase);
  vw[i] = s.vw[li];
  if (vjob) vjob[i] = j;
  if (cmap) cmap[i] = c_off[j] + (s.cmap[li] - s.cbase);
}

__global__ void __launch_bounds__(256) k_pack_edge(int32_t nj, const PackSrc* __restrict__ src,
                                                   const int32_t* __restrict__ v_off,
                                                   const int32_t* __restrict__ e_off, int32_t etot,
                                                   int32_t* __restrict__ adj, int32_t* __restrict__ w) {
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (e >= etot) return;
  const int32_t j = job_of(e_off, nj, e);
  const PackSrc s = src[j];
  const int32_t le = int32_t(e - e_off[j]);
  adj[e] = v_off[j] + (s.adj[le] - s.abase);
  w[e] = s.w[le];
}

// Coarse-side arrays of a refinement pack (partition and boundary flags of
// the next coarser level, one range per job).
__global__ void __launch_bounds__(256) k_pack_coarse(int32_t nj, const PackSrc* __restrict__ src,
                                                     const int32_t* __restrict__ c_off, int32_t ctot,
                                                     int32_t* __restrict__ cpart, int32_t* __restrict__ cbnd) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= ctot) return;
  const int32_t j = job_of(c_off, nj, i);
  const int32_t li = int32_t(i - c_off[j]);
  cpart[i] = src[j].cpart[li];
  cbnd[i] = src[j].cbnd[li];
}

int launch_pack(int32_t nj, const PackSrc* src, const int32_t* v_off, const int32_t* e_off, const int32_t* c_off,
                int32_t ntot, int32_t etot, int32_t ctot, int32_t* xadj, int32_t* adj, int32_t* w, int32_t* vw,
                int32_t* vjob, int32_t* cmap, int32_t* cpart, int32_t* cbnd, hipStream_t st) {
  if (ntot > 0)
    k_pack_vert<<<grid_for(ntot, 256), 256, 0, st>>>(nj, src, v_off, e_off, c_off, ntot, xadj, vw, vjob, cmap);
  fill<int32_t>(xadj + ntot, etot, 1, st);
  if (etot > 0) k_pack_edge<<<grid_for(etot, 256), 256, 0, st>>>(nj, src, v_off, e_off, etot, adj, w);
  if (cpart && ctot > 0)
    k_pack_coarse<<<grid_for(ctot, 256), 256, 0, st>>>(nj, src, c_off, ctot, cpart, cbnd);
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
// Coarse level of the pack -> per-job storage with local ids; jobs whose
// destination is null (level discarded) are skipped.
__global__ void __launch_bounds__(256) k_unpack_cvert(int32_t nj, const UnpackDst* __restrict__ dst,
                                                      const int32_t* __restrict__ c_off,
                                                      const int32_t* __restrict__ ce_off, int32_t nc,
                                                      const int32_t* __restrict__ cxadj,
                                                      const int32_t* __restrict__ cvw) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nc) return;
  const int32_t j = job_of(c_off, nj, i);
  const UnpackDst d = dst[j];
  if (!d.xadj) return;
  const int32_t li = int32_t(i - c_off[j]);
  d.xadj[li] = cxadj[i] - ce_off[j];
  d.vw[li] = cvw[i];
  if (i + 1 == c_off[j + 1]) d.xadj[li + 1] = cxadj[i + 1] - ce_off[j];
}

__global__ void __launch_bounds__(256) k_unpack_cedge(int32_t nj, const UnpackDst* __restrict__ dst,
                                                      const int32_t* __restrict__ c_off,
                                                      const int32_t* __restrict__ ce_off, int32_t ce,
                                                      const int32_t* __restrict__ cadj,
                                                      const int32_t* __restrict__ cw) {
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (e >= ce) return;
  const int32_t j = job_of(ce_off, nj, e);
  const UnpackDst d = dst[j];
  if (!d.adj) return;
  const int32_t le = int32_t(e - ce_off[j]);
  d.adj[le] = cadj[e] - c_off[j];
  d.w[le] = cw[e];
}

__global__ void __launch_bounds__(256) k_unpack_cmap(int32_t nj, const UnpackDst* __restrict__ dst,
                                                     const int32_t* __restrict__ v_off,
                                                     const int32_t* __restrict__ c_off, int32_t n,
                                                     const int32_t* __restrict__ cmap) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int32_t j = job_of(v_off, nj, i);
  const UnpackDst d = dst[j];
  if (!d.cmap) return;
  d.cmap[i - v_off[j]] = cmap[i] - c_off[j];
}

int launch_unpack_coarse(int32_t nj, const UnpackDst* dst, const int32_t* v_off, const int32_t* c_off,
                         const int32_t* ce_off, int32_t n, int32_t nc, int32_t ce, const int32_t* cmap,
                         const int32_t* cxadj, const int32_t* cadj, const int32_t* cw, const int32_t* cvw,
                         hipStream_t st) {
  if (nc > 0) k_unpack_cvert<<<grid_for(nc, 256), 256, 0, st>>>(nj, dst, c_off, ce_off, nc, cxadj, cvw);
  if (ce > 0) k_unpack_cedge<<<grid_for(ce, 256), 256, 0, st>>>(nj, dst, c_off, ce_off, ce, cadj, cw);
  if (n > 0) k_unpack_cmap<<<grid_for(n, 256), 256, 0, st>>>(nj, dst, v_off, c_off, n, cmap);
  return hip_status();
}

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

// ---------------------------------------------------------------- selection
// create_independent_move_buffer (gkway/uncoarsen.hpp:111-173) on the pack:
// the cap and the part weights are those of the vertex's job.  Appends the
// selected vertex ids; ctl[0] = count, ctl[1] = max gain.
__global__ void __launch_bounds__(256) k_select_b(int32_t n, const int32_t* __restrict__ xadj,
                                                  const int32_t* __restrict__ adj,
                                                  const int32_t* __restrict__ vw,
                                                  const int32_t* __restrict__ vjob,
                                                  const int32_t* __restrict__ bnd,
                                                  const int32_t* __restrict__ gain,
                                                  const int32_t* __restrict__ dest,
                                                  const int32_t* __restrict__ pw,
                                                  const int32_t* __restrict__ pw_off,
                                                  const long long* __restrict__ cap,
                                                  int32_t* __restrict__ list, int32_t* __restrict__ ctl,
                                                  int32_t* __restrict__ reset) {
  const int64_t v = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (reset && v == 0) { reset[0] = 0; reset[1] = 0; }
  bool sel = false;
  int32_t g = 0;
  if (v < n) {
    g = gain[v];
    if (bnd[v] && g > 0) {
      const int32_t j = vjob[v];
      const int32_t* jpw = pw + pw_off[j];
      const long long jc = cap[j];
      if ((long long)jpw[dest[v]] + vw[v] <= jc) {
        sel = true;
        for (int32_t e = xadj[v]; e < xadj[v + 1]; ++e) {
          const int32_t u = adj[e];
          if (u < v && gain[u] > 0 && (long long)jpw[dest[u]] + vw[u] <= jc) { sel = false; break; }
        }
      }
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
  if (sel) list[base + popc64(m & lanemask_lt())] = int32_t(v);
}

int launch_select_b(const DGraph& g, const int32_t* vjob, const int32_t* bnd, const int32_t* gain,
                    const int32_t* dest, const int32_t* pw, const int32_t* pw_off, const long long* cap,
                    int32_t* list, int32_t* ctl, int32_t* reset, hipStream_t st) {
  fill<int32_t>(ctl, 0, 2, st);
  if (g.n <= 0) return 0;
  k_select_b<<<grid_for(g.n, 256), 256, 0, st>>>(g.n, g.xadj, g.adj, g.vw, vjob, bnd, gain, dest, pw, pw_off, cap,
                                                 list, ctl, reset);
  return hip_status();
}

// key = job << (gbits + idbits) | (gmax - gain) << idbits | v: ascending order
// is (job, gain desc, id asc) -- the mvRequest order (gkway/move_request.hpp:
// 36-41) inside each job.
__global__ void __launch_bounds__(256) k_keys_b(int32_t count, const int32_t* __restrict__ list,
                                                const int32_t* __restrict__ gain,
                                                const int32_t* __restrict__ vjob, int32_t gmax, int gbits,
                                                int idbits, uint64_t* __restrict__ keys) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= count) return;
  const int32_t v = list[i];
  keys[i] = (uint64_t(uint32_t(vjob[v])) << (gbits + idbits)) | (uint64_t(uint32_t(gmax - gain[v])) << idbits) |
            uint64_t(uint32_t(v));
}

int launch_keys_b(int32_t count, const int32_t* list, const int32_t* gain, const int32_t* vjob, int32_t gmax,
                  int gbits, int idbits, uint64_t* keys, hipStream_t st) {
  if (count <= 0) return 0;
  k_keys_b<<<grid_for(count, 256), 256, 0, st>>>(count, list, gain, vjob, gmax, gbits, idbits, keys);
  return hip_status();
}

__global__ void __launch_bounds__(256) k_segments(int32_t count, const uint64_t* __restrict__ keys, int shift,
                                                  int32_t* __restrict__ seg_beg, int32_t* __restrict__ seg_end) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= count) return;
  const uint64_t j = keys[i] >> shift;
  if (i == 0 || (keys[i - 1] >> shift) != j) seg_beg[j] = int32_t(i);
  if (i + 1 == count || (keys[i + 1] >> shift) != j) seg_end[j] = int32_t(i + 1);
}

int launch_segments(int32_t nj, int32_t count, const uint64_t* keys, int shift, int32_t* seg_beg, int32_t* seg_end,
                    hipStream_t st) {
  fill<int32_t>(seg_beg, 0, nj, st);
  fill<int32_t>(seg_end, 0, nj, st);
  if (count > 0) k_segments<<<grid_for(count, 256), 256, 0, st>>>(count, keys, shift, seg_beg, seg_end);
  return hip_status();
}

// One workgroup per job: its first min(count_j, M) moves (presorted).  Per-job
// round and move counters advance for jobs that had moves.
__global__ void __launch_bounds__(kBalThreads) k_balance_b(const uint64_t* __restrict__ keys,
                                                           const int32_t* __restrict__ seg_beg,
                                                           const int32_t* __restrict__ seg_end, int32_t M,
                                                           uint64_t idmask, const int32_t* __restrict__ kk,
                                                           const long long* __restrict__ cap,
                                                           const int32_t* __restrict__ pw_off,
                                                           const int32_t* __restrict__ part,
                                                           const int32_t* __restrict__ dest,
                                                           const int32_t* __restrict__ gain,
                                                           const int32_t* __restrict__ vw, int32_t* __restrict__ pw,
                                                           int32_t* __restrict__ pos, long long* __restrict__ cut,
                                                           int32_t* __restrict__ rounds, long long* __restrict__ moved) {
  __shared__ BalanceLds L;
  extern __shared__ int32_t dpw[];
  const int32_t j = blockIdx.x;
  const int32_t b = seg_beg[j], cnt = seg_end[j] - b;
  if (cnt <= 0) {
    if (threadIdx.x == 0) pos[j] = -1;
    return;
  }
  balance_block(L, dpw, kk[j], keys + b, cnt, M, 1, idmask, part, dest, gain, vw, pw + pw_off[j], cap[j], nullptr,
                &pos[j], &cut[j]);
  if (threadIdx.x == 0) {
    rounds[j] += 1;
    moved[j] += pos[j] + 1;
  }
}

int launch_balance_b(int32_t nj, int32_t kmax, const uint64_t* keys, const int32_t* seg_beg, const int32_t* seg_end,
                     int32_t M, int idbits, const int32_t* kk, const long long* cap, const int32_t* pw_off,
                     const int32_t* part, const int32_t* dest, const int32_t* gain, const int32_t* vw, int32_t* pw,
                     int32_t* pos, long long* cut, int32_t* rounds, long long* moved, hipStream_t st) {
  if (nj <= 0) return 0;
  if (M > kBalMax) return GK_ERR_LIMIT;
  const uint64_t idmask = (1ull << idbits) - 1;
  k_balance_b<<<nj, kBalThreads, sizeof(int32_t) * kmax, st>>>(keys, seg_beg, seg_end, M, idmask, kk, cap, pw_off,
                                                               part, dest, gain, vw, pw, pos, cut, rounds, moved);
  return hip_status();
}

// apply_sequence_move (gkway/uncoarsen.hpp:290-323) over the sorted keys:
// a key at rank r of its job's segment is applied when r <= pos[job].
__global__ void __launch_bounds__(256) k_apply_b(int32_t count, const uint64_t* __restrict__ keys,
                                                 uint64_t idmask, int shift,
                                                 const int32_t* __restrict__ seg_beg,
                                                 const int32_t* __restrict__ pos,
                                                 const int32_t* __restrict__ xadj,
                                                 const int32_t* __restrict__ adj,
                                                 const int32_t* __restrict__ dest, int32_t* __restrict__ part,
                                                 int32_t* __restrict__ upd, int32_t* __restrict__ wl,
                                                 int32_t* __restrict__ wlc) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= count) return;
  const uint64_t key = keys[i];
  const int32_t j = int32_t(key >> shift);
  if (i - seg_beg[j] > pos[j]) return;
  const int32_t v = int32_t(key & idmask);
  part[v] = dest[v];
  if (atomicExch(&upd[v], 1) == 0) wl[atomicAdd(&wlc[0], 1)] = v;
  for (int32_t e = xadj[v]; e < xadj[v + 1]; ++e) {
    const int32_t u = adj[e];
    if (atomicExch(&upd[u], 1) == 0) wl[atomicAdd(&wlc[0], 1)] = u;
  }
}

int launch_apply_b(const DGraph& g, int32_t count, const uint64_t* keys, int idbits, int shift,
                   const int32_t* seg_beg, const int32_t* pos, const int32_t* dest, int32_t* part, int32_t* upd,
                   int32_t* wl, int32_t* wlc, hipStream_t st) {
  if (count <= 0) return 0;
  k_apply_b<<<grid_for(count, 256), 256, 0, st>>>(count, keys, (1ull << idbits) - 1, shift, seg_beg, pos, g.xadj, g.adj,
                                                  dest, part, upd, wl, wlc);
  return hip_status();
}

// Final cut and part weights per job (calculate_cutsize +
// calculate_partition_wgt), 8 lanes per vertex, only for jobs with done[j].
__global__ void __launch_bounds__(256) k_cut_b(int32_t n, const int32_t* __restrict__ xadj,
                                               const int32_t* __restrict__ adj, const int32_t* __restrict__ w,
                                               const int32_t* __restrict__ vw, const int32_t* __restrict__ vjob,
                                               const int32_t* __restrict__ part,
                                               const int32_t* __restrict__ pw_off,
                                               const int32_t* __restrict__ done,
                                               unsigned long long* __restrict__ cut,
                                               unsigned long long* __restrict__ pw64) {
  const int64_t gt = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int64_t v = gt / 8;
  const int sub = static_cast<int>(gt % 8);
  if (v >= n) return;
  const int32_t j = vjob[v];
  if (!done[j]) return;
  const int32_t p = part[v];
  long long c = 0;
  for (int32_t e = xadj[v] + sub; e < xadj[v + 1]; e += 8) {
    const int32_t u = adj[e];
    if (v < u && part[u] != p) c += w[e];
  }
  if (c) atomicAdd(&cut[j], (unsigned long long)c);
  if (sub == 0) atomicAdd(&pw64[pw_off[j] + p], (unsigned long long)vw[v]);
}

int launch_cut_b(const DGraph& g, const int32_t* vjob, const int32_t* part, const int32_t* pw_off,
                 const int32_t* done, unsigned long long* cut, unsigned long long* pw64, hipStream_t st) {
  if (g.n <= 0) return 0;
  k_cut_b<<<grid_for(int64_t(g.n) * 8, 256), 256, 0, st>>>(g.n, g.xadj, g.adj, g.w, g.vw, vjob, part, pw_off, done,
                                                           cut, pw64);
  return hip_status();
}

}  // namespace gkd
