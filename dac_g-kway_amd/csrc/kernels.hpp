// Launch wrappers of the gfx950 kernels (coarsen.hip, refine.hip), called by
// the host driver.  All take device pointers and a stream; all return a gk
// status code (0 = launched).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <cstdint>

#include "steal.hpp"

namespace gkd {

// Kernel-level stage timing.  While the host times a stage kernel by kernel
// (the refinement round and the coarsening stages that can dominate a run),
// kernel_events hands out a start/stop event pair for each launch made
// through gk_launch, which records them at the kernel's own start and end
// (hipExtLaunchKernelGGL), so the stage's time is the sum of its kernels'
// durations as a kernel trace reports them, without event packets between
// the launches.  Otherwise gk_launch is a plain launch.  Defined in
// driver.hip.
bool kernel_events(hipEvent_t* start, hipEvent_t* stop);

template <typename... KArgs, typename... Args>
inline void gk_launch(void (*k)(KArgs...), dim3 grid, dim3 block, uint32_t shm, hipStream_t st, Args... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (kernel_events(&e0, &e1))
    hipExtLaunchKernelGGL(k, grid, block, shm, st, e0, e1, 0u, args...);
  else
    hipLaunchKernelGGL(k, grid, block, shm, st, args...);
}

struct DGraph {
  int32_t n = 0;
  int32_t nnz = 0;
  const int32_t* xadj = nullptr;
  const int32_t* adj = nullptr;
  const int32_t* w = nullptr;
  const int32_t* vw = nullptr;
};

// Lanes per vertex for row-parallel kernels, from the average degree.
int lanes_for_degree(double avg_deg);

// ---- coarsening (coarsen.hip)
// lanes per vertex (1..64) and block size (128, 256 or 512 threads) change
// only the launch shape, never the result.
// ws (may be null): run as a persistent work-stealing launch (same result).
int launch_match(const DGraph& g, int lanes, int block, int32_t* cand, hipStream_t st,
                 const Steal* ws = nullptr);
// Pointer jumping over the candidate pseudo-forest.
int launch_jump_init(int32_t n, const int32_t* cand, int32_t* next, int32_t* dist, hipStream_t st);
// One Jacobi round; sets flags[round] = 1 when any pointer moved.  Skips all
// work when round > 0 and flags[round - 1] == 0 (already converged).
int launch_jump_round(int32_t n, const int32_t* cand, const int32_t* next_in, const int32_t* dist_in,
                      int32_t* next_out, int32_t* dist_out, int32_t* flags, int32_t round,
                      hipStream_t st);
// label[v] = largest id of v's component (root pair found via next[]).
int launch_label(int32_t n, const int32_t* cand, const int32_t* next, int32_t* compmax,
                 int32_t* label, hipStream_t st);
// keys[v] = (label << sbits) | dist, ids[v] = v.
int launch_make_keys(int32_t n, const int32_t* label, const int32_t* dist, int sbits,
                     uint64_t* keys, int32_t* ids, hipStream_t st);
// Chunking of the sorted order: cmap, chunk pointers, coarse weights; nc_dev.
int launch_cluster(int32_t n, const uint64_t* skeys, const int32_t* sids, int sbits,
                   const int32_t* vwgt, int32_t G, int32_t* gstart, int32_t* flagscan,
                   int32_t* cmap, int32_t* cptr, int32_t* cvwgt, int32_t* nc_dev,
                   int32_t* scratch, hipStream_t st);
// Contraction, phase 1: per coarse row the upper bound of its entries
// (sum of member degrees), its class (16/32/64-lane or big), appended to
// lists[cls * nc ..] with its member range in desc; mrow[m] = (row start,
// row length) of the member at order position m; the exclusive scan of the
// bounds (toff, nc+1 entries).  lists and desc: 4 * nc entries, mrow: n.
int launch_contract_bounds(const DGraph& g, int32_t nc, const int32_t* cptr, const int32_t* order,
                           int32_t* ub, int32_t* toff, int32_t* lists, int2* desc, int2* mrow,
                           int32_t* list_cnt, int32_t* scratch, hipStream_t st);
// Phase 2 for the rows of one class (lanes in {16, 32, 64}); block size 64,
// 128, 256 or 512 threads (launch shape only).  nc: coarse vertex count; the
// sort packs the lane number beside the coarse id in 32 bits when nc allows,
// else (or with force_wide) sorts (id, weight) pairs.  ws (may be null):
// the 32-bit-key rows run as a persistent work-stealing launch.
int launch_contract_rows(const DGraph& g, int lanes, int block, const int32_t* rows, const int2* desc,
                         int32_t nrows, int32_t nc, const int2* mrow, const int32_t* cmap, const int32_t* toff,
                         int32_t* tadj, int32_t* tw, int32_t* len, bool force_wide, hipStream_t st,
                         const Steal* ws = nullptr);
// Phase 2 for big rows: gather to (row-rank, neighbour) keys.
int launch_contract_big_gather(const DGraph& g, const int32_t* rows, int32_t nrows,
                               const int32_t* boff, const int32_t* cptr, const int32_t* order,
                               const int32_t* cmap, uint64_t* keys, int32_t* vals, hipStream_t st);
// Segmented merge of the sorted big-row entries into the temporary rows.
int launch_contract_big_merge(const int32_t* rows, int32_t nrows, const int32_t* boff,
                              const uint64_t* keys, const int32_t* vals, const int32_t* toff,
                              int32_t* tadj, int32_t* tw, int32_t* len, hipStream_t st);
// Phase 3: copy the merged rows to their final CSR slots.
int launch_contract_copy(int32_t nc, const int32_t* toff, const int32_t* len, const int32_t* cxadj,
                         const int32_t* tadj, const int32_t* tw, int32_t* cadj, int32_t* cw,
                         hipStream_t st);
// Big-row prefix (boff = exclusive scan of ub over the big rows).
int launch_big_offsets(const int32_t* rows, int32_t nrows, const int32_t* ub, int32_t* boff,
                       int32_t* total, int32_t* scratch, hipStream_t st);
// cxadj = exclusive scan of len (nc + 1 entries), total into *cnnz.
int launch_len_scan(int32_t nc, const int32_t* len, int32_t* cxadj, int32_t* cnnz,
                    int32_t* scratch, hipStream_t st);

// ---- refinement (refine.hip)
int launch_project(int32_t n, const int32_t* cmap, const int32_t* cpart, const int32_t* cbnd,
                   int32_t* part, int32_t* bnd, int32_t* gain, int32_t* dest, int32_t* upd,
                   hipStream_t st);
// Worklist of the v with flags[v] != 0 appended at list[*cnt...].
int launch_compact(int32_t n, const int32_t* flags, int32_t* list, int32_t* cnt, hipStream_t st);
// Gain pass over the worklist wl[0..wlc[0]); rows longer than `lanes` go
// through heavy[] (count wlc[1]) to the LDS-histogram kernel.  Clears upd of
// every processed vertex.  work (may be null) accumulates [0] vertices
// recomputed, [1] their entries.
// Part counts up to which the gain pass takes rows longer than its lane
// count itself (per-wave LDS histograms); above it they go to the heavy list
// and launch_gain_heavy.
constexpr int32_t kGainInlineK = 256;
inline bool gain_heavy_separate(int32_t k) { return k > kGainInlineK; }
// Workgroups of a gain launch at most: work holds 2 * kGainMaxBlocks slots.
constexpr unsigned kGainMaxBlocks = 2048;
// ws (may be null): run as a persistent work-stealing launch (same result).
int launch_gain_wl(const DGraph& g, int32_t k, int lanes, const int32_t* wl, int32_t* wlc, int32_t* heavy,
                   const int32_t* part, int32_t* upd, int32_t* bnd, int32_t* gain, int32_t* dest,
                   unsigned long long* work, int64_t max_items, uint64_t* cand, hipStream_t st,
                   const Steal* ws = nullptr);
// Rows collected in heavy[0..wlc[1]) by launch_gain_wl: one wave each, LDS
// histogram of the k part connectivities.  Both gain kernels keep the optional
// bitmask cand (bit v = boundary vertex with positive gain) in step.
int launch_gain_heavy(const DGraph& g, int32_t k, const int32_t* heavy, const int32_t* wlc, const int32_t* part,
                      int32_t* upd, int32_t* bnd, int32_t* gain, int32_t* dest, unsigned long long* work,
                      uint64_t* cand, hipStream_t st);
// Selection: appends keys ((~gain & 0x7fffffff) << idbits | v) of the
// selected vertices; ctl[0] = count, ctl[1] = max gain.  clear_upd zeroes
// upd[] (the reference's behaviour); reset (may be null) gets reset[0] =
// reset[1] = 0 (worklist counters).  work (may be null) accumulates
// [2] candidates, [3] rows walked, [4] entries scanned.
int launch_select(const DGraph& g, const int32_t* bnd, const int32_t* gain, const int32_t* dest,
                  const int32_t* pw, int64_t cap, int idbits, int32_t* upd, uint64_t* keys,
                  int32_t* ctl, unsigned long long* work, int clear_upd, int32_t* reset, hipStream_t st);
// Balance: first min(count, M) keys (sorted when presorted, else sorted in
// LDS); writes moves[0..pos], ctl[2] = pos; advances pw; cut -= gains.
int launch_balance(const DGraph& g, int32_t k, const uint64_t* keys, int32_t count, int32_t M,
                   bool presorted, int idbits, const int32_t* part, const int32_t* dest,
                   const int32_t* gain, int32_t* pw, int64_t cap, int32_t* moves, int32_t* ctl,
                   long long* cut, hipStream_t st);
// Applies moves[0..ctl[2]]; movers and neighbours whose upd flag goes 0 -> 1
// are appended to wl (count wlc[0]).
int launch_apply(const DGraph& g, int32_t M, const int32_t* moves, const int32_t* ctl,
                 const int32_t* dest, int32_t* part, int32_t* upd, int32_t* wl, int32_t* wlc,
                 hipStream_t st);
int launch_boundary(const DGraph& g, const int32_t* part, int32_t* bnd, hipStream_t st);
// cut and part weights pw64[k] (both overwritten); scratch holds
// cut_scratch(n, k) values, or fewer when max_slots caps it (fewer blocks;
// the sums do not depend on the block count).
int64_t cut_scratch(int64_t n, int32_t k);
int launch_cut(const DGraph& g, int32_t k, const int32_t* part, unsigned long long* cut,
               unsigned long long* pw64, unsigned long long* scratch, hipStream_t st,
               int64_t max_slots = INT64_MAX);

int launch_sum(const int32_t* a, int64_t n, unsigned long long* out, hipStream_t st);
int launch_cvwgt(int32_t nc, const int32_t* cptr, const int32_t* order, const int32_t* vw,
                 int32_t* cvw, hipStream_t st);

}  // namespace gkd
