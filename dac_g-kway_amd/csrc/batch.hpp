// Launch wrappers of the batched-form kernels (batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels.hpp"

namespace gkd {

// One job's source arrays for a pack.  Vertex-side arrays are indexed by the
// job-local vertex id, entry-side arrays by the job-local entry index; xbase /
// abase are subtracted from xadj / adjncy values (0 for job-owned levels, the
// job's offsets when the source is a slice of a caller's packed CSR).
struct PackSrc {
  const int32_t* xadj;
  const int32_t* adj;
  const int32_t* w;
  const int32_t* vw;
  const int32_t* cmap;   // refinement packs: fine -> coarse (minus cbase = local ids); null = identity
  const int32_t* cpart;  // refinement packs: coarse partition
  const int32_t* cbnd;   // refinement packs: coarse boundary flags
  const int32_t* cgain;  // resumed levels: gain and destination carried over
  const int32_t* cdest;  //   (null for a level entered by projection)
  int32_t xbase;
  int32_t abase;
  int32_t cbase;
};

// One job's destination arrays for an unpack (null members are skipped).
struct UnpackDst {
  int32_t* part;
  int32_t* bnd;
};

// Pack nj jobs: vertices at v_off, entries at e_off (nj+1 prefix arrays);
// writes xadj[ntot+1], adj/w[etot], vw/vjob[ntot]; with cmap != null also the
// packed fine->coarse map (+c_off) and, with cpart != null, the packed coarse
// partition/boundary arrays (ctot entries).
// With cgain != null also cgain/cdest: the carried values of resumed jobs, and
// dest = -1 for the others.
int launch_pack(int32_t nj, const PackSrc* src, const int32_t* v_off, const int32_t* e_off, const int32_t* c_off,
                int32_t ntot, int32_t etot, int32_t ctot, int32_t* xadj, int32_t* adj, int32_t* w, int32_t* vw,
                int32_t* vjob, int32_t* cmap, int32_t* cpart, int32_t* cbnd, int32_t* cgain, int32_t* cdest,
                hipStream_t st);
// Projection of a refinement pack: expand_partition for levels entered from
// the coarser level (gain 0, dest = part, update = boundary); resumed levels
// (cdest >= 0) take their carried gain and dest and need no update.
int launch_project_b(int32_t n, const int32_t* cmap, const int32_t* cpart, const int32_t* cbnd, const int32_t* cgain,
                     const int32_t* cdest, int32_t* part, int32_t* bnd, int32_t* gain, int32_t* dest, int32_t* upd,
                     hipStream_t st);
// cand bit v = (bnd[v] && gain[v] > 0), all (n + 63) / 64 words written.
int launch_cand_init(int32_t n, const int32_t* bnd, const int32_t* gain, uint64_t* cand, hipStream_t st);
// out[j] = total vertex weight of job j (v_off on the device, nj+1).
int launch_job_sum(int32_t nj, const int32_t* v_off, const int32_t* vw, long long* out, hipStream_t st);
// c_off[j] = first coarse id of job j (c_off[nj] = nc) from the clustering.
int launch_job_offsets(int32_t nj, const int32_t* v_off, const int32_t* order, const int32_t* cmap, int32_t nc,
                       int32_t* c_off, hipStream_t st);
// ce_off[j] = cxadj[c_off[j]].
int launch_job_eoffsets(int32_t nj, const int32_t* c_off, const int32_t* cxadj, int32_t* ce_off, hipStream_t st);
int launch_unpack_part(int32_t nj, const UnpackDst* dst, const int32_t* v_off, int32_t n, const int32_t* part,
                       const int32_t* bnd, hipStream_t st);
// Buffers of the round selection (n vertices, nwords = (n + 63) / 64):
// movable, selbits, defbits and wpos hold nwords + 1 entries, the
// last word of selbits zero (filled once); list holds n ints; scratch holds
// select_scratch_ints(n) ints (per-block selection counts).
struct SelectBufs {
  uint64_t* movable;
  uint64_t* selbits;
  uint64_t* defbits;
  int32_t* wpos;
  int32_t* list;
  int* scratch;
  // May be null: kSelectByteSlots slots to which the selection kernels add
  // their algorithmic bytes (block b into slot b % kSelectByteSlots).
  unsigned long long* abytes;
};
constexpr int kSelectByteSlots = 2048;
int select_scratch_ints(int64_t n);

// Refinement round without host round trips.  launch_select_list: the
// movable set (candidates whose move fits the job's cap), the selection
// (no smaller-id movable neighbour) as a bitmask, and its ordered list; a
// job's movable vertices with gain below thr[job] (1 before a level's first
// round) are only marked in defbits.  launch_round_b: one workgroup per job
// (best min(count, M) moves, LDS sort, balanced prefix, apply), which tests
// the deferred vertices when fewer than M were selected (into the job's
// segment of xlist, n ints) and sets the job's next threshold.  *active
// receives the number of selected moves of the round (0 once every job has
// converged) and lastcnt[j] job j's count (lastcnt nonzero before the first
// round: a job's candidate bits are cleared in the first round that selects
// nothing for it, after which it cannot change).  vjob == null: one job.
// paths (may be null): per job three counters, advanced per round: rounds
// that ran the top-M radix select, rounds whose compressed keys exceeded 32
// bits, rounds that tested deferred vertices.  order (may be null): per job
// M + 2 ints, [0] = B, [1] = pos and [2, 2 + B) the ordered moves of the
// job's latest round that selected any.  abytes (may be null): per job, the
// algorithmic bytes of its rounds are added.
int launch_select_list(const DGraph& g, const int32_t* vjob, const int32_t* gain, const int32_t* dest,
                       const int32_t* pw, const int32_t* pw_off, const long long* cap, const int32_t* thr,
                       const uint64_t* cand, const SelectBufs& sb, int32_t* reset, int32_t* active,
                       hipStream_t st);
int launch_round_b(int32_t nj, int32_t kmax, const int32_t* v_off, int32_t M, const int32_t* kk, const long long* cap,
                   const int32_t* pw_off, const DGraph& g, int32_t* part, const int32_t* dest, const int32_t* gain,
                   int32_t* pw, long long* cut, int32_t* rounds, long long* moved, const SelectBufs& sb,
                   int32_t* xlist, uint32_t* hkey, int32_t* upd, int32_t* wl, int32_t* wlc, int32_t* lastcnt,
                   int32_t* active, uint64_t* cand, int32_t* thr, long long* paths, int32_t* order,
                   long long* abytes, hipStream_t st);
// *out = largest row length of g.
int launch_max_deg(const DGraph& g, int32_t* out, hipStream_t st);
// Final cut and part weights of the jobs covered by the chunk table: chunk c
// is vertices [cs[c], ce[c]) of job cj[c] (pack ids); cut[job] and
// pw64[pw_off[job] + q] accumulate (zeroed by the caller).
int launch_cut_chunks(int32_t nchunks, const int32_t* cj, const int32_t* cs, const int32_t* ce, int32_t kmax,
                      const int32_t* kk, const DGraph& g, const int32_t* part, const int32_t* pw_off,
                      unsigned long long* cut, unsigned long long* pw64, hipStream_t st);

}  // namespace gkd
