/* gkway.h — C ABI of the MI355X (gfx950) multilevel k-way graph partitioner.
 *
 * One shared library, libgkway.so, exports everything below.  Signatures use
 * plain pointers, sizes and opaque handles only.  Device pointers are HIP
 * device addresses; `stream` arguments are hipStream_t passed as void*.
 *
 * What each entry point replaces in the reference (wanluanlee/DAC_g-kway):
 *   gk_partition_file      graph_partitioner()          gkway/graph_partitioner.hpp:222-365
 *   gk_partition           the V-cycle inside it         gkway/graph_partitioner.hpp:297-335
 *   gk_graph_read          gk::Graph::_parse             gkway/graph.hpp:70-152
 *   gk_write_partition     gk::Graph::dump_result        gkway/graph.hpp:227-240
 *   gk_initpart            metis_init_partition          gkway/metis_partition.hpp:4-99
 *   gk_k_match             matching_with_score           gkway/coarsen.hpp:39-78
 *   gk_k_group             init_group_id + HEC_union_find loop
 *                                                        gkway/coarsen.hpp:351-359,80-118,461-469
 *   gk_k_cluster           constraint_group_size + coarse numbering + cal_coarsen_vwgt
 *                                                        gkway/coarsen.hpp:361-421,120-171,472-483
 *   gk_k_contract          contraction                   gkway/coarsen.hpp:302-349
 *   gk_k_project           expand_partition              gkway/uncoarsen.hpp:325-356
 *   gk_k_gain              update_vertex_info            gkway/uncoarsen.hpp:175-237
 *   gk_k_select            create_independent_move_buffer + mergesort
 *                                                        gkway/uncoarsen.hpp:111-173,471-476
 *   gk_k_balance           find_mv_wgt_sequence + scan + find_balance_sequence
 *                                                        gkway/uncoarsen.hpp:239-288,37-80
 *   gk_k_apply             apply_sequence_move           gkway/uncoarsen.hpp:290-323
 *   gk_k_round             one pass of the refinement loop body: the above
 *                          four plus update_vertex_info  gkway/uncoarsen.hpp:456-506
 *   gk_k_cut               calculate_cutsize + calculate_partition_wgt
 *                                                        gkway/uncoarsen.hpp:398-424,
 *                                                        gkway/cuda_check.hpp:61-73
 * The reference has no C ABI at all: these stage functions are C++ templates
 * taking device pointers and a moderngpu context.  Errors are status codes
 * (never exceptions across the ABI); gk_strerror() names them and
 * gk_last_error() returns the last detailed message of the calling thread.
 */
#ifndef GKWAY_H
#define GKWAY_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GK_OK 0
#define GK_ERR_ARG -1        /* invalid argument (null pointer, k < 1, ...) */
#define GK_ERR_IO -2         /* file could not be opened / written          */
#define GK_ERR_FORMAT -3     /* malformed METIS file                        */
#define GK_ERR_GRAPH -4      /* graph violates an input invariant           */
#define GK_ERR_HIP -5        /* a HIP runtime call failed                   */
#define GK_ERR_NOMEM -6      /* device or host allocation failed            */
#define GK_ERR_LIMIT -7      /* size beyond a documented limit              */
#define GK_ERR_INTERNAL -8   /* internal consistency check failed           */
#define GK_ERR_NODEVICE -9   /* no HIP device / device code not loaded      */

/* ---- graph description --------------------------------------------------
 * Symmetric CSR: for every entry (v, u, w) there is an entry (u, v, w).
 * adjncy is 0-based.  adjwgt / vwgt may be NULL (unit weights).  Limits:
 * nnz < 2^31, total vertex weight < 2^31, total undirected edge weight < 2^31,
 * weights >= 1, no self loops.                                                */
typedef struct gk_csr {
  int32_t n;               /* vertices                                  */
  int32_t nnz;             /* directed adjacency entries (2 x edges)    */
  const int32_t* xadj;     /* n + 1                                     */
  const int32_t* adjncy;   /* nnz                                       */
  const int32_t* adjwgt;   /* nnz or NULL                               */
  const int32_t* vwgt;     /* n or NULL                                 */
} gk_csr;

/* Coarsening stop policy. */
#define GK_COARSEN_TARGET 0  /* keep a level only while nc >= max(min(n0/(20*ceil(log2 k)), 240k), 30k) */
#define GK_COARSEN_REF 1     /* the reference's rule: coarsen while n > 20k
                                (gkway/graph_partitioner.hpp:22,66,93)   */

typedef struct gk_params {
  int32_t k;               /* number of parts (>= 1)                          */
  int32_t ubfactor;        /* allowed imbalance in 1/1000 (30 = 3 %)          */
  int32_t max_group;       /* coarse vertex size cap G (reference: 6)         */
  int32_t max_moves;       /* moves applied per refinement round (ref: 1024)  */
  int32_t coarsen_policy;  /* GK_COARSEN_TARGET or GK_COARSEN_REF             */
  int32_t coarsen_to;      /* > 0 overrides the TARGET policy's floor         */
  int32_t max_levels;      /* safety cap on hierarchy depth (0 = 64)          */
  int32_t seed;            /* seed of the initial partitioner                 */
  int32_t initpart_threads;/* host threads for the initial partitioner (0=auto)*/
  int32_t flags;           /* GK_FLAG_* bits                                  */
} gk_params;

#define GK_FLAG_CHECK 1      /* validate input symmetry on entry             */
#define GK_FLAG_TRACE 2      /* record per-level hashes in gk_stats.levels   */
#define GK_FLAG_STAGE_TIME 4 /* bracket every launch with HIP events on the
                                work stream; fill gk_stats.stage_*          */
#define GK_FLAG_GRAPH_ROUNDS 8 /* capture a level's refinement rounds (four
                                between two host checks) in a hipGraph once
                                and replay it; results are identical.  Not
                                used while a round stage is event-timed.    */
#define GK_FLAG_AUTOTUNE 16  /* choose the launch shape of matching (lanes
                                per vertex, block size) and contraction
                                (block size) by timing candidates on first
                                use of a degree profile; kept in the ctx.
                                Results are identical.                      */
#define GK_FLAG_PERSISTENT 32 /* run matching, the contraction rows and the
                                gain pass as persistent work-stealing
                                launches: a grid sized to the device whose
                                waves take tiles from per-XCD queues.
                                Results are identical.                      */

/* Stage ids of gk_stats.stage_ms / stage_launches / stage_bytes. */
#define GK_ST_MATCH 0      /* candidate matching                           */
#define GK_ST_GROUP 1      /* pointer jumping + component labels           */
#define GK_ST_SORT 2       /* radix sort by (label, stamp, id)             */
#define GK_ST_CLUSTER 3    /* chunking, coarse ids, coarse vertex weights  */
#define GK_ST_CONTRACT 4   /* per-row sort + merge (rows <= 64 entries)    */
#define GK_ST_CONTRACT_BIG 5 /* gather + sort + merge of rows > 64 entries */
#define GK_ST_PROJECT 6
#define GK_ST_GAIN 7
#define GK_ST_SELECT 8
#define GK_ST_MSORT 9      /* radix sort of > 1024 selected moves          */
#define GK_ST_BALANCE 10
#define GK_ST_APPLY 11
#define GK_ST_CUT 12       /* final cut and part weights                   */
#define GK_ST_GAIN_HEAVY 13 /* gain of rows longer than the lane count     */
#define GK_ST_CONTRACT_AUX 14 /* row bounds, offset scans, final copy      */
#define GK_ST_PACK 15      /* batched form: copies of levels into a pack   */
#define GK_ST_COUNT 16

/* Fill *p with the defaults: ubfactor 30, max_group 6, max_moves 1024,
 * TARGET policy, seed 1. k is left as given. */
void gk_params_default(gk_params* p, int32_t k);

#define GK_MAX_LEVELS 64

typedef struct gk_level_stats {
  int32_t n;               /* vertices of this level                          */
  int32_t nnz;             /* directed entries of this level                  */
  int32_t sweeps;          /* union-find sweeps run to coarsen this level     */
  int32_t rounds;          /* refinement rounds run at this level             */
  int64_t moves;           /* vertices moved at this level                    */
  int64_t cut_after;       /* tracked cut after refining this level           */
  uint64_t h_cand;         /* FNV-1a of cand[] (GK_FLAG_TRACE)                */
  uint64_t h_cmap;         /* FNV-1a of cmap[]                                */
  uint64_t h_coarse;       /* FNV-1a of the coarse CSR (xadj,adjncy,adjwgt,vwgt) */
  uint64_t h_part;         /* FNV-1a of the partition after refining          */
} gk_level_stats;

typedef struct gk_stats {
  int32_t nlevels;         /* levels kept, including level 0                   */
  int32_t coarsest_n;
  int64_t cut;             /* recomputed weighted edge cut of the result       */
  int64_t cut_tracked;     /* initial cut minus applied gains (== cut)         */
  int64_t init_cut;        /* cut of the initial partition at the coarsest    */
  int64_t max_part_wgt;
  int64_t total_vwgt;
  int64_t cap;             /* maximum allowed part weight                      */
  double imbalance;        /* max part weight / (total / k)                    */
  double t_coarsen_ms, t_init_ms, t_refine_ms, t_total_ms;
  int64_t alg_bytes;       /* algorithmic HBM bytes of all stages             */
  gk_level_stats levels[GK_MAX_LEVELS];
  /* Per-stage device time (GK_FLAG_STAGE_TIME), launches and algorithmic
   * bytes (compulsory 4-byte element traffic of the stage's kernels). */
  double stage_ms[GK_ST_COUNT];
  int32_t stage_launches[GK_ST_COUNT];
  int64_t stage_bytes[GK_ST_COUNT];
  /* Refinement rounds by the ordering path they took (all levels): [0] more
   * than max_moves selected (exact top-max_moves radix select), [1] of those,
   * compressed keys wider than 32 bits (64-bit key sort), [2] rounds that
   * tested vertices deferred by the previous round's gain threshold. */
  int64_t round_paths[3];
} gk_stats;

/* ---- context ------------------------------------------------------------ */
typedef struct gk_ctx gk_ctx;
int gk_ctx_create(int device, gk_ctx** out);
int gk_ctx_destroy(gk_ctx* ctx);
/* Stages that GK_FLAG_STAGE_TIME brackets with HIP events: bit s = stage s
   (GK_ST_*); default all.  Timing one stage keeps the event overhead out of
   the others.  No counterpart in the reference (it has no stage timers). */
int gk_ctx_set_stage_mask(gk_ctx* ctx, uint32_t mask);
/* Caps the device memory the context's pool may hand out at once (bytes;
   0 = no cap, the default).  An allocation beyond it fails with
   GK_ERR_NOMEM, as a failed hipMalloc does.  No counterpart in the reference
   (it allocates per call, gkway/graph_partitioner.hpp:253-273); for sharing a
   device between tenants, and for forcing allocation failures in tests. */
int gk_ctx_set_pool_limit(gk_ctx* ctx, int64_t bytes);
const char* gk_strerror(int status);
const char* gk_last_error(void);

/* ---- whole-graph partitioning ---------------------------------------------
 * gk_partition: host CSR in, host partition out (copies included).
 * gk_partition_device: CSR already in device memory, partition written to a
 * device buffer; the work runs on `stream` (NULL = the context's stream).     */
int gk_partition(gk_ctx* ctx, const gk_csr* host_graph, const gk_params* params,
                 int32_t* part_out, gk_stats* stats);
int gk_partition_device(gk_ctx* ctx, const gk_csr* dev_graph, const gk_params* params,
                        int32_t* dev_part_out, gk_stats* stats, void* stream);
/* graph_partitioner(GRAPH_FILE, OUT_FILE, k): read, partition, write
 * "<OUT_FILE>.out" (path taken as given; the C++ wrapper prefixes the cwd). */
int gk_partition_file(gk_ctx* ctx, const char* graph_file, const char* out_file,
                      int32_t k, gk_stats* stats);

/* ---- batched form ------------------------------------------------------------
 * Many graphs partitioned together: every stage of every level is one launch
 * over all graphs in flight (graphs packed into one CSR with global ids), and
 * the host initial partitions of some graphs overlap the device work of the
 * others.  Each graph's result is bit-identical to gk_partition on that graph
 * alone (no counterpart in the reference, which partitions one graph per
 * process run).
 *
 * gk_packed_csr: graphs concatenated; graph g owns vertices [v_off[g],
 * v_off[g+1]) and entries [e_off[g], e_off[g+1]); xadj is global and adjncy
 * holds global vertex ids inside the graph's own range.  Arrays are device
 * pointers except v_off / e_off (host).                                       */
typedef struct gk_packed_csr {
  int32_t ngraphs;
  int32_t n;
  int32_t nnz;
  const int32_t* xadj;     /* device, n + 1                                   */
  const int32_t* adjncy;   /* device, nnz                                     */
  const int32_t* adjwgt;   /* device, nnz, or NULL (unit)                     */
  const int32_t* vwgt;     /* device, n, or NULL (unit)                       */
  const int32_t* v_off;    /* host, ngraphs + 1                               */
  const int32_t* e_off;    /* host, ngraphs + 1                               */
} gk_packed_csr;

/* params[g] and stats[g] per graph (stats may be NULL); dev_part_out[n]
 * receives each graph's partition (part ids local to the graph). */
int gk_partition_packed(gk_ctx* ctx, const gk_packed_csr* pack, const gk_params* params,
                        int32_t* dev_part_out, gk_stats* stats, void* stream);
/* Host graphs in, host partitions out (packed and copied inside). */
int gk_partition_batch(gk_ctx* ctx, int32_t count, const gk_csr* graphs, const gk_params* params,
                       int32_t* const* parts_out, gk_stats* stats);

/* ---- serving form --------------------------------------------------------------
 * A request loop on its own host thread with continuous batching: submitted
 * graphs wait in FIFO order until their device footprint (CSR bytes times the
 * estimated level count, plus per-vertex scratch) fits the HBM budget, are
 * then admitted and join the batch engine at the next step boundary, and
 * advance level-synchronously with every other request in flight.  Each
 * result is bit-identical to gk_partition on that graph alone.  While a server
 * is open it owns the context: make no other call on that gk_ctx.  Every
 * request of one server must use max_group and max_moves equal to the first
 * request's.  No counterpart in the reference.                                */
typedef struct gk_server gk_server;

typedef struct gk_serve_config {
  int64_t hbm_budget;      /* device bytes admitted requests may hold
                              (0 = half of the free HBM at open)             */
  int32_t max_inflight;    /* requests stepped together (0 = no limit)        */
  int32_t workers;         /* host threads for initial partitions (0 = auto)  */
} gk_serve_config;

typedef struct gk_serve_stats {
  int64_t submitted, completed, failed, queued, inflight;
  int64_t edges_done;      /* undirected edges of completed requests         */
  int64_t steps;           /* engine steps run                               */
  int64_t footprint_now;   /* admitted footprint (estimate), bytes           */
  int64_t footprint_peak;
  int64_t device_peak;     /* peak device bytes held by the context's pool   */
  double p50_ms, p99_ms, max_ms;  /* latency = completion - submission      */
  double edges_per_s;      /* edges_done / (last completion - first submit)  */
  int64_t isolated;        /* re-runs of requests alone after a failed step  */
} gk_serve_stats;

int gk_serve_open(gk_ctx* ctx, const gk_serve_config* cfg, gk_server** out);
/* Copies the host graph; *request receives the request id. */
int gk_serve_submit(gk_server* s, const gk_csr* graph, const gk_params* params, int64_t* request);
/* count requests at once (graphs[i] at params[i]; ids to requests[i]): all
 * are validated first and queued together, in order, so that the loop's next
 * admission sees the whole burst; on an error none is queued.  No
 * counterpart in the reference (no serving there): for clients that submit
 * bursts. */
int gk_serve_submit_many(gk_server* s, int32_t count, const gk_csr* graphs, const gk_params* params,
                         int64_t* requests);
/* 1 = finished (part_out[n] and *stats written, the request is released),
 * 0 = still queued or running, < 0 = the request failed (also released).
 * GK_ERR_ARG for an unknown id, a result already taken, or a request another
 * caller is waiting on: each result is taken exactly once.
 * Failures are per request: when a step shared by several requests fails
 * (an allocation beyond the device memory, say), each of them is run again
 * by itself, and only a request that fails alone reports the error. */
int gk_serve_poll(gk_server* s, int64_t request, int32_t* part_out, gk_stats* stats);
/* Blocks until the request finishes; GK_OK or its error.  The waiting
 * caller owns the request: a second wait or a poll on it gets GK_ERR_ARG. */
int gk_serve_wait(gk_server* s, int64_t request, int32_t* part_out, gk_stats* stats);
int gk_serve_query(gk_server* s, gk_serve_stats* out);
/* Footprint estimate used for admission (bytes) of a graph at parameters p. */
int64_t gk_serve_footprint(int32_t n, int32_t nnz, const gk_params* p);
/* Finishes every queued and running request, waits for the callers inside
 * gk_serve_wait to take their results, then stops the loop and frees the
 * server.  Call it once, after which no other call may use s. */
int gk_serve_close(gk_server* s);

/* ---- host I/O --------------------------------------------------------------
 * gk_graph_read allocates the arrays with malloc; free with gk_graph_free.    */
int gk_graph_read(const char* path, gk_csr* out);
void gk_graph_free(gk_csr* g);
int gk_graph_write(const char* path, const gk_csr* g);
int gk_graph_check(const gk_csr* g, int check_symmetry);
/* Binary CSR files (this library's own format, no counterpart in the reference,
   whose only input is METIS text, gkway/graph.hpp:70-152): a 64-byte header
   and 64-byte-aligned int32 sections, so a graph is mapped instead of parsed.
   gk_graph_write_bin writes one; gk_graph_map_bin maps one read-only and
   points out's arrays into the mapping (no copy; adjwgt / vwgt NULL when the
   file has none; verify != 0 checks the checksum) until gk_graph_unmap.
   gk_graph_read also accepts binary CSR files (copying, like METIS text). */
int gk_graph_write_bin(const char* path, const gk_csr* g);
int gk_graph_map_bin(const char* path, int verify, gk_csr* out, void** handle);
void gk_graph_unmap(void* handle);
int gk_write_partition(const char* path, const int32_t* part, int32_t n);

/* ---- coarsest-level initial partitioner (host) -----------------------------
 * Deterministic multilevel recursive bisection with FM refinement; the role
 * METIS_PartGraphKway plays in the reference (gkway/metis_partition.hpp:46-47).
 * cap = maximum part weight.  Writes part[n]; *cut = weighted edge cut.       */
int gk_initpart(const gk_csr* host_graph, int32_t k, int64_t cap, int32_t seed,
                int32_t nthreads, int32_t* part, int64_t* cut);
/* Same computation as gk_params-driven rule: cap for total weight T. */
int64_t gk_part_cap(int64_t total_vwgt, int32_t k, int32_t ubfactor);
/* Coarsening floor of the TARGET policy for a level-0 graph of n vertices. */
int32_t gk_coarsen_floor(int32_t n0, int32_t k, int32_t policy, int32_t override_to);

/* ---- per-stage kernels (device pointers; each is bit-compared with the
 * oracle's stage of the same name) ------------------------------------------ */
int gk_k_match(gk_ctx* ctx, const gk_csr* dg, int32_t* cand, void* stream);
/* labels: max vertex id of the component of {v, cand[v]}; stamps: hop
 * distance to that vertex; *sweeps: synchronous sweeps executed. */
int gk_k_group(gk_ctx* ctx, int32_t n, const int32_t* cand, int32_t* label,
               int32_t* stamp, int32_t* sweeps, void* stream);
/* order: vertices sorted by (label, stamp, id); cmap: coarse id per vertex;
 * cvwgt: coarse vertex weights; returns nc in *nc. */
int gk_k_cluster(gk_ctx* ctx, int32_t n, const int32_t* label, const int32_t* stamp,
                 const int32_t* vwgt, int32_t max_group, int32_t* order, int32_t* cmap,
                 int32_t* cvwgt, int32_t* nc, void* stream);
/* Coarse CSR of the contraction; cxadj needs nc+1 entries, cadjncy/cadjwgt
 * need dg->nnz entries (upper bound); *cnnz receives the coarse entry count. */
int gk_k_contract(gk_ctx* ctx, const gk_csr* dg, const int32_t* cmap, int32_t nc,
                  int32_t* cxadj, int32_t* cadjncy, int32_t* cadjwgt, int32_t* cnnz,
                  void* stream);
int gk_k_project(gk_ctx* ctx, int32_t n_fine, const int32_t* cmap,
                 const int32_t* cpart, const int32_t* cboundary, int32_t* part,
                 int32_t* boundary, int32_t* gain, int32_t* dest, int32_t* updated,
                 void* stream);
/* Recompute boundary/gain/dest for vertices with updated[v] != 0. */
int gk_k_gain(gk_ctx* ctx, const gk_csr* dg, int32_t k, const int32_t* part,
              const int32_t* updated, int32_t* boundary, int32_t* gain, int32_t* dest,
              void* stream);
/* Selected moves sorted by (gain desc, id asc), truncated to max_moves.
 * Clears updated[].  moves[i] = vertex id; *nsel = selected before truncation. */
int gk_k_select(gk_ctx* ctx, const gk_csr* dg, const int32_t* part, const int32_t* boundary,
                const int32_t* gain, const int32_t* dest, const int32_t* pw, int64_t cap,
                int32_t max_moves, int32_t* updated, int32_t* moves, int32_t* nsel,
                void* stream);
/* Largest balanced prefix of the nmoves sorted moves: *pos (0-based, -1 when
 * nmoves == 0); part weights pw[k] are advanced by the prefix. */
int gk_k_balance(gk_ctx* ctx, const gk_csr* dg, int32_t k, const int32_t* moves,
                 int32_t nmoves, const int32_t* part, const int32_t* dest, int32_t* pw,
                 int64_t cap, int32_t* pos, void* stream);
/* Apply moves[0..pos]; marks movers and their neighbours updated; returns the
 * summed gain of the applied moves in *gain_sum. */
int gk_k_apply(gk_ctx* ctx, const gk_csr* dg, const int32_t* moves, int32_t pos,
               const int32_t* gain, const int32_t* dest, int32_t* part, int32_t* updated,
               int64_t* gain_sum, void* stream);
/* nrounds refinement rounds exactly as the V-cycle runs them, with the
 * product's round kernels: per round the movable set and its independent
 * selection, the best min(selected, max_moves) moves by (gain desc, id asc)
 * (all of them, or a top-max_moves radix select), the largest balanced
 * prefix, the apply, then the gain pass over the updated vertices.  The state
 * part/bnd/gain/dest (n ints) and pw (k ints) is device memory, read and
 * advanced in place; gain[v] > 0 only where bnd[v] != 0 (as the gain pass
 * leaves it).  The first round tests every candidate; later rounds use the
 * gain threshold the previous round set, as in the V-cycle.  Per round r
 * (host arrays of nrounds, may be NULL): nmoves[r] = moves ordered (0 when
 * none was selected), pos[r] = last applied position (-1: none).
 * *gain_sum = summed gain of every applied move.  order (device, may be
 * NULL): max_moves ints, the ordered moves of the last round that selected
 * any.  paths (host, may be NULL): as gk_stats.round_paths over these rounds. */
int gk_k_round(gk_ctx* ctx, const gk_csr* dg, int32_t k, int64_t cap, int32_t max_moves, int32_t nrounds,
               int32_t* part, int32_t* bnd, int32_t* gain, int32_t* dest, int32_t* pw, int32_t* nmoves,
               int32_t* pos, int64_t* gain_sum, int32_t* order, int64_t* paths, void* stream);
/* Weighted cut over entries (v,u), v<u, part differs; part weights pw[k]. */
int gk_k_cut(gk_ctx* ctx, const gk_csr* dg, int32_t k, const int32_t* part,
             int64_t* cut, int64_t* pw, void* stream);
/* gk_k_cut without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same kernels and the
 * same sums as gk_k_cut, but the outputs are DEVICE memory the caller owns
 * (dev_cut: 1 int64, dev_pw: k int64, written in stream order; no host
 * synchronisation), and the per-block slots live in a static 2 MiB device
 * array (2^18 64-bit slots per device, held until the process ends) instead
 * of the context's pool.  k above 8192 is rejected with GK_ERR_LIMIT.  Calls
 * no hipMalloc / hipFree.  All calls on one device share that array: calls
 * that overlap on one device (other streams or threads) race, so serialise
 * them.  dg->adjwgt and dg->vwgt are required. */
int gk_k_cut_noalloc(gk_ctx* ctx, const gk_csr* dg, int32_t k, const int32_t* part,
                     int64_t* dev_cut, int64_t* dev_pw, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GKWAY_H */
