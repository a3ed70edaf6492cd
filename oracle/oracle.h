/* oracle.h — CPU oracle of the multilevel k-way partitioning path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker.
 * Nothing in the product (libgkway.so, the CLI) links or calls it.
 *
 * A single-threaded restatement of G-kway's V-cycle (wanluanlee/DAC_g-kway)
 * under the rule decisions that make it deterministic; each function cites
 * the reference code it restates.  Vertex ids are 0-based.  The initial
 * partition of the coarsest graph is an input callback: the reference calls
 * METIS there (gkway/metis_partition.hpp:46-47); tests pass oracle_initpart,
 * the restatement of the build's own partitioner for that step (initpart.cpp),
 * so no product code takes part in an oracle run.                           */
#ifndef GK_ORACLE_H
#define GK_ORACLE_H
#include <stdint.h>
#include "../include/gkway.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef int (*oracle_initpart_fn)(const gk_csr* g, int32_t k, int64_t cap, int32_t seed,
                                  int32_t nthreads, int32_t* part, int64_t* cut);

int oracle_partition(const gk_csr* g, const gk_params* p, oracle_initpart_fn initpart,
                     int32_t* part, gk_stats* stats);

void oracle_match(const gk_csr* g, int32_t* cand);
int32_t oracle_group(int32_t n, const int32_t* cand, int32_t* label, int32_t* stamp);
int32_t oracle_cluster(int32_t n, const int32_t* label, const int32_t* stamp,
                       const int32_t* vwgt, int32_t max_group, int32_t* order,
                       int32_t* cmap, int32_t* cvwgt);
int32_t oracle_contract(const gk_csr* g, const int32_t* cmap, int32_t nc, int32_t* cxadj,
                        int32_t* cadjncy, int32_t* cadjwgt);
void oracle_project(int32_t n_fine, const int32_t* cmap, const int32_t* cpart,
                    const int32_t* cboundary, int32_t* part, int32_t* boundary,
                    int32_t* gain, int32_t* dest, int32_t* updated);
void oracle_gain(const gk_csr* g, int32_t k, const int32_t* part, const int32_t* updated,
                 int32_t* boundary, int32_t* gain, int32_t* dest);
int32_t oracle_select(const gk_csr* g, const int32_t* part, const int32_t* boundary,
                      const int32_t* gain, const int32_t* dest, const int32_t* pw,
                      int64_t cap, int32_t max_moves, int32_t* updated, int32_t* moves,
                      int32_t* nsel);
int32_t oracle_balance(const gk_csr* g, int32_t k, const int32_t* moves, int32_t nmoves,
                       const int32_t* part, const int32_t* dest, int32_t* pw, int64_t cap);
int64_t oracle_apply(const gk_csr* g, const int32_t* moves, int32_t pos, const int32_t* gain,
                     const int32_t* dest, int32_t* part, int32_t* updated);
void oracle_cut(const gk_csr* g, int32_t k, const int32_t* part, int64_t* cut, int64_t* pw);
int32_t oracle_coarsen_floor(int32_t n0, int32_t k, int32_t policy, int32_t override_to);
/* Restatement of gk_initpart's rules (initpart.cpp); same signature as the
 * callback above.  nthreads is ignored (single-threaded). */
int oracle_initpart(const gk_csr* g, int32_t k, int64_t cap, int32_t seed, int32_t nthreads,
                    int32_t* part, int64_t* cut);
int64_t oracle_part_cap(int64_t total_vwgt, int32_t k, int32_t ubfactor);

#ifdef __cplusplus
}
#endif
#endif
