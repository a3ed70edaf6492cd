/* gkgen.h — seeded synthetic workload generators (libgkgen.so, host only).
 *
 * The reference ships one sample input (simple_test/delaunay_n11.graph) and
 * no generator; these produce the benchmark's graph families from a seed so
 * the GPU box can rebuild any workload without files:
 *   - 2-D Delaunay triangulations of uniform random points (exact integer
 *     predicates on a 2^24 x 2^24 grid, duplicates redrawn);
 *   - road-like graphs: Euclidean MST of uniform points plus the shortest
 *     remaining Delaunay edges up to m = round(avg_deg * n / 2);
 *   - layered circuit-like DAGs treated as undirected graphs (power-law
 *     fanout, edges to the next 1..window layers, weighted).
 * Arrays are malloc'ed; release with gkgen_free.                            */
#ifndef GKGEN_H
#define GKGEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gkgen_graph {
  int32_t n;
  int32_t nnz;
  int32_t* xadj;
  int32_t* adjncy;
  int32_t* adjwgt;   /* NULL when unit weights */
  int32_t* vwgt;     /* NULL when unit weights */
} gkgen_graph;

/* Uniform points in [0, 2^24)^2 drawn from `seed`; x[i], y[i]. */
int gkgen_points(int32_t n, uint64_t seed, int32_t* x, int32_t* y);
/* Delaunay triangulation of gkgen_points(n, seed). */
int gkgen_delaunay(int32_t n, uint64_t seed, gkgen_graph* out);
/* Delaunay triangulation of caller-given distinct points. */
int gkgen_delaunay_points(int32_t n, const int32_t* x, const int32_t* y, gkgen_graph* out);
/* Road-like graph: EMST + shortest extra Delaunay edges, avg degree avg_deg_x100/100. */
int gkgen_road(int32_t n, uint64_t seed, int32_t avg_deg_x100, gkgen_graph* out);
/* Layered DAG: n vertices in `layers` layers; fanout f in [1, max_fanout] with
 * P(f) ~ f^-exponent; each edge goes to a layer d in [1, window] ahead
 * (clipped to the last layer), target uniform in that layer; duplicates
 * dropped and the graph symmetrised.  Edge weights uniform in [1, wmax]
 * (unit when wmax <= 1); vertex weights uniform in [1, vwmax]. */
int gkgen_dag(int32_t n, int32_t layers, uint64_t seed, int32_t max_fanout,
              double exponent, int32_t window, int32_t wmax, int32_t vwmax,
              gkgen_graph* out);
void gkgen_free(gkgen_graph* g);
/* FNV-1a 64 over int32 arrays, element-wise (h ^= (uint32)x; h *= prime). */
uint64_t gkgen_fnv1a_i32(const int32_t* a, int64_t n, uint64_t h);

#ifdef __cplusplus
}
#endif
#endif
