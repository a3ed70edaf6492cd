// Host-only entry points of the C ABI: status strings, parameters, METIS
// file I/O and validation.  (Device work lives in driver.cpp.)
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/gkway.h"
#include "host_graph.hpp"

namespace gkh {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }
}  // namespace gkh

extern "C" {

const char* gk_strerror(int s) {
  switch (s) {
    case GK_OK: return "ok";
    case GK_ERR_ARG: return "invalid argument";
    case GK_ERR_IO: return "i/o error";
    case GK_ERR_FORMAT: return "malformed METIS graph file";
    case GK_ERR_GRAPH: return "graph violates an input invariant";
    case GK_ERR_HIP: return "HIP runtime error";
    case GK_ERR_NOMEM: return "out of memory";
    case GK_ERR_LIMIT: return "size beyond a supported limit";
    case GK_ERR_INTERNAL: return "internal consistency check failed";
    case GK_ERR_NODEVICE: return "no usable HIP device";
    default: return "unknown status";
  }
}

const char* gk_last_error(void) { return gkh::last_error(); }

void gk_params_default(gk_params* p, int32_t k) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->k = k;
  p->ubfactor = 30;
  p->max_group = 6;
  p->max_moves = 1024;
  p->coarsen_policy = GK_COARSEN_TARGET;
  p->coarsen_to = 0;
  p->max_levels = 0;
  p->seed = 1;
  p->initpart_threads = 0;
  p->flags = 0;
}

int32_t gk_coarsen_floor(int32_t n0, int32_t k, int32_t policy, int32_t override_to) {
  // REF: the reference coarsens while n > 20k (gkway/graph_partitioner.hpp:22,66).
  if (policy == GK_COARSEN_REF) return 20 * k;
  if (override_to > 0) return override_to;
  int lg = 0;
  while ((int64_t(1) << lg) < k) ++lg;
  if (lg < 1) lg = 1;
  // METIS-style n0 / (20 log2 k), capped at 240 coarse vertices per part
  // (larger coarsest graphs cost host time without a better final cut), and
  // at least 30 per part.
  int64_t a = int64_t(n0) / (20 * int64_t(lg));
  if (a > 240 * int64_t(k)) a = 240 * int64_t(k);
  int64_t b = 30 * int64_t(k);
  int64_t r = a > b ? a : b;
  return static_cast<int32_t>(r > INT32_MAX ? INT32_MAX : r);
}

__attribute__((visibility("hidden"))) bool gk_is_bin_csr(const char* path);  // csr_bin.cpp

int gk_graph_read(const char* path, gk_csr* out) {
  if (!path || !out) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  gkh::HostGraph g;
  if (gk_is_bin_csr(path)) {
    gk_csr m;
    void* h = nullptr;
    int rc = gk_graph_map_bin(path, 1, &m, &h);
    if (rc != GK_OK) return rc;
    g.n = m.n;
    g.xadj.assign(m.xadj, m.xadj + m.n + 1);
    g.adjncy.assign(m.adjncy, m.adjncy + m.nnz);
    if (m.adjwgt) g.adjwgt.assign(m.adjwgt, m.adjwgt + m.nnz); else g.adjwgt.assign(m.nnz, 1);
    if (m.vwgt) g.vwgt.assign(m.vwgt, m.vwgt + m.n); else g.vwgt.assign(m.n, 1);
    gk_graph_unmap(h);
  } else {
    std::string err;
    int rc = gkh::read_metis(path, g, err);
    if (rc != GK_OK) { gkh::set_error(err); return rc; }
  }
  const size_t nnz = g.adjncy.size();
  int32_t* xadj = static_cast<int32_t*>(malloc(sizeof(int32_t) * (g.n + 1)));
  int32_t* adj = static_cast<int32_t*>(malloc(sizeof(int32_t) * (nnz ? nnz : 1)));
  int32_t* w = static_cast<int32_t*>(malloc(sizeof(int32_t) * (nnz ? nnz : 1)));
  int32_t* vw = static_cast<int32_t*>(malloc(sizeof(int32_t) * (g.n ? g.n : 1)));
  if (!xadj || !adj || !w || !vw) {
    free(xadj); free(adj); free(w); free(vw);
    gkh::set_error("host allocation failed");
    return GK_ERR_NOMEM;
  }
  std::memcpy(xadj, g.xadj.data(), sizeof(int32_t) * (g.n + 1));
  if (nnz) {
    std::memcpy(adj, g.adjncy.data(), sizeof(int32_t) * nnz);
    std::memcpy(w, g.adjwgt.data(), sizeof(int32_t) * nnz);
  }
  if (g.n) std::memcpy(vw, g.vwgt.data(), sizeof(int32_t) * g.n);
  out->n = g.n;
  out->nnz = static_cast<int32_t>(nnz);
  out->xadj = xadj;
  out->adjncy = adj;
  out->adjwgt = w;
  out->vwgt = vw;
  return GK_OK;
}

void gk_graph_free(gk_csr* g) {
  if (!g) return;
  free(const_cast<int32_t*>(g->xadj));
  free(const_cast<int32_t*>(g->adjncy));
  free(const_cast<int32_t*>(g->adjwgt));
  free(const_cast<int32_t*>(g->vwgt));
  std::memset(g, 0, sizeof *g);
}

int gk_graph_write(const char* path, const gk_csr* c) {
  if (!path || !c || !c->xadj) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  gkh::HostGraph g;
  g.n = c->n;
  g.xadj.assign(c->xadj, c->xadj + c->n + 1);
  const int32_t nnz = g.xadj[c->n];
  g.adjncy.assign(c->adjncy, c->adjncy + nnz);
  g.has_ewgt = c->adjwgt != nullptr;
  g.has_vwgt = c->vwgt != nullptr;
  if (c->adjwgt) g.adjwgt.assign(c->adjwgt, c->adjwgt + nnz); else g.adjwgt.assign(nnz, 1);
  if (c->vwgt) g.vwgt.assign(c->vwgt, c->vwgt + c->n); else g.vwgt.assign(c->n, 1);
  std::string err;
  int rc = gkh::write_metis(path, g, err);
  if (rc != GK_OK) gkh::set_error(err);
  return rc;
}

int gk_graph_check(const gk_csr* c, int check_symmetry) {
  if (!c) { gkh::set_error("null graph"); return GK_ERR_ARG; }
  if (c->xadj && c->n >= 0 && c->xadj[c->n] != c->nnz) {
    gkh::set_error("xadj[n] != nnz");
    return GK_ERR_GRAPH;
  }
  std::string err;
  int rc = gkh::check_graph(c->n, c->xadj, c->adjncy, c->adjwgt, c->vwgt, check_symmetry != 0, err);
  if (rc != GK_OK) gkh::set_error(err);
  return rc;
}

int gk_write_partition(const char* path, const int32_t* part, int32_t n) {
  if (!path || (!part && n > 0)) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  std::string err;
  int rc = gkh::write_partition(path, part, n, err);
  if (rc != GK_OK) gkh::set_error(err);
  return rc;
}

}  // extern "C"
