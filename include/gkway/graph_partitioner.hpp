// graph_partitioner() — host C++ mirror of the reference's entry point
// (gkway/graph_partitioner.hpp:222 in wanluanlee/DAC_g-kway):
//   void graph_partitioner(const std::string& GRAPH_FILE,
//                          const std::string& OUT_FILE, const int NUM_PARTITIONS);
// reads a METIS graph, partitions it into NUM_PARTITIONS parts on the GPU and
// writes "<cwd>/<OUT_FILE>.out".  An overload takes full gk_params and
// returns the run statistics.  Errors throw std::runtime_error.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../gkway.h"
#include "graph.hpp"

inline gk_stats graph_partitioner(const std::string& GRAPH_FILE, const std::string& OUT_FILE,
                                  const gk_params& params, int device = 0) {
  gk::Graph graph(GRAPH_FILE);
  gk_ctx* ctx = nullptr;
  int rc = gk_ctx_create(device, &ctx);
  if (rc != GK_OK) throw std::runtime_error(std::string("gk_ctx_create: ") + gk_last_error());
  std::vector<int> part(graph.get_num_vertex());
  gk_stats st{};
  rc = gk_partition(ctx, &graph.csr(), &params, part.data(), &st);
  gk_ctx_destroy(ctx);
  if (rc != GK_OK)
    throw std::runtime_error(std::string("gk_partition: ") + gk_strerror(rc) + ": " + gk_last_error());
  graph.dump_result(part, OUT_FILE);
  return st;
}

inline void graph_partitioner(const std::string& GRAPH_FILE, const std::string& OUT_FILE,
                              const int NUM_PARTITIONS) {
  gk_params p;
  gk_params_default(&p, NUM_PARTITIONS);
  (void)graph_partitioner(GRAPH_FILE, OUT_FILE, p);
}
