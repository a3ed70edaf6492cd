// g-kway CLI: `g-kway graph_file num_partitions out_file [--eps PERMILLE]
// [--seed S] [--ref] [--json]`.  Same positional order as the reference
// (main/main.cu:11-17); writes "<cwd>/<out_file>.out" like
// gk::Graph::dump_result (gkway/graph.hpp:227-240).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/gkway/graph_partitioner.hpp"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s graph_file num_partitions out_file [--eps PERMILLE] [--seed S] [--ref] [--json]\n",
                 argv[0]);
    return 2;
  }
  const std::string graph_file = argv[1];
  const int k = std::atoi(argv[2]);
  const std::string out_file = argv[3];
  gk_params p;
  gk_params_default(&p, k);
  bool json = false;
  for (int i = 4; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--eps") && i + 1 < argc) p.ubfactor = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) p.seed = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--ref")) p.coarsen_policy = GK_COARSEN_REF;
    else if (!std::strcmp(argv[i], "--json")) json = true;
    else { std::fprintf(stderr, "unknown option %s\n", argv[i]); return 2; }
  }
  try {
    gk_stats st = graph_partitioner(graph_file, out_file, p);
    if (json) {
      std::printf("{\"n\": %d, \"k\": %d, \"cut\": %lld, \"imbalance\": %.5f, \"levels\": %d, "
                  "\"coarsest_n\": %d, \"ms_total\": %.3f, \"ms_coarsen\": %.3f, \"ms_init\": %.3f, "
                  "\"ms_refine\": %.3f}\n",
                  st.levels[0].n, k, (long long)st.cut, st.imbalance, st.nlevels, st.coarsest_n,
                  st.t_total_ms, st.t_coarsen_ms, st.t_init_ms, st.t_refine_ms);
    } else {
      std::printf("Total time: %.6fs\nCoarsen time: %.6fs\nInit partition time: %.6fs\nRefinement time: %.6fs\n",
                  st.t_total_ms / 1e3, st.t_coarsen_ms / 1e3, st.t_init_ms / 1e3, st.t_refine_ms / 1e3);
      std::printf("cut: %lld  imbalance: %.4f  levels: %d  coarsest: %d\n", (long long)st.cut, st.imbalance,
                  st.nlevels, st.coarsest_n);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "g-kway: %s\n", e.what());
    return 1;
  }
  return 0;
}
