// Host-side CSR graph container and METIS text I/O.
//
// Mirrors the data the reference's gk::Graph holds (gkway/graph.hpp:43-50):
// xadj (adjp), adjncy, adjwgt, vwgt as int arrays.  Differences, all on purpose:
//   * adjncy is 0-based here; the reference keeps the 1-based file ids
//     (gkway/graph.hpp:115,128) and subtracts 1 inside every kernel.
//   * the full METIS header is honoured (fmt digits for vertex sizes, vertex
//     weights and edge weights, ncon = 1), '%' comment lines are skipped and
//     counts are validated; the reference only recognises the token "11" and
//     reads it as "edge weights only" (gkway/graph.hpp:92-96).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gkh {

struct HostGraph {
  int32_t n = 0;
  std::vector<int32_t> xadj;    // n + 1
  std::vector<int32_t> adjncy;  // xadj[n] entries, 0-based
  std::vector<int32_t> adjwgt;  // xadj[n] entries (all 1 when unweighted)
  std::vector<int32_t> vwgt;    // n entries (all 1 when unweighted)
  bool has_ewgt = false;
  bool has_vwgt = false;
  int64_t nnz() const { return xadj.empty() ? 0 : xadj.back(); }
};

// Parse a METIS graph file.  Returns 0 on success, a negative gk status code
// on failure (message in err).
int read_metis(const std::string& path, HostGraph& g, std::string& err);

// Write a METIS graph file (fmt chosen from has_vwgt / has_ewgt).
int write_metis(const std::string& path, const HostGraph& g, std::string& err);

// Structural checks every partitioner entry point relies on: monotone xadj,
// ids in range, no self loops, weights >= 1, symmetric adjacency with equal
// weights in both directions, and weight totals that fit the int32 arithmetic
// of the kernels.  Returns 0 or a negative status code.
int check_graph(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                const int32_t* adjwgt, const int32_t* vwgt, bool check_symmetry,
                std::string& err);

// Write one 0-based part id per line (the reference's dump_result format,
// gkway/graph.hpp:227-240).
int write_partition(const std::string& path, const int32_t* part, int32_t n,
                    std::string& err);

}  // namespace gkh
