// gk::Graph — host C++ mirror of the reference's graph reader
// (gkway/graph.hpp:17-51 in wanluanlee/DAC_g-kway), implemented on the C ABI
// (gk_graph_read / gk_write_partition).  Same class name, getters and
// dump_result, with the reference's array conventions:
//   get_adjp()           xadj, n + 1 offsets                  (graph.hpp:146-148)
//   get_adjncy()         neighbour ids as in the file, 1-based (graph.hpp:115,128)
//   get_adjncy_source()  1-based source vertex of every entry  (graph.hpp:116,129)
//   get_adjwgt()         edge weights, 1 when the file has none (graph.hpp:117)
//   get_vwgt()           vertex weights, 1 when the file has none (graph.hpp:105;
//                        the reference never reads them, this reader does)
//   get_num_edge()       directed entries, 2 x the header's edge count (:90)
// csr() gives the 0-based CSR the C ABI takes.  Errors throw
// std::runtime_error, as the reference's check_cuda does.
#pragma once
#include <cstdint>
#include <filesystem>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../gkway.h"

namespace gk {

class Graph {
 public:
  explicit Graph(const std::string& input_path) : _input_path(input_path) { _parse(); }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;
  ~Graph() { gk_graph_free(&_g); }

  const int* get_adjncy() const { return _adjncy1.data(); }
  const int* get_adjncy_source() const { return _source1.data(); }
  const int* get_adjp() const { return _g.xadj; }
  const int* get_adjwgt() const { return _g.adjwgt ? _g.adjwgt : _ones_e.data(); }
  const int* get_vwgt() const { return _g.vwgt ? _g.vwgt : _ones_v.data(); }
  size_t get_num_vertex() const { return static_cast<size_t>(_g.n); }
  // Directed adjacency entries (2 x edges), as the reference's _num_edge.
  size_t get_num_edge() const { return static_cast<size_t>(_g.nnz); }
  int get_total_vertex_weight() const {
    int64_t t = 0;
    for (int32_t v = 0; v < _g.n; ++v) t += _g.vwgt ? _g.vwgt[v] : 1;
    return static_cast<int>(t);
  }
  // The graph as the C ABI takes it (0-based adjncy; NULL weights = unit).
  const gk_csr& csr() const { return _g; }

  // Writes "<cwd>/<file_name>.out", one 0-based part id per line.
  void dump_result(const std::vector<int>& vertex_partition, const std::string& file_name) const {
    const std::string out = (std::filesystem::current_path() / (file_name + ".out")).string();
    int rc = gk_write_partition(out.c_str(), vertex_partition.data(),
                                static_cast<int32_t>(vertex_partition.size()));
    if (rc != GK_OK) throw std::runtime_error(std::string("dump_result: ") + gk_last_error());
  }

 private:
  void _parse() {
    int rc = gk_graph_read(_input_path.c_str(), &_g);
    if (rc != GK_OK)
      throw std::runtime_error(std::string("cannot read ") + _input_path + ": " + gk_strerror(rc) + ": " +
                               gk_last_error());
    const size_t nnz = static_cast<size_t>(_g.nnz);
    _adjncy1.resize(nnz);
    _source1.resize(nnz);
    for (int32_t v = 0; v < _g.n; ++v)
      for (int32_t j = _g.xadj[v]; j < _g.xadj[v + 1]; ++j) {
        _adjncy1[j] = _g.adjncy[j] + 1;
        _source1[j] = v + 1;
      }
    if (!_g.adjwgt) _ones_e.assign(nnz, 1);
    if (!_g.vwgt) _ones_v.assign(static_cast<size_t>(_g.n), 1);
  }
  std::string _input_path;
  gk_csr _g{};
  std::vector<int> _adjncy1, _source1, _ones_e, _ones_v;
};

}  // namespace gk
