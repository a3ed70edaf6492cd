// METIS text reader/writer and input validation (host C++).
// Reader semantics follow the METIS manual's graph format; the reference's
// reader is gkway/graph.hpp:70-152 (see host_graph.hpp for the differences).
#include "host_graph.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <sys/mman.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <unistd.h>

#include "../../include/gkway.h"

namespace gkh {

namespace {

struct Reader {
  const char* p;
  const char* end;
  // Advance to the start of the next line that is not a comment; returns
  // false at end of file.  A line beginning with '%' is a comment.
  bool next_line(const char*& ls, const char*& le) {
    while (p < end) {
      const char* s = p;
      const char* e = static_cast<const char*>(memchr(p, '\n', end - p));
      if (!e) e = end;
      p = (e < end) ? e + 1 : end;
      if (s < e && *s == '%') continue;
      ls = s;
      le = e;
      return true;
    }
    return false;
  }
};

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Parse one non-negative decimal integer; returns false when no digits.
inline bool parse_i64(const char*& s, const char* e, int64_t& out) {
  while (s < e && is_space(*s)) ++s;
  if (s >= e) return false;
  bool neg = false;
  if (*s == '-') { neg = true; ++s; }
  if (s >= e || *s < '0' || *s > '9') return false;
  int64_t v = 0;
  while (s < e && *s >= '0' && *s <= '9') {
    v = v * 10 + (*s - '0');
    if (v > (int64_t(1) << 40)) return false;
    ++s;
  }
  if (s < e && !is_space(*s)) return false;
  out = neg ? -v : v;
  return true;
}

}  // namespace

int read_metis(const std::string& path, HostGraph& g, std::string& err) {
  int fd = open(path.c_str(), O_RDONLY);
  if (fd < 0) { err = "cannot open " + path; return GK_ERR_IO; }
  struct stat st;
  if (fstat(fd, &st) != 0) { close(fd); err = "cannot stat " + path; return GK_ERR_IO; }
  size_t sz = static_cast<size_t>(st.st_size);
  const char* base = nullptr;
  if (sz > 0) {
    void* m = mmap(nullptr, sz, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) { close(fd); err = "cannot map " + path; return GK_ERR_IO; }
    base = static_cast<const char*>(m);
  }
  close(fd);
  struct Unmap { const char* b; size_t s; ~Unmap() { if (b) munmap(const_cast<char*>(b), s); } } um{base, sz};

  Reader r{base, base + sz};
  const char *ls, *le;
  if (!r.next_line(ls, le)) { err = "empty file"; return GK_ERR_FORMAT; }
  int64_t hv[4];
  int nh = 0;
  {
    const char* s = ls;
    int64_t v;
    while (nh < 4 && parse_i64(s, le, v)) hv[nh++] = v;
    while (s < le && is_space(*s)) ++s;
    if (s != le || nh < 2) { err = "bad header line"; return GK_ERR_FORMAT; }
  }
  int64_t n = hv[0], m = hv[1];
  int64_t fmt = nh >= 3 ? hv[2] : 0;
  int64_t ncon = nh >= 4 ? hv[3] : 1;
  if (n < 0 || m < 0 || n >= (int64_t(1) << 31) || 2 * m >= (int64_t(1) << 31)) {
    err = "vertex/edge count out of range"; return GK_ERR_LIMIT;
  }
  if (fmt < 0 || fmt > 111 || (fmt % 10) > 1 || ((fmt / 10) % 10) > 1 || (fmt / 100) > 1) {
    err = "bad fmt field"; return GK_ERR_FORMAT;
  }
  if (ncon != 1) { err = "only ncon = 1 is supported"; return GK_ERR_FORMAT; }
  const bool has_vsize = (fmt / 100) == 1;
  const bool has_vw = ((fmt / 10) % 10) == 1;
  const bool has_ew = (fmt % 10) == 1;

  g.n = static_cast<int32_t>(n);
  g.has_ewgt = has_ew;
  g.has_vwgt = has_vw;
  g.xadj.assign(n + 1, 0);
  g.vwgt.assign(n, 1);
  g.adjncy.clear();
  g.adjwgt.clear();
  g.adjncy.reserve(2 * m);
  g.adjwgt.reserve(2 * m);

  for (int64_t v = 0; v < n; ++v) {
    if (!r.next_line(ls, le)) {
      // Trailing isolated vertices may be represented by missing lines only
      // when the file ends with blank lines; anything else is truncated.
      err = "file has fewer vertex lines than the header says";
      return GK_ERR_FORMAT;
    }
    const char* s = ls;
    int64_t x;
    if (has_vsize) {
      if (!parse_i64(s, le, x)) { err = "missing vertex size"; return GK_ERR_FORMAT; }
    }
    if (has_vw) {
      if (!parse_i64(s, le, x)) { err = "missing vertex weight on line of vertex " + std::to_string(v + 1); return GK_ERR_FORMAT; }
      if (x < 1 || x >= (int64_t(1) << 31)) { err = "vertex weight out of range"; return GK_ERR_GRAPH; }
      g.vwgt[v] = static_cast<int32_t>(x);
    }
    while (true) {
      if (!parse_i64(s, le, x)) break;
      if (x < 1 || x > n) { err = "neighbour id out of range at vertex " + std::to_string(v + 1); return GK_ERR_FORMAT; }
      int64_t w = 1;
      if (has_ew) {
        if (!parse_i64(s, le, w)) { err = "missing edge weight at vertex " + std::to_string(v + 1); return GK_ERR_FORMAT; }
        if (w < 1 || w >= (int64_t(1) << 31)) { err = "edge weight out of range"; return GK_ERR_GRAPH; }
      }
      if (static_cast<int64_t>(g.adjncy.size()) >= 2 * m) {
        err = "more adjacency entries than 2 x edges"; return GK_ERR_FORMAT;
      }
      g.adjncy.push_back(static_cast<int32_t>(x - 1));
      g.adjwgt.push_back(static_cast<int32_t>(w));
    }
    while (s < le && is_space(*s)) ++s;
    if (s != le) { err = "unparsable token on line of vertex " + std::to_string(v + 1); return GK_ERR_FORMAT; }
    g.xadj[v + 1] = static_cast<int32_t>(g.adjncy.size());
  }
  // Anything after the n vertex lines must be blank.
  while (r.next_line(ls, le)) {
    const char* s = ls;
    while (s < le && is_space(*s)) ++s;
    if (s != le) { err = "extra non-blank lines after the vertex lines"; return GK_ERR_FORMAT; }
  }
  if (static_cast<int64_t>(g.adjncy.size()) != 2 * m) {
    err = "adjacency entries (" + std::to_string(g.adjncy.size()) + ") != 2 x edges (" +
          std::to_string(2 * m) + ")";
    return GK_ERR_FORMAT;
  }
  return GK_OK;
}

int write_metis(const std::string& path, const HostGraph& g, std::string& err) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) { err = "cannot create " + path; return GK_ERR_IO; }
  std::vector<char> buf(1 << 20);
  setvbuf(f, buf.data(), _IOFBF, buf.size());
  int fmt = (g.has_vwgt ? 10 : 0) + (g.has_ewgt ? 1 : 0);
  if (fmt)
    fprintf(f, "%d %lld %03d\n", g.n, static_cast<long long>(g.nnz() / 2), fmt);
  else
    fprintf(f, "%d %lld\n", g.n, static_cast<long long>(g.nnz() / 2));
  char line[64];
  bool bad = false;
  for (int32_t v = 0; v < g.n; ++v) {
    bool first = true;
    if (g.has_vwgt) {
      if (fprintf(f, "%d", g.vwgt[v]) < 0) bad = true;
      first = false;
    }
    for (int32_t j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
      int len;
      if (g.has_ewgt)
        len = snprintf(line, sizeof line, first ? "%d %d" : " %d %d", g.adjncy[j] + 1, g.adjwgt[j]);
      else
        len = snprintf(line, sizeof line, first ? "%d" : " %d", g.adjncy[j] + 1);
      if (fwrite(line, 1, len, f) != size_t(len)) bad = true;
      first = false;
    }
    if (fputc('\n', f) == EOF) bad = true;
    if (bad) break;
  }
  // A failed flush of the buffer (a full disk, say) shows up in ferror or fclose.
  if (ferror(f)) bad = true;
  if (fclose(f) != 0 || bad) { err = "write failed: " + path; return GK_ERR_IO; }
  return GK_OK;
}

int check_graph(int32_t n, const int32_t* xadj, const int32_t* adjncy,
                const int32_t* adjwgt, const int32_t* vwgt, bool check_symmetry,
                std::string& err) {
  if (n < 0) { err = "n < 0"; return GK_ERR_ARG; }
  if (!xadj) { err = "xadj is null"; return GK_ERR_ARG; }
  if (xadj[0] != 0) { err = "xadj[0] != 0"; return GK_ERR_GRAPH; }
  for (int32_t v = 0; v < n; ++v)
    if (xadj[v + 1] < xadj[v]) { err = "xadj not monotone at " + std::to_string(v); return GK_ERR_GRAPH; }
  const int32_t nnz = xadj[n];
  if (nnz > 0 && !adjncy) { err = "adjncy is null"; return GK_ERR_ARG; }
  int64_t tv = 0, te = 0;
  for (int32_t v = 0; v < n; ++v) {
    int64_t w = vwgt ? vwgt[v] : 1;
    if (w < 1) { err = "vertex weight < 1 at " + std::to_string(v); return GK_ERR_GRAPH; }
    tv += w;
  }
  if (tv >= (int64_t(1) << 31)) { err = "total vertex weight >= 2^31"; return GK_ERR_LIMIT; }
  for (int32_t v = 0; v < n; ++v) {
    for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j) {
      int32_t u = adjncy[j];
      if (u < 0 || u >= n) { err = "neighbour out of range at vertex " + std::to_string(v); return GK_ERR_GRAPH; }
      if (u == v) { err = "self loop at vertex " + std::to_string(v); return GK_ERR_GRAPH; }
      int64_t w = adjwgt ? adjwgt[j] : 1;
      if (w < 1) { err = "edge weight < 1 at vertex " + std::to_string(v); return GK_ERR_GRAPH; }
      te += w;
    }
  }
  if (te / 2 >= (int64_t(1) << 31)) { err = "total edge weight >= 2^31"; return GK_ERR_LIMIT; }
  if (check_symmetry) {
    // Sort (u, v, w) triples of the transposed entries and compare.
    std::vector<std::pair<int64_t, int32_t>> fwd, bwd;
    fwd.reserve(nnz);
    bwd.reserve(nnz);
    for (int32_t v = 0; v < n; ++v)
      for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j) {
        int32_t u = adjncy[j];
        int32_t w = adjwgt ? adjwgt[j] : 1;
        fwd.emplace_back((int64_t(v) << 32) | uint32_t(u), w);
        bwd.emplace_back((int64_t(u) << 32) | uint32_t(v), w);
      }
    std::sort(fwd.begin(), fwd.end());
    std::sort(bwd.begin(), bwd.end());
    if (fwd != bwd) { err = "adjacency is not symmetric (entries or weights)"; return GK_ERR_GRAPH; }
    for (size_t i = 1; i < fwd.size(); ++i)
      if (fwd[i].first == fwd[i - 1].first) { err = "duplicate edge"; return GK_ERR_GRAPH; }
  }
  return GK_OK;
}

int write_partition(const std::string& path, const int32_t* part, int32_t n, std::string& err) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) { err = "cannot create " + path; return GK_ERR_IO; }
  std::vector<char> buf(1 << 20);
  setvbuf(f, buf.data(), _IOFBF, buf.size());
  char line[16];
  bool bad = false;
  for (int32_t v = 0; v < n && !bad; ++v) {
    int len = snprintf(line, sizeof line, "%d\n", part[v]);
    if (fwrite(line, 1, len, f) != size_t(len)) bad = true;
  }
  // A failed flush of the buffer (a full disk, say) shows up in ferror or fclose.
  if (ferror(f)) bad = true;
  if (fclose(f) != 0 || bad) { err = "write failed: " + path; return GK_ERR_IO; }
  return GK_OK;
}

}  // namespace gkh
