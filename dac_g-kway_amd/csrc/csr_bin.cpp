// Binary CSR files: the input step without text parsing.  The reference reads
// only METIS text (gkway/graph.hpp:70-152); parsing the config-2 graph takes
// seconds, and a server re-reading graphs should not pay it.  This format has
// no counterpart in the reference.
//
// Layout (little-endian):
//   bytes 0..63   header: magic "GKCSR\x01\0\0", n (int64), nnz (int64),
//                 has_adjwgt (int32), has_vwgt (int32), checksum (uint64),
//                 24 reserved bytes (zero)
//   then, each starting at a multiple of 64 bytes:
//                 xadj[n + 1], adjncy[nnz], adjwgt[nnz] (if present),
//                 vwgt[n] (if present), all int32, 0-based adjncy
// checksum: FNV-1a over the 64-bit little-endian words of the sections in
// that order (each section zero-padded to 8 bytes).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/gkway.h"

namespace gkh {
void set_error(const std::string& msg);  // abi_host.cpp
}

namespace {

constexpr char kMagic[8] = {'G', 'K', 'C', 'S', 'R', 1, 0, 0};
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull, kFnvPrime = 0x100000001b3ull;

struct Header {
  char magic[8];
  int64_t n, nnz;
  int32_t has_adjwgt, has_vwgt;
  uint64_t checksum;
  int64_t reserved[3];
};
static_assert(sizeof(Header) == 64, "header is 64 bytes");

inline int64_t align64(int64_t x) { return (x + 63) & ~int64_t(63); }

struct Sections {
  int64_t xadj, adj, w, vw, end;
};

// Section offsets; absent sections get offset 0; end = end of the last one.
Sections layout(int64_t n, int64_t nnz, bool w, bool vw) {
  Sections s{};
  s.xadj = 64;
  int64_t e = s.xadj + 4 * (n + 1);
  s.adj = align64(e);
  e = s.adj + 4 * nnz;
  if (w) {
    s.w = align64(e);
    e = s.w + 4 * nnz;
  }
  if (vw) {
    s.vw = align64(e);
    e = s.vw + 4 * n;
  }
  s.end = e;
  return s;
}

uint64_t fnv_words(uint64_t h, const int32_t* a, int64_t count) {
  const int64_t pairs = count / 2;
  for (int64_t i = 0; i < pairs; ++i) {
    uint64_t x;
    std::memcpy(&x, a + 2 * i, 8);
    h = (h ^ x) * kFnvPrime;
  }
  if (count & 1) h = (h ^ uint64_t(uint32_t(a[count - 1]))) * kFnvPrime;
  return h;
}

uint64_t checksum(const gk_csr* g, bool w, bool vw) {
  uint64_t h = fnv_words(kFnvBasis, g->xadj, int64_t(g->n) + 1);
  h = fnv_words(h, g->adjncy, g->nnz);
  if (w) h = fnv_words(h, g->adjwgt, g->nnz);
  if (vw) h = fnv_words(h, g->vwgt, g->n);
  return h;
}

struct Mapping {
  void* base;
  size_t size;
};

}  // namespace

int gk_graph_write_bin(const char* path, const gk_csr* g) {
  if (!path || !g || !g->xadj || g->n < 0 || (g->nnz > 0 && !g->adjncy)) {
    gkh::set_error("null argument");
    return GK_ERR_ARG;
  }
  if (g->xadj[g->n] != g->nnz) {
    gkh::set_error("xadj[n] != nnz");
    return GK_ERR_GRAPH;
  }
  const bool w = g->adjwgt != nullptr, vw = g->vwgt != nullptr;
  const Sections s = layout(g->n, g->nnz, w, vw);
  Header h;
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, kMagic, 8);
  h.n = g->n;
  h.nnz = g->nnz;
  h.has_adjwgt = w;
  h.has_vwgt = vw;
  h.checksum = checksum(g, w, vw);
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    gkh::set_error(std::string("cannot open ") + path);
    return GK_ERR_IO;
  }
  static const char zeros[64] = {};
  int64_t at = 0;
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  at = sizeof h;
  auto put = [&](int64_t off, const int32_t* a, int64_t count) {
    if (!ok) return;
    if (off > at) ok = std::fwrite(zeros, 1, size_t(off - at), f) == size_t(off - at);
    if (ok && count > 0) ok = std::fwrite(a, 4, size_t(count), f) == size_t(count);
    at = off + 4 * count;
  };
  put(s.xadj, g->xadj, int64_t(g->n) + 1);
  put(s.adj, g->adjncy, g->nnz);
  if (w) put(s.w, g->adjwgt, g->nnz);
  if (vw) put(s.vw, g->vwgt, g->n);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    gkh::set_error(std::string("write failed: ") + path);
    return GK_ERR_IO;
  }
  return GK_OK;
}

int gk_graph_map_bin(const char* path, int verify, gk_csr* out, void** handle) {
  if (!path || !out || !handle) {
    gkh::set_error("null argument");
    return GK_ERR_ARG;
  }
  *handle = nullptr;
  const int fd = open(path, O_RDONLY);
  if (fd < 0) {
    gkh::set_error(std::string("cannot open ") + path);
    return GK_ERR_IO;
  }
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size < int64_t(sizeof(Header))) {
    close(fd);
    gkh::set_error("not a binary CSR file (too short)");
    return GK_ERR_FORMAT;
  }
  const size_t size = static_cast<size_t>(st.st_size);
  void* base = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (base == MAP_FAILED) {
    gkh::set_error(std::string("cannot map ") + path);
    return GK_ERR_IO;
  }
  auto fail = [&](int rc, const char* msg) {
    munmap(base, size);
    gkh::set_error(msg);
    return rc;
  };
  Header h;
  std::memcpy(&h, base, sizeof h);
  if (std::memcmp(h.magic, kMagic, 8) != 0) return fail(GK_ERR_FORMAT, "not a binary CSR file (magic)");
  if (h.n < 0 || h.nnz < 0 || h.n >= (int64_t(1) << 31) || h.nnz >= (int64_t(1) << 31))
    return fail(GK_ERR_LIMIT, "vertex/entry count out of range");
  const Sections s = layout(h.n, h.nnz, h.has_adjwgt != 0, h.has_vwgt != 0);
  if (int64_t(size) < s.end) return fail(GK_ERR_FORMAT, "binary CSR file truncated");
  const char* b = static_cast<const char*>(base);
  gk_csr g;
  g.n = static_cast<int32_t>(h.n);
  g.nnz = static_cast<int32_t>(h.nnz);
  g.xadj = reinterpret_cast<const int32_t*>(b + s.xadj);
  g.adjncy = reinterpret_cast<const int32_t*>(b + s.adj);
  g.adjwgt = h.has_adjwgt ? reinterpret_cast<const int32_t*>(b + s.w) : nullptr;
  g.vwgt = h.has_vwgt ? reinterpret_cast<const int32_t*>(b + s.vw) : nullptr;
  if (g.xadj[0] != 0 || g.xadj[g.n] != g.nnz) return fail(GK_ERR_FORMAT, "binary CSR file: bad xadj ends");
  if (verify && checksum(&g, g.adjwgt != nullptr, g.vwgt != nullptr) != h.checksum)
    return fail(GK_ERR_FORMAT, "binary CSR file: checksum mismatch");
  *out = g;
  *handle = new Mapping{base, size};
  return GK_OK;
}

void gk_graph_unmap(void* handle) {
  if (!handle) return;
  Mapping* m = static_cast<Mapping*>(handle);
  munmap(m->base, m->size);
  delete m;
}

// Whether a file starts with the binary CSR magic (used by gk_graph_read).
extern "C" __attribute__((visibility("hidden"))) bool gk_is_bin_csr(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  char m[8] = {};
  const bool ok = std::fread(m, 1, 8, f) == 8 && std::memcmp(m, kMagic, 8) == 0;
  std::fclose(f);
  return ok;
}
