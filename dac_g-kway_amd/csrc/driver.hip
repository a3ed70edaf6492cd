// Host driver of the MI355X partitioner: context, device memory pool, the
// multilevel V-cycle, and the C ABI entry points that launch device work.
//
// Flow of one gk_partition_device call (the reference's graph_partitioner,
// gkway/graph_partitioner.hpp:222-365):
//   coarsening levels  -> match, pointer-jump grouping, radix sort by
//                         (label, stamp, id), chunking, contraction
//   coarsest level     -> copy to host, gk_initpart (METIS's role), copy back
//   uncoarsening       -> per finer level: projection, gain pass, rounds of
//                         select / balance / apply / gain until no move
//   final              -> weighted cut and part weights of level 0
// One stream per context; the host synchronises once per level and once per
// refinement round to read the counts that size the next launches.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gkway.h"
#include "host_graph.hpp"
#include "batch.hpp"
#include "kernels.hpp"
#include "prim.hpp"
#include "runtime.hpp"


using gkd::DGraph;

using namespace gkrt;

namespace gkd {
bool kernel_events(hipEvent_t* start, hipEvent_t* stop) { return gkrt::StageClock::kernel_pair(start, stop); }
}  // namespace gkd

namespace gkrt {

// Smallest launches worth timing candidates for: matching over at least
// 1 Mi entries, a contraction row class of at least 16 Ki rows.
constexpr int64_t kTuneMinEntries = int64_t(1) << 20;
constexpr int32_t kTuneMinRows = 1 << 14;

// ---------------------------------------------------------------- stages
// Grouping: pointer jumping to the root pair, then the component maximum.
// Returns label/dist in the given buffers, the jump rounds in *rounds.
int group_stage(gk_ctx* ctx, hipStream_t st, int32_t n, const int32_t* cand, int32_t* label,
                Buf<int32_t>& dist_out, int32_t* rounds) {
  Pool& P = ctx->pool;
  Buf<int32_t> nx[2], ds[2], flags, compmax;
  GK_TRY(nx[0].alloc(P, n));
  GK_TRY(nx[1].alloc(P, n));
  GK_TRY(ds[0].alloc(P, n));
  GK_TRY(ds[1].alloc(P, n));
  GK_TRY(flags.alloc(P, kMaxJumpRounds + 8));
  GK_TRY(compmax.alloc(P, n));
  gkd::fill<int32_t>(flags.p, 0, kMaxJumpRounds + 8, st);
  GK_STAGE(ctx, GK_ST_GROUP, 16LL * n, gkd::launch_jump_init(n, cand, nx[0].p, ds[0].p, st));
  int32_t f = -1;
  int32_t r = 0;
  constexpr int kBatch = 4;
  while (f < 0) {
    if (r >= kMaxJumpRounds) {
      gkh::set_error("candidate graph did not reduce to pairs (asymmetric input?)");
      return GK_ERR_GRAPH;
    }
    ctx->clock.begin(GK_ST_GROUP);
    for (int i = 0; i < kBatch && r + i < kMaxJumpRounds; ++i) {
      const int rr = r + i;
      GK_TRY(gkd::launch_jump_round(n, cand, nx[rr & 1].p, ds[rr & 1].p, nx[(rr + 1) & 1].p,
                                    ds[(rr + 1) & 1].p, flags.p, rr, st));
    }
    ctx->clock.end(GK_ST_GROUP);
    int32_t hf[kBatch];
    const int cnt = std::min(kBatch, kMaxJumpRounds - r);
    GK_TRY(fetch(ctx, st, {{hf, flags.p + r, sizeof(int32_t) * cnt}}));
    for (int i = 0; i < cnt; ++i)
      if (hf[i] == 0) { f = r + i; break; }
    r += cnt;
  }
  // Round f moved nothing: buffers f&1 and (f+1)&1 agree.
  int32_t* next = nx[f & 1].p;
  // Rounds 0..f ran in full (24 bytes per vertex each); later ones exit early.
  count_stage(ctx, GK_ST_GROUP, 24LL * n * (f + 1));
  GK_STAGE(ctx, GK_ST_GROUP, 32LL * n, gkd::launch_label(n, cand, next, compmax.p, label, st));
  dist_out = std::move(ds[f & 1]);
  *rounds = f;
  return GK_OK;
}

// Sort vertices by (label, stamp, id) and chunk them; fills order (sorted
// ids), cmap, cptr (nc+1) and *nc.
int cluster_stage(gk_ctx* ctx, hipStream_t st, int32_t n, const int32_t* label, const int32_t* dist,
                  int32_t max_dist_bits, int32_t G, Buf<int32_t>& order, int32_t* cmap,
                  Buf<int32_t>& cptr, int32_t* nc) {
  Pool& P = ctx->pool;
  const int sbits = max_dist_bits;
  const int lbits = bits_for(n > 0 ? n - 1 : 0);
  Buf<uint64_t> keys, keys2;
  Buf<int32_t> ids, ids2, scratch, gstart, flagscan, ncd;
  GK_TRY(keys.alloc(P, n));
  GK_TRY(keys2.alloc(P, n));
  GK_TRY(ids.alloc(P, n));
  GK_TRY(ids2.alloc(P, n));
  GK_TRY(scratch.alloc(P, gkd::sort_scratch_ints(n) + gkd::scan_scratch_ints(n)));
  GK_TRY(gstart.alloc(P, n));
  GK_TRY(flagscan.alloc(P, n));
  GK_TRY(ncd.alloc(P, 1));
  GK_TRY(cptr.alloc(P, int64_t(n) + 1));
  bool alt = false;
  const int passes = (sbits + lbits + 7) / 8;
  GK_STAGE(ctx, GK_ST_SORT, 20LL * n + 32LL * n * passes,
           (gkd::launch_make_keys(n, label, dist, sbits, keys.p, ids.p, st) ||
                    gkd::radix_sort_pairs<uint64_t>(keys.p, ids.p, keys2.p, ids2.p, n, 0, sbits + lbits, scratch.p, st, &alt)
                ? GK_ERR_HIP
                : GK_OK));
  uint64_t* sk = alt ? keys2.p : keys.p;
  Buf<int32_t>& sid = alt ? ids2 : ids;
  GK_STAGE(ctx, GK_ST_CLUSTER, 64LL * n,
           gkd::launch_cluster(n, sk, sid.p, sbits, nullptr, G, gstart.p, flagscan.p, cmap, cptr.p,
                               nullptr, ncd.p, scratch.p, st));
  GK_TRY(fetch(ctx, st, {{nc, ncd.p, sizeof(int32_t)}}));
  order = std::move(sid);
  return GK_OK;
}

// Contraction of g under (cmap, cptr, order) into the coarse level c.
int contract_stage(gk_ctx* ctx, hipStream_t st, const DGraph& g, int32_t nc, const int32_t* cmap,
                   const int32_t* cptr, const int32_t* order, Level& c) {
  Pool& P = ctx->pool;
  Buf<int32_t> ub, toff, lists, lcnt, scratch, tadj, tw, len, cnnzd;
  Buf<int2> desc, mrow;
  GK_TRY(ub.alloc(P, nc));
  GK_TRY(toff.alloc(P, int64_t(nc) + 1));
  GK_TRY(lists.alloc(P, int64_t(4) * nc));
  GK_TRY(desc.alloc(P, int64_t(4) * nc));
  GK_TRY(mrow.alloc(P, g.n));
  GK_TRY(lcnt.alloc(P, 4));
  GK_TRY(scratch.alloc(P, gkd::scan_scratch_ints(int64_t(nc) + 1) + 64));
  GK_STAGE(ctx, GK_ST_CONTRACT_AUX, 20LL * g.n + 36LL * nc,
           gkd::launch_contract_bounds(g, nc, cptr, order, ub.p, toff.p, lists.p, desc.p, mrow.p, lcnt.p,
                                       scratch.p, st));
  int32_t hc[4];
  GK_TRY(fetch(ctx, st, {{hc, lcnt.p, sizeof hc}}));
  GK_TRY(tadj.alloc(P, g.nnz));
  GK_TRY(tw.alloc(P, g.nnz));
  GK_TRY(len.alloc(P, nc));
  const int lanes[3] = {16, 32, 64};
  // GK_CONTRACT_WIDE=1 runs the row kernels that sort (id, weight) pairs,
  // which otherwise serve only coarse graphs of 2^26 - 1 vertices or more.
  const char* wide_env = std::getenv("GK_CONTRACT_WIDE");
  const bool wide = wide_env && wide_env[0] == '1';
  auto row_classes = [&]() -> int {
    for (int cls = 0; cls < 3; ++cls) {
      const int32_t* rows = lists.p + int64_t(cls) * nc;
      auto run = [&](const Tuner::Shape& x) {
        return gkd::launch_contract_rows(g, x.lanes, x.block, rows, desc.p + int64_t(cls) * nc, hc[cls], nc,
                                         mrow.p, cmap, toff.p, tadj.p, tw.p, len.p, wide, st, steal_of(ctx));
      };
      Tuner::Shape shape{lanes[cls], 256};
      if (ctx->autotune && hc[cls] >= kTuneMinRows) {
        // Block size per row class; the lanes of a class are fixed by its rows.
        const uint64_t key = Tuner::key(1 + cls, g.n, g.nnz);
        if (const Tuner::Shape* s = ctx->tuner.find(key)) {
          shape = *s;
        } else {
          const int L = lanes[cls];
          GK_TRY(ctx->tuner.tune(key, {{L, 256}, {L, 64}, {L, 128}, {L, 512}}, st, run, &shape));
          continue;  // every candidate wrote the rows
        }
      }
      GK_TRY(run(shape));
    }
    return GK_OK;
  };
  ctx->clock.begin(GK_ST_CONTRACT);
  const int rc_rows = row_classes();
  ctx->clock.end(GK_ST_CONTRACT);
  if (rc_rows != GK_OK) return rc_rows;
  count_stage(ctx, GK_ST_CONTRACT, 0);  // bytes added once the coarse entry count is known
  if (hc[3] > 0) {
    const int32_t nbig = hc[3];
    const int32_t* rows = lists.p + int64_t(3) * nc;
    Buf<int32_t> boff, btot, bscr;
    GK_TRY(boff.alloc(P, int64_t(nbig) + 1));
    GK_TRY(btot.alloc(P, 1));
    GK_TRY(bscr.alloc(P, gkd::scan_scratch_ints(int64_t(nbig) + 1) + 64));
    GK_STAGE(ctx, GK_ST_CONTRACT_BIG, 16LL * nbig,
             gkd::launch_big_offsets(rows, nbig, ub.p, boff.p, btot.p, bscr.p, st));
    int32_t eb = 0;
    GK_TRY(fetch(ctx, st, {{&eb, btot.p, sizeof eb}}));
    Buf<uint64_t> k1, k2;
    Buf<int32_t> v1, v2, sscr;
    GK_TRY(k1.alloc(P, eb));
    GK_TRY(k2.alloc(P, eb));
    GK_TRY(v1.alloc(P, eb));
    GK_TRY(v2.alloc(P, eb));
    GK_TRY(sscr.alloc(P, gkd::sort_scratch_ints(eb)));
    bool alt = false;
    const int bpasses = (32 + bits_for(nbig - 1) + 7) / 8;
    GK_STAGE(ctx, GK_ST_CONTRACT_BIG, 24LL * eb + 36LL * eb * bpasses + 24LL * eb,
             (gkd::launch_contract_big_gather(g, rows, nbig, boff.p, cptr, order, cmap, k1.p, v1.p, st) ||
                      gkd::radix_sort_pairs<uint64_t>(k1.p, v1.p, k2.p, v2.p, eb, 0, 32 + bits_for(nbig - 1),
                                                      sscr.p, st, &alt) ||
                      gkd::launch_contract_big_merge(rows, nbig, boff.p, alt ? k2.p : k1.p, alt ? v2.p : v1.p,
                                                     toff.p, tadj.p, tw.p, len.p, st)
                  ? GK_ERR_HIP
                  : GK_OK));
  }
  GK_TRY(cnnzd.alloc(P, 1));
  GK_TRY(c.xadj.alloc(P, int64_t(nc) + 1));
  GK_STAGE(ctx, GK_ST_CONTRACT_AUX, 12LL * nc, gkd::launch_len_scan(nc, len.p, c.xadj.p, cnnzd.p, scratch.p, st));
  int32_t cnnz = 0;
  GK_TRY(fetch(ctx, st, {{&cnnz, cnnzd.p, sizeof cnnz}}));
  GK_TRY(c.adj.alloc(P, cnnz));
  GK_TRY(c.w.alloc(P, cnnz));
  // Row merge: member rows and the cmap gathers read once, merged rows
  // written once (known now that cnnz is).
  if (ctx->cur) ctx->cur->stage_bytes[GK_ST_CONTRACT] += 8LL * g.n + 20LL * nc + 12LL * g.nnz + 8LL * cnnz;
  GK_STAGE(ctx, GK_ST_CONTRACT_AUX, 12LL * nc + 16LL * cnnz,
           gkd::launch_contract_copy(nc, toff.p, len.p, c.xadj.p, tadj.p, tw.p, c.adj.p, c.w.p, st));
  c.g.n = nc;
  c.g.nnz = cnnz;
  c.g.xadj = c.xadj.p;
  c.g.adj = c.adj.p;
  c.g.w = c.w.p;
  return GK_OK;
}

int match_stage(gk_ctx* ctx, hipStream_t st, const DGraph& g, int32_t* cand) {
  const int def = gkd::lanes_for_degree(g.n ? double(g.nnz) / g.n : 0.0);
  Tuner::Shape shape{def, 256};
  auto run = [&](const Tuner::Shape& x) { return gkd::launch_match(g, x.lanes, x.block, cand, st, steal_of(ctx)); };
  if (ctx->autotune && g.nnz >= kTuneMinEntries) {
    const uint64_t key = Tuner::key(0, g.n, g.nnz);
    if (const Tuner::Shape* s = ctx->tuner.find(key)) {
      shape = *s;
    } else {
      // Vertices per wavefront around the degree table's choice, three block sizes.
      std::vector<Tuner::Shape> c;
      for (int l : {def, def / 2, def * 2})
        if (l >= 1 && l <= 64)
          for (int b : {256, 128, 512}) c.push_back({l, b});
      return ctx->tuner.tune(key, c, st, run, &shape);  // every candidate wrote cand
    }
  }
  return run(shape);
}

int gain_lanes(const DGraph& g) {
  const double avg = g.n > 0 ? double(g.nnz) / g.n : 0.0;
  int s = 4;
  while (s < 64 && s < 2.0 * avg) s <<= 1;
  return s;
}

}  // namespace gkrt

namespace {

// ---------------------------------------------------------------- V-cycle
int run_vcycle(gk_ctx* ctx, hipStream_t st, const DGraph& g0in, const gk_params& prm,
               int32_t* dpart_out, gk_stats* stats) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  Pool& P = ctx->pool;
  const int32_t k = prm.k;
  const bool trace = (prm.flags & GK_FLAG_TRACE) != 0;
  gk_stats S;
  std::memset(&S, 0, sizeof S);
  ctx->cur = &S;
  ctx->clock.start((prm.flags & GK_FLAG_STAGE_TIME) != 0, st);
  struct Reset {
    gk_ctx* c;
    ~Reset() {
      c->cur = nullptr;
      c->clock.start(false, nullptr);
      gkrt::clear_launch_options(c);
    }
  } reset{ctx};
  GK_TRY(gkrt::set_launch_options(ctx, prm.flags));

  std::vector<Level> L;
  L.reserve(GK_MAX_LEVELS + 1);
  L.emplace_back();
  L[0].g = g0in;
  const int32_t n0 = g0in.n;
  if (!g0in.w && g0in.nnz > 0) {
    GK_TRY(L[0].w.alloc(P, g0in.nnz));
    gkd::fill<int32_t>(L[0].w.p, 1, g0in.nnz, st);
    L[0].g.w = L[0].w.p;
  }
  if (!g0in.vw) {
    GK_TRY(L[0].vw.alloc(P, n0));
    gkd::fill<int32_t>(L[0].vw.p, 1, n0, st);
    L[0].g.vw = L[0].vw.p;
  }
  int64_t T = 0;
  {
    Buf<unsigned long long> tw;
    GK_TRY(tw.alloc(P, 1));
    GK_TRY(gkd::launch_sum(L[0].g.vw, n0, tw.p, st));
    unsigned long long ht = 0;
    GK_TRY(fetch(ctx, st, {{&ht, tw.p, sizeof ht}}));
    T = static_cast<int64_t>(ht);
  }
  if (T >= (int64_t(1) << 31)) { gkh::set_error("total vertex weight >= 2^31"); return GK_ERR_LIMIT; }
  const int64_t cap = gk_part_cap(T, k, prm.ubfactor);
  const int32_t floor_n = gk_coarsen_floor(n0, k, prm.coarsen_policy, prm.coarsen_to);
  const int32_t maxlv = prm.max_levels > 0 ? std::min(prm.max_levels, GK_MAX_LEVELS) : GK_MAX_LEVELS;
  S.total_vwgt = T;
  S.cap = cap;

  // ---------------------------------------------------------- coarsening
  while (static_cast<int32_t>(L.size()) < maxlv) {
    Level& f = L.back();
    const int32_t n = f.g.n;
    if (!(n > floor_n)) break;
    gk_level_stats& ls = S.levels[L.size() - 1];
    Buf<int32_t> cand, label, dist, order, cptr, cmap;
    GK_TRY(cand.alloc(P, n));
    GK_TRY(label.alloc(P, n));
    GK_TRY(cmap.alloc(P, n));
    GK_STAGE(ctx, GK_ST_MATCH, 16LL * f.g.nnz + 8LL * n + 4,
             match_stage(ctx, st, f.g, cand.p));
    int32_t rounds = 0;
    GK_TRY(group_stage(ctx, st, n, cand.p, label.p, dist, &rounds));
    int32_t nc = 0;
    GK_TRY(cluster_stage(ctx, st, n, label.p, dist.p, rounds + 1, prm.max_group, order, cmap.p, cptr, &nc));
    if (nc == n || nc < k + 1) break;
    if (prm.coarsen_policy != GK_COARSEN_REF && nc < floor_n) break;
    Level c;
    GK_TRY(c.vw.alloc(P, nc));
    GK_STAGE(ctx, GK_ST_CLUSTER, 8LL * n + 12LL * nc, gkd::launch_cvwgt(nc, cptr.p, order.p, f.g.vw, c.vw.p, st));
    c.g.vw = c.vw.p;
    GK_TRY(contract_stage(ctx, st, f.g, nc, cmap.p, cptr.p, order.p, c));
    ls.sweeps = rounds;
    if (trace) {
      std::vector<int32_t> h;
      GK_TRY(to_host(st, h, cand.p, n));
      ls.h_cand = fnv_host(h);
      GK_TRY(to_host(st, h, cmap.p, n));
      ls.h_cmap = fnv_host(h);
      std::vector<int32_t> a, b, d, e;
      GK_TRY(to_host(st, a, c.g.xadj, int64_t(nc) + 1));
      GK_TRY(to_host(st, b, c.g.adj, c.g.nnz));
      GK_TRY(to_host(st, d, c.g.w, c.g.nnz));
      GK_TRY(to_host(st, e, c.g.vw, nc));
      ls.h_coarse = fnv_host(e, fnv_host(d, fnv_host(b, fnv_host(a))));
    }
    f.cmap = std::move(cmap);
    L.push_back(std::move(c));
  }
  const int32_t nl = static_cast<int32_t>(L.size());
  GK_HIP(hipStreamSynchronize(st));
  const auto t1 = clk::now();

  // ---------------------------------------------------------- initial partition
  Level& top = L.back();
  const int32_t nt = top.g.n;
  // Largest row of every level refined later (it decides whether a level
  // needs the heavy-row gain kernel and the long-row selection pass): read
  // back with the initial partition's upload, not one host sync per level.
  Buf<int32_t> dmaxall;
  std::vector<int32_t> hmaxdeg(nl, 0);
  GK_TRY(dmaxall.alloc(P, nl));
  for (int32_t l = 0; l + 1 < nl; ++l) GK_TRY(gkd::launch_max_deg(L[l].g, dmaxall.p + l, st));
  // The coarsest graph comes back through the context's pinned staging in
  // one piece (four copies, one host sync) and is partitioned from there.
  const int64_t tn = top.g.nnz;
  void* hdn = nullptr;
  GK_TRY(ctx->dl.get(st, sizeof(int32_t) * (2 * size_t(nt) + 1 + 2 * size_t(tn)), &hdn));
  int32_t* hx = static_cast<int32_t*>(hdn);
  int32_t* ha = hx + nt + 1;
  int32_t* hw = ha + tn;
  int32_t* hv = hw + tn;
  GK_HIP(hipMemcpyAsync(hx, top.g.xadj, sizeof(int32_t) * (size_t(nt) + 1), hipMemcpyDeviceToHost, st));
  if (tn > 0) {
    GK_HIP(hipMemcpyAsync(ha, top.g.adj, sizeof(int32_t) * size_t(tn), hipMemcpyDeviceToHost, st));
    GK_HIP(hipMemcpyAsync(hw, top.g.w, sizeof(int32_t) * size_t(tn), hipMemcpyDeviceToHost, st));
  }
  if (nt > 0) GK_HIP(hipMemcpyAsync(hv, top.g.vw, sizeof(int32_t) * size_t(nt), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  gk_csr hc;
  hc.n = nt;
  hc.nnz = tn;
  hc.xadj = hx;
  hc.adjncy = ha;
  hc.adjwgt = hw;
  hc.vwgt = hv;
  // Initial part, boundary flags, part weights and cut go up through pinned
  // staging as well.
  void* hup = nullptr;
  GK_TRY(ctx->up.get(st, sizeof(int32_t) * (2 * size_t(nt) + size_t(k)) + sizeof(long long), &hup));
  long long* hcut0 = static_cast<long long*>(hup);
  int32_t* hp = reinterpret_cast<int32_t*>(hcut0 + 1);
  int32_t* hb = hp + nt;
  int32_t* hpw = hb + nt;
  int64_t cut0 = 0;
  GK_TRY(gk_initpart(&hc, k, cap, prm.seed, prm.initpart_threads, hp, &cut0));
  std::fill(hb, hb + nt, 0);
  std::fill(hpw, hpw + k, 0);
  for (int32_t v = 0; v < nt; ++v) {
    hpw[hp[v]] += hv[v];
    for (int32_t j = hx[v]; j < hx[v + 1]; ++j)
      if (hp[ha[j]] != hp[v]) { hb[v] = 1; break; }
  }
  *hcut0 = cut0;
  S.init_cut = cut0;
  Buf<int32_t> partA, partB, bndA, bndB, pw;
  GK_TRY(partA.alloc(P, n0));
  GK_TRY(partB.alloc(P, n0));
  GK_TRY(bndA.alloc(P, n0));
  GK_TRY(bndB.alloc(P, n0));
  GK_TRY(pw.alloc(P, k));
  Buf<long long> dcut;
  GK_TRY(dcut.alloc(P, 1));
  if (nt > 0) {
    GK_HIP(hipMemcpyAsync(partA.p, hp, sizeof(int32_t) * size_t(nt), hipMemcpyHostToDevice, st));
    GK_HIP(hipMemcpyAsync(bndA.p, hb, sizeof(int32_t) * size_t(nt), hipMemcpyHostToDevice, st));
  }
  GK_HIP(hipMemcpyAsync(pw.p, hpw, sizeof(int32_t) * size_t(k), hipMemcpyHostToDevice, st));
  GK_HIP(hipMemcpyAsync(dcut.p, hcut0, sizeof(long long), hipMemcpyHostToDevice, st));
  GK_TRY(fetch(ctx, st, {{hmaxdeg.data(), dmaxall.p, sizeof(int32_t) * size_t(nl)}}));
  const uint64_t hpart_top = trace ? fnv_host(std::vector<int32_t>(hp, hp + nt)) : 0;
  ctx->dl.drained();  // the stream is synchronised: both staging arenas are free again
  ctx->up.drained();
  S.levels[nl - 1].cut_after = cut0;
  if (trace) S.levels[nl - 1].h_part = hpart_top;
  const auto t2 = clk::now();

  // ---------------------------------------------------------- uncoarsening
  // Rounds run back to back on the device (select_list, round_b with one
  // job, gain passes); the host reads the last round's move count every
  // kRoundsPerCheck rounds, every 8, 16, then 32 once a level has run 32
  // (a check costs a copy and a host sync).  A round with nothing selected
  // changes nothing.
  int32_t* cur_part = partA.p;
  int32_t* cur_bnd = bndA.p;
  int32_t* oth_part = partB.p;
  int32_t* oth_bnd = bndB.p;
  if (nl > 1) {
    const int32_t M = prm.max_moves;
    Buf<int32_t> xlist, dthr, wpos, sscr;
    Buf<int32_t> gain, dest, upd, heavy, hcnt, wl, list, active, drounds, lastcnt, dvoff, dpw_off, dkk;
    Buf<long long> dcap, dmoved, dpaths, dabytes;
    Buf<uint64_t> cand, movable, selbits, defbits;
    Buf<uint32_t> hkey;
    Buf<unsigned long long> work;
    GK_TRY(dpaths.alloc(P, 3));
    gkd::fill<long long>(dpaths.p, 0, 3, st);
    // gain statistics: a (vertices, entries) pair per block of the gain
    // launches (grids of at most kGainSlots blocks), summed at the end
    constexpr int kGainSlots = gkd::kGainMaxBlocks;
    GK_TRY(work.alloc(P, 2 * kGainSlots));
    gkd::fill<unsigned long long>(work.p, 0ull, 2 * kGainSlots, st);
    // the selection kernels' algorithmic bytes, likewise per block slot
    Buf<unsigned long long> selbytes;
    GK_TRY(selbytes.alloc(P, gkd::kSelectByteSlots));
    gkd::fill<unsigned long long>(selbytes.p, 0ull, gkd::kSelectByteSlots, st);
    GK_TRY(gain.alloc(P, n0));
    GK_TRY(dest.alloc(P, n0));
    GK_TRY(upd.alloc(P, n0));
    GK_TRY(heavy.alloc(P, n0));
    GK_TRY(wl.alloc(P, n0));
    GK_TRY(list.alloc(P, n0));
    GK_TRY(hkey.alloc(P, n0));
    GK_TRY(cand.alloc(P, (int64_t(n0) + 63) / 64 + 1));
    const int64_t nwords0 = (int64_t(n0) + 63) / 64;
    GK_TRY(movable.alloc(P, nwords0 + 1));
    GK_TRY(selbits.alloc(P, nwords0 + 1));
    GK_TRY(defbits.alloc(P, nwords0 + 1));
    GK_TRY(wpos.alloc(P, nwords0 + 1));
    GK_TRY(sscr.alloc(P, gkd::select_scratch_ints(n0)));
    GK_TRY(xlist.alloc(P, n0));
    GK_TRY(dthr.alloc(P, 1));
    GK_TRY(hcnt.alloc(P, 2));  // [0] gain worklist count, [1] heavy-row count
    GK_TRY(active.alloc(P, 1));
    // Per-level rounds, moves and cut after the level, read back once at the end.
    Buf<long long> dcutlv;
    GK_TRY(drounds.alloc(P, nl));
    GK_TRY(lastcnt.alloc(P, 1));
    GK_TRY(dmoved.alloc(P, nl));
    GK_TRY(dabytes.alloc(P, nl));
    GK_TRY(dcutlv.alloc(P, nl));
    GK_TRY(dvoff.alloc(P, 2));
    GK_TRY(dpw_off.alloc(P, 2));
    GK_TRY(dkk.alloc(P, 1));
    GK_TRY(dcap.alloc(P, 1));
    gkd::fill<uint64_t>(selbits.p, 0ull, nwords0 + 1, st);
    const gkd::SelectBufs sb{movable.p, selbits.p, defbits.p, wpos.p, list.p, sscr.p, selbytes.p};
    gkd::fill<int32_t>(dvoff.p, 0, 1, st);
    gkd::fill<int32_t>(dpw_off.p, 0, 1, st);
    gkd::fill<int32_t>(dpw_off.p + 1, k, 1, st);
    gkd::fill<int32_t>(dkk.p, k, 1, st);
    gkd::fill<long long>(dcap.p, static_cast<long long>(cap), 1, st);
    constexpr int kRoundsPerCheck = 4;
    for (int32_t l = nl - 2; l >= 0; --l) {
      Level& f = L[l];
      const DGraph& g = f.g;
      gk_level_stats& ls = S.levels[l];
      GK_STAGE(ctx, GK_ST_PROJECT, 32LL * g.n,
               gkd::launch_project(g.n, f.cmap.p, cur_part, cur_bnd, oth_part, oth_bnd, gain.p, dest.p,
                                   upd.p, st));
      std::swap(cur_part, oth_part);
      std::swap(cur_bnd, oth_bnd);
      const int lanes = gain_lanes(g);
      // Rows longer than the lane count: the gain pass takes them itself for
      // k <= kGainInlineK, else they go to the heavy-row kernel, whose
      // launches are skipped when there are none.
      const int32_t maxdeg = hmaxdeg[l];
      const bool heavy_rows = maxdeg > lanes && gkd::gain_heavy_separate(k);
      // After projection every gain is 0: no candidate yet.
      gkd::fill<uint64_t>(cand.p, 0ull, (int64_t(g.n) + 63) / 64 + 1, st);
      gkd::fill<int32_t>(dvoff.p + 1, g.n, 1, st);
      gkd::fill<int32_t>(drounds.p + l, 0, 1, st);
      gkd::fill<int32_t>(lastcnt.p, 1, 1, st);
      gkd::fill<int32_t>(dthr.p, 1, 1, st);
      gkd::fill<long long>(dmoved.p + l, 0, 1, st);
      gkd::fill<long long>(dabytes.p + l, 0, 1, st);
      // The projected boundary becomes the first gain worklist.
      gkd::fill<int32_t>(hcnt.p, 0, 2, st);
      GK_STAGE(ctx, GK_ST_PROJECT, 4LL * g.n, gkd::launch_compact(g.n, upd.p, wl.p, hcnt.p, st));
      GK_STAGE(ctx, GK_ST_GAIN, 0,
               gkd::launch_gain_wl(g, k, lanes, wl.p, hcnt.p, heavy.p, cur_part, upd.p, cur_bnd, gain.p, dest.p,
                                   work.p, g.n, cand.p, st, steal_of(ctx)));
      if (heavy_rows)
        GK_STAGE(ctx, GK_ST_GAIN_HEAVY, 0,
                 gkd::launch_gain_heavy(g, k, heavy.p, hcnt.p, cur_part, upd.p, cur_bnd, gain.p, dest.p, work.p,
                                        cand.p, st));
      // A round updates at most M movers and their neighbours.
      const int64_t avg_deg = g.n > 0 ? (int64_t(g.nnz) + g.n - 1) / g.n : 0;
      const int64_t wl_items = std::min<int64_t>(g.n, int64_t(M) * (1 + 2 * avg_deg));
      auto one_round = [&]() -> int {
        GK_STAGE(ctx, GK_ST_SELECT, 0,
                 gkd::launch_select_list(g, nullptr, gain.p, dest.p, pw.p, dpw_off.p, dcap.p, dthr.p, cand.p, sb,
                                         hcnt.p, active.p, st));
        GK_STAGE(ctx, GK_ST_BALANCE, 0,
                 gkd::launch_round_b(1, k, dvoff.p, M, dkk.p, dcap.p, dpw_off.p, g, cur_part, dest.p, gain.p, pw.p,
                                     dcut.p, drounds.p + l, dmoved.p + l, sb, xlist.p, hkey.p, upd.p, wl.p,
                                     hcnt.p, lastcnt.p, active.p, cand.p, dthr.p, dpaths.p, nullptr,
                                     dabytes.p + l, st));
        GK_STAGE(ctx, GK_ST_GAIN, 0,
                 gkd::launch_gain_wl(g, k, lanes, wl.p, hcnt.p, heavy.p, cur_part, upd.p, cur_bnd, gain.p, dest.p,
                                     work.p, wl_items, cand.p, st, steal_of(ctx)));
        if (heavy_rows)
          GK_STAGE(ctx, GK_ST_GAIN_HEAVY, 0,
                   gkd::launch_gain_heavy(g, k, heavy.p, hcnt.p, cur_part, upd.p, cur_bnd, gain.p, dest.p, work.p,
                                          cand.p, st));
        return GK_OK;
      };
      // GK_FLAG_GRAPH_ROUNDS: kRoundsPerCheck rounds are captured once per
      // level and every check interval (a multiple of it) replays them.
      RoundGraph rg;
      if ((prm.flags & GK_FLAG_GRAPH_ROUNDS) && !ctx->clock.timed(GK_ST_SELECT) &&
          !ctx->clock.timed(GK_ST_BALANCE) && !ctx->clock.timed(GK_ST_GAIN) && !ctx->clock.timed(GK_ST_GAIN_HEAVY))
        GK_TRY(rg.capture(ctx, st, [&]() -> int {
          for (int r = 0; r < kRoundsPerCheck; ++r) GK_TRY(one_round());
          return GK_OK;
        }));
      int per_check = kRoundsPerCheck, ran = 0;
      while (true) {
        if (rg.ready()) {
          for (int r = 0; r < per_check; r += kRoundsPerCheck) GK_TRY(rg.replay(ctx, st));
        } else {
          for (int r = 0; r < per_check; ++r) GK_TRY(one_round());
        }
        ran += per_check;
        int32_t live = 0;
        GK_TRY(fetch(ctx, st, {{&live, active.p, sizeof live}}));
        if (live == 0) break;
        if (ran >= 32 && per_check < 32) per_check *= 2;
      }
      GK_HIP(hipMemcpyAsync(dcutlv.p + l, dcut.p, sizeof(long long), hipMemcpyDeviceToDevice, st));
      if (trace) {
        std::vector<int32_t> h;
        GK_TRY(to_host(st, h, cur_part, g.n));
        ls.h_part = fnv_host(h);
      }
    }
    std::vector<long long> hcutlv(nl), hmovedlv(nl), habytes(nl);
    std::vector<int32_t> hroundslv(nl);
    long long hpaths[3] = {0, 0, 0};
    GK_TRY(fetch(ctx, st, {{hcutlv.data(), dcutlv.p, sizeof(long long) * size_t(nl - 1)},
                           {hmovedlv.data(), dmoved.p, sizeof(long long) * size_t(nl - 1)},
                           {habytes.data(), dabytes.p, sizeof(long long) * size_t(nl - 1)},
                           {hroundslv.data(), drounds.p, sizeof(int32_t) * size_t(nl - 1)},
                           {hpaths, dpaths.p, sizeof hpaths}}));
    for (int i = 0; i < 3; ++i) S.round_paths[i] = hpaths[i];
    for (int32_t l = 0; l + 1 < nl; ++l) {
      gk_level_stats& ls = S.levels[l];
      // The round kernel's algorithmic bytes, counted on the device.
      S.stage_bytes[GK_ST_BALANCE] += habytes[l];
      ls.rounds = hroundslv[l];
      ls.moves = hmovedlv[l];
      ls.cut_after = hcutlv[l];
    }
    // Data-dependent bytes of the gain passes from the device work counters.
    std::vector<unsigned long long> wk;
    GK_TRY(to_host(st, wk, work.p, 2 * kGainSlots));
    unsigned long long wv = 0, we = 0;
    for (int i = 0; i < kGainSlots; ++i) {
      wv += wk[2 * i];
      we += wk[2 * i + 1];
    }
    S.stage_bytes[GK_ST_GAIN] += int64_t(28 * wv + 12 * we);
    std::vector<unsigned long long> sbk;
    GK_TRY(to_host(st, sbk, selbytes.p, gkd::kSelectByteSlots));
    for (unsigned long long b : sbk) S.stage_bytes[GK_ST_SELECT] += int64_t(b);
  }
  GK_HIP(hipStreamSynchronize(st));
  const auto t3 = clk::now();

  // ---------------------------------------------------------- final metrics
  {
    Buf<unsigned long long> cut, pw64, cscr;
    GK_TRY(cut.alloc(P, 1));
    GK_TRY(pw64.alloc(P, k));
    GK_TRY(cscr.alloc(P, gkd::cut_scratch(L[0].g.n, k)));
    GK_STAGE(ctx, GK_ST_CUT, 16LL * n0 + 12LL * L[0].g.nnz,
             gkd::launch_cut(L[0].g, k, cur_part, cut.p, pw64.p, cscr.p, st));
    std::vector<unsigned long long> hpw64;
    unsigned long long hcut = 0;
    long long tracked = 0;
    GK_TRY(fetch(ctx, st, {{&hcut, cut.p, sizeof hcut}, {&tracked, dcut.p, sizeof tracked}}));
    GK_TRY(to_host(st, hpw64, pw64.p, k));
    S.cut = static_cast<int64_t>(hcut);
    S.cut_tracked = tracked;
    unsigned long long mx = 0;
    for (auto x : hpw64) mx = std::max(mx, x);
    S.max_part_wgt = static_cast<int64_t>(mx);
    S.imbalance = T > 0 ? double(mx) * k / double(T) : 1.0;
  }
  if (dpart_out != cur_part)
    GK_HIP(hipMemcpyAsync(dpart_out, cur_part, sizeof(int32_t) * n0, hipMemcpyDeviceToDevice, st));
  GK_HIP(hipStreamSynchronize(st));
  S.nlevels = nl;
  S.coarsest_n = nt;
  for (int32_t l = 0; l < nl; ++l) {
    S.levels[l].n = L[l].g.n;
    S.levels[l].nnz = L[l].g.nnz;
  }
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  S.t_coarsen_ms = ms(t0, t1);
  S.t_init_ms = ms(t1, t2);
  S.t_refine_ms = ms(t2, t3);
  GK_TRY(ctx->clock.collect(S));
  S.t_total_ms = ms(t0, clk::now());
  for (int i = 0; i < GK_ST_COUNT; ++i) S.alg_bytes += S.stage_bytes[i];
  if (stats) *stats = S;
  if (S.cut != S.cut_tracked) {
    gkh::set_error("tracked cut " + std::to_string(S.cut_tracked) + " != recomputed cut " +
                   std::to_string(S.cut));
    return GK_ERR_INTERNAL;
  }
  return GK_OK;
}

int check_params(const gk_params* p) {
  if (!p) { gkh::set_error("null params"); return GK_ERR_ARG; }
  if (p->k < 1 || p->k > kMaxK) { gkh::set_error("k must be in [1, 8192]"); return GK_ERR_ARG; }
  if (p->ubfactor < 0) { gkh::set_error("ubfactor < 0"); return GK_ERR_ARG; }
  if (p->max_group < 1 || p->max_group > 64) { gkh::set_error("max_group must be in [1, 64]"); return GK_ERR_ARG; }
  if (p->max_moves < 1 || p->max_moves > 1024) { gkh::set_error("max_moves must be in [1, 1024]"); return GK_ERR_ARG; }
  return GK_OK;
}

struct DeviceGuard {
  explicit DeviceGuard(int d) { (void)hipSetDevice(d); }
};

// ---------------------------------------------------------------- no-alloc cut
// gk_k_cut_noalloc keeps the cut kernel's per-block slots (one 64-bit cut
// share per block, then k 64-bit part weights per block) in `scratch` instead
// of pool memory: 2^18 slots = 2 MiB on every device the library runs on,
// held until the process ends.  Every slot a launch reads was written by the
// same launch, so nothing needs clearing between calls.
constexpr int64_t kCutScratchSlots = int64_t(1) << 18;
static_assert(kCutScratchSlots >= 2 * (int64_t(kMaxK) + 1), "two blocks of slots at the largest k");
static __device__ __align__(256) unsigned long long scratch[kCutScratchSlots];

// On the host, `scratch` names a host-side shadow, not device memory.  Its
// device address on each device is looked up once, with that device current
// (gk_ctx_create), and only that address is ever handed to kernels.
constexpr int kScratchDevices = 256;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_addr[kScratchDevices] = {};

int scratch_setup(int device) {
  if (device < 0 || device >= kScratchDevices) return GK_OK;  // gk_k_cut_noalloc rejects such devices
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_addr[device]) return GK_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    gkh::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return GK_ERR_NODEVICE;
  }
  g_scratch_addr[device] = static_cast<unsigned long long*>(p);
  return GK_OK;
}

unsigned long long* scratch_addr(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_addr[device];
}

}  // namespace

extern "C" {

int gk_ctx_create(int device, gk_ctx** out) {
  if (!out) { gkh::set_error("null out"); return GK_ERR_ARG; }
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    gkh::set_error("no HIP device");
    return GK_ERR_NODEVICE;
  }
  if (device < 0 || device >= ndev) { gkh::set_error("device out of range"); return GK_ERR_ARG; }
  GK_HIP(hipSetDevice(device));
  GK_TRY(scratch_setup(device));
  gk_ctx* c = new gk_ctx;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    gkh::set_error("hipStreamCreate failed");
    return GK_ERR_HIP;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&c->hpin), 4096, hipHostMallocDefault) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    gkh::set_error("hipHostMalloc failed");
    return GK_ERR_NOMEM;
  }
  *out = c;
  return GK_OK;
}

int gk_ctx_set_stage_mask(gk_ctx* c, uint32_t mask) {
  if (!c) return GK_ERR_ARG;
  c->clock.set_mask(mask);
  return GK_OK;
}

int gk_ctx_set_pool_limit(gk_ctx* c, int64_t bytes) {
  if (!c || bytes < 0) { gkh::set_error("bad argument"); return GK_ERR_ARG; }
  c->pool.set_limit(bytes);
  return GK_OK;
}

int gk_ctx_destroy(gk_ctx* c) {
  if (!c) return GK_OK;
  DeviceGuard dg(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->pool.release_all();
  (void)hipHostFree(c->hpin);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return GK_OK;
}

int gk_partition_device(gk_ctx* ctx, const gk_csr* dg, const gk_params* prm, int32_t* dpart,
                        gk_stats* stats, void* stream) {
  if (!ctx || !dg || !dpart) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  GK_TRY(check_params(prm));
  if (dg->n < 0 || dg->nnz < 0 || !dg->xadj || (dg->nnz > 0 && !dg->adjncy)) {
    gkh::set_error("bad graph description");
    return GK_ERR_ARG;
  }
  DeviceGuard guard(ctx->device);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  if (dg->n == 0) {
    if (stats) std::memset(stats, 0, sizeof *stats);
    return GK_OK;
  }
  DGraph g;
  g.n = dg->n;
  g.nnz = dg->nnz;
  g.xadj = dg->xadj;
  g.adj = dg->adjncy;
  g.w = dg->adjwgt;
  g.vw = dg->vwgt;
  return run_vcycle(ctx, st, g, *prm, dpart, stats);
}

int gk_partition(gk_ctx* ctx, const gk_csr* hg, const gk_params* prm, int32_t* part, gk_stats* stats) {
  if (!ctx || !hg || !part) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  GK_TRY(check_params(prm));
  GK_TRY(gk_graph_check(hg, (prm->flags & GK_FLAG_CHECK) != 0));
  DeviceGuard guard(ctx->device);
  hipStream_t st = ctx->stream;
  const int32_t n = hg->n, nnz = hg->nnz;
  if (n == 0) {
    if (stats) std::memset(stats, 0, sizeof *stats);
    return GK_OK;
  }
  Buf<int32_t> x, a, w, v, p;
  GK_TRY(x.alloc(ctx->pool, int64_t(n) + 1));
  GK_TRY(a.alloc(ctx->pool, nnz));
  GK_TRY(p.alloc(ctx->pool, n));
  GK_HIP(hipMemcpyAsync(x.p, hg->xadj, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
  if (nnz) GK_HIP(hipMemcpyAsync(a.p, hg->adjncy, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
  gk_csr d = *hg;
  d.xadj = x.p;
  d.adjncy = a.p;
  if (hg->adjwgt) {
    GK_TRY(w.alloc(ctx->pool, nnz));
    if (nnz) GK_HIP(hipMemcpyAsync(w.p, hg->adjwgt, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
    d.adjwgt = w.p;
  }
  if (hg->vwgt) {
    GK_TRY(v.alloc(ctx->pool, n));
    GK_HIP(hipMemcpyAsync(v.p, hg->vwgt, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    d.vwgt = v.p;
  }
  GK_TRY(gk_partition_device(ctx, &d, prm, p.p, stats, st));
  GK_HIP(hipMemcpyAsync(part, p.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}

int gk_partition_file(gk_ctx* ctx, const char* graph_file, const char* out_file, int32_t k,
                      gk_stats* stats) {
  if (!ctx || !graph_file || !out_file) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  gk_csr g;
  GK_TRY(gk_graph_read(graph_file, &g));
  gk_params p;
  gk_params_default(&p, k);
  std::vector<int32_t> part(g.n > 0 ? g.n : 1);
  int rc = gk_partition(ctx, &g, &p, part.data(), stats);
  if (rc == GK_OK) rc = gk_write_partition((std::string(out_file) + ".out").c_str(), part.data(), g.n);
  gk_graph_free(&g);
  return rc;
}

// ---------------------------------------------------------------- stage hooks
static DGraph dview(const gk_csr* g) {
  DGraph d;
  d.n = g->n;
  d.nnz = g->nnz;
  d.xadj = g->xadj;
  d.adj = g->adjncy;
  d.w = g->adjwgt;
  d.vw = g->vwgt;
  return d;
}

static hipStream_t pick(gk_ctx* ctx, void* s) { return s ? static_cast<hipStream_t>(s) : ctx->stream; }

int gk_k_match(gk_ctx* ctx, const gk_csr* dg, int32_t* cand, void* stream) {
  if (!ctx || !dg || !cand || !dg->adjwgt) { gkh::set_error("null argument (adjwgt required)"); return GK_ERR_ARG; }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  DGraph g = dview(dg);
  GK_TRY(gkd::launch_match(g, gkd::lanes_for_degree(g.n ? double(g.nnz) / g.n : 0.0), 256, cand, st));
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}

int gk_k_group(gk_ctx* ctx, int32_t n, const int32_t* cand, int32_t* label, int32_t* stamp,
               int32_t* sweeps, void* stream) {
  if (!ctx || !cand || !label || !stamp) { gkh::set_error("null argument"); return GK_ERR_ARG; }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  if (n <= 0) { if (sweeps) *sweeps = 0; return GK_OK; }
  Buf<int32_t> dist;
  int32_t rounds = 0;
  GK_TRY(group_stage(ctx, st, n, cand, label, dist, &rounds));
  GK_HIP(hipMemcpyAsync(stamp, dist.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
  GK_HIP(hipStreamSynchronize(st));
  if (sweeps) *sweeps = rounds;
  return GK_OK;
}

int gk_k_cluster(gk_ctx* ctx, int32_t n, const int32_t* label, const int32_t* stamp, const int32_t* vwgt,
                 int32_t max_group, int32_t* order, int32_t* cmap, int32_t* cvwgt, int32_t* nc,
                 void* stream) {
  if (!ctx || !label || !stamp || !vwgt || !order || !cmap || !cvwgt || !nc || max_group < 1) {
    gkh::set_error("bad argument");
    return GK_ERR_ARG;
  }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  if (n <= 0) { *nc = 0; return GK_OK; }
  // The stamp bit width is taken from the data (max stamp).
  std::vector<int32_t> hs;
  GK_TRY(to_host(st, hs, stamp, n));
  int32_t mx = 0;
  for (int32_t x : hs) mx = std::max(mx, x);
  Buf<int32_t> ord, cptr;
  GK_TRY(cluster_stage(ctx, st, n, label, stamp, bits_for(mx), max_group, ord, cmap, cptr, nc));
  GK_TRY(gkd::launch_cvwgt(*nc, cptr.p, ord.p, vwgt, cvwgt, st));
  GK_HIP(hipMemcpyAsync(order, ord.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}

int gk_k_contract(gk_ctx* ctx, const gk_csr* dg, const int32_t* cmap, int32_t nc, int32_t* cxadj,
                  int32_t* cadjncy, int32_t* cadjwgt, int32_t* cnnz, void* stream) {
  if (!ctx || !dg || !cmap || !cxadj || !cadjncy || !cadjwgt || !cnnz || !dg->adjwgt) {
    gkh::set_error("bad argument");
    return GK_ERR_ARG;
  }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  Pool& P = ctx->pool;
  const int32_t n = dg->n;
  // Members of each coarse vertex from a stable sort of (cmap[v], v).
  Buf<int32_t> zero, order, cptr;
  Buf<int32_t> cm_as_label;
  GK_TRY(zero.alloc(P, n));
  gkd::fill<int32_t>(zero.p, 0, n, st);
  int32_t nc2 = 0;
  Buf<int32_t> scratch_cmap;
  GK_TRY(scratch_cmap.alloc(P, n));
  GK_TRY(cluster_stage(ctx, st, n, cmap, zero.p, 1, INT32_MAX, order, scratch_cmap.p, cptr, &nc2));
  // With G = INT32_MAX every label is one chunk, so chunks are the coarse
  // vertices in id order exactly when cmap is onto [0, nc).
  if (nc2 != nc) { gkh::set_error("cmap is not onto [0, nc)"); return GK_ERR_ARG; }
  Level c;
  DGraph g = dview(dg);
  GK_TRY(contract_stage(ctx, st, g, nc, cmap, cptr.p, order.p, c));
  GK_HIP(hipMemcpyAsync(cxadj, c.g.xadj, sizeof(int32_t) * (nc + 1), hipMemcpyDeviceToDevice, st));
  if (c.g.nnz) {
    GK_HIP(hipMemcpyAsync(cadjncy, c.g.adj, sizeof(int32_t) * c.g.nnz, hipMemcpyDeviceToDevice, st));
    GK_HIP(hipMemcpyAsync(cadjwgt, c.g.w, sizeof(int32_t) * c.g.nnz, hipMemcpyDeviceToDevice, st));
  }
  GK_HIP(hipStreamSynchronize(st));
  *cnnz = c.g.nnz;
  return GK_OK;
}

int gk_k_project(gk_ctx* ctx, int32_t n, const int32_t* cmap, const int32_t* cpart, const int32_t* cbnd,
                 int32_t* part, int32_t* bnd, int32_t* gain, int32_t* dest, int32_t* upd, void* stream) {
  if (!ctx) return GK_ERR_ARG;
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  GK_TRY(gkd::launch_project(n, cmap, cpart, cbnd, part, bnd, gain, dest, upd, st));
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}

int gk_k_gain(gk_ctx* ctx, const gk_csr* dg, int32_t k, const int32_t* part, const int32_t* upd,
              int32_t* bnd, int32_t* gain, int32_t* dest, void* stream) {
  if (!ctx || !dg || !dg->adjwgt || k < 1 || k > kMaxK) { gkh::set_error("bad argument"); return GK_ERR_ARG; }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  DGraph g = dview(dg);
  const int lanes = gain_lanes(g);
  Buf<int32_t> wl, heavy, cnt, flags;
  GK_TRY(wl.alloc(ctx->pool, g.n));
  GK_TRY(heavy.alloc(ctx->pool, g.n));
  GK_TRY(cnt.alloc(ctx->pool, 2));
  GK_TRY(flags.alloc(ctx->pool, g.n));
  // The kernels clear the flags they consume: work on a copy.
  if (g.n) GK_HIP(hipMemcpyAsync(flags.p, upd, sizeof(int32_t) * g.n, hipMemcpyDeviceToDevice, st));
  gkd::fill<int32_t>(cnt.p, 0, 2, st);
  GK_TRY(gkd::launch_compact(g.n, flags.p, wl.p, cnt.p, st));
  GK_TRY(gkd::launch_gain_wl(g, k, lanes, wl.p, cnt.p, heavy.p, part, flags.p, bnd, gain, dest, nullptr, g.n, nullptr,
                              st));
  if (gkd::gain_heavy_separate(k))
    GK_TRY(gkd::launch_gain_heavy(g, k, heavy.p, cnt.p, part, flags.p, bnd, gain, dest, nullptr, nullptr, st));
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}

int gk_k_select(gk_ctx* ctx, const gk_csr* dg, const int32_t* part, const int32_t* bnd, const int32_t* gain,
                const int32_t* dest, const int32_t* pw, int64_t cap, int32_t max_moves, int32_t* upd,
                int32_t* moves, int32_t* nsel, void* stream) {
  (void)part;
  if (!ctx || !dg || !dg->vwgt || !moves || !nsel || max_moves < 1) { gkh::set_error("bad argument"); return GK_ERR_ARG; }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  DGraph g = dview(dg);
  Pool& P = ctx->pool;
  Buf<uint64_t> keys, keys2;
  Buf<int32_t> ctl, sscr;
  GK_TRY(keys.alloc(P, g.n));
  GK_TRY(keys2.alloc(P, g.n));
  GK_TRY(ctl.alloc(P, 4));
  GK_TRY(sscr.alloc(P, gkd::sort_scratch_ints(g.n)));
  const int idbits = bits_for(g.n > 0 ? g.n - 1 : 0);
  GK_TRY(gkd::launch_select(g, bnd, gain, dest, pw, cap, idbits, upd, keys.p, ctl.p, nullptr, 1, nullptr, st));
  int32_t hc[2];
  GK_TRY(fetch(ctx, st, {{hc, ctl.p, sizeof hc}}));
  *nsel = hc[0];
  if (hc[0] == 0) return GK_OK;
  bool alt = false;
  if (gkd::radix_sort_pairs<uint64_t>(keys.p, nullptr, keys2.p, nullptr, hc[0], 0, idbits + bits_for(hc[1]),
                                      sscr.p, st, &alt))
    GK_HIP(hipGetLastError());
  std::vector<uint64_t> hk;
  GK_TRY(to_host(st, hk, alt ? keys2.p : keys.p, hc[0]));
  const int32_t B = std::min(hc[0], max_moves);
  std::vector<int32_t> mv(B);
  const uint64_t mask = (1ull << idbits) - 1;
  for (int32_t i = 0; i < B; ++i) mv[i] = int32_t(hk[i] & mask);
  GK_HIP(hipMemcpyAsync(moves, mv.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}

int gk_k_balance(gk_ctx* ctx, const gk_csr* dg, int32_t k, const int32_t* moves, int32_t nmoves,
                 const int32_t* part, const int32_t* dest, int32_t* pw, int64_t cap, int32_t* pos,
                 void* stream) {
  if (!ctx || !dg || !dg->vwgt || !pos || k < 1 || k > kMaxK || nmoves > 1024) {
    gkh::set_error("bad argument");
    return GK_ERR_ARG;
  }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  if (nmoves <= 0) { *pos = -1; return GK_OK; }
  DGraph g = dview(dg);
  Pool& P = ctx->pool;
  // Keys in the given order: the vertex id with idbits = 32 (presorted).
  std::vector<int32_t> hm;
  GK_TRY(to_host(st, hm, moves, nmoves));
  std::vector<uint64_t> hk(nmoves);
  for (int32_t i = 0; i < nmoves; ++i) hk[i] = uint64_t(uint32_t(hm[i]));
  Buf<uint64_t> keys;
  Buf<int32_t> out, ctl, gz;
  Buf<long long> cut;
  GK_TRY(keys.alloc(P, nmoves));
  GK_TRY(out.alloc(P, nmoves));
  GK_TRY(ctl.alloc(P, 4));
  GK_TRY(gz.alloc(P, g.n));
  GK_TRY(cut.alloc(P, 1));
  gkd::fill<int32_t>(gz.p, 0, g.n, st);
  gkd::fill<long long>(cut.p, 0, 1, st);
  GK_HIP(hipMemcpyAsync(keys.p, hk.data(), sizeof(uint64_t) * nmoves, hipMemcpyHostToDevice, st));
  GK_TRY(gkd::launch_balance(g, k, keys.p, nmoves, nmoves, true, 32, part, dest, gz.p, pw, cap, out.p,
                             ctl.p, cut.p, st));
  int32_t hc[3];
  GK_TRY(fetch(ctx, st, {{hc, ctl.p, sizeof hc}}));
  *pos = hc[2];
  return GK_OK;
}

int gk_k_apply(gk_ctx* ctx, const gk_csr* dg, const int32_t* moves, int32_t pos, const int32_t* gain,
               const int32_t* dest, int32_t* part, int32_t* upd, int64_t* gain_sum, void* stream) {
  if (!ctx || !dg || !moves || !gain_sum) { gkh::set_error("bad argument"); return GK_ERR_ARG; }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  DGraph g = dview(dg);
  *gain_sum = 0;
  if (pos < 0) return GK_OK;
  std::vector<int32_t> hm, hg;
  GK_TRY(to_host(st, hm, moves, int64_t(pos) + 1));
  GK_TRY(to_host(st, hg, gain, g.n));
  int64_t s = 0;
  for (int32_t v : hm) s += hg[v];
  Buf<int32_t> ctl, wl, wlc;
  GK_TRY(ctl.alloc(ctx->pool, 4));
  GK_TRY(wl.alloc(ctx->pool, g.n));
  GK_TRY(wlc.alloc(ctx->pool, 2));
  gkd::fill<int32_t>(ctl.p, pos, 4, st);
  gkd::fill<int32_t>(wlc.p, 0, 2, st);
  GK_TRY(gkd::launch_apply(g, pos + 1, moves, ctl.p, dest, part, upd, wl.p, wlc.p, st));
  GK_HIP(hipStreamSynchronize(st));
  *gain_sum = s;
  return GK_OK;
}

int gk_k_round(gk_ctx* ctx, const gk_csr* dg, int32_t k, int64_t cap, int32_t max_moves, int32_t nrounds,
               int32_t* part, int32_t* bnd, int32_t* gain, int32_t* dest, int32_t* pw, int32_t* nmoves,
               int32_t* pos, int64_t* gain_sum, int32_t* order, int64_t* paths, void* stream) {
  if (!ctx || !dg || !dg->adjwgt || !dg->vwgt || !part || !bnd || !gain || !dest || !pw || !gain_sum ||
      k < 1 || k > kMaxK || max_moves < 1 || max_moves > 1024 || nrounds < 0) {
    gkh::set_error("bad argument");
    return GK_ERR_ARG;
  }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  *gain_sum = 0;
  if (paths) paths[0] = paths[1] = paths[2] = 0;
  DGraph g = dview(dg);
  if (g.n <= 0 || nrounds == 0) {
    for (int32_t r = 0; r < nrounds; ++r) {
      if (nmoves) nmoves[r] = 0;
      if (pos) pos[r] = -1;
    }
    return GK_OK;
  }
  Pool& P = ctx->pool;
  const int32_t n = g.n, M = max_moves;
  const int64_t nwords = (int64_t(n) + 63) / 64;
  // The round state of run_vcycle's refinement, for one job.
  Buf<int32_t> upd, wl, heavy, hcnt, list, xlist, wpos, sscr, dthr, active, drounds, lastcnt, dvoff, dpw_off, dkk,
      dorder, dmax;
  Buf<long long> dcap, dcut, dmoved, dpaths;
  Buf<uint64_t> cand, movable, selbits, defbits;
  Buf<uint32_t> hkey;
  GK_TRY(upd.alloc(P, n));
  GK_TRY(wl.alloc(P, n));
  GK_TRY(heavy.alloc(P, n));
  GK_TRY(hcnt.alloc(P, 2));
  GK_TRY(list.alloc(P, n));
  GK_TRY(xlist.alloc(P, n));
  GK_TRY(hkey.alloc(P, n));
  GK_TRY(wpos.alloc(P, nwords + 1));
  GK_TRY(sscr.alloc(P, gkd::select_scratch_ints(n)));
  GK_TRY(cand.alloc(P, nwords + 1));
  GK_TRY(movable.alloc(P, nwords + 1));
  GK_TRY(selbits.alloc(P, nwords + 1));
  GK_TRY(defbits.alloc(P, nwords + 1));
  GK_TRY(dthr.alloc(P, 1));
  GK_TRY(active.alloc(P, 1));
  GK_TRY(drounds.alloc(P, 1));
  GK_TRY(lastcnt.alloc(P, 1));
  GK_TRY(dvoff.alloc(P, 2));
  GK_TRY(dpw_off.alloc(P, 2));
  GK_TRY(dkk.alloc(P, 1));
  GK_TRY(dorder.alloc(P, int64_t(M) + 2));
  GK_TRY(dmax.alloc(P, 1));
  GK_TRY(dcap.alloc(P, 1));
  GK_TRY(dcut.alloc(P, 1));
  GK_TRY(dmoved.alloc(P, 1));
  GK_TRY(dpaths.alloc(P, 3));
  gkd::fill<int32_t>(upd.p, 0, n, st);
  gkd::fill<int32_t>(hcnt.p, 0, 2, st);
  gkd::fill<uint64_t>(selbits.p, 0ull, nwords + 1, st);
  gkd::fill<int32_t>(dthr.p, 1, 1, st);
  gkd::fill<int32_t>(drounds.p, 0, 1, st);
  gkd::fill<int32_t>(lastcnt.p, 1, 1, st);
  gkd::fill<int32_t>(dvoff.p, 0, 1, st);
  gkd::fill<int32_t>(dvoff.p + 1, n, 1, st);
  gkd::fill<int32_t>(dpw_off.p, 0, 1, st);
  gkd::fill<int32_t>(dpw_off.p + 1, k, 1, st);
  gkd::fill<int32_t>(dkk.p, k, 1, st);
  gkd::fill<long long>(dcap.p, static_cast<long long>(cap), 1, st);
  gkd::fill<long long>(dcut.p, 0, 1, st);
  gkd::fill<long long>(dmoved.p, 0, 1, st);
  gkd::fill<long long>(dpaths.p, 0, 3, st);
  GK_TRY(gkd::launch_cand_init(n, bnd, gain, cand.p, st));
  GK_TRY(gkd::launch_max_deg(g, dmax.p, st));
  int32_t maxdeg = 0;
  GK_TRY(fetch(ctx, st, {{&maxdeg, dmax.p, sizeof maxdeg}}));
  const int lanes = gain_lanes(g);
  const gkd::SelectBufs sb{movable.p, selbits.p, defbits.p, wpos.p, list.p, sscr.p, nullptr};
  for (int32_t r = 0; r < nrounds; ++r) {
    // B = 0, pos = -1 unless the round orders any move.
    gkd::fill<int32_t>(dorder.p, 0, 1, st);
    gkd::fill<int32_t>(dorder.p + 1, -1, 1, st);
    GK_TRY(gkd::launch_select_list(g, nullptr, gain, dest, pw, dpw_off.p, dcap.p, dthr.p, cand.p, sb, hcnt.p,
                                   active.p, st));
    GK_TRY(gkd::launch_round_b(1, k, dvoff.p, M, dkk.p, dcap.p, dpw_off.p, g, part, dest, gain, pw, dcut.p,
                               drounds.p, dmoved.p, sb, xlist.p, hkey.p, upd.p, wl.p, hcnt.p, lastcnt.p, active.p,
                               cand.p, dthr.p, dpaths.p, dorder.p, nullptr, st));
    GK_TRY(gkd::launch_gain_wl(g, k, lanes, wl.p, hcnt.p, heavy.p, part, upd.p, bnd, gain, dest, nullptr, n,
                                cand.p, st));
    if (maxdeg > lanes && gkd::gain_heavy_separate(k))
      GK_TRY(gkd::launch_gain_heavy(g, k, heavy.p, hcnt.p, part, upd.p, bnd, gain, dest, nullptr, cand.p, st));
    int32_t hdr[2] = {0, -1};
    GK_TRY(fetch(ctx, st, {{hdr, dorder.p, sizeof hdr}}));
    if (nmoves) nmoves[r] = hdr[0];
    if (pos) pos[r] = hdr[1];
    if (order && hdr[0] > 0)
      GK_HIP(hipMemcpyAsync(order, dorder.p + 2, sizeof(int32_t) * size_t(hdr[0]), hipMemcpyDeviceToDevice, st));
  }
  long long hcut = 0, hp[3] = {0, 0, 0};
  GK_TRY(fetch(ctx, st, {{&hcut, dcut.p, sizeof hcut}, {hp, dpaths.p, sizeof hp}}));
  *gain_sum = -hcut;
  if (paths)
    for (int i = 0; i < 3; ++i) paths[i] = hp[i];
  return GK_OK;
}

int gk_k_cut(gk_ctx* ctx, const gk_csr* dg, int32_t k, const int32_t* part, int64_t* cut, int64_t* pw,
             void* stream) {
  if (!ctx || !dg || !dg->adjwgt || !dg->vwgt || !cut || k < 1) { gkh::set_error("bad argument"); return GK_ERR_ARG; }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  DGraph g = dview(dg);
  Buf<unsigned long long> c, p, cscr;
  GK_TRY(c.alloc(ctx->pool, 1));
  GK_TRY(p.alloc(ctx->pool, k));
  GK_TRY(cscr.alloc(ctx->pool, gkd::cut_scratch(g.n, k)));
  GK_TRY(gkd::launch_cut(g, k, part, c.p, p.p, cscr.p, st));
  unsigned long long hc = 0;
  GK_TRY(fetch(ctx, st, {{&hc, c.p, sizeof hc}}));
  *cut = static_cast<int64_t>(hc);
  if (pw) {
    std::vector<unsigned long long> hp;
    GK_TRY(to_host(st, hp, p.p, k));
    for (int32_t q = 0; q < k; ++q) pw[q] = static_cast<int64_t>(hp[q]);
  }
  return GK_OK;
}

int gk_k_cut_noalloc(gk_ctx* ctx, const gk_csr* dg, int32_t k, const int32_t* part, int64_t* dev_cut,
                     int64_t* dev_pw, void* stream) {
  if (!ctx || !dg || !dg->adjwgt || !dg->vwgt || !dev_cut || !dev_pw || k < 1 || dg->n < 0 ||
      (dg->n > 0 && (!part || !dg->xadj || !dg->adjncy))) {
    gkh::set_error("bad argument");
    return GK_ERR_ARG;
  }
  if (k > kMaxK) {
    gkh::set_error("k = " + std::to_string(k) + " exceeds " + std::to_string(kMaxK) +
                   ", the largest k of the no-allocation cut (its scratch holds 2^18 slots)");
    return GK_ERR_LIMIT;
  }
  unsigned long long* scr = scratch_addr(ctx->device);
  if (!scr) { gkh::set_error("no scratch address for this device"); return GK_ERR_INTERNAL; }
  DeviceGuard guard(ctx->device);
  hipStream_t st = pick(ctx, stream);
  const int rc = gkd::launch_cut(dview(dg), k, part, reinterpret_cast<unsigned long long*>(dev_cut),
                                 reinterpret_cast<unsigned long long*>(dev_pw), scr, st, kCutScratchSlots);
  if (rc != GK_OK) gkh::set_error("cut launch failed");
  return rc;
}

}  // extern "C"
