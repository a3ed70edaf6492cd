// Largest balanced prefix of a move sequence, computed by one workgroup
// (find_mv_wgt_sequence + scan + find_balance_sequence, gkway/uncoarsen.hpp:
// 239-288, 37-80, plus the pw/cut update of apply_sequence_move, :305-316).
//
// Instead of the reference's dense k x B matrix of weight deltas: a part can
// only rise above the cap inside the sequence if its weight plus everything
// the sequence moves into it exceeds the cap.  Only such parts ("risky"; few
// once the partition is near balance) are examined.  Each move gives events
// (source part, position, -w) and (destination part, position, +w); the
// events of risky parts are sorted by (part, position) in LDS and a segmented
// prefix sum gives each part's weight change after each of its events.  A
// part stays above the cap from such an event up to its next event; those
// position intervals are marked in a difference array, and the answer is the
// last position no interval covers.  The cost is O(B log B), independent of
// k apart from two passes over the part weights.  Shared by the single-graph
// kernel (refine.hip) and the batched kernel with one workgroup per graph
// (batch.hip).
#pragma once
#include "dev_common.hpp"

namespace gkd {

constexpr int kBalThreads = 1024;
constexpr int kBalMax = 1024;

constexpr int kBalMaxK = 8192;  // parts

// sort_events_by_part ranks digits per wave with a 16 x 16 count table
// (16 waves x 16 four-bit digits = 256 = 64 lanes x 4 in wave 0's scan);
// event keys are part << 11 | position << 1 | kind, so positions need 10
// bits and parts the remaining 21.
static_assert(kBalThreads == 1024, "event sort assumes 16 waves of 64 lanes");
static_assert(kBalMax <= 1024, "event key holds positions in 10 bits");
static_assert(2 * kBalMax <= 2 * kBalThreads, "two events per thread");
static_assert(kBalMaxK <= (1 << 21), "event key holds parts in 21 bits");

struct BalanceLds {
  uint64_t sk[kBalMax];  // move keys; then the 2 * kBalMax event keys (uint32)
  int32_t sv[kBalMax], sw[kBalMax];
  int32_t cover[kBalMax + 1];
  unsigned long long risk[kBalMaxK / 64];
  int32_t ws[kBalThreads / 64 + 1];
  int32_t wmax[kBalThreads / 64];
  long long sgain[kBalThreads / 64];
  uint32_t ev2[2 * kBalMax];  // second buffer of the event sort
  int32_t rcnt[256];          // event sort: element counts per (digit, wave)
};

// Exclusive scan of one int per thread over a kBalThreads block; *tot gets
// the block total.  ws: >= kBalThreads / 64 ints of LDS.  One barrier: each
// wave adds up the wave totals below it itself.  ws is read after that
// barrier, so the next scan on the same ws must follow another barrier.
__device__ __forceinline__ int block_excl_scan_1024(int x, int* tot, int32_t* ws) {
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const int incl = wave_incl_scan(x);
  if (lane == 63) ws[wave] = incl;
  __syncthreads();
  int below = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kBalThreads / 64; ++w) {
    const int q = ws[w];
    below += w < wave ? q : 0;
    all += q;
  }
  *tot = all;
  return incl - x + below;
}

// Ascending bitonic sort of a[0, P) in LDS, P a power of two in
// [128, 2 * kBalThreads]; all kBalThreads threads call it, after a barrier
// that follows the writes of a[].  Wave w < P / 128 keeps a[128w, 128w + 128)
// in registers (lane l: elements 128w + l and 128w + l + 64): strides up to 64
// run in registers and wave shuffles, larger strides through LDS, so only the
// stages with stride >= 128 need a barrier.  Ends with a barrier.
template <typename T>
__device__ inline void bitonic_lds(T* a, int P) {
  const int t = threadIdx.x;
  const int lane = lane_id();
  const int wave = t >> 6;
  const bool own = wave < (P >> 7);
  const int i0 = (wave << 7) + lane;
  const int i1 = i0 + 64;
  T x0 = 0, x1 = 0;
  if (own) {
    x0 = a[i0];
    x1 = a[i1];
  }
  // Strides JS, JS / 2, ..., 1 of merge size k2 on the wave's registers.
  auto inwave = [&](auto jsc, int k2) {
    if (!own) return;
    for_pow2_down<decltype(jsc)::value>([&](auto jc) {
      constexpr int J = decltype(jc)::value;
      if constexpr (J == 64) {
        // i0 and i1 = i0 + 64 are partners; k2 >= 128 gives both the same direction
        const bool up = (i0 & k2) == 0;
        const T lo = x0 < x1 ? x0 : x1;
        const T hi = x0 < x1 ? x1 : x0;
        x0 = up ? lo : hi;
        x1 = up ? hi : lo;
      } else {
        const T y0 = lane_xor<J>(x0);
        const T y1 = lane_xor<J>(x1);
        const bool lower = (lane & J) == 0;
        // the lower element of an ascending pair keeps the minimum
        const bool mn0 = lower == ((i0 & k2) == 0);
        const bool mn1 = lower == ((i1 & k2) == 0);
        x0 = mn0 ? (y0 < x0 ? y0 : x0) : (y0 > x0 ? y0 : x0);
        x1 = mn1 ? (y1 < x1 ? y1 : x1) : (y1 > x1 ? y1 : x1);
      }
    });
  };
  inwave(std::integral_constant<int, 1>{}, 2);
  inwave(std::integral_constant<int, 2>{}, 4);
  inwave(std::integral_constant<int, 4>{}, 8);
  inwave(std::integral_constant<int, 8>{}, 16);
  inwave(std::integral_constant<int, 16>{}, 32);
  inwave(std::integral_constant<int, 32>{}, 64);
  inwave(std::integral_constant<int, 64>{}, 128);
  for (int k2 = 256; k2 <= P; k2 <<= 1) {
    if (own) {
      a[i0] = x0;
      a[i1] = x1;
    }
    __syncthreads();
    for (int j2 = k2 >> 1; j2 >= 128; j2 >>= 1) {
      for (int p = t; p < (P >> 1); p += kBalThreads) {
        const int lo = ((p & ~(j2 - 1)) << 1) | (p & (j2 - 1));
        const int hi = lo + j2;
        const T u = a[lo], v = a[hi];
        const bool up = (lo & k2) == 0;
        if (up ? u > v : u < v) {
          a[lo] = v;
          a[hi] = u;
        }
      }
      __syncthreads();
    }
    if (own) {
      x0 = a[i0];
      x1 = a[i1];
    }
    inwave(std::integral_constant<int, 64>{}, k2);
  }
  if (own) {
    a[i0] = x0;
    a[i1] = x1;
  }
  __syncthreads();
}

// Ascending sort of L.sk[0, count) (count <= kBalMax); slots [count, kBalMax)
// must hold the maximum key.  All threads of the block call it, after a
// barrier that follows the writes of L.sk.
__device__ inline void lds_bitonic(BalanceLds& L, int32_t count) {
  int P = 128;
  while (P < count) P <<= 1;
  bitonic_lds(L.sk, P);
}

// Stable sort of the event keys a[0, n) (n <= 2 * kBalThreads) by their part
// field, bits [11, 11 + pbits): LSD passes over 4-bit digits.  Thread t moves
// elements 2t and 2t + 1; an element's rank among the equal digits of its
// wave comes from ballot match masks, and the per-(digit, wave) counts in C
// (256 ints), scanned in digit-major order, give the offsets.  The events
// arrive in (position, kind) order, so the result is ordered by the whole
// key.  Returns the buffer holding the result (a or b).  All threads call it,
// after a barrier that follows the writes of a[]; it ends with a barrier.
__device__ inline uint32_t* sort_events_by_part(uint32_t* a, uint32_t* b, int n, int pbits, int32_t* C) {
  const int t = threadIdx.x;
  const int lane = lane_id();
  const int wave = t >> 6;
  const uint64_t lt = lanemask_lt(), le = lt | (1ull << lane);
  for (int s = 11; s < 11 + pbits; s += 4) {
    const bool v0 = 2 * t < n, v1 = 2 * t + 1 < n;
    const uint32_t k0 = v0 ? a[2 * t] : 0u, k1 = v1 ? a[2 * t + 1] : 0u;
    const int d0 = int((k0 >> s) & 15u), d1 = int((k1 >> s) & 15u);
    uint64_t m0[4], m1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      m0[i] = ballot(v0 && ((d0 >> i) & 1));
      m1[i] = ballot(v1 && ((d1 >> i) & 1));
    }
    const uint64_t vm0 = ballot(v0), vm1 = ballot(v1);
    // lanes whose first (second) element has digit x
    auto same0 = [&](int x) {
      uint64_t m = vm0;
#pragma unroll
      for (int i = 0; i < 4; ++i) m &= ((x >> i) & 1) ? m0[i] : ~m0[i];
      return m;
    };
    auto same1 = [&](int x) {
      uint64_t m = vm1;
#pragma unroll
      for (int i = 0; i < 4; ++i) m &= ((x >> i) & 1) ? m1[i] : ~m1[i];
      return m;
    };
    // wave order: lane by lane, first element before second
    const int r0 = popc64(same0(d0) & lt) + popc64(same1(d0) & lt);
    const int r1 = popc64(same0(d1) & le) + popc64(same1(d1) & lt);
    if (lane < 16) C[lane * 16 + wave] = popc64(same0(lane)) + popc64(same1(lane));
    __syncthreads();
    if (wave == 0) {
      int c[4], sum = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        c[i] = C[4 * lane + i];
        sum += c[i];
      }
      int run = wave_incl_scan(sum) - sum;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        C[4 * lane + i] = run;
        run += c[i];
      }
    }
    __syncthreads();
    if (v0) b[C[d0 * 16 + wave] + r0] = k0;
    if (v1) b[C[d1 * 16 + wave] + r1] = k1;
    __syncthreads();
    uint32_t* x = a;
    a = b;
    b = x;
  }
  return a;
}

// Optional inputs and outputs of balance_prefix.
struct BalanceOpt {
  const int32_t* spw = nullptr;  // the part weights already in LDS: read instead of pw
  int32_t* mdst = nullptr;       // when set, each position's destination (LDS, B entries)
  bool gain_in_key = false;      // key bits 32..62 hold 0x7fffffff - gain: gain[] is not read
  bool cut_given = false;        // thread 0's cut_now holds *cut: *cut is written, not read
  long long cut_now = 0;
};

// Largest balanced prefix of the B sorted moves in L.sk (vertex id in the
// bits of idmask): the last position pos such that, after moves 0..pos, no
// part whose weight grew is above cap.  dpw: k ints of LDS (k parts, k <=
// kBalMaxK).  Fills L.sv, writes pos to *pos_out (-1 when no position
// qualifies) and, when moves != null, the moved vertices to moves[0..pos];
// advances pw and subtracts the prefix's gains from *cut.  Returns pos to
// every thread.
__device__ inline int balance_prefix(BalanceLds& L, int32_t* dpw, int32_t k, int32_t B, uint64_t idmask,
                                     const int32_t* __restrict__ part, const int32_t* __restrict__ dest,
                                     const int32_t* __restrict__ gain, const int32_t* __restrict__ vw,
                                     int32_t* __restrict__ pw, long long cap, int32_t* __restrict__ moves,
                                     int32_t* __restrict__ pos_out, long long* __restrict__ cut,
                                     const BalanceOpt& opt = BalanceOpt{}) {
  const int t = threadIdx.x;
  const int lane = lane_id();
  auto pw_of = [&](int32_t q) -> int32_t { return opt.spw ? opt.spw[q] : pw[q]; };
  const int wave = t >> 6;
  // 1. the moves, and each part's inflow over the whole sequence
  for (int32_t q = t; q < k; q += kBalThreads) dpw[q] = 0;
  if (t == 0) L.cover[kBalMax] = 0;
  L.cover[t] = 0;
  int32_t g = 0, src = 0, dst = 0, w = 0;
  if (t < B) {
    const uint64_t key = L.sk[t];
    const int32_t v = int32_t(key & idmask);
    L.sv[t] = v;
    w = vw[v];
    L.sw[t] = w;
    g = opt.gain_in_key ? int32_t(0x7fffffffu - uint32_t(key >> 32)) : gain[v];
    src = part[v];
    dst = dest[v];
    if (opt.mdst) opt.mdst[t] = dst;
  }
  __syncthreads();
  if (t < B) atomicAdd(&dpw[dst], w);
  __syncthreads();
  // 2. risky parts: pw + inflow > cap (the only ones that can be above the
  // cap at some position with a positive change); a wave writes the bitmask
  // word of its 64 parts (no shared-word atomics)
  for (int32_t q0 = wave * 64; q0 < k; q0 += kBalThreads) {
    const int32_t q = q0 + lane;
    bool r = false;
    if (q < k) {
      const int32_t in = dpw[q];
      r = in > 0 && (long long)pw_of(q) + in > cap;
      dpw[q] = 0;
    }
    const uint64_t m = ballot(r);
    if (lane == 0) L.risk[q0 >> 6] = m;
  }
  __syncthreads();
  // 3. event keys of the risky parts (part << 11 | position << 1 | kind),
  // kind 0 = leave, 1 = enter; slots from a block scan, so in (position,
  // kind) order; then a stable sort by part
  uint32_t* ev = reinterpret_cast<uint32_t*>(L.sk);
  bool rs = false, rd = false;
  if (t < B) {
    rs = (L.risk[src >> 6] >> (src & 63)) & 1ull;
    rd = (L.risk[dst >> 6] >> (dst & 63)) & 1ull;
  }
  int nev;
  int slot = block_excl_scan_1024(int(rs) + int(rd), &nev, L.ws);
  if (rs) ev[slot++] = (uint32_t(src) << 11) | (uint32_t(t) << 1);
  if (rd) ev[slot] = (uint32_t(dst) << 11) | (uint32_t(t) << 1) | 1u;
  __syncthreads();
  if (nev > 0) ev = sort_events_by_part(ev, L.ev2, nev, k > 1 ? 32 - __clz(k - 1) : 0, L.rcnt);
  // 4. thread t holds sorted events 2t and 2t + 1; per-part inclusive sums
  const bool ha = 2 * t < nev, hb = 2 * t + 1 < nev;
  int32_t qa = 0, qb = -1, ja = 0, jb = 0, xa = 0, xb = 0;
  if (ha) {
    const uint32_t ea = ev[2 * t];
    qa = int32_t(ea >> 11);
    ja = int32_t((ea >> 1) & 1023u);
    xa = (ea & 1u) ? L.sw[ja] : -L.sw[ja];
  }
  if (hb) {
    const uint32_t eb = ev[2 * t + 1];
    qb = int32_t(eb >> 11);
    jb = int32_t((eb >> 1) & 1023u);
    xb = (eb & 1u) ? L.sw[jb] : -L.sw[jb];
  }
  int tot;
  const int ex = block_excl_scan_1024(xa + xb, &tot, L.ws);
  const int pa = ex + xa, pb = pa + xb;
  const bool heada = ha && (t == 0 || int32_t(ev[2 * t - 1] >> 11) != qa);
  const bool headb = hb && qb != qa;
  if (heada) dpw[qa] = ex;  // prefix before the part's first event
  if (headb) dpw[qb] = pa;
  __syncthreads();
  // 5. a risky part is above the cap from its event at j up to its next event
  if (ha) {
    const int32_t sa = pa - dpw[qa];
    const int32_t na = (hb && qb == qa) ? jb : B;
    if (ja < na && sa > 0 && (long long)pw_of(qa) + sa > cap) {
      atomicAdd(&L.cover[ja], 1);
      atomicAdd(&L.cover[na], -1);
    }
  }
  if (hb) {
    const int32_t sb = pb - dpw[qb];
    int32_t nb = B;
    if (2 * t + 2 < nev) {
      const uint32_t ec = ev[2 * t + 2];
      if (int32_t(ec >> 11) == qb) nb = int32_t((ec >> 1) & 1023u);
    }
    if (jb < nb && sb > 0 && (long long)pw_of(qb) + sb > cap) {
      atomicAdd(&L.cover[jb], 1);
      atomicAdd(&L.cover[nb], -1);
    }
  }
  __syncthreads();
  if (heada) dpw[qa] = 0;
  if (headb) dpw[qb] = 0;
  // 6. pos = last position no interval covers
  const int c = t < B ? L.cover[t] : 0;
  const int cov = block_excl_scan_1024(c, &tot, L.ws) + c;
  const int cand = wave_max((t < B && cov == 0) ? t : -1);
  if (lane == 0) L.wmax[wave] = cand;
  __syncthreads();
  int pos = -1;
#pragma unroll
  for (int w2 = 0; w2 < kBalThreads / 64; ++w2) pos = max(pos, L.wmax[w2]);
  // 7. apply the prefix to the part weights, moves and cut
  long long gsum = 0;
  if (t <= pos) {
    atomicAdd(&dpw[src], -w);
    atomicAdd(&dpw[dst], w);
    if (moves) moves[t] = L.sv[t];
    gsum = g;
  }
  gsum = wave_sum64(gsum);
  if (lane == 0) L.sgain[wave] = gsum;
  __syncthreads();
  for (int32_t q = t; q < k; q += kBalThreads) {
    const int32_t d = dpw[q];
    if (d != 0) pw[q] = pw_of(q) + d;
  }
  if (t == 0) {
    long long s = 0;
    for (int w2 = 0; w2 < kBalThreads / 64; ++w2) s += L.sgain[w2];
    if (opt.cut_given) *cut = opt.cut_now - s;
    else *cut -= s;
    if (pos_out) *pos_out = pos;
  }
  return pos;
}

// keys: the move keys (vertex id in the low bits, idmask); when !presorted
// the first `count` (<= kBalMax) keys are sorted here, else the first
// min(count, M) are taken in order.  Then balance_prefix.
__device__ inline void balance_block(BalanceLds& L, int32_t* dpw, int32_t k, const uint64_t* __restrict__ keys,
                                     int32_t count, int32_t M, int presorted, uint64_t idmask,
                                     const int32_t* __restrict__ part, const int32_t* __restrict__ dest,
                                     const int32_t* __restrict__ gain, const int32_t* __restrict__ vw,
                                     int32_t* __restrict__ pw, long long cap, int32_t* __restrict__ moves,
                                     int32_t* __restrict__ pos_out, long long* __restrict__ cut) {
  const int t = threadIdx.x;
  const int B = count < M ? count : M;
  L.sk[t] = (t < (presorted ? B : count)) ? keys[t] : ~0ull;
  __syncthreads();
  if (!presorted) lds_bitonic(L, count);
  balance_prefix(L, dpw, k, B, idmask, part, dest, gain, vw, pw, cap, moves, pos_out, cut);
}

}  // namespace gkd
