// Largest balanced prefix of a move sequence, computed by one workgroup
// (find_mv_wgt_sequence + scan + find_balance_sequence, gkway/uncoarsen.hpp:
// 239-288, 37-80, plus the pw/cut update of apply_sequence_move, :305-316).
//
// Instead of the reference's dense k x B matrix of weight deltas, each wave
// owns a set of parts and runs a wave64 prefix scan of that part's delta
// along the move sequence; positions where the part's weight would rise above
// the cap are OR-ed into a bitmap, and the answer is the highest clear bit.
// Shared by the single-graph kernel (refine.hip) and the batched kernel with
// one workgroup per graph (batch.hip).
#pragma once
#include "dev_common.hpp"

namespace gkd {

constexpr int kBalThreads = 1024;
constexpr int kBalMax = 1024;

struct BalanceLds {
  uint64_t sk[kBalMax];
  int32_t sv[kBalMax], ssrc[kBalMax], sdst[kBalMax], sw[kBalMax], sg[kBalMax];
  unsigned long long bad[kBalMax / 64];
  int32_t spos;
  long long sgain[kBalThreads / 64];
};

// Ascending bitonic sort of L.sk[0, P), P = next_pow2(count) (<= kBalMax);
// slots in [count, P) must hold the maximum key.  All threads of the block
// call it.
__device__ inline void lds_bitonic(BalanceLds& L, int32_t count) {
  const int t = threadIdx.x;
  int P = 2;
  while (P < count) P <<= 1;
  for (int k2 = 2; k2 <= P; k2 <<= 1) {
    for (int j2 = k2 >> 1; j2 > 0; j2 >>= 1) {
      const int ix = t ^ j2;
      if (t < P && ix > t) {
        const uint64_t a = L.sk[t], b = L.sk[ix];
        const bool up = (t & k2) == 0;
        if (up ? a > b : a < b) { L.sk[t] = b; L.sk[ix] = a; }
      }
      __syncthreads();
    }
  }
}

// Largest balanced prefix of the B sorted moves in L.sk (vertex id in the
// bits of idmask).  dpw: k ints of LDS.  Fills L.sv/ssrc/sdst/sw/sg, writes
// the prefix end to *pos_out (-1 only for B == 0) and, when moves != null,
// the moved vertices to moves[0..pos]; advances pw and subtracts the
// prefix's gains from *cut.  Returns the prefix end to every thread.
__device__ inline int balance_prefix(BalanceLds& L, int32_t* dpw, int32_t k, int32_t B, uint64_t idmask,
                                     const int32_t* __restrict__ part, const int32_t* __restrict__ dest,
                                     const int32_t* __restrict__ gain, const int32_t* __restrict__ vw,
                                     int32_t* __restrict__ pw, long long cap, int32_t* __restrict__ moves,
                                     int32_t* __restrict__ pos_out, long long* __restrict__ cut) {
  const int t = threadIdx.x;
  if (t < kBalMax / 64) L.bad[t] = 0;
  for (int32_t q = t; q < k; q += kBalThreads) dpw[q] = 0;
  if (t < B) {
    const int32_t v = int32_t(L.sk[t] & idmask);
    L.sv[t] = v;
    L.ssrc[t] = part[v];
    L.sdst[t] = dest[v];
    L.sw[t] = vw[v];
    L.sg[t] = gain[v];
  }
  __syncthreads();
  const int lane = lane_id();
  const int wave = t >> 6;
  const int nchunk = (B + 63) >> 6;
  for (int32_t q = wave; q < k; q += kBalThreads / 64) {
    const long long pq = pw[q];
    long long carry = 0;
    for (int c = 0; c < nchunk; ++c) {
      const int j = c * 64 + lane;
      int x = 0;
      if (j < B) {
        if (L.sdst[j] == q) x += L.sw[j];
        if (L.ssrc[j] == q) x -= L.sw[j];
      }
      const long long incl = carry + wave_incl_scan(x);
      const bool b = j < B && incl > 0 && pq + incl > cap;
      const uint64_t m = ballot(b);
      if (m != 0 && lane == 0) atomicOr(&L.bad[c], (unsigned long long)m);
      carry = __shfl(incl, 63, 64);
    }
  }
  __syncthreads();
  if (t == 0) {
    int pos = -1;
    for (int i = B - 1; i >= 0; --i) {
      if (((L.bad[i >> 6] >> (i & 63)) & 1ull) == 0) { pos = i; break; }
    }
    L.spos = pos;
  }
  __syncthreads();
  const int pos = L.spos;
  long long gsum = 0;
  if (t <= pos) {
    atomicAdd(&dpw[L.ssrc[t]], -L.sw[t]);
    atomicAdd(&dpw[L.sdst[t]], L.sw[t]);
    if (moves) moves[t] = L.sv[t];
    gsum = L.sg[t];
  }
  gsum = wave_sum64(gsum);
  if (lane == 0) L.sgain[wave] = gsum;
  __syncthreads();
  for (int32_t q = t; q < k; q += kBalThreads) pw[q] += dpw[q];
  if (t == 0) {
    long long s = 0;
    for (int w2 = 0; w2 < kBalThreads / 64; ++w2) s += L.sgain[w2];
    *cut -= s;
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
