// Device-wide primitives written for gfx950 (wave64): prefix scan and a
// stable LSD radix sort.  They replace the moderngpu calls of the reference
// (mgpu::scan at gkway/coarsen.hpp:325,342,412,476; mgpu::mergesort at
// gkway/coarsen.hpp:415 and gkway/uncoarsen.hpp:472; mgpu::segmented_sort at
// gkway/coarsen.hpp:338).  All arithmetic is integer; results do not depend on
// the launch schedule.
#pragma once
#include "dev_common.hpp"

namespace gkd {

// ------------------------------------------------------------------ scan
// Reduce-then-scan over tiles of kScanTile items (256 threads x 8 items).
constexpr int kScanThreads = 256;
constexpr int kScanIPT = 8;
constexpr int kScanTile = kScanThreads * kScanIPT;

// Exclusive block scan of one value per thread (256 threads); returns the
// exclusive prefix and writes the block total to *total.
__device__ __forceinline__ int block_excl_scan_256(int x, int* total, int* lds /*>= 5 ints*/) {
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  int incl = wave_incl_scan(x);
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) { int t = lds[w]; lds[w] = run; run += t; }
    lds[4] = run;
  }
  __syncthreads();
  int r = incl - x + lds[wave];
  *total = lds[4];
  __syncthreads();
  return r;
}

template <class Load>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(int64_t n, Load load, int* partial) {
  __shared__ int lds[8];
  const int64_t base = int64_t(blockIdx.x) * kScanTile;
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanIPT; ++j) {
    int64_t i = base + int64_t(j) * kScanThreads + threadIdx.x;
    if (i < n) s += load(i);
  }
  int tot;
  (void)block_excl_scan_256(s, &tot, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Exclusive scan of the tile; `carry` holds the per-block offsets (or null
// for a single block).  Store(i, exclusive, value).
template <class Load, class Store>
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(int64_t n, Load load, Store store,
                                                             const int* carry, int* total_out) {
  __shared__ int lds[8];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanIPT;
  int v[kScanIPT];
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanIPT; ++j) {
    int64_t i = base + j;
    v[j] = (i < n) ? load(i) : 0;
    s += v[j];
  }
  int tot;
  int pre = block_excl_scan_256(s, &tot, lds);
  int off = carry ? carry[blockIdx.x] : 0;
  int run = off + pre;
#pragma unroll
  for (int j = 0; j < kScanIPT; ++j) {
    int64_t i = base + j;
    if (i < n) store(i, run, v[j]);
    run += v[j];
  }
  if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) *total_out = run;
}

struct ArrayLoad {
  const int* a;
  __device__ int operator()(int64_t i) const { return a[i]; }
};
struct ArrayStoreExcl {
  int* out;
  __device__ void operator()(int64_t i, int excl, int) const { out[i] = excl; }
};

// Scratch ints needed by scan() for n items.
inline int64_t scan_scratch_ints(int64_t n) {
  int64_t s = 0;
  while (n > kScanTile) {
    n = (n + kScanTile - 1) / kScanTile;
    s += n + 64;
  }
  return s + 64;
}

// Exclusive scan of the n values load(i) (int32 wrap-around arithmetic);
// store(i, exclusive, value) gets every result, *total_dev the sum.
template <class Load, class Store>
int scan(int64_t n, Load load, Store store, int* total_dev, int* scratch, hipStream_t st) {
  if (n <= 0) {
    if (total_dev) fill<int>(total_dev, 0, 1, st);
    return hipGetLastError() == hipSuccess ? 0 : -5;
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 1) {
    k_scan_apply<<<1, kScanThreads, 0, st>>>(n, load, store, nullptr, total_dev);
    return hipGetLastError() == hipSuccess ? 0 : -5;
  }
  int* partial = scratch;
  k_scan_reduce<<<static_cast<unsigned>(nb), kScanThreads, 0, st>>>(n, load, partial);
  // Scan the partials in place (recursively), then apply.
  int rc = scan(nb, ArrayLoad{partial}, ArrayStoreExcl{partial}, nullptr, scratch + nb + 64, st);
  if (rc) return rc;
  k_scan_apply<<<static_cast<unsigned>(nb), kScanThreads, 0, st>>>(n, load, store, partial, total_dev);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

// ------------------------------------------------------------------ radix sort
constexpr int kSortThreads = 256;
constexpr int kSortIPT = 16;
constexpr int kSortTile = kSortThreads * kSortIPT;  // 4096
constexpr int kSortWaves = kSortThreads / 64;

// Scratch ints needed by radix_sort_pairs() for n items.
inline int64_t sort_scratch_ints(int64_t n) {
  int64_t nb = (n + kSortTile - 1) / kSortTile;
  return 256 * nb + 64 + scan_scratch_ints(256 * nb);
}

template <typename K>
__global__ void __launch_bounds__(kSortThreads) k_radix_hist(const K* keys, int64_t n, int shift,
                                                              int nbits, int* hist, int nblocks) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kSortTile;
  const K mask = (K(1) << nbits) - 1;
#pragma unroll 4
  for (int j = 0; j < kSortIPT; ++j) {
    int64_t i = base + int64_t(j) * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[int((keys[i] >> shift) & mask)], 1);
  }
  __syncthreads();
  hist[int64_t(threadIdx.x) * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: items of a tile are ranked in index order.  Layout: wave w
// owns the contiguous sub-tile [w*64*IPT, (w+1)*64*IPT); slot j of that wave
// holds items w*64*IPT + j*64 + lane.  Per-wave digit counters in LDS give
// the rank inside the wave; a prefix over waves and the global offsets give
// the final position.
template <typename K, bool kHasVals>
__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(const K* keys, const int* vals, K* okeys,
                                                                 int* ovals, int64_t n, int shift, int nbits,
                                                                 const int* offs, int nblocks) {
  __shared__ int cnt[kSortWaves][256];
  __shared__ int goff[256];
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  for (int d = threadIdx.x; d < 256; d += kSortThreads) {
    goff[d] = (d < (1 << nbits)) ? offs[int64_t(d) * nblocks + blockIdx.x] : 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) cnt[w][d] = 0;
  }
  __syncthreads();
  const K mask = (K(1) << nbits) - 1;
  const int64_t wbase = int64_t(blockIdx.x) * kSortTile + int64_t(wave) * 64 * kSortIPT;
  K k[kSortIPT];
  int v[kSortIPT];
  int rank[kSortIPT];
  int dig[kSortIPT];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kSortIPT; ++j) {
    const int64_t i = wbase + int64_t(j) * 64 + lane;
    const bool valid = i < n;
    k[j] = valid ? keys[i] : K(0);
    if (kHasVals) v[j] = valid ? vals[i] : 0;
    const int d = valid ? int((k[j] >> shift) & mask) : 255;
    dig[j] = d;
    // Lanes holding the same digit.
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bb = ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? bb : ~bb;
    }
    const int r = popc64(peers & lt);
    const int base = cnt[wave][d];
    // The lowest lane of each digit group advances that digit's counter
    // after every lane of the wave has read it (lockstep read, then write).
    __builtin_amdgcn_wave_barrier();
    if (valid && r == 0) cnt[wave][d] = base + popc64(peers);
    __builtin_amdgcn_wave_barrier();
    rank[j] = base + r;
  }
  __syncthreads();
  // Exclusive prefix of the per-wave counts, per digit.
  for (int d = threadIdx.x; d < 256; d += kSortThreads) {
    int run = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) { int c = cnt[w][d]; cnt[w][d] = run; run += c; }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortIPT; ++j) {
    const int64_t i = wbase + int64_t(j) * 64 + lane;
    if (i < n) {
      const int d = dig[j];
      const int64_t pos = int64_t(goff[d]) + cnt[wave][d] + rank[j];
      okeys[pos] = k[j];
      if (kHasVals) ovals[pos] = v[j];
    }
  }
}

// Stable LSD radix sort of (keys, vals) on bits [begin_bit, end_bit).  The
// sorted data ends in (keys, vals) if *in_alt == false on return, otherwise in
// (keys_alt, vals_alt).  vals may be null (keys only).
template <typename K>
int radix_sort_pairs(K* keys, int* vals, K* keys_alt, int* vals_alt, int64_t n, int begin_bit,
                     int end_bit, int* scratch, hipStream_t st, bool* in_alt) {
  *in_alt = false;
  if (n <= 1 || end_bit <= begin_bit) return 0;
  const int nb = static_cast<int>((n + kSortTile - 1) / kSortTile);
  int* hist = scratch;
  int* sscr = scratch + int64_t(256) * nb + 64;
  K* src_k = keys;
  K* dst_k = keys_alt;
  int* src_v = vals;
  int* dst_v = vals_alt;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    const int nbits = (end_bit - shift) < 8 ? (end_bit - shift) : 8;
    k_radix_hist<K><<<nb, kSortThreads, 0, st>>>(src_k, n, shift, nbits, hist, nb);
    int rc = scan(int64_t(256) * nb, ArrayLoad{hist}, ArrayStoreExcl{hist}, nullptr, sscr, st);
    if (rc) return rc;
    if (vals)
      k_radix_scatter<K, true><<<nb, kSortThreads, 0, st>>>(src_k, src_v, dst_k, dst_v, n, shift, nbits, hist, nb);
    else
      k_radix_scatter<K, false><<<nb, kSortThreads, 0, st>>>(src_k, nullptr, dst_k, nullptr, n, shift, nbits, hist, nb);
    std::swap(src_k, dst_k);
    std::swap(src_v, dst_v);
    *in_alt = !*in_alt;
  }
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

}  // namespace gkd
