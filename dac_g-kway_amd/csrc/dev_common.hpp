// Device-side helpers shared by the gfx950 kernels: wave64 lane utilities,
// launch-shape constants and error reporting.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <type_traits>

#define GK_WAVE 64

namespace gkd {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Bit mask of lanes below this one (wave64).
__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// ---------------------------------------------------------------- DPP moves
// Lane permutations and shifts as DPP source modifiers (gfx9 encoding): a
// VALU operation instead of an LDS-pipe ds_bpermute.  A "row" is 16 lanes.
namespace dpp {
constexpr int kQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]: lane ^ 1
constexpr int kQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]: lane ^ 2
constexpr int kQuadXor3 = 0x1B;       // quad_perm [3,2,1,0]: lane ^ 3
constexpr int kRowShr = 0x110;        // + n (1..15): lane - n within the row, zero below
constexpr int kRowRor8 = 0x128;       // rotate by 8 within the row: lane ^ 8
constexpr int kWaveShl1 = 0x130;      // lane + 1 (lane 63: zero)
constexpr int kWaveShr1 = 0x138;      // lane - 1 (lane 0: zero)
constexpr int kRowMirror = 0x140;     // lane ^ 15
constexpr int kRowHalfMirror = 0x141; // lane ^ 7
constexpr int kRowBcast15 = 0x142;    // lane 15 of each row to the next row
constexpr int kRowBcast31 = 0x143;    // lane 31 to rows 2 and 3
}  // namespace dpp

// x from the lane that DPP control Ctrl names; lanes of rows outside
// RowMask, and lanes whose source is outside the row or wave, get 0.
template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t x) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), Ctrl, RowMask, 0xF, true));
}

// x from lane ^ M inside aligned groups of 32 lanes (ds_swizzle, bit-mask
// mode: and-mask 31, or-mask 0, xor-mask M).
template <int M>
__device__ __forceinline__ uint32_t swz_xor(uint32_t x) {
  return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(x), 0x1F | (M << 10)));
}

// Inclusive add-scan inside aligned segments of S lanes (S = 16, 32, 64);
// every lane of a segment must be active.
template <int S>
__device__ __forceinline__ int32_t seg_incl_scan(int32_t x) {
  x += int32_t(dpp_mov<dpp::kRowShr + 1>(uint32_t(x)));
  x += int32_t(dpp_mov<dpp::kRowShr + 2>(uint32_t(x)));
  x += int32_t(dpp_mov<dpp::kRowShr + 4>(uint32_t(x)));
  x += int32_t(dpp_mov<dpp::kRowShr + 8>(uint32_t(x)));
  if (S >= 32) x += int32_t(dpp_mov<dpp::kRowBcast15, 0xA>(uint32_t(x)));
  if (S >= 64) x += int32_t(dpp_mov<dpp::kRowBcast31, 0xC>(uint32_t(x)));
  return x;
}

// Ascending sort of one uint32 per lane inside aligned segments of S lanes
// (S = 16, 32, 64; every lane of a segment active): Batcher's bitonic network
// with the mirror form of each merge (compare lane i with its mirror in the
// merged block, then half-cleaners on lane ^ j), so every compare keeps the
// minimum in the lower lane.  Partners inside a row come from DPP, lane ^ 16
// and ^ 31 from ds_swizzle, lane ^ 63 from ds_bpermute.
template <int S>
__device__ __forceinline__ uint32_t seg_sort_u32(uint32_t x) {
  const int l = lane_id();
  const bool lo1 = !(l & 1), lo2 = !(l & 2), lo4 = !(l & 4), lo8 = !(l & 8), lo16 = !(l & 16), lo32 = !(l & 32);
  auto cx = [&](uint32_t y, bool lower) { x = lower ? min(x, y) : max(x, y); };
  auto xor4 = [&]() { return dpp_mov<dpp::kQuadXor3>(dpp_mov<dpp::kRowHalfMirror>(x)); };
  cx(dpp_mov<dpp::kQuadXor1>(x), lo1);  // blocks of 2
  cx(dpp_mov<dpp::kQuadXor3>(x), lo2);  // blocks of 4
  cx(dpp_mov<dpp::kQuadXor1>(x), lo1);
  cx(dpp_mov<dpp::kRowHalfMirror>(x), lo4);  // blocks of 8
  cx(dpp_mov<dpp::kQuadXor2>(x), lo2);
  cx(dpp_mov<dpp::kQuadXor1>(x), lo1);
  cx(dpp_mov<dpp::kRowMirror>(x), lo8);  // blocks of 16
  cx(xor4(), lo4);
  cx(dpp_mov<dpp::kQuadXor2>(x), lo2);
  cx(dpp_mov<dpp::kQuadXor1>(x), lo1);
  if (S >= 32) {
    cx(swz_xor<31>(x), lo16);  // blocks of 32
    cx(dpp_mov<dpp::kRowRor8>(x), lo8);
    cx(xor4(), lo4);
    cx(dpp_mov<dpp::kQuadXor2>(x), lo2);
    cx(dpp_mov<dpp::kQuadXor1>(x), lo1);
  }
  if (S >= 64) {
    cx(static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), 63, 64)), lo32);  // the whole wave
    cx(swz_xor<16>(x), lo16);
    cx(dpp_mov<dpp::kRowRor8>(x), lo8);
    cx(xor4(), lo4);
    cx(dpp_mov<dpp::kQuadXor2>(x), lo2);
    cx(dpp_mov<dpp::kQuadXor1>(x), lo1);
  }
  return x;
}

// x from lane ^ J (J a power of two below 64): DPP moves inside a row,
// ds_swizzle for 16, ds_bpermute for 32.
template <int J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x) {
  static_assert(J == 1 || J == 2 || J == 4 || J == 8 || J == 16 || J == 32, "lane distance");
  if constexpr (J == 1) return dpp_mov<dpp::kQuadXor1>(x);
  else if constexpr (J == 2) return dpp_mov<dpp::kQuadXor2>(x);
  else if constexpr (J == 4) return dpp_mov<dpp::kQuadXor3>(dpp_mov<dpp::kRowHalfMirror>(x));
  else if constexpr (J == 8) return dpp_mov<dpp::kRowRor8>(x);
  else if constexpr (J == 16) return swz_xor<16>(x);
  else return static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), J, 64));
}
template <int J>
__device__ __forceinline__ uint64_t lane_xor(uint64_t x) {
  const uint64_t lo = lane_xor<J>(static_cast<uint32_t>(x)), hi = lane_xor<J>(static_cast<uint32_t>(x >> 32));
  return (hi << 32) | lo;
}

// f(std::integral_constant<int, J>) for J = J0, J0 / 2, ..., 1.
template <int J0, class F>
__device__ __forceinline__ void for_pow2_down(F&& f) {
  f(std::integral_constant<int, J0>{});
  if constexpr (J0 > 1) for_pow2_down<J0 / 2>(f);
}

// Wave-wide op-reduction by lane ^ J exchanges (whole wave active); every
// lane gets the result.
template <class T, class Op>
__device__ __forceinline__ T wave_reduce(T x, Op op) {
  for_pow2_down<32>([&](auto jc) { x = op(x, static_cast<T>(lane_xor<decltype(jc)::value>(x))); });
  return x;
}

// Wave-level inclusive add-scan (int32), whole wave active.
__device__ __forceinline__ int wave_incl_scan(int x) { return seg_incl_scan<64>(x); }

__device__ __forceinline__ int wave_sum(int x) {
  return static_cast<int>(wave_reduce(static_cast<uint32_t>(x), [](uint32_t a, uint32_t b) { return a + b; }));
}

__device__ __forceinline__ long long wave_sum64(long long x) {
  return static_cast<long long>(wave_reduce(static_cast<uint64_t>(x), [](uint64_t a, uint64_t b) { return a + b; }));
}

__device__ __forceinline__ int wave_max(int x) {
  return static_cast<int>(wave_reduce(static_cast<uint32_t>(x), [](uint32_t a, uint32_t b) {
    return static_cast<uint32_t>(max(static_cast<int>(a), static_cast<int>(b)));
  }));
}

// Highest lane at or below `lane` whose bit is set in m (-1 if none).
__device__ __forceinline__ int last_set_at_or_below(uint64_t m, int lane) {
  const uint64_t v = m & (~0ull >> (63 - lane));
  return v ? 63 - __clzll(static_cast<long long>(v)) : -1;
}

// XCD-aware workgroup order.  The dispatcher deals the workgroups of a
// launch to the eight XCDs in turn (workgroup b to XCD b % 8), and each XCD
// has its own L2.  xcd_tile renumbers them so that XCD x takes a contiguous
// run of tiles -- [x * per, (x + 1) * per) with per = G / 8, the first G % 8
// XCDs one tile more -- still in increasing order over time, so the rows
// that neighbouring tiles share, and the neighbour lookups of graphs
// numbered with locality, are fetched into one L2 instead of into all
// eight.  It is a bijection on [0, G), so no result changes.  Measured on
// MI355X (pack of 4096 graphs, level 0, HBM bytes fetched per launch):
// matching 2048 -> 1201 MB, contraction rows (16 / 32 / 64 lanes) 1521 /
// 2208 / 1744 -> 808 / 1372 / 1216 MB.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t G) {
  const uint32_t per = G / 8, rem = G % 8;
  const uint32_t x = b % 8, i = b / 8;
  return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}
__device__ __forceinline__ uint32_t xcd_block() { return xcd_tile(blockIdx.x, gridDim.x); }

// Fill kernel used instead of hipMemsetAsync, so the same code path can be
// captured into a HIP graph.
template <typename T>
__global__ void k_fill(T* p, T v, int64_t n) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

inline unsigned grid_for(long long items, int per_block) {
  long long g = (items + per_block - 1) / per_block;
  return static_cast<unsigned>(g < 1 ? 1 : g);
}

template <typename T>
inline void fill(T* p, T v, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  long long g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  k_fill<T><<<static_cast<unsigned>(g), 256, 0, st>>>(p, v, n);
}

}  // namespace gkd

// Host-side status plumbing: every HIP call goes through GK_HIP, which
// records the failing call and returns GK_ERR_HIP from the enclosing
// function (functions returning int status codes only).
namespace gkh {
void set_error(const std::string& msg);
}

#define GK_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      gkh::set_error(std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                     __FILE__ + ":" + std::to_string(__LINE__));                  \
      return GK_ERR_HIP;                                                          \
    }                                                                             \
  } while (0)

#define GK_TRY(expr)          \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != 0) return rc_; \
  } while (0)
