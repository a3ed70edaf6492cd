// Device-side helpers shared by the gfx950 kernels: wave64 lane utilities,
// launch-shape constants and error reporting.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

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

// Wave-level inclusive add-scan (int32), whole wave active.
__device__ __forceinline__ int wave_incl_scan(int x) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (l >= o) x += y;
  }
  return x;
}

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ long long wave_sum64(long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
  return x;
}

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
