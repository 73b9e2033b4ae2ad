// Shared helpers for the gfx950 kernels: error reporting for the C-ABI and
// cross-lane exchange inside a 64-wide wavefront.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

namespace igm {

// Last error message of the calling host thread (read by igm_last_error()).
void set_error(const std::string& msg);
const char* last_error();

#define IGM_CHECK_ARG(cond, msg)                         \
  do {                                                   \
    if (!(cond)) {                                       \
      ::igm::set_error(std::string(__func__) + ": " + (msg)); \
      return -1;                                         \
    }                                                    \
  } while (0)

#define IGM_CHECK_LAUNCH()                                                    \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      ::igm::set_error(std::string(__func__) + ": launch failed: " + hipGetErrorString(e_)); \
      return -2;                                                              \
    }                                                                         \
  } while (0)

constexpr int kWave = 64;

// Lane l receives the value held by lane l+1 (lane 63 receives 0).
// DPP wave_shl:1 -- one VALU move, no LDS traffic.
__device__ __forceinline__ float lane_from_right(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xF, 0xF, true));
}
// Lane l receives the value held by lane l-1 (lane 0 receives 0).
// DPP wave_shr:1.
__device__ __forceinline__ float lane_from_left(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ double lane_from_right(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xF, 0xF, true);
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_from_left(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, true);
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wavefront (every lane gets the total).
// Raw buffer resources with 32-bit byte offsets: an offset at or beyond the
// resource size reads 0 and drops a store, in hardware, so out-of-grid and
// masked lanes need no branch.  kOOB is above every resource size (the host
// checks that each buffer is below 2 GiB).
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr uint32_t kOOB = 0x80000000u;
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer, gfx9 data format

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kRsrcWord3);
}

// MI355X dispatches the workgroups of a launch to its 8 XCDs round-robin
// (workgroup p runs on XCD p % 8), and each XCD has an L2 of its own.
// xcd_remap turns the dispatch index p of nb workgroups into a logical index
// such that every XCD gets one contiguous range of logical indices: tiles
// that are neighbours in the logical order then share an L2, and the cache
// lines on their common edges are fetched from HBM once.  A bijection on
// [0, nb) whatever the hardware order, so only speed depends on it.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_remap(int p, int nb) {
  const int q = nb / kXcds, r = nb % kXcds;
  const int x = p % kXcds, i = p / kXcds;
  return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}
__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}
__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x = fmaxf(x, __shfl_xor(x, off, kWave));
  return x;
}

// x^e for x > 0 via the hardware log2/exp2 (fp32) or libm (fp64).
__device__ __forceinline__ float pow_pos(float x, float e) {
  return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
}
__device__ __forceinline__ double pow_pos(double x, double e) { return pow(x, e); }
// d * x^e for x, d > 0, with l2d = log2(d): one exp2 of e log2(x) + log2(d)
// in fp32 (the multiply by d folds into the exponent)
__device__ __forceinline__ float scaled_pow(float x, float e, float d, float l2d) {
  return __builtin_amdgcn_exp2f(__builtin_fmaf(e, __builtin_amdgcn_logf(x), l2d));
}
__device__ __forceinline__ double scaled_pow(double x, double e, double d, double l2d) {
  return d * pow(x, e);
}
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
// hardware square root (v_sqrt_f32, ~1 ulp) -- not the correctly rounded expansion
__device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
// clamp of x to [lo, hi] in one v_med3_f32
__device__ __forceinline__ float clamp_(float x, float lo, float hi) { return __builtin_amdgcn_fmed3f(x, lo, hi); }
__device__ __forceinline__ double clamp_(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

}  // namespace igm
