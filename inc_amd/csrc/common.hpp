// Shared host/device helpers of libinc_amd: error reporting, real-type
// traits, wavefront reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/inc_amd.h"

namespace inc {

// ---- error handling --------------------------------------------------------
void set_error(const char* fmt, ...);
void clear_error();
const char* last_error();

#define INC_REQUIRE(cond, code, ...)      \
  do {                                    \
    if (!(cond)) {                        \
      ::inc::set_error(__VA_ARGS__);      \
      return (code);                      \
    }                                     \
  } while (0)

#define INC_HIP(call)                                                       \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ::inc::set_error("HIP error %s at %s:%d", hipGetErrorString(e_),      \
                       __FILE__, __LINE__);                                 \
      return INC_EHIP;                                                      \
    }                                                                       \
  } while (0)

// Status of the last launch on this thread (launch-configuration errors).
int check_launch(const char* what);

// ---- device helpers -------------------------------------------------------
constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T rcp(T x) {
  return T(1) / x;
}
template <>
__device__ __forceinline__ float rcp<float>(float x) {
  // v_rcp_f32: 1 ulp; inputs here are normal numbers (>= 1e-12)
  return __builtin_amdgcn_rcpf(x);
}

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask) {
  return __shfl_xor(v, mask, kWave);
}

// (|u| max, first index) reduction across `width` lanes (power of two,
// <= 64) of a wavefront; ties keep the smaller index, as torch.max(dim)
// returns the first maximal element.
template <typename T>
__device__ __forceinline__ void wave_argmax(T& v, int& idx, int width) {
  for (int off = 1; off < width; off <<= 1) {
    T ov = shfl_xor(v, off);
    int oi = __shfl_xor(idx, off, kWave);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v, int width) {
  for (int off = 1; off < width; off <<= 1) v += shfl_xor(v, off);
  return v;
}

}  // namespace inc
