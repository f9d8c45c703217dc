// Resolution changes of the UNet1D correction network for gfx950: the
// 2-to-1 max pooling of its encoder and the 2x upsampling (transposed
// convolution, kernel 2, stride 2) with the skip concatenation of its
// decoder (reference models/models_1d/unet.py:5-53).
//
// The transposed convolution's arithmetic is a 1x1 convolution with twice
// the output channels (run by the library's matrix-core convolution): with
// W'[t*C + c][ci] = W[ci][c][t] and the bias repeated,
//   y2[b][t*C + c][i] = bias[c] + sum_ci W[ci][c][t] x[b][ci][i],
// and out[b][c][2i + t] = y2[b][t*C + c][i] is torch.nn.ConvTranspose1d(k=2,
// s=2).  The kernels here move those values into place: they interleave y2's
// two halves into the upsampled rows and write the skip connection's
// channels behind them, which is torch.cat((upconv(x), skip), dim=1) in one
// pass; the backward splits the concatenation's gradient the same way.
//
// All kernels are elementwise over flat indices (HBM-bound, one read and
// one write of every element); grid-stride loops over at most 8192 blocks.

#include <algorithm>

#include "common.hpp"

namespace inc {
namespace {

constexpr int kBlock = 256;

int64_t blocks_for(int64_t n) { return std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 8192)); }

// torch's max pooling picks x1 over x0 when x1 > x0 or x1 is NaN (the
// first element wins ties; a NaN wins over numbers): the CUDA kernel of
// torch.nn.MaxPool1d, max_pool2d_with_indices, scans the window with
// `val > maxval || isnan(val)`.
template <typename T>
__device__ __forceinline__ bool take_second(T x0, T x1) {
  return x1 > x0 || x1 != x1;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) maxpool2_fwd_kernel(int64_t m, const T* __restrict__ x, T* __restrict__ y) {
  for (int64_t o = int64_t(blockIdx.x) * kBlock + threadIdx.x; o < m; o += int64_t(gridDim.x) * kBlock) {
    const T x0 = x[2 * o], x1 = x[2 * o + 1];
    y[o] = take_second(x0, x1) ? x1 : x0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) maxpool2_bwd_kernel(int64_t m, const T* __restrict__ x,
                                                              const T* __restrict__ gy, T* __restrict__ gx) {
  for (int64_t o = int64_t(blockIdx.x) * kBlock + threadIdx.x; o < m; o += int64_t(gridDim.x) * kBlock) {
    const bool second = take_second(x[2 * o], x[2 * o + 1]);
    const T g = gy[o];
    gx[2 * o] = second ? T(0) : g;
    gx[2 * o + 1] = second ? g : T(0);
  }
}

// out [B][C + Cs][2n]; y2 [B][2C][n]; skip [B][Cs][2n] (Cs may be 0).
template <typename T>
__global__ void __launch_bounds__(kBlock) upsample2_cat_kernel(int64_t total, int C, int Cs, int n,
                                                               const T* __restrict__ y2, const T* __restrict__ skip,
                                                               T* __restrict__ out) {
  const int row = 2 * n;
  const int per_b = (C + Cs) * row;
  for (int64_t e = int64_t(blockIdx.x) * kBlock + threadIdx.x; e < total; e += int64_t(gridDim.x) * kBlock) {
    const int64_t b = e / per_b;
    const int r = int(e - b * per_b);
    const int c = r / row, j = r - c * row;
    out[e] = c < C ? y2[(b * 2 * C + (j & 1) * C + c) * n + (j >> 1)] : skip[(b * Cs + (c - C)) * row + j];
  }
}

// g [B][C + Cs][2n] -> gy2 [B][2C][n], gskip [B][Cs][2n] (NULL: not wanted).
template <typename T>
__global__ void __launch_bounds__(kBlock) upsample2_cat_bwd_kernel(int64_t total, int C, int Cs, int n,
                                                                   const T* __restrict__ g, T* __restrict__ gy2,
                                                                   T* __restrict__ gskip) {
  const int row = 2 * n;
  const int per_b = (C + Cs) * row;
  for (int64_t e = int64_t(blockIdx.x) * kBlock + threadIdx.x; e < total; e += int64_t(gridDim.x) * kBlock) {
    const int64_t b = e / per_b;
    const int r = int(e - b * per_b);
    const int c = r / row, j = r - c * row;
    if (c < C)
      gy2[(b * 2 * C + (j & 1) * C + c) * n + (j >> 1)] = g[e];
    else if (gskip)
      gskip[(b * Cs + (c - C)) * row + j] = g[e];
  }
}

int dtype_check(int dtype) {
  INC_REQUIRE(dtype == INC_F32 || dtype == INC_F64, INC_EINVAL, "unsupported dtype %d", dtype);
  return INC_OK;
}

template <typename T>
int maxpool2_fwd(int64_t m, const void* x, void* y, hipStream_t st) {
  hipLaunchKernelGGL(maxpool2_fwd_kernel<T>, dim3(unsigned(blocks_for(m))), dim3(kBlock), 0, st, m,
                     static_cast<const T*>(x), static_cast<T*>(y));
  return check_launch("max pooling");
}

template <typename T>
int maxpool2_bwd(int64_t m, const void* x, const void* gy, void* gx, hipStream_t st) {
  hipLaunchKernelGGL(maxpool2_bwd_kernel<T>, dim3(unsigned(blocks_for(m))), dim3(kBlock), 0, st, m,
                     static_cast<const T*>(x), static_cast<const T*>(gy), static_cast<T*>(gx));
  return check_launch("max pooling gradient");
}

template <typename T>
int upsample2_cat(int64_t total, int C, int Cs, int n, const void* y2, const void* skip, void* out, hipStream_t st) {
  hipLaunchKernelGGL(upsample2_cat_kernel<T>, dim3(unsigned(blocks_for(total))), dim3(kBlock), 0, st, total, C, Cs,
                     n, static_cast<const T*>(y2), static_cast<const T*>(skip), static_cast<T*>(out));
  return check_launch("upsampling concatenation");
}

template <typename T>
int upsample2_cat_bwd(int64_t total, int C, int Cs, int n, const void* g, void* gy2, void* gskip, hipStream_t st) {
  hipLaunchKernelGGL(upsample2_cat_bwd_kernel<T>, dim3(unsigned(blocks_for(total))), dim3(kBlock), 0, st, total, C,
                     Cs, n, static_cast<const T*>(g), static_cast<T*>(gy2), static_cast<T*>(gskip));
  return check_launch("upsampling concatenation gradient");
}

int cat_check(int dtype, int B, int C, int n, int Cs, const void* a, const void* b, const void* c) {
  if (int rc = dtype_check(dtype)) return rc;
  INC_REQUIRE(B >= 0 && C >= 1 && Cs >= 0 && n >= 1, INC_EINVAL, "upsampling concatenation: bad sizes");
  INC_REQUIRE(int64_t(C + Cs) * 2 * n < (int64_t(1) << 31), INC_EINVAL,
              "upsampling concatenation: (C + Cs) * 2n must stay below 2^31");
  INC_REQUIRE(B == 0 || (a && c && (Cs == 0 || b)), INC_EINVAL, "upsampling concatenation: missing buffer");
  return INC_OK;
}

}  // namespace
}  // namespace inc

extern "C" {

int inc_maxpool2_fwd(int dtype, int64_t rows, int n, const void* x, void* y, void* stream) {
  inc::clear_error();
  if (int rc = inc::dtype_check(dtype)) return rc;
  INC_REQUIRE(rows >= 0 && n >= 2 && n % 2 == 0, INC_EINVAL, "max pooling: rows >= 0 and an even n >= 2 (got n = %d)",
              n);
  const int64_t m = rows * (n / 2);
  if (m == 0) return INC_OK;
  INC_REQUIRE(x && y, INC_EINVAL, "max pooling: missing buffer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return dtype == INC_F32 ? inc::maxpool2_fwd<float>(m, x, y, st) : inc::maxpool2_fwd<double>(m, x, y, st);
}

int inc_maxpool2_bwd(int dtype, int64_t rows, int n, const void* x, const void* gy, void* gx, void* stream) {
  inc::clear_error();
  if (int rc = inc::dtype_check(dtype)) return rc;
  INC_REQUIRE(rows >= 0 && n >= 2 && n % 2 == 0, INC_EINVAL, "max pooling: rows >= 0 and an even n >= 2 (got n = %d)",
              n);
  const int64_t m = rows * (n / 2);
  if (m == 0) return INC_OK;
  INC_REQUIRE(x && gy && gx, INC_EINVAL, "max pooling gradient: missing buffer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return dtype == INC_F32 ? inc::maxpool2_bwd<float>(m, x, gy, gx, st) : inc::maxpool2_bwd<double>(m, x, gy, gx, st);
}

int inc_upsample2_cat(int dtype, int B, int C, int n, int Cs, const void* y2, const void* skip, void* out,
                      void* stream) {
  inc::clear_error();
  if (int rc = inc::cat_check(dtype, B, C, n, Cs, y2, skip, out)) return rc;
  const int64_t total = int64_t(B) * (C + Cs) * 2 * n;
  if (total == 0) return INC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return dtype == INC_F32 ? inc::upsample2_cat<float>(total, C, Cs, n, y2, skip, out, st)
                          : inc::upsample2_cat<double>(total, C, Cs, n, y2, skip, out, st);
}

int inc_upsample2_cat_bwd(int dtype, int B, int C, int n, int Cs, const void* g, void* gy2, void* gskip,
                          void* stream) {
  inc::clear_error();
  if (int rc = inc::dtype_check(dtype)) return rc;
  INC_REQUIRE(B >= 0 && C >= 1 && Cs >= 0 && n >= 1, INC_EINVAL, "upsampling concatenation gradient: bad sizes");
  INC_REQUIRE(int64_t(C + Cs) * 2 * n < (int64_t(1) << 31), INC_EINVAL,
              "upsampling concatenation gradient: (C + Cs) * 2n must stay below 2^31");
  const int64_t total = int64_t(B) * (C + Cs) * 2 * n;
  if (total == 0) return INC_OK;
  INC_REQUIRE(g && gy2, INC_EINVAL, "upsampling concatenation gradient: missing buffer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return dtype == INC_F32 ? inc::upsample2_cat_bwd<float>(total, C, Cs, n, g, gy2, gskip, st)
                          : inc::upsample2_cat_bwd<double>(total, C, Cs, n, g, gy2, gskip, st);
}

}  // extern "C"
