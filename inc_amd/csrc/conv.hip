// Periodic 1D convolution kernels of the INC correction network (gfx950).
//
// The correction s(u) that every INC rollout step evaluates is a stack of
// circular Conv1d layers (the reference's correction networks use 5-tap
// circular convolutions, models/models_1d/resnet.py:5-52).  These kernels
// replace the library convolutions on that path:
//   y[b,o,i] = bias[o] + sum_{c,k} W[o,c,k] x[b,c,(i+k-P) mod N],  P = K/2
// with
//   * forward: a block covers 128 positions of one row; its input tile (with
//     the periodic halo) is staged in LDS; each wavefront owns a wave-uniform
//     set of output channels whose weights come through the scalar cache;
//   * input gradient: the same kernel on the flipped, transposed weights;
//   * weight/bias gradient: per-block partial sums over a strided set of
//     (row, tile) pairs written to a caller workspace, then a second kernel
//     reduces the partials in a fixed order (deterministic).
#include <algorithm>
#include <type_traits>

#include "common.hpp"

namespace inc {
namespace {

constexpr int kPPT = 2;    // output positions per lane
constexpr int kWaves = 4;  // wavefronts per block, each with its own output channels

__device__ __forceinline__ int wrapn(int j, int n) {
  j %= n;
  return j < 0 ? j + n : j;
}

// Exact (erf) GELU, torch.nn.GELU(approximate='none'), and its derivative.
template <typename T>
__device__ __forceinline__ T gelu(T v) {
  return T(0.5) * v * (T(1) + erf(v * T(0.70710678118654752440)));
}
template <typename T>
__device__ __forceinline__ T gelu_grad(T v) {
  return T(0.5) * (T(1) + erf(v * T(0.70710678118654752440))) + v * T(0.39894228040143267794) * exp(T(-0.5) * v * v);
}

// One block: TI = 64 * kPPT positions of one row, kWaves * CPT output
// channels.  The input tile (with its periodic halo) is staged in LDS and
// shared by the block's wavefronts; every wavefront owns CPT output channels
// (wave-uniform), so its weights are uniform loads served by the scalar
// cache and each lane accumulates CPT x kPPT outputs in registers.
// FLIP: read W as the input-gradient operator W'[c_out', c_in', k] =
// W[c_in', c_out', K-1-k] (channels swapped, taps reversed).
// Epilogue, per output v = acc + bias: times mul[b,o,i] when mul is given
// (the input gradient of a layer whose input was an activation: mul holds
// the activation's derivative); with act = GELU, y = gelu(v) and
// dy = gelu'(v) (the next layer's input and the derivative its input
// gradient needs).
template <typename T, int K, int CPT, bool FLIP>
__global__ void __launch_bounds__(256) conv_fwd_kernel(int Ci, int Co, int N, const T* __restrict__ x,
                                                      const T* __restrict__ w, const T* __restrict__ bias,
                                                      T* __restrict__ y, const T* __restrict__ mul, int act,
                                                      T* __restrict__ dy) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int P = K / 2;
  constexpr int TI = kWave * kPPT;
  constexpr int XW = TI + K - 1;
  T* xs = reinterpret_cast<T*>(smem_raw);  // [Ci][XW]
  const int b = blockIdx.z;
  const int i0 = blockIdx.x * TI;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int lane = tid % kWave;
  const int o0 = (blockIdx.y * kWaves + wave) * CPT;
  const T* xb = x + size_t(b) * Ci * N;
  for (int e = tid; e < Ci * XW; e += blockDim.x) {
    int c = e / XW, j = e - c * XW;
    xs[e] = xb[size_t(c) * N + wrapn(i0 - P + j, N)];
  }
  __syncthreads();
  if (o0 >= Co) return;
  T acc[CPT][kPPT];
#pragma unroll
  for (int a = 0; a < CPT; ++a)
#pragma unroll
    for (int p = 0; p < kPPT; ++p) acc[a][p] = T(0);
  const int p0 = lane * kPPT;
  for (int c = 0; c < Ci; ++c) {
    T xv[kPPT + K - 1];
#pragma unroll
    for (int j = 0; j < kPPT + K - 1; ++j) xv[j] = xs[c * XW + p0 + j];
#pragma unroll
    for (int a = 0; a < CPT; ++a) {
      const int o = o0 + a < Co ? o0 + a : Co - 1;  // clamp: extra channels are computed and dropped
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const T wv = FLIP ? w[(size_t(c) * Co + o) * K + (K - 1 - k)] : w[(size_t(o) * Ci + c) * K + k];
#pragma unroll
        for (int p = 0; p < kPPT; ++p) acc[a][p] += wv * xv[p + k];
      }
    }
  }
#pragma unroll
  for (int a = 0; a < CPT; ++a) {
    const int o = o0 + a;
    if (o >= Co) continue;
    const T bo = bias ? bias[o] : T(0);
    const size_t row = (size_t(b) * Co + o) * N;
#pragma unroll
    for (int p = 0; p < kPPT; ++p) {
      const int i = i0 + p0 + p;
      if (i < N) {
        T v = acc[a][p] + bo;
        if (mul) v *= mul[row + i];
        if (act == INC_ACT_GELU) {
          dy[row + i] = gelu_grad(v);
          v = gelu(v);
        }
        y[row + i] = v;
      }
    }
  }
}

// Partial weight/bias gradients of a strided set of (row, tile) pairs.
// Thread t owns a 2 x 2 block of (output, input) channel pairs (blocks t,
// t + nthr, ...) and all K taps, so each staged gradient / input value
// feeds two weight pairs.  LDS rows are skewed to an odd stride so the
// threads of a wavefront (consecutive input channels) hit distinct banks.
template <typename T, int K, int MAXP>
__global__ void __launch_bounds__(256) conv_wgrad_kernel(int B, int Ci, int Co, int N, int TI, const T* __restrict__ g,
                                                        const T* __restrict__ x, T* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int P = K / 2;
  const int GW = TI | 1;                   // gradient row stride (odd)
  const int XW = (TI + K - 1) | 1;         // input row stride (odd)
  T* gs = reinterpret_cast<T*>(smem_raw);  // [Co2][GW]
  const int Co2 = (Co + 1) & ~1, Ci2 = (Ci + 1) & ~1;
  T* xs = gs + size_t(Co2) * GW;           // [Ci2][XW]
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nblk2 = (Co2 / 2) * (Ci2 / 2);  // 2 x 2 channel blocks
  T acc[MAXP][4][K];
  T bacc = T(0);
#pragma unroll
  for (int q = 0; q < MAXP; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int k = 0; k < K; ++k) acc[q][r][k] = T(0);
  const int tiles = (N + TI - 1) / TI;
  for (int job = blockIdx.x; job < B * tiles; job += gridDim.x) {
    const int b = job / tiles, i0 = (job - b * tiles) * TI;
    __syncthreads();
    const T* gb = g + size_t(b) * Co * N;
    const T* xb = x + size_t(b) * Ci * N;
    for (int e = tid; e < Co2 * TI; e += nthr) {
      int o = e / TI, j = e - o * TI;
      gs[o * GW + j] = (o < Co && i0 + j < N) ? gb[size_t(o) * N + i0 + j] : T(0);
    }
    const int xw = TI + K - 1;
    for (int e = tid; e < Ci2 * xw; e += nthr) {
      int c = e / xw, j = e - c * xw;
      xs[c * XW + j] = c < Ci ? xb[size_t(c) * N + wrapn(i0 - P + j, N)] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
      const int blk = tid + q * nthr;
      if (blk < nblk2) {
        const int o = 2 * (blk / (Ci2 / 2)), c = 2 * (blk % (Ci2 / 2));
        const T* g0 = gs + o * GW;
        const T* g1 = g0 + GW;
        const T* x0 = xs + c * XW;
        const T* x1 = x0 + XW;
        int i = 0;
        for (; i + 8 <= TI; i += 8) {
          T xa[8 + K - 1], xc[8 + K - 1], ga[8], gc[8];
#pragma unroll
          for (int j = 0; j < 8 + K - 1; ++j) { xa[j] = x0[i + j]; xc[j] = x1[i + j]; }
#pragma unroll
          for (int j = 0; j < 8; ++j) { ga[j] = g0[i + j]; gc[j] = g1[i + j]; }
#pragma unroll
          for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int k = 0; k < K; ++k) {
              acc[q][0][k] += ga[j] * xa[j + k];
              acc[q][1][k] += ga[j] * xc[j + k];
              acc[q][2][k] += gc[j] * xa[j + k];
              acc[q][3][k] += gc[j] * xc[j + k];
            }
        }
        for (; i < TI; ++i)
#pragma unroll
          for (int k = 0; k < K; ++k) {
            acc[q][0][k] += g0[i] * x0[i + k];
            acc[q][1][k] += g0[i] * x1[i + k];
            acc[q][2][k] += g1[i] * x0[i + k];
            acc[q][3][k] += g1[i] * x1[i + k];
          }
      }
    }
    if (tid < Co)
      for (int i = 0; i < TI; ++i) bacc += gs[tid * GW + i];
  }
  const int npairs = Co * Ci;
  T* out = part + size_t(blockIdx.x) * (npairs * K + Co);
#pragma unroll
  for (int q = 0; q < MAXP; ++q) {
    const int blk = tid + q * nthr;
    if (blk < nblk2) {
      const int o = 2 * (blk / (Ci2 / 2)), c = 2 * (blk % (Ci2 / 2));
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int oo = o + (r >> 1), cc = c + (r & 1);
        if (oo < Co && cc < Ci)
#pragma unroll
          for (int k = 0; k < K; ++k) out[(size_t(oo) * Ci + cc) * K + k] = acc[q][r][k];
      }
    }
  }
  if (tid < Co) out[npairs * K + tid] = bacc;
}

// Sum of nparts partial rows: a block covers 16 consecutive elements with 16
// partial-lanes each; lanes combine in a fixed order (deterministic).
template <typename T>
__global__ void __launch_bounds__(256) reduce_parts_kernel(int nparts, int width, const T* __restrict__ part,
                                                          T* __restrict__ dw, T* __restrict__ db, int nw) {
  __shared__ T red[16][17];
  const int ex = threadIdx.x % 16, pl = threadIdx.x / 16;
  const int e = blockIdx.x * 16 + ex;
  T s = T(0);
  if (e < width)
    for (int p = pl; p < nparts; p += 16) s += part[size_t(p) * width + e];
  red[pl][ex] = s;
  __syncthreads();
  if (pl == 0 && e < width) {
    T t = red[0][ex];
    for (int q = 1; q < 16; ++q) t += red[q][ex];
    if (e < nw) dw[e] = t;
    else if (db) db[e - nw] = t;
  }
}

// ---- weight gradient on the matrix cores (float32) --------------------------------
// G[o][col] = sum_pos g[b,o,pos] * X[col][pos] with col = c*K + k and
// X[c*K + k][pos] = x[b,c,(pos + k - P) mod N]; col = Ci*K is a column of ones
// (the bias gradient).  v_mfma_f32_16x16x4f32 takes A = g[16 o][4 pos] and
// B = X[4 pos][16 col]; lane l supplies A[l%16][l/16] and B[l/16][l%16] and
// holds D[4*(l/16) + r][l%16], r < 4.  A block of four wavefronts loops over
// (row, 128-position tile) jobs staged in LDS; wavefront w takes the column
// tiles w % WT, w % WT + WT, ... and the 4-position steps s with
// s % WP == w / WT (WT * WP <= 4: thin layers with few column tiles split
// positions instead).  Every wavefront's accumulators are one partial sum,
// written to slot blockIdx.x * WP + w / WT of the workspace.
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kWgTP = 128;  // positions per job

template <int MT, int NT>
__global__ void __launch_bounds__(256) conv_wgrad_mfma_kernel(int B, int Ci, int Co, int N, int K, int WT, int WP,
                                                             const float* __restrict__ g,
                                                             const float* __restrict__ x, float* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int P = K / 2;
  constexpr int GS = kWgTP + 1;             // odd row strides: the 16 rows of a fragment
  const int XS = (kWgTP + K - 1) | 1;       // read hit distinct banks
  float* gs = reinterpret_cast<float*>(smem_raw);  // [MT*16][GS]
  float* xs = gs + MT * 16 * GS;                   // [Ci][XS]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tphase = wave % WT, pphase = wave / WT;
  const bool active = pphase < WP;
  const int ncw = Ci * K, ntile = (ncw + 1 + 15) / 16;
  const int lr = lane & 15, lk = lane >> 4;
  // this lane's B-fragment source per column tile: offset into xs, or the
  // ones column (-1), or padding (-2)
  int off[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int tile = tphase + t * WT, col = tile * 16 + lr;
    off[t] = tile >= ntile || col > ncw ? -2 : (col == ncw ? -1 : (col / K) * XS + col % K);
  }
  v4f acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[m][t] = v4f{0.f, 0.f, 0.f, 0.f};
  const int tiles = (N + kWgTP - 1) / kWgTP;
  const int xw = kWgTP + K - 1;
  for (int job = blockIdx.x; job < B * tiles; job += gridDim.x) {
    const int b = job / tiles, i0 = (job - b * tiles) * kWgTP;
    const float* gb = g + size_t(b) * Co * N;
    const float* xb = x + size_t(b) * Ci * N;
    // Stage the job's tiles with every load of a thread in flight before its
    // LDS writes (the staging is latency-bound otherwise): all MT * 8
    // gradient values, then the input rows eight at a time.
    float gv[MT * 8];
#pragma unroll
    for (int q = 0; q < MT * 8; ++q) {
      const int e = tid + q * 256, o = e / kWgTP, j = e % kWgTP;
      gv[q] = (o < Co && i0 + j < N) ? gb[size_t(o) * N + i0 + j] : 0.f;
    }
    __syncthreads();  // the previous job's fragments are consumed
#pragma unroll
    for (int q = 0; q < MT * 8; ++q) {
      const int e = tid + q * 256;
      gs[(e / kWgTP) * GS + e % kWgTP] = gv[q];
    }
    for (int e0 = tid; e0 < Ci * xw; e0 += 256 * 8) {
      float xv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = e0 + q * 256;
        if (e < Ci * xw) {
          const int c = e / xw, j = e - c * xw;
          int pos = i0 - P + j;  // in [-P, N + kWgTP + K): one wrap unless N is small
          pos = pos < 0 ? pos + N : (pos >= N ? pos - N : pos);
          if (pos >= N) pos %= N;
          xv[q] = xb[size_t(c) * N + pos];
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = e0 + q * 256;
        if (e < Ci * xw) xs[(e / xw) * XS + e % xw] = xv[q];
      }
    }
    __syncthreads();
    if (active) {
      // fragments of step s + WP are read while step s's MFMAs run
      float a[MT], bv[NT], an[MT], bn[NT];
      auto load = [&](int s, float* fa, float* fb) {
        const int p = 4 * s + lk;
#pragma unroll
        for (int m = 0; m < MT; ++m) fa[m] = gs[(m * 16 + lr) * GS + p];
#pragma unroll
        for (int t = 0; t < NT; ++t) fb[t] = off[t] >= 0 ? xs[off[t] + p] : (off[t] == -1 ? 1.f : 0.f);
      };
      load(pphase, a, bv);
      for (int s = pphase; s < kWgTP / 4; s += WP) {
        if (s + WP < kWgTP / 4) load(s + WP, an, bn);
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m], bv[t], acc[m][t], 0, 0, 0);
#pragma unroll
        for (int m = 0; m < MT; ++m) a[m] = an[m];
#pragma unroll
        for (int t = 0; t < NT; ++t) bv[t] = bn[t];
      }
    }
  }
  if (!active) return;
  const int width = Co * ncw + Co;
  float* out = part + size_t(blockIdx.x * WP + pphase) * width;
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int col = (tphase + t * WT) * 16 + lr;
      if (off[t] == -2) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = m * 16 + 4 * lk + r;
        if (o >= Co) continue;
        if (col < ncw) out[size_t(o) * ncw + col] = acc[m][t][r];
        else out[size_t(Co) * ncw + o] = acc[m][t][r];
      }
    }
}

// MFMA weight gradient when the shape fits (Co <= 64, Ci*K + 1 <= 512, LDS);
// returns INC_EUNSUPPORTED (without touching the error text) otherwise.
int conv_wgrad_mfma(int B, int Ci, int Co, int N, int K, const float* g, const float* x, float* dw, float* db,
                    float* part, int64_t work_elems, hipStream_t st) {
  const int MT = (Co + 15) / 16;
  const int ncw = Ci * K, ntile = (ncw + 1 + 15) / 16;
  const int WT = ntile >= 4 ? 4 : (ntile >= 2 ? 2 : 1);
  const int ntw = (ntile + WT - 1) / WT;
  const int NT = ntw <= 4 ? ntw : (ntw <= 6 ? 6 : 8);
  const size_t lds = (size_t(MT) * 16 * (kWgTP + 1) + size_t(Ci) * ((kWgTP + K - 1) | 1)) * sizeof(float);
  if (MT > 4 || ntw > 8 || lds > 150 * 1024) return INC_EUNSUPPORTED;
  const int width = Co * ncw + Co;
  const int64_t slots = std::min<int64_t>(512, work_elems / width);  // two blocks per CU
  if (slots < 1) return INC_EUNSUPPORTED;
  const int WP = int(std::min<int64_t>(4 / WT, slots));
  const int jobs = B * ((N + kWgTP - 1) / kWgTP);
  const int nblk = int(std::min<int64_t>(jobs, slots / WP));
  dim3 grid(nblk), block(256);
#define INC_WGM(M, T) \
  hipLaunchKernelGGL((conv_wgrad_mfma_kernel<M, T>), grid, block, lds, st, B, Ci, Co, N, K, WT, WP, g, x, part)
#define INC_WGM_T(M)                  \
  switch (NT) {                       \
    case 1: INC_WGM(M, 1); break;     \
    case 2: INC_WGM(M, 2); break;     \
    case 3: INC_WGM(M, 3); break;     \
    case 4: INC_WGM(M, 4); break;     \
    case 6: INC_WGM(M, 6); break;     \
    default: INC_WGM(M, 8); break;    \
  }
  switch (MT) {
    case 1: INC_WGM_T(1); break;
    case 2: INC_WGM_T(2); break;
    case 3: INC_WGM_T(3); break;
    default: INC_WGM_T(4); break;
  }
#undef INC_WGM_T
#undef INC_WGM
  int rc = check_launch("conv1d weight gradient (matrix cores)");
  if (rc) return rc;
  hipLaunchKernelGGL(reduce_parts_kernel<float>, dim3((width + 15) / 16), dim3(256), 0, st, nblk * WP, width, part,
                     dw, db, Co * ncw);
  return check_launch("conv1d weight gradient reduction");
}

template <typename T, int K, bool FLIP>
int conv_fwd_k(int cpt, dim3 grid, size_t lds, hipStream_t st, int Ci, int Co, int N, const T* x, const T* w,
               const T* b, T* y, const T* mul, int act, T* dy) {
  dim3 block(kWaves * kWave);
#define INC_CF(C) \
  hipLaunchKernelGGL((conv_fwd_kernel<T, K, C, FLIP>), grid, block, lds, st, Ci, Co, N, x, w, b, y, mul, act, dy)
  switch (cpt) {
    case 1: INC_CF(1); break;
    case 2: INC_CF(2); break;
    case 4: INC_CF(4); break;
    case 8: INC_CF(8); break;
    case 12: INC_CF(12); break;
    default: INC_CF(16); break;
  }
#undef INC_CF
  return check_launch("conv1d");
}

// Output channels per wavefront: the smallest of 1, 2, 4, 8, 12, 16 with
// kWaves * cpt >= Co (capped at 16; wider layers use more blocks).
int conv_cpt(int Co) {
  static const int opts[] = {1, 2, 4, 8, 12, 16};
  for (int c : opts)
    if (kWaves * c >= Co) return c;
  return 16;
}

template <typename T, bool FLIP>
int conv_fwd_launch(int B, int Ci, int Co, int N, int K, const void* x, const void* w, const void* bias, void* y,
                    const void* mul, int act, void* dy, hipStream_t st) {
  const int TI = kWave * kPPT;
  size_t lds = size_t(Ci) * (TI + K - 1) * sizeof(T);
  INC_REQUIRE(lds <= 150 * 1024, INC_EUNSUPPORTED, "conv1d: %d input channels need %zu bytes of LDS", Ci, lds);
  const int cpt = conv_cpt(Co);
  dim3 grid((N + TI - 1) / TI, (Co + kWaves * cpt - 1) / (kWaves * cpt), B);
  const T* xp = static_cast<const T*>(x);
  const T* wp = static_cast<const T*>(w);
  const T* bp = static_cast<const T*>(bias);
  const T* mp = static_cast<const T*>(mul);
  T* yp = static_cast<T*>(y);
  T* dp = static_cast<T*>(dy);
  switch (K) {
    case 1: return conv_fwd_k<T, 1, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp);
    case 3: return conv_fwd_k<T, 3, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp);
    case 5: return conv_fwd_k<T, 5, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp);
    case 7: return conv_fwd_k<T, 7, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp);
    case 9: return conv_fwd_k<T, 9, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp);
    default: set_error("conv1d kernel size must be 1, 3, 5, 7 or 9 (got %d)", K); return INC_EUNSUPPORTED;
  }
}

template <typename T, int K>
int wgrad_launch_k(int B, int Ci, int Co, int N, int TI, int nblk, size_t lds, const T* g, const T* x, T* part,
                   hipStream_t st) {
  const int nblk2 = ((Co + 1) / 2) * ((Ci + 1) / 2);
  const int maxp = (nblk2 + 255) / 256;
  dim3 grid(nblk), block(256);
#define INC_WG(MP) \
  hipLaunchKernelGGL((conv_wgrad_kernel<T, K, MP>), grid, block, lds, st, B, Ci, Co, N, TI, g, x, part)
  if (maxp <= 1) INC_WG(1);
  else if (maxp <= 2) INC_WG(2);
  else if (maxp <= 3) INC_WG(3);
  else if (maxp <= 4) INC_WG(4);
  else {
    set_error("conv1d weight gradient: %d x %d channels exceed 4096 channel pairs", Co, Ci);
    return INC_EUNSUPPORTED;
  }
#undef INC_WG
  return check_launch("conv1d weight gradient");
}

template <typename T>
int conv_wgrad(int B, int Ci, int Co, int N, int K, const void* g, const void* x, void* dw, void* db, void* work,
               int64_t work_elems, hipStream_t st) {
  if constexpr (std::is_same<T, float>::value) {
    const int rc = conv_wgrad_mfma(B, Ci, Co, N, K, static_cast<const float*>(g), static_cast<const float*>(x),
                                   static_cast<float*>(dw), static_cast<float*>(db), static_cast<float*>(work),
                                   work_elems, st);
    if (rc != INC_EUNSUPPORTED) return rc;
  }
  int TI = N < 128 ? N : 128;
  size_t lds = (size_t((Co + 1) & ~1) * (TI | 1) + size_t((Ci + 1) & ~1) * ((TI + K - 1) | 1)) * sizeof(T);
  INC_REQUIRE(lds <= 150 * 1024, INC_EUNSUPPORTED, "conv1d weight gradient needs %zu bytes of LDS", lds);
  const int width = Co * Ci * K + Co;
  int jobs = B * ((N + TI - 1) / TI);
  int nblk = jobs < 256 ? jobs : 256;  // about one block per CU, each looping over several tiles
  if (int64_t(nblk) * width > work_elems) nblk = int(work_elems / width);
  INC_REQUIRE(nblk >= 1, INC_EINVAL, "conv1d weight gradient workspace too small (need >= %d elements)", width);
  const T* gp = static_cast<const T*>(g);
  const T* xp = static_cast<const T*>(x);
  T* part = static_cast<T*>(work);
  int rc;
  switch (K) {
    case 1: rc = wgrad_launch_k<T, 1>(B, Ci, Co, N, TI, nblk, lds, gp, xp, part, st); break;
    case 3: rc = wgrad_launch_k<T, 3>(B, Ci, Co, N, TI, nblk, lds, gp, xp, part, st); break;
    case 5: rc = wgrad_launch_k<T, 5>(B, Ci, Co, N, TI, nblk, lds, gp, xp, part, st); break;
    case 7: rc = wgrad_launch_k<T, 7>(B, Ci, Co, N, TI, nblk, lds, gp, xp, part, st); break;
    case 9: rc = wgrad_launch_k<T, 9>(B, Ci, Co, N, TI, nblk, lds, gp, xp, part, st); break;
    default: set_error("conv1d kernel size must be 1, 3, 5, 7 or 9 (got %d)", K); return INC_EUNSUPPORTED;
  }
  if (rc) return rc;
  hipLaunchKernelGGL(reduce_parts_kernel<T>, dim3((width + 15) / 16), dim3(256), 0, st, nblk, width, part,
                     static_cast<T*>(dw), static_cast<T*>(db), Co * Ci * K);
  return check_launch("conv1d weight gradient reduction");
}

int conv_check(int dtype, int act, int B, int Ci, int Co, int N, int K) {
  INC_REQUIRE(dtype == INC_F32 || dtype == INC_F64, INC_EINVAL, "unsupported dtype %d", dtype);
  INC_REQUIRE(act == INC_ACT_NONE || act == INC_ACT_GELU, INC_EINVAL, "unknown activation %d", act);
  INC_REQUIRE(B >= 0 && Ci >= 1 && Co >= 1 && N >= 1, INC_EINVAL, "conv1d: bad sizes");
  INC_REQUIRE(K % 2 == 1 && K <= 9 && K <= N, INC_EUNSUPPORTED, "conv1d kernel size must be odd, <= 9 and <= n");
  return INC_OK;
}

}  // namespace
}  // namespace inc

extern "C" {

int inc_conv1d_fwd_act(int dtype, int act, int B, int Ci, int Co, int N, int K, const void* x, const void* w,
                       const void* bias, void* y, void* dy, void* stream) {
  inc::clear_error();
  int rc = inc::conv_check(dtype, act, B, Ci, Co, N, K);
  if (rc || B == 0) return rc;
  INC_REQUIRE(act == INC_ACT_NONE || dy != nullptr, INC_EINVAL, "conv1d with an activation needs dy");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return dtype == INC_F32
             ? inc::conv_fwd_launch<float, false>(B, Ci, Co, N, K, x, w, bias, y, nullptr, act, dy, st)
             : inc::conv_fwd_launch<double, false>(B, Ci, Co, N, K, x, w, bias, y, nullptr, act, dy, st);
}

int inc_conv1d_bwd_data_mul(int dtype, int B, int Ci, int Co, int N, int K, const void* g, const void* w,
                            const void* m, void* gx, void* stream) {
  inc::clear_error();
  int rc = inc::conv_check(dtype, INC_ACT_NONE, B, Ci, Co, N, K);
  if (rc || B == 0) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // gx = conv(g, W') with W'[c][o][k] = W[o][c][K-1-k]: a forward pass with Co -> Ci channels,
  // times m elementwise when given
  return dtype == INC_F32
             ? inc::conv_fwd_launch<float, true>(B, Co, Ci, N, K, g, w, nullptr, gx, m, INC_ACT_NONE, nullptr, st)
             : inc::conv_fwd_launch<double, true>(B, Co, Ci, N, K, g, w, nullptr, gx, m, INC_ACT_NONE, nullptr, st);
}

int inc_conv1d_fwd(int dtype, int B, int Ci, int Co, int N, int K, const void* x, const void* w, const void* bias,
                   void* y, void* stream) {
  return inc_conv1d_fwd_act(dtype, INC_ACT_NONE, B, Ci, Co, N, K, x, w, bias, y, nullptr, stream);
}

int inc_conv1d_bwd_data(int dtype, int B, int Ci, int Co, int N, int K, const void* g, const void* w, void* gx,
                        void* stream) {
  return inc_conv1d_bwd_data_mul(dtype, B, Ci, Co, N, K, g, w, nullptr, gx, stream);
}

int inc_conv1d_bwd_weight(int dtype, int B, int Ci, int Co, int N, int K, const void* g, const void* x, void* gw,
                          void* gb, void* work, int64_t work_elems, void* stream) {
  inc::clear_error();
  int rc = inc::conv_check(dtype, INC_ACT_NONE, B, Ci, Co, N, K);
  if (rc) return rc;
  INC_REQUIRE(B > 0, INC_EINVAL, "conv1d weight gradient of an empty batch");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return dtype == INC_F32 ? inc::conv_wgrad<float>(B, Ci, Co, N, K, g, x, gw, gb, work, work_elems, st)
                          : inc::conv_wgrad<double>(B, Ci, Co, N, K, g, x, gw, gb, work, work_elems, st);
}

}  // extern "C"
