// Periodic 1D convolution kernels of the INC correction network (gfx950).
//
// The correction s(u) that every INC rollout step evaluates is a stack of
// circular Conv1d layers (the reference's correction networks use 5-tap
// circular convolutions, models/models_1d/resnet.py:5-52).  These kernels
// replace the library convolutions on that path:
//   y[b,o,i] = bias[o] + sum_{c,k} W[o,c,k] x[b,c,(i+k-P) mod N],  P = K/2
// with
//   * forward (float32, input channels a multiple of 16, 16..64 output
//     channels): the matrix cores, weights staged once per block in LDS,
//     blocks looping over position tiles with the next tile prefetched;
//   * forward (other shapes, float64): a block covers 128 positions of one
//     row; its input tile (with the periodic halo) is staged in LDS; each
//     wavefront owns a wave-uniform set of output channels whose weights
//     come through the scalar cache;
//   * input gradient: the same kernels on the flipped, transposed weights;
//   * weight/bias gradient: per-block partial sums over a strided set of
//     (row, tile) pairs written to a caller workspace, then a second kernel
//     reduces the partials in a fixed order (deterministic).
#include <algorithm>
#include <mutex>
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

// Rows of a convolution.  Uniform: B rows of N points, row b's [C][N] block
// at element C*b*N.  Ragged (tile != nullptr): row r has n[r] points and its
// [C][n[r]] block starts at element C*off[r]; the work is a list of
// kRowTile-position tiles (row, first position), of which a launch takes the
// first njobs (lists are ordered so that rows still inside their horizon come
// first).  Positions of a tile past its row's end are neither read into the
// result nor written.
constexpr int kRowTile = 128;
struct RowTiles {
  const int* tile;
  const int* n;
  const int64_t* off;
};

// How a weight gradient runs: over uniform rows or ragged row tiles (njobs
// of them); writing or adding to the per-block partial sums; with or without
// the final reduction (a caller that accumulates over several calls reduces
// once at the end, over the largest partial count any call used, *nparts).
struct WgradMode {
  RowTiles rt;
  int njobs;
  bool accumulate;
  bool reduce;
  int* nparts;
};

// One block of a weight gradient's channels: output channels [o0, o0 + Co)
// of the gC channels of g and input channels [c0, c0 + Ci) of the xC
// channels of x, where the kernel's Co and Ci are the block's sizes.  The
// whole gradient is the block {Co, Ci, 0, 0}.  rpj > 1 (the packed weight
// gradient, n = 16, 32 or 64): a job is rpj = 128 / n consecutive batch rows
// of `rows`.
struct ChanBlock {
  int gC, xC, o0, c0;
  int rpj, rows;  // packed rows (rows per job, batch rows); rpj = 1: one row per job
};

// One channel block of a matrix-core convolution: the kernel's Ci input
// channels are channels [c0, c0 + Ci) of the xC of x, its Co outputs are
// channels [o0, o0 + Co) of the yC of y (and of mul, dy); the weights are
// indexed in the full layout (flip: w[c][o] with yC per row, else w[o][c]
// with xC).  phase: 0 the whole reduction (bias, mul, activation in the
// epilogue); over several input-channel blocks 1 writes the first partial
// sum plus the bias, 2 adds a middle block's, 3 adds the last and applies
// mul and the activation.  rpj > 1 (n = 16 or 32, the packed kernel): a
// job is rpj = 64 / n consecutive batch rows of `rows`, each staged with its
// own periodic halo, so no position of the 64-position tile is idle.
struct ConvBlock {
  int xC, c0, yC, o0, phase;
  int rpj, rows;  // packed rows (rows per job, batch rows); rpj = 1: one row per job
};

struct JobPos {
  int b, i0, n;
  int64_t base;  // first point of the row (element offset C * base)
};

// The tile lists are never written by the kernels that read them: read
// through constant-address-space pointers at a wave-uniform index, they are
// scalar loads (scalar cache, counted apart from the vector memory traffic)
// instead of vector loads that wait behind every access in flight.  job is
// wave-uniform in every caller.
template <typename V>
__device__ __forceinline__ V const_load(const V* p, int i) {
  return ((const __attribute__((address_space(4))) V*)p)[i];
}

__device__ __forceinline__ JobPos job_pos(const RowTiles& rt, int job, int tiles, int TI, int N) {
  if (rt.tile) {
    job = __builtin_amdgcn_readfirstlane(job);
    const int b = const_load(rt.tile, 2 * job), i0 = const_load(rt.tile, 2 * job + 1);
    return JobPos{b, i0, const_load(rt.n, b), const_load(rt.off, b)};
  }
  const int b = job / tiles;
  return JobPos{b, (job - b * tiles) * TI, N, int64_t(b) * N};
}

// Exact (erf) GELU, torch.nn.GELU(approximate='none'), and its derivative
// from one erf evaluation: g = gelu(v) = v h with h = (1 + erf(v/sqrt 2))/2,
// returns gelu'(v) = h + v exp(-v^2/2) / sqrt(2 pi).
template <typename T>
__device__ __forceinline__ T gelu_and_grad(T v, T& g) {
  const T h = T(0.5) * (T(1) + erf(v * T(0.70710678118654752440)));
  g = v * h;
  return h + v * T(0.39894228040143267794) * exp(T(-0.5) * v * v);
}

// gelu(v) alone (no gradient wanted): the same value as gelu_and_grad's g.
template <typename T>
__device__ __forceinline__ T gelu_value(T v) {
  return v * (T(0.5) * (T(1) + erf(v * T(0.70710678118654752440))));
}

// ReLU as torch.relu: NaN stays NaN; the derivative is 1 where the output is
// positive, 0 elsewhere (NaN and 0 included), as relu's backward masks it.
template <typename T>
__device__ __forceinline__ T relu(T v) {
  return v < T(0) ? T(0) : v;
}
template <typename T>
__device__ __forceinline__ T relu_grad(T v) {
  return v > T(0) ? T(1) : T(0);
}

// float32: Phi(-|v|) = erfc(z) / 2 with z = |v| / sqrt 2 and
// erfc(z) = exp(-z^2) t Q(t), t = 1 / (1 + p z), Q a degree-9 polynomial
// least-squares fitted (relative weights) to erfc(z) exp(z^2) / t over
// z in [0, 14]: relative error 3.6e-9 in exact arithmetic.  Evaluated in
// float32, gelu is within 2 ulp of the exact value for v >= 1, 7 ulp on
// [-1, 1) and 32 ulp (exp(-v^2/2) of a rounded square) on [-6, -1), where
// (1 + erf) in float32 loses every digit; about 23 instructions where the
// library erf and exp take twice that.  exp(-v^2/2) is shared by gelu and
// gelu'.  NaN propagates; gelu(+-inf) and gelu'(+-inf) are what the erf form
// gives (+inf, NaN; NaN, NaN).
constexpr float kGeluP = 0.34f;
__device__ constexpr float kGeluQ[10] = {0.19181646453274762f, 0.1920384442751763f, 0.17825813396327506f,
                                         0.17491467764958224f, 0.05928315898984988f, 0.27570342376534757f,
                                         -0.2785340354581258f, 0.4018655397996093f, -0.24607228818300206f,
                                         0.050726484182980335f};

// Phi(v) = (1 + erf(v / sqrt 2)) / 2, with e = exp(-v^2/2)
__device__ __forceinline__ float gelu_phi(float v, float& e) {
  e = __builtin_amdgcn_exp2f(v * v * -0.72134752044448170368f);  // 2^(-v^2 log2(e) / 2)
  const float t = __builtin_amdgcn_rcpf(fmaf(kGeluP * 0.70710678118654752440f, fabsf(v), 1.0f));
  float q = kGeluQ[9];
#pragma unroll
  for (int k = 8; k >= 0; --k) q = fmaf(q, t, kGeluQ[k]);
  const float h = 0.5f * e * (t * q);  // Phi(-|v|)
  return v >= 0.0f ? 1.0f - h : h;
}

template <>
__device__ __forceinline__ float gelu_and_grad<float>(float v, float& g) {
  float e;
  const float ph = gelu_phi(v, e);
  g = v * ph;
  return fmaf(v * 0.39894228040143267794f, e, ph);
}

template <>
__device__ __forceinline__ float gelu_value<float>(float v) {
  float e;
  return v * gelu_phi(v, e);
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
// gradient needs), or y alone when dy is NULL (no gradient wanted).
template <typename T, int K, int CPT, bool FLIP>
__global__ void __launch_bounds__(256) conv_fwd_kernel(int Ci, int Co, int Nu, const T* __restrict__ x,
                                                      const T* __restrict__ w, const T* __restrict__ bias,
                                                      T* __restrict__ y, const T* __restrict__ mul, int act,
                                                      T* __restrict__ dy, RowTiles rt) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int P = K / 2;
  constexpr int TI = kWave * kPPT;
  static_assert(TI == kRowTile, "ragged tiles are one block's positions");
  constexpr int XW = TI + K - 1;
  T* xs = reinterpret_cast<T*>(smem_raw);  // [Ci][XW]
  // uniform rows: grid (tiles, channel groups, B); ragged: grid (jobs, channel groups)
  const JobPos jp = rt.tile ? job_pos(rt, blockIdx.x, 0, TI, 0)
                            : JobPos{int(blockIdx.z), int(blockIdx.x) * TI, Nu, int64_t(blockIdx.z) * Nu};
  const int N = jp.n, i0 = jp.i0;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int lane = tid % kWave;
  const int o0 = (blockIdx.y * kWaves + wave) * CPT;
  const T* xb = x + jp.base * Ci;
  // four loads of a thread in flight before their LDS writes
  constexpr int SU = 4;
  for (int e0 = tid; e0 < Ci * XW; e0 += 256 * SU) {
    T v[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const int e = e0 + u * 256;
      if (e < Ci * XW) {
        const int c = e / XW, j = e - c * XW;
        v[u] = xb[size_t(c) * N + wrapn(i0 - P + j, N)];
      }
    }
#pragma unroll
    for (int u = 0; u < SU; ++u)
      if (e0 + u * 256 < Ci * XW) xs[e0 + u * 256] = v[u];
  }
  __syncthreads();
  T acc[CPT][kPPT];
#pragma unroll
  for (int a = 0; a < CPT; ++a)
#pragma unroll
    for (int p = 0; p < kPPT; ++p) acc[a][p] = T(0);
  // Position p of a lane: with many output channels per wavefront the
  // lanes take positions lane, lane + 64, ... (every output store is 64
  // consecutive values); with few, consecutive pairs (one input window of
  // kPPT + K - 1 values serves both, fewer LDS reads per input channel).
  constexpr bool kStrided = CPT >= 4;
  auto pos = [&](int p) { return kStrided ? p * kWave + lane : lane * kPPT + p; };
  if constexpr (CPT == 1) {
    if (Co == 1 && Ci >= kWaves) {
      // One output channel (a network's last layer, or the input gradient of
      // its first): wavefront w sums input channels w, w + kWaves, ... for
      // every position of the tile, and the kWaves partial sums meet in LDS
      // (over the staged tile, Ci * XW >= kWaves * TI) in wavefront order.
      for (int c = wave; c < Ci; c += kWaves) {
        T xv[kPPT + K - 1];
#pragma unroll
        for (int j = 0; j < kPPT + K - 1; ++j) xv[j] = xs[c * XW + pos(0) + j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const T wv = FLIP ? w[size_t(c) * K + (K - 1 - k)] : w[size_t(c) * K + k];
#pragma unroll
          for (int p = 0; p < kPPT; ++p) acc[0][p] += wv * xv[p + k];
        }
      }
      __syncthreads();  // every wavefront is done with the staged tile
#pragma unroll
      for (int p = 0; p < kPPT; ++p) xs[wave * TI + pos(p)] = acc[0][p];
      __syncthreads();
      if (wave != 0) return;
#pragma unroll
      for (int p = 0; p < kPPT; ++p) {
        T v = xs[pos(p)];
#pragma unroll
        for (int q = 1; q < kWaves; ++q) v += xs[q * TI + pos(p)];
        acc[0][p] = v;
      }
    }
  }
  if (o0 >= Co) return;
  for (int c = (CPT == 1 && Co == 1 && Ci >= kWaves) ? Ci : 0; c < Ci; ++c) {
    if constexpr (kStrided) {
      T xv[kPPT][K];
#pragma unroll
      for (int p = 0; p < kPPT; ++p)
#pragma unroll
        for (int k = 0; k < K; ++k) xv[p][k] = xs[c * XW + pos(p) + k];
#pragma unroll
      for (int a = 0; a < CPT; ++a) {
        const int o = o0 + a < Co ? o0 + a : Co - 1;  // clamp: extra channels are computed and dropped
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const T wv = FLIP ? w[(size_t(c) * Co + o) * K + (K - 1 - k)] : w[(size_t(o) * Ci + c) * K + k];
#pragma unroll
          for (int p = 0; p < kPPT; ++p) acc[a][p] += wv * xv[p][k];
        }
      }
    } else {
      T xv[kPPT + K - 1];
#pragma unroll
      for (int j = 0; j < kPPT + K - 1; ++j) xv[j] = xs[c * XW + pos(0) + j];
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
  }
#pragma unroll
  for (int a = 0; a < CPT; ++a) {
    const int o = o0 + a;
    if (o >= Co) continue;
    const T bo = bias ? bias[o] : T(0);
    const size_t row = size_t(jp.base) * Co + size_t(o) * N;
#pragma unroll
    for (int p = 0; p < kPPT; ++p) {
      const int i = i0 + pos(p);
      if (i < N) {
        T v = acc[a][p] + bo;
        if (mul) v *= mul[row + i];
        if (act == INC_ACT_GELU) {
          if (dy) {
            T gv;
            dy[row + i] = gelu_and_grad(v, gv);
            v = gv;
          } else {
            v = gelu_value(v);
          }
        } else if (act == INC_ACT_RELU) {
          if (dy) dy[row + i] = relu_grad(v);
          v = relu(v);
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
__global__ void __launch_bounds__(256) conv_wgrad_kernel(int njobs, int Ci, int Co, int Nu, int TI,
                                                        const T* __restrict__ g, const T* __restrict__ x,
                                                        T* __restrict__ part, int accumulate, RowTiles rt,
                                                        ChanBlock cb) {
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
  const int tiles = (Nu + TI - 1) / TI;
  for (int job = blockIdx.x; job < njobs; job += gridDim.x) {
    const JobPos jp = job_pos(rt, job, tiles, TI, Nu);
    const int N = jp.n, i0 = jp.i0;
    __syncthreads();
    const T* gb = g + jp.base * cb.gC + size_t(cb.o0) * N;
    const T* xb = x + jp.base * cb.xC + size_t(cb.c0) * N;
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
          for (int k = 0; k < K; ++k) {
            T* dst = out + (size_t(oo) * Ci + cc) * K + k;
            *dst = accumulate ? *dst + acc[q][r][k] : acc[q][r][k];
          }
      }
    }
  }
  if (tid < Co) out[npairs * K + tid] = accumulate ? out[npairs * K + tid] + bacc : bacc;
}

// Sum of nparts partial rows: a block covers 16 consecutive elements with 16
// partial-lanes each; lanes combine in a fixed order (deterministic).  Each
// lane keeps eight running sums (rows p, p + 16, ..., p + 112 of every
// 128-row group; the rows after the last whole group go to the first), so
// eight loads are in flight rather than one.
template <typename T>
__global__ void __launch_bounds__(256) reduce_parts_kernel(int nparts, int width, const T* __restrict__ part,
                                                          T* __restrict__ dw, T* __restrict__ db, int nw, int rowlen,
                                                          int rowstride) {
  __shared__ T red[16][17];
  const int ex = threadIdx.x % 16, pl = threadIdx.x / 16;
  const int e = blockIdx.x * 16 + ex;
  T s = T(0);
  if (e < width) {
    T a[8] = {};
    int p = pl;
    for (; p + 112 < nparts; p += 128)
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] += part[size_t(p + 16 * u) * width + e];
    for (; p < nparts; p += 16) a[0] += part[size_t(p) * width + e];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += a[u];
  }
  red[pl][ex] = s;
  __syncthreads();
  if (pl == 0 && e < width) {
    T t = red[0][ex];
    for (int q = 1; q < 16; ++q) t += red[q][ex];
    if (e < nw) dw[size_t(e / rowlen) * rowstride + e % rowlen] = t;
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
static_assert(kWgTP == kRowTile, "ragged tiles are one job's positions");

// PACK (n = 16, 32 or 64): a job is cb.rpj = 128 / n whole batch rows; the
// gradient tile is the rows side by side, the input staged row after row
// with its own halo ([c][r (n + K - 1) + i + k]).
template <int MT, int NT, bool PACK = false>
__global__ void __launch_bounds__(256) conv_wgrad_mfma_kernel(int njobs, int Ci, int Co, int Nu, int K, int WT, int WP,
                                                             const float* __restrict__ g,
                                                             const float* __restrict__ x, float* __restrict__ part,
                                                             int accumulate, RowTiles rt, ChanBlock cb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int P = K / 2;
  constexpr int GS = kWgTP + 1;             // odd row strides: the 16 rows of a fragment
  const int XR = Nu + K - 1;                // staged points per packed row
  const int lgn = 31 - __builtin_clz(Nu);   // packed: n is a power of two
  const int XS = ((PACK ? cb.rpj * XR : kWgTP + K - 1)) | 1;  // read hit distinct banks
  float* gs = reinterpret_cast<float*>(smem_raw);  // [MT*16][GS]
  float* xs = gs + MT * 16 * GS;                   // [Ci][XS], then a row of ones and one of zeros
  const int tid = threadIdx.x, lane = tid & 63;
  // the bias column reads the ones row and padding columns the zeros row, so
  // every B-fragment read is an unconditional LDS read (a conditional read
  // makes the MFMAs wait for every read in flight); written once, ordered by
  // the first job's barriers
  for (int e = tid; e < 2 * XS; e += 256) xs[Ci * XS + e] = e < XS ? 1.f : 0.f;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tphase = wave % WT, pphase = wave / WT;
  const bool active = pphase < WP;
  const int ncw = Ci * K, ntile = (ncw + 1 + 15) / 16;
  const int lr = lane & 15, lk = lane >> 4;
  // this lane's B-fragment source per column tile: offset into xs of its
  // input row and tap, of the ones row (the bias column) or of the zeros row
  // (padding)
  int off[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int tile = tphase + t * WT, col = tile * 16 + lr;
    off[t] = tile >= ntile || col > ncw ? Ci * XS + XS : (col == ncw ? Ci * XS : (col / K) * XS + col % K);
  }
  v4f acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[m][t] = v4f{0.f, 0.f, 0.f, 0.f};
  const int tiles = (Nu + kWgTP - 1) / kWgTP;
  const int xw = PACK ? cb.rpj * XR : kWgTP + K - 1;
  for (int job = blockIdx.x; job < njobs; job += gridDim.x) {
    const JobPos jp = job_pos(rt, job, tiles, kWgTP, Nu);
    const int N = jp.n, i0 = jp.i0;
    const float* gb = g + jp.base * cb.gC + size_t(cb.o0) * N;
    const float* xb = x + jp.base * cb.xC + size_t(cb.c0) * N;
    // Stage the job's tiles with every load of a thread in flight before its
    // LDS writes (the staging is latency-bound otherwise): all MT * 8
    // gradient values, then the input rows eight at a time.
    float gv[MT * 8];
#pragma unroll
    for (int q = 0; q < MT * 8; ++q) {
      const int e = tid + q * 256, o = e / kWgTP, j = e % kWgTP;
      if constexpr (PACK) {
        const int b = job * cb.rpj + (j >> lgn);
        gv[q] = (o < Co && b < cb.rows) ? g[(size_t(b) * cb.gC + cb.o0 + o) * Nu + (j & (Nu - 1))] : 0.f;
      } else {
        gv[q] = (o < Co && i0 + j < N) ? gb[size_t(o) * N + i0 + j] : 0.f;
      }
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
          if constexpr (PACK) {
            const int r = j / XR;
            int pos = j - r * XR - P;  // in [-P, n + P)
            pos = pos < 0 ? pos + Nu : (pos >= Nu ? pos - Nu : pos);
            const int b = min(job * cb.rpj + r, cb.rows - 1);  // past the batch: the gradient rows are 0
            xv[q] = x[(size_t(b) * cb.xC + cb.c0 + c) * Nu + pos];
          } else {
            int pos = i0 - P + j;  // in [-P, N + kWgTP + K): one wrap unless N is small
            pos = pos < 0 ? pos + N : (pos >= N ? pos - N : pos);
            if (pos >= N) pos %= N;
            xv[q] = xb[size_t(c) * N + pos];
          }
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
        const int px = PACK ? p + (p >> lgn) * (K - 1) : p;  // position of the packed row's point
#pragma unroll
        for (int m = 0; m < MT; ++m) fa[m] = gs[(m * 16 + lr) * GS + p];
#pragma unroll
        for (int t = 0; t < NT; ++t) fb[t] = xs[off[t] + px];
      };
      load(pphase, a, bv);
      for (int s = pphase; s < kWgTP / 4; s += WP) {
        load(min(s + WP, kWgTP / 4 - 1), an, bn);  // unconditional: the last one is unused
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
      const int tile = tphase + t * WT, col = tile * 16 + lr;
      if (tile >= ntile || col > ncw) continue;  // padding
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = m * 16 + 4 * lk + r;
        if (o >= Co) continue;
        float* dst = col < ncw ? out + size_t(o) * ncw + col : out + size_t(Co) * ncw + o;
        *dst = accumulate ? *dst + acc[m][t][r] : acc[m][t][r];
      }
    }
}

// MFMA weight gradient when the shape fits (Co <= 64, Ci*K + 1 <= 512, LDS);
// returns INC_EUNSUPPORTED (without touching the error text) otherwise.
// The same weight gradient with the job tiles copied into LDS
// asynchronously (global_load_lds: each lane's value lands at the copy's LDS
// base + 4 * lane, with no register in between), double-buffered: the next
// job's tiles stream in while this job's MFMAs read the other buffer, so no
// staging sits between the jobs' matrix work.  Rows of whole 128-position
// tiles or ragged row tiles (not packed short rows).  Fragments, wavefront
// roles and partial-sum slots are those of conv_wgrad_mfma_kernel; the input
// tile is stored flat, xs[c * xw + j] (xw = 128 + K - 1: one copy per 64
// values), followed by the ones and zeros rows.  Gradient values past the
// row's end or past Co are written as zeros by the lanes that copy nothing.
// NW = 8 wavefronts (WT * WP = 8) where the two buffers leave room for one
// block per CU, so that every SIMD still holds two wavefronts.
template <int MT, int NT, int NW>
__global__ void __launch_bounds__(64 * NW) conv_wgrad_async_kernel(int njobs, int Ci, int Co, int Nu, int K, int WT,
                                                              int WP, const float* __restrict__ g,
                                                              const float* __restrict__ x, float* __restrict__ part,
                                                              int accumulate, RowTiles rt, ChanBlock cb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int P = K / 2;
  constexpr int GS = kWgTP + 1;  // odd row stride: the 16 rows of a fragment read hit distinct banks
  const int xw = kWgTP + K - 1;
  const int gsz = MT * 16 * GS, xsz = (Ci + 2) * xw, bsz = gsz + xsz;
  float* bufs = reinterpret_cast<float*>(smem_raw);  // two of [gs: MT*16][GS] [xs: Ci+2][xw]
  const int tid = threadIdx.x, lane = tid & 63;
  for (int e = tid; e < 2 * xw; e += 64 * NW) {
    const float v = e < xw ? 1.f : 0.f;
    bufs[gsz + Ci * xw + e] = v;
    bufs[bsz + gsz + Ci * xw + e] = v;
  }
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tphase = wave % WT, pphase = wave / WT;
  const bool active = pphase < WP;
  const int ncw = Ci * K, ntile = (ncw + 1 + 15) / 16;
  const int lr = lane & 15, lk = lane >> 4;
  int off[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int tile = tphase + t * WT, col = tile * 16 + lr;
    off[t] = tile >= ntile || col > ncw ? Ci * xw + xw : (col == ncw ? Ci * xw : (col / K) * xw + col % K);
  }
  v4f acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[m][t] = v4f{0.f, 0.f, 0.f, 0.f};
  const int tiles = (Nu + kWgTP - 1) / kWgTP;
  const int nxc = (Ci * xw + 63) / 64;
  const float rxw = 1.f / float(xw);
  // copies of job `job` into buffer `b`, shared round-robin by the wavefronts
  auto issue = [&](int job, int b) __attribute__((always_inline)) {
    const JobPos jp = job_pos(rt, job, tiles, kWgTP, Nu);
    const int N = jp.n, i0 = jp.i0;
    const float* gb = g + jp.base * cb.gC + size_t(cb.o0) * N;
    const float* xb = x + jp.base * cb.xC + size_t(cb.c0) * N;
    float* gsb = bufs + b * bsz;
    float* xsb = gsb + gsz;
    for (int ii = wave; ii < MT * 32; ii += NW) {  // gradient: MT*16 rows of two 64-position halves
      const int o = ii >> 1, j = ((ii & 1) << 6) + lane;
      float* dst = gsb + o * GS + ((ii & 1) << 6);
      if (o < Co && i0 + j < N) __builtin_amdgcn_global_load_lds(gb + size_t(o) * N + i0 + j, dst, 4, 0, 0);
      else dst[lane] = 0.f;
    }
    for (int ii = wave; ii < nxc; ii += NW) {  // input: Ci rows of xw points with their periodic halo
      const int e = ii * 64 + lane;
      if (e < Ci * xw) {
        // c = e / xw from a float reciprocal (e < 2^14): (e + 0.5) rxw is within
        // 2^-21 of (e + 0.5) / xw, whose fraction stays 0.5 / xw from the integers
        const int c = int((float(e) + 0.5f) * rxw), j = e - c * xw;
        int pos = i0 - P + j;  // in [-P, N + kWgTP + K): one wrap unless N is small
        pos = pos < 0 ? pos + N : (pos >= N ? pos - N : pos);
        if (pos >= N) pos %= N;
        __builtin_amdgcn_global_load_lds(xb + size_t(c) * N + pos, xsb + ii * 64, 4, 0, 0);
      }
    }
  };
  int b = 0;
  if (int(blockIdx.x) < njobs) issue(blockIdx.x, 0);
  for (int job = blockIdx.x; job < njobs; job += gridDim.x) {
    __builtin_amdgcn_s_waitcnt(0);  // this wavefront's copies of the job have landed
    __syncthreads();                 // and every wavefront's; the other buffer's last readers are done
    if (job + int(gridDim.x) < njobs) issue(job + int(gridDim.x), b ^ 1);
    if (active) {
      const float* gs = bufs + b * bsz;
      const float* xs = gs + gsz;
      float a[MT], bv[NT], an[MT], bn[NT];
      auto load = [&](int s, float* fa, float* fb) __attribute__((always_inline)) {
        const int p = 4 * s + lk;
#pragma unroll
        for (int m = 0; m < MT; ++m) fa[m] = gs[(m * 16 + lr) * GS + p];
#pragma unroll
        for (int t = 0; t < NT; ++t) fb[t] = xs[off[t] + p];
      };
      load(pphase, a, bv);
      for (int s = pphase; s < kWgTP / 4; s += WP) {
        load(min(s + WP, kWgTP / 4 - 1), an, bn);  // unconditional: the last one is unused
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
    b ^= 1;
  }
  if (!active) return;
  const int width = Co * ncw + Co;
  float* out = part + size_t(blockIdx.x * WP + pphase) * width;
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int tile = tphase + t * WT, col = tile * 16 + lr;
      if (tile >= ntile || col > ncw) continue;  // padding
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = m * 16 + 4 * lk + r;
        if (o >= Co) continue;
        float* dst = col < ncw ? out + size_t(o) * ncw + col : out + size_t(Co) * ncw + o;
        *dst = accumulate ? *dst + acc[m][t][r] : acc[m][t][r];
      }
    }
}

int cu_count();
int blocks_per_cu(const void* fn, size_t lds, int threads = 256);

int conv_wgrad_mfma(int B, int Ci, int Co, int N, int K, const float* g, const float* x, float* dw, float* db,
                    float* part, int64_t work_elems, hipStream_t st, const WgradMode& md, ChanBlock cb) {
  const int MT = (Co + 15) / 16;
  const int ncw = Ci * K, ntile = (ncw + 1 + 15) / 16;
  const int WT = ntile >= 4 ? 4 : (ntile >= 2 ? 2 : 1);
  const int ntw = (ntile + WT - 1) / WT;
  const int NT = ntw <= 4 ? ntw : (ntw <= 6 ? 6 : 8);
  // rows of 16, 32 or 64 points: 128 / n whole rows per job
  const bool pack = !md.rt.tile && (N == 16 || N == 32 || N == 64);
  cb.rpj = pack ? kWgTP / N : 1;
  cb.rows = B;
  const int XS = (pack ? cb.rpj * (N + K - 1) : kWgTP + K - 1) | 1;
  const size_t lds = (size_t(MT) * 16 * (kWgTP + 1) + size_t(Ci) * XS + 2 * size_t(XS)) * sizeof(float);
  if (MT > 4 || ntw > 8 || lds > 150 * 1024) return INC_EUNSUPPORTED;
  const int width = Co * ncw + Co;
  const int64_t slots = std::min<int64_t>(512, work_elems / width);  // two blocks per CU
  if (slots < 1) return INC_EUNSUPPORTED;
  const int WP = int(std::min<int64_t>(4 / WT, slots));
  const int jobs = md.rt.tile ? md.njobs : (pack ? (B + cb.rpj - 1) / cb.rpj : B * ((N + kWgTP - 1) / kWgTP));
  const int nblk = int(std::min<int64_t>(jobs, slots / WP));
  if (nblk < 1) {  // no active rows: the partial sums stay as they are
    if (md.nparts) *md.nparts = std::max(*md.nparts, 0);
    return INC_OK;
  }
  const int accumulate = md.accumulate ? 1 : 0;
  // unpacked rows: the asynchronous, double-buffered kernel when two
  // buffers fit (grid: what the CUs hold at once)
  const size_t lds_async = 2 * (size_t(MT) * 16 * (kWgTP + 1) + size_t(Ci + 2) * (kWgTP + K - 1)) * sizeof(float);
  if (!pack && lds_async <= 150 * 1024) {
    const void* fn = nullptr;
    // one block per CU at this LDS size: eight wavefronts, positions split in two more phases
    const bool eight = lds_async > 75 * 1024 && 8 / WT >= 2;
    const int NWv = eight ? 8 : 4, WPa = eight ? int(std::min<int64_t>(8 / WT, slots)) : WP;
#define INC_WGA(M, T)                                                                          \
  fn = eight ? reinterpret_cast<const void*>(&conv_wgrad_async_kernel<M, T, 8>)              \
             : reinterpret_cast<const void*>(&conv_wgrad_async_kernel<M, T, 4>)
#define INC_WGA_T(M)                \
  switch (NT) {                     \
    case 1: INC_WGA(M, 1); break;   \
    case 2: INC_WGA(M, 2); break;   \
    case 3: INC_WGA(M, 3); break;   \
    case 4: INC_WGA(M, 4); break;   \
    case 6: INC_WGA(M, 6); break;   \
    default: INC_WGA(M, 8); break;  \
  }
    switch (MT) {
      case 1: INC_WGA_T(1); break;
      case 2: INC_WGA_T(2); break;
      case 3: INC_WGA_T(3); break;
      default: INC_WGA_T(4); break;
    }
#undef INC_WGA_T
#undef INC_WGA
    const int per_cu = blocks_per_cu(fn, lds_async, 64 * NWv);
    if (per_cu >= 1) {
      const int ablk = int(std::min<int64_t>({int64_t(jobs), slots / WPa, int64_t(cu_count()) * per_cu}));
      if (md.nparts) *md.nparts = std::max(*md.nparts, ablk * WPa);
      int nj = jobs, ci = Ci, co = Co, nu = N, kk = K, wt = WT, wp = WPa, accu = accumulate;
      RowTiles rt = md.rt;
      void* args[] = {&nj, &ci, &co, &nu, &kk, &wt, &wp, &g, &x, &part, &accu, &rt, &cb};
      INC_HIP(hipLaunchKernel(fn, dim3(ablk), dim3(64 * NWv), args, lds_async, st));
      int rc = check_launch("conv1d weight gradient (matrix cores, asynchronous tiles)");
      if (rc || !md.reduce) return rc;
      hipLaunchKernelGGL(reduce_parts_kernel<float>, dim3((width + 15) / 16), dim3(256), 0, st, ablk * WPa, width,
                         part, dw, db, Co * ncw, ncw, cb.xC * K);
      return check_launch("conv1d weight gradient reduction");
    }
  }
  if (md.nparts) *md.nparts = std::max(*md.nparts, nblk * WP);
  dim3 grid(nblk), block(256);
#define INC_WGM(M, T)                                                                                         \
  if (pack)                                                                                                   \
    hipLaunchKernelGGL((conv_wgrad_mfma_kernel<M, T, true>), grid, block, lds, st, jobs, Ci, Co, N, K, WT, WP, \
                       g, x, part, accumulate, md.rt, cb);                                                   \
  else                                                                                                        \
    hipLaunchKernelGGL((conv_wgrad_mfma_kernel<M, T>), grid, block, lds, st, jobs, Ci, Co, N, K, WT, WP, g, x, \
                       part, accumulate, md.rt, cb)
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
  if (rc || !md.reduce) return rc;
  hipLaunchKernelGGL(reduce_parts_kernel<float>, dim3((width + 15) / 16), dim3(256), 0, st, nblk * WP, width, part,
                     dw, db, Co * ncw, ncw, cb.xC * K);
  return check_launch("conv1d weight gradient reduction");
}

// ---- forward / input gradient on the matrix cores (float32) -----------------------
// D[o][i] = sum_r Wr[o][r] Xr[r][i] with r = k*Ci + c and Xr[k*Ci + c][i] =
// x[b,c,(i+k-P) mod N].  A block of four wavefronts stages the weights once,
// as ws[o][r] (o padded to CT*16 rows with zeros), then loops over (row,
// TI-position tile) jobs, staging each input tile position-major:
// xs[j][c] = x[b,c,(i0-P+j) mod N], j < TI + K - 1.  Wavefront w owns
// positions i0 + 16*PW*w .. + 16*PW and all CT*16 output channels.  The
// reduction runs in chunks of 16 values of r; Ci % 16 == 0 puts a chunk in
// one tap k (uniform across the wavefront).  In chunk s, lane quarter
// q = l/16 feeds r = 16s + 4q + j to MFMA j (j < 4), so its A operands are
// ws[o][16s+4q .. +3] and its B operands xs[i+k][c0+4q .. +3] (c0 = 16s mod
// Ci): one 16-byte LDS read per fragment.  Row strides are 4 mod 16 words,
// so the 16 rows of a fragment read cover every bank.  The next chunk's
// fragments are read while the current chunk's MFMAs run (two register sets
// used in turn).
// flip: the input-gradient operator, as in conv_fwd_kernel; the epilogue is
// conv_fwd_kernel's (bias, mul, GELU with its derivative).
// PACK (PW = 1, n = 16 or 32): a job is cb.rpj = 64 / n whole batch rows,
// staged row after row, each as n + K - 1 rows of xs with its own halo;
// position p of the tile is point p % n of row p / n.
template <int CT, int PW, bool PACK = false>
__global__ void __launch_bounds__(256) conv_fwd_mfma_kernel(int njobs, int Ci, int Co, int Nu, int K, int flip,
                                                            const float* __restrict__ x,
                                                            const float* __restrict__ w,
                                                            const float* __restrict__ bias, float* __restrict__ y,
                                                            const float* __restrict__ mul, int act,
                                                            float* __restrict__ dy, RowTiles rt, ConvBlock cb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int TI = 64 * PW;
  static_assert(!PACK || PW == 1, "packed rows fill a 64-position tile");
  const int P = K / 2;
  const int XR = Nu + K - 1;  // staged rows per packed batch row
  const int R = Ci * K, RS = R + 4, CS = Ci + 4, XW = PACK ? cb.rpj * XR : TI + K - 1;
  float* ws = reinterpret_cast<float*>(smem_raw);  // [CT*16][RS]
  float* xs = ws + CT * 16 * RS;                   // [XW + 1][CS] (row XW spare)
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, q = lane >> 4;
  // weights: thread t takes the (o, c) pairs t, t + 256, ... and their K
  // taps, which are contiguous in w (both layouts); two pairs' loads in
  // flight before the LDS writes (a load-store loop would wait for every
  // load in turn, and per-element index divisions cost more than the copy)
  for (int p0 = tid; p0 < CT * 16 * Ci; p0 += 512) {
    float v[2][9];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int p = p0 + u * 256, o = p / Ci, c = p - o * Ci;
      const bool ok = p < CT * 16 * Ci && o < Co;
      const float* src = w + (flip ? size_t(cb.c0 + c) * cb.yC + cb.o0 + o : size_t(cb.o0 + o) * cb.xC + cb.c0 + c) * K;
#pragma unroll
      for (int k = 0; k < 9; ++k) v[u][k] = (ok && k < K) ? src[flip ? K - 1 - k : k] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int p = p0 + u * 256, o = p / Ci, c = p - o * Ci;
      if (p >= CT * 16 * Ci) continue;
      float* dst = ws + o * RS + c;
#pragma unroll
      for (int k = 0; k < 9; ++k)
        if (k < K) dst[k * Ci] = v[u][k];
    }
  }
  const int tiles = (Nu + TI - 1) / TI, nchunk = R / 16, c4n = Ci / 4;
  const int items = c4n * XW;
  const int p0 = wave * 16 * PW;
  // A tile is c4n * XW items (channel quad c4, row j): item e = tid + 256 u
  // of this thread is quad e / XW at row e % XW, fetched as four coalesced
  // loads x[b, 4c4 .. 4c4+3, pos(j)] into registers and staged as one
  // 16-byte LDS write.  The next job's tile is fetched before this job's
  // MFMAs and staged after them, so its HBM latency overlaps the matrix work.
  // items <= 16 * (64 PW + 8) <= 256 U (Ci <= 64, K <= 9); packed: 16 * (64 + 4 * 8)
  constexpr int U = PACK ? 6 : 4 * PW + 1;
  // A thread's items are the same in every job: quad and row of item u,
  // computed once (the global channel offset clamped to the last quad for
  // items past the tile, which are loaded and not staged)
  int jr[U], cq[U];
  {
    int c4 = tid / XW, j = tid - c4 * XW;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      jr[u] = j;
      cq[u] = c4;
      j += 256;
      while (j >= XW) {
        j -= XW;
        ++c4;
      }
    }
  }
  // packed: each item's batch-row offset within the job and its point, the
  // same in every job
  int prow[PACK ? U : 1], ppos[PACK ? U : 1];
  if constexpr (PACK) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      prow[u] = jr[u] / XR;
      int pos = jr[u] - prow[u] * XR - P;  // in [-P, Nu + P)
      ppos[u] = pos < 0 ? pos + Nu : (pos >= Nu ? pos - Nu : pos);
    }
  }
  auto fetch = [&](int job, v4f* v) __attribute__((always_inline)) {
    if constexpr (PACK) {
      const int b0 = job * cb.rpj;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int b = min(b0 + prow[u], cb.rows - 1);  // rows past the batch are loaded, never stored
        const float* s = x + (size_t(b) * cb.xC + cb.c0 + 4 * min(cq[u], c4n - 1)) * Nu + ppos[u];
        v[u] = v4f{s[0], s[Nu], s[2 * Nu], s[3 * Nu]};
      }
      return;
    }
    const JobPos jp = job_pos(rt, job, tiles, TI, Nu);
    const int N = jp.n, i0 = jp.i0;
    const float* xb = x + jp.base * cb.xC + size_t(cb.c0) * N;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // unconditional loads (items past the tile read quad c4n - 1 and are
      // not staged): a conditional load's value is copied at the branch
      // join, and that copy would wait for it
      int pos = i0 - P + jr[u];  // in [-P, N + TI + K): one wrap unless N is small
      pos = pos < 0 ? pos + N : (pos >= N ? pos - N : pos);
      if (pos >= N) pos %= N;
      const float* s = xb + size_t(4 * min(cq[u], c4n - 1)) * N + pos;
      v[u] = v4f{s[0], s[N], s[2 * size_t(N)], s[3 * size_t(N)]};
    }
  };
  auto stage = [&](const v4f* v) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (tid + 256 * u < items) *reinterpret_cast<v4f*>(xs + jr[u] * CS + 4 * cq[u]) = v[u];
  };
  // Epilogue straight from the accumulators: lane (lr, q) of tile (m, t)
  // holds D[16m + 4q + r][16 (wave PW + t) + lr], so a store writes 16
  // consecutive positions of four channels; mul is read the same way, all
  // of a lane's values in flight at once (clamped addresses, unused past
  // the edges).  Staging the tile through LDS for whole-row stores cost
  // two barriers per job and measured 4-7% slower.
  // element (m, t, r) of this lane in job `job` (row length N, first
  // position i0, output base ybase): its address, or false past the edges
  auto addr = [&](int job, int N, int i0, size_t ybase, int m, int t, int r,
                  size_t& at) __attribute__((always_inline)) {
    const int o = m * 16 + 4 * q + r, i = i0 + p0 + 16 * t + lr;
    if constexpr (PACK) {
      const int b = job * cb.rpj + (p0 + lr) / Nu;
      at = (size_t(b) * cb.yC + cb.o0 + o) * Nu + (p0 + lr) % Nu;
      return o < Co && b < cb.rows;
    } else {
      at = ybase + size_t(o) * N + i;
      return o < Co && i < N;
    }
  };
  // Activation and stores of a job's values, the activation chosen once
  // (a branch per element would leave the compiler unable to count the
  // loads still in flight at each element, and it would wait for all).
  auto store_all = [&](float (*v)[PW][4], int job, int N, int i0, size_t ybase) __attribute__((always_inline)) {
    auto each = [&](auto&& f) __attribute__((always_inline)) {
#pragma unroll
      for (int m = 0; m < CT; ++m)
#pragma unroll
        for (int t = 0; t < PW; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            size_t at;
            if (addr(job, N, i0, ybase, m, t, r, at)) f(at, v[m][t][r]);
          }
    };
    if (act == INC_ACT_GELU && dy) {
      each([&](size_t at, float x) __attribute__((always_inline)) {
        float gv;
        dy[at] = gelu_and_grad(x, gv);
        y[at] = gv;
      });
    } else if (act == INC_ACT_GELU) {
      each([&](size_t at, float x) __attribute__((always_inline)) { y[at] = gelu_value(x); });
    } else if (act == INC_ACT_RELU && dy) {
      each([&](size_t at, float x) __attribute__((always_inline)) {
        dy[at] = relu_grad(x);
        y[at] = relu(x);
      });
    } else if (act == INC_ACT_RELU) {
      each([&](size_t at, float x) __attribute__((always_inline)) { y[at] = relu(x); });
    } else {
      each([&](size_t at, float x) __attribute__((always_inline)) { y[at] = x; });
    }
  };
  // the same address clamped into the output (for loads whose values past
  // the edges go unused)
  auto clamped = [&](int job, int N, int i0, size_t ybase, int m, int t, int r) __attribute__((always_inline)) {
    const int o = min(m * 16 + 4 * q + r, Co - 1), i = min(i0 + p0 + 16 * t + lr, N - 1);
    if constexpr (PACK) {
      const int b = min(job * cb.rpj + (p0 + lr) / Nu, cb.rows - 1);
      return (size_t(b) * cb.yC + cb.o0 + o) * Nu + (p0 + lr) % Nu;
    }
    return ybase + size_t(o) * N + i;
  };
  // the whole reduction in this launch (phase 0): bias, mul, activation.
  // Its loads come before the job's matrix work and reach here long done
  // (a load issued after a store is waited for together with it: one
  // in-order counter for both): this lane's bias values once per block,
  // and the job's mul values (1 without mul) in pmv.
  float bv[CT][4];
#pragma unroll
  for (int m = 0; m < CT; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r) bv[m][r] = bias ? bias[min(m * 16 + 4 * q + r, Co - 1)] : 0.f;
  float pmv[CT][PW][4];
  auto load_mul = [&](int job, int N, int i0, size_t ybase) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < CT; ++m)
#pragma unroll
      for (int t = 0; t < PW; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) pmv[m][t][r] = mul ? mul[clamped(job, N, i0, ybase, m, t, r)] : 1.f;
  };
  auto epilogue = [&](const v4f (*ac)[PW], const float (*mv)[PW][4], int job, int N, int i0,
                      size_t ybase) __attribute__((always_inline)) {
    float v[CT][PW][4];
#pragma unroll
    for (int m = 0; m < CT; ++m)
#pragma unroll
      for (int t = 0; t < PW; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) v[m][t][r] = (ac[m][t][r] + bv[m][r]) * mv[m][t][r];
    store_all(v, job, N, i0, ybase);
  };
  // A job's phase-0 epilogue runs in the next job, after that job's tile is
  // staged: its stores (a lane stores only inside the edges, a count the
  // compiler cannot know) then drain during the next job's matrix work,
  // where the wait for every outstanding access at the top of each job
  // would otherwise wait for them.  pacc and the p* fields hold the pending
  // job (pjob < 0: none).
  v4f pacc[CT][PW];
  int pjob = -1, pN = 0, pi0 = 0;
  size_t pybase = 0;
  v4f nx[U];
  if (int(blockIdx.x) < njobs) fetch(blockIdx.x, nx);
  for (int job = blockIdx.x; job < njobs; job += gridDim.x) {
    const JobPos jp = job_pos(rt, job, tiles, TI, Nu);
    const int N = jp.n, i0 = jp.i0;
    const size_t ybase = size_t(jp.base) * cb.yC + size_t(cb.o0) * N;
    __syncthreads();  // the weights are staged and the previous tile consumed
    stage(nx);
    __syncthreads();
    if (pjob >= 0) epilogue(pacc, pmv, pjob, pN, pi0, pybase);
    // unconditional for the same reason as the loads (the last job fetches
    // its own tile again, unused)
    fetch(min(job + int(gridDim.x), njobs - 1), nx);
    if (cb.phase == 0) load_mul(job, N, i0, ybase);
    v4f acc[CT][PW];
#pragma unroll
    for (int m = 0; m < CT; ++m)
#pragma unroll
      for (int t = 0; t < PW; ++t) acc[m][t] = v4f{0.f, 0.f, 0.f, 0.f};
    // chunk s: the xs offset of its tap and first channel, k*CS + c0 (wave-uniform)
    int c0 = 0, xo = 0;
    auto next = [&]() __attribute__((always_inline)) {
      c0 += 16;
      xo += 16;
      if (c0 == Ci) {
        c0 = 0;
        xo += CS - Ci;
      }
    };
    const float* wl = ws + lr * RS + 4 * q;
    const float* xl = xs + (PACK ? ((p0 + lr) / Nu) * XR + (p0 + lr) % Nu : p0 + lr) * CS + 4 * q;
    auto load = [&](int s, v4f* fa, v4f* fb) __attribute__((always_inline)) {
#pragma unroll
      for (int m = 0; m < CT; ++m) fa[m] = *reinterpret_cast<const v4f*>(wl + m * 16 * RS + 16 * s);
#pragma unroll
      for (int t = 0; t < PW; ++t) fb[t] = *reinterpret_cast<const v4f*>(xl + 16 * t * CS + xo);
    };
    auto mma = [&](const v4f* fa, const v4f* fb) __attribute__((always_inline)) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int m = 0; m < CT; ++m)
#pragma unroll
          for (int t = 0; t < PW; ++t)
            acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[m][j], fb[t][j], acc[m][t], 0, 0, 0);
    };
    // The reads are unconditional (a conditional read defeats the compiler's
    // count of LDS reads in flight, and each MFMA would then wait for the
    // prefetch): after the last chunk they read the ws row that follows and
    // the spare xs row XW, and the values are never used.  The scheduling
    // barriers keep each prefetch ahead of the MFMAs it overlaps (the
    // scheduler otherwise sinks it to its first use).
    v4f a0[CT], b0[PW], a1[CT], b1[PW];
    load(0, a0, b0);
    for (int s = 0; s < nchunk; s += 2) {
      next();
      load(s + 1, a1, b1);
      __builtin_amdgcn_sched_barrier(0);
      mma(a0, b0);
      __builtin_amdgcn_sched_barrier(0);
      if (s + 1 >= nchunk) break;
      next();
      load(s + 2, a0, b0);
      __builtin_amdgcn_sched_barrier(0);
      mma(a1, b1);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (cb.phase == 0) {
#pragma unroll
      for (int m = 0; m < CT; ++m)
#pragma unroll
        for (int t = 0; t < PW; ++t) pacc[m][t] = acc[m][t];
      pjob = job;
      pN = N;
      pi0 = i0;
      pybase = ybase;
    } else {  // one input-channel block of several, accumulated into y in stream order
      // the loads first (the running sum in y or the bias, and mul; clamped
      // addresses past the edges, values unused), then the stores
      float pv[CT][PW][4], mv[CT][PW][4];
#pragma unroll
      for (int m = 0; m < CT; ++m)
#pragma unroll
        for (int t = 0; t < PW; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const size_t at = clamped(job, N, i0, ybase, m, t, r);
            pv[m][t][r] = cb.phase == 1 ? (bias ? bias[min(m * 16 + 4 * q + r, Co - 1)] : 0.f) : y[at];
            mv[m][t][r] = (cb.phase == 3 && mul) ? mul[at] : 1.f;
          }
      float v[CT][PW][4];
#pragma unroll
      for (int m = 0; m < CT; ++m)
#pragma unroll
        for (int t = 0; t < PW; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) v[m][t][r] = (acc[m][t][r] + pv[m][t][r]) * mv[m][t][r];
      if (cb.phase == 3) {  // the last block: multiplier and activation
        store_all(v, job, N, i0, ybase);
      } else {  // the first block (with the bias) or a middle one
#pragma unroll
        for (int m = 0; m < CT; ++m)
#pragma unroll
          for (int t = 0; t < PW; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              size_t at;
              if (addr(job, N, i0, ybase, m, t, r, at)) y[at] = v[m][t][r];
            }
      }
    }
  }
  if (pjob >= 0) epilogue(pacc, pmv, pjob, pN, pi0, pybase);
}

int cu_count() {
  static int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || n < 1)
      n = 256;
    return n;
  }();
  return cus;
}

// Matrix-core forward (or input gradient) when the shape fits: Ci a
// multiple of 16 up to 64, 16 <= Co <= 64 and the LDS tiles fit; INC_EUNSUPPORTED
// (without touching the error text) otherwise.  Positions per wavefront:
// 16*PW with PW = 2 when N > 64 and two blocks' LDS fit a CU, else 1.
// Blocks loop over jobs; the grid is what the CUs hold at once (the
// occupancy query: a block that started late would extend the tail).
// Resident blocks per CU of a kernel at a dynamic LDS size, queried once per
// (kernel, LDS size) and cached: the query is a runtime call, and a call
// that is being captured into a hipGraph then only enqueues its kernel.
int blocks_per_cu(const void* fn, size_t lds, int threads) {
  struct Entry {
    const void* fn;
    size_t lds;
    int per_cu;
  };
  static std::mutex mu;
  static Entry cache[64];
  static int used = 0;
  std::lock_guard<std::mutex> lock(mu);
  for (int i = 0; i < used; ++i)
    if (cache[i].fn == fn && cache[i].lds == lds) return cache[i].per_cu;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds) != hipSuccess) return 0;
  if (used < 64) cache[used++] = Entry{fn, lds, per_cu};
  return per_cu;
}

int conv_fwd_mfma_block(int B, int Ci, int Co, int N, int K, bool flip, const float* x, const float* w,
                        const float* bias, float* y, const float* mul, int act, float* dy, hipStream_t st,
                        RowTiles rt, int rjobs, ConvBlock cb) {
  if (Ci % 16 != 0 || Ci > 64 || Co < 16 || Co > 64) return INC_EUNSUPPORTED;
  const int CT = (Co + 15) / 16;
  // rows of 16 or 32 points: 64 / n whole rows per 64-position tile
  const bool pack = !rt.tile && (N == 16 || N == 32);
  cb.rpj = pack ? 64 / N : 1;
  cb.rows = B;
  auto lds_of = [&](int pw) {  // staged weights and the input tile (with its spare row)
    const size_t xrows = pack ? size_t(cb.rpj) * (N + K - 1) + 1 : size_t(64 * pw + K);
    return (size_t(CT) * 16 * (Ci * K + 4) + xrows * (Ci + 4)) * sizeof(float);
  };
  // ragged rows come as kRowTile = 128-position tiles: two 16-position groups per wavefront
  const int PW = rt.tile ? 2 : ((N > 64 && lds_of(2) <= 80 * 1024) ? 2 : 1);
  const size_t lds = lds_of(PW);
  if (lds > 150 * 1024) return INC_EUNSUPPORTED;
  const int64_t jobs = rt.tile ? rjobs
                               : (pack ? (int64_t(B) + cb.rpj - 1) / cb.rpj : int64_t(B) * ((N + 64 * PW - 1) / (64 * PW)));
  const void* fn;
  if (pack) {
    switch (CT) {
      case 1: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<1, 1, true>); break;
      case 2: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<2, 1, true>); break;
      case 3: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<3, 1, true>); break;
      default: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<4, 1, true>); break;
    }
  } else switch (CT * 2 + PW - 1) {
    case 2: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<1, 1>); break;
    case 3: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<1, 2>); break;
    case 4: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<2, 1>); break;
    case 5: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<2, 2>); break;
    case 6: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<3, 1>); break;
    case 7: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<3, 2>); break;
    case 8: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<4, 1>); break;
    default: fn = reinterpret_cast<const void*>(&conv_fwd_mfma_kernel<4, 2>); break;
  }
  const int per_cu = blocks_per_cu(fn, lds);
  if (per_cu < 1) return INC_EUNSUPPORTED;
  if (jobs < 1) return INC_OK;
  const int nblk = int(std::min<int64_t>(jobs, int64_t(cu_count()) * per_cu));
  int fl = flip ? 1 : 0;
  int nj = int(jobs);
  void* args[] = {&nj, &Ci, &Co, &N, &K, &fl, &x, &w, &bias, &y, &mul, &act, &dy, &rt, &cb};
  INC_HIP(hipLaunchKernel(fn, dim3(nblk), dim3(256), args, lds, st));
  return check_launch("conv1d (matrix cores)");
}

// Matrix-core forward (or input gradient).  A layer of more than 64 input
// or output channels runs as channel blocks: output blocks of equal size
// (<= 64, >= 16) are independent launches; input blocks of 64 (Ci a
// multiple of 16) accumulate into y in stream order, the last one applying
// mul and the activation.  INC_EUNSUPPORTED when a block does not fit.
int conv_fwd_mfma(int B, int Ci, int Co, int N, int K, bool flip, const float* x, const float* w,
                  const float* bias, float* y, const float* mul, int act, float* dy, hipStream_t st,
                  RowTiles rt, int rjobs) {
  if (Ci <= 64 && Co <= 64)
    return conv_fwd_mfma_block(B, Ci, Co, N, K, flip, x, w, bias, y, mul, act, dy, st, rt, rjobs,
                               ConvBlock{Ci, 0, Co, 0, 0, 1, B});
  if (Ci % 16 != 0 || Co < 16 || rt.tile) return INC_EUNSUPPORTED;
  const int nob = (Co + 63) / 64, cob = (Co + nob - 1) / nob;
  const int nib = (Ci + 63) / 64;
  for (int o0 = 0; o0 < Co; o0 += cob) {
    const int co = std::min(cob, Co - o0);
    for (int ib = 0; ib < nib; ++ib) {
      const int c0 = 64 * ib, ci = std::min(64, Ci - c0);
      const int phase = nib == 1 ? 0 : (ib == 0 ? 1 : (ib == nib - 1 ? 3 : 2));
      const int rc = conv_fwd_mfma_block(B, ci, co, N, K, flip, x, w, bias ? bias + o0 : nullptr, y, mul, act, dy, st,
                                         rt, rjobs, ConvBlock{Ci, c0, Co, o0, phase, 1, B});
      if (rc) return rc;
    }
  }
  return INC_OK;
}

template <typename T, int K, bool FLIP>
int conv_fwd_k(int cpt, dim3 grid, size_t lds, hipStream_t st, int Ci, int Co, int N, const T* x, const T* w,
               const T* b, T* y, const T* mul, int act, T* dy, RowTiles rt) {
  dim3 block(kWaves * kWave);
#define INC_CF(C) \
  hipLaunchKernelGGL((conv_fwd_kernel<T, K, C, FLIP>), grid, block, lds, st, Ci, Co, N, x, w, b, y, mul, act, dy, rt)
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
                    const void* mul, int act, void* dy, hipStream_t st, RowTiles rt = RowTiles{}, int rjobs = 0) {
  if constexpr (std::is_same<T, float>::value) {
    const int rc = conv_fwd_mfma(B, Ci, Co, N, K, FLIP, static_cast<const float*>(x), static_cast<const float*>(w),
                                 static_cast<const float*>(bias), static_cast<float*>(y),
                                 static_cast<const float*>(mul), act, static_cast<float*>(dy), st, rt, rjobs);
    if (rc != INC_EUNSUPPORTED) return rc;
  }
  const int TI = kWave * kPPT;
  size_t lds = size_t(Ci) * (TI + K - 1) * sizeof(T);
  INC_REQUIRE(lds <= 150 * 1024, INC_EUNSUPPORTED, "conv1d: %d input channels need %zu bytes of LDS", Ci, lds);
  const int cpt = conv_cpt(Co);
  const unsigned cg = (Co + kWaves * cpt - 1) / (kWaves * cpt);
  if (rt.tile && rjobs < 1) return INC_OK;
  dim3 grid = rt.tile ? dim3(rjobs, cg, 1) : dim3((N + TI - 1) / TI, cg, B);
  const T* xp = static_cast<const T*>(x);
  const T* wp = static_cast<const T*>(w);
  const T* bp = static_cast<const T*>(bias);
  const T* mp = static_cast<const T*>(mul);
  T* yp = static_cast<T*>(y);
  T* dp = static_cast<T*>(dy);
  switch (K) {
    case 1: return conv_fwd_k<T, 1, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp, rt);
    case 3: return conv_fwd_k<T, 3, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp, rt);
    case 5: return conv_fwd_k<T, 5, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp, rt);
    case 7: return conv_fwd_k<T, 7, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp, rt);
    case 9: return conv_fwd_k<T, 9, FLIP>(cpt, grid, lds, st, Ci, Co, N, xp, wp, bp, yp, mp, act, dp, rt);
    default: set_error("conv1d kernel size must be 1, 3, 5, 7 or 9 (got %d)", K); return INC_EUNSUPPORTED;
  }
}

template <typename T, int K>
int wgrad_launch_k(int njobs, int Ci, int Co, int N, int TI, int nblk, size_t lds, const T* g, const T* x, T* part,
                   hipStream_t st, int accumulate, RowTiles rt, const ChanBlock& cb) {
  const int nblk2 = ((Co + 1) / 2) * ((Ci + 1) / 2);
  const int maxp = (nblk2 + 255) / 256;
  dim3 grid(nblk), block(256);
#define INC_WG(MP)                                                                                            \
  hipLaunchKernelGGL((conv_wgrad_kernel<T, K, MP>), grid, block, lds, st, njobs, Ci, Co, N, TI, g, x, part, \
                     accumulate, rt, cb)
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

// ---- weight gradient of a layer with one channel on one side ----------------
// The network's first layer (1 -> C) and last layer (C -> 1): C*K + C (or
// C*K + 1) entries summed over every position, a streaming reduction that
// reads each gradient and input value once (HBM-bound), without matrix-core
// tiles or LDS staging.  Wavefront w of the 8 owns channels [w*CPW, w*CPW +
// CPW) of the wide side; its 64 lanes take consecutive positions of a (row,
// 128-position) job, so every load is coalesced (the K shifted input reads
// of a lane hit the lines its neighbours load):
//   ONE_IN (Ci = 1): acc[c][k] += g[c][i] x[(i + k - P) mod n], bias[c] += g[c][i]
//   else   (Co = 1): acc[c][k] += g[i] x[c][(i + k - P) mod n], bias += g[i]
// A lane keeps its sums over all of the block's jobs; at the end the 64
// lanes of a wavefront are summed by a shuffle tree (a fixed order) and the
// block writes, or adds to, its partial row as the other weight-gradient
// kernels do.
constexpr int kThinWaves = 8;

template <typename T, int CPW, bool ONE_IN, int KM>
__global__ void __launch_bounds__(64 * kThinWaves) conv_wgrad_thin_kernel(int njobs, int C, int Nu, int K,
                                                                       const T* __restrict__ g,
                                                                       const T* __restrict__ x,
                                                                       T* __restrict__ part, int accumulate,
                                                                       RowTiles rt) {
  const int lane = threadIdx.x & 63;
  const int c0 = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)) * CPW;
  const int P = K / 2;
  T acc[CPW][KM], bacc[CPW];
#pragma unroll
  for (int cc = 0; cc < CPW; ++cc) {
    bacc[cc] = T(0);
#pragma unroll
    for (int k = 0; k < KM; ++k) acc[cc][k] = T(0);
  }
  constexpr int H = kRowTile / 64;  // positions per lane and job
  const int tiles = (Nu + kRowTile - 1) / kRowTile;
  for (int job = blockIdx.x; job < njobs; job += gridDim.x) {
    const JobPos jp = job_pos(rt, job, tiles, kRowTile, Nu);
    const int N = jp.n;
    const T* gb = g + jp.base * (ONE_IN ? C : 1);
    const T* xb = x + jp.base * (ONE_IN ? 1 : C);
    // The one-channel side is read as a K-point window per position, the
    // wide side once per channel and position.  Written out, dW[c][k] sums
    // g[c][i] x[i + k - P] (ONE_IN) or x[c][j] g[j - k + P] (with j = i + k -
    // P, every position once per row), so the window is x at i + k - P or
    // g at i - k + P.  Every load is unconditional (positions past the row
    // read its last point and are masked), so all of them are in flight at
    // once.
    int ic[H];
    bool ok[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const int i = jp.i0 + 64 * h + lane;
      ok[h] = i < N;
      ic[h] = ok[h] ? i : N - 1;
    }
    T wv[H][KM], cv[H][CPW];
#pragma unroll
    for (int h = 0; h < H; ++h) {
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        int q = ONE_IN ? ic[h] + k - P : ic[h] - k + P;  // in [-P, N + P): one wrap (K <= N)
        q = q < 0 ? q + N : (q >= N ? q - N : q);
        wv[h][k] = k < K ? (ONE_IN ? xb : gb)[q] : T(0);
      }
#pragma unroll
      for (int cc = 0; cc < CPW; ++cc)
        cv[h][cc] = c0 + cc < C ? (ONE_IN ? gb : xb)[size_t(c0 + cc) * N + ic[h]] : T(0);
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if constexpr (!ONE_IN) {
#pragma unroll
        for (int k = 0; k < KM; ++k) wv[h][k] = ok[h] ? wv[h][k] : T(0);
      }
      if (ONE_IN) {
#pragma unroll
        for (int cc = 0; cc < CPW; ++cc) {
          const T v = ok[h] ? cv[h][cc] : T(0);
          bacc[cc] += v;
#pragma unroll
          for (int k = 0; k < KM; ++k)
            if (k < K) acc[cc][k] += v * wv[h][k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < KM; ++k)
          if (k == P) bacc[0] += wv[h][k];  // g[i]
#pragma unroll
        for (int cc = 0; cc < CPW; ++cc)
#pragma unroll
          for (int k = 0; k < KM; ++k)
            if (k < K) acc[cc][k] += cv[h][cc] * wv[h][k];
      }
    }
  }
  const int width = C * K + (ONE_IN ? C : 1);
  T* out = part + size_t(blockIdx.x) * width;
#pragma unroll
  for (int cc = 0; cc < CPW; ++cc) {
#pragma unroll
    for (int k = 0; k < KM; ++k)
      if (k < K) acc[cc][k] = wave_sum(acc[cc][k], kWave);
    bacc[cc] = wave_sum(bacc[cc], kWave);
  }
  if (lane != 0) return;
#pragma unroll
  for (int cc = 0; cc < CPW; ++cc) {
    const int c = c0 + cc;
    if (c >= C) break;
#pragma unroll
    for (int k = 0; k < KM; ++k)
      if (k < K) {
        T* dst = out + c * K + k;
        *dst = accumulate ? *dst + acc[cc][k] : acc[cc][k];
      }
    if (ONE_IN) {
      T* dst = out + C * K + c;
      *dst = accumulate ? *dst + bacc[cc] : bacc[cc];
    }
  }
  if (!ONE_IN && c0 == 0) {
    T* dst = out + C * K;
    *dst = accumulate ? *dst + bacc[0] : bacc[0];
  }
}

// Weight gradient of a 1 -> C or C -> 1 layer (C <= 64) on the streaming
// kernel above.  INC_EUNSUPPORTED for other shapes.
template <typename T>
int conv_wgrad_thin(int B, int Ci, int Co, int N, int K, const T* g, const T* x, T* dw, T* db, T* part,
                    int64_t work_elems, hipStream_t st, const WgradMode& md) {
  const bool one_in = Ci == 1;
  const int C = one_in ? Co : Ci;
  if (!(Ci == 1 || Co == 1) || C > 8 * kThinWaves) return INC_EUNSUPPORTED;
  const int width = Co * Ci * K + Co;
  const int jobs = md.rt.tile ? md.njobs : B * ((N + kRowTile - 1) / kRowTile);
  const int64_t slots = std::min<int64_t>(512, work_elems / width);
  INC_REQUIRE(jobs < 1 || slots >= 1, INC_EINVAL, "conv1d weight gradient workspace too small (need >= %d elements)",
              width);
  const int nblk = int(std::min<int64_t>(jobs, slots));
  if (md.nparts) *md.nparts = std::max(*md.nparts, nblk);
  if (nblk < 1) return INC_OK;
  const int cpw = (C + kThinWaves - 1) / kThinWaves;
  const int acc = md.accumulate ? 1 : 0;
  dim3 grid(nblk), block(64 * kThinWaves);
  // taps held per lane: 7 covers K <= 7 (fewer registers, more resident blocks), 9 the rest
#define INC_WT_K(CPW, KM)                                                                                           \
  if (one_in)                                                                                                       \
    hipLaunchKernelGGL((conv_wgrad_thin_kernel<T, CPW, true, KM>), grid, block, 0, st, jobs, C, N, K, g, x, part,    \
                       acc, md.rt);                                                                                 \
  else                                                                                                              \
    hipLaunchKernelGGL((conv_wgrad_thin_kernel<T, CPW, false, KM>), grid, block, 0, st, jobs, C, N, K, g, x, part,   \
                       acc, md.rt)
#define INC_WT(CPW)          \
  do {                       \
    if (K <= 7) {            \
      INC_WT_K(CPW, 7);      \
    } else {                 \
      INC_WT_K(CPW, 9);      \
    }                        \
  } while (0)
  if (cpw <= 1) INC_WT(1);
  else if (cpw <= 2) INC_WT(2);
  else if (cpw <= 3) INC_WT(3);
  else if (cpw <= 4) INC_WT(4);
  else if (cpw <= 6) INC_WT(6);
  else INC_WT(8);
#undef INC_WT
#undef INC_WT_K
  int rc = check_launch("conv1d weight gradient (one-channel side)");
  if (rc || !md.reduce) return rc;
  hipLaunchKernelGGL(reduce_parts_kernel<T>, dim3((width + 15) / 16), dim3(256), 0, st, nblk, width, part, dw, db,
                     Co * Ci * K, Ci * K, Ci * K);
  return check_launch("conv1d weight gradient reduction");
}

template <typename T>
int conv_wgrad_block(int B, int Ci, int Co, int N, int K, const void* g, const void* x, void* dw, void* db, void* work,
                     int64_t work_elems, hipStream_t st, const WgradMode& md, const ChanBlock& cb) {
  if (cb.o0 == 0 && cb.c0 == 0 && cb.gC == Co && cb.xC == Ci) {
    const int rc = conv_wgrad_thin<T>(B, Ci, Co, N, K, static_cast<const T*>(g), static_cast<const T*>(x),
                                      static_cast<T*>(dw), static_cast<T*>(db), static_cast<T*>(work), work_elems, st,
                                      md);
    if (rc != INC_EUNSUPPORTED) return rc;
  }
  if constexpr (std::is_same<T, float>::value) {
    const int rc = conv_wgrad_mfma(B, Ci, Co, N, K, static_cast<const float*>(g), static_cast<const float*>(x),
                                   static_cast<float*>(dw), static_cast<float*>(db), static_cast<float*>(work),
                                   work_elems, st, md, cb);
    if (rc != INC_EUNSUPPORTED) return rc;
  }
  int TI = md.rt.tile ? kRowTile : (N < 128 ? N : 128);
  size_t lds = (size_t((Co + 1) & ~1) * (TI | 1) + size_t((Ci + 1) & ~1) * ((TI + K - 1) | 1)) * sizeof(T);
  INC_REQUIRE(lds <= 150 * 1024, INC_EUNSUPPORTED, "conv1d weight gradient needs %zu bytes of LDS", lds);
  const int width = Co * Ci * K + Co;
  const int jobs = md.rt.tile ? md.njobs : B * ((N + TI - 1) / TI);
  int nblk = jobs < 256 ? jobs : 256;  // about one block per CU, each looping over several tiles
  if (int64_t(nblk) * width > work_elems) nblk = int(work_elems / width);
  INC_REQUIRE(jobs < 1 || nblk >= 1, INC_EINVAL, "conv1d weight gradient workspace too small (need >= %d elements)",
              width);
  if (md.nparts) *md.nparts = std::max(*md.nparts, nblk);
  if (nblk < 1) return INC_OK;
  const T* gp = static_cast<const T*>(g);
  const T* xp = static_cast<const T*>(x);
  T* part = static_cast<T*>(work);
  const int acc = md.accumulate ? 1 : 0;
  int rc;
  switch (K) {
    case 1: rc = wgrad_launch_k<T, 1>(jobs, Ci, Co, N, TI, nblk, lds, gp, xp, part, st, acc, md.rt, cb); break;
    case 3: rc = wgrad_launch_k<T, 3>(jobs, Ci, Co, N, TI, nblk, lds, gp, xp, part, st, acc, md.rt, cb); break;
    case 5: rc = wgrad_launch_k<T, 5>(jobs, Ci, Co, N, TI, nblk, lds, gp, xp, part, st, acc, md.rt, cb); break;
    case 7: rc = wgrad_launch_k<T, 7>(jobs, Ci, Co, N, TI, nblk, lds, gp, xp, part, st, acc, md.rt, cb); break;
    case 9: rc = wgrad_launch_k<T, 9>(jobs, Ci, Co, N, TI, nblk, lds, gp, xp, part, st, acc, md.rt, cb); break;
    default: set_error("conv1d kernel size must be 1, 3, 5, 7 or 9 (got %d)", K); return INC_EUNSUPPORTED;
  }
  if (rc || !md.reduce) return rc;
  hipLaunchKernelGGL(reduce_parts_kernel<T>, dim3((width + 15) / 16), dim3(256), 0, st, nblk, width, part,
                     static_cast<T*>(dw), static_cast<T*>(db), Co * Ci * K, Ci * K, cb.xC * K);
  return check_launch("conv1d weight gradient reduction");
}

// Weight and bias gradients; a layer with more channels than one launch
// holds (accumulators in registers, tiles in LDS) runs as blocks of at most
// 64 output x 64 input channels, each reduced into its part of dw (the
// blocks are ordered on the stream, so they share the workspace; the bias
// gradient comes from the blocks of the first input channels).
template <typename T>
int conv_wgrad(int B, int Ci, int Co, int N, int K, const void* g, const void* x, void* dw, void* db, void* work,
               int64_t work_elems, hipStream_t st, const WgradMode& md = WgradMode{RowTiles{}, 0, false, true, nullptr}) {
  int rc = conv_wgrad_block<T>(B, Ci, Co, N, K, g, x, dw, db, work, work_elems, st, md, ChanBlock{Co, Ci, 0, 0});
  if (rc != INC_EUNSUPPORTED || md.accumulate || !md.reduce || (Co <= 64 && Ci <= 64)) return rc;
  constexpr int kB = 64;
  T* dwp = static_cast<T*>(dw);
  T* dbp = static_cast<T*>(db);
  for (int o0 = 0; o0 < Co; o0 += kB)
    for (int c0 = 0; c0 < Ci; c0 += kB) {
      rc = conv_wgrad_block<T>(B, std::min(kB, Ci - c0), std::min(kB, Co - o0), N, K, g, x,
                               dwp + (size_t(o0) * Ci + c0) * K, c0 == 0 && dbp ? dbp + o0 : nullptr, work, work_elems,
                               st, md, ChanBlock{Co, Ci, o0, c0});
      if (rc) return rc;
    }
  clear_error();
  return INC_OK;
}

int conv_check(int dtype, int act, int B, int Ci, int Co, int N, int K) {
  INC_REQUIRE(dtype == INC_F32 || dtype == INC_F64, INC_EINVAL, "unsupported dtype %d", dtype);
  INC_REQUIRE(act == INC_ACT_NONE || act == INC_ACT_GELU || act == INC_ACT_RELU, INC_EINVAL, "unknown activation %d",
              act);
  INC_REQUIRE(B >= 0 && Ci >= 1 && Co >= 1 && N >= 1, INC_EINVAL, "conv1d: bad sizes");
  INC_REQUIRE(K % 2 == 1 && K <= 9 && K <= N, INC_EUNSUPPORTED, "conv1d kernel size must be odd, <= 9 and <= n");
  return INC_OK;
}

int rows_check(int dtype, int act, const inc_conv_rows* r, int Ci, int Co, int K) {
  INC_REQUIRE(r != nullptr, INC_EINVAL, "null rows descriptor");
  INC_REQUIRE(dtype == INC_F32 || dtype == INC_F64, INC_EINVAL, "unsupported dtype %d", dtype);
  INC_REQUIRE(act == INC_ACT_NONE || act == INC_ACT_GELU || act == INC_ACT_RELU, INC_EINVAL, "unknown activation %d",
              act);
  INC_REQUIRE(r->count >= 0 && r->njobs >= 0 && Ci >= 1 && Co >= 1, INC_EINVAL, "conv1d rows: bad sizes");
  INC_REQUIRE(K % 2 == 1 && K <= 9, INC_EUNSUPPORTED, "conv1d kernel size must be odd and <= 9");
  INC_REQUIRE(r->njobs == 0 || (r->tile && r->n && r->offset), INC_EINVAL, "conv1d rows: tile, n, offset missing");
  return INC_OK;
}

RowTiles row_tiles(const inc_conv_rows* r) { return RowTiles{r->tile, r->n, r->offset}; }

// ---- the INC rollout of a ragged batch --------------------------------------------
// out = a x + y + z (y, z optional), elementwise over n values.
template <typename T>
__global__ void __launch_bounds__(256) lincomb_kernel(int64_t n, T a, const T* __restrict__ x,
                                                     const T* __restrict__ y, const T* __restrict__ z,
                                                     T* __restrict__ out) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
    T v = a * x[i];
    if (y) v += y[i];
    if (z) v += z[i];
    out[i] = v;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) fill_kernel(int64_t n, T v, T* __restrict__ out) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) out[i] = v;
}

template <typename T>
int lincomb(int64_t n, T a, const T* x, const T* y, const T* z, T* out, hipStream_t st) {
  if (n <= 0) return INC_OK;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(lincomb_kernel<T>, dim3(unsigned(blocks)), dim3(256), 0, st, n, a, x, y, z, out);
  return check_launch("ragged INC elementwise");
}

template <typename T>
int fill(int64_t n, T v, T* out, hipStream_t st) {
  if (n <= 0) return INC_OK;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_kernel<T>, dim3(unsigned(blocks)), dim3(256), 0, st, n, v, out);
  return check_launch("ragged INC fill");
}

// Partial-sum slots per layer for the accumulated weight gradients (the
// matrix-core kernel uses at most 512, the vector-ALU kernel 256).
constexpr int kIncSlots = 512;

// Workspace of an INC rollout over `tot` state values (the packed ragged
// batch, or batch x n), in elements of the state's type.  forward: two
// activation buffers [C][tot], the correction [tot] and the scaled last
// layer (C K + 1); adjoint: the recomputed activations and their GELU
// derivatives ((L - 1) each, not with stored activations), two cotangent
// buffers [C][tot], four [tot] vectors, the scaled last layer and every
// layer's partial sums.
enum IncMode { kIncForward = 0, kIncAdjoint = 1, kIncAdjointStored = 2 };

struct IncLayout {
  int64_t act[2], s, wl, y[8], dy[8], g[2], lam_s, lam, lam_sol, gx, part[8], total;
};

IncLayout inc_layout(int64_t tot, const inc_cnn_desc* cnn, int mode) {
  IncLayout L{};
  const int64_t C = cnn->channels, K = cnn->kernel;
  int64_t o = 0;
  auto take = [&](int64_t n) {
    const int64_t at = o;
    o += (n + 63) / 64 * 64;  // 256- or 512-byte aligned pieces
    return at;
  };
  L.wl = take(C * K + 1);
  if (mode == kIncForward) {
    L.act[0] = take(C * tot);
    L.act[1] = take(C * tot);
    L.s = take(tot);
  } else {
    if (mode == kIncAdjoint)
      for (int l = 0; l < cnn->layers - 1; ++l) {
        L.y[l] = take(C * tot);
        L.dy[l] = take(C * tot);
      }
    L.g[0] = take(C * tot);
    L.g[1] = take(C * tot);
    L.lam_s = take(tot);
    L.lam = take(tot);
    L.lam_sol = take(tot);
    L.gx = take(tot);
    for (int l = 0; l < cnn->layers; ++l) {
      const int64_t ci = l == 0 ? 1 : C, co = l == cnn->layers - 1 ? 1 : C;
      L.part[l] = take(kIncSlots * (co * ci * K + co));
    }
  }
  L.total = o;
  return L;
}

// The workspace kind of the *_inc_workspace calls: 0 forward, 1 adjoint
// evaluating the network again, 2 adjoint reading stored activations.
int inc_mode_check(int kind) {
  INC_REQUIRE(kind >= kIncForward && kind <= kIncAdjointStored, INC_EINVAL,
              "workspace kind must be 0, 1 or 2 (got %d)", kind);
  return INC_OK;
}

int cnn_check(const inc_cnn_desc* cnn) {
  INC_REQUIRE(cnn->layers >= 2 && cnn->layers <= 8, INC_EUNSUPPORTED, "the network needs 2..8 layers (got %d)",
              cnn->layers);
  INC_REQUIRE(cnn->channels >= 1 && cnn->channels <= 64, INC_EUNSUPPORTED, "network width must be 1..64");
  INC_REQUIRE(cnn->kernel % 2 == 1 && cnn->kernel <= 9, INC_EUNSUPPORTED, "kernel size must be odd and <= 9");
  for (int l = 0; l < cnn->layers; ++l) INC_REQUIRE(cnn->weight[l] && cnn->bias[l], INC_EINVAL, "missing layer %d", l);
  return INC_OK;
}

int inc_check(const inc_ragged_desc* d, const inc_cnn_desc* cnn, const inc_conv_rows* r) {
  INC_REQUIRE(d && cnn && r, INC_EINVAL, "null descriptor");
  INC_REQUIRE(d->dtype == INC_F32 || d->dtype == INC_F64, INC_EINVAL, "unsupported dtype %d", d->dtype);
  const int rc = cnn_check(cnn);
  if (rc) return rc;
  INC_REQUIRE(2 * d->nmin >= cnn->kernel, INC_EUNSUPPORTED, "every trajectory needs >= K/2 points");
  INC_REQUIRE(r->count == d->count, INC_EINVAL, "rows (%d) and trajectories (%d) differ", r->count, d->count);
  return rows_check(d->dtype, INC_ACT_GELU, r, 1, 1, cnn->kernel);
}

int ks_inc_check(const inc_ks_desc* d, const inc_cnn_desc* cnn) {
  INC_REQUIRE(d && cnn, INC_EINVAL, "null descriptor");
  INC_REQUIRE(d->dtype == INC_F32 || d->dtype == INC_F64, INC_EINVAL, "unsupported dtype %d", d->dtype);
  INC_REQUIRE(d->batch >= 0 && d->substeps >= 1, INC_EINVAL, "bad KS descriptor");
  const int rc = cnn_check(cnn);
  if (rc) return rc;
  INC_REQUIRE(2 * d->n >= cnn->kernel, INC_EUNSUPPORTED, "the resolution needs >= K/2 points");
  return INC_OK;
}

int ks_stages(const inc_ks_desc* d) {
  return (d->scheme == INC_ETD1 ? 1 : (d->scheme == INC_ETRK2 ? 2 : 4)) * d->substeps;
}

// Where the network runs at step k: the ragged rows (the first active[k]
// tiles of rt, every tile when active is NULL) or a uniform batch of B rows
// of n points (rt empty).
struct NetRows {
  RowTiles rt;
  int B, n;
  const int32_t* active;
  int njobs;
  int jobs(int k) const { return rt.tile ? (active ? active[k] : njobs) : 1; }
};

// The network on the state `in` at step k: hidden activations to y[l] (with
// their GELU derivatives to dy[l] when dy is given), the scaled correction
// to s.
template <typename T>
int inc_network(const inc_cnn_desc* cnn, const T* wl, const T* in, T* const* y, T* const* dy, T* s, const NetRows& nr,
                int jobs, hipStream_t st) {
  const int C = cnn->channels, K = cnn->kernel, nl = cnn->layers;
  const T* x = in;
  for (int l = 0; l < nl - 1; ++l) {
    const int ci = l == 0 ? 1 : C;
    const int rc = conv_fwd_launch<T, false>(nr.B, ci, C, nr.n, K, x, cnn->weight[l], cnn->bias[l], y[l], nullptr,
                                             INC_ACT_GELU, dy ? dy[l] : nullptr, st, nr.rt, jobs);
    if (rc) return rc;
    x = y[l];
  }
  if (s == nullptr) return INC_OK;  // the adjoint's recomputation needs the hidden layers only
  return conv_fwd_launch<T, false>(nr.B, C, 1, nr.n, K, x, wl, wl + C * K, s, nullptr, INC_ACT_NONE, nullptr, st,
                                   nr.rt, jobs);
}

// The solver step of an INC rollout, one step at a time on step-major
// buffers: the ragged Burgers batch (its forcing clock at step k) ...
struct RaggedStep {
  inc_ragged_desc d;
  int64_t tot;
  int ns;
  int fwd(int k, const void* in, const void* s, void* out, void* stages, hipStream_t st) {
    d.step0 = k;
    return inc_burgers_ragged_fwd(&d, in, s, out, 1, nullptr, stages, st);
  }
  int adj(int k, const void* stages, const void* lam_in, void* lam_sol, void* lam_s, hipStream_t st) {
    d.step0 = k;
    return inc_burgers_ragged_adj(&d, stages, lam_in, lam_sol, lam_s, 1, st);
  }
};

// ... or the KS batch ([batch][n] per step, stages [batch][nst][n]).
struct KsStep {
  const inc_ks_desc* d;
  int64_t tot;
  int ns;
  int fwd(int, const void* in, const void* s, void* out, void* stages, hipStream_t st) {
    return inc_ks_fwd(d, in, s, out, 1, nullptr, stages, st);
  }
  int adj(int, const void* stages, const void* lam_in, void* lam_sol, void* lam_s, hipStream_t st) {
    return inc_ks_adj(d, stages, lam_in, lam_sol, lam_s, 1, st);
  }
};

// Stored activations of an INC rollout: per step and hidden layer l the
// activation and its GELU derivative, C x tot values each, at
// acts + ((k (L - 1) + l) 2 + {0, 1}) C tot.
template <typename T>
T* stored_act(T* acts, int k, int l, int which, const inc_cnn_desc* cnn, int64_t tot) {
  const int64_t ct = int64_t(cnn->channels) * tot;
  return acts + ((int64_t(k) * (cnn->layers - 1) + l) * 2 + which) * ct;
}

// acts (optional): the forward also writes every step's hidden activations
// and GELU derivatives there (the per-step slots replace the two reused
// activation buffers, so the only extra traffic is the derivatives), and
// the adjoint reads them instead of evaluating the network again.
template <typename T, typename Step>
int inc_fwd(Step& sv, const inc_cnn_desc* cnn, const NetRows& nr, const T* u0, int steps, T* traj, T* corr,
            T* stages, T* acts, T* work, hipStream_t st) {
  const int64_t tot = sv.tot;
  const IncLayout L = inc_layout(tot, cnn, kIncForward);
  const int C = cnn->channels, K = cnn->kernel, nl = cnn->layers;
  const int ns = sv.ns;
  T* wl = work + L.wl;
  int rc = lincomb<T>(int64_t(C) * K, T(cnn->scale), static_cast<const T*>(cnn->weight[nl - 1]), nullptr, nullptr,
                      wl, st);
  if (!rc) rc = lincomb<T>(1, T(cnn->scale), static_cast<const T*>(cnn->bias[nl - 1]), nullptr, nullptr, wl + C * K,
                           st);
  if (!rc && corr) rc = fill<T>(int64_t(steps) * tot, T(0), corr, st);
  if (rc) return rc;
  T* bufs[8];
  T* dbufs[8];
  for (int l = 0; l < nl - 1; ++l) bufs[l] = work + L.act[l & 1];
  const T* in = u0;
  for (int k = 0; k < steps; ++k) {
    const int jobs = nr.jobs(k);
    T* s = corr ? corr + int64_t(k) * tot : work + L.s;
    if (acts)
      for (int l = 0; l < nl - 1; ++l) {
        bufs[l] = stored_act(acts, k, l, 0, cnn, tot);
        dbufs[l] = stored_act(acts, k, l, 1, cnn, tot);
      }
    rc = inc_network<T>(cnn, wl, in, bufs, acts ? dbufs : nullptr, s, nr, jobs, st);
    if (rc) return rc;
    T* out = traj + int64_t(k) * tot;
    rc = sv.fwd(k, in, s, out, stages ? stages + int64_t(k) * ns * tot : nullptr, st);
    if (rc) return rc;
    in = out;
  }
  return INC_OK;
}

// g_corr (optional, [steps][tot]): cotangent of the corrections, added to
// the solver's cotangent of each step's correction.  acts (optional): what
// inc_fwd stored there; without it the network is evaluated again at every
// step's input state.
template <typename T, typename Step>
int inc_adj(Step& sv, const inc_cnn_desc* cnn, const NetRows& nr, const T* u0, const T* traj, const T* stages,
            const T* acts, const T* g_traj, const T* g_corr, T* g_u0, T* g_w, int steps, T* work, hipStream_t st) {
  const int64_t tot = sv.tot;
  const IncLayout L = inc_layout(tot, cnn, acts ? kIncAdjointStored : kIncAdjoint);
  const int C = cnn->channels, K = cnn->kernel, nl = cnn->layers;
  const int ns = sv.ns;
  T* wl = work + L.wl;
  int rc = lincomb<T>(int64_t(C) * K, T(cnn->scale), static_cast<const T*>(cnn->weight[nl - 1]), nullptr, nullptr,
                      wl, st);
  if (!rc) rc = lincomb<T>(1, T(cnn->scale), static_cast<const T*>(cnn->bias[nl - 1]), nullptr, nullptr, wl + C * K,
                           st);
  // every layer's partial sums and the network's input cotangent start at
  // zero (rows past their horizon are never written by the network's passes)
  if (!rc) rc = fill<T>(L.total - L.part[0], T(0), work + L.part[0], st);
  if (!rc) rc = fill<T>(tot, T(0), work + L.gx, st);
  if (rc) return rc;
  T* y[8] = {};
  T* dy[8] = {};
  if (!acts)
    for (int l = 0; l < nl - 1; ++l) {
      y[l] = work + L.y[l];
      dy[l] = work + L.dy[l];
    }
  T* lam = work + L.lam;
  T* lam_s = work + L.lam_s;
  T* lam_sol = work + L.lam_sol;
  T* gx = work + L.gx;
  T* g[2] = {work + L.g[0], work + L.g[1]};
  const RowTiles rt = nr.rt;
  const int B = nr.B, n = nr.n;
  int nparts[8] = {};
  int64_t width[8];
  for (int l = 0; l < nl; ++l) {
    const int ci = l == 0 ? 1 : C, co = l == nl - 1 ? 1 : C;
    width[l] = int64_t(co) * ci * K + co;
  }
  const T* lam_in = g_traj + int64_t(steps - 1) * tot;  // cotangent of the last state
  for (int k = steps - 1; k >= 0; --k) {
    const int jobs = nr.jobs(k);
    const T* in = k == 0 ? u0 : traj + int64_t(k - 1) * tot;
    // the solver step's transpose: cotangents of its input state and of s
    rc = sv.adj(k, stages + int64_t(k) * ns * tot, lam_in, lam_sol, lam_s, st);
    if (!rc && g_corr) rc = lincomb<T>(tot, T(1), lam_s, g_corr + int64_t(k) * tot, nullptr, lam_s, st);
    if (rc) return rc;
    // the step's hidden activations with their GELU derivatives: stored by
    // the forward, or the network again at the step's input state
    if (acts) {
      for (int l = 0; l < nl - 1; ++l) {
        y[l] = const_cast<T*>(stored_act(acts, k, l, 0, cnn, tot));
        dy[l] = const_cast<T*>(stored_act(acts, k, l, 1, cnn, tot));
      }
    } else {
      rc = inc_network<T>(cnn, wl, in, y, dy, nullptr, nr, jobs, st);
      if (rc) return rc;
    }
    if (jobs > 0) {
      // last layer (C -> 1): weight gradient of lam_s x y[L-2] (scaled at the end), input cotangent
      // scale W^T lam_s times the GELU derivative
      WgradMode md{rt, jobs, true, false, &nparts[nl - 1]};
      rc = conv_wgrad<T>(B, C, 1, n, K, lam_s, y[nl - 2], nullptr, nullptr, work + L.part[nl - 1],
                         kIncSlots * width[nl - 1], st, md);
      if (!rc) rc = conv_fwd_launch<T, true>(B, 1, C, n, K, lam_s, wl, nullptr, g[0], dy[nl - 2], INC_ACT_NONE,
                                             nullptr, st, rt, jobs);
      int cur = 0;
      for (int l = nl - 2; l >= 1 && !rc; --l) {
        WgradMode mh{rt, jobs, true, false, &nparts[l]};
        rc = conv_wgrad<T>(B, C, C, n, K, g[cur], y[l - 1], nullptr, nullptr, work + L.part[l],
                           kIncSlots * width[l], st, mh);
        if (!rc) rc = conv_fwd_launch<T, true>(B, C, C, n, K, g[cur], cnn->weight[l], nullptr, g[cur ^ 1], dy[l - 1],
                                               INC_ACT_NONE, nullptr, st, rt, jobs);
        cur ^= 1;
      }
      if (!rc) {
        WgradMode m0{rt, jobs, true, false, &nparts[0]};
        rc = conv_wgrad<T>(B, 1, C, n, K, g[cur], in, nullptr, nullptr, work + L.part[0], kIncSlots * width[0], st,
                           m0);
      }
      if (!rc) rc = conv_fwd_launch<T, true>(B, C, 1, n, K, g[cur], cnn->weight[0], nullptr, gx, nullptr,
                                             INC_ACT_NONE, nullptr, st, rt, jobs);
      if (rc) return rc;
    }
    // cotangent of this step's input state: through the solver, through the
    // network, and the loss's own cotangent of that state (for k > 0)
    T* dst = k == 0 ? g_u0 : lam;
    rc = lincomb<T>(tot, T(1), lam_sol, gx, k == 0 ? nullptr : g_traj + int64_t(k - 1) * tot, dst, st);
    if (rc) return rc;
    lam_in = lam;
  }
  // reduce every layer's partial sums into g_w = [w0, b0, w1, b1, ...]; the
  // last layer's times the output scale
  int64_t at = 0;
  for (int l = 0; l < nl; ++l) {
    const int nw = int(width[l]) - (l == nl - 1 ? 1 : C);
    if (nparts[l] > 0) {
      hipLaunchKernelGGL(reduce_parts_kernel<T>, dim3(unsigned((width[l] + 15) / 16)), dim3(256), 0, st, nparts[l],
                         int(width[l]), work + L.part[l], g_w + at, g_w + at + nw, nw, nw, nw);
      rc = check_launch("ragged INC weight-gradient reduction");
    } else {
      rc = fill<T>(width[l], T(0), g_w + at, st);
    }
    if (rc) return rc;
    if (l == nl - 1 && cnn->scale != 1.0)
      rc = lincomb<T>(width[l], T(cnn->scale), g_w + at, nullptr, nullptr, g_w + at, st);
    if (rc) return rc;
    at += width[l];
  }
  return INC_OK;
}

// ---- allocation-free weight gradient ------------------------------------------
// The partial sums of inc_conv1d_bwd_weight_noalloc go to this file-scope
// device array instead of a caller workspace.  Each GPU the library runs on
// holds its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: the device address of each GPU's
// copy comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address is passed to the kernels.
constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB per GPU
__device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (nullptr and the
// error text set on failure).
void* scratch_address() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("conv1d weight gradient (no allocation): no usable current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("conv1d weight gradient (no allocation): hipGetSymbolAddress failed: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace inc

extern "C" {

int inc_conv1d_fwd_act(int dtype, int act, int B, int Ci, int Co, int N, int K, const void* x, const void* w,
                       const void* bias, void* y, void* dy, void* stream) {
  inc::clear_error();
  int rc = inc::conv_check(dtype, act, B, Ci, Co, N, K);
  if (rc || B == 0) return rc;
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

int inc_conv1d_bwd_weight_noalloc(int dtype, int B, int Ci, int Co, int N, int K, const void* g, const void* x,
                                  void* gw, void* gb, void* stream) {
  inc::clear_error();
  int rc = inc::conv_check(dtype, INC_ACT_NONE, B, Ci, Co, N, K);
  if (rc) return rc;
  INC_REQUIRE(B > 0, INC_EINVAL, "conv1d weight gradient of an empty batch");
  const int64_t esize = dtype == INC_F32 ? 4 : 8;
  const int64_t capacity = int64_t(inc::kScratchBytes) / esize;
  const int64_t width = int64_t(Co) * Ci * K + Co;
  INC_REQUIRE(width <= capacity, INC_EINVAL,
              "conv1d weight gradient (no allocation): one partial row of Co*Ci*K + Co = %lld elements exceeds the "
              "%zu-byte scratch array",
              static_cast<long long>(width), inc::kScratchBytes);
  void* work = inc::scratch_address();
  if (work == nullptr) return INC_EHIP;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return dtype == INC_F32 ? inc::conv_wgrad<float>(B, Ci, Co, N, K, g, x, gw, gb, work, capacity, st)
                          : inc::conv_wgrad<double>(B, Ci, Co, N, K, g, x, gw, gb, work, capacity, st);
}

int inc_conv1d_rows_fwd_act(int dtype, int act, const inc_conv_rows* r, int Ci, int Co, int K, const void* x,
                            const void* w, const void* bias, void* y, void* dy, void* stream) {
  inc::clear_error();
  int rc = inc::rows_check(dtype, act, r, Ci, Co, K);
  if (rc || r->njobs == 0) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const inc::RowTiles rt = inc::row_tiles(r);
  return dtype == INC_F32
             ? inc::conv_fwd_launch<float, false>(0, Ci, Co, 0, K, x, w, bias, y, nullptr, act, dy, st, rt, r->njobs)
             : inc::conv_fwd_launch<double, false>(0, Ci, Co, 0, K, x, w, bias, y, nullptr, act, dy, st, rt, r->njobs);
}

int inc_conv1d_rows_bwd_data_mul(int dtype, const inc_conv_rows* r, int Ci, int Co, int K, const void* g,
                                 const void* w, const void* m, void* gx, void* stream) {
  inc::clear_error();
  int rc = inc::rows_check(dtype, INC_ACT_NONE, r, Ci, Co, K);
  if (rc || r->njobs == 0) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const inc::RowTiles rt = inc::row_tiles(r);
  return dtype == INC_F32
             ? inc::conv_fwd_launch<float, true>(0, Co, Ci, 0, K, g, w, nullptr, gx, m, INC_ACT_NONE, nullptr, st, rt,
                                                 r->njobs)
             : inc::conv_fwd_launch<double, true>(0, Co, Ci, 0, K, g, w, nullptr, gx, m, INC_ACT_NONE, nullptr, st, rt,
                                                  r->njobs);
}

int inc_conv1d_rows_bwd_weight(int dtype, const inc_conv_rows* r, int Ci, int Co, int K, const void* g, const void* x,
                               void* gw, void* gb, void* work, int64_t work_elems, void* stream) {
  inc::clear_error();
  int rc = inc::rows_check(dtype, INC_ACT_NONE, r, Ci, Co, K);
  if (rc) return rc;
  INC_REQUIRE(r->njobs > 0, INC_EINVAL, "conv1d weight gradient of an empty set of rows");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const inc::WgradMode md{inc::row_tiles(r), r->njobs, false, true, nullptr};
  return dtype == INC_F32 ? inc::conv_wgrad<float>(0, Ci, Co, 0, K, g, x, gw, gb, work, work_elems, st, md)
                          : inc::conv_wgrad<double>(0, Ci, Co, 0, K, g, x, gw, gb, work, work_elems, st, md);
}

int inc_ragged_inc_workspace(const inc_ragged_desc* d, const inc_cnn_desc* cnn, int adjoint, int64_t* elems) {
  inc::clear_error();
  INC_REQUIRE(d && cnn && elems, INC_EINVAL, "null argument");
  INC_REQUIRE(cnn->layers >= 2 && cnn->layers <= 8 && cnn->channels >= 1 && cnn->kernel >= 1, INC_EINVAL,
              "bad network description");
  const int rc = inc::inc_mode_check(adjoint);
  if (rc) return rc;
  *elems = inc::inc_layout(d->total, cnn, adjoint).total;
  return INC_OK;
}

int inc_ragged_inc_fwd(const inc_ragged_desc* d, const inc_cnn_desc* cnn, const inc_conv_rows* rows,
                       const int32_t* active_tiles, const void* u0, int steps, void* traj, void* corr, void* stages,
                       void* acts, void* work, int64_t work_elems, void* stream) {
  inc::clear_error();
  int rc = inc::inc_check(d, cnn, rows);
  if (rc) return rc;
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(u0 && traj && work, INC_EINVAL, "u0, traj and work are required");
  INC_REQUIRE(work_elems >= inc::inc_layout(d->total, cnn, inc::kIncForward).total, INC_EINVAL,
              "workspace too small");
  if (d->count == 0) return INC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  inc::RaggedStep sv{*d, d->total, d->scheme == INC_SSPRK3 ? 3 : 1};
  const inc::NetRows nr{inc::row_tiles(rows), 0, 0, active_tiles, rows->njobs};
  if (d->dtype == INC_F32)
    return inc::inc_fwd<float>(sv, cnn, nr, static_cast<const float*>(u0), steps, static_cast<float*>(traj),
                               static_cast<float*>(corr), static_cast<float*>(stages), static_cast<float*>(acts),
                               static_cast<float*>(work), st);
  return inc::inc_fwd<double>(sv, cnn, nr, static_cast<const double*>(u0), steps, static_cast<double*>(traj),
                              static_cast<double*>(corr), static_cast<double*>(stages), static_cast<double*>(acts),
                              static_cast<double*>(work), st);
}

int inc_ragged_inc_adj(const inc_ragged_desc* d, const inc_cnn_desc* cnn, const inc_conv_rows* rows,
                       const int32_t* active_tiles, const void* u0, const void* traj, const void* stages,
                       const void* acts, const void* g_traj, void* g_u0, void* g_w, int steps, void* work,
                       int64_t work_elems, void* stream) {
  inc::clear_error();
  int rc = inc::inc_check(d, cnn, rows);
  if (rc) return rc;
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(u0 && traj && stages && g_traj && g_u0 && g_w && work, INC_EINVAL, "missing buffer");
  INC_REQUIRE(work_elems >= inc::inc_layout(d->total, cnn, acts ? inc::kIncAdjointStored : inc::kIncAdjoint).total,
              INC_EINVAL, "workspace too small");
  if (d->count == 0) return INC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  inc::RaggedStep sv{*d, d->total, d->scheme == INC_SSPRK3 ? 3 : 1};
  const inc::NetRows nr{inc::row_tiles(rows), 0, 0, active_tiles, rows->njobs};
  if (d->dtype == INC_F32)
    return inc::inc_adj<float>(sv, cnn, nr, static_cast<const float*>(u0), static_cast<const float*>(traj),
                               static_cast<const float*>(stages), static_cast<const float*>(acts),
                               static_cast<const float*>(g_traj), nullptr, static_cast<float*>(g_u0),
                               static_cast<float*>(g_w), steps, static_cast<float*>(work), st);
  return inc::inc_adj<double>(sv, cnn, nr, static_cast<const double*>(u0), static_cast<const double*>(traj),
                              static_cast<const double*>(stages), static_cast<const double*>(acts),
                              static_cast<const double*>(g_traj), nullptr, static_cast<double*>(g_u0),
                              static_cast<double*>(g_w), steps, static_cast<double*>(work), st);
}

int inc_ks_inc_workspace(const inc_ks_desc* d, const inc_cnn_desc* cnn, int adjoint, int64_t* elems) {
  inc::clear_error();
  INC_REQUIRE(d && cnn && elems, INC_EINVAL, "null argument");
  INC_REQUIRE(cnn->layers >= 2 && cnn->layers <= 8 && cnn->channels >= 1 && cnn->kernel >= 1, INC_EINVAL,
              "bad network description");
  const int rc = inc::inc_mode_check(adjoint);
  if (rc) return rc;
  *elems = inc::inc_layout(int64_t(d->batch) * d->n, cnn, adjoint).total;
  return INC_OK;
}

int inc_ks_inc_fwd(const inc_ks_desc* d, const inc_cnn_desc* cnn, const void* u0, int steps, void* traj, void* corr,
                   void* stages, void* acts, void* work, int64_t work_elems, void* stream) {
  inc::clear_error();
  int rc = inc::ks_inc_check(d, cnn);
  if (rc) return rc;
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(u0 && traj && work, INC_EINVAL, "u0, traj and work are required");
  const int64_t tot = int64_t(d->batch) * d->n;
  INC_REQUIRE(work_elems >= inc::inc_layout(tot, cnn, inc::kIncForward).total, INC_EINVAL, "workspace too small");
  if (d->batch == 0) return INC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  inc::KsStep sv{d, tot, inc::ks_stages(d)};
  const inc::NetRows nr{inc::RowTiles{}, d->batch, d->n, nullptr, 0};
  if (d->dtype == INC_F32)
    return inc::inc_fwd<float>(sv, cnn, nr, static_cast<const float*>(u0), steps, static_cast<float*>(traj),
                               static_cast<float*>(corr), static_cast<float*>(stages), static_cast<float*>(acts),
                               static_cast<float*>(work), st);
  return inc::inc_fwd<double>(sv, cnn, nr, static_cast<const double*>(u0), steps, static_cast<double*>(traj),
                              static_cast<double*>(corr), static_cast<double*>(stages), static_cast<double*>(acts),
                              static_cast<double*>(work), st);
}

int inc_ks_inc_adj(const inc_ks_desc* d, const inc_cnn_desc* cnn, const void* u0, const void* traj,
                   const void* stages, const void* acts, const void* g_traj, const void* g_corr, void* g_u0,
                   void* g_w, int steps, void* work, int64_t work_elems, void* stream) {
  inc::clear_error();
  int rc = inc::ks_inc_check(d, cnn);
  if (rc) return rc;
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(u0 && traj && stages && g_traj && g_u0 && g_w && work, INC_EINVAL, "missing buffer");
  const int64_t tot = int64_t(d->batch) * d->n;
  INC_REQUIRE(work_elems >= inc::inc_layout(tot, cnn, acts ? inc::kIncAdjointStored : inc::kIncAdjoint).total,
              INC_EINVAL, "workspace too small");
  if (d->batch == 0) return INC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  inc::KsStep sv{d, tot, inc::ks_stages(d)};
  const inc::NetRows nr{inc::RowTiles{}, d->batch, d->n, nullptr, 0};
  if (d->dtype == INC_F32)
    return inc::inc_adj<float>(sv, cnn, nr, static_cast<const float*>(u0), static_cast<const float*>(traj),
                               static_cast<const float*>(stages), static_cast<const float*>(acts),
                               static_cast<const float*>(g_traj), static_cast<const float*>(g_corr),
                               static_cast<float*>(g_u0), static_cast<float*>(g_w), steps, static_cast<float*>(work),
                               st);
  return inc::inc_adj<double>(sv, cnn, nr, static_cast<const double*>(u0), static_cast<const double*>(traj),
                              static_cast<const double*>(stages), static_cast<const double*>(acts),
                              static_cast<const double*>(g_traj), static_cast<const double*>(g_corr),
                              static_cast<double*>(g_u0), static_cast<double*>(g_w), steps,
                              static_cast<double*>(work), st);
}

}  // extern "C"
