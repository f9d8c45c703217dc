// Burgers WENO5 / FDM step kernels for gfx950 (forward and adjoint).
//
// Replaces the eager op sequence of BurgersSolverTorch (reference
// solvers/solver_1d.py:44-195) and weno_reconstruction (:4-42).
//
// Layout: a trajectory is a row of n points in HBM; a team of `team`
// threads (a power of two <= 256) owns one trajectory and each thread owns
// PPT consecutive points, kept in registers across stages and steps.  The
// stage state is staged through LDS so neighbours can read the periodic
// 6-point WENO window; the Lax-Friedrichs alpha = max|u| (and its argmax,
// for the adjoint) is a wavefront shuffle reduction plus an LDS combine
// across the team's wavefronts.
#include <type_traits>

#include "common.hpp"

namespace inc {
namespace {

// ---- WENO5 tables ----------------------------------------------------------
// Candidate k of direction d: q (value), a and b (linear forms with
// beta = 13/12 a^2 + 1/4 b^2) over 5 stencil taps.  Taps are offsets
// -2..2 for 'left' (solver_1d.py:14-22) and -1..3 for 'right'
// (solver_1d.py:24-33; the reference's right reconstruction is the left one
// shifted by a cell, with beta_2's b form mirrored).
__device__ constexpr double kQ[2][3][5] = {
    {{1. / 3, -7. / 6, 11. / 6, 0, 0}, {0, -1. / 6, 5. / 6, 1. / 3, 0}, {0, 0, 1. / 3, 5. / 6, -1. / 6}},
    {{1. / 3, -7. / 6, 11. / 6, 0, 0}, {0, -1. / 6, 5. / 6, 1. / 3, 0}, {0, 0, 1. / 3, 5. / 6, -1. / 6}}};
__device__ constexpr double kA[2][3][5] = {
    {{1, -2, 1, 0, 0}, {0, 1, -2, 1, 0}, {0, 0, 1, -2, 1}},
    {{1, -2, 1, 0, 0}, {0, 1, -2, 1, 0}, {0, 0, 1, -2, 1}}};
__device__ constexpr double kB[2][3][5] = {
    {{1, -4, 3, 0, 0}, {0, 1, 0, -1, 0}, {0, 0, 3, -4, 1}},
    {{1, -4, 3, 0, 0}, {0, -1, 0, 1, 0}, {0, 0, 1, -4, 3}}};
__device__ constexpr double kGamma[3] = {0.1, 0.6, 0.3};
constexpr double kEps = 1e-6;  // solver_1d.py:5

template <int D, int K, typename T>
__device__ __forceinline__ T lin_q(const T* f) {
  T acc = T(0);
#pragma unroll
  for (int m = 0; m < 5; ++m)
    if (kQ[D][K][m] != 0.0) acc += T(kQ[D][K][m]) * f[m];
  return acc;
}
template <int D, int K, typename T>
__device__ __forceinline__ T lin_a(const T* f) {
  T acc = T(0);
#pragma unroll
  for (int m = 0; m < 5; ++m)
    if (kA[D][K][m] != 0.0) acc += T(kA[D][K][m]) * f[m];
  return acc;
}
template <int D, int K, typename T>
__device__ __forceinline__ T lin_b(const T* f) {
  T acc = T(0);
#pragma unroll
  for (int m = 0; m < 5; ++m)
    if (kB[D][K][m] != 0.0) acc += T(kB[D][K][m]) * f[m];
  return acc;
}

template <typename T>
struct WenoParts {
  T q[3], a[3], b[3], r[3], w[3], F;  // r_k = 1 / (eps + beta_k)
};

// Nonlinear weights and reconstructed value for one 5-tap stencil.
// alpha_k = gamma_k / d_k^2 is formed from r_k = 1/d_k (one reciprocal per
// candidate, which the VJP reuses).
template <int D, typename T>
__device__ __forceinline__ void weno_parts(const T* f, WenoParts<T>& o) {
  o.q[0] = lin_q<D, 0>(f); o.q[1] = lin_q<D, 1>(f); o.q[2] = lin_q<D, 2>(f);
  o.a[0] = lin_a<D, 0>(f); o.a[1] = lin_a<D, 1>(f); o.a[2] = lin_a<D, 2>(f);
  o.b[0] = lin_b<D, 0>(f); o.b[1] = lin_b<D, 1>(f); o.b[2] = lin_b<D, 2>(f);
  T al[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T beta = T(13.0 / 12.0) * o.a[k] * o.a[k] + T(0.25) * o.b[k] * o.b[k];
    o.r[k] = rcp(T(kEps) + beta);
    al[k] = T(kGamma[k]) * (o.r[k] * o.r[k]);
  }
  T inv = rcp(al[0] + al[1] + al[2]);
  o.w[0] = al[0] * inv; o.w[1] = al[1] * inv; o.w[2] = al[2] * inv;
  o.F = o.w[0] * o.q[0] + o.w[1] * o.q[1] + o.w[2] * o.q[2];
}

template <int D, typename T>
__device__ __forceinline__ T weno_value(const T* f) {
  WenoParts<T> o;
  weno_parts<D>(f, o);
  return o.F;
}

// df[m] = gbar * dF/df_m.  With w_k = alpha_k / sum alpha,
// alpha_k = gamma_k / d_k^2, d_k = eps + beta_k:
//   dF/df_m = sum_k w_k dq_k/df_m + sum_k c_k dbeta_k/df_m,
//   c_k = -2 w_k (q_k - F) / d_k,
//   dbeta_k/df_m = 13/6 a_k da_k/df_m + 1/2 b_k db_k/df_m.
// With e_k = gbar w_k (q_k - F) / d_k the beta terms' weights are
// gbar c_k 13/6 a_k = -13/3 e_k a_k and gbar c_k 1/2 b_k = -e_k b_k.
template <int D, typename T>
__device__ __forceinline__ void weno_vjp(const T* f, T gbar, T* df) {
  WenoParts<T> o;
  weno_parts<D>(f, o);
  T pw[3], pa[3], pb[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pw[k] = gbar * o.w[k];
    const T e = pw[k] * ((o.q[k] - o.F) * o.r[k]);
    pa[k] = e * (T(-13.0 / 3.0) * o.a[k]);
    pb[k] = -e * o.b[k];
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (kQ[D][k][m] != 0.0) acc += T(kQ[D][k][m]) * pw[k];
      if (kA[D][k][m] != 0.0) acc += T(kA[D][k][m]) * pa[k];
      if (kB[D][k][m] != 0.0) acc += T(kB[D][k][m]) * pb[k];
    }
    df[m] = acc;
  }
}

__device__ __forceinline__ int wrap(int j, int n) {
  return j < 0 ? j + n : (j >= n ? j - n : j);
}

// ---- standalone WENO reconstruction (weno_reconstruction) ------------------
template <typename T, int D>
__global__ void __launch_bounds__(256) weno_fwd_kernel(const T* __restrict__ f, T* __restrict__ out,
                                                       int64_t rows, int n) {
  int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= rows * n) return;
  int64_t r = gid / n;
  int i = int(gid - r * n);
  const T* row = f + r * n;
  T taps[5];
  const int o0 = D == 0 ? -2 : -1;
#pragma unroll
  for (int m = 0; m < 5; ++m) taps[m] = row[wrap(i + o0 + m, n)];
  out[gid] = weno_value<D>(taps);
}

// g_f[j] = sum over reconstructions i whose stencil covers j.
template <typename T, int D>
__global__ void __launch_bounds__(256) weno_adj_kernel(const T* __restrict__ f, const T* __restrict__ g,
                                                       T* __restrict__ gf, int64_t rows, int n) {
  int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= rows * n) return;
  int64_t r = gid / n;
  int j = int(gid - r * n);
  const T* row = f + r * n;
  const T* grow = g + r * n;
  const int o0 = D == 0 ? -2 : -1;
  T acc = T(0);
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    // reconstruction i = j - (o0 + m) has tap m at point j
    int i = wrap(j - (o0 + m), n);
    T taps[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) taps[q] = row[wrap(i + o0 + q, n)];
    T df[5];
    weno_vjp<D>(taps, grow[i], df);
    acc += df[m];
  }
  gf[gid] = acc;
}

// ---- Burgers step kernels ---------------------------------------------------
template <typename T>
struct BArgs {
  int batch, n, J, team, steps, scheme, source, parts;
  T dt, inv_dx, inv_dx2, half_inv_dx, nu, beta, t0, two_pi_over_L;
  const T* nu_ptr;
  const T* beta_ptr;
  const T* t_ptr;
  const T* x;
  const T* A;
  const T* omega;
  const T* phi;
  const T* l;
  const T* u;
  const T* s;
  T* out;
  T* traj;
  T* stages;
  const T* g_traj;
  T* g_s;
  // fused MSE to a target trajectory [B, steps, n]: the forward accumulates
  // sum (u - target)^2 into loss_parts[b][4] (one partial per wavefront of
  // the team); the adjoint adds *loss_scale * (traj - target) to the cotangent
  // of every output, reading the stored trajectory from traj
  const T* target;
  T* loss_parts;
  const T* loss_scale;  // device scalar
  // adaptive sub-stepping (cfl > 0, holding CFL * dx): per step and
  // trajectory, nsub sub-steps of sizes dts, stages indexed [b][step][sub][stage]
  T cfl;
  int max_sub;
  int* nsub;
  T* dts;
  // [batch] cotangent of the viscosity (per trajectory), written by the
  // adjoint kernels when non-null: sum over stages of cb . (D2 u)
  T* g_nu;
};

constexpr int kBlock = 256;
constexpr int kMaxWaves = kBlock / kWave;
constexpr int kLossSlots = 4;  // partial loss sums per trajectory (wavefronts per team <= 4)

template <typename T>
__device__ __forceinline__ bool better(T v, int i, T ov, int oi) {
  // first maximal element; NaN wins (torch.max propagates NaN)
  bool vn = v != v, on = ov != ov;
  if (vn || on) return on && (!vn || oi < i);
  return ov > v || (ov == v && oi < i);
}

template <typename T>
struct TeamCtx {
  int lt, team, wave_in_block, first_wave, i0, n, b;
  bool active;
};

// Team-wide argmax of |u| over the team's points.  Writes per-wave partials
// to LDS before the barrier the caller issues; finish after that barrier.
// The common case (no NaN among the team's values) reduces the maximum alone
// and takes the index from the first lane that holds it: lanes hold
// consecutive points, so that lane's own first maximal index is the team's
// first maximal index, as torch.max(dim) returns.  A NaN anywhere (it wins,
// first NaN index) takes the full (value, index) shuffle reduction.
template <typename T, int PPT>
__device__ __forceinline__ void argmax_begin(const TeamCtx<T>& c, const T* u, T& v, int& idx,
                                             T* red_v, int* red_i) {
  v = T(-1);
  idx = 0x7fffffff;
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    int i = c.i0 + p;
    if (i < c.n) {
      T a = fabs(u[p]);
      if (better(v, idx, a, i)) { v = a; idx = i; }
    }
  }
  const int width = c.team < kWave ? c.team : kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int base = lane & ~(width - 1);  // first lane of this team's part of the wavefront
  const uint64_t team_mask = (width == kWave ? ~0ull : ((1ull << width) - 1)) << base;
  if ((__ballot(v != v) & team_mask) == 0) {
    T m = v;
    for (int off = 1; off < width; off <<= 1) m = fmax(m, shfl_xor(m, off));
    const uint64_t hit = __ballot(v == m) & team_mask;
    const int first = __ffsll(static_cast<unsigned long long>(hit)) - 1;
    idx = __shfl(idx, first, kWave);
    v = m;
  } else {
    for (int off = 1; off < width; off <<= 1) {
      T ov = shfl_xor(v, off);
      int oi = __shfl_xor(idx, off, kWave);
      if (better(v, idx, ov, oi)) { v = ov; idx = oi; }
    }
  }
  if (c.team > kWave && (threadIdx.x & (kWave - 1)) == 0) {
    red_v[c.wave_in_block] = v;
    red_i[c.wave_in_block] = idx;
  }
}

template <typename T>
__device__ __forceinline__ void argmax_end(const TeamCtx<T>& c, T& v, int& idx, const T* red_v,
                                           const int* red_i) {
  if (c.team > kWave) {
    v = red_v[c.first_wave];
    idx = red_i[c.first_wave];
    for (int w = 1; w < c.team / kWave; ++w) {
      T ov = red_v[c.first_wave + w];
      int oi = red_i[c.first_wave + w];
      if (better(v, idx, ov, oi)) { v = ov; idx = oi; }
    }
  }
}

template <typename T>
__device__ __forceinline__ T source_at(const BArgs<T>& a, int b, int i, T u, T t, T xi) {
  if (a.source == INC_SRC_LINEAR) {
    T be = a.beta_ptr ? a.beta_ptr[b] : a.beta;
    return be * u;
  }
  if (a.source == INC_SRC_LEARNING) {
    // solver_1d.py:82-84
    T acc = T(0);
    for (int j = 0; j < a.J; ++j) {
      int o = b * a.J + j;
      T arg = a.omega[o] * t + (a.two_pi_over_L * a.l[o]) * xi + a.phi[o];
      acc += a.A[o] * sin(arg);
    }
    return acc;
  }
  return T(0);
}

// One spatial right-hand side evaluation for the team's own points:
// R = convection + diffusion + source (+ s), restricted to `parts`.
// Contains two block barriers; the caller issues the third before LDS reuse.
template <typename T, int PPT, int CONV>
__device__ __forceinline__ void rhs_eval(const BArgs<T>& a, const TeamCtx<T>& c, const T* u,
                                         const T* sv, const T* xv, T t, T nu, int parts, T* us,
                                         T* Fs, T* red_v, int* red_i, T* R) {
  constexpr int W = PPT + 5;
#pragma unroll
  for (int p = 0; p < PPT; ++p)
    if (c.i0 + p < c.n) us[c.i0 + p] = u[p];
  T alpha = T(0);
  int amax = 0;
  if (CONV == INC_WENO) argmax_begin<T, PPT>(c, u, alpha, amax, red_v, red_i);
  __syncthreads();
  if (CONV == INC_WENO) argmax_end(c, alpha, amax, red_v, red_i);

  T win[W];
  if (c.active && c.i0 < c.n) {
#pragma unroll
    for (int k = 0; k < W; ++k) win[k] = us[wrap(c.i0 - 2 + k, c.n)];
  }
  T F[PPT];
  if (CONV == INC_WENO && (parts & INC_PART_CONVECTION) && c.active && c.i0 < c.n) {
    T fp[W], fm[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
      T f = T(0.5) * win[k] * win[k];
      fp[k] = T(0.5) * (f + alpha * win[k]);
      fm[k] = T(0.5) * (f - alpha * win[k]);
    }
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      F[p] = weno_value<0>(fp + p) + weno_value<1>(fm + p + 1);
      if (c.i0 + p < c.n) Fs[c.i0 + p] = F[p];
    }
  }
  __syncthreads();
  if (!(c.active && c.i0 < c.n)) return;
  T Fm1 = T(0);
  if (CONV == INC_WENO && (parts & INC_PART_CONVECTION)) Fm1 = Fs[wrap(c.i0 - 1, c.n)];
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    T r = T(0);
    if (parts & INC_PART_CONVECTION) {
      if (CONV == INC_WENO) {
        r = -(F[p] - (p == 0 ? Fm1 : F[p - 1])) * a.inv_dx;
      } else {
        // -u * (D1 u), central difference (solver_1d.py:87-93,125-127)
        r = -win[p + 2] * ((win[p + 3] - win[p + 1]) * a.half_inv_dx);
      }
    }
    if (parts & INC_PART_DIFFUSION) r += nu * ((win[p + 1] - T(2) * win[p + 2] + win[p + 3]) * a.inv_dx2);
    if (parts & INC_PART_SOURCE) {
      r += source_at(a, c.b, c.i0 + p, win[p + 2], t, xv[p]);
      if (sv) r += sv[p];
    }
    R[p] = r;
  }
}

// cotangent of the right-hand side: ub = (dR/du)^T cb for the team's points.
// With want_nu, this thread's points add cb . dR/dnu = cb . (D2 u) to *nub.
template <typename T, int PPT, int CONV>
__device__ __forceinline__ void rhs_vjp(const BArgs<T>& a, const TeamCtx<T>& c, const T* u,
                                        const T* cb, T nu, int parts, T* us, T* cs, T* C,
                                        T* red_v, int* red_i, T* ub, bool want_nu = false,
                                        T* nub = nullptr) {
  constexpr int W = PPT + 5;
  const int n = c.n;
  const bool own = c.active && c.i0 < n;
#pragma unroll
  for (int p = 0; p < PPT; ++p)
    if (c.i0 + p < n) { us[c.i0 + p] = u[p]; cs[c.i0 + p] = cb[p]; }
  T alpha = T(0);
  int amax = 0;
  const bool conv = parts & INC_PART_CONVECTION;
  if (CONV == INC_WENO && conv) argmax_begin<T, PPT>(c, u, alpha, amax, red_v, red_i);
  __syncthreads();
  if (CONV == INC_WENO && conv) argmax_end(c, alpha, amax, red_v, red_i);

  T uw[W], cw[PPT + 2];
  T abar = T(0);
  if (own) {
#pragma unroll
    for (int k = 0; k < W; ++k) uw[k] = us[wrap(c.i0 - 2 + k, n)];
#pragma unroll
    for (int k = 0; k < PPT + 2; ++k) cw[k] = cs[wrap(c.i0 - 1 + k, n)];
  }
  if (want_nu && own && (parts & INC_PART_DIFFUSION)) {
    // the forward's diffusion term is nu * ((u_{i-1} - 2 u_i + u_{i+1}) * inv_dx2)
    T acc = T(0);
#pragma unroll
    for (int p = 0; p < PPT; ++p) acc += cb[p] * ((uw[p + 1] - T(2) * uw[p + 2] + uw[p + 3]) * a.inv_dx2);
    *nub += acc;
  }
  if (own && conv) {
    if (CONV == INC_WENO) {
      T fp[W], fm[W];
#pragma unroll
      for (int k = 0; k < W; ++k) {
        T f = T(0.5) * uw[k] * uw[k];
        fp[k] = T(0.5) * (f + alpha * uw[k]);
        fm[k] = T(0.5) * (f - alpha * uw[k]);
      }
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        const int i = c.i0 + p;
        // N_i = -(F_i - F_{i-1})/dx  =>  Fbar_i = (c_{i+1} - c_i)/dx; the
        // halves of d f+/du = (u + alpha)/2, d f-/du = (u - alpha)/2,
        // d f+/dalpha = u/2 and d f-/dalpha = -u/2 are folded into it
        T Fbar = (cw[p + 2] - cw[p + 1]) * (T(0.5) * a.inv_dx);
        T dfp[5], dfm[5];
        weno_vjp<0>(fp + p, Fbar, dfp);      // taps i-2..i+2
        weno_vjp<1>(fm + p + 1, Fbar, dfm);  // taps i-1..i+3
#pragma unroll
        for (int m = -2; m <= 3; ++m) {
          T gp = (m <= 2) ? dfp[m + 2] : T(0);
          T gm = (m >= -1) ? dfm[m + 1] : T(0);
          T uu = uw[p + 2 + m];
          T contrib = gp * (uu + alpha) + gm * (uu - alpha);
          abar += uu * (gp - gm);
          if (i < n) C[(m + 2) * n + i] = contrib;
        }
      }
    } else {
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        const int i = c.i0 + p;
        if (i < n) {
          T ci = cw[p + 1];
          T ui = uw[p + 2];
          C[0 * n + i] = ci * ui * a.half_inv_dx;                       // to u_{i-1}
          C[1 * n + i] = -ci * (uw[p + 3] - uw[p + 1]) * a.half_inv_dx;  // to u_i
          C[2 * n + i] = -ci * ui * a.half_inv_dx;                      // to u_{i+1}
        }
      }
    }
  }
  if (CONV == INC_WENO && conv) {
    const int width = c.team < kWave ? c.team : kWave;
    abar = wave_sum(abar, width);
    if (c.team > kWave && (threadIdx.x & (kWave - 1)) == 0) red_v[kMaxWaves + c.wave_in_block] = abar;
  }
  __syncthreads();
  if (CONV == INC_WENO && conv && c.team > kWave) {
    abar = red_v[kMaxWaves + c.first_wave];
    for (int w = 1; w < c.team / kWave; ++w) abar += red_v[kMaxWaves + c.first_wave + w];
  }
  if (!own) return;
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    const int i = c.i0 + p;
    T g = T(0);
    if (conv) {
      if (CONV == INC_WENO) {
#pragma unroll
        for (int m = -2; m <= 3; ++m) g += C[(m + 2) * n + wrap(i - m, n)];
        if (i == amax) {
          T ui = uw[p + 2];
          T sg = ui > T(0) ? T(1) : (ui < T(0) ? T(-1) : T(0));
          g += abar * sg;
        }
      } else {
        g += C[0 * n + wrap(i + 1, n)] + C[1 * n + i] + C[2 * n + wrap(i - 1, n)];
      }
    }
    if (parts & INC_PART_DIFFUSION) g += nu * ((cw[p] - T(2) * cw[p + 1] + cw[p + 2]) * a.inv_dx2);
    if ((parts & INC_PART_SOURCE) && a.source == INC_SRC_LINEAR) {
      T be = a.beta_ptr ? a.beta_ptr[c.b] : a.beta;
      g += be * cw[p + 1];
    }
    ub[p] = g;
  }
}

template <typename T>
__device__ __forceinline__ TeamCtx<T> make_ctx(int team, int n, int batch, int ppt) {
  TeamCtx<T> c;
  c.team = team;
  int tpb = kBlock / team;
  int tid = threadIdx.x;
  int t = tid / team;
  c.lt = tid - t * team;
  c.b = blockIdx.x * tpb + t;
  c.active = c.b < batch;
  c.wave_in_block = tid / kWave;
  c.first_wave = (t * team) / kWave;
  c.i0 = c.lt * ppt;
  c.n = n;
  return c;
}

// Sum of v over the team, written to out[c.b] by the team's first thread.
// Block-uniform call (it holds a barrier when teams span wavefronts); uses
// red_v[0, kMaxWaves), which no other reader may still hold.
template <typename T>
__device__ __forceinline__ void team_sum_store(const TeamCtx<T>& c, T v, T* red_v, T* out) {
  const int width = c.team < kWave ? c.team : kWave;
  v = wave_sum(v, width);
  if (c.team > kWave) {
    if ((threadIdx.x & (kWave - 1)) == 0) red_v[c.wave_in_block] = v;
    __syncthreads();
    v = red_v[c.first_wave];
    for (int w = 1; w < c.team / kWave; ++w) v += red_v[c.first_wave + w];
  }
  if (c.active && c.lt == 0) out[c.b] = v;
}

// Shared-memory layout per team (nb = number of n-sized arrays).
template <typename T>
__device__ __forceinline__ T* team_lds(T* base, int team, int n, int nb) {
  int t = threadIdx.x / team;
  return base + size_t(t) * nb * n;
}

template <typename T, int PPT, int CONV>
__global__ void __launch_bounds__(kBlock) burgers_fwd_kernel(BArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T red_v[2 * kMaxWaves];
  __shared__ int red_i[kMaxWaves];
  TeamCtx<T> c = make_ctx<T>(a.team, a.n, a.batch, PPT);
  T* us = team_lds(reinterpret_cast<T*>(smem_raw), a.team, a.n, 2);
  T* Fs = us + a.n;
  const int n = a.n;
  const bool own = c.active && c.i0 < n;
  const int ns = a.scheme == INC_SSPRK3 ? 3 : 1;
  const T nu = a.nu_ptr ? (c.active ? a.nu_ptr[c.b] : T(0)) : a.nu;
  const T t0 = a.t_ptr ? (c.active ? a.t_ptr[c.b] : T(0)) : a.t0;
  T u0[PPT], uc[PPT], sv[PPT], xv[PPT], R[PPT];
  T lsum = T(0);  // squared error to the target over this thread's points (fused MSE)
  const size_t row = size_t(c.b) * n;
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    int i = c.i0 + p;
    bool ok = own && i < n;
    u0[p] = ok ? a.u[row + i] : T(0);
    sv[p] = (ok && a.s) ? a.s[row + i] : T(0);
    xv[p] = (ok && a.source == INC_SRC_LEARNING) ? a.x[i] : T(0);
  }
  // One (sub-)step of size h from u0 with the forcing clock at ts0; stage
  // states go to sdst when given.  on == false: the team takes part in the
  // block's barriers only and its state is left as it is.
  auto substep = [&](T h, T ts0, T* sdst, bool on) __attribute__((always_inline)) {
#pragma unroll
    for (int p = 0; p < PPT; ++p) uc[p] = u0[p];
    for (int st = 0; st < ns; ++st) {
      if (sdst && on && own) {
        T* dst = sdst + size_t(st) * n;
#pragma unroll
        for (int p = 0; p < PPT; ++p)
          if (c.i0 + p < n) dst[c.i0 + p] = uc[p];
      }
      const T ts = st == 0 ? ts0 : (st == 1 ? ts0 + h : ts0 + T(0.5) * h);
      rhs_eval<T, PPT, CONV>(a, c, uc, a.s ? sv : nullptr, xv, ts, nu, a.parts, us, Fs, red_v,
                             red_i, R);
      __syncthreads();
      if (on) {
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          T v = uc[p] + h * R[p];  // forward-Euler stage u + h*R(u)
          if (ns == 1 || st == 0) uc[p] = v;
          else if (st == 1) uc[p] = T(0.75) * u0[p] + T(0.25) * v;
          else uc[p] = T(1.0 / 3.0) * u0[p] + T(2.0 / 3.0) * v;
        }
      }
    }
    if (on) {
#pragma unroll
      for (int p = 0; p < PPT; ++p) u0[p] = uc[p];
    }
  };
  for (int step = 0; step < a.steps; ++step) {
    const T t = t0 + T(step) * a.dt;
    if (a.cfl <= T(0)) {
      substep(a.dt, t, a.stages ? a.stages + (size_t(c.b) * a.steps + step) * ns * n : nullptr, true);
    } else {
      // Adaptive (solver_1d.py:148-179 with max|u| per trajectory): sub-steps
      // h = min(remaining, cfl * dx / max|u|), the max_sub-th covering the
      // rest.  The loop is block-uniform: a team that has covered its dt
      // keeps joining the barriers until every team of the block has.
      T remaining = a.dt, tl = t;
      int j = 0;
      bool going = c.active;
      for (;;) {
        T m;
        int mi;
        argmax_begin<T, PPT>(c, u0, m, mi, red_v, red_i);
        __syncthreads();
        argmax_end(c, m, mi, red_v, red_i);
        T h = T(0);
        if (going) {
          h = remaining;
          if (m > T(0)) {
            const T allowed = a.cfl / m;  // (cfl * dx) / max|u|, the reference's order
            if (allowed < h) h = allowed;
          }
          if (j == a.max_sub - 1) h = remaining;
        }
        if (!__syncthreads_or(going)) break;
        T* sdst = (a.stages && going) ? a.stages + ((size_t(c.b) * a.steps + step) * a.max_sub + j) * ns * n
                                      : nullptr;
        substep(h, tl, sdst, going);
        if (going) {
          if (c.lt == 0) a.dts[(size_t(c.b) * a.steps + step) * a.max_sub + j] = h;
          tl += h;
          remaining -= h;
          ++j;
          going = remaining > T(0) && j < a.max_sub;
        }
      }
      if (c.active && c.lt == 0) a.nsub[size_t(c.b) * a.steps + step] = j;
    }
    if (a.traj && own) {
      T* dst = a.traj + (size_t(c.b) * a.steps + step) * n;
#pragma unroll
      for (int p = 0; p < PPT; ++p)
        if (c.i0 + p < n) dst[c.i0 + p] = u0[p];
    }
    if (a.target && own) {
      const T* tg = a.target + (size_t(c.b) * a.steps + step) * n;
#pragma unroll
      for (int p = 0; p < PPT; ++p)
        if (c.i0 + p < n) {
          const T dd = u0[p] - tg[c.i0 + p];
          lsum += dd * dd;
        }
    }
  }
  if (own) {
#pragma unroll
    for (int p = 0; p < PPT; ++p)
      if (c.i0 + p < n) a.out[row + c.i0 + p] = u0[p];
  }
  if (a.target) {
    // team sum: per wavefront by shuffles, one slot per wavefront of the team
    const int w = c.team < kWave ? c.team : kWave;
    lsum = wave_sum(lsum, w);
    if (c.active && c.lt % w == 0) a.loss_parts[size_t(c.b) * kLossSlots + c.lt / w] = lsum;
    if (c.active && c.lt == 0)
      for (int k = c.team > kWave ? c.team / kWave : 1; k < kLossSlots; ++k)
        a.loss_parts[size_t(c.b) * kLossSlots + k] = T(0);
  }
}

// ---- INC rollout with the correction network inside the kernel ---------------------
// simulate_rollout(model, solver, u, steps, 'INC') (trainer/trainer_1d.py:154-183)
// for a circular Conv1d + GELU correction network (models/models_1d/resnet.py:5-52
// stacks such layers): before every step s = scale * CNN(u), held over the
// step's stages, as solver.correction_s.  One block per trajectory; the
// network runs on the block's own state in LDS, so no activation leaves the
// CU and the whole rollout is one launch.
constexpr int kCnnMaxLayers = 8;

template <typename T>
struct CnnArgs {
  int layers;  // >= 2: Conv(1->C), GELU, (layers-2) x [Conv(C->C), GELU], Conv(C->1)
  T scale;
  const T* w[kCnnMaxLayers];  // [Co][Ci][K]
  const T* b[kCnnMaxLayers];  // [Co] or NULL
};

// Exact (erf) GELU.
template <typename T>
__device__ __forceinline__ T gelu_exact(T v) {
  return T(0.5) * v * (T(1) + erf(v * T(0.70710678118654752440)));
}

// s[p] = scale * CNN(u) at this thread's points i0 + p (every thread of the
// block calls it; u is in us, written and synchronised by the caller).
// Each thread computes every output channel of its own points, so a layer's
// weights are wave-uniform (scalar cache) and its outputs accumulate in
// registers; buf ([C][n] in LDS) holds one layer's activations, overwritten
// in place between barriers.  The last layer's output stays in registers,
// in the solver's layout.
template <typename T, int C, int K, int PPT>
__device__ __forceinline__ void cnn_eval(const CnnArgs<T>& cn, const T* us, T* buf, int n, int i0, bool own,
                                         T* s) {
  constexpr int P = K / 2;
  constexpr int WN = PPT + K - 1;
  auto at = [&](int j) {  // periodic index of i0 - P + j, |i0 - P + j| < 2n
    int i = i0 - P + j;
    return i < 0 ? i + n : (i >= n ? i - n : i);
  };
  T acc[C][PPT];
  {
    T win[WN];
#pragma unroll
    for (int j = 0; j < WN; ++j) win[j] = own ? us[at(j)] : T(0);
    const T* w = cn.w[0];
#pragma unroll
    for (int o = 0; o < C; ++o) {
      const T bo = cn.b[0] ? cn.b[0][o] : T(0);
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        T v = bo;
#pragma unroll
        for (int k = 0; k < K; ++k) v += w[o * K + k] * win[p + k];
        acc[o][p] = gelu_exact(v);
      }
    }
  }
  auto store = [&]() {
    if (own) {
#pragma unroll
      for (int o = 0; o < C; ++o)
#pragma unroll
        for (int p = 0; p < PPT; ++p)
          if (i0 + p < n) buf[o * n + i0 + p] = acc[o][p];
    }
  };
  store();
  __syncthreads();
  for (int l = 1; l < cn.layers - 1; ++l) {
    const T* w = cn.w[l];
    const T* bl = cn.b[l];
#pragma unroll
    for (int o = 0; o < C; ++o)
#pragma unroll
      for (int p = 0; p < PPT; ++p) acc[o][p] = bl ? bl[o] : T(0);
    for (int c = 0; c < C; ++c) {
      T win[WN];
#pragma unroll
      for (int j = 0; j < WN; ++j) win[j] = own ? buf[c * n + at(j)] : T(0);
#pragma unroll
      for (int o = 0; o < C; ++o)
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const T wv = w[(o * C + c) * K + k];
#pragma unroll
          for (int p = 0; p < PPT; ++p) acc[o][p] += wv * win[p + k];
        }
    }
#pragma unroll
    for (int o = 0; o < C; ++o)
#pragma unroll
      for (int p = 0; p < PPT; ++p) acc[o][p] = gelu_exact(acc[o][p]);
    __syncthreads();  // every thread has read the layer's input
    store();
    __syncthreads();
  }
  const T* w = cn.w[cn.layers - 1];
  const T bl = cn.b[cn.layers - 1] ? cn.b[cn.layers - 1][0] : T(0);
  T v[PPT];
#pragma unroll
  for (int p = 0; p < PPT; ++p) v[p] = bl;
  for (int c = 0; c < C; ++c) {
    T win[WN];
#pragma unroll
    for (int j = 0; j < WN; ++j) win[j] = own ? buf[c * n + at(j)] : T(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const T wv = w[c * K + k];
#pragma unroll
      for (int p = 0; p < PPT; ++p) v[p] += wv * win[p + k];
    }
  }
#pragma unroll
  for (int p = 0; p < PPT; ++p) s[p] = cn.scale * v[p];
}

// ---- the same network with its hidden layers on the matrix cores (float) ----------
typedef float v4f __attribute__((ext_vector_type(4)));

// One hidden layer, C -> C, in place on act (position-major: act[i][c], row
// stride C + 4), weights ws[o][k*C + c] (row stride C*K + 4) staged in LDS.
// D[o][i] = sum_r W[o][r] X[r][i] with r = k*C + c, as v_mfma_f32_16x16x4f32
// in chunks of 16 r (one tap per chunk: C % 16 == 0); lane (lr, q) feeds
// A = ws[o][16s+4q .. +3] and B = act[(i+k-P) mod n][c0+4q .. +3], one
// 16-byte LDS read each, and holds D[4q + r][lr].  Wavefront w computes the
// 16-position tiles w*PW .. w*PW + PW - 1 (positions past n are computed
// and dropped).  Reads of the next chunk are issued before the current
// chunk's MFMAs (two register sets, unconditional reads: the read after
// the last chunk lands inside LDS and is unused).  bias + GELU on the way
// back; the caller synchronises before the next reader.
template <int C, int K, int PW>
__device__ __forceinline__ void cnn_hidden_mfma(const float* ws, const float* bias, float* act, int n) {
  constexpr int CT = C / 16, P = K / 2, CS = C + 4, RS = C * K + 4, NCH = C * K / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, q = lane >> 4;
  int ib[PW];
#pragma unroll
  for (int t = 0; t < PW; ++t) ib[t] = (wave * PW + t) * 16 + lr;
  auto row = [&](int i) {  // (i) mod n for i in (-n, 3n)
    i = i < 0 ? i + n : i;
    i = i >= n ? i - n : i;
    return i >= n ? i - n : i;
  };
  v4f acc[CT][PW];
#pragma unroll
  for (int m = 0; m < CT; ++m)
#pragma unroll
    for (int t = 0; t < PW; ++t) acc[m][t] = v4f{0.f, 0.f, 0.f, 0.f};
  int k = 0, c0 = 0;
  const float* wl = ws + lr * RS + 4 * q;
  auto load = [&](int s, v4f* fa, v4f* fb) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < CT; ++m) fa[m] = *reinterpret_cast<const v4f*>(wl + m * 16 * RS + 16 * s);
#pragma unroll
    for (int t = 0; t < PW; ++t)
      fb[t] = *reinterpret_cast<const v4f*>(act + row(ib[t] + k - P) * CS + c0 + 4 * q);
  };
  auto next = [&]() __attribute__((always_inline)) {
    c0 += 16;
    if (c0 == C) {
      c0 = 0;
      ++k;
    }
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
  v4f a0[CT], b0[PW], a1[CT], b1[PW];
  load(0, a0, b0);
  for (int s = 0; s < NCH; s += 2) {
    next();
    load(s + 1, a1, b1);
    __builtin_amdgcn_sched_barrier(0);
    mma(a0, b0);
    __builtin_amdgcn_sched_barrier(0);
    if (s + 1 >= NCH) break;
    next();
    load(s + 2, a0, b0);
    __builtin_amdgcn_sched_barrier(0);
    mma(a1, b1);
    __builtin_amdgcn_sched_barrier(0);
  }
  __syncthreads();  // every wavefront has read the layer's input
#pragma unroll
  for (int m = 0; m < CT; ++m) {
    const int o = m * 16 + 4 * q;
    const v4f bo = bias ? *reinterpret_cast<const v4f*>(bias + o) : v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < PW; ++t)
      if (ib[t] < n) {
        v4f v = acc[m][t] + bo;
        *reinterpret_cast<v4f*>(act + ib[t] * CS + o) =
            v4f{gelu_exact(v[0]), gelu_exact(v[1]), gelu_exact(v[2]), gelu_exact(v[3])};
      }
  }
}

// s = scale * CNN(u) with the activations position-major in act and the
// hidden layers' weights staged in ws (layer l at ws + (l-1) * C * (C*K+4)).
template <int C, int K, int PPT, int PW>
__device__ __forceinline__ void cnn_eval_mfma(const CnnArgs<float>& cn, const float* ws, const float* us, float* act,
                                              int n, int i0, bool own, float* s) {
  constexpr int P = K / 2, WN = PPT + K - 1, CS = C + 4;
  auto at = [&](int j) {
    int i = i0 - P + j;
    return i < 0 ? i + n : (i >= n ? i - n : i);
  };
  {
    float win[WN];
#pragma unroll
    for (int j = 0; j < WN; ++j) win[j] = own ? us[at(j)] : 0.f;
    const float* w = cn.w[0];
#pragma unroll
    for (int o4 = 0; o4 < C / 4; ++o4) {
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int o = 4 * o4 + e;
          float a = cn.b[0] ? cn.b[0][o] : 0.f;
#pragma unroll
          for (int k = 0; k < K; ++k) a += w[o * K + k] * win[p + k];
          v[e] = gelu_exact(a);
        }
        if (own && i0 + p < n)
          *reinterpret_cast<v4f*>(act + (i0 + p) * CS + 4 * o4) = v4f{v[0], v[1], v[2], v[3]};
      }
    }
  }
  __syncthreads();
  for (int l = 1; l < cn.layers - 1; ++l) {
    cnn_hidden_mfma<C, K, PW>(ws + size_t(l - 1) * C * (C * K + 4), cn.b[l], act, n);
    __syncthreads();
  }
  const float* w = cn.w[cn.layers - 1];
  const float bl = cn.b[cn.layers - 1] ? cn.b[cn.layers - 1][0] : 0.f;
  float v[PPT];
#pragma unroll
  for (int p = 0; p < PPT; ++p) v[p] = bl;
#pragma unroll
  for (int j = 0; j < WN; ++j) {
    const float* xr = act + (own ? at(j) : 0) * CS;
#pragma unroll
    for (int c4 = 0; c4 < C / 4; ++c4) {
      const v4f x = *reinterpret_cast<const v4f*>(xr + 4 * c4);
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          const int k = j - p;
          if (k >= 0 && k < K) v[p] += w[(4 * c4 + e) * K + k] * x[e];
        }
    }
  }
#pragma unroll
  for (int p = 0; p < PPT; ++p) s[p] = cn.scale * v[p];
}

// Block = one trajectory's team (blockDim.x threads, PPT points each).
// Writes traj[b][step] and, when corr is given, corr[b][step] = the
// correction of that step (simulate_rollout's second result).
// PW > 0 (float): hidden layers on the matrix cores with PW 16-position
// tiles per wavefront, their weights staged in LDS once for all steps.
template <typename T, int PPT, int C, int K, int PW>
__global__ void __launch_bounds__(kBlock) burgers_inc_fwd_kernel(BArgs<T> a, CnnArgs<T> cn, T* corr) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T red_v[2 * kMaxWaves];
  __shared__ int red_i[kMaxWaves];
  TeamCtx<T> c;
  c.team = blockDim.x;
  c.lt = threadIdx.x;
  c.b = blockIdx.x;
  c.active = true;
  c.wave_in_block = threadIdx.x / kWave;
  c.first_wave = 0;
  c.i0 = c.lt * PPT;
  c.n = a.n;
  const int n = a.n;
  T* us = reinterpret_cast<T*>(smem_raw);
  T* Fs = us + n;
  T* buf = Fs + n;  // PW == 0: [C][n]; PW > 0: staged hidden weights, then act [n][C + 4]
  const bool own = c.i0 < n;
  const int ns = a.scheme == INC_SSPRK3 ? 3 : 1;
  const T nu = a.nu_ptr ? a.nu_ptr[c.b] : a.nu;
  const T t0 = a.t_ptr ? a.t_ptr[c.b] : a.t0;
  T* act = buf;
  if constexpr (PW > 0) {
    // hidden layers' weights, once for the whole rollout: ws[l-1][o][k*C + c]
    constexpr int RS = C * K + 4;
    const int nh = cn.layers - 2;
    for (int e = threadIdx.x; e < nh * C * C; e += blockDim.x) {
      const int l = e / (C * C), oc = e - l * C * C, o = oc / C, cc = oc - o * C;
      const T* src = cn.w[l + 1] + size_t(oc) * K;
      T* dst = buf + size_t(l) * C * RS + o * RS + cc;
#pragma unroll
      for (int k = 0; k < K; ++k) dst[k * C] = src[k];
    }
    act = buf + size_t(nh) * C * RS;
  }
  const size_t row = size_t(c.b) * n;
  T u0[PPT], uc[PPT], sv[PPT], xv[PPT], R[PPT];
  T lsum = T(0);  // squared errors to the target (fused MSE)
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    const int i = c.i0 + p;
    const bool ok = own && i < n;
    u0[p] = ok ? a.u[row + i] : T(0);
    xv[p] = (ok && a.source == INC_SRC_LEARNING) ? a.x[i] : T(0);
  }
  for (int step = 0; step < a.steps; ++step) {
    // correction of this step from the current state
#pragma unroll
    for (int p = 0; p < PPT; ++p)
      if (c.i0 + p < n) us[c.i0 + p] = u0[p];
    __syncthreads();
    if constexpr (PW > 0) cnn_eval_mfma<C, K, PPT, PW>(cn, buf, us, act, n, c.i0, own, sv);
    else cnn_eval<T, C, K, PPT>(cn, us, buf, n, c.i0, own, sv);
    __syncthreads();  // us and the activations are free again
    const size_t at = (size_t(c.b) * a.steps + step) * n;
    if (corr && own) {
#pragma unroll
      for (int p = 0; p < PPT; ++p)
        if (c.i0 + p < n) corr[at + c.i0 + p] = sv[p];
    }
    // one step of the solver with the correction held over its stages
    const T ts0 = t0 + T(step) * a.dt;
#pragma unroll
    for (int p = 0; p < PPT; ++p) uc[p] = u0[p];
    for (int st = 0; st < ns; ++st) {
      if (a.stages && own) {  // stage states for the training adjoint: [b][step][stage][n]
        T* dst = a.stages + ((size_t(c.b) * a.steps + step) * ns + st) * n;
#pragma unroll
        for (int p = 0; p < PPT; ++p)
          if (c.i0 + p < n) dst[c.i0 + p] = uc[p];
      }
      const T ts = st == 0 ? ts0 : (st == 1 ? ts0 + a.dt : ts0 + T(0.5) * a.dt);
      rhs_eval<T, PPT, INC_WENO>(a, c, uc, sv, xv, ts, nu, a.parts, us, Fs, red_v, red_i, R);
      __syncthreads();
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        T v = uc[p] + a.dt * R[p];
        if (ns == 1 || st == 0) uc[p] = v;
        else if (st == 1) uc[p] = T(0.75) * u0[p] + T(0.25) * v;
        else uc[p] = T(1.0 / 3.0) * u0[p] + T(2.0 / 3.0) * v;
      }
    }
#pragma unroll
    for (int p = 0; p < PPT; ++p) u0[p] = uc[p];
    if (a.traj && own) {
#pragma unroll
      for (int p = 0; p < PPT; ++p)
        if (c.i0 + p < n) a.traj[at + c.i0 + p] = u0[p];
    }
    if (a.target && own) {
#pragma unroll
      for (int p = 0; p < PPT; ++p)
        if (c.i0 + p < n) {
          const T dd = u0[p] - a.target[at + c.i0 + p];
          lsum += dd * dd;
        }
    }
  }
  if (own) {
#pragma unroll
    for (int p = 0; p < PPT; ++p)
      if (c.i0 + p < n) a.out[row + c.i0 + p] = u0[p];
  }
  if (a.target) {
    // one partial per wavefront of the block, unused slots zero
    lsum = wave_sum(lsum, kWave);
    if (c.lt % kWave == 0) a.loss_parts[size_t(c.b) * kLossSlots + c.lt / kWave] = lsum;
    if (c.lt == 0)
      for (int k = blockDim.x / kWave; k < kLossSlots; ++k) a.loss_parts[size_t(c.b) * kLossSlots + k] = T(0);
  }
}

// ---- INC training rollout: the network's transpose inside the adjoint kernel ----------
// Gradient of simulate_rollout(model, solver, u0, steps, 'INC')
// (trainer/trainer_1d.py:154-183) with respect to u0 and the correction
// network's weights, for a circular Conv1d + GELU network of width C
// (float).  One block of 256 threads per trajectory runs backwards in time.
// Per step: the solver step's transpose from its stored stage states gives
// the cotangents of the step's input state and of its correction s; the
// network is evaluated again at that input state (stage 0) with every
// layer's pre-activation kept in LDS; its transpose adds the state's
// cotangent through s and the weight gradients.  The hidden layers run on
// the matrix cores; their weight gradients accumulate over all steps in MFMA
// tiles held in registers, the first and last layers' and the biases' in
// per-thread partial sums, and each trajectory writes its own gradient row.
constexpr int kTrainBlock = 256;

// The GELU and its derivative as the adjoint re-evaluates them: gelu(v) =
// v Phi(v), gelu'(v) = Phi(v) + v phi(v), Phi(v) = (1 + erf(v / sqrt 2)) / 2,
// phi(v) = exp(-v^2 / 2) / sqrt(2 pi).  erf by Abramowitz & Stegun 7.1.26
// (Hastings): erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2),
// t = 1 / (1 + p x), x >= 0, |error| <= 1.5e-7, odd extension for x < 0.
// exp(-x^2) at x = |v| / sqrt 2 is phi's exponential as well, so both come
// from one exp and one reciprocal, with no branch (the library erff takes a
// divergent branch and ~30 instructions).  The forward kernels keep erff;
// the recomputed activations differ from theirs by a few 1e-7, well
// inside the gradients' float32 rounding.  Returns gelu(v), writes gelu'(v).
__device__ __forceinline__ float gelu_with_grad(float v, float& d) {
  const float x = fabsf(v) * 0.70710678118654752440f;
  const float t = __builtin_amdgcn_rcpf(1.f + 0.3275911f * x);
  const float e = __expf(-x * x);
  const float poly =
      t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
  const float ea = 1.f - poly * e;  // erf(x)
  const float ph = 0.5f * (1.f + (v < 0.f ? -ea : ea));
  d = ph + v * 0.39894228040143267794f * e;
  return v * ph;
}

__device__ __forceinline__ float gelu_adj(float v) {
  float d;
  return gelu_with_grad(v, d);
}

// One wavefront's part of a C -> C circular convolution of `in` (position-
// major, row stride C + 4) on the matrix cores: acc[m][t] holds
// D[16m + 4q + r][16 (wave*PW + t) + lr] (lane (lr, q), r = 0..3).
// TRANS == false: D[o][i] = sum_{k,c} W[o][c][k] in[i + k - P][c];
// TRANS == true (the input gradient): D[c][i] = sum_{k,o} W[o][c][k]
// in[i - k + P][o], i.e. the flipped transposed kernel, whose A operand is
// read from the same staged weights ws[o][k*C + c] (row stride C*K + 4)
// with four scalar reads per chunk.  HALO: `in` carries P wrapped rows on
// either side (rows -P .. n + P - 1), so no row index wraps.  Positions past
// n are computed from clamped rows and dropped by the caller.
template <int C, int K, int PW, bool TRANS, bool HALO>
__device__ __forceinline__ void conv_tiles(const float* ws, const float* in, int n, v4f (&acc)[C / 16][PW]) {
  constexpr int CT = C / 16, P = K / 2, CS = C + 4, RS = C * K + 4, NCH = C * K / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, q = lane >> 4;
  int rb[PW];
#pragma unroll
  for (int t = 0; t < PW; ++t) rb[t] = min((wave * PW + t) * 16 + lr, n - 1) - P;
  auto row = [&](int i) { return i < 0 ? i + n : (i >= n ? i - n : i); };  // i in (-n, 2n)
#pragma unroll
  for (int m = 0; m < CT; ++m)
#pragma unroll
    for (int t = 0; t < PW; ++t) acc[m][t] = v4f{0.f, 0.f, 0.f, 0.f};
  int k = 0, c0 = 0;
  auto load = [&](v4f* fa, v4f* fb) __attribute__((always_inline)) {
    const int kw = min(k, K - 1);  // the read after the last chunk stays inside the weights (unused)
    if constexpr (!TRANS) {
      const float* wl = ws + lr * RS + kw * C + c0 + 4 * q;
#pragma unroll
      for (int m = 0; m < CT; ++m) fa[m] = *reinterpret_cast<const v4f*>(wl + m * 16 * RS);
    } else {
      const float* wl = ws + (c0 + 4 * q) * RS + (K - 1 - kw) * C + lr;
#pragma unroll
      for (int m = 0; m < CT; ++m)
        fa[m] = v4f{wl[m * 16], wl[RS + m * 16], wl[2 * RS + m * 16], wl[3 * RS + m * 16]};
    }
#pragma unroll
    for (int t = 0; t < PW; ++t)
      fb[t] = *reinterpret_cast<const v4f*>(in + (HALO ? rb[t] + kw : row(rb[t] + k)) * CS + c0 + 4 * q);
  };
  auto next = [&]() __attribute__((always_inline)) {
    c0 += 16;
    if (c0 == C) {
      c0 = 0;
      ++k;
    }
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
  // two register sets in turn, the next chunk's reads issued before the
  // current chunk's MFMAs; the reads are unconditional (a conditional read
  // makes every MFMA wait for the reads in flight); the one after the last
  // chunk reads valid rows and weights and is unused
  v4f a0[CT], b0[PW], a1[CT], b1[PW];
  load(a0, b0);
  for (int s = 0; s < NCH; s += 2) {
    next();
    load(a1, b1);
    __builtin_amdgcn_sched_barrier(0);
    mma(a0, b0);
    __builtin_amdgcn_sched_barrier(0);
    if (s + 1 >= NCH) break;
    next();
    load(a0, b0);
    __builtin_amdgcn_sched_barrier(0);
    mma(a1, b1);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// This wavefront's tiles of one hidden layer's weight gradient, added to
// acc: dW[o][c][k] += sum_j gy[j][o] x[j + k - P][c] (gy, x position-major,
// row stride C + 4).  The layer's (C/16) x (C*K/16) tiles of D[o][k*C + c]
// go round the four wavefronts (tile W + 4t of wavefront W); lane (lr, q)
// holds D[16m + 4q + r][16ct + lr].  The reduction runs over positions, four
// per MFMA, all tiles of the wavefront interleaved.  W is a template
// argument so that every tile's row, column and tap are constants, and x
// carries P wrapped rows on either side (rows -P .. n + P - 1): each operand
// is one LDS read at a constant offset from the lane's row, with no wrap.
template <int C, int K, int TPW, int W>
__device__ __forceinline__ void wgrad_wave(const float* gy, const float* x, int n, v4f (&acc)[TPW]) {
  constexpr int P = K / 2, CS = C + 4, NCOL = C * K / 16, TT = (C / 16) * NCOL;
  const int lane = threadIdx.x & 63, lr = lane & 15, q = lane >> 4;
  auto quad = [&](int j0, bool tail) __attribute__((always_inline)) {
    const int j = j0 + q, jc = tail ? min(j, n - 1) : j;
    const float* ga = gy + jc * CS + lr;
    const float* xb[K];
#pragma unroll
    for (int k = 0; k < K; ++k) xb[k] = x + (jc + k - P) * CS + lr;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int id = W + 4 * t;
      if (id >= TT) continue;
      const int m = id / NCOL, col = (id - m * NCOL) * 16, k = col / C, c0 = col - k * C;
      float av = ga[m * 16];
      if (tail && j >= n) av = 0.f;
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, xb[k][c0], acc[t], 0, 0, 0);
    }
  };
  // two quads per iteration, so the second's reads can issue under the
  // first's MFMAs; then the remaining whole quad and the ragged tail
  const int n4 = n & ~3, n8 = n & ~7;
  int j0 = 0;
  for (; j0 < n8; j0 += 8) {
    quad(j0, false);
    quad(j0 + 4, false);
  }
  if (j0 < n4) quad(j0, false);
  if (n4 < n) quad(n4, true);
}

template <int C, int K, int TPW>
__device__ __forceinline__ void wgrad_tiles(const float* gy, const float* x, int n, v4f (&acc)[TPW]) {
  switch (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) {
    case 0: wgrad_wave<C, K, TPW, 0>(gy, x, n, acc); break;
    case 1: wgrad_wave<C, K, TPW, 1>(gy, x, n, acc); break;
    case 2: wgrad_wave<C, K, TPW, 2>(gy, x, n, acc); break;
    default: wgrad_wave<C, K, TPW, 3>(gy, x, n, acc); break;
  }
}

template <int C, int K, int NH, int PW>
__global__ void __launch_bounds__(kTrainBlock) burgers_inc_adj_kernel(BArgs<float> a, CnnArgs<float> cn,
                                                                      const float* __restrict__ g_corr,
                                                                      float* __restrict__ g_w) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ float red_v[2 * kMaxWaves];
  __shared__ int red_i[kMaxWaves];
  constexpr int P = K / 2, CS = C + 4, RS = C * K + 4, Q = C / 4;
  constexpr int NCOL = C * K / 16, TT = (C / 16) * NCOL, TPW = (TT + 3) / 4;
  static_assert(kTrainBlock % Q == 0, "a thread's channel quad must not change between items");
  // per-thread partial sums (channel quad fixed per thread): first layer
  // dW0[4][K], db0[4]; last layer dWL[4][K], dbL; hidden biases db[NH][4]
  constexpr int V0 = 4 * K + 4, VL = 4 * K + 1, V = V0 + VL + 4 * NH;
  const int n = a.n, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  TeamCtx<float> c;
  c.team = kTrainBlock;
  c.lt = tid;
  c.b = blockIdx.x;
  c.active = true;
  c.wave_in_block = wave;
  c.first_wave = 0;
  c.i0 = tid;
  c.n = n;
  const bool own = tid < n;
  float* ws = reinterpret_cast<float*>(smem_raw);  // [NH][C][RS]
  float* Y = ws + NH * C * RS;                      // [NH + 1][n][CS]: pre-activations, then their cotangents
  float* Xb = Y + (NH + 1) * n * CS;                // [n + 2P][CS]: activations (the solver's scratch before them)
  float* X = Xb + P * CS;                           // row j of the activations, rows -P .. n + P - 1 (wrapped)
  float* G = Xb + (n + 2 * P) * CS;                 // [n][CS]: cotangent of the activations
  float* un = G + n * CS;                           // [n]: the step's input state
  float* gb = un + n;                               // [n]: cotangent of the network's output
  float* us = Xb;
  float* cs = Xb + n;
  float* Cw = Xb + 2 * n;
  // an activation row goes to X with its wrapped copy when it lies within P
  // of either end (n >= 8 > 2P: at most one copy)
  auto xstore = [&](int j, int col, v4f v) __attribute__((always_inline)) {
    *reinterpret_cast<v4f*>(X + j * CS + col) = v;
    if (j < P) *reinterpret_cast<v4f*>(X + (j + n) * CS + col) = v;
    if (j >= n - P) *reinterpret_cast<v4f*>(X + (j - n) * CS + col) = v;
  };
  for (int e = tid; e < NH * C * C; e += kTrainBlock) {
    const int l = e / (C * C), oc = e - l * C * C, o = oc / C, cc = oc - o * C;
    const float* src = cn.w[l + 1] + size_t(oc) * K;
    float* dst = ws + size_t(l) * C * RS + o * RS + cc;
#pragma unroll
    for (int k = 0; k < K; ++k) dst[k * C] = src[k];
  }
  auto row = [&](int i) { return i < 0 ? i + n : (i >= n ? i - n : i); };
  const int quad = tid % Q;
  const int nitems = n * Q;
  float pw0[4][K], pb0[4], pwL[4][K], pbL = 0.f, pbh[NH > 0 ? NH : 1][4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    pb0[e] = 0.f;
#pragma unroll
    for (int k = 0; k < K; ++k) pw0[e][k] = pwL[e][k] = 0.f;
#pragma unroll
    for (int l = 0; l < (NH > 0 ? NH : 1); ++l) pbh[l][e] = 0.f;
  }
  v4f wacc[NH > 0 ? NH : 1][TPW];
#pragma unroll
  for (int l = 0; l < (NH > 0 ? NH : 1); ++l)
#pragma unroll
    for (int t = 0; t < TPW; ++t) wacc[l][t] = v4f{0.f, 0.f, 0.f, 0.f};

  const int ns = a.scheme == INC_SSPRK3 ? 3 : 1;
  const float nu = a.nu_ptr ? a.nu_ptr[c.b] : a.nu;
  const float dt = a.dt;
  const int parts = INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE;
  float lam, gs, st[3], cb, r, b1, b2;
  // cotangent of output `step`: g_traj, plus *loss_scale * (traj - target)
  // with a target (the fused MSE)
  const float lsc = a.target ? *a.loss_scale : 0.f;
  auto gload = [&](int step) {
    const size_t at = (size_t(c.b) * a.steps + step) * n + tid;
    float g = (own && a.g_traj) ? a.g_traj[at] : 0.f;
    if (own && a.target) g += lsc * (a.traj[at] - a.target[at]);
    return g;
  };
  lam = gload(a.steps - 1);
  for (int step = a.steps - 1; step >= 0; --step) {
    // ---- the solver step's transpose (its stage states from the forward)
    const float* src = a.stages + (size_t(c.b) * a.steps + step) * ns * n;
    st[0] = own ? src[tid] : 0.f;
    st[1] = (own && ns == 3) ? src[n + tid] : 0.f;
    st[2] = (own && ns == 3) ? src[2 * n + tid] : 0.f;
    gs = 0.f;
    if (ns == 1) {
      cb = dt * lam;
      rhs_vjp<float, 1, INC_WENO>(a, c, &st[0], &cb, nu, parts, us, cs, Cw, red_v, red_i, &r);
      __syncthreads();
      gs += cb;
      lam += r;
    } else {
      cb = (2.f / 3.f) * dt * lam;
      rhs_vjp<float, 1, INC_WENO>(a, c, &st[2], &cb, nu, parts, us, cs, Cw, red_v, red_i, &r);
      __syncthreads();
      gs += cb;
      b2 = (2.f / 3.f) * lam + r;
      lam = (1.f / 3.f) * lam;
      cb = 0.25f * dt * b2;
      rhs_vjp<float, 1, INC_WENO>(a, c, &st[1], &cb, nu, parts, us, cs, Cw, red_v, red_i, &r);
      __syncthreads();
      gs += cb;
      b1 = 0.25f * b2 + r;
      lam += 0.75f * b2;
      cb = dt * b1;
      rhs_vjp<float, 1, INC_WENO>(a, c, &st[0], &cb, nu, parts, us, cs, Cw, red_v, red_i, &r);
      __syncthreads();
      gs += cb;
      lam += b1 + r;
    }
    // ---- the network at the step's input state, pre-activations kept
    if (own) {
      un[tid] = st[0];
      const float gc = g_corr ? g_corr[(size_t(c.b) * a.steps + step) * n + tid] : 0.f;
      gb[tid] = cn.scale * (gs + gc);
    }
    __syncthreads();
    {
      const float* w0 = cn.w[0];
      for (int e = tid; e < nitems; e += kTrainBlock) {
        const int j = e / Q;
        float y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) y[u] = cn.b[0] ? cn.b[0][4 * quad + u] : 0.f;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float uv = un[row(j + k - P)];
#pragma unroll
          for (int u = 0; u < 4; ++u) y[u] += w0[(4 * quad + u) * K + k] * uv;
        }
        *reinterpret_cast<v4f*>(Y + j * CS + 4 * quad) = v4f{y[0], y[1], y[2], y[3]};
        xstore(j, 4 * quad, v4f{gelu_adj(y[0]), gelu_adj(y[1]), gelu_adj(y[2]), gelu_adj(y[3])});
      }
    }
    __syncthreads();
#pragma unroll
    for (int l = 1; l <= NH; ++l) {
      v4f acc[C / 16][PW];
      conv_tiles<C, K, PW, false, true>(ws + size_t(l - 1) * C * RS, X, n, acc);
      __syncthreads();  // every wavefront has read X
      const int lr = lane & 15, q = lane >> 4;
#pragma unroll
      for (int m = 0; m < C / 16; ++m) {
        const int o = m * 16 + 4 * q;
        const v4f bo = cn.b[l] ? *reinterpret_cast<const v4f*>(cn.b[l] + o) : v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < PW; ++t) {
          const int i = (wave * PW + t) * 16 + lr;
          if (i < n) {
            const v4f v = acc[m][t] + bo;
            *reinterpret_cast<v4f*>(Y + (size_t(l) * n + i) * CS + o) = v;
            xstore(i, o, v4f{gelu_adj(v[0]), gelu_adj(v[1]), gelu_adj(v[2]), gelu_adj(v[3])});
          }
        }
      }
      __syncthreads();
    }
    // ---- transpose of the last layer (C -> 1): X holds its input.  From
    // here on a pre-activation slot is overwritten by the GELU derivative at
    // it (its one later use), one layer ahead of the pass that reads it.
    {
      const float* wL = cn.w[NH + 1];
      float* Yt = Y + size_t(NH) * n * CS;
      for (int e = tid; e < nitems; e += kTrainBlock) {
        const int j = e / Q;
        const v4f xv = *reinterpret_cast<const v4f*>(X + j * CS + 4 * quad);
        {
          const v4f yv = *reinterpret_cast<const v4f*>(Yt + j * CS + 4 * quad);
          float dv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) gelu_with_grad(yv[u], dv[u]);
          *reinterpret_cast<v4f*>(Yt + j * CS + 4 * quad) = v4f{dv[0], dv[1], dv[2], dv[3]};
        }
        float g[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float gv = gb[row(j - k + P)];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            pwL[u][k] += xv[u] * gv;
            g[u] += wL[(4 * quad + u) * K + k] * gv;
          }
        }
        pbL += gb[j];
        *reinterpret_cast<v4f*>(G + j * CS + 4 * quad) = v4f{g[0], g[1], g[2], g[3]};
      }
    }
    __syncthreads();
    // ---- hidden layers, last to first
#pragma unroll
    for (int l = NH; l >= 1; --l) {
      float* Yl = Y + size_t(l) * n * CS;
      float* Yp = Y + size_t(l - 1) * n * CS;
      for (int e = tid; e < nitems; e += kTrainBlock) {
        const int j = e / Q;
        const v4f gv = *reinterpret_cast<const v4f*>(G + j * CS + 4 * quad);
        const v4f dl = *reinterpret_cast<const v4f*>(Yl + j * CS + 4 * quad);  // gelu'(Y_l)
        const v4f yp = *reinterpret_cast<const v4f*>(Yp + j * CS + 4 * quad);
        v4f gy;
        float ap[4], dp[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          gy[u] = gv[u] * dl[u];
          pbh[l - 1][u] += gy[u];
          ap[u] = gelu_with_grad(yp[u], dp[u]);
        }
        *reinterpret_cast<v4f*>(Yl + j * CS + 4 * quad) = gy;
        *reinterpret_cast<v4f*>(Yp + j * CS + 4 * quad) = v4f{dp[0], dp[1], dp[2], dp[3]};
        xstore(j, 4 * quad, v4f{ap[0], ap[1], ap[2], ap[3]});
      }
      __syncthreads();
      wgrad_tiles<C, K, TPW>(Yl, X, n, wacc[l - 1]);
      v4f acc[C / 16][PW];
      conv_tiles<C, K, PW, true, false>(ws + size_t(l - 1) * C * RS, Yl, n, acc);
      const int lr = lane & 15, q = lane >> 4;
#pragma unroll
      for (int m = 0; m < C / 16; ++m)
#pragma unroll
        for (int t = 0; t < PW; ++t) {
          const int i = (wave * PW + t) * 16 + lr;
          if (i < n) *reinterpret_cast<v4f*>(G + i * CS + m * 16 + 4 * q) = acc[m][t];
        }
      __syncthreads();
    }
    // ---- first layer (1 -> C): its cotangent, weight gradient and the state's cotangent
    for (int e = tid; e < nitems; e += kTrainBlock) {
      const int j = e / Q;
      const v4f gv = *reinterpret_cast<const v4f*>(G + j * CS + 4 * quad);
      const v4f dv = *reinterpret_cast<const v4f*>(Y + j * CS + 4 * quad);  // gelu'(Y_0)
      v4f gy;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        gy[u] = gv[u] * dv[u];
        pb0[u] += gy[u];
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float uv = un[row(j + k - P)];
#pragma unroll
        for (int u = 0; u < 4; ++u) pw0[u][k] += gy[u] * uv;
      }
      *reinterpret_cast<v4f*>(Y + j * CS + 4 * quad) = gy;
    }
    __syncthreads();
    if (own) {
      const float* w0 = cn.w[0];
      float g = 0.f;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float* yr = Y + row(tid - k + P) * CS;
#pragma unroll
        for (int q4 = 0; q4 < Q; ++q4) {
          const v4f gy = *reinterpret_cast<const v4f*>(yr + 4 * q4);
#pragma unroll
          for (int u = 0; u < 4; ++u) g += w0[(4 * q4 + u) * K + k] * gy[u];
        }
      }
      lam += g;
    }
    if (step > 0) lam += gload(step - 1);
    __syncthreads();  // Y and un are read above; the next step's transpose writes X
  }
  if (own) a.out[size_t(c.b) * n + tid] = lam;

  // ---- gradient row of this trajectory: [w0, b0, w1, b1, ..., wL, bL] as
  // the Conv1d tensors are laid out
  float* gw = g_w + size_t(c.b) * (size_t(C) * K + C + size_t(NH) * (size_t(C) * C * K + C) + C * K + 1);
  constexpr int OFF0 = 0, OFFH = C * K + C, HSZ = C * C * K + C, OFFL = OFFH + NH * HSZ;
  {
    const int lr = lane & 15, q = lane >> 4;
#pragma unroll
    for (int l = 0; l < NH; ++l)
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        const int id = wave + 4 * t;
        if (id >= TT) continue;
        const int m = id / NCOL, col = (id - m * NCOL) * 16 + lr, k = col / C, cc = col - k * C;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int o = m * 16 + 4 * q + rr;
          gw[OFFH + l * HSZ + (o * C + cc) * K + k] = wacc[l][t][rr];
        }
      }
  }
  // per-thread partials: sum over the threads of one channel quad (lanes
  // l, l + Q, ... of each wavefront by shuffles, then the four wavefronts)
  float v[V];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      v[u * K + k] = pw0[u][k];
      v[V0 + u * K + k] = pwL[u][k];
    }
    v[4 * K + u] = pb0[u];
#pragma unroll
    for (int l = 0; l < NH; ++l) v[V0 + VL + 4 * l + u] = pbh[l][u];
  }
  v[V0 + 4 * K] = pbL;
#pragma unroll
  for (int i = 0; i < V; ++i)
    for (int off = Q; off < kWave; off <<= 1) v[i] += shfl_xor(v[i], off);
  float* red = Y;  // [4 waves][Q][V]
  if (lane < Q) {
#pragma unroll
    for (int i = 0; i < V; ++i) red[(wave * Q + lane) * V + i] = v[i];
  }
  __syncthreads();
  for (int e = tid; e < Q * V; e += kTrainBlock) {
    const int qd = e / V, i = e - qd * V;
    const float s = red[qd * V + i] + red[(Q + qd) * V + i] + red[(2 * Q + qd) * V + i] + red[(3 * Q + qd) * V + i];
    if (i < 4 * K) {
      const int u = i / K, k = i - u * K;
      gw[OFF0 + (4 * qd + u) * K + k] = s;
    } else if (i < V0) {
      gw[OFF0 + C * K + 4 * qd + (i - 4 * K)] = s;
    } else if (i < V0 + 4 * K) {
      const int u = (i - V0) / K, k = (i - V0) - u * K;
      gw[OFFL + (4 * qd + u) * K + k] = s;
    } else if (i == V0 + 4 * K) {
      if (qd == 0) gw[OFFL + C * K] = s;
    } else {
      const int l = (i - V0 - VL) / 4, u = (i - V0 - VL) - 4 * l;
      gw[OFFH + l * HSZ + C * C * K + 4 * qd + u] = s;
    }
  }
}

// Right-hand side parts of one state (calc_nonlinear / calc_linear / source).
template <typename T, int PPT, int CONV>
__global__ void __launch_bounds__(kBlock) burgers_rhs_kernel(BArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T red_v[2 * kMaxWaves];
  __shared__ int red_i[kMaxWaves];
  TeamCtx<T> c = make_ctx<T>(a.team, a.n, a.batch, PPT);
  T* us = team_lds(reinterpret_cast<T*>(smem_raw), a.team, a.n, 2);
  T* Fs = us + a.n;
  const int n = a.n;
  const bool own = c.active && c.i0 < n;
  const T nu = a.nu_ptr ? (c.active ? a.nu_ptr[c.b] : T(0)) : a.nu;
  const T t0 = a.t_ptr ? (c.active ? a.t_ptr[c.b] : T(0)) : a.t0;
  T u[PPT], sv[PPT], xv[PPT], R[PPT];
  const size_t row = size_t(c.b) * n;
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    int i = c.i0 + p;
    bool ok = own && i < n;
    u[p] = ok ? a.u[row + i] : T(0);
    sv[p] = (ok && a.s) ? a.s[row + i] : T(0);
    xv[p] = (ok && a.source == INC_SRC_LEARNING) ? a.x[i] : T(0);
  }
  rhs_eval<T, PPT, CONV>(a, c, u, a.s ? sv : nullptr, xv, t0, nu, a.parts, us, Fs, red_v, red_i, R);
  if (own) {
#pragma unroll
    for (int p = 0; p < PPT; ++p)
      if (c.i0 + p < n) a.out[row + c.i0 + p] = R[p];
  }
}

template <typename T, int PPT, int CONV>
__global__ void __launch_bounds__(kBlock) burgers_rhs_adj_kernel(BArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T red_v[2 * kMaxWaves];
  __shared__ int red_i[kMaxWaves];
  TeamCtx<T> c = make_ctx<T>(a.team, a.n, a.batch, PPT);
  T* us = team_lds(reinterpret_cast<T*>(smem_raw), a.team, a.n, 8);
  T* cs = us + a.n;
  T* C = cs + a.n;
  const int n = a.n;
  const bool own = c.active && c.i0 < n;
  const T nu = a.nu_ptr ? (c.active ? a.nu_ptr[c.b] : T(0)) : a.nu;
  T u[PPT], g[PPT], ub[PPT];
  const size_t row = size_t(c.b) * n;
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    int i = c.i0 + p;
    bool ok = own && i < n;
    u[p] = ok ? a.u[row + i] : T(0);
    g[p] = ok ? a.g_traj[row + i] : T(0);
  }
  T nubar = T(0);
  rhs_vjp<T, PPT, CONV>(a, c, u, g, nu, a.parts, us, cs, C, red_v, red_i, ub, a.g_nu != nullptr, &nubar);
  if (own) {
#pragma unroll
    for (int p = 0; p < PPT; ++p)
      if (c.i0 + p < n) a.out[row + c.i0 + p] = ub[p];
  }
  if (a.g_nu) team_sum_store(c, nubar, red_v, a.g_nu);
}

// Reverse sweep over `steps` steps from the saved stage states.
// NU: also the cotangent of nu (a separate instance: the accumulator's
// register would take the plain adjoint past the count that keeps six
// wavefronts per SIMD, 84 -> 88 VGPRs, 13.9 -> 18.9 ms at B=8192, N=256,
// 200 SSP-RK3 steps).
template <typename T, int PPT, int CONV, bool NU>
__global__ void __launch_bounds__(kBlock) burgers_adj_kernel(BArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T red_v[2 * kMaxWaves];
  __shared__ int red_i[kMaxWaves];
  TeamCtx<T> c = make_ctx<T>(a.team, a.n, a.batch, PPT);
  T* us = team_lds(reinterpret_cast<T*>(smem_raw), a.team, a.n, 8);
  T* cs = us + a.n;
  T* C = cs + a.n;
  const int n = a.n;
  const bool own = c.active && c.i0 < n;
  const int ns = a.scheme == INC_SSPRK3 ? 3 : 1;
  const T nu = a.nu_ptr ? (c.active ? a.nu_ptr[c.b] : T(0)) : a.nu;
  const T dt = a.dt;
  const int parts = INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE;
  T lam[PPT], gs[PPT], st[3][PPT], cb[PPT], r[PPT], b1[PPT], b2[PPT];
  __shared__ T nu_acc[NU ? kBlock : 1];  // per-thread cotangent of nu (see rhs_vjp)
  constexpr bool want_nu = NU;
  T* const nub = &nu_acc[NU ? threadIdx.x : 0];
  if (want_nu) *nub = T(0);
  // cotangent of output `step`: g_traj, plus *loss_scale * (traj - target)
  // with a target
  const T lsc = a.target ? *a.loss_scale : T(0);
  auto gload = [&](int step, T* dst) {
    const size_t at = (size_t(c.b) * a.steps + step) * n;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      int i = c.i0 + p;
      dst[p] = (own && i < n && a.g_traj) ? a.g_traj[at + i] : T(0);
      if (a.target && own && i < n) dst[p] += lsc * (a.traj[at + i] - a.target[at + i]);
    }
  };
  gload(a.steps - 1, lam);
#pragma unroll
  for (int p = 0; p < PPT; ++p) gs[p] = T(0);
  // Transpose of one (sub-)step of size h whose stage states start at src.
  // on == false: barriers only, the team's cotangents are left as they are.
  // (always inlined: as a called function its arrays would live in scratch)
  auto substep_vjp = [&](T h, const T* src, bool on) __attribute__((always_inline)) {
    for (int k = 0; k < ns; ++k) {
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        int i = c.i0 + p;
        st[k][p] = (on && own && i < n) ? src[size_t(k) * n + i] : T(0);
      }
    }
    if (ns == 1) {
#pragma unroll
      for (int p = 0; p < PPT; ++p) cb[p] = on ? h * lam[p] : T(0);
      rhs_vjp<T, PPT, CONV>(a, c, st[0], cb, nu, parts, us, cs, C, red_v, red_i, r, want_nu, nub);
      __syncthreads();
      if (on) {
#pragma unroll
        for (int p = 0; p < PPT; ++p) { gs[p] += cb[p]; lam[p] = lam[p] + r[p]; }
      }
    } else {
      // u' = u0/3 + 2/3 (u2 + h R(u2))
#pragma unroll
      for (int p = 0; p < PPT; ++p) cb[p] = on ? T(2.0 / 3.0) * h * lam[p] : T(0);
      rhs_vjp<T, PPT, CONV>(a, c, st[2], cb, nu, parts, us, cs, C, red_v, red_i, r, want_nu, nub);
      __syncthreads();
      if (on) {
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          gs[p] += cb[p];
          b2[p] = T(2.0 / 3.0) * lam[p] + r[p];
          lam[p] = T(1.0 / 3.0) * lam[p];  // running cotangent of u0
        }
      }
      // u2 = 3/4 u0 + 1/4 (u1 + h R(u1))
#pragma unroll
      for (int p = 0; p < PPT; ++p) cb[p] = on ? T(0.25) * h * b2[p] : T(0);
      rhs_vjp<T, PPT, CONV>(a, c, st[1], cb, nu, parts, us, cs, C, red_v, red_i, r, want_nu, nub);
      __syncthreads();
      if (on) {
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          gs[p] += cb[p];
          b1[p] = T(0.25) * b2[p] + r[p];
          lam[p] += T(0.75) * b2[p];
        }
      }
      // u1 = u0 + h R(u0)
#pragma unroll
      for (int p = 0; p < PPT; ++p) cb[p] = on ? h * b1[p] : T(0);
      rhs_vjp<T, PPT, CONV>(a, c, st[0], cb, nu, parts, us, cs, C, red_v, red_i, r, want_nu, nub);
      __syncthreads();
      if (on) {
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          gs[p] += cb[p];
          lam[p] += b1[p] + r[p];
        }
      }
    }
  };
  for (int step = a.steps - 1; step >= 0; --step) {
    if (a.cfl <= T(0)) {
      substep_vjp(dt, a.stages + (size_t(c.b) * a.steps + step) * ns * n, true);
    } else {
      // the step's sub-steps in reverse; block-uniform over the sub-step index
      const size_t sb = size_t(c.b) * a.steps + step;
      const int nj = c.active ? a.nsub[sb] : 0;
      for (int jj = a.max_sub - 1; jj >= 0; --jj) {
        const bool on = jj < nj;
        if (!__syncthreads_or(on)) continue;
        const T h = on ? a.dts[sb * a.max_sub + jj] : T(0);
        substep_vjp(h, on ? a.stages + (sb * a.max_sub + jj) * ns * n : nullptr, on);
      }
    }
    if (step > 0) {
      gload(step - 1, cb);
#pragma unroll
      for (int p = 0; p < PPT; ++p) lam[p] += cb[p];
    }
  }
  if (own) {
    const size_t row = size_t(c.b) * n;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      int i = c.i0 + p;
      if (i < n) {
        a.out[row + i] = lam[p];
        if (a.g_s) a.g_s[row + i] = gs[p];
      }
    }
  }
  if (want_nu) team_sum_store(c, *nub, red_v, a.g_nu);
}

// ---- ragged batches ------------------------------------------------------------
// Trajectories of different length n_b, viscosity, dt and horizon packed end to
// end (trajectory b occupies [off_b, off_b + n_b) of every packed array).  One
// block of `team` threads owns one trajectory; the per-thread point count is
// chosen per trajectory (uniform within the block).  Blocks are issued in the
// order of `order` (longest first), so the hardware dispatcher acts as the work
// queue.  Packed per-step buffers (trajectory, stages) have stride `total`.
template <typename T>
struct RArgs {
  int count, scheme, steps, step0, team;
  int64_t total;
  const int* n;
  const int64_t* off;
  const int* order;
  const int* horizon;
  const T* nu;
  const T* dt;
  const T* dx;
  const T* u;
  const T* s;
  T* out;
  T* traj;
  T* stages;
  const T* g_traj;
  T* g_s;
};

template <typename T>
__device__ __forceinline__ BArgs<T> ragged_bargs(const RArgs<T>& r, int b, int n) {
  BArgs<T> a{};
  a.batch = 1;
  a.n = n;
  a.team = r.team;
  a.scheme = r.scheme;
  a.source = INC_SRC_NONE;
  a.parts = INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE;
  a.dt = r.dt[b];
  T dx = r.dx[b];
  a.inv_dx = T(1) / dx;
  a.inv_dx2 = T(1) / (dx * dx);
  a.half_inv_dx = T(0.5) / dx;
  a.nu = r.nu[b];
  return a;
}

template <typename T>
__device__ __forceinline__ TeamCtx<T> ragged_ctx(int n, int ppt) {
  TeamCtx<T> c;
  c.team = blockDim.x;
  c.lt = threadIdx.x;
  c.b = 0;
  c.active = true;
  c.wave_in_block = threadIdx.x / kWave;
  c.first_wave = 0;
  c.i0 = c.lt * ppt;
  c.n = n;
  return c;
}

template <typename T, int PPT, int CONV>
__device__ void ragged_fwd_body(const RArgs<T>& r, int b, T* us, T* Fs, T* red_v, int* red_i) {
  const int n = r.n[b];
  const int64_t off = r.off[b];
  const int hz = r.horizon ? r.horizon[b] : (r.step0 + r.steps);
  BArgs<T> a = ragged_bargs(r, b, n);
  TeamCtx<T> c = ragged_ctx<T>(n, PPT);
  const bool own = c.i0 < n;
  const int ns = r.scheme == INC_SSPRK3 ? 3 : 1;
  T u0[PPT], uc[PPT], sv[PPT], xv[PPT], R[PPT];
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    int i = c.i0 + p;
    bool ok = own && i < n;
    u0[p] = ok ? r.u[off + i] : T(0);
    sv[p] = (ok && r.s) ? r.s[off + i] : T(0);
    xv[p] = T(0);
  }
  for (int step = 0; step < r.steps; ++step) {
    if (r.step0 + step < hz) {
#pragma unroll
      for (int p = 0; p < PPT; ++p) uc[p] = u0[p];
      for (int st = 0; st < ns; ++st) {
        if (r.stages && own) {
          T* dst = r.stages + (int64_t(step) * ns + st) * r.total + off;
#pragma unroll
          for (int p = 0; p < PPT; ++p)
            if (c.i0 + p < n) dst[c.i0 + p] = uc[p];
        }
        rhs_eval<T, PPT, CONV>(a, c, uc, r.s ? sv : nullptr, xv, T(0), a.nu, a.parts, us, Fs, red_v, red_i, R);
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          T v = uc[p] + a.dt * R[p];
          if (ns == 1 || st == 0) uc[p] = v;
          else if (st == 1) uc[p] = T(0.75) * u0[p] + T(0.25) * v;
          else uc[p] = T(1.0 / 3.0) * u0[p] + T(2.0 / 3.0) * v;
        }
      }
#pragma unroll
      for (int p = 0; p < PPT; ++p) u0[p] = uc[p];
    }
    if (r.traj && own) {
      T* dst = r.traj + int64_t(step) * r.total + off;
#pragma unroll
      for (int p = 0; p < PPT; ++p)
        if (c.i0 + p < n) dst[c.i0 + p] = u0[p];
    }
  }
  if (own) {
#pragma unroll
    for (int p = 0; p < PPT; ++p)
      if (c.i0 + p < n) r.out[off + c.i0 + p] = u0[p];
  }
}

template <typename T, int PPT, int CONV>
__device__ void ragged_adj_body(const RArgs<T>& r, int b, T* us, T* cs, T* C, T* red_v, int* red_i) {
  const int n = r.n[b];
  const int64_t off = r.off[b];
  const int hz = r.horizon ? r.horizon[b] : (r.step0 + r.steps);
  BArgs<T> a = ragged_bargs(r, b, n);
  TeamCtx<T> c = ragged_ctx<T>(n, PPT);
  const bool own = c.i0 < n;
  const int ns = r.scheme == INC_SSPRK3 ? 3 : 1;
  const T dt = a.dt, nu = a.nu;
  T lam[PPT], gs[PPT], st[3][PPT], cb[PPT], rr[PPT], b1[PPT], b2[PPT];
  auto gload = [&](int step, T* dst) {
    const T* src = r.g_traj + int64_t(step) * r.total + off;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      int i = c.i0 + p;
      dst[p] = (own && i < n) ? src[i] : T(0);
    }
  };
  gload(r.steps - 1, lam);
#pragma unroll
  for (int p = 0; p < PPT; ++p) gs[p] = T(0);
  for (int step = r.steps - 1; step >= 0; --step) {
    if (r.step0 + step < hz) {
      for (int k = 0; k < ns; ++k) {
        const T* src = r.stages + (int64_t(step) * ns + k) * r.total + off;
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          int i = c.i0 + p;
          st[k][p] = (own && i < n) ? src[i] : T(0);
        }
      }
      if (ns == 1) {
#pragma unroll
        for (int p = 0; p < PPT; ++p) cb[p] = dt * lam[p];
        rhs_vjp<T, PPT, CONV>(a, c, st[0], cb, nu, a.parts, us, cs, C, red_v, red_i, rr);
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PPT; ++p) { gs[p] += cb[p]; lam[p] += rr[p]; }
      } else {
#pragma unroll
        for (int p = 0; p < PPT; ++p) cb[p] = T(2.0 / 3.0) * dt * lam[p];
        rhs_vjp<T, PPT, CONV>(a, c, st[2], cb, nu, a.parts, us, cs, C, red_v, red_i, rr);
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          gs[p] += cb[p];
          b2[p] = T(2.0 / 3.0) * lam[p] + rr[p];
          lam[p] = T(1.0 / 3.0) * lam[p];
        }
#pragma unroll
        for (int p = 0; p < PPT; ++p) cb[p] = T(0.25) * dt * b2[p];
        rhs_vjp<T, PPT, CONV>(a, c, st[1], cb, nu, a.parts, us, cs, C, red_v, red_i, rr);
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          gs[p] += cb[p];
          b1[p] = T(0.25) * b2[p] + rr[p];
          lam[p] += T(0.75) * b2[p];
        }
#pragma unroll
        for (int p = 0; p < PPT; ++p) cb[p] = dt * b1[p];
        rhs_vjp<T, PPT, CONV>(a, c, st[0], cb, nu, a.parts, us, cs, C, red_v, red_i, rr);
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          gs[p] += cb[p];
          lam[p] += b1[p] + rr[p];
        }
      }
    }
    if (step > 0) {
      gload(step - 1, cb);
#pragma unroll
      for (int p = 0; p < PPT; ++p) lam[p] += cb[p];
    }
  }
  if (own) {
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      int i = c.i0 + p;
      if (i < n) {
        r.out[off + i] = lam[p];
        if (r.g_s) r.g_s[off + i] = gs[p];
      }
    }
  }
}

__device__ __forceinline__ int ragged_ppt(int n, int team) {
  int need = (n + team - 1) / team;
  int ppt = 1;
  while (ppt < need) ppt <<= 1;
  return ppt;
}

template <typename T, int CONV>
__global__ void __launch_bounds__(kBlock) burgers_ragged_fwd_kernel(RArgs<T> r) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T red_v[2 * kMaxWaves];
  __shared__ int red_i[kMaxWaves];
  const int b = r.order ? r.order[blockIdx.x] : int(blockIdx.x);
  const int n = r.n[b];
  T* us = reinterpret_cast<T*>(smem_raw);
  T* Fs = us + n;
  switch (ragged_ppt(n, blockDim.x)) {
    case 1: ragged_fwd_body<T, 1, CONV>(r, b, us, Fs, red_v, red_i); break;
    case 2: ragged_fwd_body<T, 2, CONV>(r, b, us, Fs, red_v, red_i); break;
    case 4: ragged_fwd_body<T, 4, CONV>(r, b, us, Fs, red_v, red_i); break;
    case 8: ragged_fwd_body<T, 8, CONV>(r, b, us, Fs, red_v, red_i); break;
    default: ragged_fwd_body<T, 16, CONV>(r, b, us, Fs, red_v, red_i); break;
  }
}

template <typename T, int CONV>
__global__ void __launch_bounds__(kBlock) burgers_ragged_adj_kernel(RArgs<T> r) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T red_v[2 * kMaxWaves];
  __shared__ int red_i[kMaxWaves];
  const int b = r.order ? r.order[blockIdx.x] : int(blockIdx.x);
  const int n = r.n[b];
  T* us = reinterpret_cast<T*>(smem_raw);
  T* cs = us + n;
  T* C = cs + n;
  switch (ragged_ppt(n, blockDim.x)) {
    case 1: ragged_adj_body<T, 1, CONV>(r, b, us, cs, C, red_v, red_i); break;
    case 2: ragged_adj_body<T, 2, CONV>(r, b, us, cs, C, red_v, red_i); break;
    case 4: ragged_adj_body<T, 4, CONV>(r, b, us, cs, C, red_v, red_i); break;
    case 8: ragged_adj_body<T, 8, CONV>(r, b, us, cs, C, red_v, red_i); break;
    default: ragged_adj_body<T, 16, CONV>(r, b, us, cs, C, red_v, red_i); break;
  }
}

template <typename T>
int ragged_launch(const inc_ragged_desc* d, const void* u, const void* s, void* out, int steps, void* traj,
                  void* stages, const void* g_traj, void* g_s, bool adj, void* stream) {
  INC_REQUIRE(d->scheme == INC_EULER || d->scheme == INC_SSPRK3, INC_EINVAL, "unknown time scheme %d", d->scheme);
  INC_REQUIRE(d->convection == INC_WENO || d->convection == INC_FDM, INC_EINVAL, "Convection scheme not recognized");
  INC_REQUIRE(d->count >= 0 && d->total >= 0, INC_EINVAL, "negative sizes");
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(d->nmin >= 8, INC_EINVAL, "every trajectory needs >= 8 points (smallest is %d)", d->nmin);
  const int team = d->team > 0 ? d->team : 64;
  INC_REQUIRE(team >= kWave && team <= kBlock && (team & (team - 1)) == 0, INC_EUNSUPPORTED,
              "ragged team must be 64, 128 or 256 threads (got %d)", team);
  INC_REQUIRE(d->nmax <= 16 * team, INC_EUNSUPPORTED, "n=%d exceeds 16 points per thread of a %d-thread team",
              d->nmax, team);
  INC_REQUIRE(d->n && d->offset && d->nu && d->dt && d->dx, INC_EINVAL, "ragged descriptor arrays missing");
  if (d->count == 0) return INC_OK;
  INC_REQUIRE(!adj || stages, INC_EINVAL, "adjoint needs the saved stages");
  size_t lds = size_t(adj ? 8 : 2) * d->nmax * sizeof(T);
  INC_REQUIRE(lds <= 160 * 1024 - 1024, INC_EUNSUPPORTED, "n=%d needs %zu bytes of LDS", d->nmax, lds);
  RArgs<T> r{};
  r.count = d->count;
  r.scheme = d->scheme;
  r.steps = steps;
  r.step0 = d->step0;
  r.team = team;
  r.total = d->total;
  r.n = d->n;
  r.off = d->offset;
  r.order = d->order;
  r.horizon = d->horizon;
  r.nu = static_cast<const T*>(d->nu);
  r.dt = static_cast<const T*>(d->dt);
  r.dx = static_cast<const T*>(d->dx);
  r.u = static_cast<const T*>(u);
  r.s = static_cast<const T*>(s);
  r.out = static_cast<T*>(out);
  r.traj = static_cast<T*>(traj);
  r.stages = static_cast<T*>(stages);
  r.g_traj = static_cast<const T*>(g_traj);
  r.g_s = static_cast<T*>(g_s);
  void* fn;
  if (adj)
    fn = d->convection == INC_WENO ? reinterpret_cast<void*>(&burgers_ragged_adj_kernel<T, INC_WENO>)
                                   : reinterpret_cast<void*>(&burgers_ragged_adj_kernel<T, INC_FDM>);
  else
    fn = d->convection == INC_WENO ? reinterpret_cast<void*>(&burgers_ragged_fwd_kernel<T, INC_WENO>)
                                   : reinterpret_cast<void*>(&burgers_ragged_fwd_kernel<T, INC_FDM>);
  void* args[] = {&r};
  INC_HIP(hipLaunchKernel(fn, dim3(d->count), dim3(team), args, lds, static_cast<hipStream_t>(stream)));
  return INC_OK;
}

int ragged_dispatch(const inc_ragged_desc* d, const void* u, const void* s, void* out, int steps, void* traj,
                    void* stages, const void* g_traj, void* g_s, bool adj, void* stream) {
  clear_error();
  INC_REQUIRE(d != nullptr, INC_EINVAL, "null descriptor");
  if (d->dtype == INC_F32) return ragged_launch<float>(d, u, s, out, steps, traj, stages, g_traj, g_s, adj, stream);
  if (d->dtype == INC_F64) return ragged_launch<double>(d, u, s, out, steps, traj, stages, g_traj, g_s, adj, stream);
  set_error("unsupported dtype %d", d->dtype);
  return INC_EINVAL;
}

// ---- host side ---------------------------------------------------------------
int simd_count() {
  static int simds = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    return 4 * cus;
  }();
  return simds;
}

// Points per thread and team size.  An explicit team is honoured (the
// smallest PPT that covers n).  The default takes, among PPT <= ppt_max with
// a team of at most 256 threads, the largest PPT that still gives every SIMD
// 4 wavefronts, or the smallest PPT when the batch cannot: more points per
// thread amortise the stencil halo, more wavefronts hide LDS and barrier
// latency.  ppt_max is 4, and lower for the WENO adjoint, whose exact VJP
// needs ~170 (PPT 2) and 256 (PPT 4) float32 registers (measured on MI355X,
// B=8192 N=256: adjoint 15.6 ms at PPT 2 against 26.9 ms at PPT 4).  Above
// 1024 points the team stays at 256 threads with 8 or 16 points each.
int pick_ppt(int n, int team_req, int* team_out, int batch, int ppt_max) {
  static const int ppts[] = {1, 2, 4, 8, 16};
  auto team_for = [&](int ppt) {
    int need = (n + ppt - 1) / ppt, team = 1;
    while (team < need) team <<= 1;
    return team;
  };
  if (team_req > 0) {
    if (team_req > kBlock || (team_req & (team_req - 1)) != 0) return -1;
    for (int ppt : ppts) {
      if (ppt * team_req >= n) {
        *team_out = team_req;
        return ppt;
      }
    }
    return -1;
  }
  int best = -1;
  for (int ppt : ppts) {
    const int team = team_for(ppt);
    if (team > kBlock) continue;  // needs more points per thread
    if (best < 0) best = ppt;     // the smallest feasible PPT
    if (ppt > ppt_max && best != ppt) break;
    const long long waves = (static_cast<long long>(batch) * team + kWave - 1) / kWave;
    if (waves >= 4LL * simd_count()) best = ppt;
  }
  if (best < 0) return -1;
  *team_out = team_for(best);
  return best;
}

template <typename T>
int fill_args(const inc_burgers_desc* d, BArgs<T>& a) {
  INC_REQUIRE(d->n >= 8, INC_EINVAL, "resolution must be >= 8 (got %d)", d->n);
  INC_REQUIRE(d->batch >= 0, INC_EINVAL, "batch must be >= 0");
  INC_REQUIRE(d->scheme == INC_EULER || d->scheme == INC_SSPRK3, INC_EINVAL,
              "unknown time scheme %d", d->scheme);
  INC_REQUIRE(d->convection == INC_WENO || d->convection == INC_FDM, INC_EINVAL,
              "Convection scheme not recognized");
  INC_REQUIRE(d->source >= INC_SRC_NONE && d->source <= INC_SRC_LEARNING, INC_EINVAL,
              "Unknown equation type.");
  INC_REQUIRE(d->source != INC_SRC_LEARNING || (d->x && d->A && d->omega && d->phi && d->l) || d->J == 0,
              INC_EINVAL, "'learning' source needs x, A, omega, phi, l");
  INC_REQUIRE(d->dx > 0, INC_EINVAL, "dx must be positive");
  a = BArgs<T>{};
  a.batch = d->batch;
  a.n = d->n;
  a.J = d->J;
  a.scheme = d->scheme;
  a.source = d->source;
  a.parts = INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE;
  a.dt = T(d->dt);
  a.inv_dx = T(1.0 / d->dx);
  a.inv_dx2 = T(1.0 / (d->dx * d->dx));
  a.half_inv_dx = T(0.5 / d->dx);
  a.nu = T(d->nu);
  a.beta = T(d->beta);
  a.t0 = T(d->t0);
  a.two_pi_over_L = T(2.0 * 3.14159265358979323846 / d->L);
  a.nu_ptr = static_cast<const T*>(d->nu_ptr);
  a.beta_ptr = static_cast<const T*>(d->beta_ptr);
  a.t_ptr = static_cast<const T*>(d->t_ptr);
  a.x = static_cast<const T*>(d->x);
  a.A = static_cast<const T*>(d->A);
  a.omega = static_cast<const T*>(d->omega);
  a.phi = static_cast<const T*>(d->phi);
  a.l = static_cast<const T*>(d->l);
  return INC_OK;
}

enum class Kind { Fwd, Rhs, RhsAdj, Adj, AdjNu };

template <typename T, int PPT, int CONV>
void* kernel_ptr(Kind k) {
  switch (k) {
    case Kind::Fwd: return reinterpret_cast<void*>(&burgers_fwd_kernel<T, PPT, CONV>);
    case Kind::Rhs: return reinterpret_cast<void*>(&burgers_rhs_kernel<T, PPT, CONV>);
    case Kind::RhsAdj: return reinterpret_cast<void*>(&burgers_rhs_adj_kernel<T, PPT, CONV>);
    case Kind::AdjNu: return reinterpret_cast<void*>(&burgers_adj_kernel<T, PPT, CONV, true>);
    default: return reinterpret_cast<void*>(&burgers_adj_kernel<T, PPT, CONV, false>);
  }
}

template <typename T, int CONV>
void* kernel_for_ppt(Kind k, int ppt) {
  switch (ppt) {
    case 1: return kernel_ptr<T, 1, CONV>(k);
    case 2: return kernel_ptr<T, 2, CONV>(k);
    case 4: return kernel_ptr<T, 4, CONV>(k);
    case 8: return kernel_ptr<T, 8, CONV>(k);
    default: return kernel_ptr<T, 16, CONV>(k);
  }
}

template <typename T>
int launch(Kind kind, const inc_burgers_desc* d, BArgs<T>& a, void* stream) {
  int team = 0;
  if (kind == Kind::Adj && a.g_nu) kind = Kind::AdjNu;
  const bool weno_vjp = kind != Kind::Fwd && kind != Kind::Rhs && d->convection == INC_WENO;
  const int ppt_max = weno_vjp ? (std::is_same<T, float>::value ? 2 : 1) : 4;
  int ppt = pick_ppt(d->n, d->team, &team, d->batch, ppt_max);
  INC_REQUIRE(ppt > 0 && team >= 1 && team <= kBlock && (team & (team - 1)) == 0, INC_EUNSUPPORTED,
              "no launch shape for n=%d team=%d (team must be a power of two <= 256, n <= 16*team)",
              d->n, d->team);
  a.team = team;
  if (a.batch == 0) return INC_OK;
  const int tpb = kBlock / team;
  const int nb = (kind == Kind::Fwd || kind == Kind::Rhs) ? 2 : 8;
  size_t lds = size_t(tpb) * nb * d->n * sizeof(T);
  INC_REQUIRE(lds <= 160 * 1024 - 1024, INC_EUNSUPPORTED,
              "n=%d needs %zu bytes of LDS per block (team=%d); use a larger team", d->n, lds, team);
  void* fn = d->convection == INC_WENO ? kernel_for_ppt<T, INC_WENO>(kind, ppt)
                                       : kernel_for_ppt<T, INC_FDM>(kind, ppt);
  dim3 grid((a.batch + tpb - 1) / tpb), block(kBlock);
  void* args[] = {&a};
  INC_HIP(hipLaunchKernel(fn, grid, block, args, lds, static_cast<hipStream_t>(stream)));
  return INC_OK;
}

template <typename T>
int weno_launch(int direction, const void* f, const void* g, void* out, int64_t rows, int n,
                void* stream, bool adj) {
  INC_REQUIRE(direction == 0 || direction == 1, INC_EINVAL, "Direction must be 'left' or 'right'.");
  INC_REQUIRE(n >= 5, INC_EINVAL, "weno_reconstruction needs at least 5 points (got %d)", n);
  int64_t total = rows * int64_t(n);
  if (total == 0) return INC_OK;
  dim3 grid(unsigned((total + 255) / 256)), block(256);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const T* fp = static_cast<const T*>(f);
  T* op = static_cast<T*>(out);
  if (!adj) {
    if (direction == 0) hipLaunchKernelGGL((weno_fwd_kernel<T, 0>), grid, block, 0, s, fp, op, rows, n);
    else hipLaunchKernelGGL((weno_fwd_kernel<T, 1>), grid, block, 0, s, fp, op, rows, n);
  } else {
    const T* gp = static_cast<const T*>(g);
    if (direction == 0) hipLaunchKernelGGL((weno_adj_kernel<T, 0>), grid, block, 0, s, fp, gp, op, rows, n);
    else hipLaunchKernelGGL((weno_adj_kernel<T, 1>), grid, block, 0, s, fp, gp, op, rows, n);
  }
  return check_launch("weno");
}

struct Adaptive {
  double cfl;  // <= 0: fixed steps of dt
  int max_sub;
  void* nsub;
  void* dts;
};

// Fused MSE arguments (see BArgs): target trajectory, forward partial sums,
// adjoint scale.
struct Mse {
  const void* target = nullptr;
  void* parts = nullptr;
  const void* scale = nullptr;  // device scalar of the element type
};

template <typename T>
int burgers_call(Kind kind, const inc_burgers_desc* d, int parts, const void* u, const void* s,
                 void* out, int steps, void* traj, void* stages, const void* g_traj, void* g_s,
                 void* stream, const Adaptive& ad, const Mse& mse, void* g_nu) {
  BArgs<T> a;
  int rc = fill_args<T>(d, a);
  if (rc) return rc;
  a.cfl = T(ad.cfl > 0 ? ad.cfl * d->dx : 0);
  a.max_sub = ad.max_sub;
  a.nsub = static_cast<int*>(ad.nsub);
  a.dts = static_cast<T*>(ad.dts);
  a.parts = parts;
  a.steps = steps;
  a.u = static_cast<const T*>(u);
  a.s = static_cast<const T*>(s);
  a.out = static_cast<T*>(out);
  a.traj = static_cast<T*>(traj);
  a.stages = static_cast<T*>(stages);
  a.g_traj = static_cast<const T*>(g_traj);
  a.g_s = static_cast<T*>(g_s);
  a.target = static_cast<const T*>(mse.target);
  a.loss_parts = static_cast<T*>(mse.parts);
  a.loss_scale = static_cast<const T*>(mse.scale);
  a.g_nu = static_cast<T*>(g_nu);
  return launch<T>(kind, d, a, stream);
}

int dispatch(Kind kind, const inc_burgers_desc* d, int parts, const void* u, const void* s, void* out,
             int steps, void* traj, void* stages, const void* g_traj, void* g_s, void* stream,
             const Adaptive& ad = Adaptive{0.0, 0, nullptr, nullptr}, const Mse& mse = Mse{},
             void* g_nu = nullptr) {
  clear_error();
  INC_REQUIRE(d != nullptr, INC_EINVAL, "null descriptor");
  if (d->dtype == INC_F32)
    return burgers_call<float>(kind, d, parts, u, s, out, steps, traj, stages, g_traj, g_s, stream, ad, mse, g_nu);
  if (d->dtype == INC_F64)
    return burgers_call<double>(kind, d, parts, u, s, out, steps, traj, stages, g_traj, g_s, stream, ad, mse,
                                g_nu);
  set_error("unsupported dtype %d", d->dtype);
  return INC_EINVAL;
}

template <typename T, int PPT, int C>
void* inc_fwd_kernel_k(int k) {
  return k == 5 ? reinterpret_cast<void*>(&burgers_inc_fwd_kernel<T, PPT, C, 5, 0>)
                : reinterpret_cast<void*>(&burgers_inc_fwd_kernel<T, PPT, C, 7, 0>);
}

// matrix-core hidden layers: PW 16-position tiles per wavefront (1, 2, 4 or 8)
template <int PPT, int C, int K>
void* inc_fwd_mfma_kernel_pw(int pw) {
  switch (pw) {
    case 1: return reinterpret_cast<void*>(&burgers_inc_fwd_kernel<float, PPT, C, K, 1>);
    case 2: return reinterpret_cast<void*>(&burgers_inc_fwd_kernel<float, PPT, C, K, 2>);
    case 4: return reinterpret_cast<void*>(&burgers_inc_fwd_kernel<float, PPT, C, K, 4>);
    default: return reinterpret_cast<void*>(&burgers_inc_fwd_kernel<float, PPT, C, K, 8>);
  }
}

// The instances that keep every value in registers (the others spill, and
// their shapes take the vector-ALU network): one point per thread, any
// width and kernel size except 48 x 7 at 8 tiles; two points per thread,
// width 32 and kernel size 5.  nullptr otherwise.
void* inc_fwd_mfma_kernel(int ppt, int c, int k, int pw) {
  if (ppt == 2) return (c == 32 && k == 5) ? inc_fwd_mfma_kernel_pw<2, 32, 5>(pw) : nullptr;
  if (c == 32) return k == 5 ? inc_fwd_mfma_kernel_pw<1, 32, 5>(pw) : inc_fwd_mfma_kernel_pw<1, 32, 7>(pw);
  if (k == 5) return inc_fwd_mfma_kernel_pw<1, 48, 5>(pw);
  return pw == 8 ? nullptr : inc_fwd_mfma_kernel_pw<1, 48, 7>(pw);
}

template <typename T>
int inc_fwd_launch(const inc_burgers_desc* d, const inc_cnn_desc* cd, const void* u, void* out, int steps,
                   void* traj, void* corr, void* stream, void* stages = nullptr, const void* target = nullptr,
                   void* loss_parts = nullptr) {
  BArgs<T> a;
  int rc = fill_args<T>(d, a);
  if (rc) return rc;
  INC_REQUIRE(d->convection == INC_WENO, INC_EUNSUPPORTED, "the in-kernel correction network runs WENO convection");
  INC_REQUIRE(cd->layers >= 2 && cd->layers <= kCnnMaxLayers, INC_EINVAL, "correction network needs 2..%d layers",
              kCnnMaxLayers);
  INC_REQUIRE((cd->channels == 32 || cd->channels == 48) && (cd->kernel == 5 || cd->kernel == 7), INC_EUNSUPPORTED,
              "in-kernel correction network: width 32 or 48, kernel size 5 or 7 (got %d, %d)", cd->channels,
              cd->kernel);
  INC_REQUIRE(cd->kernel <= d->n, INC_EINVAL, "kernel size exceeds the resolution");
  // float: up to 2 points per thread; double: 1 (two would spill registers).
  // The automatic choice keeps one point per thread while a 256-thread team
  // covers n: more points per thread shrink the team, and the network's
  // tiles then crowd onto fewer wavefronts (B=4096, N=128: 23.4 ms for 100
  // steps at one point per thread against 40.0 ms at two).
  constexpr int kPptMax = std::is_same<T, float>::value ? 2 : 1;
  int team = 0;
  const int ppt = pick_ppt(d->n, d->team, &team, d->batch, d->n <= kBlock ? 1 : kPptMax);
  INC_REQUIRE(ppt >= 1 && ppt <= kPptMax, INC_EUNSUPPORTED,
              "in-kernel correction network: n=%d needs at most %d point(s) per thread of a 256-thread team", d->n,
              kPptMax);
  if (team < kWave) team = kWave;
  size_t lds = size_t(2 + cd->channels) * d->n * sizeof(T);
  // float: hidden layers on the matrix cores when their weights and the
  // position-major activations fit in LDS
  int pw = 0;
  if (std::is_same<T, float>::value && cd->layers > 2) {
    const int C = cd->channels, K = cd->kernel;
    const int tiles = (d->n + 15) / 16;
    auto pw_for = [&](int tm) {  // 16-position tiles per wavefront, rounded up to an instance's
      const int waves = tm / kWave, need = (tiles + waves - 1) / waves;
      return need <= 1 ? 1 : (need <= 2 ? 2 : (need <= 4 ? 4 : 8));
    };
    const size_t lm = (size_t(2) * d->n + size_t(cd->layers - 2) * C * (C * K + 4) + size_t(d->n) * (C + 4)) *
                      sizeof(float);
    // An automatic team widens to 256 threads when the batch would leave
    // SIMDs without a wavefront: the network's tiles then spread over four
    // wavefronts (the solver's extra threads idle through its stages).
    // Measured on MI355X, N=64: B=256 1.59x faster; at one wavefront per
    // SIMD or more the narrower team is as fast or faster (B=4096, N=128:
    // 2x), so it stays.
    const long long waves_now = static_cast<long long>(d->batch) * (team / kWave);
    if (d->team == 0 && team < kBlock && waves_now < simd_count() && lm <= 150 * 1024 &&
        inc_fwd_mfma_kernel(ppt, C, K, pw_for(kBlock)))
      team = kBlock;
    const int need = pw_for(team);
    if (lm <= 150 * 1024 && inc_fwd_mfma_kernel(ppt, C, K, need) != nullptr) {
      pw = need;
      lds = lm;
    }
  }
  INC_REQUIRE(lds <= 150 * 1024, INC_EUNSUPPORTED, "in-kernel correction network: n=%d, width %d need %zu bytes of LDS",
              d->n, cd->channels, lds);
  CnnArgs<T> cn{};
  cn.layers = cd->layers;
  cn.scale = T(cd->scale);
  for (int l = 0; l < cd->layers; ++l) {
    INC_REQUIRE(cd->weight[l] != nullptr, INC_EINVAL, "correction network layer %d has no weight", l);
    cn.w[l] = static_cast<const T*>(cd->weight[l]);
    cn.b[l] = static_cast<const T*>(cd->bias[l]);
  }
  a.team = team;
  a.steps = steps;
  a.parts = INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE;
  a.u = static_cast<const T*>(u);
  a.s = nullptr;
  a.out = static_cast<T*>(out);
  a.traj = static_cast<T*>(traj);
  a.stages = static_cast<T*>(stages);
  a.target = static_cast<const T*>(target);
  a.loss_parts = static_cast<T*>(loss_parts);
  if (a.batch == 0) return INC_OK;
  void* fn;
  if constexpr (kPptMax == 2) {
    if (pw > 0)
      fn = inc_fwd_mfma_kernel(ppt, cd->channels, cd->kernel, pw);
    else if (ppt == 2)
      fn = cd->channels == 32 ? inc_fwd_kernel_k<T, 2, 32>(cd->kernel) : inc_fwd_kernel_k<T, 2, 48>(cd->kernel);
    else
      fn = cd->channels == 32 ? inc_fwd_kernel_k<T, 1, 32>(cd->kernel) : inc_fwd_kernel_k<T, 1, 48>(cd->kernel);
  } else {
    fn = cd->channels == 32 ? inc_fwd_kernel_k<T, 1, 32>(cd->kernel) : inc_fwd_kernel_k<T, 1, 48>(cd->kernel);
  }
  T* cp = static_cast<T*>(corr);
  void* args[] = {&a, &cn, &cp};
  INC_HIP(hipLaunchKernel(fn, dim3(a.batch), dim3(team), args, lds, static_cast<hipStream_t>(stream)));
  return INC_OK;
}

// Training adjoint instances: width 32, K 5 or 7, 0..2 hidden layers, 1 or 2
// position tiles per wavefront (n <= 64, n <= 128).
template <int K, int NH>
void* inc_adj_kernel_pw(int pw) {
  return pw == 1 ? reinterpret_cast<void*>(&burgers_inc_adj_kernel<32, K, NH, 1>)
                 : reinterpret_cast<void*>(&burgers_inc_adj_kernel<32, K, NH, 2>);
}

template <int K>
void* inc_adj_kernel_k(int nh, int pw) {
  return nh == 0 ? inc_adj_kernel_pw<K, 0>(pw) : (nh == 1 ? inc_adj_kernel_pw<K, 1>(pw) : inc_adj_kernel_pw<K, 2>(pw));
}

// LDS of burgers_inc_adj_kernel: staged hidden weights, nh + 1 pre-activation
// arrays, activations ([n + K - 1][C + 4], with wrapped rows) and their
// cotangent ([n][C + 4]), the input state and the output's cotangent; the
// final reduction reuses it from the pre-activations on.
size_t inc_adj_lds(int n, int C, int K, int nh) {
  const size_t cs = C + 4, rs = size_t(C) * K + 4, ws = size_t(nh) * C * rs;
  const size_t v = (4 * K + 4) + (4 * K + 1) + 4 * nh;
  const size_t body = ws + size_t(nh + 3) * n * cs + size_t(K / 2) * 2 * cs + 2 * size_t(n);
  const size_t red = ws + 4 * size_t(C / 4) * v;
  return (body > red ? body : red) * sizeof(float);
}

int inc_train_adj_launch(const inc_burgers_desc* d, const inc_cnn_desc* cd, const void* stages, const void* g_traj,
                         const void* g_corr, void* g_u, void* g_w, int steps, void* stream,
                         const void* traj = nullptr, const void* target = nullptr, const void* scale = nullptr) {
  INC_REQUIRE(d->dtype == INC_F32, INC_EUNSUPPORTED, "the training adjoint with the network inside runs float32");
  BArgs<float> a;
  int rc = fill_args<float>(d, a);
  if (rc) return rc;
  INC_REQUIRE(d->convection == INC_WENO, INC_EUNSUPPORTED, "the in-kernel correction network runs WENO convection");
  INC_REQUIRE(cd->channels == 32 && (cd->kernel == 5 || cd->kernel == 7) && cd->layers >= 2 && cd->layers <= 4,
              INC_EUNSUPPORTED,
              "training adjoint with the network inside: width 32, kernel size 5 or 7, 2..4 layers (got %d, %d, %d)",
              cd->channels, cd->kernel, cd->layers);
  INC_REQUIRE(d->n <= 128, INC_EUNSUPPORTED, "training adjoint with the network inside: n <= 128 (got %d)", d->n);
  INC_REQUIRE(stages != nullptr && g_u != nullptr && g_w != nullptr, INC_EINVAL,
              "the training adjoint needs stages, g_u and g_w");
  const int nh = cd->layers - 2, pw = d->n <= 64 ? 1 : 2;
  const size_t lds = inc_adj_lds(d->n, 32, cd->kernel, nh);
  INC_REQUIRE(lds <= 160 * 1024 - 256, INC_EUNSUPPORTED,
              "training adjoint with the network inside: n=%d, %d layers, K=%d need %zu bytes of LDS", d->n,
              cd->layers, cd->kernel, lds);
  CnnArgs<float> cn{};
  cn.layers = cd->layers;
  cn.scale = float(cd->scale);
  for (int l = 0; l < cd->layers; ++l) {
    INC_REQUIRE(cd->weight[l] != nullptr, INC_EINVAL, "correction network layer %d has no weight", l);
    cn.w[l] = static_cast<const float*>(cd->weight[l]);
    cn.b[l] = static_cast<const float*>(cd->bias[l]);
  }
  a.team = kTrainBlock;
  a.steps = steps;
  a.stages = static_cast<float*>(const_cast<void*>(stages));
  a.g_traj = static_cast<const float*>(g_traj);
  a.out = static_cast<float*>(g_u);
  a.traj = static_cast<float*>(const_cast<void*>(traj));
  a.target = static_cast<const float*>(target);
  a.loss_scale = static_cast<const float*>(scale);
  if (a.batch == 0) return INC_OK;
  void* fn = cd->kernel == 5 ? inc_adj_kernel_k<5>(nh, pw) : inc_adj_kernel_k<7>(nh, pw);
  const float* gc = static_cast<const float*>(g_corr);
  float* gw = static_cast<float*>(g_w);
  void* args[] = {&a, &cn, &gc, &gw};
  INC_HIP(hipLaunchKernel(fn, dim3(a.batch), dim3(kTrainBlock), args, lds, static_cast<hipStream_t>(stream)));
  return INC_OK;
}

}  // namespace
}  // namespace inc

using inc::Kind;

extern "C" {

int inc_weno_fwd(int dtype, int direction, const void* f, void* out, int64_t rows, int n,
                 void* stream) {
  inc::clear_error();
  if (dtype == INC_F32) return inc::weno_launch<float>(direction, f, nullptr, out, rows, n, stream, false);
  if (dtype == INC_F64) return inc::weno_launch<double>(direction, f, nullptr, out, rows, n, stream, false);
  inc::set_error("unsupported dtype %d", dtype);
  return INC_EINVAL;
}

int inc_weno_adj(int dtype, int direction, const void* f, const void* g_out, void* g_f, int64_t rows,
                 int n, void* stream) {
  inc::clear_error();
  if (dtype == INC_F32) return inc::weno_launch<float>(direction, f, g_out, g_f, rows, n, stream, true);
  if (dtype == INC_F64) return inc::weno_launch<double>(direction, f, g_out, g_f, rows, n, stream, true);
  inc::set_error("unsupported dtype %d", dtype);
  return INC_EINVAL;
}

int inc_burgers_rhs(const inc_burgers_desc* d, int parts, const void* u, const void* s, void* out,
                    void* stream) {
  return inc::dispatch(Kind::Rhs, d, parts, u, s, out, 1, nullptr, nullptr, nullptr, nullptr, stream);
}

int inc_burgers_rhs_adj(const inc_burgers_desc* d, int parts, const void* u, const void* g, void* g_u,
                        void* stream) {
  return inc::dispatch(Kind::RhsAdj, d, parts, u, nullptr, g_u, 1, nullptr, nullptr, g, nullptr, stream);
}

int inc_burgers_rhs_adj_nu(const inc_burgers_desc* d, int parts, const void* u, const void* g, void* g_u,
                           void* g_nu, void* stream) {
  return inc::dispatch(Kind::RhsAdj, d, parts, u, nullptr, g_u, 1, nullptr, nullptr, g, nullptr, stream,
                       inc::Adaptive{0.0, 0, nullptr, nullptr}, inc::Mse{}, g_nu);
}

int inc_burgers_fwd(const inc_burgers_desc* d, const void* u, const void* s, void* u_out, int steps,
                    void* traj, void* stages, void* stream) {
  if (steps < 1) {
    inc::set_error("steps must be >= 1 (got %d)", steps);
    return INC_EINVAL;
  }
  return inc::dispatch(Kind::Fwd, d, INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE, u, s,
                       u_out, steps, traj, stages, nullptr, nullptr, stream);
}

int inc_burgers_ragged_fwd(const inc_ragged_desc* d, const void* u, const void* s, void* u_out, int steps,
                           void* traj, void* stages, void* stream) {
  return inc::ragged_dispatch(d, u, s, u_out, steps, traj, stages, nullptr, nullptr, false, stream);
}

int inc_burgers_ragged_adj(const inc_ragged_desc* d, const void* stages, const void* g_traj, void* g_u, void* g_s,
                           int steps, void* stream) {
  return inc::ragged_dispatch(d, nullptr, nullptr, g_u, steps, nullptr, const_cast<void*>(stages), g_traj, g_s,
                              true, stream);
}

int inc_burgers_adj(const inc_burgers_desc* d, const void* stages, const void* g_traj, void* g_u,
                    void* g_s, int steps, void* stream) {
  return inc_burgers_adj_nu(d, stages, g_traj, g_u, g_s, nullptr, steps, stream);
}

int inc_burgers_adj_nu(const inc_burgers_desc* d, const void* stages, const void* g_traj, void* g_u,
                       void* g_s, void* g_nu, int steps, void* stream) {
  if (steps < 1) {
    inc::set_error("steps must be >= 1 (got %d)", steps);
    return INC_EINVAL;
  }
  return inc::dispatch(Kind::Adj, d, INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE,
                       nullptr, nullptr, g_u, steps, nullptr, const_cast<void*>(stages), g_traj, g_s,
                       stream, inc::Adaptive{0.0, 0, nullptr, nullptr}, inc::Mse{}, g_nu);
}

int inc_burgers_fwd_mse(const inc_burgers_desc* d, const void* u, const void* s, void* u_out, int steps,
                        void* traj, void* stages, const void* target, void* loss_parts, void* stream) {
  inc::clear_error();
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(target != nullptr && loss_parts != nullptr, INC_EINVAL, "the MSE forward needs target and loss_parts");
  inc::Mse m;
  m.target = target;
  m.parts = loss_parts;
  return inc::dispatch(Kind::Fwd, d, INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE, u, s, u_out,
                       steps, traj, stages, nullptr, nullptr, stream, inc::Adaptive{0.0, 0, nullptr, nullptr}, m);
}

int inc_burgers_adj_mse(const inc_burgers_desc* d, const void* stages, const void* traj, const void* target,
                        const void* scale, const void* g_traj, void* g_u, void* g_s, int steps, void* stream) {
  return inc_burgers_adj_mse_nu(d, stages, traj, target, scale, g_traj, g_u, g_s, nullptr, steps, stream);
}

int inc_burgers_adj_mse_nu(const inc_burgers_desc* d, const void* stages, const void* traj, const void* target,
                           const void* scale, const void* g_traj, void* g_u, void* g_s, void* g_nu, int steps,
                           void* stream) {
  inc::clear_error();
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(stages != nullptr && traj != nullptr && target != nullptr && scale != nullptr, INC_EINVAL,
              "the MSE adjoint needs stages, traj, target and scale");
  inc::Mse m;
  m.target = target;
  m.scale = scale;
  return inc::dispatch(Kind::Adj, d, INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE, nullptr,
                       nullptr, g_u, steps, const_cast<void*>(traj), const_cast<void*>(stages), g_traj, g_s,
                       stream, inc::Adaptive{0.0, 0, nullptr, nullptr}, m, g_nu);
}

int inc_burgers_inc_fwd(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* u, void* u_out, int steps,
                        void* traj, void* corr, void* stream) {
  inc::clear_error();
  INC_REQUIRE(d != nullptr && cnn != nullptr, INC_EINVAL, "null descriptor");
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  if (d->dtype == INC_F32) return inc::inc_fwd_launch<float>(d, cnn, u, u_out, steps, traj, corr, stream);
  if (d->dtype == INC_F64) return inc::inc_fwd_launch<double>(d, cnn, u, u_out, steps, traj, corr, stream);
  inc::set_error("unsupported dtype %d", d->dtype);
  return INC_EINVAL;
}

int inc_burgers_inc_train_fwd(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* u, void* u_out,
                              int steps, void* traj, void* corr, void* stages, void* stream) {
  inc::clear_error();
  INC_REQUIRE(d != nullptr && cnn != nullptr, INC_EINVAL, "null descriptor");
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(stages != nullptr, INC_EINVAL, "the training forward needs stages");
  if (d->dtype == INC_F32) return inc::inc_fwd_launch<float>(d, cnn, u, u_out, steps, traj, corr, stream, stages);
  if (d->dtype == INC_F64) return inc::inc_fwd_launch<double>(d, cnn, u, u_out, steps, traj, corr, stream, stages);
  inc::set_error("unsupported dtype %d", d->dtype);
  return INC_EINVAL;
}

int inc_burgers_inc_train_adj(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* stages,
                              const void* g_traj, const void* g_corr, void* g_u, void* g_w, int steps,
                              void* stream) {
  inc::clear_error();
  INC_REQUIRE(d != nullptr && cnn != nullptr, INC_EINVAL, "null descriptor");
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  return inc::inc_train_adj_launch(d, cnn, stages, g_traj, g_corr, g_u, g_w, steps, stream);
}

int inc_burgers_inc_train_fwd_mse(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* u, void* u_out,
                                  int steps, void* traj, void* corr, void* stages, const void* target,
                                  void* loss_parts, void* stream) {
  inc::clear_error();
  INC_REQUIRE(d != nullptr && cnn != nullptr, INC_EINVAL, "null descriptor");
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(stages != nullptr && traj != nullptr, INC_EINVAL, "the training forward needs stages and traj");
  INC_REQUIRE(target != nullptr && loss_parts != nullptr, INC_EINVAL, "the MSE forward needs target and loss_parts");
  if (d->dtype == INC_F32)
    return inc::inc_fwd_launch<float>(d, cnn, u, u_out, steps, traj, corr, stream, stages, target, loss_parts);
  if (d->dtype == INC_F64)
    return inc::inc_fwd_launch<double>(d, cnn, u, u_out, steps, traj, corr, stream, stages, target, loss_parts);
  inc::set_error("unsupported dtype %d", d->dtype);
  return INC_EINVAL;
}

int inc_burgers_inc_train_adj_mse(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* stages,
                                  const void* traj, const void* target, const void* scale, const void* g_traj,
                                  const void* g_corr, void* g_u, void* g_w, int steps, void* stream) {
  inc::clear_error();
  INC_REQUIRE(d != nullptr && cnn != nullptr, INC_EINVAL, "null descriptor");
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(traj != nullptr && target != nullptr && scale != nullptr, INC_EINVAL,
              "the MSE adjoint needs traj, target and scale");
  return inc::inc_train_adj_launch(d, cnn, stages, g_traj, g_corr, g_u, g_w, steps, stream, traj, target, scale);
}

int inc_burgers_adaptive_fwd(const inc_burgers_desc* d, double cfl, int max_sub, const void* u, const void* s,
                             void* u_out, int steps, void* traj, void* nsub, void* dts, void* stages,
                             void* stream) {
  inc::clear_error();
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(cfl > 0, INC_EINVAL, "adaptive CFL number must be positive (got %g)", cfl);
  INC_REQUIRE(max_sub >= 1 && max_sub <= 1000, INC_EINVAL, "max_sub must be in [1, 1000] (got %d)", max_sub);
  INC_REQUIRE(nsub != nullptr && dts != nullptr, INC_EINVAL, "adaptive stepping needs nsub and dts");
  return inc::dispatch(Kind::Fwd, d, INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE, u, s, u_out,
                       steps, traj, stages, nullptr, nullptr, stream, inc::Adaptive{cfl, max_sub, nsub, dts});
}

int inc_burgers_adaptive_adj(const inc_burgers_desc* d, int max_sub, const void* nsub, const void* dts,
                             const void* stages, const void* g_traj, void* g_u, void* g_s, int steps,
                             void* stream) {
  return inc_burgers_adaptive_adj_nu(d, max_sub, nsub, dts, stages, g_traj, g_u, g_s, nullptr, steps, stream);
}

int inc_burgers_adaptive_adj_nu(const inc_burgers_desc* d, int max_sub, const void* nsub, const void* dts,
                                const void* stages, const void* g_traj, void* g_u, void* g_s, void* g_nu,
                                int steps, void* stream) {
  inc::clear_error();
  INC_REQUIRE(steps >= 1, INC_EINVAL, "steps must be >= 1 (got %d)", steps);
  INC_REQUIRE(max_sub >= 1 && max_sub <= 1000, INC_EINVAL, "max_sub must be in [1, 1000] (got %d)", max_sub);
  INC_REQUIRE(nsub != nullptr && dts != nullptr && stages != nullptr && g_traj != nullptr, INC_EINVAL,
              "the adaptive adjoint needs nsub, dts, stages and g_traj");
  // cfl only selects the adaptive sweep here: the sub-step sizes are read from dts
  return inc::dispatch(Kind::Adj, d, INC_PART_CONVECTION | INC_PART_DIFFUSION | INC_PART_SOURCE, nullptr,
                       nullptr, g_u, steps, nullptr, const_cast<void*>(stages), g_traj, g_s, stream,
                       inc::Adaptive{1.0, max_sub, const_cast<void*>(nsub), const_cast<void*>(dts)}, inc::Mse{},
                       g_nu);
}

}  // extern "C"
