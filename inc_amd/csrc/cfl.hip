// Per-trajectory stability numbers of the Burgers step for gfx950: the
// advective Courant number max_i |u_i| dt / dx and the viscous number
// nu dt / dx^2 of every trajectory, from the state on the GPU with no host
// synchronisation.  The reference reads one batch-wide max |u| back to the
// host to size its adaptive sub-steps (solvers/solver_1d.py:158-161); these
// numbers let a caller flag the trajectories of a fixed-step rollout that
// left the stable region (a time step too large for their speed or
// viscosity) before or after running it.
//
// One block of 256 threads per trajectory; a NaN anywhere in a trajectory
// makes its Courant number NaN (the maximum propagates NaN), so a check of
// the form !(c <= limit) flags it.

#include "common.hpp"

namespace inc {
namespace {

constexpr int kThreads = 256;

// maximum of |a| and |b| that keeps a NaN
template <typename T>
__device__ __forceinline__ T nanmax(T a, T b) {
  return (b > a || b != b) ? b : a;
}

template <typename T>
__device__ __forceinline__ T block_nanmax(T v) {
  __shared__ T red[kThreads / kWave];
  for (int off = 1; off < kWave; off <<= 1) v = nanmax(v, shfl_xor(v, off));
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
  __syncthreads();
  v = red[0];
  for (int i = 1; i < kThreads / kWave; ++i) v = nanmax(v, red[i]);
  return v;
}

// uniform batch: row b of u is [n] at b * n
template <typename T>
__global__ void __launch_bounds__(kThreads) burgers_cfl_kernel(int n, T dt, T dx, T nu, const T* __restrict__ nu_ptr,
                                                               const T* __restrict__ u, T* __restrict__ out) {
  const int b = blockIdx.x;
  const T* row = u + size_t(b) * n;
  T m = T(0);
  for (int i = threadIdx.x; i < n; i += kThreads) m = nanmax(m, fabs(row[i]));
  m = block_nanmax(m);
  if (threadIdx.x == 0) {
    const T v = nu_ptr ? nu_ptr[b] : nu;
    out[2 * b] = m * dt / dx;
    out[2 * b + 1] = v * dt / (dx * dx);
  }
}

// ragged batch: trajectory b is [n[b]] at offset[b], with its own dt, dx, nu
template <typename T>
__global__ void __launch_bounds__(kThreads) ragged_cfl_kernel(const int32_t* __restrict__ nv,
                                                              const int64_t* __restrict__ off,
                                                              const T* __restrict__ dt, const T* __restrict__ dx,
                                                              const T* __restrict__ nu, const T* __restrict__ u,
                                                              T* __restrict__ out) {
  const int b = blockIdx.x;
  const int n = nv[b];
  const T* row = u + off[b];
  T m = T(0);
  for (int i = threadIdx.x; i < n; i += kThreads) m = nanmax(m, fabs(row[i]));
  m = block_nanmax(m);
  if (threadIdx.x == 0) {
    out[2 * b] = m * dt[b] / dx[b];
    out[2 * b + 1] = nu[b] * dt[b] / (dx[b] * dx[b]);
  }
}

}  // namespace
}  // namespace inc

extern "C" {

int inc_burgers_cfl(const inc_burgers_desc* d, const void* u, void* cfl, void* stream) {
  inc::clear_error();
  INC_REQUIRE(d != nullptr, INC_EINVAL, "null descriptor");
  INC_REQUIRE(d->dtype == INC_F32 || d->dtype == INC_F64, INC_EINVAL, "unsupported dtype %d", d->dtype);
  INC_REQUIRE(d->batch >= 0 && d->n >= 1 && d->dx > 0 && d->dt >= 0, INC_EINVAL,
              "Burgers stability numbers: bad descriptor (batch %d, n %d, dx %g, dt %g)", d->batch, d->n, d->dx,
              d->dt);
  if (d->batch == 0) return INC_OK;
  INC_REQUIRE(u && cfl, INC_EINVAL, "Burgers stability numbers: missing buffer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d->dtype == INC_F32)
    hipLaunchKernelGGL(inc::burgers_cfl_kernel<float>, dim3(d->batch), dim3(inc::kThreads), 0, st, d->n,
                       float(d->dt), float(d->dx), float(d->nu), static_cast<const float*>(d->nu_ptr),
                       static_cast<const float*>(u), static_cast<float*>(cfl));
  else
    hipLaunchKernelGGL(inc::burgers_cfl_kernel<double>, dim3(d->batch), dim3(inc::kThreads), 0, st, d->n, d->dt,
                       d->dx, d->nu, static_cast<const double*>(d->nu_ptr), static_cast<const double*>(u),
                       static_cast<double*>(cfl));
  return inc::check_launch("Burgers stability numbers");
}

int inc_burgers_ragged_cfl(const inc_ragged_desc* d, const void* u, void* cfl, void* stream) {
  inc::clear_error();
  INC_REQUIRE(d != nullptr, INC_EINVAL, "null descriptor");
  INC_REQUIRE(d->dtype == INC_F32 || d->dtype == INC_F64, INC_EINVAL, "unsupported dtype %d", d->dtype);
  INC_REQUIRE(d->count >= 0, INC_EINVAL, "ragged stability numbers: bad descriptor");
  if (d->count == 0) return INC_OK;
  INC_REQUIRE(u && cfl && d->n && d->offset && d->dt && d->dx && d->nu, INC_EINVAL,
              "ragged stability numbers: missing buffer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d->dtype == INC_F32)
    hipLaunchKernelGGL(inc::ragged_cfl_kernel<float>, dim3(d->count), dim3(inc::kThreads), 0, st, d->n, d->offset,
                       static_cast<const float*>(d->dt), static_cast<const float*>(d->dx),
                       static_cast<const float*>(d->nu), static_cast<const float*>(u), static_cast<float*>(cfl));
  else
    hipLaunchKernelGGL(inc::ragged_cfl_kernel<double>, dim3(d->count), dim3(inc::kThreads), 0, st, d->n, d->offset,
                       static_cast<const double*>(d->dt), static_cast<const double*>(d->dx),
                       static_cast<const double*>(d->nu), static_cast<const double*>(u), static_cast<double*>(cfl));
  return inc::check_launch("ragged stability numbers");
}

}  // extern "C"
