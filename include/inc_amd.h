/*
 * inc_amd.h -- C ABI of the MI355X INC 1D solver-step library (libinc_amd.so).
 *
 * The reference exposes this path as a Python class API with no FFI
 * (solvers/solver_1d.py: weno_reconstruction :4-42, BurgersSolverTorch
 * :44-195, KSSolverTorch :198-288).  Each entry point below replaces the
 * eager-PyTorch op sequence of one of those methods; the Python package
 * inc_amd binds them with ctypes behind classes of the same names.
 *
 * Conventions
 *  - All array arguments are device pointers (HBM), row-major, contiguous
 *    along x: a batch of trajectories is [batch, n].  `dtype` selects
 *    float32 (INC_F32) or float64 (INC_F64) for every floating array of a
 *    call, except the KS domain sizes, which are always float64.
 *  - Every call is asynchronous on `stream` (a hipStream_t; NULL = the
 *    null stream), allocates nothing and enqueues kernels on it with no host
 *    synchronisation.  Besides kernel launches it makes only device queries
 *    whose results are cached (the CU count, and the resident blocks of a
 *    matrix-core convolution at its LDS size), so a call made once outside a
 *    hipGraph capture can be captured afterwards: the ragged and KS INC
 *    rollouts (inc_ragged_inc_fwd/_adj) and the KS rollout with its adjoint
 *    are validated under capture and replay.
 *  - Return value: INC_OK (0) or an error code; inc_last_error() returns a
 *    thread-local message for the last failing call of the calling thread.
 *  - Calls that overlap on one stream are ordered; calls on different
 *    streams that write the same buffers race (the caller orders them).
 */
#ifndef INC_AMD_H
#define INC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { INC_OK = 0, INC_EINVAL = 1, INC_EHIP = 2, INC_EUNSUPPORTED = 3 };
enum { INC_F32 = 0, INC_F64 = 1 };

/* Burgers: time schemes, convection schemes, source kinds, RHS parts. */
enum { INC_EULER = 0, INC_SSPRK3 = 1 };
enum { INC_WENO = 0, INC_FDM = 1 };
enum { INC_SRC_NONE = 0, INC_SRC_LINEAR = 1, INC_SRC_LEARNING = 2 };
enum { INC_PART_CONVECTION = 1, INC_PART_DIFFUSION = 2, INC_PART_SOURCE = 4 };

/* KS time schemes. */
enum { INC_ETD1 = 0, INC_ETRK2 = 1, INC_ETDRK4 = 2 };

/* Library version (major*10000 + minor*100 + patch). */
int inc_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* inc_last_error(void);

/*
 * WENO5 reconstruction with the reference's stencils.
 * Replaces weno_reconstruction(f, direction), solver_1d.py:4-42.
 * direction: 0 = 'left' (value at x_{i+1/2}), 1 = 'right' (the reference's
 * shifted stencil, value at x_{i+3/2}).  f, out: [rows, n], periodic in n.
 */
int inc_weno_fwd(int dtype, int direction, const void* f, void* out,
                 int64_t rows, int n, void* stream);
/* Vector-Jacobian product: g_f = (dF/df)^T g_out. */
int inc_weno_adj(int dtype, int direction, const void* f, const void* g_out,
                 void* g_f, int64_t rows, int n, void* stream);

/*
 * Description of a Burgers problem (constructor parameters of
 * BurgersSolverTorch, solver_1d.py:45-66, plus the forcing parameters read
 * by _compute_delta, solver_1d.py:74-85).
 */
typedef struct inc_burgers_desc {
  int32_t dtype;      /* INC_F32 / INC_F64 */
  int32_t scheme;     /* INC_EULER (reference) / INC_SSPRK3 */
  int32_t convection; /* INC_WENO / INC_FDM */
  int32_t source;     /* INC_SRC_* (equation_type, solver_1d.py:104-114) */
  int32_t batch;      /* trajectories */
  int32_t n;          /* grid points (resolution) */
  int32_t J;          /* forcing modes ('learning') */
  int32_t team;       /* threads per trajectory (0 = library default) */
  double dt, dx, L;
  double nu;          /* viscosity, used when nu_ptr == NULL */
  double beta;        /* 'linear' source coefficient, when beta_ptr == NULL */
  double t0;          /* time of the first step, when t_ptr == NULL */
  const void* nu_ptr;   /* [batch] per-trajectory viscosity or NULL */
  const void* beta_ptr; /* [batch] or NULL */
  const void* t_ptr;    /* [batch] start time per trajectory or NULL */
  const void* x;        /* [n] grid coordinates ('learning' only) */
  const void* A;        /* [batch, J] forcing amplitudes */
  const void* omega;    /* [batch, J] */
  const void* phi;      /* [batch, J] */
  const void* l;        /* [batch, J] */
} inc_burgers_desc;

/*
 * Spatial right-hand side parts (bitmask of INC_PART_*):
 * convection = calc_nonlinear_weno / calc_nonlinear_FDM (solver_1d.py:125-142),
 * diffusion = calc_linear (:144-146), source = _compute_source (+ s if non-NULL).
 * u, s, out: [batch, n]; the source is evaluated at time t0 / t_ptr.
 */
int inc_burgers_rhs(const inc_burgers_desc* d, int parts, const void* u,
                    const void* s, void* out, void* stream);
/* g_u = (d rhs_parts / du)^T g. */
int inc_burgers_rhs_adj(const inc_burgers_desc* d, int parts, const void* u,
                        const void* g, void* g_u, void* stream);
/*
 * The *_nu adjoints below also return the cotangent of the viscosity, the
 * gradient autograd gives for a nu tensor that requires grad (calc_linear,
 * solver_1d.py:144-146: nu * (u @ D2_T)):
 *   g_nu  [batch]  sum over the right-hand-side evaluations of g . (D2 u),
 *                  per trajectory (a scalar nu's cotangent is its sum)
 * Without INC_PART_DIFFUSION in parts it is zero.
 */
int inc_burgers_rhs_adj_nu(const inc_burgers_desc* d, int parts, const void* u,
                           const void* g, void* g_u, void* g_nu, void* stream);

/*
 * Advance `steps` steps of size dt from u (BurgersSolverTorch.step,
 * solver_1d.py:181-195; Euler as in the reference or SSP-RK3), with the
 * correction s ([batch, n] or NULL) held fixed over the steps and the
 * forcing clock starting at t0 / t_ptr.
 *   u_out  [batch, n]               final state
 *   traj   [batch, steps, n]        every step's output, or NULL
 *   stages [batch, steps, ns, n]    stage states for the adjoint, or NULL
 *          (ns = 1 for Euler, 3 for SSP-RK3)
 */
int inc_burgers_fwd(const inc_burgers_desc* d, const void* u, const void* s,
                    void* u_out, int steps, void* traj, void* stages,
                    void* stream);
/*
 * Adjoint of inc_burgers_fwd (the exact transpose of the reference's autograd
 * through the same step): from the saved stages and the cotangents g_traj
 * [batch, steps, n] of every step's output, returns g_u [batch, n] (cotangent
 * of the input state) and g_s [batch, n] (of the fixed correction s; NULL to
 * skip).
 */
int inc_burgers_adj(const inc_burgers_desc* d, const void* stages,
                    const void* g_traj, void* g_u, void* g_s, int steps,
                    void* stream);
/* inc_burgers_adj plus g_nu [batch] (see inc_burgers_rhs_adj_nu). */
int inc_burgers_adj_nu(const inc_burgers_desc* d, const void* stages,
                       const void* g_traj, void* g_u, void* g_s, void* g_nu,
                       int steps, void* stream);

/*
 * Rollout with its MSE to a target trajectory fused in (simulate_rollout,
 * trainer/trainer_1d.py:154-183, followed by the loss at :723,
 * MSELoss(results, target)):
 *   target      [batch, steps, n] target trajectory
 *   loss_parts  [batch, 4] written by the forward: partial sums whose total is
 *               sum over steps and points of (u - target)^2 (one partial per
 *               wavefront of a trajectory's team, unused slots zero)
 * inc_burgers_fwd_mse is inc_burgers_fwd plus loss_parts; traj is still
 * written when given (the adjoint reads it).
 * inc_burgers_adj_mse is inc_burgers_adj with the cotangent of output t
 *   g_traj[t] (optional, NULL for none) + scale[0] * (traj[t] - target[t]),
 * scale a device scalar of the element type (2 dL/dloss / (batch*steps*n)
 * for the mean), read by the kernel, so no host synchronisation.
 */
int inc_burgers_fwd_mse(const inc_burgers_desc* d, const void* u, const void* s,
                        void* u_out, int steps, void* traj, void* stages,
                        const void* target, void* loss_parts, void* stream);
int inc_burgers_adj_mse(const inc_burgers_desc* d, const void* stages,
                        const void* traj, const void* target, const void* scale,
                        const void* g_traj, void* g_u, void* g_s, int steps,
                        void* stream);
/* inc_burgers_adj_mse plus g_nu [batch] (see inc_burgers_rhs_adj_nu). */
int inc_burgers_adj_mse_nu(const inc_burgers_desc* d, const void* stages,
                           const void* traj, const void* target, const void* scale,
                           const void* g_traj, void* g_u, void* g_s, void* g_nu,
                           int steps, void* stream);

/*
 * INC rollout with the correction network inside the kernel:
 * simulate_rollout(model, solver, u, steps, 'INC') (trainer/trainer_1d.py:154-183)
 * for a stack of circular Conv1d layers with GELU between them (the
 * reference's correction networks stack such convolutions,
 * models/models_1d/resnet.py:5-52):
 *   s = scale * conv_{L-1}(gelu(... gelu(conv_0(u)) ...)),
 *   conv_0: 1 -> C channels, conv_1..L-2: C -> C, conv_{L-1}: C -> 1,
 * evaluated from the current state before every step and held over the
 * step's stages (solver.correction_s).  One launch for the whole rollout; no
 * gradient (validation / test rollouts).
 *   weight[l]  [Co][Ci][K] of the state's element type; bias[l] [Co] or NULL
 *   traj       [batch, steps, n] states after each step, or NULL
 *   corr       [batch, steps, n] corrections of each step, or NULL
 * Supported: WENO convection, C = 32 or 48, K = 5 or 7, 2..8 layers, n up to
 * 512 (float) or 256 (double) (LDS: (C + 2) * n elements per block).
 */
typedef struct inc_cnn_desc {
  int32_t layers;
  int32_t channels;
  int32_t kernel;
  int32_t reserved;
  double scale;
  const void* weight[8];
  const void* bias[8];
} inc_cnn_desc;

int inc_burgers_inc_fwd(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* u,
                        void* u_out, int steps, void* traj, void* corr, void* stream);

/*
 * Training form of the same rollout: the simulate_rollout loop of
 * CorrectionTrainer.process_batch (trainer/trainer_1d.py:699-727, rollout at
 * :154-183) differentiated through the solver and the network.
 * inc_burgers_inc_train_fwd is inc_burgers_inc_fwd that also writes every
 * step's stage states:
 *   stages [batch, steps, ns, n] (ns = 3 for SSP-RK3, 1 for Euler)
 * inc_burgers_inc_train_adj is its adjoint.  From the stages, the cotangents
 * of the trajectory (g_traj [batch, steps, n] or NULL) and of the
 * corrections (g_corr [batch, steps, n] or NULL) it writes the cotangent of
 * the initial state (g_u [batch, n]) and each trajectory's gradient of the
 * network's parameters:
 *   g_w [batch, P], row b = [w0, b0, w1, b1, ..., w_{L-1}, b_{L-1}] in the
 *   layout of the weight and bias tensors, P = 2CK + C + 1 + (L-2)(C C K + C);
 * the batch gradient is the sum of the rows.  The network is evaluated again
 * at every step's input state inside the kernel; no activation is stored.
 * Supported by the adjoint: float, WENO convection, fixed time step, C = 32,
 * K = 5 or 7, 2..4 layers, 8 <= n <= 128 and the block's LDS within 160 KiB:
 * the hidden layers' weights (L - 2) * C * (C K + 4) and the activations
 * ((L + 1) n + K - 1) * (C + 4) + 2 n floats.  Other shapes return
 * INC_EUNSUPPORTED.
 */
int inc_burgers_inc_train_fwd(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* u,
                              void* u_out, int steps, void* traj, void* corr, void* stages,
                              void* stream);
int inc_burgers_inc_train_adj(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* stages,
                              const void* g_traj, const void* g_corr, void* g_u, void* g_w, int steps,
                              void* stream);
/*
 * The same training rollout with the trainer's MSE to a target trajectory
 * fused in (CorrectionTrainer.process_batch, trainer/trainer_1d.py:716-723:
 * loss_fn(results, target) after simulate_rollout).  The forward also writes
 * loss_parts [batch, 4], partial sums whose total is the sum over steps and
 * points of (traj - target)^2 (one partial per wavefront, unused slots
 * zero); traj is required (the adjoint reads it).  The adjoint adds
 * scale[0] * (traj[t] - target[t]) to the cotangent of every output, scale a
 * device scalar (2 dL/dloss / (batch * steps * n) for the mean), so the
 * backward needs no host synchronisation and no cotangent tensor.
 */
int inc_burgers_inc_train_fwd_mse(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* u, void* u_out,
                                  int steps, void* traj, void* corr, void* stages, const void* target,
                                  void* loss_parts, void* stream);
int inc_burgers_inc_train_adj_mse(const inc_burgers_desc* d, const inc_cnn_desc* cnn, const void* stages,
                                  const void* traj, const void* target, const void* scale, const void* g_traj,
                                  const void* g_corr, void* g_u, void* g_w, int steps, void* stream);

/*
 * Adaptive sub-stepping per trajectory (BurgersSolverTorch.step_adaptive,
 * solver_1d.py:148-179, which sub-steps the whole batch to
 * dt_j = min(remaining, CFL * dx / max|u|) with max|u| read back to the host
 * every sub-step): here max|u| is each trajectory's own, reduced on the
 * device, so trajectories sub-step independently and nothing synchronises
 * with the host.  Each step of dt takes at most max_sub (<= 1000) sub-steps,
 * the last one covering what remains.  Gradients treat the sub-step sizes as
 * constants, as the reference's .item() does.
 *   nsub   [batch, steps] int32       sub-steps taken (output)
 *   dts    [batch, steps, max_sub]    their sizes (output, d->dtype)
 *   stages [batch, steps, max_sub, ns, n] stage states for the adjoint, or NULL
 */
int inc_burgers_adaptive_fwd(const inc_burgers_desc* d, double cfl, int max_sub,
                             const void* u, const void* s, void* u_out, int steps,
                             void* traj, void* nsub, void* dts, void* stages,
                             void* stream);
/* Adjoint of inc_burgers_adaptive_fwd from its nsub, dts and stages. */
int inc_burgers_adaptive_adj(const inc_burgers_desc* d, int max_sub, const void* nsub,
                             const void* dts, const void* stages, const void* g_traj,
                             void* g_u, void* g_s, int steps, void* stream);
/* inc_burgers_adaptive_adj plus g_nu [batch] (see inc_burgers_rhs_adj_nu). */
int inc_burgers_adaptive_adj_nu(const inc_burgers_desc* d, int max_sub, const void* nsub,
                                const void* dts, const void* stages, const void* g_traj,
                                void* g_u, void* g_s, void* g_nu, int steps, void* stream);

/*
 * A ragged batch of Burgers trajectories with different resolution,
 * viscosity, time step and horizon, packed end to end: trajectory b occupies
 * [offset[b], offset[b] + n[b]) of every packed [total] array.  Batched form
 * of BurgersSolverTorch.step (solver_1d.py:181-195) over the uneven batches
 * of the INC training set (one trajectory per sample, data_loader_1d.py:8-67);
 * the source is the correction s only.
 */
typedef struct inc_ragged_desc {
  int32_t dtype;      /* INC_F32 / INC_F64 */
  int32_t scheme;     /* INC_EULER / INC_SSPRK3 */
  int32_t convection; /* INC_WENO / INC_FDM */
  int32_t count;      /* trajectories */
  int64_t total;      /* sum of n[b] */
  int32_t nmin, nmax; /* smallest / largest n[b] (n[b] >= 8; LDS sizing) */
  int32_t team;       /* threads per trajectory: 64, 128 or 256 (0 = 64) */
  int32_t step0;      /* rollout step index of this call's first step */
  const int32_t* n;       /* [count] points per trajectory (device) */
  const int64_t* offset;  /* [count] start in the packed arrays (device) */
  const int32_t* order;   /* [count] issue order (longest first), or NULL */
  const int32_t* horizon; /* [count] steps a trajectory takes; at rollout step
                             k >= horizon[b] its state is carried unchanged;
                             NULL = every step */
  const void* nu;         /* [count] viscosity (dtype) */
  const void* dt;         /* [count] time step (dtype) */
  const void* dx;         /* [count] grid spacing (dtype) */
} inc_ragged_desc;

/*
 * `steps` steps of every trajectory of a ragged batch in one launch, with the
 * correction s ([total] or NULL) fixed over the steps.  u, u_out: [total];
 * traj: [steps, total] or NULL; stages: [steps, ns, total] or NULL.
 */
int inc_burgers_ragged_fwd(const inc_ragged_desc* d, const void* u, const void* s, void* u_out, int steps,
                           void* traj, void* stages, void* stream);
/* Adjoint of inc_burgers_ragged_fwd: g_traj [steps, total] -> g_u [total], g_s [total] (or NULL). */
/*
 * Stability numbers of one Burgers step from the state u, per trajectory:
 * cfl [batch, 2] (the state's dtype) = (max_i |u_i| dt / dx, nu dt / dx^2),
 * computed on the GPU with no host synchronisation.  The reference reads the
 * batch-wide max |u| back to the host to size its adaptive sub-steps
 * (solvers/solver_1d.py:158-161); a fixed-step caller uses these to flag the
 * trajectories whose step leaves the stable region.  A NaN in a trajectory
 * makes its Courant number NaN.  inc_burgers_ragged_cfl does the same for a
 * ragged batch ([count, 2] in the packed order, each trajectory with its own
 * n, dt, dx and nu).
 */
int inc_burgers_cfl(const inc_burgers_desc* d, const void* u, void* cfl, void* stream);
int inc_burgers_ragged_cfl(const inc_ragged_desc* d, const void* u, void* cfl, void* stream);
int inc_burgers_ragged_adj(const inc_ragged_desc* d, const void* stages, const void* g_traj, void* g_u, void* g_s,
                           int steps, void* stream);

/*
 * Description of a KS problem (KSSolverTorch, solver_1d.py:198-212).
 * 4 <= n <= 1024: powers of two from 16 run the register-transpose FFT
 * (radix chosen by `radix`); float32 n = 3 * 2^k or 5 * 2^k (96 .. 768) a
 * radix-3/5 register step before power-of-two register transforms; every
 * other n a mixed-radix transform through LDS (`radix` ignored for both).
 */
typedef struct inc_ks_desc {
  int32_t dtype;      /* INC_F32 / INC_F64 */
  int32_t scheme;     /* INC_ETD1 / INC_ETRK2 (reference) / INC_ETDRK4 */
  int32_t batch;
  int32_t n;
  int32_t substeps;   /* sub-steps per step() call: 2 = reference double step */
  int32_t dealias;    /* 2/3-rule mask on the nonlinear term (ETDRK4) */
  int32_t radix;      /* points per lane of the FFT: 2, 4 or 8 (float64: 2 or 4) with n/radix <= 256;
                         0 = library default (2 while radix-2 teams need at most two wavefronts
                         per SIMD, else 4) */
  int32_t reserved;
  double dt;
  const void* table;        /* coefficient tables from inc_ks_tables */
  const void* twiddle;      /* FFT twiddles written by inc_ks_tables */
  const int32_t* table_idx; /* [batch] table per trajectory, or NULL = table 0
                               for all (then trajectories are FFT-packed in pairs) */
} inc_ks_desc;

/* Floats per wavenumber in a KS coefficient table. */
#define INC_KS_TABLE_STRIDE 8
/*
 * Build coefficient tables, one per domain size in L ([count] float64,
 * device): table_out holds count*(n/2+1)*INC_KS_TABLE_STRIDE elements of
 * dtype (the tables, row f = |wavenumber index|) followed by 2n elements of
 * FFT twiddles exp(-2 pi i k/n), whose address goes into inc_ks_desc.twiddle.
 * ETD1/ETRK2: E, phi1, phi2 with the reference's branches
 * (solver_1d.py:245-246, :259-267); ETDRK4: E, E2, Q, f1, f2, f3 by the
 * contour integral; all evaluated in float64 on the device; the nonlinear
 * factor k/2 (times the dealias mask) completes each row.
 */
int inc_ks_tables(int dtype, int scheme, int n, double dt, int dealias,
                  const double* L, int count, void* table_out, void* stream);
/*
 * Advance `steps` step() calls (each `substeps` sub-steps) from u with the
 * fixed correction s (or NULL).  Same output buffers as inc_burgers_fwd;
 * stages: [batch, steps, substeps*ns, n] with ns = 1 (ETD1), 2 (ETRK2),
 * 4 (ETDRK4) physical stage states.
 */
int inc_ks_fwd(const inc_ks_desc* d, const void* u, const void* s, void* u_out,
               int steps, void* traj, void* stages, void* stream);
/* Adjoint of inc_ks_fwd; same contract as inc_burgers_adj. */
int inc_ks_adj(const inc_ks_desc* d, const void* stages, const void* g_traj,
               void* g_u, void* g_s, int steps, void* stream);
/* Fused-MSE rollout; same contract as inc_burgers_fwd_mse / _adj_mse. */
int inc_ks_fwd_mse(const inc_ks_desc* d, const void* u, const void* s, void* u_out,
                   int steps, void* traj, void* stages, const void* target,
                   void* loss_parts, void* stream);
int inc_ks_adj_mse(const inc_ks_desc* d, const void* stages, const void* traj,
                   const void* target, const void* scale, const void* g_traj,
                   void* g_u, void* g_s, int steps, void* stream);

/*
 * Circular 1D convolution of the INC correction network s(u) (the reference's
 * correction networks stack 5-tap circular Conv1d layers,
 * models/models_1d/resnet.py:5-52; torch.nn.Conv1d(padding_mode='circular')).
 *   y[b,o,i] = bias[o] + sum_{c,k} w[o,c,k] x[b,c,(i+k-K/2) mod n]
 * x [B, Ci, n], w [Co, Ci, K] (K odd <= 9), bias [Co] or NULL, y [B, Co, n].
 */
int inc_conv1d_fwd(int dtype, int B, int Ci, int Co, int n, int K, const void* x, const void* w,
                   const void* bias, void* y, void* stream);
/* Input gradient gx [B, Ci, n] from the output gradient g [B, Co, n]. */
int inc_conv1d_bwd_data(int dtype, int B, int Ci, int Co, int n, int K, const void* g, const void* w,
                        void* gx, void* stream);
/*
 * Weight and bias gradients gw [Co, Ci, K], gb [Co] (or NULL), summed over
 * the batch in a fixed order; work: caller scratch of work_elems elements
 * (>= Co*Ci*K + Co; up to 512 times that is used when available).
 */
int inc_conv1d_bwd_weight(int dtype, int B, int Ci, int Co, int n, int K, const void* g, const void* x,
                          void* gw, void* gb, void* work, int64_t work_elems, void* stream);
/*
 * inc_conv1d_bwd_weight without a workspace argument and without any device
 * allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The per-block partial sums go to a 32 MiB array that
 * the library holds on each GPU it runs on until the process ends; a call
 * whose partial row (Co*Ci*K + Co elements) does not fit returns INC_EINVAL.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (different streams, or host threads) race and must be ordered by the
 * caller.  Results equal inc_conv1d_bwd_weight's given a workspace of the
 * same size.
 */
int inc_conv1d_bwd_weight_noalloc(int dtype, int B, int Ci, int Co, int n, int K, const void* g, const void* x,
                                  void* gw, void* gb, void* stream);

/*
 * Activation fused into the convolution that produces it (the correction
 * networks put GELU between their convolutions, models/models_1d/resnet.py:36-52):
 * with v = conv(x) + bias, y = act(v) and dy = act'(v) [B, Co, n].  y feeds the
 * next layer (and its weight gradient); dy is what the input gradient of that
 * next layer multiplies by (inc_conv1d_bwd_data_mul).  act: INC_ACT_NONE (dy
 * unused, the plain inc_conv1d_fwd), INC_ACT_GELU (exact erf form, as
 * torch.nn.GELU()) or INC_ACT_RELU (torch.relu, dy = 1 where y > 0 else 0:
 * ResNet1D's and UNet1D's activation, models/models_1d/resnet.py:5-20).  dy
 * may be NULL when no gradient is wanted: y alone is written, without the
 * derivative's evaluation and store.
 */
enum { INC_ACT_NONE = 0, INC_ACT_GELU = 1, INC_ACT_RELU = 2 };
int inc_conv1d_fwd_act(int dtype, int act, int B, int Ci, int Co, int n, int K, const void* x,
                       const void* w, const void* bias, void* y, void* dy, void* stream);
/* gx = (conv^T g) * m elementwise, m [B, Ci, n] (NULL: the plain inc_conv1d_bwd_data). */
int inc_conv1d_bwd_data_mul(int dtype, int B, int Ci, int Co, int n, int K, const void* g,
                            const void* w, const void* m, void* gx, void* stream);

/*
 * Ragged rows of the correction network's convolutions: row r (one trajectory
 * of a ragged batch) has n[r] >= K/2 points, and its [C][n[r]] block of a
 * C-channel array starts at element C * offset[r], so the packed [total]
 * state of inc_ragged_desc is the C = 1 case and every layer of the network
 * runs over all resolutions in one launch.  The work is a list of
 * INC_CONV_ROW_TILE-position tiles (row, first position) in device memory; a
 * call processes the first njobs tiles (a list ordered by horizon, longest
 * first, leaves out the trajectories whose rollout has ended by taking a
 * prefix).  Same arithmetic as inc_conv1d_fwd_act / _bwd_data_mul /
 * _bwd_weight on each row.
 */
#define INC_CONV_ROW_TILE 128
typedef struct inc_conv_rows {
  int32_t count;         /* rows */
  int32_t njobs;         /* tiles processed: the first njobs of tile */
  const int32_t* tile;   /* [2 * njobs] (row, first position) pairs */
  const int32_t* n;      /* [count] points per row */
  const int64_t* offset; /* [count] first point of each row in the packed layout */
} inc_conv_rows;

int inc_conv1d_rows_fwd_act(int dtype, int act, const inc_conv_rows* r, int Ci, int Co, int K, const void* x,
                            const void* w, const void* bias, void* y, void* dy, void* stream);
int inc_conv1d_rows_bwd_data_mul(int dtype, const inc_conv_rows* r, int Ci, int Co, int K, const void* g,
                                 const void* w, const void* m, void* gx, void* stream);
int inc_conv1d_rows_bwd_weight(int dtype, const inc_conv_rows* r, int Ci, int Co, int K, const void* g, const void* x,
                               void* gw, void* gb, void* work, int64_t work_elems, void* stream);

/*
 * Resolution changes of the UNet1D correction network
 * (models/models_1d/unet.py:5-53).
 *
 * inc_maxpool2_fwd is torch.nn.MaxPool1d(kernel_size=2, stride=2)
 * (unet.py:11,13) over `rows` rows of n points (n even): y [rows, n/2],
 * y[r][i] = x[r][2i + 1] if x[r][2i + 1] > x[r][2i] or it is NaN, else
 * x[r][2i] (torch's window scan: the first element wins ties).
 * inc_maxpool2_bwd routes gy to the element the forward picked, recomputed
 * from x; the other element of each pair gets 0.
 */
int inc_maxpool2_fwd(int dtype, int64_t rows, int n, const void* x, void* y, void* stream);
int inc_maxpool2_bwd(int dtype, int64_t rows, int n, const void* x, const void* gy, void* gx, void* stream);
/*
 * torch.cat((ConvTranspose1d(k=2, s=2)(x), skip), dim=1) (unet.py:16-21,
 * :37-42) from y2, the transposed convolution computed as a 1x1 convolution
 * with 2C output channels (W'[t*C + c][ci] = W[ci][c][t], bias repeated):
 *   out [B, C + Cs, 2n]: out[b][c][2i + t] = y2[b][t*C + c][i] (c < C),
 *                        out[b][C + c][j] = skip[b][c][j]
 * y2 [B, 2C, n]; skip [B, Cs, 2n] (Cs = 0: no skip connection).
 * inc_upsample2_cat_bwd splits the gradient g [B, C + Cs, 2n] back into
 * gy2 [B, 2C, n] and gskip [B, Cs, 2n] (NULL: not wanted).
 */
int inc_upsample2_cat(int dtype, int B, int C, int n, int Cs, const void* y2, const void* skip, void* out,
                      void* stream);
int inc_upsample2_cat_bwd(int dtype, int B, int C, int n, int Cs, const void* g, void* gy2, void* gskip,
                          void* stream);

/*
 * INC training rollout of a ragged batch: simulate_rollout(model, solver, u0,
 * steps, 'INC') (trainer/trainer_1d.py:154-183) over the uneven batch of
 * inc_burgers_ragged_fwd (one solver per batch in trainer_1d.py:338-358),
 * with a network of circular Conv1d layers 1 -> C -> ... -> C -> 1 and GELU
 * between them (inc_cnn_desc: width C <= 64, K odd <= 9, 2..8 layers,
 * s = scale * net(u); the reference's correction networks stack such
 * convolutions, models/models_1d/resnet.py:5-52).  Step k runs the network
 * as one ragged convolution per layer over the first active_tiles[k] tiles of
 * rows (host array [steps]; NULL = rows->njobs every step), then one ragged
 * solver step with s as its correction; trajectories past their horizon keep
 * their state and get no correction.
 *   traj   [steps, total]        states after each step
 *   corr   [steps, total] or NULL corrections of each step (0 past a horizon)
 *   stages [steps, ns, total]    stage states (ns = 3 SSP-RK3, 1 Euler), or
 *                                NULL when no adjoint follows
 * inc_ragged_inc_adj is its adjoint: from u0, traj, stages and the
 * trajectory cotangent g_traj [steps, total] it writes g_u0 [total] and the
 * network's parameter gradient summed over the batch, g_w [P] =
 * [w0, b0, w1, b1, ...] in the weight and bias tensors' layouts.  The
 * weight gradients are summed over steps in a fixed order (deterministic).
 * The hidden activations at every step either come from acts or are
 * evaluated again (acts NULL in both calls):
 *   acts [steps, L - 1, 2, C * total]  per step and hidden layer the GELU
 *                                      output and its derivative, in the
 *                                      convolutions' layout; written by the
 *                                      forward, read by the adjoint, which
 *                                      must get the buffer the forward
 *                                      filled with the same descriptor,
 *                                      network, rows, u0 and steps
 * Stored activations replace the adjoint's repeat of the network's hidden
 * layers (one convolution per layer and step) at the cost of that buffer.
 * work: workspace of work_elems elements of the state's type, at least what
 * inc_ragged_inc_workspace reports for kind 0 (forward), 1 (adjoint
 * evaluating the network again) or 2 (adjoint reading acts).
 */
int inc_ragged_inc_workspace(const inc_ragged_desc* d, const inc_cnn_desc* cnn, int adjoint, int64_t* elems);
int inc_ragged_inc_fwd(const inc_ragged_desc* d, const inc_cnn_desc* cnn, const inc_conv_rows* rows,
                       const int32_t* active_tiles, const void* u0, int steps, void* traj, void* corr, void* stages,
                       void* acts, void* work, int64_t work_elems, void* stream);
int inc_ragged_inc_adj(const inc_ragged_desc* d, const inc_cnn_desc* cnn, const inc_conv_rows* rows,
                       const int32_t* active_tiles, const void* u0, const void* traj, const void* stages,
                       const void* acts, const void* g_traj, void* g_u0, void* g_w, int steps, void* work,
                       int64_t work_elems, void* stream);

/*
 * INC training rollout of a KS batch in one call: simulate_rollout(model,
 * solver, u0, steps, 'INC', L) (trainer/trainer_1d.py:154-183) with a
 * KSSolverTorch (solvers/solver_1d.py:198-288) and the network of
 * inc_ragged_inc_fwd (inc_cnn_desc: circular Conv1d layers with GELU between
 * them, s = scale * net(u)).  Step k evaluates the network on the batch's
 * state (one convolution launch per layer over batch x n) and runs one
 * inc_ks_fwd step with s as its correction.  Step-major buffers:
 *   traj   [steps, batch, n]           states after each step
 *   corr   [steps, batch, n] or NULL   corrections of each step
 *   stages [steps, batch, nst, n]      stage states (nst as for inc_ks_fwd),
 *                                      or NULL when no adjoint follows
 * inc_ks_inc_adj is its adjoint, as inc_ragged_inc_adj: from u0, traj,
 * stages, the trajectory cotangent g_traj [steps, batch, n] and the
 * corrections' cotangent g_corr [steps, batch, n] (or NULL) it writes g_u0
 * [batch, n] and the parameter gradient g_w [P] = [w0, b0, w1, b1, ...]
 * summed over the batch and the steps in a fixed order.  acts (or NULL in
 * both calls: the network evaluated again at every step's input state) and
 * the workspace kinds as for inc_ragged_inc_*, with total = batch * n.
 */
int inc_ks_inc_workspace(const inc_ks_desc* d, const inc_cnn_desc* cnn, int adjoint, int64_t* elems);
int inc_ks_inc_fwd(const inc_ks_desc* d, const inc_cnn_desc* cnn, const void* u0, int steps, void* traj, void* corr,
                   void* stages, void* acts, void* work, int64_t work_elems, void* stream);
int inc_ks_inc_adj(const inc_ks_desc* d, const inc_cnn_desc* cnn, const void* u0, const void* traj,
                   const void* stages, const void* acts, const void* g_traj, const void* g_corr, void* g_u0,
                   void* g_w, int steps, void* work, int64_t work_elems, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* INC_AMD_H */
