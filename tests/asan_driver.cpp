// Host-side argument validation of libinc_amd under AddressSanitizer.
//
// Built by tests/test_asan_host.py against a host-only, ASan-instrumented
// build of the C-ABI sources (no device code, so nothing here may launch a
// kernel): every call below is rejected, or completes, before any HIP call.
// It checks the status codes and the thread-local error text; ASan aborts
// the process on an out-of-bounds access, a use after free or a double free
// in the host code (the error-text formatting, descriptor checks, workspace
// size arithmetic).
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

#include "../include/inc_amd.h"

static int failures = 0;

static void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s (last error: %s)\n", what, inc_last_error());
    ++failures;
  }
}

static bool has(const char* needle) { return std::strstr(inc_last_error(), needle) != nullptr; }

int main() {
  expect(inc_version() >= 100, "version");

  // WENO direction
  expect(inc_weno_fwd(INC_F32, 7, nullptr, nullptr, 1, 16, nullptr) == INC_EINVAL && has("left"), "weno direction");
  expect(inc_weno_fwd(INC_F32, 0, nullptr, nullptr, 0, 16, nullptr) == INC_OK && inc_last_error()[0] == 0,
         "empty weno call clears the error");

  // Burgers descriptors
  inc_burgers_desc d{};
  d.dtype = INC_F32;
  d.batch = 1;
  d.n = 4;
  d.dx = 0.1;
  expect(inc_burgers_fwd(&d, nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr) == INC_EINVAL &&
             has("resolution"),
         "burgers resolution");
  d.n = 64;
  d.source = 9;
  expect(inc_burgers_fwd(&d, nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr) == INC_EINVAL,
         "burgers source");
  d.source = INC_SRC_NONE;
  d.dt = 0.01;
  expect(inc_burgers_adaptive_fwd(&d, 0.0, 8, nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr, nullptr,
                                  nullptr) == INC_EINVAL &&
             has("CFL"),
         "adaptive CFL");
  expect(inc_burgers_fwd_mse(&d, nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr, nullptr, nullptr) ==
                 INC_EINVAL &&
             has("target"),
         "mse target");
  expect(inc_burgers_cfl(&d, nullptr, nullptr, nullptr) == INC_EINVAL && has("missing buffer"),
         "stability numbers need buffers");
  d.dx = 0.0;
  expect(inc_burgers_cfl(&d, nullptr, nullptr, nullptr) == INC_EINVAL && has("bad descriptor"),
         "stability numbers grid spacing");
  d.dx = 0.1;
  d.batch = 0;
  expect(inc_burgers_cfl(&d, nullptr, nullptr, nullptr) == INC_OK, "empty stability numbers");
  d.batch = 1;

  // the network inside the Burgers rollout
  inc_cnn_desc cn{};
  cn.layers = 4;
  cn.channels = 48;
  cn.kernel = 5;
  cn.scale = 1.0;
  const void* one = reinterpret_cast<const void*>(1);
  void* wone = reinterpret_cast<void*>(1);
  expect(inc_burgers_inc_train_adj(&d, &cn, one, nullptr, nullptr, wone, wone, 1, nullptr) == INC_EUNSUPPORTED &&
             has("width 32"),
         "train adjoint width");

  // KS
  inc_ks_desc k{};
  k.dtype = INC_F32;
  k.scheme = INC_ETDRK4;
  k.batch = 1;
  k.n = 2048;
  k.substeps = 1;
  expect(inc_ks_fwd(&k, nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr) == INC_EUNSUPPORTED &&
             has("resolution"),
         "ks resolution");
  k.n = 3;
  expect(inc_ks_fwd(&k, nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr) == INC_EUNSUPPORTED,
         "ks resolution below the smallest transform");
  k.n = 64;
  k.scheme = INC_ETD1;
  k.table = one;
  expect(inc_ks_adj(&k, nullptr, nullptr, nullptr, nullptr, 1, nullptr) == INC_EINVAL, "ks adjoint stages");

  // ragged batches
  inc_ragged_desc r{};
  r.dtype = INC_F32;
  r.count = 1;
  r.total = 4;
  r.nmin = 4;
  r.nmax = 4;
  expect(inc_burgers_ragged_fwd(&r, nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr) == INC_EINVAL &&
             has("8 points"),
         "ragged minimum size");
  expect(inc_burgers_ragged_cfl(&r, nullptr, nullptr, nullptr) == INC_EINVAL && has("missing buffer"),
         "ragged stability numbers need buffers");

  // workspace sizes: host arithmetic only
  r.total = 2048 * 400;
  r.count = 2048;
  r.nmin = 64;
  r.nmax = 1024;
  cn.channels = 32;
  cn.layers = 4;
  for (int l = 0; l < 4; ++l) cn.weight[l] = cn.bias[l] = one;
  int64_t elems = 0;
  expect(inc_ragged_inc_workspace(&r, &cn, 1, &elems) == INC_OK && elems > 4 * 32 * int64_t(r.total),
         "ragged INC workspace");
  int64_t stored = 0;
  expect(inc_ragged_inc_workspace(&r, &cn, 2, &stored) == INC_OK && stored < elems,
         "adjoint workspace reading stored activations is smaller");
  expect(inc_ragged_inc_workspace(&r, &cn, 3, &stored) == INC_EINVAL && has("kind"), "workspace kind");
  // the KS INC adjoint: the smaller workspace serves only the adjoint that
  // reads stored activations (rejected before any device call otherwise)
  inc_ks_desc kd{};
  kd.dtype = INC_F32;
  kd.scheme = INC_ETDRK4;
  kd.batch = 4;
  kd.n = 64;
  kd.substeps = 1;
  kd.table = one;
  int64_t full = 0, small = 0;
  expect(inc_ks_inc_workspace(&kd, &cn, 1, &full) == INC_OK && inc_ks_inc_workspace(&kd, &cn, 2, &small) == INC_OK &&
             small < full,
         "KS INC adjoint workspaces");
  expect(inc_ks_inc_adj(&kd, &cn, one, one, one, nullptr, one, nullptr, wone, wone, 3, wone, small, nullptr) ==
                 INC_EINVAL &&
             has("workspace"),
         "adjoint without stored activations needs the larger workspace");
  cn.layers = 9;
  expect(inc_ragged_inc_workspace(&r, &cn, 0, &elems) == INC_EINVAL, "network depth");

  // convolutions, pooling, upsampling
  expect(inc_conv1d_fwd_act(INC_F32, 7, 1, 1, 1, 16, 3, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) ==
                 INC_EINVAL &&
             has("activation"),
         "conv activation");
  expect(inc_conv1d_fwd(INC_F32, 1, 1, 1, 16, 4, nullptr, nullptr, nullptr, nullptr, nullptr) == INC_EUNSUPPORTED,
         "conv even kernel");
  expect(inc_conv1d_fwd(INC_F32, 1, 1, 1, 4, 5, nullptr, nullptr, nullptr, nullptr, nullptr) == INC_EUNSUPPORTED,
         "conv kernel longer than the row");
  expect(inc_conv1d_fwd(9, 1, 1, 1, 16, 3, nullptr, nullptr, nullptr, nullptr, nullptr) == INC_EINVAL,
         "conv dtype");
  expect(inc_conv1d_bwd_weight_noalloc(INC_F32, 1, 2048, 2048, 64, 9, nullptr, nullptr, nullptr, nullptr,
                                       nullptr) == INC_EINVAL &&
             has("scratch"),
         "no-alloc weight gradient over its scratch array");
  expect(inc_maxpool2_fwd(INC_F32, 4, 3, nullptr, nullptr, nullptr) == INC_EINVAL && has("even"), "pool odd rows");
  expect(inc_maxpool2_fwd(INC_F32, 0, 16, nullptr, nullptr, nullptr) == INC_OK, "empty pooling");
  expect(inc_upsample2_cat(INC_F32, 1, 0, 8, 4, nullptr, nullptr, nullptr, nullptr) == INC_EINVAL,
         "upsampling channels");
  expect(inc_upsample2_cat(INC_F32, 0, 4, 8, 4, nullptr, nullptr, nullptr, nullptr) == INC_OK, "empty upsampling");

  // the error text is per thread: another thread starts clean and its own
  // failure does not reach this thread
  expect(inc_weno_fwd(INC_F32, 7, nullptr, nullptr, 1, 16, nullptr) == INC_EINVAL, "weno direction again");
  std::string other;
  std::thread t([&] {
    other = inc_last_error();
    inc_maxpool2_fwd(INC_F32, 4, 3, nullptr, nullptr, nullptr);
  });
  t.join();
  expect(other.empty(), "fresh thread has no error text");
  expect(has("left"), "this thread keeps its own error text");

  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("asan driver: all checks passed\n");
  return 0;
}
