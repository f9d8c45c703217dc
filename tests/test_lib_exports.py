"""CPU checks of the C ABI library: it loads, exports every symbol the header
declares with the ctypes signature the package binds, and maps argument errors
(no kernel launches, no GPU needed)."""
import ctypes
import os
import re

import pytest

import inc_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "inc_amd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(inc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_what_the_binding_binds():
    assert header_functions() == sorted(L.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = L.load()
    for name in header_functions():
        assert hasattr(lib, name), name
    assert lib.inc_version() >= 100


def test_struct_sizes_match_header_layout():
    # int32 x8 + double x6 + pointer x8 = 32 + 48 + 64
    assert ctypes.sizeof(L.BurgersDesc) == 144
    # int32 x8 + double + pointer x3 = 32 + 8 + 24
    assert ctypes.sizeof(L.KSDesc) == 64
    # int32 x4 + int64 + int32 x4 + pointer x7 = 16 + 8 + 16 + 56
    assert ctypes.sizeof(L.RaggedDesc) == 96


def test_argument_errors_map_to_codes_without_launch():
    lib = L.load()
    # bad WENO direction: rejected before any launch
    rc = lib.inc_weno_fwd(L.INC_F32, 7, None, None, 1, 16, None)
    assert rc == L.INC_EINVAL and b"left" in lib.inc_last_error()
    with pytest.raises(ValueError):
        L.check(rc, "weno")
    # Burgers grid too small
    d = L.BurgersDesc(dtype=L.INC_F32, batch=1, n=4, dx=0.1)
    rc = lib.inc_burgers_fwd(ctypes.byref(d), None, None, None, 1, None, None, None)
    assert rc == L.INC_EINVAL and b"resolution" in lib.inc_last_error()
    # unknown equation type
    d = L.BurgersDesc(dtype=L.INC_F32, batch=1, n=64, dx=0.1, source=9)
    assert lib.inc_burgers_fwd(ctypes.byref(d), None, None, None, 1, None, None, None) == L.INC_EINVAL
    # KS resolution outside [4, 1024] (every size inside has a transform)
    for n in (3, 1025):
        k = L.KSDesc(dtype=L.INC_F32, scheme=L.INC_ETDRK4, batch=1, n=n, substeps=1)
        rc = lib.inc_ks_fwd(ctypes.byref(k), None, None, None, 1, None, None, None)
        assert rc == L.INC_EUNSUPPORTED and b"resolution" in lib.inc_last_error()
    # adjoint without saved stages
    k = L.KSDesc(dtype=L.INC_F32, scheme=L.INC_ETD1, batch=1, n=64, substeps=1, table=1)
    assert lib.inc_ks_adj(ctypes.byref(k), None, None, None, None, 1, None) == L.INC_EINVAL
    # ragged batch with a trajectory below 8 points
    r = L.RaggedDesc(dtype=L.INC_F32, count=1, total=4, nmin=4, nmax=4)
    rc = lib.inc_burgers_ragged_fwd(ctypes.byref(r), None, None, None, 1, None, None, None)
    assert rc == L.INC_EINVAL and b"8 points" in lib.inc_last_error()
    # adaptive sub-stepping: CFL number and sub-step cap are checked before any launch
    d = L.BurgersDesc(dtype=L.INC_F32, batch=1, n=64, dx=0.1, dt=0.01)
    rc = lib.inc_burgers_adaptive_fwd(ctypes.byref(d), 0.0, 8, None, None, None, 1, None, None, None, None, None)
    assert rc == L.INC_EINVAL and b"CFL" in lib.inc_last_error()
    rc = lib.inc_burgers_adaptive_fwd(ctypes.byref(d), 0.5, 0, None, None, None, 1, None, None, None, None, None)
    assert rc == L.INC_EINVAL and b"max_sub" in lib.inc_last_error()
    rc = lib.inc_burgers_adaptive_adj(ctypes.byref(d), 8, None, None, None, None, None, None, 1, None)
    assert rc == L.INC_EINVAL and b"nsub" in lib.inc_last_error()
    # convolution with an unknown activation
    rc = lib.inc_conv1d_fwd_act(L.INC_F32, 7, 1, 1, 1, 16, 3, None, None, None, None, None, None)
    assert rc == L.INC_EINVAL and b"activation" in lib.inc_last_error()
    # fused MSE rollouts: the forward needs a target and loss partials, the
    # adjoint the stored trajectory, the target and the device scale
    d = L.BurgersDesc(dtype=L.INC_F32, batch=1, n=64, dx=0.1, dt=0.01)
    rc = lib.inc_burgers_fwd_mse(ctypes.byref(d), None, None, None, 1, None, None, None, None, None)
    assert rc == L.INC_EINVAL and b"target" in lib.inc_last_error()
    rc = lib.inc_burgers_adj_mse(ctypes.byref(d), 1, 1, 1, None, None, None, None, 1, None)
    assert rc == L.INC_EINVAL and b"scale" in lib.inc_last_error()
    k = L.KSDesc(dtype=L.INC_F32, scheme=L.INC_ETDRK4, batch=1, n=64, substeps=1, table=1)
    rc = lib.inc_ks_fwd_mse(ctypes.byref(k), None, None, None, 1, None, None, 1, None, None)
    assert rc == L.INC_EINVAL and b"loss_parts" in lib.inc_last_error()
    rc = lib.inc_ks_adj_mse(ctypes.byref(k), 1, None, 1, 1, None, None, None, 1, None)
    assert rc == L.INC_EINVAL and b"traj" in lib.inc_last_error()
    # training rollout with the network inside: the forward needs stages; the
    # adjoint runs float32, width 32, up to 4 layers and n <= 128
    d = L.BurgersDesc(dtype=L.INC_F32, batch=1, n=64, dx=0.1, dt=0.01)
    cn = L.CnnDesc(layers=4, channels=32, kernel=5, scale=1.0)
    rc = lib.inc_burgers_inc_train_fwd(ctypes.byref(d), ctypes.byref(cn), None, None, 1, None, None, None, None)
    assert rc == L.INC_EINVAL and b"stages" in lib.inc_last_error()
    rc = lib.inc_burgers_inc_train_adj(ctypes.byref(d), ctypes.byref(cn), None, None, None, None, None, 1, None)
    assert rc == L.INC_EINVAL and b"stages" in lib.inc_last_error()
    cn48 = L.CnnDesc(layers=4, channels=48, kernel=5, scale=1.0)
    rc = lib.inc_burgers_inc_train_adj(ctypes.byref(d), ctypes.byref(cn48), 1, None, None, 1, 1, 1, None)
    assert rc == L.INC_EUNSUPPORTED and b"width 32" in lib.inc_last_error()
    d64 = L.BurgersDesc(dtype=L.INC_F64, batch=1, n=64, dx=0.1, dt=0.01)
    rc = lib.inc_burgers_inc_train_adj(ctypes.byref(d64), ctypes.byref(cn), 1, None, None, 1, 1, 1, None)
    assert rc == L.INC_EUNSUPPORTED and b"float32" in lib.inc_last_error()
    d256 = L.BurgersDesc(dtype=L.INC_F32, batch=1, n=256, dx=0.1, dt=0.01)
    rc = lib.inc_burgers_inc_train_adj(ctypes.byref(d256), ctypes.byref(cn), 1, None, None, 1, 1, 1, None)
    assert rc == L.INC_EUNSUPPORTED and b"n <= 128" in lib.inc_last_error()
    # successful-path state is reset by the next call
    assert lib.inc_weno_fwd(L.INC_F32, 0, None, None, 0, 16, None) == L.INC_OK
    assert lib.inc_last_error() == b""


def test_product_refuses_cpu_tensors():
    import torch
    import inc_amd
    with pytest.raises(ValueError, match="GPU"):
        inc_amd.weno_reconstruction(torch.zeros(2, 16), "left")
    with pytest.raises(ValueError):
        inc_amd.weno_reconstruction(torch.zeros(2, 16), "up")


def test_fused_training_shape_rule_matches_the_library():
    # The Python predicate that routes a training rollout to the fused kernels
    # must accept exactly what inc_burgers_inc_train_adj accepts.  With batch 0
    # the library validates every argument and returns before any launch.
    from inc_amd.solver_1d import inc_train_shape_supported
    lib = L.load()
    seen = {True: 0, False: 0}
    for c in (16, 32, 48):
        for k in (3, 5, 7):
            for layers in (2, 3, 4, 5):
                for n in (8, 16, 37, 64, 65, 96, 100, 112, 120, 128, 129, 256):
                    d = L.BurgersDesc(dtype=L.INC_F32, batch=0, n=n, dx=0.1, dt=0.01, scheme=L.INC_SSPRK3)
                    cn = L.CnnDesc(layers=layers, channels=c, kernel=k, scale=1.0)
                    for i in range(layers):
                        cn.weight[i] = 1
                    rc = lib.inc_burgers_inc_train_adj(ctypes.byref(d), ctypes.byref(cn), 1, None, None, 1, 1, 1,
                                                       None)
                    want = inc_train_shape_supported(c, k, layers, n)
                    assert (rc == L.INC_OK) == want, (c, k, layers, n, rc, lib.inc_last_error())
                    seen[want] += 1
    assert seen[True] > 10 and seen[False] > 10
