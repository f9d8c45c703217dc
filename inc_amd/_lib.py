"""ctypes binding of libinc_amd.so (declarations in include/inc_amd.h).

The library is loaded on first use.  There is no fallback: if the shared
library is missing or a tensor is not on the GPU, the call raises.
"""
import ctypes
import os
import threading

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libinc_amd.so")

INC_OK, INC_EINVAL, INC_EHIP, INC_EUNSUPPORTED = 0, 1, 2, 3
INC_F32, INC_F64 = 0, 1
INC_EULER, INC_SSPRK3 = 0, 1
INC_WENO, INC_FDM = 0, 1
INC_SRC_NONE, INC_SRC_LINEAR, INC_SRC_LEARNING = 0, 1, 2
INC_PART_CONVECTION, INC_PART_DIFFUSION, INC_PART_SOURCE = 1, 2, 4
INC_ETD1, INC_ETRK2, INC_ETDRK4 = 0, 1, 2
INC_KS_TABLE_STRIDE = 8
INC_ACT_NONE, INC_ACT_GELU, INC_ACT_RELU = 0, 1, 2

_vp = ctypes.c_void_p


class BurgersDesc(ctypes.Structure):
    """inc_burgers_desc (include/inc_amd.h)."""
    _fields_ = [
        ("dtype", ctypes.c_int32), ("scheme", ctypes.c_int32),
        ("convection", ctypes.c_int32), ("source", ctypes.c_int32),
        ("batch", ctypes.c_int32), ("n", ctypes.c_int32),
        ("J", ctypes.c_int32), ("team", ctypes.c_int32),
        ("dt", ctypes.c_double), ("dx", ctypes.c_double), ("L", ctypes.c_double),
        ("nu", ctypes.c_double), ("beta", ctypes.c_double), ("t0", ctypes.c_double),
        ("nu_ptr", _vp), ("beta_ptr", _vp), ("t_ptr", _vp), ("x", _vp),
        ("A", _vp), ("omega", _vp), ("phi", _vp), ("l", _vp),
    ]


class KSDesc(ctypes.Structure):
    """inc_ks_desc (include/inc_amd.h)."""
    _fields_ = [
        ("dtype", ctypes.c_int32), ("scheme", ctypes.c_int32),
        ("batch", ctypes.c_int32), ("n", ctypes.c_int32),
        ("substeps", ctypes.c_int32), ("dealias", ctypes.c_int32),
        ("radix", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("table", _vp), ("twiddle", _vp), ("table_idx", _vp),
    ]


class RaggedDesc(ctypes.Structure):
    """inc_ragged_desc (include/inc_amd.h)."""
    _fields_ = [
        ("dtype", ctypes.c_int32), ("scheme", ctypes.c_int32), ("convection", ctypes.c_int32),
        ("count", ctypes.c_int32), ("total", ctypes.c_int64),
        ("nmin", ctypes.c_int32), ("nmax", ctypes.c_int32), ("team", ctypes.c_int32), ("step0", ctypes.c_int32),
        ("n", _vp), ("offset", _vp), ("order", _vp), ("horizon", _vp), ("nu", _vp), ("dt", _vp), ("dx", _vp),
    ]


class CnnDesc(ctypes.Structure):
    """inc_cnn_desc (include/inc_amd.h)."""
    _fields_ = [
        ("layers", ctypes.c_int32), ("channels", ctypes.c_int32), ("kernel", ctypes.c_int32),
        ("reserved", ctypes.c_int32), ("scale", ctypes.c_double),
        ("weight", _vp * 8), ("bias", _vp * 8),
    ]


class ConvRows(ctypes.Structure):
    """inc_conv_rows (include/inc_amd.h)."""
    _fields_ = [
        ("count", ctypes.c_int32), ("njobs", ctypes.c_int32),
        ("tile", _vp), ("n", _vp), ("offset", _vp),
    ]


INC_CONV_ROW_TILE = 128

#: every symbol include/inc_amd.h declares, with its ctypes signature
SIGNATURES = {
    "inc_version": (ctypes.c_int, []),
    "inc_last_error": (ctypes.c_char_p, []),
    "inc_weno_fwd": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int64, ctypes.c_int, _vp]),
    "inc_weno_adj": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_int, _vp]),
    "inc_burgers_rhs": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.c_int, _vp, _vp, _vp, _vp]),
    "inc_burgers_rhs_adj": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.c_int, _vp, _vp, _vp, _vp]),
    "inc_burgers_rhs_adj_nu": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "inc_burgers_fwd": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "inc_burgers_adj": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "inc_burgers_adj_nu": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "inc_burgers_fwd_mse": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp,
                                           _vp, _vp]),
    "inc_burgers_adj_mse": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                           ctypes.c_int, _vp]),
    "inc_burgers_adj_mse_nu": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                              ctypes.c_int, _vp]),
    "inc_burgers_inc_fwd": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.POINTER(CnnDesc), _vp, _vp,
                                           ctypes.c_int, _vp, _vp, _vp]),
    "inc_burgers_inc_train_fwd": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.POINTER(CnnDesc), _vp, _vp,
                                                 ctypes.c_int, _vp, _vp, _vp, _vp]),
    "inc_burgers_inc_train_adj": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.POINTER(CnnDesc), _vp, _vp, _vp,
                                                 _vp, _vp, ctypes.c_int, _vp]),
    "inc_burgers_inc_train_fwd_mse": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.POINTER(CnnDesc), _vp, _vp,
                                                     ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inc_burgers_inc_train_adj_mse": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.POINTER(CnnDesc), _vp, _vp,
                                                     _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "inc_burgers_adaptive_fwd": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.c_double, ctypes.c_int, _vp, _vp,
                                                _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "inc_burgers_adaptive_adj": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                                _vp, ctypes.c_int, _vp]),
    "inc_burgers_adaptive_adj_nu": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), ctypes.c_int, _vp, _vp, _vp, _vp,
                                                   _vp, _vp, _vp, ctypes.c_int, _vp]),
    "inc_burgers_ragged_fwd": (ctypes.c_int, [ctypes.POINTER(RaggedDesc), _vp, _vp, _vp, ctypes.c_int, _vp, _vp,
                                              _vp]),
    "inc_burgers_ragged_adj": (ctypes.c_int, [ctypes.POINTER(RaggedDesc), _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "inc_burgers_cfl": (ctypes.c_int, [ctypes.POINTER(BurgersDesc), _vp, _vp, _vp]),
    "inc_burgers_ragged_cfl": (ctypes.c_int, [ctypes.POINTER(RaggedDesc), _vp, _vp, _vp]),
    "inc_ks_tables": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     _vp, ctypes.c_int, _vp, _vp]),
    "inc_ks_fwd": (ctypes.c_int, [ctypes.POINTER(KSDesc), _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "inc_ks_adj": (ctypes.c_int, [ctypes.POINTER(KSDesc), _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "inc_ks_fwd_mse": (ctypes.c_int, [ctypes.POINTER(KSDesc), _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "inc_ks_adj_mse": (ctypes.c_int, [ctypes.POINTER(KSDesc), _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "inc_conv1d_fwd": (ctypes.c_int, [ctypes.c_int] * 6 + [_vp, _vp, _vp, _vp, _vp]),
    "inc_conv1d_bwd_data": (ctypes.c_int, [ctypes.c_int] * 6 + [_vp, _vp, _vp, _vp]),
    "inc_conv1d_bwd_weight": (ctypes.c_int, [ctypes.c_int] * 6 + [_vp, _vp, _vp, _vp, _vp, ctypes.c_int64, _vp]),
    "inc_conv1d_bwd_weight_noalloc": (ctypes.c_int, [ctypes.c_int] * 6 + [_vp, _vp, _vp, _vp, _vp]),
    "inc_conv1d_fwd_act": (ctypes.c_int, [ctypes.c_int] * 7 + [_vp, _vp, _vp, _vp, _vp, _vp]),
    "inc_conv1d_bwd_data_mul": (ctypes.c_int, [ctypes.c_int] * 6 + [_vp, _vp, _vp, _vp, _vp]),
    "inc_conv1d_rows_fwd_act": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ConvRows)]
                                + [ctypes.c_int] * 3 + [_vp] * 6),
    "inc_conv1d_rows_bwd_data_mul": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ConvRows)] + [ctypes.c_int] * 3
                                     + [_vp] * 5),
    "inc_conv1d_rows_bwd_weight": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ConvRows)] + [ctypes.c_int] * 3
                                   + [_vp] * 5 + [ctypes.c_int64, _vp]),
    "inc_maxpool2_fwd": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int, _vp, _vp, _vp]),
    "inc_maxpool2_bwd": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "inc_upsample2_cat": (ctypes.c_int, [ctypes.c_int] * 5 + [_vp, _vp, _vp, _vp]),
    "inc_upsample2_cat_bwd": (ctypes.c_int, [ctypes.c_int] * 5 + [_vp, _vp, _vp, _vp]),
    "inc_ragged_inc_workspace": (ctypes.c_int, [ctypes.POINTER(RaggedDesc), ctypes.POINTER(CnnDesc), ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_int64)]),
    "inc_ragged_inc_fwd": (ctypes.c_int, [ctypes.POINTER(RaggedDesc), ctypes.POINTER(CnnDesc),
                                          ctypes.POINTER(ConvRows), _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                          ctypes.c_int64, _vp]),
    "inc_ragged_inc_adj": (ctypes.c_int, [ctypes.POINTER(RaggedDesc), ctypes.POINTER(CnnDesc),
                                          ctypes.POINTER(ConvRows), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                          ctypes.c_int, _vp, ctypes.c_int64, _vp]),
    "inc_ks_inc_workspace": (ctypes.c_int, [ctypes.POINTER(KSDesc), ctypes.POINTER(CnnDesc), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int64)]),
    "inc_ks_inc_fwd": (ctypes.c_int, [ctypes.POINTER(KSDesc), ctypes.POINTER(CnnDesc), _vp, ctypes.c_int, _vp, _vp,
                                      _vp, _vp, _vp, ctypes.c_int64, _vp]),
    "inc_ks_inc_adj": (ctypes.c_int, [ctypes.POINTER(KSDesc), ctypes.POINTER(CnnDesc), _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp, _vp, ctypes.c_int, _vp, ctypes.c_int64, _vp]),
}

_lib = None
_lock = threading.Lock()


def load(path=None):
    """Load (once) and return the ctypes handle of libinc_amd.so."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(
                f"libinc_amd.so not found at {p}; build it with `python -m inc_amd.build`")
        lib = ctypes.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def check(rc, what):
    if rc == INC_OK:
        return
    msg = load().inc_last_error().decode(errors="replace")
    if rc in (INC_EINVAL, INC_EUNSUPPORTED):
        raise ValueError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg}")


def dtype_code(dt):
    if dt == torch.float32:
        return INC_F32
    if dt == torch.float64:
        return INC_F64
    raise ValueError(f"inc_amd kernels take float32 or float64 tensors (got {dt})")


def require_gpu(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise ValueError(
                "inc_amd runs the solver step on the GPU (HIP); got a tensor on "
                f"{t.device}. Move the state to 'cuda'.")


def require_like(ref, *tensors, what="parameter"):
    """Every tensor on the GPU with ref's dtype.  The kernels read weights
    and biases through raw pointers of the state's element type, so a
    float32 weight under a float64 state, or a host tensor, must be refused
    before any launch (torch.nn.Conv1d raises on such a mix as well)."""
    require_gpu(ref, *tensors)
    for t in tensors:
        if t is not None and t.dtype != ref.dtype:
            raise TypeError(f"{what} has dtype {t.dtype} but the input is {ref.dtype}; "
                            "convert the model or the state so that they match")


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
