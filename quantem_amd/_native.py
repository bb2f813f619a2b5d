"""ctypes binding of the C-ABI library ``libqem_ptycho.so`` (declared in ``include/qem_ptycho.h``).

The library must have been built (``quantem_amd.build.build()``); there is no fallback: every
compute entry point of this package runs the HIP kernels or raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

from .build import LIB_PATH

QEM_OK = 0
QEM_ERR_ARG = 1
QEM_ERR_NULL = 2
QEM_ERR_ALIGN = 3
QEM_ERR_WORKSPACE = 4
QEM_ERR_LAUNCH = 5
QEM_ERR_UNSUPPORTED = 6

QEM_FLAG_STORE_WAVES = 1 << 0
QEM_FLAG_GRAD_POSITIONS = 1 << 1
QEM_FLAG_GRAD_DESCAN = 1 << 2
QEM_FLAG_ANALYTIC = 1 << 3
QEM_FLAG_GRAD_PROPAGATORS = 1 << 4
QEM_FLAG_FULL_PROPAGATORS = 1 << 5

QEM_STATIC_WORKSPACE_BYTES = 64 << 20  # workspace of qem_ptycho_forward_loss_static

LOSS_TYPES = {
    "l2_amplitude": 0,
    "l1_amplitude": 1,
    "l2_intensity": 2,
    "l1_intensity": 3,
    "poisson": 4,
}

SUPPORTED_ROI = (32, 64, 128, 256)  # square pattern sizes of the fixed-size multislice kernels
GENERIC_ROI_MAX = 512  # generic-size kernels: Nr x Nc, 2..512 per axis, prime factors 2, 3, 5, 7


def roi_supported(nr: int, nc: int) -> bool:
    """Whether the library runs Nr x Nc patterns (the fixed-size square kernels or the generic ones)."""
    if nr == nc and nr in SUPPORTED_ROI:
        return True

    def smooth(n: int) -> bool:
        if n < 2 or n > GENERIC_ROI_MAX:
            return False
        for r in (2, 3, 5, 7):
            while n % r == 0:
                n //= r
        return n == 1

    return smooth(nr) and smooth(nc)
FFT_ROI = (32, 64, 128)  # sizes of the standalone fft2 entry point

EXPORTED_SYMBOLS = (
    "qem_ptycho_workspace_bytes",
    "qem_ptycho_forward_loss",
    "qem_ptycho_forward_loss_static",
    "qem_ptycho_adjoint",
    "qem_ptycho_projection_update",
    "qem_ptycho_propagator_gradient",
    "qem_fft2_c64",
    "qem_last_error",
    "qem_abi_version",
    "qem_profile_begin",
    "qem_profile_end",
    "qem_pattern_features",
    "qem_conv3x3_forward",
    "qem_conv3x3_wgrad_workspace_bytes",
    "qem_conv3x3_wgrad",
    "qem_adam_step_f32",
    "qem_adamw_step_f32",
    "qem_sgd_step_f32",
    "qem_object_constraints_workspace_bytes",
    "qem_object_constraints_forward",
    "qem_object_constraints_backward",
    "qem_pattern_com",
    "qem_probe_orthogonalize_workspace_bytes",
    "qem_probe_orthogonalize_forward",
    "qem_probe_orthogonalize_backward",
    "qem_center_amplitudes_workspace_bytes",
    "qem_center_amplitudes",
)


class QemDesc(ctypes.Structure):
    _fields_ = [
        ("num_slices", ctypes.c_int32),
        ("num_modes", ctypes.c_int32),
        ("batch", ctypes.c_int32),
        ("roi", ctypes.c_int32),
        ("obj_rows", ctypes.c_int32),
        ("obj_cols", ctypes.c_int32),
        ("num_gpts", ctypes.c_int32),
        ("loss_type", ctypes.c_int32),
        ("mean_intensity", ctypes.c_float),
        ("flags", ctypes.c_uint32),
        ("segment_slices", ctypes.c_int32),
        ("roi_cols", ctypes.c_int32),
    ]


_lib = None
_lock = threading.Lock()

_vp = ctypes.c_void_p
_fp = ctypes.c_void_p  # float* passed as raw device address


def _declare(lib: ctypes.CDLL) -> None:
    dp = ctypes.POINTER(QemDesc)
    lib.qem_ptycho_workspace_bytes.argtypes = [dp, ctypes.POINTER(ctypes.c_size_t)]
    lib.qem_ptycho_workspace_bytes.restype = ctypes.c_int
    lib.qem_ptycho_forward_loss.argtypes = [dp] + [_vp] * 10 + [ctypes.c_size_t, _vp]
    lib.qem_ptycho_forward_loss.restype = ctypes.c_int
    lib.qem_ptycho_forward_loss_static.argtypes = [dp] + [_vp] * 10
    lib.qem_ptycho_forward_loss_static.restype = ctypes.c_int
    lib.qem_ptycho_adjoint.argtypes = [dp] + [_vp] * 11 + [ctypes.c_size_t, _vp]
    lib.qem_ptycho_adjoint.restype = ctypes.c_int
    lib.qem_ptycho_projection_update.argtypes = [dp] + [_vp] * 10 + [ctypes.c_size_t, _vp]
    lib.qem_ptycho_projection_update.restype = ctypes.c_int
    lib.qem_ptycho_propagator_gradient.argtypes = [dp, _vp, _vp, ctypes.c_size_t, _vp]
    lib.qem_ptycho_propagator_gradient.restype = ctypes.c_int
    lib.qem_fft2_c64.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp]
    lib.qem_fft2_c64.restype = ctypes.c_int
    lib.qem_last_error.argtypes = []
    lib.qem_last_error.restype = ctypes.c_char_p
    lib.qem_profile_begin.argtypes = [ctypes.c_int32]
    lib.qem_profile_begin.restype = ctypes.c_int
    lib.qem_profile_end.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32)]
    lib.qem_profile_end.restype = ctypes.c_int
    i32 = ctypes.c_int32
    lib.qem_pattern_features.argtypes = [i32, i32, i32, _vp, _vp, _vp, _vp]
    lib.qem_pattern_features.restype = ctypes.c_int
    lib.qem_conv3x3_forward.argtypes = [i32] * 5 + [_vp, _vp, _vp, i32, i32, _vp, _vp]
    lib.qem_conv3x3_forward.restype = ctypes.c_int
    lib.qem_conv3x3_wgrad_workspace_bytes.argtypes = [i32] * 5 + [ctypes.POINTER(ctypes.c_size_t)]
    lib.qem_conv3x3_wgrad_workspace_bytes.restype = ctypes.c_int
    lib.qem_conv3x3_wgrad.argtypes = [i32] * 5 + [_vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]
    lib.qem_conv3x3_wgrad.restype = ctypes.c_int
    lib.qem_adam_step_f32.argtypes = [ctypes.c_int64, _vp, _vp, _vp, _vp] + [ctypes.c_double] * 6 + [i32, _vp]
    lib.qem_adam_step_f32.restype = ctypes.c_int
    lib.qem_adamw_step_f32.argtypes = lib.qem_adam_step_f32.argtypes
    lib.qem_adamw_step_f32.restype = ctypes.c_int
    lib.qem_sgd_step_f32.argtypes = [ctypes.c_int64, _vp, _vp, _vp] + [ctypes.c_double] * 4 + [i32, ctypes.c_double,
                                                                                              i32, _vp]
    lib.qem_sgd_step_f32.restype = ctypes.c_int
    lib.qem_object_constraints_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]
    lib.qem_object_constraints_workspace_bytes.restype = ctypes.c_int
    lib.qem_object_constraints_forward.argtypes = [ctypes.c_int64, i32, _vp, _vp, _vp, ctypes.c_size_t, _vp]
    lib.qem_object_constraints_forward.restype = ctypes.c_int
    lib.qem_object_constraints_backward.argtypes = [ctypes.c_int64, i32, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]
    lib.qem_object_constraints_backward.restype = ctypes.c_int
    lib.qem_pattern_com.argtypes = [i32, i32, _vp, _vp, _vp, _vp]
    lib.qem_pattern_com.restype = ctypes.c_int
    lib.qem_center_amplitudes_workspace_bytes.argtypes = [i32, i32, ctypes.POINTER(ctypes.c_size_t)]
    lib.qem_center_amplitudes_workspace_bytes.restype = ctypes.c_int
    lib.qem_center_amplitudes.argtypes = [i32, i32, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]
    lib.qem_center_amplitudes.restype = ctypes.c_int
    i64 = ctypes.c_int64
    lib.qem_probe_orthogonalize_workspace_bytes.argtypes = [i32, i64, ctypes.POINTER(ctypes.c_size_t)]
    lib.qem_probe_orthogonalize_workspace_bytes.restype = ctypes.c_int
    lib.qem_probe_orthogonalize_forward.argtypes = [i32, i64, _vp, _vp, _vp, ctypes.c_size_t, _vp]
    lib.qem_probe_orthogonalize_forward.restype = ctypes.c_int
    lib.qem_probe_orthogonalize_backward.argtypes = [i32, i64, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]
    lib.qem_probe_orthogonalize_backward.restype = ctypes.c_int
    lib.qem_abi_version.argtypes = []
    lib.qem_abi_version.restype = ctypes.c_int32


def lib_path() -> str:
    return os.environ.get("QEM_PTYCHO_LIB", LIB_PATH)


def load() -> ctypes.CDLL:
    """Loads the HIP library once; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            path = lib_path()
            if not os.path.exists(path):
                raise RuntimeError(
                    f"HIP library {path} is missing: build it with `python -m quantem_amd.build` "
                    "(this package has no CPU fallback)"
                )
            lib = ctypes.CDLL(path)
            _declare(lib)
            _lib = lib
        return _lib


def last_error() -> str:
    msg = load().qem_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    """Raises the exception class the reference raises for the same kind of failure."""
    if rc == QEM_OK:
        return
    msg = f"{what}: {last_error()} (code {rc})"
    if rc in (QEM_ERR_ARG, QEM_ERR_NULL, QEM_ERR_ALIGN, QEM_ERR_WORKSPACE, QEM_ERR_UNSUPPORTED):
        raise ValueError(msg)
    raise RuntimeError(msg)


class KernelTimer:
    """Context manager timing the fused forward / adjoint kernels with HIP events (library hook)."""

    def __init__(self, max_launches: int = 4096):
        self.max_launches = max_launches
        self.fwd_ms = self.adj_ms = 0.0
        self.n_fwd = self.n_adj = 0

    def __enter__(self):
        check(load().qem_profile_begin(self.max_launches), "qem_profile_begin")
        return self

    def __exit__(self, *exc):
        f, a = ctypes.c_float(0), ctypes.c_float(0)
        nf, na = ctypes.c_int32(0), ctypes.c_int32(0)
        check(load().qem_profile_end(ctypes.byref(f), ctypes.byref(nf), ctypes.byref(a), ctypes.byref(na)),
              "qem_profile_end")
        self.fwd_ms, self.adj_ms = float(f.value), float(a.value)
        self.n_fwd, self.n_adj = int(nf.value), int(na.value)
        return False
