"""Loader for the HIP library ``lib/libfl_correction_amd.so`` (the C ABI of
``include/fl_correction_amd.h``).

There is no fallback: if the library has not been built, or no HIP device is
present when an op runs, the op raises.  PyTorch only provides device memory
and the current stream; every computation runs in the library's kernels.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libfl_correction_amd.so")

_lock = threading.Lock()
_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_SZ = ctypes.c_size_t
_F = ctypes.c_float
_D = ctypes.c_double
_INT = ctypes.c_int

# name -> (restype, argtypes); must mirror include/fl_correction_amd.h
SIGNATURES = {
    "fl_last_error": (ctypes.c_char_p, []),
    "fl_version": (_INT, []),
    "fl_forward_emission_batch": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    "fl_backward_emission_batch": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    "fl_forward_emission": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _P]),
    "fl_backward_emission": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _P]),
    "fl_huber_grad_batch": (_INT, [_P, _P, _I64, _I64, _I64, _I64, _F, _P]),
    "fl_mlem_batch": (_INT, [_P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _D, _D, _P, _SZ, _P]),
    "fl_mlem_workspace_bytes": (_SZ, [_I64, _I64, _I64, _I64, _I64]),
    "fl_mlem_batch_noalloc": (_INT, [_P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _D, _D, _P]),
    "fl_slice_major": (_INT, [_P, _P, _I64, _I64, _I64, _P]),
    "fl_backproject_sm_scratch_bytes": (_SZ, [_I64, _I64, _I64, _I64]),
    "fl_backward_emission_sm": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _INT, _P, _SZ, _P]),
    "fl_mlem_ratio_sm": (_INT, [_P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _P]),
    "fl_mlem_sensitivity_sm": (_INT, [_P, _P, _P, _P, _I64, _I64, _I64, _I64, _P, _SZ, _P]),
    "fl_mlem_update_sm": (_INT, [_P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _D, _D, _P, _SZ, _P]),
    "fl_mlem_workspace_bytes_sm": (_SZ, [_I64, _I64, _I64, _I64, _I64]),
    "fl_mlem_sensitivity": (_INT, [_P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    "fl_mlem_ratio": (_INT, [_P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    "fl_mlem_update": (_INT, [_P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _D, _D, _P]),
    "fl_cal_atten": (_INT, [_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _P]),
    "fl_rotate": (_INT, [_P, _P, _I64, _I64, _I64, _F, _F, _P]),
    "fl_incident_atten": (_INT, [_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _P]),
    "fl_mu_xrf_rot": (_INT, [_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _INT, _P]),
    "fl_cone_atten": (_INT, [_P, _P, _I64, _I64, _I64, _I64, _P, _P, _P, _P, _I64, _P]),
    "fl_xrf_combine": (_INT, [_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _INT, _P]),
    "fl_cal_atten_adjoint": (_INT, [_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _P]),
    "fl_cone_atten_adjoint": (_INT, [_P, _P, _I64, _I64, _I64, _I64, _P, _P, _P, _P, _I64, _P]),
    "fl_backward_emission_accumulate": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    "fl_sample_grad": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _P]),
    "fl_xrf_grad_prep": (_INT, [_P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _INT, _P]),
    "fl_incident_grad": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _P]),
    "fl_mu_adjoint": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _INT, _P]),
    "fl_rotate_adjoint": (_INT, [_P, _P, _I64, _I64, _I64, _F, _F, _INT, _P]),
    "fl_mlem_multi_workspace_bytes": (_SZ, [_P, _I64]),
    "fl_mlem_multi_workspace_bytes_sm": (_SZ, [_P, _I64]),
    "fl_mlem_multi_create": (_INT, [_P, _I64, _D, _D, _P, _SZ, _P, ctypes.POINTER(_P)]),
    "fl_mlem_multi_run": (_INT, [_P, _I64, _INT, _P]),
    "fl_mlem_multi_destroy": (_INT, [_P]),
    "fl_cone_sweep_workspace_bytes": (_SZ, [_I64, _I64, _I64, _I64, _P, _P, _P, _I64]),
    "fl_cone_sweep": (_INT, [_P, _P, _I64, _I64, _I64, _I64, _P, _P, _P, _P, _I64, _INT, _P, _SZ, _P]),
    "fl_frac_sum": (_INT, [_P, _P, _I64, _I64, _INT, _P]),
    "fl_adaptive_mask": (_INT, [_P, _P, _I64, _I64, _I64, _INT, _INT, _INT, _INT, _P, _P, _P]),
    "fl_frac_ratio": (_INT, [_P, _P, _P, _I64, _I64, _INT, _INT, _P]),
    "fl_frac_final": (_INT, [_P, _P, _P, _P, _P, _I64, _I64, _INT, _P]),
}


class MLEMProblem(ctypes.Structure):
    """fl_mlem_problem of include/fl_correction_amd.h."""
    _fields_ = [
        ("atten", _P), ("em_cs", _P), ("theta", _P), ("I", _P), ("C", _P),
        ("A", _I64), ("B", _I64), ("H", _I64), ("W", _I64),
        ("fwd_angles", ctypes.c_int32), ("bwd_rows", ctypes.c_int32),
    ]


class FLError(RuntimeError):
    """A call into libfl_correction_amd returned an error code."""


def load():
    """Load the library once and bind every exported entry point."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m fl_correction_amd.build` "
                "(hipcc, gfx950). There is no CPU fallback.")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def call(name, *args):
    """Invoke entry point `name`; raise FLError on a non-zero return code."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.fl_last_error().decode(errors="replace")
        raise FLError(f"{name} failed (code {rc}): {msg}")
    return rc


def ptr(t):
    """Device address of a tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(device):
    """hipStream_t handle of torch's current stream on `device`."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
