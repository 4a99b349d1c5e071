"""ctypes binding of libmeamd.so (the C ABI declared in include/motion_estimation_amd.h).

The library is the only compute path: if it is missing or a GPU is absent
the calls raise; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmeamd.so")

_lock = threading.Lock()
_lib = None

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_longlong
dp = C.c_void_p  # device pointers are passed as void*

# name -> argtypes (restype int unless noted)
_SIGS = {
    "me_version": [],
    "me_rigid_transform": [vp, i64, i32, vp, i64, vp, i64, i32, vp, vp, dp, dp, dp, dp, i64, dp, i32, i32, dp, vp],
    "me_rigid_deriv": [vp, i64, vp, i64, vp, i64, i32, vp, vp, dp, dp, dp, dp, i64, dp, i32, vp],
    "me_rigid_deriv_chain": [vp, i64, vp, vp, vp, i64, vp, i64, vp, i64, i32, vp, vp, dp, dp, dp, dp, i64, dp, i32,
                             i32, vp],
    "me_fft_axis": [vp, i64, vp, i64, i32, vp, vp, i32, i32, vp],
    "me_encode_forward": [vp, i64, vp, i32, i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32, vp],
    "me_encode_forward_static": [vp, i64, vp, i32, i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, i32, vp],
    "me_encode_adjoint": [vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, i64, vp],
    "me_encode_tuning_set": [vp, i32, i32, i32],
    "me_tune_encode_forward": [vp, i64, vp, i32, i32, vp, vp, vp, vp, vp, i32, vp, i32, vp, vp, vp],
    "me_permute_shift": [vp, vp, vp, vp, vp, vp, vp, C.c_float, vp],
    "me_coil_ssq": [vp, i32, i64, C.c_float, C.c_float, vp, vp, vp],
    "me_gather_hybrid": [vp, i32, vp, vp, vp, vp, vp, i32, i64, C.c_float, vp, vp],
    "me_max_abs": [vp, i64, vp, vp],
    "me_scatter_hybrid": [vp, i32, vp, vp, vp, vp, vp, i32, i64, vp, vp],
    "me_dot_re": [vp, vp, vp, i64, vp, vp, vp],
    "me_cg_init": [vp, vp, vp, vp, vp, i64, vp, vp, vp],
    "me_cg_step": [vp, vp, vp, vp, vp, i64, C.c_double, vp, vp, vp, vp, i64, vp],
    "me_csub": [vp, vp, vp, i64, vp],
    "me_dot_re_batched": [vp, i64, vp, i64, i64, i32, vp, vp, i64, vp],
    "me_lm_ga_accumulate": [vp, i64, i64, i32, vp, i32, C.c_double, i32, vp, i64, vp],
    "me_lm_ga_solve": [vp, vp, vp, vp, i32, vp, vp, i32, vp],
    "me_sum_batch": [vp, i64, i32, i64, vp, i32, vp],
    "me_axpy_real": [vp, C.c_double, vp, i64, vp],
    "me_coil_gram": [vp, i32, i64, vp, C.c_double, vp, vp, vp],
    "me_whitening_matrix": [vp, i32, C.c_double, vp, vp],
    "me_hermitian_eig": [vp, i32, i32, vp, vp, vp],
    "me_coil_mix": [vp, i32, i32, vp, i64, vp, vp],
    "me_espirit_calibration_matrix": [vp, vp, vp, vp, vp],
    "me_line_tuning_set": [vp, vp, i32, i32],
    "me_abs_diff_norm": [vp, vp, i64, vp, vp, vp, vp],
    "me_gradient_max": [vp, vp, vp, vp],
    "me_gradient_entropy": [vp, vp, vp, vp, vp, vp],
    "me_tune_line_passes": [vp, vp, i64, i32, vp, vp, i32, vp, vp, vp],
    "me_graph_begin": [vp],
    "me_graph_end": [vp, vp],
    "me_graph_abort": [vp],
    "me_graph_launch": [vp, vp],
    "me_graph_destroy": [vp],
    "me_gram_rows": [vp, i32, i64, vp, vp],
    "me_jacobi_eig": [vp, vp, i32, i32, C.c_double, vp, vp, vp, vp],
    "me_espirit_projector": [vp, i32, i32, vp, i32, vp, vp, vp],
    "me_zgemm": [i32, i32, i32, i32, i32, vp, vp, vp, vp],
    "me_random_normal": [vp, i64, C.c_ulonglong, vp],
    "me_scale_columns": [vp, i64, i32, vp, vp],
    "me_espirit_maps": [vp, i32, vp, vp, C.c_double, i32, vp, vp, vp],
    "me_rigid_deriv_apply": [vp, i64, vp, i64, vp, vp, i64, vp, i64, vp, i64, i32, vp, vp, dp, dp, dp, dp, i64, dp,
                             i32, i32, vp],
    "me_rigid_deriv_apply_dot": [vp, i64, vp, i64, vp, vp, i64, vp, i64, vp, i64, vp, i64, i32, vp, vp, dp, dp, dp,
                                 dp, i64, dp, i32, i32, vp, vp, i32, vp],
    "me_rigid_apply_chain": [vp, i64, vp, vp, i64, vp, i64, i32, vp, vp, dp, dp, dp, dp, i64, dp, vp],
    "me_mask_mul": [vp, vp, i64, vp],
    "me_lm_gram": [vp, vp, i64, vp, vp, i32, i32, vp, vp, vp],
    "me_lm_solve": [vp, vp, vp, i32, vp, vp, i32, vp],
    "me_lm_accept": [vp, vp, vp, i32, vp, vp, vp, vp, C.c_double, i32, vp],
    "me_lm_mean": [vp, i32, i32, vp, vp],
    "me_disorder_order": [i32, i32, i32, i32, i32, C.c_ulonglong, vp, vp, i64, vp, vp, vp],
}
# entry points returning sizes rather than a status
_SIZES = {
    "me_encode_partials_size": ([i32, i32, i32, i32], C.c_longlong),
    "me_vec_partials_size": ([], C.c_int),
    "me_lm_partials_size": ([i32], C.c_longlong),
    "me_encode_coil_groups": ([i32], C.c_int),
    "me_dot_batched_partials_size": ([i32], C.c_longlong),
    "me_coil_gram_partials_size": ([i32], C.c_longlong),
    "me_espirit_hankel_size": ([vp, vp], C.c_longlong),
    "me_line_tuning_get": ([vp, vp, i32], C.c_int),
    "me_encode_tuning_get": ([vp, i32, i32], C.c_int),
    "me_metric_partials_size": ([], C.c_int),
    "me_jacobi_workspace_size": ([i32], C.c_longlong),
    "me_rigid_dot_partials_size": ([i32, vp, vp], C.c_longlong),
    "me_jacobi_set_lds_rows": ([i32], C.c_int),
    "me_espirit_maps_workspace_size": ([i32, vp, vp], C.c_longlong),
}
_OPTIONAL = {}
# entry points returning nothing
_VOID = {"me_line_tuning_clear": [], "me_encode_tuning_clear": []}


def declared_symbols():
    """Names of every entry point the Python layer binds."""
    return sorted(["me_last_error", *_SIGS.keys(), *_SIZES.keys(), *_OPTIONAL.keys(), *_VOID.keys()])


def load():
    """Load libmeamd.so once; raises ImportError if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
        lib = C.CDLL(LIB_PATH)
        lib.me_last_error.restype = C.c_char_p
        lib.me_last_error.argtypes = []
        for name, args in {**_SIGS, **_OPTIONAL}.items():
            fn = getattr(lib, name)
            fn.restype = C.c_int
            fn.argtypes = args
        for name, args in _VOID.items():
            fn = getattr(lib, name)
            fn.restype = None
            fn.argtypes = args
        for name, (args, res) in _SIZES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def call(name, *args):
    """Invoke an entry point; non-zero status raises RuntimeError(message)."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.me_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({rc}): {msg}")


def size(name, *args):
    """Value of a size-query entry point."""
    return getattr(load(), name)(*args)


def stream(device=None):
    """Current HIP stream handle (torch's current stream on the device)."""
    return vp(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    """Device (or host) address of a tensor, None for None."""
    if t is None:
        return None
    return vp(t.data_ptr())


def int3(vals):
    return (C.c_int * 3)(*[int(v) for v in vals])


def ll3(vals):
    return (C.c_longlong * 3)(*[int(v) for v in vals])


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("motion_estimation_amd needs a HIP device (MI355X); none is visible")
    return torch.device("cuda", torch.cuda.current_device())
