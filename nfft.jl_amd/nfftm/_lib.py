"""ctypes binding of libnfftm.so (include/nfftm.h).

The library is loaded from this package directory only; there is no CPU
fallback, so importing the host API without the built library raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libnfftm.so")

OK, ERR_ARG, ERR_DIM, ERR_UNSUPPORTED, ERR_HIP, ERR_FFT, ERR_OOM = range(7)
F32, F64, MIXED = 0, 1, 2
FULL, TENSOR, LINEAR, POLYNOMIAL = 1, 2, 3, 4
KAISER_BESSEL = 0
FLAG_BLOCKING, FLAG_TIMING, FLAG_GRAPHS, FLAG_ROCFFT, FLAG_SEPARATE_FOLD, FLAG_SPARSE = 1, 2, 4, 8, 16, 32
BATCH_AUTOTUNE, BATCH_NO_GRAPH = 1, 2
R32, R64, C32, C64, R16, C16 = range(6)  # element types of the typed stage operators

# every symbol include/nfftm.h declares
EXPORTS = (
    "nfftm_plan_create", "nfftm_plan_create_dims", "nfftm_nodes", "nfftm_forward", "nfftm_adjoint",
    "nfftm_deconvolve", "nfftm_deconvolve_transpose", "nfftm_convolve",
    "nfftm_convolve_transpose", "nfftm_convolve_typed", "nfftm_convolve_transpose_typed", "nfftm_fft", "nfftm_sizes", "nfftm_grid",
    "nfftm_timing", "nfftm_set_timing", "nfftm_profile", "nfftm_profile_read",
    "nfftm_set_stream", "nfftm_plan_copy", "nfftm_binning", "nfftm_grad_nodes", "nfftm_grad_nodes_grid",
    "nfftm_plan_tune",
    "nfftm_tile_shape", "nfftm_plan_set_tiles", "nfftm_batch_create", "nfftm_batch_info", "nfftm_batch_forward",
    "nfftm_batch_adjoint", "nfftm_batch_destroy", "nfftm_sdc", "nfftm_sdc_noalloc", "nfftm_toeplitz_create",
    "nfftm_toeplitz_kernel", "nfftm_toeplitz_apply", "nfftm_toeplitz_destroy",
    "nfftm_plan_destroy", "nfftm_last_error", "nfftm_version", "nfftm_lds_issue_cycles",
)

_lib = None


class LibraryMissing(ImportError):
    pass


def load():
    """Load libnfftm.so and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise LibraryMissing(
            f"{LIBPATH} is not built; run __graft_entry__.build() (make -C nfft.jl_amd/csrc)")
    L = ctypes.CDLL(LIBPATH)
    vp, i, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    pi64 = ctypes.POINTER(ctypes.c_int64)
    pd = ctypes.POINTER(ctypes.c_double)
    L.nfftm_plan_create.argtypes = [ctypes.POINTER(vp), i, i, pi64, i64, pd, i, d, i, i, i, vp]
    L.nfftm_plan_create_dims.argtypes = [ctypes.POINTER(vp), i, i, pi64, i, i, i64, pd, i, d, i, i, i, vp]
    L.nfftm_nodes.argtypes = [vp, pd, i64]
    L.nfftm_grad_nodes.argtypes = [vp, vp, vp, vp]
    L.nfftm_grad_nodes_grid.argtypes = [vp, vp, vp, vp]
    L.nfftm_plan_tune.argtypes = [vp]
    L.nfftm_binning.argtypes = [vp, ctypes.POINTER(i), vp, vp]
    L.nfftm_plan_set_tiles.argtypes = [vp, i, i]
    L.nfftm_tile_shape.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i)]
    L.nfftm_batch_create.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(vp), ctypes.POINTER(i), i]
    L.nfftm_batch_info.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i)]
    L.nfftm_batch_forward.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), vp]
    L.nfftm_batch_adjoint.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), vp]
    L.nfftm_batch_destroy.argtypes = [vp]
    L.nfftm_batch_destroy.restype = None
    L.nfftm_sdc.argtypes = [vp, i, vp, i]
    L.nfftm_lds_issue_cycles.argtypes = [pd, pd, ctypes.POINTER(i)]
    L.nfftm_sdc_noalloc.argtypes = [vp, i, vp, i]
    L.nfftm_toeplitz_create.argtypes = [ctypes.POINTER(vp), i, i, pi64, vp]
    L.nfftm_toeplitz_kernel.argtypes = [vp, vp, vp]
    L.nfftm_toeplitz_apply.argtypes = [vp, vp, vp]
    L.nfftm_toeplitz_destroy.argtypes = [vp]
    L.nfftm_toeplitz_destroy.restype = None
    for name in ("nfftm_forward", "nfftm_adjoint", "nfftm_deconvolve",
                 "nfftm_deconvolve_transpose", "nfftm_convolve", "nfftm_convolve_transpose"):
        getattr(L, name).argtypes = [vp, vp, vp]
    L.nfftm_convolve_typed.argtypes = [vp, vp, i, i64, vp, i, i64]
    L.nfftm_convolve_transpose_typed.argtypes = [vp, vp, i, i64, vp, i, i64]
    L.nfftm_fft.argtypes = [vp, vp, i]
    L.nfftm_sizes.argtypes = [vp, pi64, pi64, pd]
    L.nfftm_grid.argtypes = [vp, ctypes.POINTER(vp)]
    L.nfftm_timing.argtypes = [vp, pd]
    L.nfftm_set_timing.argtypes = [vp, i]
    L.nfftm_profile.argtypes = [vp, i]
    L.nfftm_profile_read.argtypes = [vp, pd, pi64]
    L.nfftm_set_stream.argtypes = [vp, vp]
    L.nfftm_plan_copy.argtypes = [vp, ctypes.POINTER(vp)]
    L.nfftm_plan_destroy.argtypes = [vp]
    L.nfftm_plan_destroy.restype = None
    L.nfftm_last_error.restype = ctypes.c_char_p
    L.nfftm_last_error.argtypes = []
    L.nfftm_version.restype = ctypes.c_char_p
    L.nfftm_version.argtypes = []
    for name in EXPORTS:
        if name not in ("nfftm_plan_destroy", "nfftm_last_error", "nfftm_version",
                        "nfftm_batch_destroy", "nfftm_toeplitz_destroy"):
            getattr(L, name).restype = i
    _lib = L
    return L
