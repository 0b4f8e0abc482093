"""ctypes binding of ``libtt_amd.so`` (the C ABI declared in include/tt.h).

There is no fallback: if the library is missing or no GPU is visible, every
compute entry point raises.  ``load()`` only needs the shared object and does
not touch the GPU, so CPU-only tests can check the exported symbols.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtt_amd.so")

TT_OK, TT_ERR_INVALID, TT_ERR_NOMEM, TT_ERR_DEVICE, TT_ERR_UNSUPPORTED = range(5)
(TT_KERNEL_SINO_T05, TT_KERNEL_SINO_SORT, TT_KERNEL_CIRCUS, TT_KERNEL_AUX,
 TT_KERNEL_BATCH, TT_KERNEL_GATHER, TT_KERNEL_CIRCUS_VJP,
 TT_KERNEL_BATCH_BACKWARD, TT_KERNEL_NOS, TT_KERNEL_P3_DFT,
 TT_KERNEL_RECORDS) = range(11)

# Every symbol include/tt.h declares; tests check the library exports them.
EXPORTED = (
    "tt_context_create", "tt_context_destroy", "tt_context_set_stream",
    "tt_context_synchronize", "tt_last_error", "tt_pad_geometry", "tt_pad",
    "tt_angle_count", "tt_sinograms", "tt_sinograms_host", "tt_circus",
    "tt_circus_host", "tt_profile_enable", "tt_profile_read",
    "tt_forward_batched", "tt_context_set_autotune", "tt_batched_launch_shape",
    "tt_circus_backward", "tt_sinograms_backward", "tt_backward",
    "tt_resize", "tt_zscore", "tt_nearest_orthonormal_sinogram",
    "tt_nearest_orthonormal_sinograms", "tt_backward_batched", "tt_saved_bytes",
    "tt_circus_noalloc", "tt_saved_offsets", "tt_context_set_tiles", "tt_batched_tiles",
    # test hooks (include/tt_testing.h)
    "tt_testing_fail_allocations", "tt_testing_short_sqrt", "tt_testing_nos_product",
)


class TTError(RuntimeError):
    """A tt_* entry point returned a status other than TT_OK."""

    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} failed (status {status}): {msg}")
        self.status = status


class tt_pfunc(C.Structure):
    _fields_ = [("code", C.c_int32), ("order", C.c_int32), ("center", C.c_int32)]


class tt_image(C.Structure):
    _fields_ = [("d_image", C.c_void_p), ("n", C.c_int32), ("d_sino", C.c_void_p),
                ("d_circus", C.c_void_p), ("d_saved", C.c_void_p)]


class tt_image_grad(C.Structure):
    _fields_ = [("d_image", C.c_void_p), ("n", C.c_int32), ("d_dcircus", C.c_void_p),
                ("d_dimage", C.c_void_p), ("d_sino", C.c_void_p), ("d_saved", C.c_void_p)]


class tt_tiles(C.Structure):
    _fields_ = [("line_angles", C.c_int32), ("line_columns", C.c_int32),
                ("gather_tile", C.c_int32), ("gather_waves", C.c_int32)]


class tt_sino_aux(C.Structure):
    _fields_ = [("median_x", C.c_void_p), ("median_sqrt", C.c_void_p),
                ("t6_rank", C.c_void_p), ("t6_src", C.c_void_p),
                ("t7_rank", C.c_void_p), ("t7_src", C.c_void_p)]


_lib = None
_lock = threading.Lock()


def load():
    """Load libtt_amd.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m tracetransform_amd.build` (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp, i32, f32 = C.c_void_p, C.c_int32, C.c_float
        ip = C.POINTER(C.c_int)
        L.tt_last_error.restype = C.c_char_p
        L.tt_last_error.argtypes = []
        L.tt_context_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.tt_context_destroy.argtypes = [vp]
        L.tt_context_set_stream.argtypes = [vp, vp]
        L.tt_context_synchronize.argtypes = [vp]
        L.tt_pad_geometry.argtypes = [C.c_int, C.c_int, ip, ip, ip]
        L.tt_pad.argtypes = [vp, vp, C.c_int, C.c_int, vp]
        L.tt_angle_count.argtypes = [f32]
        L.tt_angle_count.restype = C.c_int
        L.tt_sinograms.argtypes = [vp, vp, C.c_int, C.c_int, f32,
                                   C.POINTER(i32), C.c_int, vp,
                                   C.POINTER(tt_sino_aux)]
        L.tt_sinograms_host.argtypes = [vp, vp, C.c_int, C.c_int, f32,
                                        C.POINTER(i32), C.c_int, vp]
        L.tt_circus.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int,
                                C.POINTER(tt_pfunc), C.c_int, vp, vp]
        L.tt_circus_host.argtypes = [vp, vp, C.c_int, C.c_int,
                                     C.POINTER(tt_pfunc), C.c_int, vp]
        L.tt_circus_noalloc.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(tt_pfunc), C.c_int, vp, vp]
        L.tt_forward_batched.argtypes = [vp, C.POINTER(tt_image), C.c_int, C.c_int,
                                         f32, C.POINTER(i32), C.c_int,
                                         C.POINTER(tt_pfunc), C.c_int]
        L.tt_context_set_autotune.argtypes = [vp, C.c_int]
        L.tt_batched_launch_shape.argtypes = [vp, ip, ip]
        L.tt_context_set_tiles.argtypes = [vp, C.POINTER(tt_tiles)]
        L.tt_batched_tiles.argtypes = [vp, C.POINTER(tt_tiles)]
        L.tt_circus_backward.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(tt_pfunc), C.c_int, vp, vp]
        L.tt_sinograms_backward.argtypes = [vp, vp, C.c_int, C.c_int, f32,
                                            C.POINTER(i32), C.c_int, vp, vp]
        L.tt_backward.argtypes = [vp, vp, C.c_int, C.c_int, f32, C.POINTER(i32),
                                  C.c_int, C.POINTER(tt_pfunc), C.c_int, vp, vp]
        L.tt_resize.argtypes = [vp, vp, C.c_int, C.c_int, vp, C.c_int, C.c_int]
        L.tt_zscore.argtypes = [vp, vp, C.c_int, C.c_int, vp]
        L.tt_nearest_orthonormal_sinogram.argtypes = [vp, vp, C.c_int, C.c_int, vp,
                                                      C.c_int, ip, ip]
        L.tt_backward_batched.argtypes = [vp, C.POINTER(tt_image_grad), C.c_int,
                                          C.c_int, f32, C.POINTER(i32), C.c_int,
                                          C.POINTER(tt_pfunc), C.c_int]
        L.tt_saved_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
        L.tt_saved_bytes.restype = C.c_size_t
        szp = C.POINTER(C.c_size_t)
        L.tt_saved_offsets.argtypes = [C.c_int, C.c_int, C.c_int, szp, szp, szp, szp, szp]
        L.tt_testing_fail_allocations.argtypes = [C.c_longlong]
        L.tt_testing_short_sqrt.argtypes = [vp, C.c_uint32, C.c_uint64,
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
        L.tt_testing_nos_product.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_int, vp, vp, vp, C.POINTER(C.c_double), vp,
                                             vp]
        L.tt_profile_enable.argtypes = [vp, C.c_int]
        L.tt_profile_read.argtypes = [vp, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_int64)]
        for name in EXPORTED:
            if name not in ("tt_last_error", "tt_angle_count", "tt_saved_bytes"):
                getattr(L, name).restype = C.c_int
        _lib = L
        return L


def check(fn: str, status: int) -> None:
    if status != TT_OK:
        msg = load().tt_last_error().decode(errors="replace")
        if status == TT_ERR_INVALID:
            raise ValueError(f"{fn}: {msg}")
        raise TTError(fn, status, msg)
