"""ctypes binding of librxhip.so (declared in include/rxhip.h).

The library is built in-tree by renderx_amd/build.py.  Loading fails loudly if
it is missing: there is no fallback implementation of any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librxhip.so")
# RENDERX_AMD_LIB selects another build of the same library (e.g. one compiled
# with a different -DRX_LDS_DEPTH) without touching the in-tree one.
LIB_PATH = os.environ.get("RENDERX_AMD_LIB", LIB_PATH)

RX_OK = 0
STATUS = {0: "RX_OK", -1: "RX_ERR_INVALID", -2: "RX_ERR_HIP", -3: "RX_ERR_NO_MEMORY", -4: "RX_ERR_STATE",
          -5: "RX_ERR_NO_DEVICE", -6: "RX_ERR_OVERFLOW"}

# every symbol include/rxhip.h declares, with (restype, argtypes)
P = C.c_void_p
U32, I32, U64, F32 = C.c_uint32, C.c_int32, C.c_uint64, C.c_float
SIGNATURES = {
    "rx_tree_build": (C.c_int, [U32, P, U32, P, U32, C.POINTER(P)]),
    "rx_tree_free": (None, [P]),
    "rx_tree_info": (C.c_int, [P, P, P, P]),
    "rx_tree_serialize": (C.c_int, [P, P, P, P, P, P, P]),
    "rx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "rx_destroy": (None, [P]),
    "rx_last_error": (C.c_char_p, [P]),
    "rx_synchronize": (C.c_int, [P]),
    "rx_debug_fail_device_alloc": (C.c_int, [I32]),
    "rx_load_tree": (C.c_int, [P, U32, P, P, P, U32, P, P, P]),
    "rx_load_legality": (C.c_int, [P, U32, P, P]),
    "rx_intersect": (C.c_int, [P, U32, P, P, P, P, P, U32, P, P, P, P]),
    "rx_intersect_dev": (C.c_int, [P, U32, P, P, P, P, P, U32, P, P, P, P]),
    "rx_intersect_dev_noalloc": (C.c_int, [P, U32, P, P, P, P, P, U32, P, P, P, P]),
    "rx_last_kernel_ms": (C.c_int, [P, C.POINTER(F32)]),
    "rx_copy_kernel": (C.c_int, [P, P, P, U64]),
    "rx_load_scene": (C.c_int, [P, U32, P, U32, P, U32, P, U32]),
    "rx_render": (C.c_int, [P, P, P, P, P, P, P]),
    "rx_set_render_tuning": (C.c_int, [P, P]),
    "rx_get_render_tuning": (C.c_int, [P, P]),
    "rx_autotune_render": (C.c_int, [P, P, P, P, P]),
    "rx_batch_create": (C.c_int, [P, C.POINTER(P)]),
    "rx_batch_destroy": (None, [P]),
    "rx_batch_add_scene": (C.c_int, [P, U32, P, U32, P, U32, P, U32, P, P, P]),
    "rx_batch_render": (C.c_int, [P, P]),
    "rx_batch_autotune": (C.c_int, [P, U32, P, P]),
    "rx_batch_get_image": (C.c_int, [P, U32, P, P, P]),
    "rx_save_pfm": (C.c_int, [C.c_char_p, I32, I32, P]),
    "rx_save_pfm_snapshot": (C.c_int, [C.c_char_p, U32, U32, I32, I32, P]),
    "rx_read_pfm": (C.c_int, [C.c_char_p, C.POINTER(I32), C.POINTER(I32), P, I32]),
    "rx_image_rmse": (C.c_int, [I32, I32, P, P, P, C.POINTER(C.c_double)]),
    "rx_image_relmse": (C.c_int, [I32, I32, P, P, F32, C.POINTER(C.c_double)]),
    "rx_scene_load": (C.c_int, [C.c_char_p, C.POINTER(P), C.c_char_p, U32]),
    "rx_scene_file_free": (None, [P]),
    "rx_scene_file_sizes": (C.c_int, [P, P, P, P]),
    "rx_scene_file_get": (C.c_int, [P, P, P, P, P, P, C.c_char_p, U32, C.POINTER(C.c_double)]),
    "rx_guide_init": (C.c_int, [P, P]),
    "rx_guide_set_bounds": (C.c_int, [P, P, P]),
    "rx_guide_set_weights": (C.c_int, [P, P]),
    "rx_guide_get_weights": (C.c_int, [P, P]),
    "rx_guide_eval": (C.c_int, [P, U32, P, P, P]),
    "rx_guide_eval_mixture": (C.c_int, [P, U32, P, P, P]),
    "rx_guide_train_batch": (C.c_int, [P, U32, P, P, P, P, P, P, P]),
}
GUIDE_PARAMS = 17648

_lib = None


class RxError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RxError(f"{LIB_PATH} is not built: run renderx_amd/build.py (no fallback path exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(ctx, code, what=""):
    if code != RX_OK:
        msg = ""
        if ctx:
            m = lib().rx_last_error(ctx)
            msg = m.decode() if m else ""
        raise RxError(f"{what}: {STATUS.get(code, code)} {msg}")


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
