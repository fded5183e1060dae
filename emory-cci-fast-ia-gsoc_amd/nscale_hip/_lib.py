"""ctypes binding of libnscale_hip.so (C ABI in include/nscale_hip.h)."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libnscale_hip.so"

NSX_OK = 0
NSX_ERR_INVALID = 1
NSX_ERR_UNSUPPORTED = 2
NSX_ERR_HIP = 3
NSX_ERR_NOMEM = 4
NSX_ERR_BUDGET = 5
NSX_ERR_NOT_CONVERGED = 6

NSX_U8, NSX_I32, NSX_F32 = 0, 1, 2
NSX_TUNE_VISIT_SWEEPS = 0
NSX_TUNE_LEVELSET_LEVELS = 1
NSX_TUNE_ENGINE = 2
NSX_TUNE_FILL_SCANS = 3
NSX_TUNE_DRAIN = 4
NSX_TUNE_DT_SOLVE = 5
NSX_ENGINE_ROUNDS = 0
NSX_ENGINE_TREE = 1
NSX_ENGINE_AUTO = 2
NSX_ENGINE_PLANES = 3
NSX_OPT_I32_INVERT = 0
NSX_I32_INVERT_WRAP = 0
NSX_I32_INVERT_SATURATE = 1
NSX_OPT_DT_ORDER = 1
NSX_DT_ORDER_QUEUE = 0
NSX_DT_ORDER_TILES = 1
NSX_MEM_DEVICE, NSX_MEM_HOST = 0, 1


class NsxImg(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("pitch_bytes", C.c_int64),
        ("rows", C.c_int32),
        ("cols", C.c_int32),
        ("dtype", C.c_int32),
        ("mem", C.c_int32),
    ]


P = C.POINTER
_IMG = P(NsxImg)
_CTX = C.c_void_p
_I32P = P(C.c_int32)

# name -> (restype, argtypes); every function declared in include/nscale_hip.h
SIGNATURES = {
    "nsx_ctx_create": (C.c_int, [P(C.c_void_p), C.c_int, C.c_size_t]),
    "nsx_ctx_destroy": (None, [_CTX]),
    "nsx_last_error": (C.c_char_p, [_CTX]),
    "nsx_ctx_synchronize": (C.c_int, [_CTX]),
    "nsx_ctx_workspace_bytes": (C.c_size_t, [_CTX]),
    "nsx_last_recon_stats": (C.c_int, [_CTX, C.c_int, P(C.c_int64)]),
    "nsx_last_round_log": (C.c_int, [_CTX, C.c_int, P(C.c_int32), C.c_int, P(C.c_int)]),
    "nsx_ctx_set_timing": (C.c_int, [_CTX, C.c_int]),
    "nsx_ctx_read_timing": (C.c_int, [_CTX, P(C.c_double), P(C.c_int64), C.c_int]),
    "nsx_ctx_read_tile_counts": (C.c_int, [_CTX, P(C.c_int64), P(C.c_int64), C.c_int]),
    "nsx_ctx_set_tuning": (C.c_int, [_CTX, C.c_int, C.c_int]),
    "nsx_ctx_get_tuning": (C.c_int, [_CTX, C.c_int, P(C.c_int)]),
    "nsx_ctx_drain_failures": (C.c_int, [_CTX, P(C.c_int64)]),
    "nsx_ctx_dt_resumes": (C.c_int, [_CTX, P(C.c_int64)]),
    "nsx_image_check": (C.c_int, [P(NsxImg), P(C.c_char_p)]),
    "nsx_ctx_set_option": (C.c_int, [_CTX, C.c_int, C.c_int]),
    "nsx_ctx_get_option": (C.c_int, [_CTX, C.c_int, P(C.c_int)]),
    "nsx_imreconstruct": (C.c_int, [_CTX, _IMG, _IMG, _IMG, C.c_int, C.c_void_p]),
    "nsx_imreconstruct_noalloc": (C.c_int, [_CTX, _IMG, _IMG, _IMG, C.c_int, C.c_void_p]),
    "nsx_imreconstruct_binary": (C.c_int, [_CTX, _IMG, _IMG, _IMG, C.c_int, C.c_void_p]),
    "nsx_fill_holes": (C.c_int, [_CTX, _IMG, _IMG, C.c_int, C.c_int, C.c_void_p]),
    "nsx_imfill": (C.c_int, [_CTX, _IMG, _IMG, _IMG, C.c_int, C.c_int, C.c_void_p]),
    "nsx_bwselect": (C.c_int, [_CTX, _IMG, _IMG, _IMG, C.c_int, C.c_void_p]),
    "nsx_imhmin": (C.c_int, [_CTX, _IMG, C.c_double, _IMG, C.c_int, C.c_void_p]),
    "nsx_bounding_box": (C.c_int, [_CTX, _IMG, C.c_int, _I32P, C.c_int, _I32P, C.c_void_p]),
    "nsx_local_maxima": (C.c_int, [_CTX, _IMG, _IMG, C.c_int, C.c_void_p]),
    "nsx_local_minima": (C.c_int, [_CTX, _IMG, _IMG, C.c_int, C.c_void_p]),
    "nsx_distance_transform": (C.c_int, [_CTX, _IMG, _IMG, C.c_int, C.c_void_p]),
    "nsx_bwlabel": (C.c_int, [_CTX, _IMG, _IMG, C.c_int, C.c_int, _I32P, C.c_void_p]),
    "nsx_bwareaopen": (C.c_int, [_CTX, _IMG, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _IMG, _I32P,
                                 C.c_void_p]),
    "nsx_relabel": (C.c_int, [_CTX, _IMG, C.c_int, _I32P, C.c_void_p]),
    "nsx_threshold_gt": (C.c_int, [_CTX, _IMG, C.c_double, _IMG, C.c_void_p]),
    "nsx_segment_chain": (C.c_int, [_CTX, _IMG, _IMG, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                    _IMG, _IMG, _IMG, _I32P, C.c_void_p]),
    "nsx_recon_label": (C.c_int, [_CTX, _IMG, _IMG, C.c_int, C.c_double, C.c_int, _IMG, _IMG, _I32P,
                                  C.c_void_p]),
    "nsx_batch_create": (C.c_int, [_CTX, P(C.c_void_p)]),
    "nsx_batch_add": (C.c_int, [C.c_void_p, _IMG, _IMG, _IMG, _IMG, C.c_int, C.c_double, C.c_int]),
    "nsx_batch_finalize": (C.c_int, [C.c_void_p, C.c_int]),
    "nsx_batch_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nsx_batch_counts": (C.c_int, [C.c_void_p, _I32P, C.c_int]),
    "nsx_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, _I32P, C.c_int]),
    "nsx_batch_get_tuning": (C.c_int, [C.c_void_p, C.c_int, _I32P, C.c_int]),
    "nsx_batch_set_drain": (C.c_int, [C.c_void_p, C.c_int]),
    "nsx_batch_set_overlap": (C.c_int, [C.c_void_p, C.c_int]),
    "nsx_batch_context": (C.c_void_p, [C.c_void_p]),
    "nsx_batch_destroy": (None, [C.c_void_p]),
    "nsx_server_create": (C.c_int, [C.c_int, C.c_size_t, C.c_int, P(C.c_void_p)]),
    "nsx_server_submit": (C.c_int, [C.c_void_p, _IMG, _IMG, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                    C.c_int, _IMG, _IMG, _IMG, P(C.c_uint64)]),
    "nsx_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, _I32P, P(C.c_double)]),
    "nsx_server_stats": (C.c_int, [C.c_void_p, P(C.c_int64)]),
    "nsx_server_set_sharing": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_int64]),
    "nsx_server_batch_stats": (C.c_int, [C.c_void_p, P(C.c_int64)]),
    "nsx_server_context": (C.c_void_p, [C.c_void_p, C.c_int]),
    "nsx_server_last_error": (C.c_char_p, [C.c_void_p]),
    "nsx_server_destroy": (None, [C.c_void_p]),
    "nsx_splitmix64_next": (C.c_uint64, [P(C.c_uint64)]),
    "nsx_gen_disks_u8": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_long, C.c_int, C.c_uint64]),
    "nsx_gen_marker_u8": (None, [C.c_void_p, C.c_void_p, C.c_long, C.c_int]),
    "nsx_gen_scale_i32": (None, [C.c_void_p, C.c_void_p, C.c_long, C.c_int]),
    "nsx_gen_scale_f32": (None, [C.c_void_p, C.c_void_p, C.c_long]),
    "nsx_gen_maze_u8": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint64]),
    "nsx_version": (C.c_char_p, []),
}

_lib = None


def load() -> C.CDLL:
    """Loads the in-tree library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
