"""ctypes binding of libglass_hip.so (include/glass_hip.h).

The library is built in-tree by pyglass_amd/build.py.  Loading fails loudly
when it is missing: there is no CPU fallback behind this API.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libglass_hip.so")

# Every symbol declared in include/glass_hip.h, with (restype, argtypes).
_vp = C.c_void_p
_i32 = C.c_int32
_pi32 = C.POINTER(C.c_int32)
_pf32 = C.POINTER(C.c_float)
SIGNATURES = {
    "glass_hip_last_error": (C.c_char_p, []),
    "glass_hip_version": (C.c_char_p, []),
    "glass_hip_device_count": (C.c_int, []),
    "glass_hip_graph_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "glass_hip_graph_save": (C.c_int, [_vp, C.c_char_p]),
    "glass_hip_graph_create": (C.c_int, [_i32, _i32, _vp, _i32, _vp, C.POINTER(_vp)]),
    "glass_hip_graph_set_upper": (C.c_int, [_vp, _i32, _i32, _vp, _vp]),
    "glass_hip_graph_info": (C.c_int, [_vp, _pi32, _pi32, _pi32, _pi32]),
    "glass_hip_graph_copy": (C.c_int, [_vp, _vp, _vp]),
    "glass_hip_graph_free": (None, [_vp]),
    "glass_hip_build_graph": (C.c_int, [C.c_char_p, _vp, _i32, _i32, C.c_char_p, _i32, _i32, _i32, C.POINTER(_vp),
                                        C.POINTER(C.c_double)]),
    "glass_hip_searcher_create": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "glass_hip_searcher_set_data": (C.c_int, [_vp, _vp, _i32, _i32]),
    "glass_hip_searcher_set_ef": (C.c_int, [_vp, _i32]),
    "glass_hip_searcher_get_ef": (_i32, [_vp]),
    "glass_hip_searcher_search": (C.c_int, [_vp, _vp, _i32, _vp, _vp]),
    "glass_hip_searcher_search_batch": (C.c_int, [_vp, _vp, _i32, _i32, _vp, _vp]),
    "glass_hip_searcher_optimize": (C.c_int, [_vp, _i32]),
    "glass_hip_searcher_last_avg_dist_cmps": (C.c_double, [_vp]),
    "glass_hip_searcher_free": (None, [_vp]),
    "glass_hip_searcher_search_batch_efk": (C.c_int, [_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    "glass_hip_searcher_search_batch_dev": (C.c_int, [_vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "glass_hip_searcher_search_batch_dev_noalloc": (C.c_int, [_vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "glass_hip_searcher_search_batch_efk_dev": (C.c_int, [_vp, _vp, _i32, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _vp,
                                                          _vp]),
    "glass_hip_searcher_set_graph_capture": (C.c_int, [_vp, _i32]),
    "glass_hip_searcher_set_refine_factor": (C.c_int, [_vp, C.c_float]),
    "glass_hip_searcher_set_launch": (C.c_int, [_vp, _i32, _i32, _i32]),
    "glass_hip_searcher_set_visited_capacity": (C.c_int, [_vp, _i32]),
    "glass_hip_searcher_set_query_waves": (C.c_int, [_vp, _i32]),
    "glass_hip_searcher_get_query_waves": (C.c_int, [_vp, _pi32, _pi32]),
    "glass_hip_searcher_set_gather": (C.c_int, [_vp, _i32]),
    "glass_hip_searcher_get_launch": (C.c_int, [_vp, _pi32, _pi32, _pi32, _pi32]),
    "glass_hip_searcher_kernel_info": (C.c_int, [_vp, _pi32, _pi32, _pi32]),
    "glass_hip_device_info": (C.c_int, [C.c_char_p, _i32, _pi32, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _pi32]),
    "glass_hip_searcher_info": (C.c_int, [_vp, _pi32, _pi32, _pi32, _pi32, _pf32]),
    "glass_hip_searcher_get_codes": (C.c_int, [_vp, _vp]),
    "glass_hip_searcher_get_calibration": (C.c_int, [_vp, _vp, _pi32]),
    "glass_hip_searcher_sync": (C.c_int, [_vp]),
    "glass_hip_quant_train": (C.c_int, [C.c_char_p, C.c_char_p, _vp, _i32, _i32, _vp, _pi32]),
    "glass_hip_clustering_create": (C.c_int, [_i32, _i32, C.c_char_p, C.POINTER(_vp)]),
    "glass_hip_clustering_free": (None, [_vp]),
    "glass_hip_clustering_fit": (C.c_int, [_vp, _vp, _i32, _i32]),
    "glass_hip_clustering_predict": (C.c_int, [_vp, _vp, _i32, _i32, _vp]),
    "glass_hip_clustering_transform": (C.c_int, [_vp, _vp, _i32, _i32]),
    "glass_hip_clustering_centroids": (C.c_int, [_vp, _vp, _pi32, _pi32]),
    "glass_hip_server_create": (C.c_int, [_vp, _i32, _i32, _i32, C.POINTER(_vp)]),
    "glass_hip_server_submit": (C.c_int, [_vp, _vp, _i32, _i32, _i32, C.POINTER(C.c_int64)]),
    "glass_hip_server_poll": (C.c_int, [_vp, C.c_int64, _vp, _vp, _pi32]),
    "glass_hip_server_wait": (C.c_int, [_vp, C.c_int64, _vp, _vp]),
    "glass_hip_server_cancel": (C.c_int, [_vp, C.c_int64]),
    "glass_hip_server_replay": (C.c_int, [_vp, _vp, C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "glass_hip_server_info": (C.c_int, [_vp, _pi32, _pi32, _pi32]),
    "glass_hip_server_stats": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "glass_hip_server_free": (None, [_vp]),
}

_lib = None


class GlassError(RuntimeError):
    pass


def load():
    """Load libglass_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m pyglass_amd.build` "
                          "(the HIP path has no CPU fallback)")
    # PyTorch-ROCm ships its own libamdhip64.  Loading torch first lets this
    # library bind to that same HIP runtime, so device pointers and streams
    # from torch tensors are valid here and torch still sees the GPUs; loaded
    # the other way round, torch would bind to the system runtime and find no
    # device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int):
    if rc != 0:
        msg = load().glass_hip_last_error().decode(errors="replace")
        raise GlassError(f"glass_hip error {rc}: {msg}")
