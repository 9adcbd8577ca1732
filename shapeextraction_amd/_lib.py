"""ctypes binding of ``libshapes_hip.so`` (declarations mirror include/shapes_c.h).

This module only loads the library and declares signatures; it never falls
back to another implementation.  A missing library raises ``ImportError``
with the build instruction.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libshapes_hip.so")
HEADER_PATH = Path(__file__).resolve().parent.parent / "include" / "shapes_c.h"

SX_OK, SX_EINVAL, SX_ENOMEM, SX_EHIP, SX_ERANGE, SX_EDATA, SX_ESTATE, SX_EINTERNAL = range(8)
SX_ROOT_TEXTBOOK, SX_ROOT_REFERENCE = 0, 1
SX_KNN_PLAIN, SX_KNN_RADIUS = 0, 1
SX_PREC_FP32, SX_PREC_FP64 = 0, 1
SX_KNN_ALGO_TILES, SX_KNN_ALGO_LANE = 0, 1
SX_SORT_ONESWEEP, SX_SORT_THREE_KERNEL, SX_SORT_HYBRID = 0, 1, 2

STATUS_NAMES = {
    SX_OK: "SX_OK",
    SX_EINVAL: "SX_EINVAL",
    SX_ENOMEM: "SX_ENOMEM",
    SX_EHIP: "SX_EHIP",
    SX_ERANGE: "SX_ERANGE",
    SX_EDATA: "SX_EDATA",
    SX_ESTATE: "SX_ESTATE",
    SX_EINTERNAL: "SX_EINTERNAL",
}


class SxError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


class OTNode(C.Structure):
    """OT::OTNode (src/Octree.hpp:19-35), 56 bytes."""

    _fields_ = [
        ("children", C.c_int32 * 8),
        ("corner", C.c_float * 3),
        ("cell_size", C.c_float),
        ("child_node_mask", C.c_int32),
        ("child_leaf_mask", C.c_int32),
    ]


class IndexView(C.Structure):
    _fields_ = [
        ("n_input", C.c_int64),
        ("n_unique", C.c_int64),
        ("n_oct", C.c_int64),
        ("min_coord", C.c_float),
        ("max_coord", C.c_float),
        ("code_bits", C.c_int),
        ("codes", C.c_void_p),
        ("has_leaf_left", C.c_void_p),
        ("has_leaf_right", C.c_void_p),
        ("prefix_n", C.c_void_p),
        ("left_child", C.c_void_p),
        ("parent", C.c_void_p),
        ("oct_nodes", C.c_void_p),
        ("points", C.c_void_p),
        ("sorted_perm", C.c_void_p),
        ("rep", C.c_void_p),
        ("unique_of_input", C.c_void_p),
        ("d_counts", C.c_void_p),
        ("d_bounds", C.c_void_p),
    ]


class BatchView(C.Structure):
    _fields_ = [
        ("n_clouds", C.c_int32),
        ("n_input", C.c_int64),
        ("n_slots", C.c_int64),
        ("n_unique", C.c_int64),
        ("cloud_offsets", C.c_void_p),
        ("cloud_base", C.c_void_p),
        ("cloud_size", C.c_void_p),
        ("cloud_bounds", C.c_void_p),
        ("codes", C.c_void_p),
        ("rep", C.c_void_p),
        ("points", C.c_void_p),
        ("prefix_n", C.c_void_p),
        ("left_child", C.c_void_p),
        ("parent", C.c_void_p),
        ("n_oct", C.c_int64),
        ("oct_nodes", C.c_void_p),
        ("cloud_oct_base", C.c_void_p),
        ("cloud_oct_size", C.c_void_p),
        ("has_leaf_left", C.c_void_p),
        ("has_leaf_right", C.c_void_p),
    ]


P = C.c_void_p
I = C.c_int
I64 = C.c_int64
SZ = C.c_size_t
F = C.c_float
U64 = C.c_uint64
PSZ = C.POINTER(C.c_size_t)

# name -> (restype, argtypes)
SIGNATURES = {
    "sx_last_error": (C.c_char_p, []),
    "sx_abi_version": (I, []),
    "sx_device_count": (I, [C.POINTER(C.c_int)]),
    "sx_set_device": (I, [I]),
    "sx_device_name": (I, [I, C.c_char_p, SZ]),
    "sx_malloc": (I, [C.POINTER(C.c_void_p), SZ]),
    "sx_free": (I, [P]),
    "sx_host_malloc": (I, [C.POINTER(C.c_void_p), SZ]),
    "sx_host_free": (I, [P]),
    "sx_memcpy_htod": (I, [P, P, SZ, P]),
    "sx_memcpy_dtoh": (I, [P, P, SZ, P]),
    "sx_memcpy_dtod": (I, [P, P, SZ, P]),
    "sx_fill_u8": (I, [P, C.c_uint8, SZ, P]),
    "sx_stream_create": (I, [C.POINTER(C.c_void_p)]),
    "sx_stream_destroy": (I, [P]),
    "sx_stream_sync": (I, [P]),
    "sx_device_sync": (I, []),
    "sx_event_create": (I, [C.POINTER(C.c_void_p)]),
    "sx_event_destroy": (I, [P]),
    "sx_event_record": (I, [P, P]),
    "sx_event_elapsed_ms": (I, [P, P, C.POINTER(C.c_float)]),
    "sx_fnv64": (U64, [P, SZ]),
    "sx_gen_sphere": (I, [U64, I64, F, P, P, P]),
    "sx_gen_scene": (I, [U64, I64, I, F, F, F, P, P, P]),
    "sx_gen_gmm_scene": (I, [U64, I64, I, I, F, F, F, P, P, P, C.POINTER(C.c_float)]),
    "sx_gen_batch_sizes": (I, [U64, I, I64, I64, P]),
    "sx_gen_degenerate": (I, [U64, I64, I, F, F, P, P, P, P, P]),
    "sx_bbox": (I, [P, P, P, I64, P, P, P, PSZ, P]),
    "sx_morton_encode": (I, [P, P, P, I64, P, I, P, P]),
    "sx_morton_decode": (I, [P, P, I64, P, I, P, P]),
    "sx_sort_pairs": (I, [P, P, I64, I, P, P, P, PSZ, P]),
    "sx_sort_check": (I, [P, I64, P]),
    "sx_sort_set_spin_limit": (I, [C.c_uint32]),
    "sx_unique": (I, [P, P, I64, P, P, P, P, P, PSZ, P]),
    "sx_build_radix_tree": (I, [P, P, I64, I, I, P, P, P, P, P, P, P, P]),
    "sx_build_octree": (I, [P, P, I64, I, P, P, P, P, P, P, P, I64, P, P, PSZ, P]),
    "sx_check_counts": (I, [P, I64, P, I64, P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), P]),
    "sx_index_create": (I, [C.POINTER(C.c_void_p), I64, I, I, F]),
    "sx_index_destroy": (I, [P]),
    "sx_index_build": (I, [P, P, P, P, I64, P]),
    "sx_index_build_codes": (I, [P, P, I64, F, F, P]),
    "sx_index_get_view": (I, [P, C.POINTER(IndexView), P]),
    "sx_index_to_input_order": (I, [P, P, I64, P, P]),
    "sx_index_set_bucket": (I, [P, I]),
    "sx_index_set_algorithms": (I, [P, I, I]),
    "sx_index_knn_reruns": (I, [P, C.POINTER(C.c_int64), I, P]),
    "sx_index_set_knn_stats": (I, [P, I]),
    "sx_index_knn_candidates": (I, [P, C.POINTER(C.c_int64), I, P]),
    "sx_knn": (I, [P, I64, I64, I, I, P, P, P]),
    "sx_knn_points": (I, [P, P, I64, I, P, P]),
    "sx_normals": (I, [P, P, I64, I64, I, I, F, F, F, P, P, P]),
    "sx_estimate_normals": (I, [P, I64, I64, I, I, P, I, P, P, P]),
    "sx_estimate_normals_noalloc": (I, [P, I64, I64, I, I, P, I, P, P]),
    "sx_mlp_create": (I, [C.POINTER(C.c_void_p), I64]),
    "sx_mlp_destroy": (I, [P]),
    "sx_mlp_num_params": (I, []),
    "sx_mlp_set_params": (I, [P, P, P]),
    "sx_mlp_get_params": (I, [P, P, P]),
    "sx_mlp_init": (I, [P, U64, P]),
    "sx_mlp_set_normalizer": (I, [P, P, P, F, P]),
    "sx_mlp_forward": (I, [P, P, I64, P, P]),
    "sx_mlp_predict_radius2": (I, [P, P, I64, P, P]),
    "sx_mlp_backward": (I, [P, P, P, I64, P, P, P]),
    "sx_mlp_adam_step": (I, [P, P, P, I64, F, F, P, P]),
    "sx_radius_features": (I, [P, I64, I64, P, P]),
    "sx_radius_predict": (I, [P, P, I64, I64, P, P]),
    # batched clouds
    "sx_batch_create": (I, [C.POINTER(P), I64, I, I]),
    "sx_batch_destroy": (I, [P]),
    "sx_batch_set_clouds": (I, [P, C.POINTER(C.c_int64), I, P]),
    "sx_batch_set_tuning": (I, [P, I, I, I]),
    "sx_batch_set_octree": (I, [P, F]),
    "sx_batch_build": (I, [P, P, P, P, P]),
    "sx_batch_get_view": (I, [P, C.POINTER(BatchView), P]),
    "sx_batch_knn": (I, [P, I, P, I64, P]),
    "sx_batch_knn_radius": (I, [P, I, P, P, I64, P]),
    "sx_batch_radius_features": (I, [P, P, I64, P]),
    "sx_batch_radius_predict": (I, [P, P, P, I64, P]),
    "sx_batch_normals": (I, [P, P, I, I, F, F, F, P, P, I64, P]),
    "sx_batch_deferred": (I, [P, C.POINTER(C.c_int64), I, P]),
    "sx_batch_graph_create": (I, [P, P, P, P, I, P, P, I64, C.POINTER(P), P]),
    "sx_batch_graph_launch": (I, [P, P]),
    "sx_batch_graph_destroy": (I, [P]),
    "sx_batch_autotune": (I, [P, I, P, I64, C.POINTER(C.c_int), I, C.POINTER(C.c_int), I, I, C.POINTER(C.c_float),
                              C.POINTER(C.c_int), C.POINTER(C.c_int), P]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (once).  Raises ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing; build it with `python -c \"import __graft_entry__ as g; g.build()\"`"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != SX_OK:
        msg = load().sx_last_error()
        raise SxError(rc, (what + ": " if what else "") + (msg.decode() if msg else ""))


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise SxError on failure."""
    check(getattr(load(), name)(*args), name)
