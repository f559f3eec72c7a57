"""ctypes binding of libfkdtree_hip.so (the C ABI of include/fkdtree/fkd_capi.h).

Loading fails loudly when the library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfkdtree_hip.so")

P = C.c_void_p
U32 = C.c_uint32
U64 = C.c_uint64
I32 = C.c_int
SZ = C.c_size_t

# name: (restype, argtypes)
_SIGS = {
    "fkd_last_error": (C.c_char_p, []),
    "fkd_create": (I32, [I32, U32, P, P, C.POINTER(P)]),
    "fkd_create_events": (I32, [I32, U32, P, P, P, C.POINTER(P)]),
    "fkd_create_device": (I32, [I32, U32, P, P, P, P, C.POINTER(P)]),
    "fkd_destroy": (None, [P]),
    "fkd_destroy_async": (I32, [P, P]),
    "fkd_build": (I32, [P, P]),
    "fkd_get_tree": (I32, [P, P, P, P]),
    "fkd_tree_device": (I32, [P, C.POINTER(P), C.POINTER(P)]),
    "fkd_tree_info": (I32, [P, C.POINTER(I32), C.POINTER(U32), C.POINTER(U64)]),
    "fkd_trim": (I32, [P, P]),
    "fkd_boxes_around": (I32, [P, C.c_float, P, P, P, P]),
    "fkd_serialize": (I32, [P, P, U64, C.POINTER(U64), P]),
    "fkd_deserialize": (I32, [P, U64, I32, P, C.POINTER(P)]),
    "fkd_set_tuning": (I32, [P, C.c_char_p, U32]),
    "fkd_get_tuning": (I32, [P, C.c_char_p, C.POINTER(U32)]),
    "fkd_device_bytes": (I32, [P, C.POINTER(U64)]),
    "fkd_search_count": (I32, [P, U32, P, P, P, P, P]),
    "fkd_search_fill": (I32, [P, U32, P, P, P, P, P, U64, P]),
    "fkd_search_stats": (I32, [P, U32, P, P, P, P, P, P]),
    "fkd_search": (I32, [P, U32, P, P, P, P, P, U64, P]),
    "fkd_search_noalloc": (I32, [P, U32, P, P, P, P, P, U64, P]),
    "fkd_search_small": (I32, [P, U32, P, U32, P, P, P]),
    "fkd_box_server_create": (I32, [P, U32, U32, U32, C.POINTER(P)]),
    "fkd_box_server_search": (I32, [P, P, P, U32, C.POINTER(U32)]),
    "fkd_box_server_stats": (I32, [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(I32)]),
    "fkd_box_server_timing": (I32, [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(U64)]),
    "fkd_box_server_destroy": (None, [P]),
    "fkd_search_status": (I32, [P, P, C.POINTER(U64), C.POINTER(I32)]),
    "fkd_trace_events": (I32, [P, P, P]),
    "fkd_graph_create": (I32, [P, U32, P, P, P, P, P, U64, P, C.POINTER(P)]),
    "fkd_graph_launch": (I32, [P, P]),
    "fkd_graph_destroy": (None, [P]),
    "fkd_search_host": (I32, [P, U32, P, P, P, P, P]),
    "fkd_fetch_results": (I32, [P, P, P]),
    "fkd_device_info": (I32, [C.c_char_p, SZ, C.POINTER(I32), C.POINTER(SZ), C.POINTER(SZ)]),
    "fkd_malloc": (I32, [C.POINTER(P), SZ]),
    "fkd_free": (I32, [P]),
    "fkd_malloc_async": (I32, [C.POINTER(P), SZ, P]),
    "fkd_free_async": (I32, [P, P]),
    "fkd_pool_info": (I32, [C.POINTER(U64), C.POINTER(U64)]),
    "fkd_release_cached": (I32, []),
    "fkd_host_alloc": (I32, [C.POINTER(P), SZ]),
    "fkd_host_free": (I32, [P]),
    "fkd_memcpy_htod": (I32, [P, P, SZ, P]),
    "fkd_memcpy_dtoh": (I32, [P, P, SZ, P]),
    "fkd_copy_u32": (I32, [P, P, U64, P]),
    "fkd_gather_u64": (I32, [P, P, P, U64, P]),
    "fkd_stream_create": (I32, [C.POINTER(P)]),
    "fkd_stream_destroy": (I32, [P]),
    "fkd_stream_sync": (I32, [P]),
    "fkd_event_create": (I32, [C.POINTER(P)]),
    "fkd_event_destroy": (I32, [P]),
    "fkd_event_record": (I32, [P, P]),
    "fkd_event_elapsed_ms": (I32, [P, P, C.POINTER(C.c_float)]),
    "fkd_event_query": (I32, [P]),
    "fkd_shape_subtree": (I32, [U32, U32, U32, C.POINTER(U32), C.POINTER(U32)]),
    "fkd_shape_locate": (I32, [U32, U32, U32, C.POINTER(U32)]),
}

EXPORTS = tuple(_SIGS)

_lib = None


class FKDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(python -m fkdtree_amd.build_lib)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise FKDError(rc, lib().fkd_last_error().decode(errors="replace"))
