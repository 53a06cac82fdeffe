"""Loader for libsubway_amd.so and the ctypes signatures of its C ABI.

The library is built in-tree (``make -C subway_amd``) and loaded from
``subway_amd/lib``; a missing library raises ImportError -- there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SUBWAY_AMD_LIB points at another build of the same library (e.g. an
# out-of-tree build directory); the default is the in-tree build.
LIB_PATH = os.environ.get("SUBWAY_AMD_LIB") or os.path.join(HERE, "lib", "libsubway_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `make -C subway_amd` "
                      "(or __graft_entry__.build())")

lib = ctypes.CDLL(LIB_PATH)


class SubwayError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


def check(rc: int) -> None:
    if rc != 0:
        raise SubwayError(rc, lib.subway_last_error().decode(errors="replace"))


_vp = ctypes.c_void_p
_i32, _u32, _u64 = ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64
_u8p, _u32p, _u64p = ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)

_SIGS = {
    "subway_version": (_i32, []),
    "subway_last_error": (ctypes.c_char_p, []),
    "subway_default_options": (None, [_vp]),
    "subway_device_count": (_i32, []),
    "subway_shutdown": (None, []),
    "subway_struct_sizes": (None, [_u32p, _u32p, _u32p]),
    "subway_load": (_i32, [ctypes.c_char_p, _i32, ctypes.POINTER(_vp)]),
    "subway_graph_from_csr": (_i32, [_u64, _u64, _u64p, _u32p, _u32p, ctypes.POINTER(_vp)]),
    "subway_graph_rmat": (_i32, [_i32, _i32, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u64, _i32,
                                 _i32, _u32, _u32, ctypes.POINTER(_vp)]),
    "subway_graph_rmat_device": (_i32, [_i32, _i32, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u64,
                                        _i32, _i32, _u32, _u32, ctypes.POINTER(_vp)]),
    "subway_graph_split_view": (_i32, [_vp, ctypes.POINTER(_u64p), ctypes.POINTER(_u32p), ctypes.POINTER(_u8p)]),
    "subway_graph_reorder": (_i32, [_vp, _i32]),
    "subway_graph_order": (_i32, [_vp, ctypes.POINTER(_u32p), ctypes.POINTER(_u32p)]),
    "subway_graph_build_blocks": (_i32, [_vp, _i32]),
    "subway_graph_blocks_view": (_i32, [_vp, ctypes.POINTER(_u64p), ctypes.POINTER(_u32p), _u64p, _u32p]),
    "subway_graph_split_tail": (_i32, [_vp, _u64p]),
    "subway_graph_save": (_i32, [_vp, ctypes.c_char_p]),
    "subway_graph_info": (_i32, [_vp, _u64p, _u64p, _i32p, _u32p, _u32p]),
    "subway_graph_export": (_i32, [_vp, _u64p, _u32p, _u32p]),
    "subway_graph_free": (None, [_vp]),
    "subway_run": (_i32, [_vp, _vp, _vp, _vp, _vp, _u32]),
    "subway_session_create": (_i32, [_vp, _vp, ctypes.POINTER(_vp)]),
    "subway_session_run": (_i32, [_vp, _u32, _vp, _vp, _vp, _u32]),
    "subway_session_values": (_vp, [_vp]),
    "subway_session_set_policy": (_i32, [_vp, _vp]),
    "subway_session_free": (None, [_vp]),
    "subway_session_run_batch": (_i32, [_vp, _u32p, _i32, _u32p, _vp, _vp, _u32]),
    "subway_session_tune_batch": (_i32, [_vp, _u32p, _i32, _i32p, ctypes.POINTER(ctypes.c_double)]),
    "subway_generate": (_i32, [_u64, _u8p, _u8p, _u32p, _u32p, _u64p, _u64p]),
    "subway_generate_noalloc": (_i32, [_u64, _vp, _vp, _vp, _vp, _vp, _u64p]),
    "subway_generate_pr": (_i32, [_u64, _f32p, ctypes.c_float, _u32p, _u32p, _u64p, _u64p]),
    "subway_partition": (_i32, [_u64, _u64p, _u32p, _u64, _u64, _u64p, _u64p]),
    "subway_gather": (_i32, [_vp, _u64, _u32p, _vp, _u64p]),
    "subway_h2d_peak": (_i32, [_u64, _i32, ctypes.POINTER(ctypes.c_double)]),
    "subway_probe_copy": (_i32, [_u64, ctypes.POINTER(ctypes.c_double)]),
    "subway_mlp_forward": (_i32, [_f32p, _f32p, _u64, _f32p]),
    "subway_mlp_grad": (_i32, [_f32p, _f32p, _f32p, _f32p, _u64, _f32p, _f32p]),
    "subway_mlp_train": (_i32, [_f32p, _f32p, _f32p, _f32p, _u64, _i32, ctypes.c_float, _f32p]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_SIGS)
