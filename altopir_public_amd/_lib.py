"""ctypes binding of the HIP engine library (C ABI of include/altopir_engine.h).

The library is built in-tree by `python -m altopir_public_amd.build`.  Loading fails
loudly when it is missing: there is no CPU fallback for the answer engine.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libaltopir_hip.so"

# every entry point declared in include/altopir_engine.h
SYMBOLS = (
    "sw_amx_init", "sw_alloc_zeroed", "sw_alloc", "sw_free", "sw_tile_db", "sw_pack_build", "alto_pack_build",
    "alto_engine_hint0", "alto_hint0", "alto_inspiring_offline", "alto_scheme_setup",
    "alto_engine_new", "alto_engine_free", "alto_request_len", "alto_response_len", "alto_answer",
    "alto_answer_device", "alto_answer_device_noalloc", "alto_engine_device_bytes", "alto_last_error", "alto_last_timings",
    "alto_variant_count", "alto_variant_name", "alto_tune_set", "alto_tune_get", "alto_set_graphs",
    "alto_engine_new_device", "alto_engine_new_records", "alto_context_new", "alto_context_bytes", "alto_context_free", "alto_engine_context", "alto_context_reserve",
    "alto_answer_gather", "alto_context_timings", "alto_context_stage_totals", "alto_stage_totals",
    "alto_context_set_graphs", "alto_context_graphs", "alto_set_fusion", "alto_fused", "alto_set_graph_timing",
    "matMulVecPacked", "matMulVecPacked2", "matMulVecPacked4", "matMulVecPacked8", "alto_matmul_vec_packed",
    "alto_fast_batched_dot_product",
    # serving loop (bound in serve.py)
    "alto_server_new", "alto_server_submit", "alto_server_poll", "alto_server_release", "alto_server_contexts",
    "alto_server_get_stats", "alto_server_error", "alto_server_free", "alto_server_replay", "alto_server_set_graphs",
    "alto_server_set_upload", "alto_server_upload_mode",
    # TCP front end (bound in net.py)
    "alto_net_new", "alto_net_port", "alto_net_wait", "alto_net_get_stats", "alto_net_free", "alto_net_load",
)

_lib = None

P = C.c_void_p
SZ = C.c_size_t


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"HIP engine library {LIB_PATH} is not built; run `python -m altopir_public_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    lib.sw_amx_init.restype = C.c_int
    lib.sw_amx_init.argtypes = []
    lib.sw_alloc_zeroed.restype = P
    lib.sw_alloc_zeroed.argtypes = [SZ]
    lib.sw_alloc.restype = P
    lib.sw_alloc.argtypes = [SZ]
    lib.sw_free.restype = None
    lib.sw_free.argtypes = [P]
    lib.sw_tile_db.restype = None
    lib.sw_tile_db.argtypes = [P, P, SZ, SZ, SZ, C.c_int]
    lib.sw_pack_build.restype = None
    lib.sw_pack_build.argtypes = [P, P, P, P, P, SZ, SZ, SZ, SZ, C.c_uint64, C.c_int]
    lib.alto_engine_hint0.restype = C.c_int
    lib.alto_engine_hint0.argtypes = [P, P, SZ, P]
    lib.alto_hint0.restype = C.c_int
    lib.alto_hint0.argtypes = [P, C.c_int, SZ, SZ, SZ, SZ, P, SZ, P]
    lib.alto_fast_batched_dot_product.restype = C.c_int
    lib.alto_fast_batched_dot_product.argtypes = [P, SZ, P, SZ, SZ, P, C.c_int, SZ, SZ, SZ, C.c_uint64, C.c_uint64]
    lib.alto_scheme_setup.restype = C.c_int
    lib.alto_scheme_setup.argtypes = [P, P, P, P, C.c_int, SZ, SZ, SZ, P, SZ, P, P, SZ, SZ]
    lib.alto_inspiring_offline.restype = C.c_int
    lib.alto_inspiring_offline.argtypes = [P, P, P, P, P, SZ, SZ, P, P, SZ, SZ]
    lib.alto_pack_build.restype = C.c_int
    lib.alto_pack_build.argtypes = [P, P, P, P, P, SZ, SZ, SZ, SZ, C.c_uint64, C.c_int]
    lib.alto_engine_new.restype = P
    lib.alto_engine_new.argtypes = [P, P, SZ, SZ, SZ, P, SZ, SZ, P, SZ, SZ, C.c_uint64, C.c_int, P, P,
                                    C.c_uint64, C.c_int]
    lib.alto_engine_free.restype = None
    lib.alto_engine_free.argtypes = [P]
    lib.alto_request_len.restype = SZ
    lib.alto_request_len.argtypes = [P]
    lib.alto_response_len.restype = SZ
    lib.alto_response_len.argtypes = [P]
    lib.alto_answer.restype = C.c_int
    lib.alto_answer.argtypes = [P, P, P, SZ, P, P, C.c_int, P]
    lib.alto_answer_device.restype = C.c_int
    lib.alto_answer_device.argtypes = [P, P, SZ, SZ, P, SZ, P, P]
    lib.alto_answer_device_noalloc.restype = C.c_int
    lib.alto_answer_device_noalloc.argtypes = [P, P, SZ, SZ, P, SZ, P, P]
    lib.alto_engine_device_bytes.restype = SZ
    lib.alto_engine_device_bytes.argtypes = [P]
    lib.alto_last_error.restype = C.c_char_p
    lib.alto_last_error.argtypes = []
    lib.alto_last_timings.restype = C.c_int
    lib.alto_last_timings.argtypes = [P, P, C.c_int]
    lib.alto_variant_count.restype = C.c_int
    lib.alto_variant_count.argtypes = []
    lib.alto_variant_name.restype = C.c_char_p
    lib.alto_variant_name.argtypes = [C.c_int]
    lib.alto_tune_set.restype = C.c_int
    lib.alto_tune_set.argtypes = [P, C.c_int, SZ, C.c_int]
    lib.alto_tune_get.restype = C.c_int
    lib.alto_tune_get.argtypes = [P, C.c_int, SZ]
    lib.alto_engine_new_device.restype = P
    lib.alto_engine_new_device.argtypes = [P, SZ, SZ, SZ, P, SZ, SZ, P, SZ, SZ, C.c_uint64, C.c_int, P, P,
                                           C.c_uint64, C.c_int]
    lib.alto_engine_new_records.restype = P
    lib.alto_engine_new_records.argtypes = [P, C.c_int, SZ, SZ, SZ, P, SZ, SZ, P, SZ, SZ, C.c_uint64, C.c_int, P, P,
                                            C.c_uint64, C.c_int]
    lib.alto_context_new.restype = P
    lib.alto_context_new.argtypes = [P]
    lib.alto_context_bytes.restype = SZ
    lib.alto_context_bytes.argtypes = [P, SZ]
    lib.alto_context_free.restype = None
    lib.alto_context_free.argtypes = [P]
    lib.alto_engine_context.restype = P
    lib.alto_engine_context.argtypes = [P]
    lib.alto_context_reserve.restype = C.c_int
    lib.alto_context_reserve.argtypes = [P, SZ]
    lib.alto_set_graph_timing.restype = C.c_int
    lib.alto_set_graph_timing.argtypes = [P, C.c_int]
    lib.alto_set_fusion.restype = C.c_int
    lib.alto_set_fusion.argtypes = [P, C.c_int]
    lib.alto_fused.restype = C.c_int
    lib.alto_fused.argtypes = [P, SZ]
    lib.alto_context_set_graphs.restype = C.c_int
    lib.alto_context_set_graphs.argtypes = [P, C.c_int]
    lib.alto_context_graphs.restype = C.c_int
    lib.alto_context_graphs.argtypes = [P]
    lib.alto_answer_gather.restype = C.c_int
    lib.alto_answer_gather.argtypes = [P, P, SZ, P, SZ, P, SZ, P, P, P]
    lib.alto_context_timings.restype = C.c_int
    lib.alto_context_timings.argtypes = [P, P, C.c_int]
    lib.alto_context_stage_totals.restype = C.c_int
    lib.alto_context_stage_totals.argtypes = [P, P, C.c_int, C.c_int]
    lib.alto_stage_totals.restype = C.c_int
    lib.alto_stage_totals.argtypes = [P, P, C.c_int, C.c_int]
    for name in ("matMulVecPacked", "matMulVecPacked2", "matMulVecPacked4", "matMulVecPacked8"):
        getattr(lib, name).restype = None
        getattr(lib, name).argtypes = [P, P, P, SZ, SZ]
    lib.alto_matmul_vec_packed.restype = C.c_int
    lib.alto_matmul_vec_packed.argtypes = [P, P, P, SZ, SZ, SZ]
    lib.alto_set_graphs.restype = C.c_int
    lib.alto_set_graphs.argtypes = [P, C.c_int]
    _lib = lib
    return lib


def last_error() -> str:
    return load().alto_last_error().decode(errors="replace")
