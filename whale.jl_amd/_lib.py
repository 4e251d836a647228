"""ctypes bindings of include/whale_amd.h (libwhale_amd.so, built in-tree)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwhale_amd.so")
SYNTH_PATH = os.path.join(HERE, "libwhale_synth.so")

WHALE_OK = 0
FP64_PLAIN, FP32_SCALED, FP64_SCALED = 0, 1, 2
KEEP_ELL = 0x100
BATCHED, PER_FAMILY = 0, 1
NODE_LEAF, NODE_SPECIATION, NODE_WGD, NODE_ROOT = 0, 1, 2, 3
COND_ROOT, COND_NONEXTINCT, COND_NONE = 0, 1, 2

MODES = {"fp64": FP64_PLAIN, "fp32": FP32_SCALED, "fp64s": FP64_SCALED}
SIZE_CLASSES = 6  # WHALE_SIZE_CLASSES

_lib = None
_synth = None

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
dblp = C.POINTER(C.c_double)
vp = C.c_void_p


class WhaleError(RuntimeError):
    pass


def lib():
    """The native library; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WhaleError(f"{LIB_PATH} is missing: build it with `python whale.jl_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        L.whale_last_error.restype = C.c_char_p
        L.whale_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.whale_ctx_destroy.argtypes = [vp]
        L.whale_ctx_stream.restype = vp
        L.whale_ctx_stream.argtypes = [vp]
        L.whale_ctx_set_timing.argtypes = [vp, C.c_int32]
        L.whale_ctx_record.argtypes = [vp, C.c_int32]
        L.whale_ctx_elapsed.argtypes = [vp, C.c_int32, C.c_int32, dblp]
        L.whale_last_timings.argtypes = [vp, dblp]
        L.whale_model_create.argtypes = [vp, C.c_int32, i32p, i32p, i32p, i32p, i32p, dblp, dblp, i32p, i32p,
                                         C.c_int32, i32p, C.c_int32, C.POINTER(vp)]
        L.whale_model_destroy.argtypes = [vp]
        L.whale_model_set_params.argtypes = [vp, dblp, dblp, dblp, C.c_int32, C.c_double, dblp]
        L.whale_model_table_offsets.argtypes = [vp, i32p]
        L.whale_model_tables.argtypes = [vp, dblp, dblp, dblp]
        L.whale_ccd_parse.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_char_p), i32p, C.c_int32, C.POINTER(vp)]
        L.whale_ccd_parse_many.argtypes = [C.POINTER(C.c_char_p), C.POINTER(C.c_size_t), C.c_int32,
                                           C.POINTER(C.c_char_p), i32p, C.c_int32, C.c_int32, C.POINTER(vp),
                                           C.POINTER(C.c_int32)]
        L.whale_ccd_shape.argtypes = [vp, i32p, i32p]
        L.whale_ccd_export.argtypes = [vp, i32p, i32p, i32p, i32p, i32p, dblp]
        L.whale_ccd_destroy.argtypes = [vp]
        L.whale_batch_create.argtypes = [vp, C.c_int32, i64p, i32p, i32p, i64p, i32p, i32p, dblp, C.POINTER(vp)]
        L.whale_batch_destroy.argtypes = [vp]
        L.whale_batch_stats.argtypes = [vp, i64p]
        L.whale_batch_fallbacks.argtypes = [vp, i32p]
        L.whale_batch_hbm_families.argtypes = [vp, i32p]
        L.whale_batch_set_lds_limit.argtypes = [vp, C.c_int64]
        L.whale_batch_autotune.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_char_p, i64p, dblp]
        L.whale_batch_set_tuning.argtypes = [vp, i64p]
        L.whale_batch_set_ell_budget.argtypes = [vp, C.c_int64]
        L.whale_batch_ell_bytes.argtypes = [vp, i64p]
        L.whale_batch_set_deterministic.argtypes = [vp, C.c_int32]
        L.whale_batch_get_tuning.argtypes = [vp, i64p]
        L.whale_batch_pooled.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
        L.whale_batch_set_seeds.argtypes = [vp, dblp]
        L.whale_mixture_eval.argtypes = [vp, C.POINTER(vp), C.c_int32, dblp, C.c_int32, C.c_int32, dblp, dblp,
                                         dblp, dblp, dblp]
        L.whale_batch_launches.argtypes = [vp, C.c_int32, i64p]
        L.whale_logpdf.argtypes = [vp, C.c_int32, C.c_int32, dblp, dblp]
        L.whale_logpdf_grad.argtypes = [vp, C.c_int32, C.c_int32, dblp, dblp, dblp, dblp, dblp, dblp]
        L.whale_logpdf_noalloc.argtypes = [vp, C.c_int32, vp, vp]
        L.whale_download_ell.argtypes = [vp, C.c_int32, dblp, C.c_int64, i64p]
        L.whale_backtrack.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_int32, i32p, i32p]
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != WHALE_OK:
        msg = lib().whale_last_error().decode(errors="replace")
        if rc == 2:
            raise ValueError(f"{what}: {msg}")
        raise WhaleError(f"{what} failed ({rc}): {msg}")


def synth():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise WhaleError(f"{SYNTH_PATH} is missing: build it with `python whale.jl_amd/build.py`")
        S = C.CDLL(SYNTH_PATH)
        S.ws_species_tree.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_char_p, C.c_int]
        S.ws_families.restype = vp
        S.ws_families.argtypes = [C.c_char_p, C.c_uint64, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                  C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        S.ws_error.restype = C.c_char_p
        S.ws_error.argtypes = [vp]
        S.ws_count.argtypes = [vp]
        S.ws_text.restype = C.c_char_p
        S.ws_text.argtypes = [vp, C.c_int]
        S.ws_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        S.ws_free.argtypes = [vp]
        _synth = S
    return _synth


def ptr(a, t):
    return a.ctypes.data_as(t)
