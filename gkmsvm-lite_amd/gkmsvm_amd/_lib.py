"""ctypes binding of libgkmhip.so (the C-ABI in include/gkmhip.h).

The library is loaded from this package's ``lib/`` directory; there is no
fallback: a missing library or a machine without a ROCm GPU raises.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgkmhip.so")

P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
D = ctypes.c_double

# name -> argtypes (every function returns int status unless listed in _RESTYPE)
SIGNATURES = {
    "gkm_last_error": [],
    "gkm_version": [],
    "gkm_planes_words": [I],
    "gkm_sync": [P],
    "gkm_zero_u64": [P, I64, P],
    "gkm_pack_planes": [P, P, I, I, I, P, P, P],
    "gkm_pack_windows": [P, P, I, I, I, P, P, P],
    "gkm_decode_ascii": [P, P, P, I, I, P, P, P],
    "gkm_cross_hist": [P, P, I, I, P, P, P, P, I, I, I, I, P, I, I, I, I64, I64, P, P],
    "gkm_self_hist": [P, P, P, I, I, I, I, P, I, I, P, P],
    "gkm_combine_kraw": [P, I64, I, P, P, P],
    "gkm_combine_gram": [P, I, I, I, P, P, P, I, I, D, P, P],
    "gkm_combine_score": [P, I, I, I, P, P, P, P, I, I, D, D, P, P],
    "gkm_combine_impact": [P, I, I, I, P, P, P, P, I, I, D, P, P],
    "gkm_explain_coef": [P, I, I, I, P, P, P, P, I, I, D, P, P],
    "gkm_mutant_diag": [P, P, P, P, I, I, I, P, P, P, P, P],
    "gkm_explain": [I, P, P, I, I, I, P, P, P, I, I, I, I, I, P, I, I, P, P, P, P, P],
    "gkm_index_explain": [I, P, P, I, I, I, P, P, P, P, P, P, I, I, P, I, I, I, I, P, P, P, P,
                          P],
    "gkm_lmer_explain": [I, P, P, I, I, I, P, P, I, P, P, I, I, P, I, P, P, P, P],
    "gkm_lmer_explain_sorted": [I, P, P, I, P, I, I, I, P, P, I, P, P, I, I, P, I, P, P, P, P],
    "gkm_index_hist": [P, P, P, I, I, P, P, P, P, I, P, P, I, I, I, P, I, I, I, I, I, I64, I64,
                       P, P, P],
    "gkm_index_reduce": [P, I, P, I, I, I, I, I, I, I, P, P],
    "gkm_combine_fixed": [P, I, I, I, P, P, P, I, I, D, P, P],
    "gkm_index_psum_bytes": [I, I, I, I, I],
    "gkm_index_score": [P, P, P, I, I, P, P, P, P, I, P, P, I, I, I, P, I, I, I, I, I, P, P, P,
                        P, I, I, D, P, P],
    "gkm_index_queue_bytes": [I, I, I, I],
    "gkm_index_ism_bytes": [I, I, I, I],
    "gkm_index_ism": [P, P, I, I, P, P, P, P, I, P, P, I, I, I, P, I, I, I, I, P, P, P, P, P, I,
                      I, I, D, I, P, P, P],
    "gkm_index_score_queue": [P, P, P, I, I, P, P, P, P, I, P, P, I, I, I, P, I, I, I, I, I, P,
                              I, I, P, P, P, I, I, D, P, P],
    "gkm_index_score_queue_pairs": [P, P, P, I, I, P, P, P, P, I, P, P, I, I, I, P, I, I, I, I,
                                    I, P, I, I, P, P, P, I, I, D, P, I, P],
    "gkm_index_hist_queue": [P, P, P, I, I, P, P, P, P, I, P, P, I, I, I, P, I, I, I, I, I,
                             I, I, P, P, P, P],
    "gkm_scratch_bytes": [],
    "gkm_index_partial_bytes": [I, I, I, I, I, I, I, I],
    "gkm_index_hist_noalloc": [P, P, P, I, I, P, P, P, P, I, P, P, I, I, I, P, I, I, I, I, I,
                               I64, I64, P, P],
    "gkm_ism": [P, P, P, I, I, I, P, P, P, I, I, I, I, P, I, I, P, P, P, P, P, I, I, D, P,
                P],
    "gkm_lmer_weights": [P, P, P, P, I, P, P],
    "gkm_wht4": [P, I, I, P],
    "gkm_table_spectrum": [P, I, P, D, P],
    "gkm_explain_fill": [P, I, P, P, D, P, P],
    "gkm_table_score": [P, P, I, I, I, P, P, P, I, D, P, P],
    "gkm_table_explain": [P, P, I, I, I, I, P, P, P, I, I, P, P],
    "gkm_table_ism": [P, P, P, I, I, I, I, P, P, P, P, P, I, P, P],
    "gkm_deltasvm_weights": [P, I, D, P, I, I, P, P],
    "gkm_deltasvm_score": [P, P, I, I, I, I, P, I, P, I, D, P, P],
    "gkm_peak_read": [P, I64, P, I, P],
    "gkm_peak_valu": [I, P, I, P],
    "gkm_peak_gather": [P, I64, P, I, I, P],
}
_RESTYPE = {"gkm_last_error": ctypes.c_char_p, "gkm_scratch_bytes": ctypes.c_int64,
            "gkm_index_partial_bytes": ctypes.c_int64, "gkm_index_psum_bytes": ctypes.c_int64,
            "gkm_index_queue_bytes": ctypes.c_int64, "gkm_index_ism_bytes": ctypes.c_int64}

_lib = None


class GkmError(RuntimeError):
    """A libgkmhip call returned an error status."""


def load():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built: run gkmsvm-lite_amd/gkmsvm_amd/build_ext.py "
                "(the engine has no CPU fallback)"
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = lib
    return _lib


def call(name, *args):
    """Call a status-returning entry point; raise GkmError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.gkm_last_error().decode(errors="replace")
        raise GkmError(f"{name} failed ({rc}): {msg}")
    return rc
