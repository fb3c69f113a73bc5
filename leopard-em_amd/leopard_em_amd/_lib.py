"""ctypes binding of the tdtm C ABI (``include/tdtm.h``).

The shared library ``libtdtm.so`` is built in-tree next to this file by
``leopard_em_amd.build.build()``.  There is no fallback: every product entry point
raises if the library or a GPU is missing.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_int32, c_int64, c_void_p

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtdtm.so")

_lib = None

_f32p = POINTER(c_float)
_i32p = POINTER(c_int32)

# name -> (restype, argtypes); must match include/tdtm.h exactly
SIGNATURES = {
    "tdtm_abi_version": (c_int, []),
    "tdtm_last_error": (c_char_p, []),
    "tdtm_match_plan_create": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, POINTER(c_void_p)]),
    "tdtm_match_plan_destroy": (c_int, [c_void_p]),
    "tdtm_match_plan_bytes": (c_int64, [c_void_p]),
    "tdtm_match_plan_set_option": (c_int, [c_void_p, c_char_p, c_int]),
    "tdtm_match_plan_set_image_size": (c_int, [c_void_p, c_int, c_int]),
    "tdtm_match_plan_phase_times": (c_int, [c_void_p, c_void_p, c_void_p]),
    "tdtm_match_set_image": (c_int, [c_void_p, c_void_p, c_void_p]),
    "tdtm_match_set_filters": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "tdtm_stats_init": (c_int, [c_int64] + [c_void_p] * 6 + [c_void_p]),
    "tdtm_match_step": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int] + [c_void_p] * 6 + [c_void_p]),
    "tdtm_match_run": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int] + [c_void_p] * 6 + [c_void_p]),
    "tdtm_match_finalize": (c_int, [c_int64, c_int64] + [c_void_p] * 6 + [c_void_p, c_int, c_void_p, c_int]
                            + [c_void_p] * 7 + [c_void_p]),
    "tdtm_stats_merge": (c_int, [c_int64] + [c_void_p] * 12 + [c_void_p]),
    "tdtm_refine_particle": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p]),
    "tdtm_refine_batch": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int64] + [c_void_p] * 18 + [c_void_p]),
    "tdtm_proposal_create": (c_int, [c_int, c_int, POINTER(c_void_p)]),
    "tdtm_proposal_destroy": (c_int, [c_void_p]),
    "tdtm_proposal_num_params": (c_int64, [c_void_p]),
    "tdtm_proposal_set_params": (c_int, [c_void_p, c_void_p, c_void_p]),
    "tdtm_proposal_get_params": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "tdtm_proposal_forward": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "tdtm_proposal_backward": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "tdtm_proposal_adam_step": (c_int, [c_void_p, c_float, c_float, c_float, c_float, c_void_p]),
    "tdtm_rfft2": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "tdtm_volume_rfft": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "tdtm_irfft2_any_workspace_bytes": (c_int64, [c_int, c_int]),
    "tdtm_irfft2_any": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_void_p]),
    "tdtm_prepare_image_workspace_bytes": (c_int64, [c_int, c_int, c_int]),
    "tdtm_prepare_image": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, ctypes.c_double, c_void_p, c_void_p,
                                   c_void_p, c_void_p]),
    "tdtm_prepare_image_noalloc": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, ctypes.c_double, c_void_p,
                                           c_void_p, c_void_p]),
    "tdtm_ctf_filter_stack": (c_int, [c_int, c_int, c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tdtm_ctf_filter_stack_mag": (c_int, [c_int, c_int, c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tdtm_match_plan_set_magnification": (c_int, [c_void_p, c_void_p]),
    "tdtm_prepare_image_filtered_workspace_bytes": (c_int64, [c_int, c_int, c_int]),
    "tdtm_prepare_image_filtered": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, ctypes.c_double, c_int,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, c_int, c_void_p,
                                            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
}


class TdtmError(RuntimeError):
    """Error reported by the tdtm library."""


def load(path: str | None = None):
    """Load (once) and return the ctypes library handle; raises if it is not built.
    ``TDTM_LIBRARY`` in the environment names another build of the library to load."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("TDTM_LIBRARY") or LIB_PATH
    if not os.path.exists(p):
        raise TdtmError(f"tdtm HIP library not built: {p} is missing (run __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.tdtm_abi_version() != 1:
        raise TdtmError("tdtm ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().tdtm_last_error()
        raise TdtmError(msg.decode() if msg else f"tdtm error {rc}")


def ptr(t) -> c_void_p:
    """Raw device pointer of a contiguous torch tensor."""
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return c_void_p(t.data_ptr())
