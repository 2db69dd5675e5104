"""ctypes binding of libmmhip.so, the C ABI declared in include/mm_hermite.h.

The library is built in-tree (``mrmustard_amd/_build/libmmhip.so``) by
``mrmustard_amd.build.build_library``.  There is no fallback: if the library
is missing or cannot be loaded, every entry point raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmmhip.so")

MM_OK = 0
MM_ERR_INVALID = -1
MM_ERR_HIP = -2
MM_ERR_WORKSPACE = -3
MM_ERR_UNSUPPORTED = -4

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_sz = ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/mm_hermite.h
SIGNATURES = {
    "mm_last_error": (ctypes.c_char_p, []),
    "mm_version": (ctypes.c_char_p, []),
    "mm_init": (_int, [_int]),
    "mm_hermite_fwd_workspace": (_sz, [_vp, _int, _int]),
    "mm_hermite_fwd": (_int, [_vp, _vp, _vp, _vp, _int, _int, _vp, _vp, _sz, _vp]),
    "mm_hermite_batch_fwd_workspace": (_sz, [_i64, _vp, _int, _int]),
    "mm_hermite_batch_fwd": (_int, [_vp, _vp, _vp, _i64, _vp, _int, _int, _vp, _vp, _sz, _vp]),
    "mm_hermite_vjp_workspace": (_sz, [_vp, _int]),
    "mm_hermite_vjp": (_int, [_vp, _vp, _vp, _vp, _int, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mm_hermite_batch_vjp_workspace": (_sz, [_i64, _vp, _int]),
    "mm_hermite_batch_vjp": (_int, [_vp, _vp, _vp, _i64, _vp, _int, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mm_hermite_batch_vjp_static": (_int, [_vp, _vp, _vp, _i64, _vp, _int, _vp, _vp, _vp, _vp]),
    "mm_hermite_fused_workspace": (_sz, [_vp, _int]),
    "mm_hermite_fused": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mm_hermite_batch_fused_workspace": (_sz, [_i64, _vp, _int]),
    "mm_hermite_batch_fused": (_int, [_vp, _vp, _vp, _vp, _i64, _vp, _int, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mm_hermite_vjp_recompute_workspace": (_sz, [_vp, _int]),
    "mm_hermite_vjp_recompute": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mm_hermite_binomial_workspace": (_sz, [_vp, _int]),
    "mm_hermite_binomial": (_int, [_vp, _vp, _vp, _vp, _int, ctypes.c_double, _i64, _vp, _vp, _vp, _sz, _vp]),
    "mm_zgemm": (_int, [_vp, _i64, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _vp]),
    "mm_zgemm_batched": (_int, [_vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp]),
    "mm_autoshape_workspace": (_sz, [_i64, _int]),
    "mm_autoshape": (_int, [_vp, _vp, _vp, _i64, _int, ctypes.c_double, _int, _int, _vp, _vp, _sz, _vp]),
    "mm_hermite_diagonal_workspace": (_sz, [_vp, _int, _i64]),
    "mm_hermite_diagonal": (_int, [_vp, _vp, _vp, _int, _vp, _int, _i64, _int, _vp, _vp, _sz, _vp]),
    "mm_hermite_diagonal_vjp_workspace": (_sz, [_vp, _int]),
    "mm_hermite_diagonal_vjp": (_int, [_vp, _vp, _vp, _vp, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mm_hermite_1leftover_workspace": (_sz, [_i64, _vp, _int]),
    "mm_hermite_1leftover": (_int, [_vp, _vp, _vp, _i64, _vp, _int, _int, _vp, _vp, _sz, _vp]),
    "mm_hermite_fwd_contract_workspace": (_sz, [_vp, _int]),
    "mm_hermite_fwd_contract": (_int, [_vp, _vp, _vp, _vp, _int, _int, _vp, _vp, _vp, _sz, _vp]),
    "mm_hermite_batch_fwd_contract_workspace": (_sz, [_vp, _int]),
    "mm_hermite_batch_fwd_contract": (_int, [_vp, _vp, _vp, _i64, _vp, _int, _int, _vp, _i64, _i64, _vp, _vp, _sz, _vp]),
    "mm_hermite_1leftover_vjp_workspace": (_sz, [_i64, _vp, _int]),
    "mm_hermite_1leftover_vjp": (_int, [_vp, _vp, _vp, _i64, _vp, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mm_plan_create": (_int, [_i64, _vp, _vp, _int, _vp, _vp]),
    "mm_plan_destroy": (_int, [_vp]),
    "mm_plan_info": (_int, [_vp, _vp, _vp, _vp]),
    "mm_plan_describe": (_int, [_vp, _vp]),
    "mm_plan_run": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mm_plan_run_graph": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}


class MMTune(ctypes.Structure):
    """``mm_tune`` of include/mm_hermite.h."""

    _fields_ = [
        ("block_threads", ctypes.c_int),
        ("blocks_per_cu", ctypes.c_int),
        ("wave_max_slab", ctypes.c_int),
        ("reserved", ctypes.c_int),
    ]


_lock = threading.Lock()
_lib = None


def load(path: str | None = None):
    """Load libmmhip.so once; raise RuntimeError if it is absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(
                f"libmmhip.so not found at {p}: build it first "
                "(python -c 'import __graft_entry__ as g; g.build()')"
            )
        L = ctypes.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        return L


def check(rc: int, what: str) -> None:
    """Map a negative status to the reference's exception types."""
    if rc == MM_OK:
        return
    msg = load().mm_last_error().decode(errors="replace")
    if rc in (MM_ERR_INVALID, MM_ERR_UNSUPPORTED):
        raise ValueError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg} (status {rc})")
