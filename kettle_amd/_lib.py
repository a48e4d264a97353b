"""ctypes binding of libkq's C-ABI (include/kq.h).

The library is built in-tree (``python -m kettle_amd.build``) into ``kettle_amd/lib/libkq.so``.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int8, c_int32, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libkq.so"

KQ_OK = 0
KQ_ERR_INVALID = 1
KQ_ERR_HIP = 2
KQ_ERR_ALLOC = 3
KQ_ERR_RUNTIME = 4
KQ_ERR_OUT_OF_RANGE = 5

KQ_C128 = 0
KQ_C64 = 1

# kq_op as a numpy record: matches `typedef struct kq_op` (88 bytes, 8-byte aligned)
OP_DTYPE = np.dtype(
    [("gate", "<i4"), ("q0", "<i4"), ("q1", "<i4"), ("aux", "<i4"), ("angle", "<f8"), ("u", "<f8", (8,))],
    align=True,
)
assert OP_DTYPE.itemsize == 88

# kq_pass_stats[KQ_PASS_CLASSES]
PASS_STATS_DTYPE = np.dtype([("ms", "<f8"), ("launches", "<u8"), ("bytes", "<f8"), ("gates", "<u8"),
                             ("min_ms", "<f8"), ("max_ms", "<f8")])
PASS_CLASSES = ("tile_low", "tile_high", "sweep", "adjoint_reverse", "pauli_apply")

LOGGER_FN = ctypes.CFUNCTYPE(None, c_void_p, c_int32, POINTER(c_double), POINTER(c_int8), c_size_t)


class KqError(RuntimeError):
    """A libkq call failed; ``status`` is the kq.h status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


_SIGNATURES = {
    "kq_last_error": (c_char_p, []),
    "kq_init": (c_int, [c_int]),
    "kq_device_info": (c_int, [c_char_p, c_size_t, POINTER(c_int), POINTER(c_size_t)]),
    "kq_synchronize": (c_int, []),
    "kq_sv_create": (c_int, [POINTER(c_void_p), c_uint32, c_int]),
    "kq_dm_create": (c_int, [POINTER(c_void_p), c_uint32, c_int]),
    "kq_state_destroy": (c_int, [c_void_p]),
    "kq_state_info": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_int), POINTER(c_int)]),
    "kq_state_device_ptr": (c_int, [c_void_p, POINTER(c_void_p), POINTER(c_size_t)]),
    "kq_state_upload": (c_int, [c_void_p, c_void_p]),
    "kq_state_download": (c_int, [c_void_p, c_void_p]),
    "kq_state_set_basis": (c_int, [c_void_p, c_uint64]),
    "kq_state_copy": (c_int, [c_void_p, c_void_p]),
    "kq_simulate": (
        c_int,
        [c_void_p, c_uint32, c_void_p, c_size_t, c_void_p, c_size_t, c_int, c_int32, c_void_p, c_size_t, LOGGER_FN, c_void_p],
    ),
    "kq_apply_gates": (c_int, [c_void_p, c_void_p, c_size_t]),
    "kq_set_fusion": (c_int, [c_int, c_int, c_int, c_double]),
    "kq_get_fusion": (c_int, [POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_double)]),
    "kq_last_plan_stats": (c_int, [POINTER(c_size_t), POINTER(c_size_t), POINTER(c_size_t)]),
    "kq_qubit_weights": (c_int, [c_void_p, c_uint32, c_void_p]),
    "kq_collapse": (c_int, [c_void_p, c_uint32, c_int, c_double]),
    "kq_measure": (c_int, [c_void_p, c_uint32, c_int, c_int32, POINTER(c_int)]),
    "kq_probabilities": (c_int, [c_void_p, c_void_p, c_void_p]),
    "kq_sample": (c_int, [c_void_p, c_void_p, c_uint64, c_int, c_int32, c_void_p]),
    "kq_sample_probabilities": (c_int, [c_void_p, c_uint64, c_uint64, c_int, c_int32, c_void_p]),
    "kq_sample_probabilities_noalloc": (c_int, [c_void_p, c_uint64, c_uint64, c_int, c_int32, c_void_p]),
    "kq_pauli_expectation": (c_int, [c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "kq_apply_pauli_string": (c_int, [c_void_p, c_size_t, c_void_p, c_void_p]),
    "kq_dm_trace": (c_int, [c_void_p, c_void_p]),
    "kq_dm_trace_of_square": (c_int, [c_void_p, c_void_p]),
    "kq_dm_partial_trace": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "kq_dm_from_statevector": (c_int, [c_void_p, c_void_p]),
    "kq_energy_gradient": (
        c_int,
        [c_void_p, c_void_p, c_size_t, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t,
         POINTER(c_double), c_void_p, c_uint32, c_double, POINTER(c_size_t)],
    ),
    "kq_profile_begin": (c_int, []),
    "kq_profile_end": (c_int, [c_void_p]),
    "kq_debug_pair_map_1q": (c_int, [c_uint32, c_uint32, c_void_p, c_void_p]),
    "kq_debug_pair_map_2q": (c_int, [c_uint32, c_uint32, c_uint32, c_int, c_void_p]),
    "kq_set_jit": (c_int, [c_int]),
    "kq_stream_probe": (c_int, [c_uint32, c_void_p]),
    "kq_stream_probe_shapes": (c_int, [c_uint32, c_void_p]),
    "kq_debug_dense_block": (c_int, [c_void_p, c_void_p, c_int, POINTER(c_double)]),
    "kq_debug_mfma16_layout": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "kq_debug_set_faults": (c_int, [c_size_t, c_int, c_int]),
    "kq_sample_circuit_shots": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t, c_void_p, c_size_t, c_size_t, c_void_p,
                                        c_uint64, c_int, c_int32, c_uint64, c_void_p, POINTER(c_uint64)]),
    "kq_set_fusion_low_qubits": (c_int, [c_int]),
    "kq_set_fusion_relabel": (c_int, [c_int]),
    "kq_set_fusion_autotune": (c_int, [c_int]),
    "kq_fusion_autotune_info": (c_int, [POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_size_t), POINTER(c_size_t)]),
    "kq_debug_pass": (c_int, [c_uint32, c_int, c_void_p, c_size_t, c_size_t, c_size_t, c_char_p, c_size_t,
                              POINTER(c_size_t), c_void_p, c_size_t, POINTER(c_size_t), c_void_p, POINTER(c_size_t)]),
    "kq_debug_pass_from_basis": (c_int, [c_uint32, c_int, c_void_p, c_size_t, c_size_t, c_size_t, c_uint64, c_char_p,
                                          c_size_t, POINTER(c_size_t), c_void_p, c_size_t, POINTER(c_size_t), c_void_p,
                                          POINTER(c_size_t)]),
    "kq_debug_check_relabel": (c_int, [c_uint32, c_void_p, c_size_t, c_void_p, POINTER(c_size_t), POINTER(c_size_t)]),
    "kq_set_jit_tuning": (c_int, [c_int, c_int, c_int]),
    "kq_set_jit_priority": (c_int, [c_int]),
    "kq_set_jit_nontemporal": (c_int, [c_int]),
    "kq_jit_info": (c_int, [POINTER(c_int), POINTER(c_size_t), POINTER(c_size_t), POINTER(c_size_t), POINTER(c_double),
                            POINTER(c_size_t)]),
    "kq_debug_jit_compile": (c_int, [c_uint32, c_int, c_void_p, c_size_t, c_size_t, POINTER(c_size_t), POINTER(c_size_t)]),
    "kq_debug_adjoint_jit_compile": (c_int, [c_uint32, c_int, c_void_p, c_size_t, c_size_t, POINTER(c_size_t),
                                             POINTER(c_size_t)]),
    "kq_debug_jit_source": (c_int, [c_uint32, c_int, c_void_p, c_size_t, c_size_t, c_size_t, c_char_p, c_size_t,
                                    POINTER(c_size_t), POINTER(c_size_t)]),
    "kq_batch_create": (c_int, [POINTER(c_void_p), c_int, c_int]),
    "kq_batch_add": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t, c_size_t, c_size_t, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, POINTER(c_size_t)]),
    "kq_batch_set_parameters": (c_int, [c_void_p, c_size_t, c_void_p]),
    "kq_batch_set_all_parameters": (c_int, [c_void_p, c_void_p]),
    "kq_batch_run": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "kq_batch_info": (c_int, [c_void_p, POINTER(c_size_t), POINTER(c_size_t), POINTER(c_size_t), POINTER(c_size_t)]),
    "kq_batch_set_drift_recompute": (c_int, [c_void_p, c_double]),
    "kq_batch_plan_info": (c_int, [c_void_p, POINTER(c_size_t), POINTER(c_size_t), POINTER(c_double)]),
    "kq_batch_recomputed": (c_int, [c_void_p, POINTER(c_size_t), POINTER(c_size_t)]),
    "kq_batch_destroy": (c_int, [c_void_p]),
    "kq_debug_plan": (c_int, [c_uint32, c_int, c_int, c_void_p, c_size_t, c_size_t, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p]),
}

_lib = None


def exported_symbols() -> list[str]:
    """The entry points kq.h declares (the binding's table)."""
    return sorted(_SIGNATURES)


def load() -> ctypes.CDLL:
    """Load libkq.so (once).  Raises if it has not been built: there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"libkq.so not found at {LIB_PATH}; build it with `python -m kettle_amd.build`")
    lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != KQ_OK:
        msg = load().kq_last_error().decode(errors="replace")
        if status == KQ_ERR_OUT_OF_RANGE:
            raise IndexError(msg)
        raise KqError(status, msg)


def ptr(a: np.ndarray) -> c_void_p:
    return c_void_p(a.ctypes.data) if a is not None and a.size else c_void_p(0)


__all__ = [
    "LIB_PATH", "OP_DTYPE", "LOGGER_FN", "KqError", "load", "check", "ptr", "exported_symbols",
    "KQ_C128", "KQ_C64", "KQ_OK", "KQ_ERR_INVALID", "KQ_ERR_HIP", "KQ_ERR_ALLOC", "KQ_ERR_RUNTIME",
    "KQ_ERR_OUT_OF_RANGE", "c_uint8",
]
