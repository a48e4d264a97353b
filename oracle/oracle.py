"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke check and the benchmark's CPU-baseline leg import this module; the
product (kettle_amd) never does.  Inputs are the same flattened kq_op arrays libkq consumes.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int8, c_int32, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None

LOGGER_FN = ctypes.CFUNCTYPE(None, c_void_p, c_int32, POINTER(c_double), POINTER(c_int8), c_size_t)

_SIG = {
    "oracle_last_error": (c_char_p, []),
    "oracle_pair_map_1q": (c_int, [c_uint32, c_uint32, c_void_p, c_void_p]),
    "oracle_pair_at_1q": (c_int, [c_uint32, c_uint32, c_uint64, POINTER(c_uint64), POINTER(c_uint64)]),
    "oracle_pair_map_2q": (c_int, [c_uint32, c_uint32, c_uint32, c_int, c_void_p]),
    "oracle_pair_at_2q": (c_int, [c_uint32, c_uint32, c_uint32, c_uint64, POINTER(c_uint64), POINTER(c_uint64)]),
    "oracle_sv_simulate": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_size_t, c_void_p, c_int, c_int32,
                                   c_void_p, c_size_t, LOGGER_FN, c_void_p]),
    "oracle_dm_simulate": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_size_t, c_void_p, c_int, c_int32,
                                   c_void_p, c_size_t]),
    "oracle_probabilities": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p]),
    "oracle_sample_probabilities": (c_int, [c_void_p, c_uint64, c_uint64, c_int, c_int32, c_void_p]),
    "oracle_cdf": (c_int, [c_void_p, c_uint64, c_void_p, POINTER(c_double)]),
    "oracle_pauli_expectation": (c_int, [c_void_p, c_uint32, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p]),
    "oracle_apply_pauli_string": (c_int, [c_void_p, c_uint32, c_size_t, c_void_p, c_void_p]),
    "oracle_sv_to_dm": (c_int, [c_void_p, c_uint32, c_void_p]),
    "oracle_dm_trace": (c_int, [c_void_p, c_uint32, c_void_p]),
    "oracle_dm_trace_of_square": (c_int, [c_void_p, c_uint32, c_void_p]),
    "oracle_dm_partial_trace": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t, c_void_p]),
}


class OracleError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing; run `make -C oracle`")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIG.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a):
    return c_void_p(a.ctypes.data) if a is not None and a.size else c_void_p(0)


def _check(st):
    if st != 0:
        msg = load().oracle_last_error().decode(errors="replace")
        if st == 5:
            raise IndexError(msg)
        raise OracleError(st, msg)


def zero_state(n):
    a = np.zeros(1 << n, np.complex128)
    a[0] = 1.0
    return a


def sv_simulate(amps, n, ops, preds=None, seed=None, n_bits=0, circuit_n=None, logger=None):
    """-> (amps, creg).  ``amps`` is copied; kettle semantics incl. M and control flow."""
    a = np.array(amps, dtype=np.complex128, copy=True)
    ops = np.ascontiguousarray(ops)
    preds = np.zeros(0, np.int32) if preds is None else np.ascontiguousarray(preds, np.int32)
    creg = np.full(max(n_bits, 1), -1, np.int8)
    cb = LOGGER_FN(logger) if logger else LOGGER_FN()
    _check(load().oracle_sv_simulate(_ptr(a), n, n if circuit_n is None else circuit_n, _ptr(ops), len(ops),
                                     _ptr(preds), 0 if seed is None else 1, 0 if seed is None else seed,
                                     _ptr(creg), n_bits, cb, None))
    return a, creg[:n_bits]


def dm_simulate(rho_vec, n, ops, preds=None, seed=None, n_bits=0, circuit_n=None):
    r = np.array(rho_vec, dtype=np.complex128, copy=True)
    ops = np.ascontiguousarray(ops)
    preds = np.zeros(0, np.int32) if preds is None else np.ascontiguousarray(preds, np.int32)
    creg = np.full(max(n_bits, 1), -1, np.int8)
    _check(load().oracle_dm_simulate(_ptr(r), n, n if circuit_n is None else circuit_n, _ptr(ops), len(ops),
                                     _ptr(preds), 0 if seed is None else 1, 0 if seed is None else seed,
                                     _ptr(creg), n_bits))
    return r, creg[:n_bits]


def probabilities(amps, n, noise=None):
    a = np.ascontiguousarray(amps, np.complex128)
    out = np.empty(1 << n)
    nz = None if noise is None else np.ascontiguousarray(noise, np.float64)
    _check(load().oracle_probabilities(_ptr(a), n, _ptr(nz) if nz is not None else c_void_p(0), _ptr(out)))
    return out


def sample_probabilities(probs, shots, seed):
    p = np.ascontiguousarray(probs, np.float64)
    out = np.empty(shots, np.uint64)
    _check(load().oracle_sample_probabilities(_ptr(p), p.size, shots, 1, seed, _ptr(out)))
    return out


def cdf(probs):
    p = np.ascontiguousarray(probs, np.float64)
    out = np.empty_like(p)
    up = c_double()
    _check(load().oracle_cdf(_ptr(p), p.size, _ptr(out), ctypes.byref(up)))
    return out, up.value


def pauli_expectation(amps, n, packed):
    offs, qs, ks, ph, co = packed
    a = np.ascontiguousarray(amps, np.complex128)
    out = np.zeros(2)
    _check(load().oracle_pauli_expectation(_ptr(a), n, len(ph), _ptr(offs), _ptr(qs), _ptr(ks), _ptr(ph), _ptr(co),
                                           _ptr(out)))
    return complex(out[0], out[1])


def apply_pauli_string(amps, n, qubits, kinds):
    a = np.array(amps, np.complex128, copy=True)
    qs = np.ascontiguousarray(qubits, np.uint32)
    ks = np.ascontiguousarray(kinds, np.uint8)
    _check(load().oracle_apply_pauli_string(_ptr(a), n, len(qs), _ptr(qs), _ptr(ks)))
    return a


def sv_to_dm(amps, n):
    a = np.ascontiguousarray(amps, np.complex128)
    out = np.empty(1 << (2 * n), np.complex128)
    _check(load().oracle_sv_to_dm(_ptr(a), n, _ptr(out)))
    return out


def dm_trace(rho_vec, n):
    r = np.ascontiguousarray(rho_vec, np.complex128)
    out = np.zeros(2)
    _check(load().oracle_dm_trace(_ptr(r), n, _ptr(out)))
    return complex(out[0], out[1])


def dm_trace_of_square(rho_vec, n):
    r = np.ascontiguousarray(rho_vec, np.complex128)
    out = np.zeros(2)
    _check(load().oracle_dm_trace_of_square(_ptr(r), n, _ptr(out)))
    return complex(out[0], out[1])


def dm_partial_trace(rho_vec, n, qubits):
    r = np.ascontiguousarray(rho_vec, np.complex128)
    qs = np.ascontiguousarray(qubits, np.uint32)
    side = 1 << (n - len(qs))
    out = np.empty(side * side, np.complex128)
    _check(load().oracle_dm_partial_trace(_ptr(r), n, _ptr(qs), len(qs), _ptr(out)))
    return out


def pair_map_1q(n, t):
    np_ = 1 << (n - 1)
    s0 = np.empty(np_, np.uint64)
    s1 = np.empty(np_, np.uint64)
    _check(load().oracle_pair_map_1q(n, t, _ptr(s0), _ptr(s1)))
    return s0, s1


def pair_at_1q(n, t, k):
    a, b = c_uint64(), c_uint64()
    _check(load().oracle_pair_at_1q(n, t, k, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def pair_map_2q(n, c, t, four_way=False):
    np_ = 0 if n < 2 else 1 << (n - 2)
    out = np.empty(np_ * (4 if four_way else 2), np.uint64)
    _check(load().oracle_pair_map_2q(n, c, t, 1 if four_way else 0, _ptr(out)))
    return out.reshape(np_, 4 if four_way else 2)


def pair_at_2q(n, c, t, k):
    a, b = c_uint64(), c_uint64()
    _check(load().oracle_pair_at_2q(n, c, t, k, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def time_gate_reps(n, ops, reps):
    """Seconds of each of `reps` runs of every op on one n-qubit state: array (len(ops), reps)."""
    load().oracle_time_gate_reps.restype = c_int
    load().oracle_time_gate_reps.argtypes = [c_uint32, c_void_p, c_size_t, c_uint32, c_void_p]
    ops = np.ascontiguousarray(ops)
    out = np.zeros((len(ops), reps))
    _check(load().oracle_time_gate_reps(n, _ptr(ops), len(ops), reps, _ptr(out)))
    return out


def time_gates(n, ops):
    """Seconds the oracle spends applying ``ops`` to |0...0> of n qubits (gate loop only)."""
    load().oracle_time_gates.restype = c_int
    load().oracle_time_gates.argtypes = [c_uint32, c_void_p, c_size_t, POINTER(c_double)]
    ops = np.ascontiguousarray(ops)
    sec = c_double()
    _check(load().oracle_time_gates(n, _ptr(ops), len(ops), ctypes.byref(sec)))
    return sec.value
