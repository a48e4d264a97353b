"""State handles and kettle's free functions over libkq.

Mirrors kettle's read-out API:
  ``simulate``                        simulation/simulate.hpp:41, simulate_density_matrix.hpp:46
  ``calculate_probabilities_raw``     calculations/probabilities.hpp:43-46
  ``perform_measurements_as_memory``  calculations/measurements.hpp:42-60
  ``perform_measurements_as_counts``  calculations/measurements.hpp:92-97 (and _raw / _marginal)
  ``expectation_value``               operator/pauli/pauli_operator.hpp:82-84
plus ``energy_and_gradient`` (adjoint sweep; kettle has no gradient).
"""
from __future__ import annotations

import enum
from collections import Counter
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import KQ_C128, KQ_C64, LOGGER_FN, check, load, ptr
from .circuit import OP_LOG_CREG, OP_LOG_STATE, QuantumCircuit


def _p(x):
    return _lib.ctypes.c_void_p(x)


class _State:
    """Owning handle of a device state (statevector or vectorised density matrix)."""

    _create = None

    def __init__(self, n_qubits: int, precision: int = KQ_C128):
        lib = load()
        h = _lib.ctypes.c_void_p()
        check(getattr(lib, self._create)(_lib.ctypes.byref(h), int(n_qubits), int(precision)))
        self._h = h
        self._n = int(n_qubits)
        self._precision = int(precision)

    @property
    def handle(self):
        return self._h

    def n_qubits(self) -> int:
        return self._n

    def n_states(self) -> int:
        return 1 << self._n

    @property
    def precision(self) -> int:
        return self._precision

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            load().kq_state_destroy(self._h)
            self._h = _lib.ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _n_values(self) -> int:
        raise NotImplementedError

    def upload(self, amps: np.ndarray) -> None:
        a = np.ascontiguousarray(amps, dtype=np.complex128).reshape(-1)
        if a.size != self._n_values():
            raise ValueError(f"expected {self._n_values()} complex values, got {a.size}")
        check(load().kq_state_upload(self._h, ptr(a)))

    def download(self) -> np.ndarray:
        out = np.empty(self._n_values(), dtype=np.complex128)
        check(load().kq_state_download(self._h, ptr(out)))
        return out

    def set_basis(self, index: int) -> None:
        check(load().kq_state_set_basis(self._h, int(index)))

    def device_ptr(self) -> tuple[int, int]:
        p = _lib.ctypes.c_void_p()
        b = _lib.ctypes.c_size_t()
        check(load().kq_state_device_ptr(self._h, _lib.ctypes.byref(p), _lib.ctypes.byref(b)))
        return p.value, b.value


class Statevector(_State):
    """``ket::Statevector`` resident in HBM; starts in |0...0> like ``Statevector(n)``."""

    _create = "kq_sv_create"

    def __init__(self, n_qubits_or_amps, precision: int = KQ_C128):
        if isinstance(n_qubits_or_amps, (int, np.integer)):
            super().__init__(int(n_qubits_or_amps), precision)
        else:
            amps = np.asarray(n_qubits_or_amps, dtype=np.complex128).reshape(-1)
            n = int(amps.size).bit_length() - 1
            if amps.size < 2 or (1 << n) != amps.size:
                raise RuntimeError("The provided coefficients must have a size equal to a power of 2.\n")
            super().__init__(n, precision)
            self.upload(amps)

    def _n_values(self) -> int:
        return 1 << self._n


class DensityMatrix(_State):
    """``ket::DensityMatrix`` (column-major, as Eigen stores it) resident in HBM as vec(rho)."""

    _create = "kq_dm_create"

    def _n_values(self) -> int:
        return 1 << (2 * self._n)

    def matrix(self) -> np.ndarray:
        """rho as a 2-D array (rows = first index)."""
        N = 1 << self._n
        return self.download().reshape(N, N).T

    def upload_matrix(self, rho: np.ndarray) -> None:
        self.upload(np.asarray(rho, dtype=np.complex128).T.reshape(-1))

    def trace(self) -> complex:
        out = np.zeros(2)
        check(load().kq_dm_trace(self._h, ptr(out)))
        return complex(out[0], out[1])

    def trace_of_square(self) -> complex:
        """Trace of rho * rho (density_matrix.cpp:164-168), FP64 on the device."""
        out = np.zeros(2)
        check(load().kq_dm_trace_of_square(self._h, ptr(out)))
        return complex(out[0], out[1])

    def is_pure(self, tolerance: float = 1.0e-8) -> bool:
        """kettle's |Trace(rho^2) - 1|^2 <= tolerance (std::norm, density_matrix.cpp:151-155)."""
        d = self.trace_of_square() - 1.0
        return d.real * d.real + d.imag * d.imag <= tolerance

    @staticmethod
    def from_statevector(sv: Statevector) -> "DensityMatrix":
        dm = DensityMatrix(sv.n_qubits(), sv.precision)
        check(load().kq_dm_from_statevector(dm.handle, sv.handle))
        return dm


def statevector_to_density_matrix(sv: Statevector) -> DensityMatrix:
    return DensityMatrix.from_statevector(sv)


def partial_trace(dm: DensityMatrix, qubit_indices: Sequence[int]) -> np.ndarray:
    """kettle's ``partial_trace`` (density_matrix.cpp:231-278): traces out ``qubit_indices``."""
    qs = np.asarray(sorted(set(int(q) for q in qubit_indices)), dtype=np.uint32)
    side = 1 << (dm.n_qubits() - len(qs))
    out = np.empty(side * side, dtype=np.complex128)
    check(load().kq_dm_partial_trace(dm.handle, ptr(qs), len(qs), ptr(out)))
    return out.reshape(side, side).T


@dataclass
class SimulationResult:
    classical_register: list[int | None]
    loggers: list = field(default_factory=list)


def simulate(circuit: QuantumCircuit, state: _State, prng_seed: int | None = None) -> SimulationResult:
    """``ket::simulate(circuit, state, seed)``; returns the classical register and logger snapshots."""
    ops, preds, _ = circuit.flatten()
    n_bits = circuit.n_bits()
    creg = np.full(max(n_bits, 1), -1, dtype=np.int8)
    logs: list = []

    def _logger(_user, kind, amps, cr, nb):
        if kind == OP_LOG_STATE:
            n = state.n_states()
            logs.append(("statevector", np.ctypeslib.as_array(amps, shape=(2 * n,)).view(np.complex128).copy()))
        elif kind == OP_LOG_CREG:
            reg = np.ctypeslib.as_array(cr, shape=(nb,)).copy() if nb else np.zeros(0, np.int8)
            logs.append(("classical_register", [None if v < 0 else int(v) for v in reg]))

    cb = LOGGER_FN(_logger)
    check(
        load().kq_simulate(
            state.handle, circuit.n_qubits(), ptr(ops), len(ops), ptr(preds), len(preds),
            0 if prng_seed is None else 1, 0 if prng_seed is None else int(prng_seed), ptr(creg), n_bits, cb, None,
        )
    )
    reg = [None if v < 0 else int(v) for v in creg[:n_bits]]
    return SimulationResult(reg, logs)


def apply_gates(state: _State, ops: np.ndarray) -> None:
    """Run a straight-line op array (no measurement / control flow) through the fused path."""
    ops = np.ascontiguousarray(ops)
    check(load().kq_apply_gates(state.handle, ptr(ops), len(ops)))


def plan_stats() -> tuple[int, int, int]:
    c = _lib.ctypes
    a, b, d = c.c_size_t(), c.c_size_t(), c.c_size_t()
    check(load().kq_last_plan_stats(c.byref(a), c.byref(b), c.byref(d)))
    return a.value, b.value, d.value


@dataclass
class FusionPlan:
    """Host-only fusion plan of a straight-line gate list (``kq_debug_plan``)."""
    order: np.ndarray       # record index of each applied gate, in execution order
    pass_of: np.ndarray     # pass that applies it
    pass_kind: np.ndarray   # per pass: 0 low-qubit tile, 1 tile with high qubits, 2 single-gate sweep
    pass_cost: np.ndarray   # per pass: arithmetic cost (general 2x2 gate = 1)


def debug_plan(ops: np.ndarray, n_qubits: int, precision: int = KQ_C128, is_density: bool = False,
               lds_bytes: int = 160 * 1024) -> FusionPlan:
    n_rec = len(ops) * (2 if is_density else 1)
    order = np.zeros(max(n_rec, 1), dtype=np.uint32)
    pass_of = np.zeros_like(order)
    kind = np.zeros(max(n_rec, 1), dtype=np.int32)
    cost = np.zeros(max(n_rec, 1), dtype=np.float64)
    n_passes = _lib.ctypes.c_size_t()
    check(load().kq_debug_plan(int(n_qubits), int(precision), int(bool(is_density)), ptr(ops), len(ops), int(lds_bytes),
                               ptr(order), ptr(pass_of), ptr(kind), ptr(cost), _lib.ctypes.byref(n_passes)))
    npass = n_passes.value
    return FusionPlan(order[:n_rec], pass_of[:n_rec], kind[:npass], cost[:npass])


def set_fusion_low_qubits(low_qubits: int) -> None:
    """Qubits 0..low_qubits-1 belong to every tile pass (default 5)."""
    check(load().kq_set_fusion_low_qubits(int(low_qubits)))


def set_fusion_relabel(on: bool) -> None:
    """Whether statevector gate passes may move qubits onto other index bits (default on)."""
    check(load().kq_set_fusion_relabel(int(bool(on))))


def set_fusion_autotune(on: bool) -> None:
    """Plan tuner (default on; explicit stage / low-qubit settings switch it off): per gate structure,
    the tile settings whose plan a cost model prices lowest."""
    check(load().kq_set_fusion_autotune(int(bool(on))))


def fusion_autotune_info() -> dict:
    c = _lib.ctypes
    on, low, st, tried, cached = c.c_int(), c.c_int(), c.c_int(), c.c_size_t(), c.c_size_t()
    check(load().kq_fusion_autotune_info(c.byref(on), c.byref(low), c.byref(st), c.byref(tried), c.byref(cached)))
    return {"on": bool(on.value), "low_qubits": low.value, "max_stages": st.value, "plans_tried": tried.value,
            "cached_choices": cached.value}


def set_jit(mode: int) -> None:
    """Run-time compiled tile passes: 0 off, 1 states of >= 18 qubits (default), 2 every tile pass."""
    check(load().kq_set_jit(int(mode)))


def set_jit_tuning(min_waves_per_simd: int = 3, blocks_per_cu: int = 0, normalise: bool = True) -> None:
    """Register cap (min waves per SIMD, default 3, 0 = compiler's choice), persistent blocks per CU (0 = one
    block per tile) and pivot normalisation of one-qubit gates in the run-time compiled pass kernels."""
    check(load().kq_set_jit_tuning(int(min_waves_per_simd), int(blocks_per_cu), int(bool(normalise))))


def set_jit_priority(mode: int) -> None:
    """Wave priority in the run-time pass kernels: 0 off, 1 raised while a wave issues its tile's loads,
    2 also while it issues its stores."""
    check(load().kq_set_jit_priority(int(mode)))


def set_jit_nontemporal(mode: int) -> None:
    """Nontemporal hint on the run-time pass kernels' HBM traffic: 0 off, 1 loads, 2 loads and stores,
    3 (default) loads and stores of vectors larger than the 256 MiB Infinity Cache."""
    check(load().kq_set_jit_nontemporal(int(mode)))


def jit_info() -> dict:
    c = _lib.ctypes
    mode, comp, hits, fails, secs, unc = c.c_int(), c.c_size_t(), c.c_size_t(), c.c_size_t(), c.c_double(), c.c_size_t()
    check(load().kq_jit_info(c.byref(mode), c.byref(comp), c.byref(hits), c.byref(fails), c.byref(secs), c.byref(unc)))
    return {"mode": mode.value, "compiled": comp.value, "cache_hits": hits.value, "failures": fails.value,
            "compile_seconds": secs.value, "uncapped": unc.value}


def debug_jit_compile(ops: np.ndarray, n_qubits: int, precision: int = KQ_C128, lds_bytes: int = 160 * 1024) -> int:
    """Host-only: generate and compile the run-time pass kernels of a gate list; returns their count."""
    c = _lib.ctypes
    nk, nf = c.c_size_t(), c.c_size_t()
    check(load().kq_debug_jit_compile(int(n_qubits), int(precision), ptr(ops), len(ops), int(lds_bytes), c.byref(nk),
                                      c.byref(nf)))
    return nk.value


def debug_jit_source(ops: np.ndarray, n_qubits: int, pass_index: int = 0, precision: int = KQ_C128,
                     lds_bytes: int = 160 * 1024) -> tuple[str, int]:
    """Host-only: (generated source of tile pass pass_index, number of tile passes)."""
    c = _lib.ctypes
    ln, nk = c.c_size_t(), c.c_size_t()
    args = (int(n_qubits), int(precision), ptr(ops), len(ops), int(lds_bytes), int(pass_index))
    check(load().kq_debug_jit_source(*args, None, 0, c.byref(ln), c.byref(nk)))
    buf = c.create_string_buffer(ln.value + 1)
    check(load().kq_debug_jit_source(*args, buf, ln.value + 1, c.byref(ln), c.byref(nk)))
    return buf.value.decode(), nk.value


def set_fusion(max_tile_qubits: int, max_gates_per_pass: int = 0, max_stages_per_pass: int = 0,
               max_cost_per_pass: float = 0.0) -> None:
    """Tile size 2^max_tile_qubits (0: no fusion, -1: largest that fits LDS); <= 0 keeps a bound."""
    check(load().kq_set_fusion(int(max_tile_qubits), int(max_gates_per_pass), int(max_stages_per_pass),
                               float(max_cost_per_pass)))


# ---- probabilities and sampling ----------------------------------------------------------------------

class QuantumNoise:
    """kettle's ``QuantumNoise`` (probabilities.cpp:20-47): per-qubit bit-flip probability."""

    def __init__(self, n_qubits: int):
        self._noise = np.zeros(n_qubits, dtype=np.float64)

    def set(self, index: int, noise: float) -> None:
        if index >= len(self._noise):
            raise RuntimeError("ERROR: Out-of-bounds access for QuantumNoise probability.")
        if not (0.0 <= noise <= 1.0):
            raise RuntimeError("ERROR: Cannot set probability in QuantumNoise object outside of [0, 1].")
        self._noise[index] = noise

    def get(self, index: int) -> float:
        if index >= len(self._noise):
            raise RuntimeError("ERROR: Out-of-bounds access for QuantumNoise probability.")
        return float(self._noise[index])

    def array(self) -> np.ndarray:
        return self._noise


def calculate_probabilities_raw(state: _State, noise: QuantumNoise | None = None) -> np.ndarray:
    out = np.empty(state.n_states(), dtype=np.float64)
    nz = None if noise is None else np.ascontiguousarray(noise.array())
    check(load().kq_probabilities(state.handle, ptr(nz) if nz is not None else _p(0), ptr(out)))
    return out


def _seed(seed):
    return (0, 0) if seed is None else (1, int(seed))


def perform_measurements_as_memory(state_or_probs, n_shots: int, noise: QuantumNoise | None = None,
                                   seed: int | None = None) -> np.ndarray:
    out = np.empty(int(n_shots), dtype=np.uint64)
    hs, sv = _seed(seed)
    if isinstance(state_or_probs, _State):
        nz = None if noise is None else np.ascontiguousarray(noise.array())
        check(load().kq_sample(state_or_probs.handle, ptr(nz) if nz is not None else _p(0), int(n_shots), hs, sv, ptr(out)))
    else:
        p = np.ascontiguousarray(state_or_probs, dtype=np.float64)
        check(load().kq_sample_probabilities(ptr(p), p.size, int(n_shots), hs, sv, ptr(out)))
    return out


def state_index_to_bitstring(index: int, n_qubits: int, marginal: Sequence[int] = ()) -> str:
    """Little-endian bitstring: character i is bit i; marginal qubits print as 'x'
    (qubit_state_conversion.cpp:37-83, marginal.cpp:19-43)."""
    marg = set(marginal)
    return "".join("x" if i in marg else ("1" if (index >> i) & 1 else "0") for i in range(n_qubits))


def perform_measurements_as_counts_raw(state_or_probs, n_shots: int, noise=None, seed=None) -> dict[int, int]:
    mem = perform_measurements_as_memory(state_or_probs, n_shots, noise, seed)
    return dict(sorted(Counter(int(i) for i in mem).items()))


DEFAULT_SHOT_CACHE_BYTES = 8 << 30


def perform_measurements_as_memory_circuit(circuit: QuantumCircuit, state: Statevector, n_shots: int,
                                           noise: QuantumNoise | None = None, seed: int | None = None,
                                           max_cached_bytes: int = DEFAULT_SHOT_CACHE_BYTES) -> tuple[np.ndarray, int]:
    """Per-shot re-simulation of ``circuit`` from ``state`` (measurements.cpp:178-208): one index per
    shot, and the number of branch states simulated (shots with the same mid-circuit outcome history
    share one branch)."""
    if circuit.n_qubits() != state.n_qubits():
        raise RuntimeError("Invalid simulation; circuit and state have different number of qubits.")
    ops, preds, _ = circuit.flatten()
    out = np.empty(int(n_shots), dtype=np.uint64)
    hs, sv = _seed(seed)
    nz = None if noise is None else np.ascontiguousarray(noise.array())
    nb = _lib.ctypes.c_uint64()
    check(load().kq_sample_circuit_shots(state.handle, circuit.n_qubits(), ptr(ops), len(ops), ptr(preds), len(preds),
                                         circuit.n_bits(), ptr(nz) if nz is not None else _p(0), int(n_shots), hs, sv,
                                         int(max_cached_bytes), ptr(out), _lib.ctypes.byref(nb)))
    return out, nb.value


def perform_measurements_as_counts_marginal(state_or_probs, n_shots, marginal_qubits: Sequence[int] = (),
                                            noise=None, seed=None, *args) -> dict[str, int]:
    """kettle's overloads: (state | raw probabilities, shots, marginal, noise, seed), and
    (circuit, state, shots, marginal, noise, seed), which re-simulates the circuit per shot."""
    if isinstance(state_or_probs, QuantumCircuit):
        circuit, state, shots = state_or_probs, n_shots, marginal_qubits
        marginal = noise if noise is not None else ()
        noise_, seed_ = (seed, args[0] if args else None)
        n = circuit.n_qubits()
        if any(q >= n for q in marginal):
            raise RuntimeError("ERROR: marginal qubit index out of range.")
        idx, _ = perform_measurements_as_memory_circuit(circuit, state, shots, noise_, seed_)
        counts: Counter = Counter()
        for i in idx.tolist():
            counts[state_index_to_bitstring(i, n, marginal)] += 1
        return dict(sorted(counts.items()))
    if isinstance(state_or_probs, _State):
        n = state_or_probs.n_qubits()
    else:
        size = len(state_or_probs)
        if size == 0 or size & (size - 1):
            raise RuntimeError("The number of probabilities must be a power of 2.\n")
        n = size.bit_length() - 1
    if any(q >= n for q in marginal_qubits):
        raise RuntimeError("ERROR: marginal qubit index out of range.")
    raw = perform_measurements_as_counts_raw(state_or_probs, n_shots, noise, seed)
    counts: Counter = Counter()
    for idx, c in raw.items():
        counts[state_index_to_bitstring(idx, n, marginal_qubits)] += c
    return dict(sorted(counts.items()))


def perform_measurements_as_counts(state, n_shots: int, noise=None, seed=None) -> dict[str, int]:
    return perform_measurements_as_counts_marginal(state, n_shots, (), noise, seed)


# ---- Pauli strings ------------------------------------------------------------------------------------

class PauliTerm(enum.IntEnum):
    I = 0
    X = 1
    Y = 2
    Z = 3


class PauliPhase(enum.IntEnum):
    PLUS_ONE = 0
    PLUS_EYE = 1
    MINUS_ONE = 2
    MINUS_EYE = 3


class SparsePauliString:
    """kettle's ``SparsePauliString`` (sparse_pauli_string.hpp:42-144): (qubit, term) pairs."""

    def __init__(self, terms, n_qubits: int | None = None, phase: PauliPhase = PauliPhase.PLUS_ONE):
        if n_qubits is None:
            # dense form: one PauliTerm per qubit
            self._n = len(terms)
            self._terms = [(i, PauliTerm(t)) for i, t in enumerate(terms)]
        else:
            self._n = int(n_qubits)
            self._terms = [(int(q), PauliTerm(t)) for q, t in terms]
        if self._n == 0:
            raise RuntimeError("ERROR: the number of qubits in the SparsePauliString must be positive.\n")
        self._phase = PauliPhase(phase)

    def n_qubits(self) -> int:
        return self._n

    def terms(self) -> list[tuple[int, PauliTerm]]:
        return list(self._terms)

    def phase(self) -> PauliPhase:
        return self._phase

    def set_phase(self, phase: PauliPhase) -> None:
        self._phase = PauliPhase(phase)

    def add(self, qubit: int, term: PauliTerm) -> None:
        """kettle's ``SparsePauliString::add`` (sparse_pauli_string.cpp): a new qubit only."""
        if qubit >= self._n:
            raise RuntimeError("ERROR: cannot perform operation on SparsePauliString with index beyond qubit range.\n")
        if any(q == qubit for q, _ in self._terms):
            raise RuntimeError("ERROR: Pauli term is already present in the string\n")
        self._terms.append((int(qubit), PauliTerm(term)))


class PauliOperator:
    """kettle's ``PauliOperator`` (pauli_operator.hpp): sum of coefficient * string."""

    def __init__(self, weighted=None, n_qubits: int | None = None):
        self._strings: list[tuple[complex, SparsePauliString]] = []
        if weighted:
            for c, s in weighted:
                self.add(c, s)
            n_set = {s.n_qubits() for _, s in self._strings}
            if len(n_set) != 1:
                raise RuntimeError(
                    "ERROR: construction of `PauliOperator` with WeightedPauliString instances requires\n"
                    "all pauli strings to have the same number of qubits.\n"
                )
            self._n = n_set.pop()
        else:
            self._n = int(n_qubits or 0)

    def n_qubits(self) -> int:
        return self._n

    def add(self, coefficient: complex, pauli_string: SparsePauliString) -> None:
        self._strings.append((complex(coefficient), pauli_string))

    def weighted_pauli_strings(self):
        return list(self._strings)

    def size(self) -> int:
        return len(self._strings)

    def packed(self):
        """-> (term_offsets u32, term_qubits u32, term_kinds u8, phases u8, coeffs f64[2S])."""
        offs = [0]
        qs: list[int] = []
        ks: list[int] = []
        ph: list[int] = []
        co: list[float] = []
        for c, s in self._strings:
            for q, t in s.terms():
                qs.append(q)
                ks.append(int(t))
            offs.append(len(qs))
            ph.append(int(s.phase()))
            co.extend([c.real, c.imag])
        return (np.asarray(offs, np.uint32), np.asarray(qs, np.uint32), np.asarray(ks, np.uint8),
                np.asarray(ph, np.uint8), np.asarray(co, np.float64))


class PackedPauliOperator:
    """A Pauli operator held directly in kq_pauli_expectation's packing (term offsets, term qubits,
    term kinds, phases, coefficients as re/im pairs); interchangeable with ``PauliOperator`` wherever
    only ``packed()``, ``size()`` and ``n_qubits()`` are used."""

    def __init__(self, n_qubits: int, offsets, qubits, kinds, phases, coeffs):
        self._n = int(n_qubits)
        self._p = (np.ascontiguousarray(offsets, np.uint32), np.ascontiguousarray(qubits, np.uint32),
                   np.ascontiguousarray(kinds, np.uint8), np.ascontiguousarray(phases, np.uint8),
                   np.ascontiguousarray(coeffs, np.float64))

    def n_qubits(self) -> int:
        return self._n

    def size(self) -> int:
        return len(self._p[3])

    def packed(self):
        return self._p

    def to_operator(self) -> "PauliOperator":
        offs, qs, ks, ph, co = self._p
        op = PauliOperator(n_qubits=self._n)
        for s in range(self.size()):
            terms = [(int(qs[j]), int(ks[j])) for j in range(offs[s], offs[s + 1])]
            op.add(complex(co[2 * s], co[2 * s + 1]), SparsePauliString(terms, self._n, PauliPhase(int(ph[s]))))
        return op


def expectation_value(op, state: Statevector) -> complex:
    """kettle's ``expectation_value`` for a ``PauliOperator`` or a single ``SparsePauliString``."""
    if isinstance(op, SparsePauliString):
        op = PauliOperator([(1.0, op)])
    offs, qs, ks, ph, co = op.packed()
    out = np.zeros(2)
    check(load().kq_pauli_expectation(state.handle, op.size(), ptr(offs), ptr(qs), ptr(ks), ptr(ph), ptr(co), ptr(out)))
    return complex(out[0], out[1])


def simulate_pauli_string(pauli_string: SparsePauliString, state: Statevector) -> None:
    """``ket::simulate(SparsePauliString, Statevector&)`` (simulate_pauli.cpp:124-128)."""
    if pauli_string.n_qubits() != state.n_qubits():
        raise RuntimeError("Invalid simulation; SparsePauliString and state have different number of qubits.")
    terms = pauli_string.terms()
    qs = np.asarray([q for q, _ in terms], np.uint32)
    ks = np.asarray([int(t) for _, t in terms], np.uint8)
    check(load().kq_apply_pauli_string(state.handle, len(terms), ptr(qs), ptr(ks)))


# ---- adjoint gradient ------------------------------------------------------------------------------------

@dataclass
class EnergyGradient:
    energy: float
    gradient: np.ndarray
    restores: int


def energy_and_gradient(circuit: QuantumCircuit, hamiltonian: PauliOperator, params: Sequence[str] | None = None,
                        precision: int = KQ_C64, state: Statevector | None = None, checkpoint_every: int = 0,
                        drift_tol: float = 1.0e-5) -> EnergyGradient:
    """E = <psi(theta)|H|psi(theta)> and dE/dtheta for the circuit's parameters (adjoint sweep).

    ``params`` fixes the gradient order (default: the circuit's parameter creation order).
    The sweep starts from ``state`` (default |0...0>) and leaves psi(theta) in it."""
    ops, _, order = circuit.flatten(params)
    if any(int(g) >= 30 for g in ops["gate"]):
        raise RuntimeError("energy_and_gradient: the circuit must be straight-line (no M, control flow or loggers)")
    if state is None:
        state = Statevector(circuit.n_qubits(), precision)
    offs, qs, ks, ph, co = hamiltonian.packed()
    grad = np.zeros(len(order), dtype=np.float64)
    e = _lib.ctypes.c_double()
    nr = _lib.ctypes.c_size_t()
    check(
        load().kq_energy_gradient(
            state.handle, ptr(ops), len(ops), hamiltonian.size(), ptr(offs), ptr(qs), ptr(ks), ptr(ph), ptr(co),
            len(order), _lib.ctypes.byref(e), ptr(grad), int(checkpoint_every), float(drift_tol), _lib.ctypes.byref(nr),
        )
    )
    return EnergyGradient(e.value, grad, nr.value)
