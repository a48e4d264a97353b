"""Batched energy + adjoint gradient of many independent parameterised circuits (kq_batch_*).

    b = EnergyGradientBatch(KQ_C64)
    for circuit, hamiltonian in instances:
        b.add(circuit, hamiltonian)              # parameter order: the circuit's creation order
    r = b.run()                                  # r.energies[i], r.gradients[i], r.drift[i]
    b.set_parameters(i, new_values); r = b.run() # later runs replay a captured hipGraph

Each instance is kettle's VQE cost function (vqe_basic.cpp:61-82): a straight-line parameterised
circuit on |0...0> and the expectation of a Pauli operator, here with its full gradient.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import KQ_C64, check, load, ptr
from .api import PauliOperator
from .circuit import QuantumCircuit


@dataclass
class BatchResult:
    energies: np.ndarray           # (n_instances,)
    gradients: list[np.ndarray]    # per instance, in its parameter order
    drift: np.ndarray              # max |psi - |0...0>| after the reverse sweep, per instance


class EnergyGradientBatch:
    """precision: the reverse sweep's; forward_precision: the forward sweep's and the energies'
    (default: the same; KQ_C128 with KQ_C64 gives FP64-exact energies)."""

    def __init__(self, precision: int = KQ_C64, forward_precision: int | None = None):
        self.precision = precision
        self.forward_precision = precision if forward_precision is None else forward_precision
        self._h = ctypes.c_void_p()
        check(load().kq_batch_create(ctypes.byref(self._h), int(precision), int(self.forward_precision)))
        self._n_params: list[int] = []

    def __del__(self):
        self.close()

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            load().kq_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __len__(self) -> int:
        return len(self._n_params)

    def add(self, circuit: QuantumCircuit, hamiltonian: PauliOperator, params: Sequence[str] | None = None) -> int:
        ops, _, order = circuit.flatten(params)
        if any(int(g) >= 30 for g in ops["gate"]):
            raise RuntimeError("batch: circuits must be straight-line (no M, control flow or loggers)")
        offs, qs, ks, ph, co = hamiltonian.packed()
        idx = ctypes.c_size_t()
        check(load().kq_batch_add(self._h, circuit.n_qubits(), ptr(ops), len(ops), len(order), hamiltonian.size(),
                                  ptr(offs), ptr(qs), ptr(ks), ptr(ph), ptr(co), ctypes.byref(idx)))
        self._n_params.append(len(order))
        return idx.value

    def add_ops(self, n_qubits: int, ops: np.ndarray, n_params: int, hamiltonian: PauliOperator) -> int:
        """An instance given as a flat op array (aux = parameter slot, angle = its initial value)."""
        ops = np.ascontiguousarray(ops)
        offs, qs, ks, ph, co = hamiltonian.packed()
        idx = ctypes.c_size_t()
        check(load().kq_batch_add(self._h, int(n_qubits), ptr(ops), len(ops), int(n_params), hamiltonian.size(),
                                  ptr(offs), ptr(qs), ptr(ks), ptr(ph), ptr(co), ctypes.byref(idx)))
        self._n_params.append(int(n_params))
        return idx.value

    def set_parameters(self, index: int, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.size != self._n_params[index]:
            raise ValueError(f"instance {index} has {self._n_params[index]} parameters, got {v.size}")
        check(load().kq_batch_set_parameters(self._h, int(index), ptr(v)))

    def set_all_parameters(self, values) -> None:
        """Every instance's parameters (an optimiser step): one array per instance, in add order."""
        parts = [np.asarray(v, dtype=np.float64).ravel() for v in values]
        if len(parts) != len(self._n_params):
            raise ValueError(f"{len(self._n_params)} instances, got {len(parts)} parameter arrays")
        for i, (v, p) in enumerate(zip(parts, self._n_params)):
            if v.size != p:
                raise ValueError(f"instance {i} has {p} parameters, got {v.size}")
        flat = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(0))
        check(load().kq_batch_set_all_parameters(self._h, ptr(flat)))

    def run(self, use_graph: bool = True) -> BatchResult:
        n = len(self._n_params)
        energies = np.zeros(n)
        drift = np.zeros(n)
        grads = np.zeros(max(sum(self._n_params), 1))
        check(load().kq_batch_run(self._h, int(bool(use_graph)), ptr(energies), ptr(grads), ptr(drift)))
        out = []
        o = 0
        for p in self._n_params:
            out.append(grads[o:o + p].copy())
            o += p
        return BatchResult(energies, out, drift)

    def plan_info(self) -> dict:
        """Forward / reverse tile passes over all instances and one run's algorithmic bytes."""
        a, b, c = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_double()
        check(load().kq_batch_plan_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"forward_passes": a.value, "reverse_passes": b.value, "algorithmic_bytes": c.value}

    def set_drift_recompute(self, drift_tol: float | None) -> None:
        """After each later run, instances whose drift exceeds drift_tol get their energy and gradient
        from the single-state complex128 adjoint instead (None turns it off)."""
        check(load().kq_batch_set_drift_recompute(self._h, -1.0 if drift_tol is None else float(drift_tol)))

    def recomputed(self) -> tuple[int, int]:
        """(instances recomputed by the last run, in total)."""
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        check(load().kq_batch_recomputed(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def info(self) -> dict:
        a, b, c, d = (ctypes.c_size_t() for _ in range(4))
        check(load().kq_batch_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
        return {"instances": a.value, "parameters": b.value, "device_bytes": c.value, "launches": d.value}


__all__ = ["EnergyGradientBatch", "BatchResult", "_lib"]
