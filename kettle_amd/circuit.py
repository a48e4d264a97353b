"""A Python mirror of kettle's ``ket::QuantumCircuit`` builder that flattens to libkq's op array.

Builder names, argument order and validation messages follow kettle
(include/kettle/circuit/circuit.hpp:20-419, source/kettle_internal/circuit/circuit.cpp):
1-qubit gates take ``target``, controlled gates ``(control, target)``, ``add_m_gate(qubit, bit)``.
``flatten()`` produces the ``kq_op`` records, predicate table and parameter order that
``kq_simulate`` / ``kq_apply_gates`` / ``kq_energy_gradient`` consume.
"""
from __future__ import annotations

import enum
import math
import uuid
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from ._lib import OP_DTYPE


class Gate(enum.IntEnum):
    """kettle's ``enum class Gate`` (include/kettle/gates/primitive_gate.hpp:19-52)."""

    H = 0
    X = 1
    Y = 2
    Z = 3
    S = 4
    SDAG = 5
    T = 6
    TDAG = 7
    SX = 8
    SXDAG = 9
    RX = 10
    RY = 11
    RZ = 12
    P = 13
    CH = 14
    CX = 15
    CY = 16
    CZ = 17
    CS = 18
    CSDAG = 19
    CT = 20
    CTDAG = 21
    CSX = 22
    CSXDAG = 23
    CRX = 24
    CRY = 25
    CRZ = 26
    CP = 27
    U = 28
    CU = 29
    M = 30


OP_IF = 64
OP_IF_ELSE = 65
OP_LOG_STATE = 66
OP_LOG_CREG = 67

ONE_TARGET = {Gate.H, Gate.X, Gate.Y, Gate.Z, Gate.S, Gate.SDAG, Gate.T, Gate.TDAG, Gate.SX, Gate.SXDAG}
ONE_TARGET_ANGLE = {Gate.RX, Gate.RY, Gate.RZ, Gate.P}
CONTROLLED = {Gate.CH, Gate.CX, Gate.CY, Gate.CZ, Gate.CS, Gate.CSDAG, Gate.CT, Gate.CTDAG, Gate.CSX, Gate.CSXDAG}
CONTROLLED_ANGLE = {Gate.CRX, Gate.CRY, Gate.CRZ, Gate.CP}


class ControlFlowBooleanKind(enum.IntEnum):
    IF = 0
    IF_NOT = 1


@dataclass
class ControlFlowPredicate:
    """kettle's ``ControlFlowPredicate`` (control_flow_predicate.cpp:12-38)."""

    bit_indices_to_check: list[int]
    expected_bits: list[int]
    control_kind: ControlFlowBooleanKind = ControlFlowBooleanKind.IF

    def __post_init__(self):
        if len(self.bit_indices_to_check) != len(self.expected_bits):
            raise RuntimeError(
                "There is a mismatch in the number of bits to check, and the number of expected bits:\n"
                f"bit_indices_to_check.size() = {len(self.bit_indices_to_check)}\n"
                f"expected_bits.size() = {len(self.expected_bits)}\n"
            )
        if len(self.bit_indices_to_check) == 0:
            raise RuntimeError("The control flow predicate must depend on at least one bit.\n")
        if not all(b in (0, 1) for b in self.expected_bits):
            raise RuntimeError("The expected bits must all be 0 or 1.\n")


@dataclass
class Matrix2X2:
    """kettle's ``Matrix2X2`` (include/kettle/common/matrix2x2.hpp)."""

    elem00: complex
    elem01: complex
    elem10: complex
    elem11: complex

    def as_array(self) -> np.ndarray:
        return np.array([[self.elem00, self.elem01], [self.elem10, self.elem11]], dtype=np.complex128)


def ParameterID() -> str:  # noqa: N802 - kettle's type name
    """A fresh parameter id (kettle draws 16 random bytes, parameter.hpp:23-25)."""
    return uuid.uuid4().hex


@dataclass
class _GateEl:
    gate: Gate
    q0: int
    q1: int = 0
    angle: float = 0.0
    param: str | None = None
    matrix: Matrix2X2 | None = None


@dataclass
class _IfEl:
    predicate: ControlFlowPredicate
    body: "QuantumCircuit"
    else_body: "QuantumCircuit | None" = None


@dataclass
class _LogEl:
    kind: int


@dataclass
class _ParamData:
    value: float | None
    name: str
    count: int = 1


class QuantumCircuit:
    """Circuit builder with kettle's method names (circuit.hpp:20-419)."""

    def __init__(self, n_qubits: int, n_bits: int | None = None):
        self._n_qubits = int(n_qubits)
        self._n_bits = int(n_qubits if n_bits is None else n_bits)
        self._elements: list = []
        self._params: dict[str, _ParamData] = {}
        self._param_count = 0
        self._version = 0          # bumped by every change of the element list
        self._flat_cache = None    # flatten()'s arrays of the last straight-line structure

    def _append(self, el) -> None:
        self._elements.append(el)
        self._version += 1

    # -- sizes -------------------------------------------------------------------------------
    def n_qubits(self) -> int:
        return self._n_qubits

    def n_bits(self) -> int:
        return self._n_bits

    def n_circuit_elements(self) -> int:
        return len(self._elements)

    @property
    def elements(self) -> list:
        return self._elements

    # -- validation (circuit.cpp:1002-1027) ---------------------------------------------------
    def _check_qubit(self, index: int, qubit_name: str, gate_name: str) -> None:
        if index < 0 or index >= self._n_qubits:
            raise RuntimeError(
                f"The index for the {qubit_name} at which the '{gate_name}' gate is applied, is out of bounds.\n"
                f"n_qubits             = {self._n_qubits}\nprovided qubit index = {index}\n"
            )

    def _check_bit(self, index: int) -> None:
        if index < 0 or index >= self._n_bits:
            raise RuntimeError(
                "The bit index at which the 'M' gate is applied, is out of bounds.\n"
                f"n_bits             = {self._n_bits}\nprovided bit index = {index}\n"
            )

    # -- parameters (circuit.cpp:83-90, 1170-1233) --------------------------------------------
    def _new_param(self, value: float | None, pid: str | None = None) -> str:
        pid = pid or ParameterID()
        self._params[pid] = _ParamData(value, f"theta{self._param_count}")
        self._param_count += 1
        return pid

    def _use_param(self, pid: str) -> str:
        if pid in self._params:
            self._params[pid].count += 1
        else:
            self._new_param(None, pid)
        return pid

    def set_parameter_value(self, pid: str, angle: float) -> None:
        if pid not in self._params:
            raise IndexError("ERROR: no parameter found with the provided id.\n")
        self._params[pid].value = float(angle)

    def parameter_values(self) -> dict[str, float | None]:
        return {k: v.value for k, v in self._params.items()}

    # -- generic builders -----------------------------------------------------------------------
    def _one(self, gate: Gate, target) -> None:
        if isinstance(target, Iterable):
            for t in target:
                self._one(gate, t)
            return
        self._check_qubit(int(target), "qubit", gate.name)
        self._append(_GateEl(gate, int(target)))

    def _one_angle(self, gate: Gate, target, angle=None, parameterized=None):
        if isinstance(target, Iterable) and not isinstance(target, str):
            for t, a in target:
                self._one_angle(gate, t, a)
            return None
        self._check_qubit(int(target), "qubit", gate.name)
        if isinstance(angle, str):
            pid = self._use_param(angle)
            self._append(_GateEl(gate, int(target), param=pid))
            return None
        if parameterized:
            pid = self._new_param(float(angle))
            self._append(_GateEl(gate, int(target), angle=float(angle), param=pid))
            return pid
        self._append(_GateEl(gate, int(target), angle=float(angle)))
        return None

    def _ctrl(self, gate: Gate, control, target=None) -> None:
        if target is None:
            for c, t in control:
                self._ctrl(gate, c, t)
            return
        self._check_qubit(int(control), "control qubit", gate.name)
        self._check_qubit(int(target), "target qubit", gate.name)
        self._append(_GateEl(gate, int(control), int(target)))

    def _ctrl_angle(self, gate: Gate, control, target=None, angle=None, parameterized=None):
        if target is None:
            for c, t, a in control:
                self._ctrl_angle(gate, c, t, a)
            return None
        self._check_qubit(int(control), "control qubit", gate.name)
        self._check_qubit(int(target), "target qubit", gate.name)
        if isinstance(angle, str):
            pid = self._use_param(angle)
            self._append(_GateEl(gate, int(control), int(target), param=pid))
            return None
        if parameterized:
            pid = self._new_param(float(angle))
            self._append(_GateEl(gate, int(control), int(target), angle=float(angle), param=pid))
            return pid
        self._append(_GateEl(gate, int(control), int(target), angle=float(angle)))
        return None

    # -- kettle's builder names -------------------------------------------------------------------
    def add_h_gate(self, t): self._one(Gate.H, t)
    def add_x_gate(self, t): self._one(Gate.X, t)
    def add_y_gate(self, t): self._one(Gate.Y, t)
    def add_z_gate(self, t): self._one(Gate.Z, t)
    def add_s_gate(self, t): self._one(Gate.S, t)
    def add_sdag_gate(self, t): self._one(Gate.SDAG, t)
    def add_t_gate(self, t): self._one(Gate.T, t)
    def add_tdag_gate(self, t): self._one(Gate.TDAG, t)
    def add_sx_gate(self, t): self._one(Gate.SX, t)
    def add_sxdag_gate(self, t): self._one(Gate.SXDAG, t)

    def add_rx_gate(self, t, angle=None, parameterized=None): return self._one_angle(Gate.RX, t, angle, parameterized)
    def add_ry_gate(self, t, angle=None, parameterized=None): return self._one_angle(Gate.RY, t, angle, parameterized)
    def add_rz_gate(self, t, angle=None, parameterized=None): return self._one_angle(Gate.RZ, t, angle, parameterized)
    def add_p_gate(self, t, angle=None, parameterized=None): return self._one_angle(Gate.P, t, angle, parameterized)

    def add_ch_gate(self, c, t=None): self._ctrl(Gate.CH, c, t)
    def add_cx_gate(self, c, t=None): self._ctrl(Gate.CX, c, t)
    def add_cy_gate(self, c, t=None): self._ctrl(Gate.CY, c, t)
    def add_cz_gate(self, c, t=None): self._ctrl(Gate.CZ, c, t)
    def add_cs_gate(self, c, t=None): self._ctrl(Gate.CS, c, t)
    def add_csdag_gate(self, c, t=None): self._ctrl(Gate.CSDAG, c, t)
    def add_ct_gate(self, c, t=None): self._ctrl(Gate.CT, c, t)
    def add_ctdag_gate(self, c, t=None): self._ctrl(Gate.CTDAG, c, t)
    def add_csx_gate(self, c, t=None): self._ctrl(Gate.CSX, c, t)
    def add_csxdag_gate(self, c, t=None): self._ctrl(Gate.CSXDAG, c, t)

    def add_crx_gate(self, c, t=None, angle=None, parameterized=None): return self._ctrl_angle(Gate.CRX, c, t, angle, parameterized)
    def add_cry_gate(self, c, t=None, angle=None, parameterized=None): return self._ctrl_angle(Gate.CRY, c, t, angle, parameterized)
    def add_crz_gate(self, c, t=None, angle=None, parameterized=None): return self._ctrl_angle(Gate.CRZ, c, t, angle, parameterized)
    def add_cp_gate(self, c, t=None, angle=None, parameterized=None): return self._ctrl_angle(Gate.CP, c, t, angle, parameterized)

    def add_u_gate(self, matrix: Matrix2X2, target) -> None:
        if isinstance(target, Iterable):
            for t in target:
                self.add_u_gate(matrix, t)
            return
        self._check_qubit(int(target), "qubit", "U")
        self._append(_GateEl(Gate.U, int(target), matrix=matrix))

    def add_cu_gate(self, matrix: Matrix2X2, control, target=None) -> None:
        if target is None:
            for c, t in control:
                self.add_cu_gate(matrix, c, t)
            return
        self._check_qubit(int(control), "control qubit", "CU")
        self._check_qubit(int(target), "target qubit", "CU")
        self._append(_GateEl(Gate.CU, int(control), int(target), matrix=matrix))

    def add_m_gate(self, target, bit=None) -> None:
        """``add_m_gate(q)`` stores the outcome in bit q; ``add_m_gate(q, bit)`` in ``bit``."""
        if isinstance(target, Iterable):
            for item in target:
                if isinstance(item, Iterable):
                    self.add_m_gate(*item)
                else:
                    self.add_m_gate(item)
            return
        bit = int(target) if bit is None else int(bit)
        self._check_qubit(int(target), "qubit", "M")
        self._check_bit(bit)
        self._append(_GateEl(Gate.M, int(target), bit))

    # compound helpers that expand to primitive gates before simulation (circuit.cpp:801-903)
    def add_swap_gate(self, t0, t1=None) -> None:
        if t1 is None:
            for a, b in t0:
                self.add_swap_gate(a, b)
            return
        if t0 == t1:
            raise RuntimeError("Cannot swap a index with itself")
        self.add_cx_gate(t0, t1)
        self.add_cx_gate(t1, t0)
        self.add_cx_gate(t0, t1)

    def _fourier_swaps(self, indices: Sequence[int]) -> None:
        left, right = 0, len(indices) - 1
        while right > left:
            self.add_swap_gate(indices[left], indices[right])
            left += 1
            right -= 1

    def add_qft_gate(self, indices: Sequence[int]) -> None:
        indices = list(indices)
        size = len(indices)
        for i_pre in range(size):
            target = indices[i_pre]
            self.add_h_gate(target)
            denom = 2
            for c_pre in range(i_pre + 1, size):
                angle = 2.0 * math.pi / float(1 << denom)
                self.add_cp_gate(indices[c_pre], target, angle)
                denom += 1
        self._fourier_swaps(indices)

    def add_iqft_gate(self, indices: Sequence[int]) -> None:
        indices = list(indices)
        size = len(indices)
        self._fourier_swaps(indices)
        for i_pre in reversed(range(size)):
            target = indices[i_pre]
            denom = size - i_pre
            for c_pre in reversed(range(i_pre + 1, size)):
                angle = 2.0 * math.pi / float(1 << denom)
                self.add_cp_gate(indices[c_pre], target, -angle)
                denom -= 1
            self.add_h_gate(target)

    # control flow (circuit.cpp:905-985)
    def _merge_params(self, sub: "QuantumCircuit") -> None:
        for pid, data in sub._params.items():
            if pid not in self._params:
                self._params[pid] = _ParamData(data.value, data.name, data.count)
                continue
            cur = self._params[pid]
            cur.count += data.count
            if cur.value is None and data.value is not None:
                cur.value = data.value
            elif cur.value is not None and data.value is not None and abs(cur.value - data.value) > 1.0e-6:
                raise RuntimeError("ERROR: found two parameter instances with same id but different non-optional values\n")

    def add_if_statement(self, predicate, subcircuit: "QuantumCircuit") -> None:
        if not isinstance(predicate, ControlFlowPredicate):
            predicate = ControlFlowPredicate([int(predicate)], [1], ControlFlowBooleanKind.IF)
        for b in predicate.bit_indices_to_check:
            self._check_bit(b)
        self._merge_params(subcircuit)
        self._append(_IfEl(predicate, subcircuit))

    def add_if_not_statement(self, bit: int, subcircuit: "QuantumCircuit") -> None:
        self.add_if_statement(ControlFlowPredicate([int(bit)], [0], ControlFlowBooleanKind.IF), subcircuit)

    def add_if_else_statement(self, predicate, if_sub: "QuantumCircuit", else_sub: "QuantumCircuit") -> None:
        if not isinstance(predicate, ControlFlowPredicate):
            predicate = ControlFlowPredicate([int(predicate)], [1], ControlFlowBooleanKind.IF)
        for b in predicate.bit_indices_to_check:
            self._check_bit(b)
        self._append(_IfEl(predicate, if_sub, else_sub))

    def add_if_not_else_statement(self, bit: int, if_sub: "QuantumCircuit", else_sub: "QuantumCircuit") -> None:
        self.add_if_else_statement(ControlFlowPredicate([int(bit)], [0], ControlFlowBooleanKind.IF), if_sub, else_sub)

    def add_classical_register_circuit_logger(self) -> None:
        self._append(_LogEl(OP_LOG_CREG))

    def add_statevector_circuit_logger(self) -> None:
        self._append(_LogEl(OP_LOG_STATE))

    def pop_back(self) -> None:
        if not self._elements:
            raise RuntimeError("ERROR: called `pop_back()` on a `QuantumCircuit` instance with no circuit elements.\n")
        self._elements.pop()
        self._version += 1

    # -- flattening ---------------------------------------------------------------------------------
    def parameter_order(self) -> list[str]:
        """Parameter ids in creation order: the index used in ``kq_op.aux``."""
        return list(self._params.keys())

    def flatten(self, param_order: Sequence[str] | None = None):
        """-> (ops: OP_DTYPE array, preds: int32 array, param_order).

        Evaluates every parameter once, like ``create_parameter_values_map``
        (parameter_expression.cpp:84-97): an uninitialised parameter raises kettle's error.
        """
        order = list(param_order) if param_order is not None else self.parameter_order()
        slot = {pid: i for i, pid in enumerate(order)}
        values = {}
        for pid, data in self._params.items():
            if data.value is None:
                raise RuntimeError("ERROR: cannot perform simulation with an uninitialized parameter value.\n")
            values[pid] = data.value
        # a straight-line circuit flattened before with the same elements and parameter order: only the
        # parameterised angles change (an optimiser's next call)
        key = (self._version, len(self._elements), tuple(order))
        cache = self._flat_cache
        if cache is not None and cache[0] == key:
            _, base, where, pids, preds_c = cache
            arr = base.copy()
            if len(where):
                arr["angle"][where] = [values[pid] for pid in pids]
            return arr, preds_c.copy(), order
        ops: list[tuple] = []
        preds: list[int] = []
        self._flatten_into(ops, preds, values, slot)
        arr = np.zeros(len(ops), dtype=OP_DTYPE)
        if ops:
            # one column at a time (a record-by-record fill costs more than the simulation of small states)
            g, q0, q1, aux, angle, u = zip(*ops)
            arr["gate"] = g
            arr["q0"] = q0
            arr["q1"] = q1
            arr["aux"] = aux
            arr["angle"] = angle
            arr["u"] = u
        preds_a = np.asarray(preds, dtype=np.int32)
        if all(isinstance(el, _GateEl) for el in self._elements):
            where = np.array([i for i, el in enumerate(self._elements) if el.param is not None], dtype=np.int64)
            pids = [self._elements[i].param for i in where]
            self._flat_cache = (key, arr.copy(), where, pids, preds_a.copy())
        return arr, preds_a, order

    def _flatten_into(self, ops, preds, values, slot) -> None:
        zero_u = (0.0,) * 8
        for el in self._elements:
            if isinstance(el, _GateEl):
                angle = el.angle
                aux = -1
                if el.param is not None:
                    angle = values[el.param]
                    aux = slot.get(el.param, -1)
                u = zero_u
                if el.matrix is not None:
                    m = el.matrix
                    u = (m.elem00.real, m.elem00.imag, m.elem01.real, m.elem01.imag,
                         m.elem10.real, m.elem10.imag, m.elem11.real, m.elem11.imag)
                ops.append((int(el.gate), el.q0, el.q1, aux, float(angle), u))
            elif isinstance(el, _IfEl):
                off = len(preds)
                p = el.predicate
                preds.extend([int(p.control_kind), len(p.bit_indices_to_check)])
                for b, e in zip(p.bit_indices_to_check, p.expected_bits):
                    preds.extend([int(b), int(e)])
                head = len(ops)
                ops.append(None)
                el.body._flatten_into(ops, preds, values, slot)
                len_if = len(ops) - head - 1
                if el.else_body is None:
                    ops[head] = (OP_IF, off, len_if, 0, 0.0, zero_u)
                else:
                    el.else_body._flatten_into(ops, preds, values, slot)
                    len_else = len(ops) - head - 1 - len_if
                    ops[head] = (OP_IF_ELSE, off, len_if, len_else, 0.0, zero_u)
            elif isinstance(el, _LogEl):
                ops.append((el.kind, 0, 0, 0, 0.0, zero_u))


def extend_circuit(left: QuantumCircuit, right: QuantumCircuit) -> None:
    """kettle's ``extend_circuit`` (append_circuits.hpp:8): append ``right``'s elements."""
    if left.n_qubits() != right.n_qubits():
        raise RuntimeError("ERROR: cannot extend circuits with different numbers of qubits.\n")
    left._merge_params(right)
    left._elements.extend(right._elements)
    left._version += 1
