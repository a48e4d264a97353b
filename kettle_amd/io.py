"""Text formats at the boundary of the simulator.

* kettle's Hamiltonian reader ``read_pauli_operator`` (io/read_pauli_operator.cpp:29-69 in the
  reference): one weighted Pauli string per line, ``<re> <im> : (q, P) (q, P) ...`` with P in
  {X, Y, Z}; a line with no terms is the identity string.
* kettle's tangelo-style circuit reader and writer ``read_tangelo_circuit`` /
  ``write_tangelo_circuit`` (io/read_tangelo_file.cpp:241-376, io/write_tangelo_file.cpp:24-213,
  io/io_control_flow.cpp:13-139): one gate per line (``RX        target : [5]   parameter : 1.57``),
  U / CU followed by two matrix lines, ``IF BITS[..] == [..]`` / ``ELSE`` blocks indented by 4 spaces.
* kettle's statevector snapshots ``save_statevector`` / ``load_statevector``
  (io/statevector.cpp:74-169) and the numpy text reader ``read_numpy_statevector``
  (io/numpy_statevector.cpp:35-76).
"""
from __future__ import annotations

import enum
import io
import os
import re
from typing import TextIO

import numpy as np

from .api import PauliOperator, PauliTerm, SparsePauliString
from .circuit import (CONTROLLED, CONTROLLED_ANGLE, ONE_TARGET, ONE_TARGET_ANGLE, ControlFlowBooleanKind,
                      ControlFlowPredicate, Gate, Matrix2X2, QuantumCircuit, _GateEl, _IfEl, _LogEl)

_TERMS = {"X": PauliTerm.X, "Y": PauliTerm.Y, "Z": PauliTerm.Z}
_TERM_RE = re.compile(r"\(\s*(\d+)\s*,\s*(\S)\s*\)")


def read_pauli_operator(source: str | os.PathLike | TextIO, n_qubits: int) -> PauliOperator:
    """``source``: a path, or an open text stream."""
    if isinstance(source, (str, os.PathLike)):
        try:
            with open(source, encoding="utf-8") as f:
                return _read(f, n_qubits)
        except FileNotFoundError as e:
            raise OSError(f"ERROR: unable to open: {os.fspath(source)}") from e
    return _read(source, n_qubits)


def _read(stream: TextIO | io.StringIO, n_qubits: int) -> PauliOperator:
    op = PauliOperator(n_qubits=n_qubits)
    for line in stream:
        if not line.strip():
            continue
        head, _, tail = line.partition(":")
        re_part, im_part = head.split()[:2]
        s = SparsePauliString([], n_qubits)
        for q, t in _TERM_RE.findall(tail):
            if t not in _TERMS:
                raise KeyError(f"ERROR: unknown Pauli term '{t}'")
            s.add(int(q), _TERMS[t])
        op.add(complex(float(re_part), float(im_part)), s)
    return op


# ---- tangelo-style circuits ------------------------------------------------------------------------------

CONTROL_FLOW_INDENT = 4                      # io_control_flow.hpp:19
_TANGELO_NAMES = {"CPHASE": "CP", "CNOT": "CX", "PHASE": "P"}   # read_tangelo_file.cpp:23-37
_INT_RE = re.compile(r"\d+")
_FLOAT_RE = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


def _bracket_ints(text: str) -> list[int]:
    """The integers inside each ``[...]`` of a gate line, in order."""
    out = []
    for group in re.findall(r"\[([^\]]*)\]", text):
        out.extend(int(v) for v in _INT_RE.findall(group))
    return out


def _angle(text: str) -> float:
    return float(_FLOAT_RE.findall(text.split("parameter", 1)[1])[0])


def _matrix(lines: list[str], i: int) -> Matrix2X2:
    a = [float(v) for v in _FLOAT_RE.findall(lines[i])]
    b = [float(v) for v in _FLOAT_RE.findall(lines[i + 1])]
    return Matrix2X2(complex(a[0], a[1]), complex(a[2], a[3]), complex(b[0], b[1]), complex(b[2], b[3]))


def _index_list(text: str) -> tuple[list[int], str]:
    """``[i, j, ...]`` at the start of ``text`` (after any prefix up to '['); returns the rest."""
    start = text.index("[")
    end = text.index("]", start)
    body = text[start + 1:end]
    return [int(v) for v in re.findall(r"-?\d+", body)], text[end + 1:]


def _predicate(text: str) -> ControlFlowPredicate:
    bits, rest = _index_list(text)
    sign = rest.split()[0] if rest.split() else ""
    if sign == "==":
        kind = ControlFlowBooleanKind.IF
    elif sign == "!=":
        kind = ControlFlowBooleanKind.IF_NOT
    else:
        raise RuntimeError("Invalid comparison sign found; must be '==' or '!='\n")
    expected, _ = _index_list(rest)
    return ControlFlowPredicate(bits, expected, kind)


def _read_block(n_qubits: int, lines: list[str], i: int, indent: int | None) -> tuple[QuantumCircuit, int]:
    circuit = QuantumCircuit(n_qubits)
    while i < len(lines):
        line = lines[i]
        if indent is not None and not line.startswith(" " * indent):
            return circuit, i
        i += 1
        words = line.split()
        if not words:
            continue
        name = words[0]
        if name == "IF":
            predicate = _predicate(line.split("IF", 1)[1])
            body, i = _read_block(n_qubits, lines, i, CONTROL_FLOW_INDENT)
            circuit.add_if_statement(predicate, body)
            continue
        if name == "ELSE":
            els = circuit.elements
            if not els or not isinstance(els[-1], _IfEl) or els[-1].else_body is not None:
                raise RuntimeError(
                    "ERROR: encountered an 'ELSE' statement, but no previous matching 'IF' statement was found.\n")
            stmt = els[-1]
            circuit.pop_back()
            else_body, i = _read_block(n_qubits, lines, i, CONTROL_FLOW_INDENT)
            circuit.add_if_else_statement(stmt.predicate, stmt.body, else_body)
            continue
        name = _TANGELO_NAMES.get(name, name)
        if name == "SWAP":
            q0, q1 = _bracket_ints(line)[:2]
            circuit.add_swap_gate(q0, q1)
            continue
        try:
            gate = Gate[name]
        except KeyError:
            raise RuntimeError(f"Unknown gate found in `read_tangelo_file()` : {name}\n") from None
        ints = _bracket_ints(line)
        if gate in ONE_TARGET:
            circuit._one(gate, ints[0])
        elif gate in ONE_TARGET_ANGLE:
            circuit._one_angle(gate, ints[0], _angle(line))
        elif gate in CONTROLLED:
            circuit._ctrl(gate, ints[1], ints[0])
        elif gate in CONTROLLED_ANGLE:
            circuit._ctrl_angle(gate, ints[1], ints[0], _angle(line))
        elif gate == Gate.M:
            circuit.add_m_gate(ints[0], ints[1])
        elif gate == Gate.U:
            circuit.add_u_gate(_matrix(lines, i), ints[0])
            i += 2
        else:
            circuit.add_cu_gate(_matrix(lines, i), ints[1], ints[0])
            i += 2
    return circuit, i


def read_tangelo_circuit(n_qubits: int, source: str | os.PathLike | TextIO, n_skip_lines: int = 0) -> QuantumCircuit:
    """``source``: a path, or an open text stream; the first ``n_skip_lines`` lines are skipped.
    Tangelo's CNOT / CPHASE / PHASE are read as CX / CP / P, SWAP as three CX gates; an unknown
    gate name raises RuntimeError."""
    if isinstance(source, (str, os.PathLike)):
        try:
            with open(source, encoding="utf-8") as f:
                text = f.read()
        except FileNotFoundError as e:
            raise OSError(f"ERROR: unable to read tangelo circuit from : '{os.fspath(source)}'\n") from e
    else:
        text = source.read()
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    circuit, _ = _read_block(int(n_qubits), lines, int(n_skip_lines), None)
    return circuit


def _matrix_element(x: float) -> str:
    return format(x, " .17f")[:16]            # write_tangelo_file.cpp:24-32


def _matrix_lines(m: Matrix2X2) -> str:
    def pair(z: complex) -> str:
        return f"[{_matrix_element(z.real)}, {_matrix_element(z.imag)}]"
    return f"    {pair(m.elem00)}   {pair(m.elem01)}\n    {pair(m.elem10)}   {pair(m.elem11)}\n"


def _predicate_text(p: ControlFlowPredicate) -> str:
    sign = " == " if p.control_kind == ControlFlowBooleanKind.IF else " != "
    def lst(v):
        return "[" + ", ".join(str(int(x)) for x in v) + "]"
    return "IF BITS" + lst(p.bit_indices_to_check) + sign + lst(p.expected_bits)


def _write(circuit: QuantumCircuit, out: list[str], indent: int) -> None:
    pad = " " * indent
    for el in circuit.elements:
        if isinstance(el, _LogEl):
            continue
        if isinstance(el, _IfEl):
            out.append(_predicate_text(el.predicate) + "\n")
            _write(el.body, out, CONTROL_FLOW_INDENT)
            if el.else_body is not None:
                out.append("ELSE\n")
                _write(el.else_body, out, CONTROL_FLOW_INDENT)
            continue
        assert isinstance(el, _GateEl)
        g = el.gate
        head = pad + f"{g.name:<10}"
        # a parameterised gate carries the placeholder angle 0.0, as in the reference (gate_create.cpp:27,84)
        angle = 0.0 if el.param is not None else el.angle
        if g in ONE_TARGET:
            out.append(f"{head}target : [{el.q0}]\n")
        elif g in ONE_TARGET_ANGLE:
            out.append(f"{head}target : [{el.q0}]   parameter : {angle:.16f}\n")
        elif g in CONTROLLED:
            out.append(f"{head}target : [{el.q1}]   control : [{el.q0}]\n")
        elif g in CONTROLLED_ANGLE:
            out.append(f"{head}target : [{el.q1}]   control : [{el.q0}]   parameter : {angle:.16f}\n")
        elif g == Gate.M:
            out.append(f"{head}target : [{el.q0}]   bit : [{el.q1}]\n")
        elif g == Gate.U:
            out.append(f"{head}target : [{el.q0}]\n" + _matrix_lines(el.matrix))
        else:
            out.append(f"{head}target : [{el.q1}]   control : [{el.q0}]\n" + _matrix_lines(el.matrix))


def write_tangelo_circuit(circuit: QuantumCircuit, dest: str | os.PathLike | TextIO | None = None) -> str:
    """Writes to ``dest`` (a path or a text stream) and returns the text; circuit loggers are skipped."""
    out: list[str] = []
    _write(circuit, out, 0)
    text = "".join(out)
    if isinstance(dest, (str, os.PathLike)):
        try:
            with open(dest, "w", encoding="utf-8") as f:
                f.write(text)
        except OSError as e:
            raise OSError(f"ERROR: unable to open file to write tangelo-style circuit: '{os.fspath(dest)}'\n") from e
    elif dest is not None:
        dest.write(text)
    return text


# ---- statevector snapshots -------------------------------------------------------------------------------

class Endian(enum.Enum):
    """kettle's ``Endian`` (include/kettle/state/endian.hpp)."""

    LITTLE = "LITTLE"
    BIG = "BIG"


def bit_reversal(n_qubits: int) -> np.ndarray:
    """Index i -> i with its n_qubits low bits reversed (kettle's ``endian_flip``)."""
    idx = np.arange(1 << n_qubits, dtype=np.int64)
    out = np.zeros_like(idx)
    for b in range(n_qubits):
        out |= ((idx >> b) & 1) << (n_qubits - 1 - b)
    return out


def _qubits_of(n_states: int) -> int:
    n = int(n_states).bit_length() - 1
    if n_states < 2 or (1 << n) != n_states:
        raise RuntimeError("The provided coefficients must have a size equal to a power of 2.\n")
    return n


def _snapshot_number(x: float) -> str:
    return (" " if x >= 0.0 else "") + f"{x:.14f}"      # statevector.cpp:45-66


def statevector_text(amps: np.ndarray, endian: Endian = Endian.LITTLE) -> str:
    """The text ``save_statevector`` writes for these (little-endian indexed) amplitudes."""
    a = np.asarray(amps, dtype=np.complex128).reshape(-1)
    n = _qubits_of(a.size)
    if endian == Endian.BIG:
        a = a[bit_reversal(n)]
    lines = [f"ENDIANNESS: {endian.value}", f"NUMBER OF STATES: {a.size}"]
    lines += [f"{_snapshot_number(z.real)}   {_snapshot_number(z.imag)}" for z in a.tolist()]
    return "\n".join(lines) + "\n"


def parse_statevector_text(text: str) -> np.ndarray:
    """Little-endian indexed amplitudes of a ``save_statevector`` text."""
    words = text.split()
    if len(words) < 6:
        raise RuntimeError("ERROR: statevector text is truncated\n")
    try:
        endian = Endian(words[1])
    except ValueError:
        raise RuntimeError("UNREACHABLE: dev error in `string_to_endian()`") from None
    n_states = int(words[5])
    vals = np.array(words[6:6 + 2 * n_states], dtype=np.float64)
    if vals.size != 2 * n_states:
        raise RuntimeError("ERROR: statevector text is truncated\n")
    amps = vals[0::2] + 1j * vals[1::2]
    if endian == Endian.BIG:
        amps = amps[bit_reversal(_qubits_of(n_states))]
    return amps


_FLOAT = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
_NUMPY_COMPLEX = re.compile(r"\(\s*(" + _FLOAT + r")\s*(" + _FLOAT + r")j\s*\)")


def parse_numpy_statevector(text: str, input_endian: Endian = Endian.LITTLE) -> np.ndarray:
    """Little-endian indexed amplitudes of a numpy text dump: n on the first line, then 2^n lines
    ``(re+imj)``."""
    head, _, body = text.partition("\n")
    n = int(head.split()[0])
    pairs = _NUMPY_COMPLEX.findall(body)
    if len(pairs) < (1 << n):
        raise RuntimeError("ERROR: numpy statevector text is truncated\n")
    amps = np.array([complex(float(r), float(i)) for r, i in pairs[:1 << n]], dtype=np.complex128)
    if input_endian == Endian.BIG:
        amps = amps[bit_reversal(n)]
    return amps


def _read_text(source, what: str) -> str:
    if isinstance(source, (str, os.PathLike)):
        try:
            with open(source, encoding="utf-8") as f:
                return f.read()
        except FileNotFoundError as e:
            raise OSError(f"ERROR: unable to open file {what}: \n'{os.fspath(source)}'\n") from e
    return source.read()


def save_statevector(dest: str | os.PathLike | TextIO, state, endian: Endian = Endian.LITTLE) -> None:
    """``state``: a device Statevector (downloaded first) or an array of amplitudes."""
    from .api import Statevector

    amps = state.download() if isinstance(state, Statevector) else state
    text = statevector_text(amps, endian)
    if isinstance(dest, (str, os.PathLike)):
        try:
            with open(dest, "w", encoding="utf-8") as f:
                f.write(text)
        except OSError as e:
            raise OSError(f"ERROR: unable to open file to save statevector: \n'{os.fspath(dest)}'\n") from e
    else:
        dest.write(text)


def load_statevector(source: str | os.PathLike | TextIO, precision: int | None = None):
    """A device Statevector holding the snapshot's amplitudes."""
    from .api import KQ_C128, Statevector

    amps = parse_statevector_text(_read_text(source, "to load statevector"))
    return Statevector(amps, KQ_C128 if precision is None else precision)


def read_numpy_statevector(source: str | os.PathLike | TextIO, input_endian: Endian = Endian.LITTLE,
                           precision: int | None = None):
    from .api import KQ_C128, Statevector

    amps = parse_numpy_statevector(_read_text(source, "for statevector"), input_endian)
    return Statevector(amps, KQ_C128 if precision is None else precision)
