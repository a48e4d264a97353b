"""Tangelo-style circuit files (read_tangelo_circuit / write_tangelo_circuit).

Known answers transcribed from the reference's test/source/io/read_tangelo_file_test.cpp:28-267 and
write_tangelo_file_test.cpp:14-179; expected matrix-element text comes from Python's own correctly
rounded formatting of the same values (the reference cuts "% .17f" output to 16 characters).  The
device test reads a controlled-rotation stream of the shape the reference's QPE drivers feed
(qpe_tangelo_project/qpe_tangelo_simulate.cpp:63) and compares libkq with the oracle."""
from __future__ import annotations

import io
import math

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd.circuit import Gate, Matrix2X2, _GateEl, _IfEl
from kettle_amd.io import read_tangelo_circuit, write_tangelo_circuit
from kat_cases import basis
from oracle import oracle as orc


def gate(c: ka.QuantumCircuit, i: int) -> _GateEl:
    el = c.elements[i]
    assert isinstance(el, _GateEl)
    return el


def test_single_and_multiple_gates():
    c = read_tangelo_circuit(9, io.StringIO("H         target : [4]   \n"), 0)
    assert c.n_circuit_elements() == 1 and gate(c, 0).gate == Gate.H and gate(c, 0).q0 == 4

    text = ("Circuit object. Size 339200                            \n"
            "                                                       \n"
            "H         target : [4]                                 \n"
            "RX        target : [5]   parameter : 1.5707963267948966\n"
            "CNOT      target : [4]   control : [2]                 \n"
            "RZ        target : [5]   parameter : 12.533816585267923\n")
    c = read_tangelo_circuit(9, io.StringIO(text), 2)
    assert c.n_circuit_elements() == 4
    assert (gate(c, 1).gate, gate(c, 1).q0, gate(c, 1).angle) == (Gate.RX, 5, 1.5707963267948966)
    assert (gate(c, 2).gate, gate(c, 2).q0, gate(c, 2).q1) == (Gate.CX, 2, 4)
    assert (gate(c, 3).gate, gate(c, 3).angle) == (Gate.RZ, 12.533816585267923)


def test_swap_phase_and_controlled_angles():
    c = read_tangelo_circuit(13, io.StringIO("SWAP      target : [12, 9]\n"), 0)
    g = [gate(c, i) for i in range(3)]
    assert c.n_circuit_elements() == 3 and all(x.gate == Gate.CX for x in g)
    assert g[0].q0 == g[1].q1 == g[2].q0 and g[0].q1 == g[1].q0 == g[2].q1

    c = read_tangelo_circuit(13, io.StringIO("PHASE     target : [11]   parameter : -1.3474016644659843\n"), 0)
    assert (gate(c, 0).gate, gate(c, 0).q0, gate(c, 0).angle) == (Gate.P, 11, -1.3474016644659843)

    for text, want in [("CPHASE    target : [9]   control : [12]   parameter : -0.39269908169872414\n", Gate.CP),
                       ("CRX    target : [9]   control : [12]   parameter : -0.39269908169872414\n", Gate.CRX),
                       ("CRY    target : [9]   control : [12]   parameter : -0.39269908169872414\n", Gate.CRY),
                       ("CRZ    target : [9]   control : [12]   parameter : -0.39269908169872414\n", Gate.CRZ)]:
        c = read_tangelo_circuit(13, io.StringIO(text), 0)
        g = gate(c, 0)
        assert (g.gate, g.q0, g.q1, g.angle) == (want, 12, 9, -0.39269908169872414)


def test_measure_and_unitaries():
    c = read_tangelo_circuit(13, io.StringIO("M         target : [1]   bit : [4]\n"), 0)
    assert (gate(c, 0).gate, gate(c, 0).q0, gate(c, 0).q1) == (Gate.M, 1, 4)
    rows = ("    [1.234, -4.321]   [2.345, -5.432]\n"
            "    [3.456, -6.543]   [4.567, -7.654]\n")
    want = np.array([[1.234 - 4.321j, 2.345 - 5.432j], [3.456 - 6.543j, 4.567 - 7.654j]])
    c = read_tangelo_circuit(10, io.StringIO("U         target : [1]\n" + rows), 0)
    assert c.n_circuit_elements() == 1 and gate(c, 0).gate == Gate.U and gate(c, 0).q0 == 1
    np.testing.assert_array_equal(gate(c, 0).matrix.as_array(), want)
    c = read_tangelo_circuit(10, io.StringIO("CU        target : [1]   control : [2]\n" + rows), 0)
    g = gate(c, 0)
    assert (g.gate, g.q0, g.q1) == (Gate.CU, 2, 1)
    np.testing.assert_array_equal(g.matrix.as_array(), want)


def test_reader_errors(tmp_path):
    with pytest.raises(RuntimeError, match="Unknown gate found"):
        read_tangelo_circuit(3, io.StringIO("FOO       target : [1]\n"), 0)
    with pytest.raises(RuntimeError, match="'ELSE' statement"):
        read_tangelo_circuit(3, io.StringIO("ELSE\n    X         target : [0]\n"), 0)
    with pytest.raises(RuntimeError, match="comparison sign"):
        read_tangelo_circuit(3, io.StringIO("IF BITS[0] <= [1]\n    X         target : [0]\n"), 0)
    with pytest.raises(OSError, match="unable to read tangelo circuit"):
        read_tangelo_circuit(3, tmp_path / "missing.txt", 0)


def test_writer_line_formats():
    c = ka.QuantumCircuit(6)
    value = 1.1234567812345677
    c.add_h_gate(5)
    c.add_sx_gate(0)
    c.add_csx_gate(3, 0)
    c.add_ch_gate(3, 5)
    c.add_rx_gate(1, value)
    c.add_p_gate(0, -value)
    c.add_crx_gate(3, 1, value)
    c.add_cp_gate(1, 0, -value)
    c.add_m_gate(1, 3)
    assert write_tangelo_circuit(c) == (
        "H         target : [5]\n"
        "SX        target : [0]\n"
        "CSX       target : [0]   control : [3]\n"
        "CH        target : [5]   control : [3]\n"
        "RX        target : [1]   parameter : 1.1234567812345677\n"
        "P         target : [0]   parameter : -1.1234567812345677\n"
        "CRX       target : [1]   control : [3]   parameter : 1.1234567812345677\n"
        "CP        target : [0]   control : [1]   parameter : -1.1234567812345677\n"
        "M         target : [1]   bit : [3]\n")

    u = ka.QuantumCircuit(2)
    u.add_u_gate(Matrix2X2(1.234 - 4.321j, 2.345 - 5.432j, 3.456 - 6.543j, 4.567 - 7.654j), 1)
    assert write_tangelo_circuit(u) == (
        "U         target : [1]\n"
        "    [ 1.2339999999999, -4.3209999999999]   [ 2.3450000000000, -5.4320000000000]\n"
        "    [ 3.4559999999999, -6.5430000000000]   [ 4.5670000000000, -7.6539999999999]\n")

    plain, logged = ka.QuantumCircuit(2), ka.QuantumCircuit(2)
    for circ in (plain, logged):
        circ.add_x_gate(0)
        if circ is logged:
            circ.add_classical_register_circuit_logger()
        circ.add_h_gate(1)
    assert write_tangelo_circuit(plain) == write_tangelo_circuit(logged)


def test_parameterised_gate_written_with_placeholder_angle():
    c = ka.QuantumCircuit(2)
    c.add_ry_gate(0, 0.75, parameterized=True)
    assert write_tangelo_circuit(c) == "RY        target : [0]   parameter : 0.0000000000000000\n"


def test_round_trip_with_control_flow(tmp_path):
    def sub(build):
        s = ka.QuantumCircuit(3)
        build(s)
        return s

    original = ka.QuantumCircuit(3)
    original.add_x_gate([0, 1])
    original.add_h_gate([0, 1, 2])
    original.add_m_gate([0, 1])
    original.add_if_statement(0, sub(lambda s: s.add_x_gate([0, 2])))
    original.add_y_gate(0)
    original.add_z_gate(1)
    original.add_if_else_statement(1, sub(lambda s: (s.add_x_gate(1), s.add_h_gate(2))),
                                   sub(lambda s: (s.add_cx_gate(1, 2), s.add_h_gate(2))))
    original.add_if_statement(ka.ControlFlowPredicate([0, 1], [1, 0], ka.ControlFlowBooleanKind.IF_NOT),
                              sub(lambda s: s.add_crz_gate(2, 0, 0.375)))
    original.add_y_gate(0)
    original.add_u_gate(Matrix2X2(0.6, 0.8j, 0.8j, 0.6), 2)
    path = tmp_path / "circuit.txt"
    text = write_tangelo_circuit(original, path)
    assert "IF BITS[0, 1] != [1, 0]\n    CRZ       target : [0]   control : [2]   parameter : 0.3750000000000000\n" in text
    back = read_tangelo_circuit(3, path, 0)
    ops_a, preds_a, _ = original.flatten()
    ops_b, preds_b, _ = back.flatten()
    assert len(ops_a) == len(ops_b) and np.array_equal(preds_a, preds_b)
    for a, b in zip(ops_a, ops_b):
        assert (a["gate"], a["q0"], a["q1"]) == (b["gate"], b["q0"], b["q1"])
        np.testing.assert_allclose(a["u"], b["u"], atol=1e-12, rtol=0)   # matrix text is cut to 13 decimals
    assert isinstance(back.elements[10], _IfEl) and back.elements[10].else_body is not None


def qpe_like_text(n_ancilla: int, n_unitary: int, n_steps: int, seed: int) -> str:
    """A Trotterised controlled-evolution stream in tangelo's own names (CNOT, CPHASE, PHASE)."""
    rng = np.random.default_rng(seed)
    n = n_ancilla + n_unitary
    lines = [f"H         target : [{a}]   " for a in range(n_ancilla)]
    for a in range(n_ancilla):
        for _ in range(n_steps * (1 << min(a, 2))):
            for q in range(n_ancilla, n - 1):
                lines.append(f"CNOT      target : [{q + 1}]   control : [{q}]   ")
                lines.append(f"CRZ       target : [{q + 1}]   control : [{a}]   parameter : {rng.uniform(-4, 4)!r}")
                lines.append(f"CNOT      target : [{q + 1}]   control : [{q}]   ")
            q = int(rng.integers(n_ancilla, n))
            lines.append(f"CRX       target : [{q}]   control : [{a}]   parameter : {rng.uniform(-4, 4)!r}")
            lines.append(f"PHASE     target : [{a}]   parameter : {rng.uniform(-4, 4)!r}")
    for a in range(n_ancilla // 2):
        lines.append(f"SWAP      target : [{a}, {n_ancilla - 1 - a}]")
    for j in range(n_ancilla):
        for k in range(j):
            lines.append(f"CPHASE    target : [{j}]   control : [{k}]   parameter : {-math.pi / (1 << (j - k))!r}")
        lines.append(f"H         target : [{j}]   ")
    return "\n".join(lines) + "\n"


def test_qpe_like_stream_reads_to_the_intended_gates():
    text = qpe_like_text(3, 4, 1, 0)
    c = read_tangelo_circuit(7, io.StringIO(text), 0)
    names = [line.split()[0] for line in text.splitlines()]
    n_expected = sum(3 if nm == "SWAP" else 1 for nm in names)
    assert c.n_circuit_elements() == n_expected
    assert {gate(c, i).gate for i in range(n_expected)} == {Gate.H, Gate.CX, Gate.CRZ, Gate.CRX, Gate.P, Gate.CP}


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(ka.KQ_C128, 1e-12), (ka.KQ_C64, 2e-5)], ids=["c128", "c64"])
def test_qpe_like_file_simulates_like_the_oracle(tmp_path, precision, tol):
    n_anc, n_uni = 5, 9
    n = n_anc + n_uni
    path = tmp_path / "qpe.txt"
    path.write_text(qpe_like_text(n_anc, n_uni, 2, 1))
    c = read_tangelo_circuit(n, path, 0)
    init = orc.zero_state(n)
    sv = ka.Statevector(n, precision)
    ka.simulate(c, sv)
    ops, _, _ = c.flatten()
    ref, _ = orc.sv_simulate(init, n, ops)
    assert np.max(np.abs(sv.download() - ref)) <= tol
    assert abs(np.sum(np.abs(ref) ** 2) - 1.0) <= 1e-12
