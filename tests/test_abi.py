"""CPU-side checks of the C-ABI boundary and the host logic (no device compute)."""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd import _lib
from kettle_amd.circuit import Gate, OP_IF, OP_IF_ELSE, OP_LOG_CREG, OP_LOG_STATE

ROOT = Path(__file__).resolve().parent.parent


def declared_functions() -> set[str]:
    text = (ROOT / "include" / "kq.h").read_text()
    return set(re.findall(r"^\s*(?:int|const char\*)\s+(kq_\w+)\s*\(", text, flags=re.M))


def test_header_and_binding_agree():
    assert declared_functions() == set(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (kq_\w+)", out))
    assert declared_functions() <= exported


def test_library_targets_gfx950():
    """The device code object inside libkq.so is built for gfx950."""
    data = _lib.LIB_PATH.read_bytes()
    assert b"gfx950" in data


def test_op_record_layout():
    """kq_op in numpy matches the C struct: 4 int32, 1 double, 8 doubles = 88 bytes."""
    assert _lib.OP_DTYPE.itemsize == 88
    assert _lib.OP_DTYPE.fields["angle"][1] == 16
    assert _lib.OP_DTYPE.fields["u"][1] == 24


def test_flatten_layout_of_control_flow():
    c = ka.QuantumCircuit(3)
    c.add_h_gate(0)
    c.add_m_gate(0)
    sub = ka.QuantumCircuit(3)
    sub.add_x_gate(1)
    sub.add_cx_gate(1, 2)
    els = ka.QuantumCircuit(3)
    els.add_rz_gate(2, 0.25)
    c.add_if_else_statement(ka.ControlFlowPredicate([0], [1], ka.ControlFlowBooleanKind.IF_NOT), sub, els)
    c.add_statevector_circuit_logger()
    c.add_classical_register_circuit_logger()
    ops, preds, _ = c.flatten()
    assert [int(g) for g in ops["gate"]] == [Gate.H, Gate.M, OP_IF_ELSE, Gate.X, Gate.CX, Gate.RZ, OP_LOG_STATE, OP_LOG_CREG]
    head = ops[2]
    assert (head["q0"], head["q1"], head["aux"]) == (0, 2, 1)
    assert preds.tolist() == [1, 1, 0, 1]
    assert (ops[4]["q0"], ops[4]["q1"]) == (1, 2)
    assert ops[5]["angle"] == 0.25


def test_parameters_flatten_and_uninitialised_error():
    c = ka.QuantumCircuit(2)
    pid = c.add_rx_gate(0, 0.5, True)
    c.add_crz_gate(0, 1, pid)
    other = ka.ParameterID()
    c.add_ry_gate(1, other)
    with pytest.raises(RuntimeError, match="uninitialized parameter"):
        c.flatten()
    c.set_parameter_value(other, 1.5)
    ops, _, order = c.flatten()
    assert order == [pid, other]
    assert ops["aux"].tolist() == [0, 0, 1]
    assert ops["angle"].tolist() == [0.5, 0.5, 1.5]
    with pytest.raises(IndexError):
        c.set_parameter_value("missing", 0.0)


def test_builder_validation_messages():
    c = ka.QuantumCircuit(2, 1)
    with pytest.raises(RuntimeError, match="is out of bounds"):
        c.add_h_gate(2)
    with pytest.raises(RuntimeError, match="bit index at which the 'M' gate"):
        c.add_m_gate(1)
    with pytest.raises(RuntimeError, match="Cannot swap"):
        c.add_swap_gate(1, 1)


def test_qft_expansion_counts():
    """add_qft_gate (circuit.cpp:850-872): n H, n(n-1)/2 CP, floor(n/2) swaps of 3 CX."""
    n = 7
    c = ka.QuantumCircuit(n)
    c.add_qft_gate(list(range(n)))
    ops, _, _ = c.flatten()
    gates = [int(g) for g in ops["gate"]]
    assert gates.count(Gate.H) == n
    assert gates.count(Gate.CP) == n * (n - 1) // 2
    assert gates.count(Gate.CX) == 3 * (n // 2)
    first_cp = ops[1]
    assert (first_cp["q0"], first_cp["q1"]) == (1, 0)
    assert first_cp["angle"] == pytest.approx(2 * np.pi / 4)


def test_n_local_order_matches_kettle():
    """n_local_test.cpp:250-282: per repetition the rotation blocks over all qubits, then the
    linear entanglement ladder; SkipLastRotationLayer drops the final rotation layer."""
    from kettle_amd.workloads import NLocalEntanglement, n_local

    c, ids = n_local(3, [Gate.RY, Gate.RZ], [Gate.CX], NLocalEntanglement.LINEAR, 2, False)
    ops, _, order = c.flatten()
    gates = [(Gate(int(o["gate"])).name, int(o["q0"]), int(o["q1"])) for o in ops]
    rot = [("RY", q, 0) for q in range(3)] + [("RZ", q, 0) for q in range(3)]
    ent = [("CX", 0, 1), ("CX", 1, 2)]
    assert gates == rot + ent + rot + ent + rot
    assert len(ids) == 18 and order == ids
    c2, ids2 = n_local(24, [Gate.RY, Gate.RZ], [Gate.CX], NLocalEntanglement.LINEAR, 20, True)
    assert len(ids2) == 960
    assert c2.n_circuit_elements() == 960 + 20 * 23


def test_state_creation_without_gpu_reports_hip_error():
    """Off the GPU the boundary returns a status and a message instead of crashing."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = ctypes.c_void_p()
    st = lib.kq_sv_create(ctypes.byref(h), 4, _lib.KQ_C128)
    assert st == _lib.KQ_ERR_HIP
    assert len(lib.kq_last_error()) > 0


def test_invalid_arguments_are_rejected_before_device_work():
    lib = _lib.load()
    assert lib.kq_set_fusion(20, 0, 0, 0.0) == _lib.KQ_ERR_INVALID
    assert b"LDS" in lib.kq_last_error()
    assert lib.kq_state_upload(None, None) == _lib.KQ_ERR_INVALID
    assert lib.kq_simulate(None, 1, None, 0, None, 0, 0, 0, None, 0, _lib.LOGGER_FN(), None) == _lib.KQ_ERR_INVALID
    assert lib.kq_sample_probabilities(None, 0, 1, 0, 0, None) == _lib.KQ_ERR_INVALID


def test_injected_fault_fires_once_after_the_given_calls():
    """kq_debug_set_faults: the call after the next fail_after_calls ones returns the forced status with
    its message, once; the setter itself never fails (host-only calls, no device needed)."""
    lib = _lib.load()
    try:
        assert lib.kq_debug_set_faults(0, _lib.KQ_ERR_RUNTIME, 2) == _lib.KQ_OK
        assert lib.kq_set_fusion_low_qubits(5) == _lib.KQ_OK
        assert lib.kq_set_fusion_low_qubits(5) == _lib.KQ_OK
        assert lib.kq_set_fusion_low_qubits(5) == _lib.KQ_ERR_RUNTIME
        assert lib.kq_last_error() == b"injected fault"
        assert lib.kq_set_fusion_low_qubits(5) == _lib.KQ_OK
        assert lib.kq_debug_set_faults(0, 0, -1) == _lib.KQ_ERR_INVALID
    finally:
        lib.kq_debug_set_faults(0, 0, 0)
        lib.kq_set_fusion_autotune(1)


def test_flatten_cache_follows_parameters_and_elements():
    """flatten() of a straight-line circuit reuses its arrays for the same elements and parameter order
    and refills only the parameterised angles; new values, appended and popped elements and another
    parameter order all give what a fresh flatten gives."""
    import numpy as np
    import kettle_amd as ka
    from kettle_amd import workloads

    w = workloads.vqe(n=6, layers=3, n_random=4, seed=3)
    c = w.circuit

    def fresh(order=None):
        c._flat_cache = None
        return c.flatten(order)[0]

    first = c.flatten(w.params)[0]
    assert np.array_equal(first, fresh(w.params))
    rng = np.random.default_rng(5)
    for pid in w.params:
        c.set_parameter_value(pid, float(rng.uniform(-3, 3)))
    again = c.flatten(w.params)[0]
    assert not np.array_equal(again["angle"], first["angle"]) and np.array_equal(again, fresh(w.params))
    c.add_rx_gate(2, 0.25)
    grown = c.flatten(w.params)[0]
    assert len(grown) == len(first) + 1 and np.array_equal(grown, fresh(w.params))
    c.pop_back()
    assert np.array_equal(c.flatten(w.params)[0], fresh(w.params))
    rev = list(reversed(w.params))
    assert np.array_equal(c.flatten(rev)[0], fresh(rev))
    assert not np.array_equal(c.flatten(rev)[0]["aux"], c.flatten(w.params)[0]["aux"])
