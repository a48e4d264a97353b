"""Full-size checks of the 30- and 33-qubit circuits (the large-statevector workload at its real sizes:
30 qubits complex128 = 16 GiB, 33 qubits complex64 = 64 GiB), through size-independent properties, since
the oracle cannot run them and the states are never copied to the host:

- the circuit followed by its inverse (every gate replaced by its conjugate transpose, in reverse
  order) returns |0...0>: <Z_q> = 1 on every qubit and the norm is 1;
- the QFT of |0...0> is the uniform superposition: <X_q> = 1 and <Z_q> = 0 on every qubit.

Expectation values are FP64 reductions on the device.  Tolerances: 1e-9 (complex128, ~2800 gates)
and 2e-3 (complex64: FP32 rounding over ~2800 gates)."""
from __future__ import annotations

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd import workloads
from kettle_amd.circuit import Gate
from test_plan import CONTROLLED, as_u_op, op_matrix, to_ops

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _tuned():
    ka.set_fusion_autotune(True)
    yield
    ka.set_fusion_autotune(True)


def inverse_ops(ops: np.ndarray) -> np.ndarray:
    """The inverse circuit: reverse order, each gate as U / CU with the conjugate-transpose matrix."""
    recs = []
    for op in ops[::-1]:
        g = Gate(int(op["gate"]))
        ctrl = g in CONTROLLED
        t = int(op["q1"] if ctrl else op["q0"])
        c = int(op["q0"]) if ctrl else None
        recs.append(as_u_op(op_matrix(op).conj().T, t, c))
    return to_ops(recs)


def mean_pauli(sv: ka.Statevector, n: int, kind: ka.PauliTerm) -> float:
    op = ka.PauliOperator(n_qubits=n)
    for q in range(n):
        op.add(1.0 / n, ka.SparsePauliString([(q, int(kind))], n))
    return ka.expectation_value(op, sv).real


def norm_squared(sv: ka.Statevector, n: int) -> float:
    op = ka.PauliOperator(n_qubits=n)
    op.add(1.0, ka.SparsePauliString([], n))
    return ka.expectation_value(op, sv).real


@pytest.mark.parametrize("n,precision,tol", [(30, ka.KQ_C128, 1e-9), (33, ka.KQ_C64, 2e-3)], ids=["30q-c128", "33q-c64"])
def test_circuit_then_inverse_returns_zero_state(n, precision, tol):
    ops, _, _ = workloads.qft_plus_layers(n=n, layers=20).flatten()
    sv = ka.Statevector(n, precision)
    ka.apply_gates(sv, ops)
    ka.apply_gates(sv, inverse_ops(ops))
    assert abs(norm_squared(sv, n) - 1.0) <= tol
    assert abs(mean_pauli(sv, n, ka.PauliTerm.Z) - 1.0) <= tol


@pytest.mark.parametrize("n,precision,tol", [(30, ka.KQ_C128, 1e-10), (33, ka.KQ_C64, 1e-4)], ids=["30q-c128", "33q-c64"])
def test_qft_of_zero_state_is_uniform(n, precision, tol):
    c = ka.QuantumCircuit(n)
    c.add_qft_gate(list(range(n)))
    sv = ka.Statevector(n, precision)
    ka.simulate(c, sv)
    assert abs(mean_pauli(sv, n, ka.PauliTerm.X) - 1.0) <= tol
    assert abs(mean_pauli(sv, n, ka.PauliTerm.Z)) <= tol
