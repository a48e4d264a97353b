"""Full-size checks of the 30- and 33-qubit circuits (the large-statevector workload at its real sizes:
30 qubits complex128 = 16 GiB, 33 qubits complex64 = 64 GiB), through size-independent properties, since
the oracle cannot run them and the states are never copied to the host:

- the circuit followed by its inverse (every gate replaced by its conjugate transpose, in reverse
  order) returns |0...0>: <Z_q> = 1 on every qubit and the norm is 1;
- the QFT of |0...0> is the uniform superposition: <X_q> = 1 and <Z_q> = 0 on every qubit.

Expectation values are FP64 reductions on the device.  Tolerances: 1e-9 (complex128, ~2800 gates)
and 2e-3 (complex64: FP32 rounding over ~2800 gates)."""
from __future__ import annotations

import math

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


def reduced_spectrum(psi: np.ndarray, n: int, keep: list[int]) -> np.ndarray:
    """Eigenvalues of the reduced density matrix of |psi> on `keep` (independent of index order)."""
    idx = np.arange(1 << n)
    a = np.zeros(1 << n, dtype=np.int64)
    for j, q in enumerate(keep):
        a |= ((idx >> q) & 1) << j
    rest_bits = [q for q in range(n) if q not in keep]
    r = np.zeros(1 << n, dtype=np.int64)
    for j, q in enumerate(rest_bits):
        r |= ((idx >> q) & 1) << j
    m = np.zeros((1 << len(keep), 1 << len(rest_bits)), dtype=np.complex128)
    m[a, r] = psi
    return np.sort(np.linalg.eigvalsh(m @ m.conj().T))


@pytest.mark.parametrize("n", [12, 14])
def test_density_matrix_full_size(n):
    """Density matrices of 12 and 14 qubits (vec(rho) of 2^24 / 2^28 amplitudes): the depth-30 random
    circuit on |0..0><0..0| against the statevector path on the same circuit (itself checked against the
    oracle): trace 1, diagonal = |psi|^2, and the spectrum of the 2-qubit partial trace equals that of
    |psi>'s reduced state (1e-12); with mid-circuit measurement and if-else every 10 layers, trace 1 and a
    Hermitian, positive 2-qubit partial trace."""
    c = workloads.random_circuit(n, 30, seed=n)
    dm = ka.DensityMatrix(n)
    ka.simulate(c, dm)
    sv = ka.Statevector(n)
    ka.simulate(c, sv)
    psi = sv.download()
    assert abs(dm.trace() - 1.0) <= 1e-12
    assert abs(dm.trace_of_square() - 1.0) <= 1e-12 and dm.is_pure()
    p = ka.calculate_probabilities_raw(dm)
    assert np.max(np.abs(p - np.abs(psi) ** 2)) <= 1e-12
    keep = [0, n - 1]
    pt = ka.partial_trace(dm, [q for q in range(n) if q not in keep])
    assert np.max(np.abs(np.sort(np.linalg.eigvalsh(pt)) - reduced_spectrum(psi, n, keep))) <= 1e-12
    cm = workloads.density_matrix_circuit(n=n, depth=30, seed=n, measure_every=10)
    dm2 = ka.DensityMatrix(n)
    res = ka.simulate(cm, dm2, 5)
    assert res.classical_register[0] in (0, 1)
    assert abs(dm2.trace() - 1.0) <= 1e-12
    pt2 = ka.partial_trace(dm2, [q for q in range(n) if q not in keep])
    assert np.max(np.abs(pt2 - pt2.conj().T)) <= 1e-12
    assert np.min(np.linalg.eigvalsh(pt2)) >= -1e-12


@pytest.mark.parametrize("n", [12, 14])
@pytest.mark.parametrize("seed", [5, 6])
def test_density_matrix_measurements_replayed_on_statevector(n, seed):
    """The depth-30 density-matrix circuit with M on qubit 0 and an if-else every 10 layers, each outcome
    kept in its own bit: the statevector path replays the recorded outcome history (gates, collapse onto
    the recorded outcome renormalised by 1 / sqrt(p), the branch the outcome selects), and rho must be
    |psi><psi| of the replayed state: its diagonal is |psi|^2, it is pure, and its 2-qubit partial trace
    has |psi>'s reduced spectrum (1e-12).  A wrong branch or collapse scale fails these."""
    from kettle_amd import _lib
    cm = workloads.density_matrix_circuit(n=n, depth=30, seed=n, measure_every=10, distinct_bits=True)
    dm = ka.DensityMatrix(n)
    res = ka.simulate(cm, dm, seed)
    outcomes = res.classical_register[:3]
    assert all(b in (0, 1) for b in outcomes)
    lib = _lib.load()
    sv = ka.Statevector(n)
    segs = workloads.density_matrix_segments(n=n, depth=30, seed=n, measure_every=10)
    for k, seg in enumerate(segs):
        ka.simulate(seg, sv)
        if k < len(outcomes):
            b = outcomes[k]
            w = np.zeros(2)
            _lib.check(lib.kq_qubit_weights(sv.handle, 0, _lib.ptr(w)))
            _lib.check(lib.kq_collapse(sv.handle, 0, int(b), float(1.0 / np.sqrt(w[b]))))
            branch = ka.QuantumCircuit(n)
            if b == 1:
                branch.add_x_gate(0)
            else:
                branch.add_rz_gate(0, math.pi / 4)
            ka.simulate(branch, sv)
    psi = sv.download()
    assert abs(np.linalg.norm(psi) - 1.0) <= 1e-12
    assert abs(dm.trace() - 1.0) <= 1e-12
    assert abs(dm.trace_of_square() - 1.0) <= 1e-12
    p = ka.calculate_probabilities_raw(dm)
    assert np.max(np.abs(p - np.abs(psi) ** 2)) <= 1e-12
    keep = [0, n - 1]
    pt = ka.partial_trace(dm, [q for q in range(n) if q not in keep])
    assert np.max(np.abs(np.sort(np.linalg.eigvalsh(pt)) - reduced_spectrum(psi, n, keep))) <= 1e-12
