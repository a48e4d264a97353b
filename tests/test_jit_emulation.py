"""The run-time generated tile kernels executed on the host (tests/jit_emulation): one OS thread per
GPU thread and a real barrier per workgroup barrier, so races between threads of a workgroup show
up as wrong amplitudes.  Compared with the oracle: complex128 within 1e-12, complex64 within 2e-5.
Covers every generated-code feature (normalised one-qubit forms, deferred controlled X gates,
phase-angle merged runs, LDS remaps, qubit moves between passes and the restore passes)."""
from __future__ import annotations

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd import workloads
from kat_cases import basis, random_gate_circuit
from oracle import oracle as orc
from jit_emulation import emulate


@pytest.fixture(autouse=True)
def _restore():
    yield
    ka.set_fusion_relabel(True)
    ka.set_fusion_low_qubits(5)


def check(circuit, n, precision, tol):
    ops, _, _ = circuit.flatten()
    out = emulate.run(ops, n, basis("0" * n), precision)
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops)
    assert np.max(np.abs(out - ref)) <= tol


@pytest.mark.parametrize("relabel", [False, True], ids=["fixed", "moves"])
def test_layered_circuit_16q(relabel):
    ka.set_fusion_relabel(relabel)
    check(workloads.qft_plus_layers(n=16, layers=8), 16, ka.KQ_C128, 1e-12)


def test_layered_circuit_flipped_runs_16q():
    """Tiles of six low qubits and up to 16 stages: merged diagonal runs and one-qubit gates meet controlled
    X gates still pending on their register bit and run in the flipped (X G X) form."""
    ka.set_fusion_low_qubits(6)
    ka.set_fusion(-1, 0, 16, 0.0)
    try:
        check(workloads.qft_plus_layers(n=16, layers=12), 16, ka.KQ_C128, 1e-12)
    finally:
        ka.set_fusion(-1, 0, 12, 0.0)
        ka.set_fusion_autotune(True)


@pytest.mark.parametrize("n", [13, 15])
def test_random_gates_complex128(n):
    check(random_gate_circuit(n, 250, np.random.default_rng(9100 + n)), n, ka.KQ_C128, 1e-12)


def test_random_gates_complex64_low_four():
    ka.set_fusion_low_qubits(4)
    check(random_gate_circuit(14, 200, np.random.default_rng(9200)), 14, ka.KQ_C64, 2e-5)


@pytest.mark.parametrize("low,stages,layers", [(6, 16, 8), (4, 12, 12)], ids=["runs-of-64", "runs-of-8"])
def test_last_pass_stores_out_of_place_in_qubit_order(low, stages, layers):
    """The last pass of a gate run with qubit moves stores every qubit back on its own bit, out of place
    into a second buffer: with the tile holding the bits bound for index bits 0..5 (stores in runs of 64
    amplitudes), and with room for only bits 0..2 of them (runs of 8).  The emulated state, read in qubit
    order with no restore pass after it, equals the oracle's."""
    ka.set_fusion_low_qubits(low)
    ka.set_fusion(-1, 0, stages, 0.0)
    try:
        ops, _, _ = workloads.qft_plus_layers(n=16, layers=layers).flatten()
        passes = emulate._passes(np.ascontiguousarray(ops), 16, ka.KQ_C128)
        assert passes[-1]["to_other"] and not any(p["to_other"] for p in passes[:-1])
        out = emulate.run(ops, 16, basis("0" * 16), ka.KQ_C128)
    finally:
        ka.set_fusion(-1, 0, 12, 0.0)
        ka.set_fusion_autotune(True)
    ref, _ = orc.sv_simulate(basis("0" * 16), 16, ops)
    assert np.max(np.abs(out - ref)) <= 1e-12


def untouched_qubit_circuit(n, groups, layers, seed):
    """Random one-qubit gates and CX ladders on the qubits of `groups` only: the others keep their
    basis value through the run."""
    rng = np.random.default_rng(seed)
    c = ka.QuantumCircuit(n)
    qs = [q for lo, hi in groups for q in range(lo, hi)]
    for layer in range(layers):
        for q in qs:
            c.add_u_gate(workloads.random_unitary2x2(rng), q)
        for a, b in zip(qs[layer % 2::2], qs[layer % 2 + 1::2]):
            c.add_cx_gate(a, b)
    return c


@pytest.mark.parametrize("n,groups,layers,b,want_modes", [
    (17, [(0, 14)], 4, 0x1C001, [1, 2]),        # the second pass skips tiles in place
    (17, [(2, 16)], 6, 0x18003, [1, 2]),        # the second pass stores out of place, zeros where known
    (17, [(0, 5), (8, 17)], 4, 0xE0, [1, 0]),
], ids=["skip-in-place", "skip-out-of-place", "basis-then-load"])
def test_run_from_pending_basis_state(n, groups, layers, b, want_modes):
    """A run from a basis state not yet written: the buffer starts as NaN, the first pass builds the
    state from the basis index (reading nothing), later passes skip the tiles known to be zero; the
    result equals the oracle's run from |b> and holds no NaN."""
    ops, _, _ = untouched_qubit_circuit(n, groups, layers, n + layers).flatten()
    out, modes = emulate.run_from_basis(ops, n, b)
    assert modes == want_modes
    ref, _ = orc.sv_simulate(basis(format(b, f"0{n}b")[::-1]), n, ops)
    assert not np.isnan(out).any()
    assert np.max(np.abs(out - ref)) <= 1e-12
