"""Adjoint energy gradient (kq_energy_gradient) against the oracle's FP64 energies.

The gradient has no counterpart in kettle; it is pinned by mathematics on the oracle:
  - central finite differences of the oracle's energy (h = 1e-5, FP64),
  - the exact parameter-shift rule dE/dtheta = [E(theta + pi/2) - E(theta - pi/2)] / 2 for
    gates exp(-i theta G / 2) with G^2 = 1 (RX, RY, RZ, and P / CP up to a phase),
applied gate by gate and summed over the gates that share a parameter.
Tolerances: complex128 energy 1e-12, gradient 1e-8 (FD) / 1e-10 (shift);
complex64 energy 1e-6 (FP64 accumulation), gradient 1e-4."""
from __future__ import annotations

import math

import numpy as np
import pytest

import torch  # noqa: F401  (torch before libkq in the autograd tests)

import kettle_amd as ka
from kettle_amd import autograd as kag
from kettle_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SHIFTABLE = {int(ka.Gate.RX), int(ka.Gate.RY), int(ka.Gate.RZ), int(ka.Gate.P), int(ka.Gate.CP)}


def oracle_energy(ops, n, packed) -> float:
    amps, _ = orc.sv_simulate(orc.zero_state(n), n, ops)
    return orc.pauli_expectation(amps, n, packed).real


def oracle_fd(ops, n, packed, n_params, h=1e-5):
    g = np.zeros(n_params)
    for p in range(n_params):
        plus, minus = ops.copy(), ops.copy()
        sel = ops["aux"] == p
        plus["angle"][sel] += h
        minus["angle"][sel] -= h
        g[p] = (oracle_energy(plus, n, packed) - oracle_energy(minus, n, packed)) / (2 * h)
    return g


def oracle_shift(ops, n, packed, n_params):
    g = np.zeros(n_params)
    for i in np.nonzero(ops["aux"] >= 0)[0]:
        assert int(ops[i]["gate"]) in SHIFTABLE
        plus, minus = ops.copy(), ops.copy()
        plus["angle"][i] += math.pi / 2
        minus["angle"][i] -= math.pi / 2
        g[ops[i]["aux"]] += 0.5 * (oracle_energy(plus, n, packed) - oracle_energy(minus, n, packed))
    return g


def mixed_circuit(n: int, seed: int, controlled_rotations: bool):
    """Fixed gates of every class, parameterised RX/RY/RZ/P (and CRX/CRY/CRZ/CP), one parameter
    shared by two gates."""
    rng = np.random.default_rng(seed)
    c = ka.QuantumCircuit(n)
    ids = []
    for q in range(n):
        c.add_h_gate(q)
    for layer in range(3):
        for q in range(n):
            kind = ["rx", "ry", "rz", "p"][(q + layer) % 4]
            ids.append(getattr(c, f"add_{kind}_gate")(q, float(rng.uniform(-math.pi, math.pi)), True))
        for q in range(layer % 2, n - 1, 2):
            c.add_cx_gate(q, q + 1)
        c.add_sx_gate(layer % n)
        c.add_u_gate(workloads.random_unitary2x2(rng), (layer + 1) % n)
        kinds = ["crx", "cry", "crz", "cp"] if controlled_rotations else ["cp"]
        for j, kind in enumerate(kinds):
            a, b = (j + layer) % n, (j + layer + 2) % n
            ids.append(getattr(c, f"add_{kind}_gate")(a, b, float(rng.uniform(-math.pi, math.pi)), True))
    c.add_ry_gate(n - 1, ids[0])          # parameter 0 drives two gates
    return c, ids


def _packed(h):
    return h.packed()


@pytest.mark.parametrize("controlled", [False, True], ids=["shiftable", "controlled-rotations"])
def test_adjoint_complex128_vs_oracle(controlled):
    n = 6
    c, ids = mixed_circuit(n, 3, controlled)
    h = workloads.random_pauli_operator(n, 24, np.random.default_rng(4))
    ops, _, _ = c.flatten(ids)
    packed = _packed(h)
    r = ka.energy_and_gradient(c, h, ids, precision=ka.KQ_C128)
    e_ref = oracle_energy(ops, n, packed)
    assert abs(r.energy - e_ref) <= 1e-12
    fd = oracle_fd(ops, n, packed, len(ids))
    assert np.max(np.abs(r.gradient - fd)) <= 1e-8
    if not controlled:
        shift = oracle_shift(ops, n, packed, len(ids))
        assert np.max(np.abs(r.gradient - shift)) <= 1e-10


@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_adjoint_compiled_reverse_passes(precision):
    """Reverse passes as run-time compiled kernels (every tile pass compiled) against the oracle's energy and
    parameter shift and against the interpreted reverse passes: every gate class, parameterised RX / RY /
    RZ / P and controlled rotations, a shared parameter, checkpoints with restores forced on."""
    n = 8
    c, ids = mixed_circuit(n, 12, True)
    h = workloads.random_pauli_operator(n, 30, np.random.default_rng(13))
    ops, _, _ = c.flatten(ids)
    tol_e, tol_g = (1e-12, 1e-9) if precision == ka.KQ_C128 else (1e-6, 1e-4)
    ka.set_fusion(-1, 12, 0, 0.0)
    try:
        ka.set_jit(2)
        compiled = ka.energy_and_gradient(c, h, ids, precision=precision, checkpoint_every=10, drift_tol=0.0)
        ka.set_jit(0)
        interpreted = ka.energy_and_gradient(c, h, ids, precision=precision, checkpoint_every=10, drift_tol=0.0)
    finally:
        ka.set_jit(1)
        ka.set_fusion(-1, 4096, 12, 1.0e9)
        ka.set_fusion_autotune(True)
    e_ref = oracle_energy(ops, n, _packed(h))
    fd = oracle_fd(ops, n, _packed(h), len(ids))
    assert abs(compiled.energy - e_ref) <= tol_e
    assert np.max(np.abs(compiled.gradient - fd)) <= tol_g
    assert np.max(np.abs(compiled.gradient - interpreted.gradient)) <= (1e-12 if precision == ka.KQ_C128 else 1e-5)
    assert compiled.restores > 0


def test_adjoint_repeated_structure_new_angles():
    """The second call on the same gate structure with other angles takes the cached forward plan with the
    new coefficients: its energy matches the oracle's at the new angles and its gradient the device's
    complex128 parameter shift (20 qubits: compiled passes, qubit moves)."""
    n = 20
    w = workloads.vqe(n=n, layers=2, n_random=12, seed=4)
    ops0, _, _ = w.circuit.flatten(w.params)
    ka.energy_and_gradient(w.circuit, w.hamiltonian, w.params, precision=ka.KQ_C128)
    rng = np.random.default_rng(77)
    for pid in w.params:
        w.circuit.set_parameter_value(pid, float(rng.uniform(-math.pi, math.pi)))
    ops, _, _ = w.circuit.flatten(w.params)
    assert not np.array_equal(ops["angle"], ops0["angle"])
    r = ka.energy_and_gradient(w.circuit, w.hamiltonian, w.params, precision=ka.KQ_C128)
    assert abs(r.energy - oracle_energy(ops, n, _packed(w.hamiltonian))) <= 1e-12
    sv = ka.Statevector(n, ka.KQ_C128)

    def energy(o):
        sv.set_basis(0)
        ka.apply_gates(sv, o)
        return ka.expectation_value(w.hamiltonian, sv).real

    shift = np.zeros(len(w.params))
    for i in np.nonzero(ops["aux"] >= 0)[0]:
        plus, minus = ops.copy(), ops.copy()
        plus["angle"][i] += math.pi / 2
        minus["angle"][i] -= math.pi / 2
        shift[ops[i]["aux"]] += 0.5 * (energy(plus) - energy(minus))
    assert np.max(np.abs(r.gradient - shift)) <= 1e-10


def test_adjoint_complex64_error_bounds():
    n = 8
    c, ids = mixed_circuit(n, 5, True)
    h = workloads.random_pauli_operator(n, 40, np.random.default_rng(6))
    ops, _, _ = c.flatten(ids)
    e_ref = oracle_energy(ops, n, _packed(h))
    r64 = ka.energy_and_gradient(c, h, ids, precision=ka.KQ_C64)
    r128 = ka.energy_and_gradient(c, h, ids, precision=ka.KQ_C128)
    assert abs(r64.energy - e_ref) <= 1e-6
    assert np.max(np.abs(r64.gradient - r128.gradient)) <= 1e-4


def test_checkpoint_restores_keep_the_gradient():
    """drift_tol = 0 restores psi at every checkpoint where uncomputation drifted at all; the
    gradient must agree with the run that never restores.  Passes of at most 8 gates, so that the
    checkpoints (at pass boundaries) are many."""
    n = 7
    c, ids = mixed_circuit(n, 8, True)
    h = workloads.random_pauli_operator(n, 16, np.random.default_rng(9))
    ka.set_fusion(-1, 8, 0, 0.0)
    try:
        base = ka.energy_and_gradient(c, h, ids, precision=ka.KQ_C128, checkpoint_every=5, drift_tol=1.0)
        forced = ka.energy_and_gradient(c, h, ids, precision=ka.KQ_C128, checkpoint_every=5, drift_tol=0.0)
        none = ka.energy_and_gradient(c, h, ids, precision=ka.KQ_C128, checkpoint_every=10**6, drift_tol=0.0)
    finally:
        ka.set_fusion(-1, 4096, 12, 1.0e9)
        ka.set_fusion_autotune(True)
    assert base.restores == 0 and none.restores == 0
    assert forced.restores > 0
    assert abs(forced.energy - base.energy) <= 1e-14
    assert np.max(np.abs(forced.gradient - base.gradient)) <= 1e-12
    assert np.max(np.abs(none.gradient - base.gradient)) <= 1e-12


def test_adjoint_leaves_psi_theta_in_the_state():
    n = 6
    c, ids = mixed_circuit(n, 12, False)
    h = workloads.random_pauli_operator(n, 8, np.random.default_rng(13))
    st = ka.Statevector(n, ka.KQ_C128)
    ka.energy_and_gradient(c, h, ids, state=st)
    ops, _, _ = c.flatten(ids)
    ref, _ = orc.sv_simulate(orc.zero_state(n), n, ops)
    assert np.max(np.abs(st.download() - ref)) <= 1e-12


def test_adjoint_rejects_non_unitary_ops():
    c = ka.QuantumCircuit(2)
    c.add_rx_gate(0, 0.3, True)
    c.add_m_gate(0)
    h = ka.PauliOperator(n_qubits=2)
    h.add(1.0, ka.SparsePauliString([(0, 3)], 2))
    with pytest.raises(RuntimeError):
        ka.energy_and_gradient(c, h)


def test_vqe_ansatz_10q_against_parameter_shift():
    w = workloads.vqe(n=10, layers=3, n_random=20, seed=1)
    ops, _, _ = w.circuit.flatten(w.params)
    packed = _packed(w.hamiltonian)
    shift = oracle_shift(ops, 10, packed, len(w.params))
    e_ref = oracle_energy(ops, 10, packed)
    r = ka.energy_and_gradient(w.circuit, w.hamiltonian, w.params, precision=ka.KQ_C64)
    assert abs(r.energy - e_ref) <= 1e-6
    assert np.max(np.abs(r.gradient - shift)) <= 1e-4
    r128 = ka.energy_and_gradient(w.circuit, w.hamiltonian, w.params, precision=ka.KQ_C128)
    assert abs(r128.energy - e_ref) <= 1e-12
    assert np.max(np.abs(r128.gradient - shift)) <= 1e-10


# the known answer of the reference's read_pauli_operator_test.cpp (6 qubits)
READ_KAT = (" 1.100000000000e+01    0.000000000000e+00   :                                                         \n"
            "-1.875000000000e-01    0.000000000000e+00   :   (0, X)   (1, X)   (2, X)   (3, X)   (4, X)   (5, X)   \n"
            " 1.875000000000e-01    0.000000000000e+00   :   (0, X)   (1, X)   (2, X)   (3, Z)                     \n")


def _random_hamiltonian_text(n: int, n_strings: int, rng) -> str:
    """A Hamiltonian file in the reference's text format ("(re im) : (q, P) ..." per string)."""
    lines = []
    for _ in range(n_strings):
        k = int(rng.integers(1, n + 1))
        qs = rng.permutation(n)[:k]
        terms = "   ".join(f"({int(q)}, {'XYZ'[int(rng.integers(3))]})" for q in qs)
        lines.append(f"{rng.normal(): .12e}    {0.0: .12e}   :   {terms}")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("source", ["reference-kat", "random-20"])
def test_read_hamiltonian_drives_device_energy_and_gradient(source, tmp_path):
    """Hamiltonian text read with read_pauli_operator (read_pauli_operator.cpp:29-69, the caller pattern of
    vqe_dipolar_linear_rotors.cpp:156) feeds the device expectation value and the adjoint energy gradient:
    the energy against the oracle's expectation of the read strings, the gradient against the oracle's
    parameter shift."""
    import io as _io
    from kettle_amd.io import read_pauli_operator
    rng = np.random.default_rng(21)
    if source == "reference-kat":
        n, text = 6, READ_KAT
    else:
        n, text = 8, _random_hamiltonian_text(8, 20, rng)
    path = tmp_path / "hamiltonian.txt"
    path.write_text(text)
    h = read_pauli_operator(path, n)
    assert h.size() == len(text.strip().splitlines())
    assert read_pauli_operator(_io.StringIO(text), n).packed()[0].tolist() == h.packed()[0].tolist()
    c, ids = workloads.n_local(n, [ka.Gate.RY, ka.Gate.RZ], [ka.Gate.CX], workloads.NLocalEntanglement.LINEAR, 2, True)
    theta = rng.uniform(-math.pi, math.pi, size=len(ids))
    for pid, t in zip(ids, theta):
        c.set_parameter_value(pid, float(t))
    ops, _, _ = c.flatten(ids)
    packed = _packed(h)
    e_ref = oracle_energy(ops, n, packed)
    sv = ka.Statevector(n)
    ka.simulate(c, sv)
    assert abs(ka.expectation_value(h, sv).real - e_ref) <= 1e-12
    r = ka.energy_and_gradient(c, h, ids, precision=ka.KQ_C128)
    assert abs(r.energy - e_ref) <= 1e-12
    shift = oracle_shift(ops, n, packed, len(ids))
    assert np.max(np.abs(r.gradient - shift)) <= 1e-10


def test_vqe_24q_960_params_full_gradient():
    """The full 24-qubit, 960-parameter, 248-string VQE (the adjoint workload at its real size):
      - the complex128 device energy equals the oracle's FP64 energy of the same workload within 1e-11
        (tests/golden/vqe24_oracle.json, written by tests/golden/make_vqe24.py in the build container), and
        the device parameter shift at the fixture's parameters equals the oracle's within 1e-10;
      - every one of the 960 complex128 adjoint gradient entries equals the complex128 device parameter
        shift (1921 energies) within 1e-10;
      - the complex64 adjoint energy is within 1e-6 of the oracle's, and its 960 entries within 1e-4."""
    import json
    from pathlib import Path
    golden = json.loads((Path(__file__).resolve().parent / "golden" / "vqe24_oracle.json").read_text())
    w = workloads.vqe(n=24, layers=20, n_random=200)
    assert len(w.params) == 960 and w.hamiltonian.size() == 248 and golden["n_params"] == 960
    ops, _, _ = w.circuit.flatten(w.params)
    sv = ka.Statevector(24, ka.KQ_C128)

    def energy(o):
        sv.set_basis(0)
        ka.apply_gates(sv, o)
        return ka.expectation_value(w.hamiltonian, sv).real

    e0 = energy(ops)
    assert abs(e0 - golden["energy"]) <= 1e-11
    shift = np.zeros(960)
    for i in np.nonzero(ops["aux"] >= 0)[0]:
        plus, minus = ops.copy(), ops.copy()
        plus["angle"][i] += math.pi / 2
        minus["angle"][i] -= math.pi / 2
        shift[ops[i]["aux"]] += 0.5 * (energy(plus) - energy(minus))
    for p, g in golden["parameter_shift"].items():
        assert abs(shift[int(p)] - g) <= 1e-10, p
    r128 = ka.energy_and_gradient(w.circuit, w.hamiltonian, w.params, precision=ka.KQ_C128)
    assert abs(r128.energy - golden["energy"]) <= 1e-11
    assert np.max(np.abs(r128.gradient - shift)) <= 1e-10
    r64 = ka.energy_and_gradient(w.circuit, w.hamiltonian, w.params, precision=ka.KQ_C64)
    assert abs(r64.energy - golden["energy"]) <= 1e-6
    assert np.max(np.abs(r64.gradient - shift)) <= 1e-4


def test_torch_autograd_function():
    n = 6
    c, ids = mixed_circuit(n, 21, True)
    h = workloads.random_pauli_operator(n, 12, np.random.default_rng(22))
    theta0 = np.array([c.parameter_values()[pid] for pid in ids])
    ops0, _, _ = c.flatten(ids)
    fd = oracle_fd(ops0, n, _packed(h), len(ids))
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    e = kag.energy(theta, c, h, ids, precision=ka.KQ_C128)
    e.backward()
    ref = ka.energy_and_gradient(c, h, ids, precision=ka.KQ_C128)
    assert abs(e.item() - ref.energy) <= 1e-14
    assert np.max(np.abs(theta.grad.numpy() - ref.gradient)) <= 1e-14
    # the gradient drives the energy down: one small step along -grad lowers E
    with torch.no_grad():
        step = theta - 1e-3 * theta.grad
    e2 = kag.energy(step, c, h, ids, precision=ka.KQ_C128)
    assert e2.item() < e.item()
    assert np.max(np.abs(theta.grad.numpy() - fd)) <= 1e-8
