"""Device parity: libkq (through the C-ABI) against the CPU oracle on the same inputs.

Tolerances: complex128 amplitudes, probabilities and energies within 1e-12 max-abs of the oracle;
complex64 within 2e-5 (amplitudes) / 1e-6 (energies, FP64 accumulation); index maps and sampled
indices bit-exact."""
from __future__ import annotations

import math

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd import workloads
from kettle_amd.circuit import Gate
from kat_cases import KATS, basis, random_gate_circuit
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL128 = 1e-12
TOL64 = 2e-5


@pytest.fixture(autouse=True)
def _fusion_default():
    ka.set_fusion(-1, 4096, 12, 1.0e9)
    yield
    ka.set_fusion(-1, 4096, 12, 1.0e9)


def gpu_run(circuit, initial, precision=ka.KQ_C128, seed=None):
    sv = ka.Statevector(circuit.n_qubits(), precision)
    sv.upload(initial)
    res = ka.simulate(circuit, sv, seed)
    return sv.download(), res


def oracle_run(circuit, initial, seed=None):
    ops, preds, _ = circuit.flatten()
    return orc.sv_simulate(initial, circuit.n_qubits(), ops, preds, seed=seed, n_bits=circuit.n_bits())


# ---- statevector gates ---------------------------------------------------------------------------------

@pytest.mark.parametrize("fusion", [0, -1], ids=["sweep", "fused"])
def test_known_answers_on_device(fusion):
    ka.set_fusion(fusion)
    for kat in KATS:
        c = ka.QuantumCircuit(kat.n)
        kat.build(c)
        out, _ = gpu_run(c, basis(kat.initial))
        np.testing.assert_allclose(out, kat.expected, atol=1e-14, rtol=0, err_msg=kat.name)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 13, 15])
@pytest.mark.parametrize("fusion", [0, 4, 8, -1], ids=["sweep", "k4", "k8", "kauto"])
def test_random_all_gates_complex128(n, fusion):
    ka.set_fusion(fusion)
    rng = np.random.default_rng(1000 + n)
    c = random_gate_circuit(n, 200, rng)
    init = basis("0" * n)
    out, _ = gpu_run(c, init)
    ref, _ = oracle_run(c, init)
    assert np.max(np.abs(out - ref)) <= TOL128


@pytest.mark.parametrize("n", [3, 9, 14])
def test_random_all_gates_complex64(n):
    rng = np.random.default_rng(2000 + n)
    c = random_gate_circuit(n, 150, rng)
    init = basis("0" * n)
    out, _ = gpu_run(c, init, ka.KQ_C64)
    ref, _ = oracle_run(c, init)
    assert np.max(np.abs(out - ref)) <= TOL64


def test_plain_statevector_workload_20q():
    """20 qubits, depth 40, complex128: amplitudes, 4096-shot counts and <H> of 64 strings."""
    w = workloads.plain_statevector(n=20, depth=40)
    ops, _, _ = w.circuit.flatten()
    sv = ka.Statevector(20)
    ka.simulate(w.circuit, sv)
    out = sv.download()
    ref, _ = orc.sv_simulate(orc.zero_state(20), 20, ops)
    assert np.max(np.abs(out - ref)) <= TOL128
    e = ka.expectation_value(w.hamiltonian, sv)
    e_ref = orc.pauli_expectation(ref, 20, w.hamiltonian.packed())
    assert abs(e - e_ref) <= TOL128
    mem = ka.perform_measurements_as_memory(sv, w.shots, seed=0)
    mem_ref = orc.sample_probabilities(orc.probabilities(ref, 20), w.shots, 0)
    assert np.array_equal(mem, mem_ref)


def test_input_state_upload_and_qubit_mismatch():
    c = ka.QuantumCircuit(3)
    c.add_h_gate(0)
    sv = ka.Statevector(2)
    with pytest.raises(ka.KqError, match="different number of qubits"):
        ka.simulate(c, sv)
    rng = np.random.default_rng(3)
    a = rng.normal(size=8) + 1j * rng.normal(size=8)
    a /= np.linalg.norm(a)
    sv = ka.Statevector(a)
    np.testing.assert_array_equal(sv.download(), a)


def test_plan_fuses_gates():
    w = workloads.qft_plus_layers(n=18, layers=4)
    sv = ka.Statevector(18)
    ka.simulate(w, sv)
    n_gates, n_passes, n_tiles = ka.plan_stats()
    assert n_gates == w.n_circuit_elements()
    assert n_passes < n_gates / 4
    ops, _, _ = w.flatten()
    ref, _ = orc.sv_simulate(orc.zero_state(18), 18, ops)
    assert np.max(np.abs(sv.download() - ref)) <= TOL128


# ---- index maps --------------------------------------------------------------------------------------------

def test_pair_maps_bit_exact():
    lib = ka._lib.load()
    for n in range(1, 13):
        for t in range(n):
            np_ = 1 << (n - 1)
            s0 = np.empty(np_, np.uint64)
            s1 = np.empty(np_, np.uint64)
            ka._lib.check(lib.kq_debug_pair_map_1q(n, t, ka._lib.ptr(s0), ka._lib.ptr(s1)))
            r0, r1 = orc.pair_map_1q(n, t)
            assert np.array_equal(s0, r0) and np.array_equal(s1, r1), (n, t)
        for c in range(n):
            for t in range(n):
                if c == t:
                    continue
                for four in (0, 1):
                    out = np.empty((1 << (n - 2)) * (4 if four else 2), np.uint64)
                    ka._lib.check(lib.kq_debug_pair_map_2q(n, c, t, four, ka._lib.ptr(out)))
                    ref = orc.pair_map_2q(n, c, t, four_way=bool(four)).reshape(-1)
                    assert np.array_equal(out, ref), (n, c, t, four)


# ---- measurement, control flow, loggers ------------------------------------------------------------------

@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 5, 6, 7])
def test_mid_circuit_measurement_and_control_flow(seed):
    n = 6
    rng = np.random.default_rng(50 + seed)
    c = random_gate_circuit(n, 30, rng)
    c.add_m_gate(0, 0)
    c.add_m_gate(3, 1)
    a = ka.QuantumCircuit(n)
    a.add_x_gate(2)
    a.add_h_gate(5)
    b = ka.QuantumCircuit(n)
    b.add_rz_gate(2, math.pi / 4)
    c.add_if_else_statement(0, a, b)
    d = ka.QuantumCircuit(n)
    d.add_cx_gate(1, 4)
    c.add_if_statement(ka.ControlFlowPredicate([0, 1], [1, 0], ka.ControlFlowBooleanKind.IF_NOT), d)
    for q in range(n):
        c.add_ry_gate(q, 0.3 * (q + 1))
    c.add_m_gate(5, 2)
    init = basis("0" * n)
    out, res = gpu_run(c, init, seed=seed)
    ref, creg = oracle_run(c, init, seed=seed)
    assert res.classical_register[:3] == [int(v) for v in creg[:3]]
    assert np.max(np.abs(out - ref)) <= TOL128


def test_loggers():
    c = ka.QuantumCircuit(2)
    c.add_h_gate(0)
    c.add_statevector_circuit_logger()
    c.add_m_gate(0)
    c.add_classical_register_circuit_logger()
    sv = ka.Statevector(2)
    res = ka.simulate(c, sv, 4)
    assert [k for k, _ in res.loggers] == ["statevector", "classical_register"]
    np.testing.assert_allclose(res.loggers[0][1], [1 / math.sqrt(2), 1 / math.sqrt(2), 0, 0], atol=1e-15)
    assert res.loggers[1][1][0] == res.classical_register[0]


def test_qubit_weights_and_collapse_primitives():
    lib = ka._lib.load()
    a = np.zeros(4, complex)
    a[0] = a[1] = a[3] = 1 / math.sqrt(3)
    sv = ka.Statevector(a)
    w = np.zeros(2)
    ka._lib.check(lib.kq_qubit_weights(sv.handle, 0, ka._lib.ptr(w)))
    np.testing.assert_allclose(w, [1 / 3, 2 / 3], atol=1e-15)
    ka._lib.check(lib.kq_collapse(sv.handle, 0, 1, math.sqrt(1.5)))
    np.testing.assert_allclose(sv.download(), [0, 1 / math.sqrt(2), 0, 1 / math.sqrt(2)], atol=1e-15)


# ---- probabilities and sampling ------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 4, 10, 14])
def test_probabilities_noise_and_exact_samples(n):
    rng = np.random.default_rng(60 + n)
    c = random_gate_circuit(n, 50, rng)
    init = basis("0" * n)
    ref, _ = oracle_run(c, init)
    sv = ka.Statevector(c.n_qubits())
    ka.simulate(c, sv)
    p = ka.calculate_probabilities_raw(sv)
    np.testing.assert_allclose(p, orc.probabilities(ref, n), atol=TOL128, rtol=0)
    noise = ka.QuantumNoise(n)
    for q in range(n):
        noise.set(q, 0.05 * (q % 3))
    pn = ka.calculate_probabilities_raw(sv, noise)
    np.testing.assert_allclose(pn, orc.probabilities(ref, n, noise.array()), atol=TOL128, rtol=0)
    for seed in (0, 1, 123):
        mem = ka.perform_measurements_as_memory(sv, 3000, seed=seed)
        mem_ref = orc.sample_probabilities(orc.probabilities(ref, n), 3000, seed)
        assert np.array_equal(mem, mem_ref)
        mem = ka.perform_measurements_as_memory(sv, 1000, noise, seed=seed)
        mem_ref = orc.sample_probabilities(orc.probabilities(ref, n, noise.array()), 1000, seed)
        assert np.array_equal(mem, mem_ref)


def test_sampling_chi_square_and_bitstrings():
    from scipy.stats import chisquare

    n = 6
    rng = np.random.default_rng(9)
    c = random_gate_circuit(n, 40, rng)
    sv = ka.Statevector(n)
    ka.simulate(c, sv)
    p = ka.calculate_probabilities_raw(sv)
    raw = ka.perform_measurements_as_counts_raw(sv, 200000, seed=4)
    counts = np.zeros(1 << n)
    for k, v in raw.items():
        counts[k] = v
    mask = p > 1e-12
    assert chisquare(counts[mask], p[mask] / p[mask].sum() * counts.sum()).pvalue > 1e-4
    bits = ka.perform_measurements_as_counts(sv, 500, seed=4)
    raw = ka.perform_measurements_as_counts_raw(sv, 500, seed=4)
    assert sum(bits.values()) == 500
    for k, v in raw.items():
        assert bits[ka.state_index_to_bitstring(k, n)] == v
    marg = ka.perform_measurements_as_counts_marginal(sv, 500, [1, 4], seed=4)
    assert all(s[1] == "x" and s[4] == "x" for s in marg)


def test_sample_from_raw_probabilities_edge_cases():
    # point mass, trailing zeros, exact duplicates in the CDF
    for p in ([1.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.5, 0.5], [0.25, 0.0, 0.0, 0.75, 0.0, 0.0, 0.0, 0.0]):
        p = np.array(p)
        for seed in (0, 5):
            got = ka.perform_measurements_as_memory(p, 2000, seed=seed)
            assert np.array_equal(got, orc.sample_probabilities(p, 2000, seed))


# ---- Pauli strings --------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 4, 12])
@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_pauli_expectation(n, precision):
    rng = np.random.default_rng(70 + n)
    c = random_gate_circuit(n, 60, rng)
    ref, _ = oracle_run(c, basis("0" * n))
    op = ka.PauliOperator(n_qubits=n)
    for _ in range(40):
        nt = int(rng.integers(0, 2 * n + 1))
        terms = [(int(rng.integers(n)), int(rng.integers(0, 4))) for _ in range(nt)]
        op.add(complex(rng.normal(), rng.normal()), ka.SparsePauliString(terms, n, ka.PauliPhase(int(rng.integers(4)))))
    sv = ka.Statevector(n, precision)
    sv.upload(ref)
    e = ka.expectation_value(op, sv)
    e_ref = orc.pauli_expectation(ref, n, op.packed())
    tol = TOL128 if precision == ka.KQ_C128 else 1e-5
    assert abs(e - e_ref) <= tol * max(1.0, abs(e_ref))
    terms = [(q, int(rng.integers(1, 4))) for q in range(n)]
    s = ka.SparsePauliString(terms, n)
    sv.upload(ref)
    ka.simulate_pauli_string(s, sv)
    want = orc.apply_pauli_string(ref, n, [q for q, _ in terms], [k for _, k in terms])
    assert np.max(np.abs(sv.download() - want)) <= (TOL128 if precision == ka.KQ_C128 else TOL64)


# ---- density matrix ----------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 4, 6])
@pytest.mark.parametrize("fusion", [0, -1], ids=["sweep", "fused"])
def test_density_matrix_gates(n, fusion):
    ka.set_fusion(fusion)
    rng = np.random.default_rng(80 + n)
    c = random_gate_circuit(n, 60, rng)
    ops, _, _ = c.flatten()
    rho0 = orc.sv_to_dm(basis("0" * n), n)
    ref, _ = orc.dm_simulate(rho0, n, ops)
    dm = ka.DensityMatrix(n)
    ka.simulate(c, dm)
    assert np.max(np.abs(dm.download() - ref)) <= TOL128
    assert abs(dm.trace() - orc.dm_trace(ref, n)) <= TOL128


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_density_matrix_measurement_and_if_else(seed):
    c = workloads.density_matrix_circuit(n=5, depth=12, seed=seed, measure_every=4)
    ops, preds, _ = c.flatten()
    ref, creg = orc.dm_simulate(orc.sv_to_dm(basis("00000"), 5), 5, ops, preds, seed=seed, n_bits=5)
    dm = ka.DensityMatrix(5)
    res = ka.simulate(c, dm, seed)
    assert res.classical_register[0] == int(creg[0])
    assert np.max(np.abs(dm.download() - ref)) <= TOL128
    for keep in ([0, 1], [2], [0, 3, 4]):
        traced = [q for q in range(5) if q not in keep]
        pt = ka.partial_trace(dm, traced)
        side = 1 << len(keep)
        want = orc.dm_partial_trace(ref, 5, traced).reshape(side, side).T
        assert np.max(np.abs(pt - want)) <= TOL128
    p = ka.calculate_probabilities_raw(dm)
    np.testing.assert_allclose(p, np.real(ref.reshape(32, 32).T.diagonal()), atol=TOL128)


@pytest.mark.parametrize("n", [3, 5, 8])
@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_dm_trace_of_square_and_is_pure(n, precision):
    """trace_of_square on the device (FP64 accumulation) against the oracle: a simulated pure state, an
    uploaded random mixed state, the maximally mixed state (2^-n) and a non-Hermitian matrix; is_pure
    follows kettle's std::norm(tr - 1) <= tol."""
    tol = TOL128 if precision == ka.KQ_C128 else 1e-6
    rng = np.random.default_rng(300 + n)
    c = random_gate_circuit(n, 40, rng)
    dm = ka.DensityMatrix(n, precision)
    ka.simulate(c, dm)
    ref, _ = orc.dm_simulate(orc.sv_to_dm(basis("0" * n), n), n, c.flatten()[0])
    assert abs(dm.trace_of_square() - orc.dm_trace_of_square(ref, n)) <= tol
    assert abs(dm.trace_of_square() - 1.0) <= tol and dm.is_pure()
    N = 1 << n
    a = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    mixed = a @ a.conj().T
    mixed /= np.trace(mixed)
    for m in (mixed, np.eye(N) / N, a / N):
        dm.upload(m.T.reshape(-1))
        want = orc.dm_trace_of_square(m.T.reshape(-1), n)
        assert abs(dm.trace_of_square() - want) <= tol * max(1.0, abs(want))
    dm.upload((np.eye(N) / N).T.reshape(-1))
    assert abs(dm.trace_of_square() - 2.0 ** -n) <= tol
    assert not dm.is_pure() and dm.is_pure(tolerance=1.0)


def test_dm_from_statevector():
    rng = np.random.default_rng(11)
    a = rng.normal(size=16) + 1j * rng.normal(size=16)
    a /= np.linalg.norm(a)
    sv = ka.Statevector(a)
    dm = ka.statevector_to_density_matrix(sv)
    assert np.max(np.abs(dm.download() - orc.sv_to_dm(a, 4))) <= 1e-15
    dm2 = ka.DensityMatrix(4)
    dm2.set_basis(5)
    m = dm2.matrix()
    assert m[5, 5] == 1.0 and np.count_nonzero(m) == 1


# ---- qubit placement: SWAP gates as relabels, then every read path -------------------------------------------

@pytest.mark.parametrize("n", [9, 14, 17])
@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_swaps_and_moves_then_reads(n, precision):
    """Random layers with SWAP gates (run as relabels where qubits may sit on any index bit), near
    misses that are not SWAPs, and a QFT (its swap network is SWAP gates); then the amplitudes, the
    device pointer's bytes, probabilities, a Pauli expectation and a second circuit on the same
    state (it starts from the placement the first one left), each against the oracle."""
    from kettle_amd.workloads import add_brick_ladder, add_random_1q_layer
    rng = np.random.default_rng(900 + n + precision)
    c = ka.QuantumCircuit(n)
    for layer in range(6):
        add_random_1q_layer(c, rng)
        add_brick_ladder(c, rng, layer % 2)
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        c.add_swap_gate(a, b)
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        c.add_cx_gate(a, b)
        c.add_cx_gate(b, a)
        c.add_cx_gate(b, a)
    c.add_qft_gate(list(range(n)))
    ref, _ = oracle_run(c, basis("0" * n))
    tol = TOL128 if precision == ka.KQ_C128 else TOL64
    # the device pointer's bytes (the state is put back in qubit order first), then a download
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    sv = ka.Statevector(n, precision)
    ka.simulate(c, sv)
    ptr, nbytes = sv.device_ptr()
    assert nbytes == (1 << n) * (16 if precision == ka.KQ_C128 else 8)
    host = np.empty(1 << n, np.complex128 if precision == ka.KQ_C128 else np.complex64)
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(ctypes.c_void_p(host.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes), 2) == 0
    assert np.max(np.abs(host.astype(np.complex128) - ref)) <= tol
    assert np.max(np.abs(sv.download() - ref)) <= tol
    # probabilities after a fresh run (the state left permuted again)
    sv2 = ka.Statevector(n, precision)
    ka.simulate(c, sv2)
    p = ka.calculate_probabilities_raw(sv2)
    assert np.max(np.abs(p - np.abs(ref) ** 2)) <= (1e-12 if precision == ka.KQ_C128 else 1e-6)
    # a Pauli expectation after a fresh run
    sv3 = ka.Statevector(n, precision)
    ka.simulate(c, sv3)
    op = ka.PauliOperator(n_qubits=n)
    for _ in range(12):
        terms = [(int(rng.integers(n)), int(rng.integers(0, 4))) for _ in range(int(rng.integers(1, n + 1)))]
        op.add(complex(rng.normal(), 0.0), ka.SparsePauliString(terms, n))
    e = ka.expectation_value(op, sv3)
    e_ref = orc.pauli_expectation(ref, n, op.packed())
    assert abs(e - e_ref) <= (TOL128 if precision == ka.KQ_C128 else 1e-5) * max(1.0, abs(e_ref))
    # a second circuit on a state still permuted by the first
    sv4 = ka.Statevector(n, precision)
    ka.simulate(c, sv4)
    ka.simulate(c, sv4)
    ref2, _ = oracle_run(c, ref)
    assert np.max(np.abs(sv4.download() - ref2)) <= 2 * tol
