"""CPU checks that pin the oracle: the reference's known answers, brute-force dense linear algebra,
closed forms and invariants.  A plausible slip in the oracle (dropped term, wrong sign, swapped
operand, wrong index order) fails at least one of these."""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd.circuit import Gate
from kat_cases import (ALL_GATES, KATS, R, basis, dense_unitary, gate_matrix, pauli_matrix, random_gate_circuit,
                       to_m2)
from oracle import oracle as orc

GOLDEN = Path(__file__).resolve().parent / "golden"


def run_oracle(circuit: ka.QuantumCircuit, initial: np.ndarray, seed=None):
    ops, preds, _ = circuit.flatten()
    return orc.sv_simulate(initial, circuit.n_qubits(), ops, preds, seed=seed, n_bits=circuit.n_bits())


# ---- simulate(): known answers ------------------------------------------------------------------

@pytest.mark.parametrize("kat", KATS, ids=[k.name for k in KATS])
def test_oracle_known_answers(kat):
    c = ka.QuantumCircuit(kat.n)
    kat.build(c)
    out, _ = run_oracle(c, basis(kat.initial))
    np.testing.assert_allclose(out, kat.expected, atol=1e-14, rtol=0)


@pytest.mark.parametrize("g", [g for g in ALL_GATES if g not in (Gate.U, Gate.CU)], ids=lambda g: g.name)
def test_oracle_u_gate_mimics_named_gate(g):
    """U / CU with a named gate's matrix reproduce that gate (simulate_test.cpp:890-1212)."""
    rng = np.random.default_rng(int(g))
    theta = float(rng.uniform(0, 2 * math.pi))
    n = 3
    init = rng.normal(size=8) + 1j * rng.normal(size=8)
    init /= np.linalg.norm(init)
    named = ka.QuantumCircuit(n)
    mimic = ka.QuantumCircuit(n)
    m = to_m2(gate_matrix(g, theta))
    if g <= Gate.SXDAG:
        getattr(named, f"add_{g.name.lower()}_gate")(1)
        mimic.add_u_gate(m, 1)
    elif g <= Gate.P:
        getattr(named, f"add_{g.name.lower()}_gate")(1, theta)
        mimic.add_u_gate(m, 1)
    elif g <= Gate.CSXDAG:
        getattr(named, f"add_{g.name.lower()}_gate")(2, 0)
        mimic.add_cu_gate(m, 2, 0)
    else:
        getattr(named, f"add_{g.name.lower()}_gate")(2, 0, theta)
        mimic.add_cu_gate(m, 2, 0)
    a, _ = run_oracle(named, init)
    b, _ = run_oracle(mimic, init)
    np.testing.assert_allclose(a, b, atol=1e-14, rtol=0)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_oracle_matches_dense_kronecker_simulation(n):
    """Every gate kind on random qubits vs the product of 2^n x 2^n Kronecker-embedded matrices."""
    rng = np.random.default_rng(100 + n)
    c = random_gate_circuit(n, 60, rng)
    ops, _, _ = c.flatten()
    init = basis("0" * n)
    out, _ = orc.sv_simulate(init, n, ops)
    ref = dense_unitary(ops, n) @ init
    np.testing.assert_allclose(out, ref, atol=1e-12, rtol=0)


def test_oracle_norm_and_inverse():
    """U then U^dagger returns the input (each gate's inverse as a U gate)."""
    n = 5
    rng = np.random.default_rng(7)
    c = random_gate_circuit(n, 80, rng)
    ops, _, _ = c.flatten()
    u = dense_unitary(ops, n)
    init = rng.normal(size=32) + 1j * rng.normal(size=32)
    init /= np.linalg.norm(init)
    out, _ = orc.sv_simulate(init, n, ops)
    assert abs(np.linalg.norm(out) - 1.0) < 1e-13
    np.testing.assert_allclose(u.conj().T @ out, init, atol=1e-12)


# ---- pair generators -----------------------------------------------------------------------------

def _bits_to_index(s: str) -> int:
    return sum(1 << i for i, ch in enumerate(s) if ch == "1")


def test_pair_generator_known_sets():
    """Expected pair sets of gate_pair_generator_test.cpp:67-398 (transcribed to golden JSON)."""
    cases = json.loads((GOLDEN / "pair_generator_kat.json").read_text())
    assert len(cases) >= 10
    for case in cases:
        expected = sorted((_bits_to_index(a), _bits_to_index(b)) for a, b in case["pairs"])
        if case["kind"] == "1q":
            s0, s1 = orc.pair_map_1q(case["n"], case["target"])
            got = sorted(zip(s0.tolist(), s1.tolist()))
        else:
            m = orc.pair_map_2q(case["n"], case["control"], case["target"])
            got = sorted(map(tuple, m.tolist()))
        assert got == expected, case


@pytest.mark.parametrize("n", range(1, 11))
def test_pair_generator_closed_form_and_set_state(n):
    """Order: i1 (high part) fastest, s0 = i0 + 2*i1*2^t; the set is every index with bit t = 0;
    set_state(k) + next() gives the k-th pair of the sequential walk."""
    rng = np.random.default_rng(n)
    for t in range(n):
        s0, s1 = orc.pair_map_1q(n, t)
        i1max = 1 << (n - t - 1)
        k = np.arange(1 << (n - 1), dtype=np.uint64)
        i0, i1 = k // np.uint64(i1max), k % np.uint64(i1max)
        np.testing.assert_array_equal(s0, i0 + np.uint64(2) * i1 * np.uint64(1 << t))
        np.testing.assert_array_equal(s1, s0 + np.uint64(1 << t))
        assert sorted(s0.tolist()) == [i for i in range(1 << n) if not (i >> t) & 1]
        for kk in rng.integers(0, 1 << (n - 1), size=4):
            assert orc.pair_at_1q(n, t, int(kk)) == (int(s0[kk]), int(s1[kk]))
    if n >= 2:
        for c in range(n):
            for t in range(n):
                if c == t:
                    continue
                m = orc.pair_map_2q(n, c, t)
                want = sorted(i for i in range(1 << n) if (i >> c) & 1 and not (i >> t) & 1)
                assert sorted(m[:, 0].tolist()) == want
                np.testing.assert_array_equal(m[:, 1], m[:, 0] + np.uint64(1 << t))
                four = orc.pair_map_2q(n, c, t, four_way=True)
                np.testing.assert_array_equal(four[:, 2], m[:, 0])
                np.testing.assert_array_equal(four[:, 3], m[:, 1])
                kk = int(rng.integers(0, 1 << (n - 2)))
                assert orc.pair_at_2q(n, c, t, kk) == tuple(int(v) for v in m[kk])


# ---- measurement and control flow ----------------------------------------------------------------

def test_measurement_of_basis_states_is_deterministic():
    for bits in ["00", "10", "01", "11"]:
        c = ka.QuantumCircuit(2)
        c.add_m_gate(0)
        c.add_m_gate(1)
        out, creg = run_oracle(c, basis(bits), seed=3)
        assert list(creg) == [int(bits[0]), int(bits[1])]
        np.testing.assert_allclose(out, basis(bits), atol=0)


def test_measurement_collapses_and_renormalises():
    # (|00> + |10> + |11>)/sqrt3 measured on qubit 0: p0 = 1/3, p1 = 2/3
    a = np.zeros(4, complex)
    a[0] = a[1] = a[3] = 1 / math.sqrt(3)
    outcomes = []
    for seed in range(200):
        c = ka.QuantumCircuit(2)
        c.add_m_gate(0)
        out, creg = run_oracle(c, a, seed=seed)
        outcomes.append(int(creg[0]))
        if creg[0] == 0:
            np.testing.assert_allclose(out, basis("00"), atol=1e-15)
        else:
            np.testing.assert_allclose(out, [0, R, 0, R], atol=1e-15)
    frac = np.mean(outcomes)
    assert 0.55 < frac < 0.78   # 2/3 within ~4 sigma for 200 draws


def test_control_flow_if_and_if_not():
    """control_flow_test.cpp:12-100: M(0) then X(1) only when bit 0 is 1 / 0."""
    for prep, expected in [(True, "11"), (False, "00")]:
        c = ka.QuantumCircuit(2)
        if prep:
            c.add_x_gate(0)
        c.add_m_gate(0)
        sub = ka.QuantumCircuit(2)
        sub.add_x_gate(1)
        c.add_if_statement(0, sub)
        out, _ = run_oracle(c, basis("00"))
        np.testing.assert_allclose(out, basis(expected), atol=0)
    c = ka.QuantumCircuit(2)
    c.add_m_gate(0)
    sub = ka.QuantumCircuit(2)
    sub.add_x_gate(0)
    c.add_if_not_statement(0, sub)
    out, _ = run_oracle(c, basis("00"))
    np.testing.assert_allclose(out, basis("10"), atol=0)


def test_control_flow_if_else_nested():
    c = ka.QuantumCircuit(3)
    c.add_x_gate(0)
    c.add_m_gate(0)
    c.add_m_gate(1)
    inner_if = ka.QuantumCircuit(3)
    inner_if.add_x_gate(2)
    inner_else = ka.QuantumCircuit(3)
    inner_else.add_h_gate(2)
    outer_if = ka.QuantumCircuit(3)
    outer_if.add_if_else_statement(ka.ControlFlowPredicate([1], [0], ka.ControlFlowBooleanKind.IF), inner_if, inner_else)
    outer_else = ka.QuantumCircuit(3)
    outer_else.add_x_gate(1)
    c.add_if_else_statement(0, outer_if, outer_else)
    out, _ = run_oracle(c, basis("000"))
    np.testing.assert_allclose(out, basis("101"), atol=0)


def test_predicate_on_unmeasured_bit_raises():
    c = ka.QuantumCircuit(2)
    sub = ka.QuantumCircuit(2)
    sub.add_x_gate(1)
    c.add_if_statement(1, sub)
    with pytest.raises(orc.OracleError, match="There is no measured bit at classical register 1"):
        run_oracle(c, basis("00"))


def test_qubit_count_mismatch_raises():
    c = ka.QuantumCircuit(3)
    c.add_h_gate(0)
    ops, preds, _ = c.flatten()
    with pytest.raises(orc.OracleError, match="different number of qubits"):
        orc.sv_simulate(basis("00"), 2, ops, preds, circuit_n=3)


# ---- probabilities, noise and sampling ----------------------------------------------------------------

def test_probabilities_and_noise_closed_form():
    rng = np.random.default_rng(0)
    a = rng.normal(size=8) + 1j * rng.normal(size=8)
    a /= np.linalg.norm(a)
    p = orc.probabilities(a, 3)
    np.testing.assert_allclose(p, np.abs(a) ** 2, atol=1e-16)
    noise = np.array([0.1, 0.0, 0.25])
    pn = orc.probabilities(a, 3, noise)
    ref = np.abs(a) ** 2
    for q, e in enumerate(noise):
        new = ref.copy()
        for i in range(8):
            new[i] = (1 - e) * ref[i] + e * ref[i ^ (1 << q)]
        ref = new
    np.testing.assert_allclose(pn, ref, atol=1e-15)
    assert abs(pn.sum() - 1.0) < 1e-14


def test_sampler_cdf_and_upper_bound():
    """CDF = running sum; draws in [0, last - 1e-4 * (last - last value below it))."""
    p = np.array([0.1, 0.0, 0.4, 0.5, 0.0, 0.0])
    cdf, upper = orc.cdf(p)
    np.testing.assert_allclose(cdf, np.cumsum(p), atol=1e-16)
    assert upper == pytest.approx(1.0 - 1e-4 * (1.0 - 0.5), abs=1e-15)
    # a point mass: everything equals last, second-last is 0.0
    cdf, upper = orc.cdf(np.array([1.0, 0.0, 0.0, 0.0]))
    assert upper == pytest.approx(1.0 - 1e-4, abs=1e-16)
    idx = orc.sample_probabilities(np.array([1.0, 0.0, 0.0, 0.0]), 100, 5)
    assert set(idx.tolist()) == {0}


def test_sampler_chi_square():
    from scipy.stats import chisquare

    rng = np.random.default_rng(1)
    p = rng.random(16)
    p /= p.sum()
    idx = orc.sample_probabilities(p, 200000, 11)
    counts = np.bincount(idx.astype(np.int64), minlength=16)
    assert chisquare(counts, p * counts.sum()).pvalue > 1e-4
    # zero-probability states are never drawn
    p[3] = 0.0
    p /= p.sum()
    idx = orc.sample_probabilities(p, 20000, 12)
    assert 3 not in set(idx.tolist())


# ---- Pauli strings ---------------------------------------------------------------------------------------

def _expect(op: ka.PauliOperator, amps, n):
    return orc.pauli_expectation(amps, n, op.packed())


def test_pauli_one_qubit_eigenstates():
    """pauli_operator_test.cpp:18-62: <P> = +-1 on P's eigenstates, 0 on the others."""
    states = {
        "x+": np.array([R, R], complex), "x-": np.array([R, -R], complex),
        "y+": np.array([R, 1j * R]), "y-": np.array([R, -1j * R]),
        "z+": basis("0"), "z-": basis("1"),
    }
    for term, name in [(1, "x"), (2, "y"), (3, "z")]:
        op = ka.PauliOperator([(1.0, ka.SparsePauliString([(0, term)], 1))])
        for sname, amps in states.items():
            want = 0.0
            if sname[0] == name:
                want = 1.0 if sname[1] == "+" else -1.0
            assert abs(_expect(op, amps, 1) - want) < 1e-15, (term, sname)


def test_pauli_operator_known_answers():
    """pauli_operator_test.cpp:103-201."""
    T = ka.PauliTerm
    z0 = basis("0")
    op = ka.PauliOperator([(1.0, ka.SparsePauliString([T.Z])), (1.0, ka.SparsePauliString([T.X]))])
    assert abs(_expect(op, z0, 1) - 1.0) < 1e-15
    plus = np.array([R, R], complex)
    op = ka.PauliOperator([(1.0, ka.SparsePauliString([T.Z])), (-2.0, ka.SparsePauliString([T.X]))])
    assert abs(_expect(op, plus, 1) - (-2.0)) < 1e-15
    bell = np.array([R, 0, 0, R], complex)
    op = ka.PauliOperator([(1.0, ka.SparsePauliString([T.Z, T.Z])), (1.0, ka.SparsePauliString([T.X, T.X]))])
    assert abs(_expect(op, bell, 2) - 2.0) < 1e-15
    ghz = np.zeros(8, complex)
    ghz[0] = ghz[7] = R
    op = ka.PauliOperator([(1.0, ka.SparsePauliString([T.Z, T.Z, T.I])), (1.0, ka.SparsePauliString([T.X, T.X, T.X]))])
    assert abs(_expect(op, ghz, 3) - 2.0) < 1e-15
    pp = np.full(4, 0.5, complex)
    op = ka.PauliOperator([(1.0, ka.SparsePauliString([T.X, T.I])), (-1j, ka.SparsePauliString([T.I, T.X]))])
    assert abs(_expect(op, pp, 2) - (1.0 - 1.0j)) < 1e-15


@pytest.mark.parametrize("n", [1, 3, 6])
def test_pauli_expectation_vs_dense(n):
    """Random strings (including repeated qubits and phases) vs dense operator products."""
    rng = np.random.default_rng(20 + n)
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a /= np.linalg.norm(a)
    op = ka.PauliOperator(n_qubits=n)
    dense = np.zeros((1 << n, 1 << n), complex)
    phases = [1, 1j, -1, -1j]
    for s in range(12):
        nt = int(rng.integers(0, 2 * n + 1))
        terms = [(int(rng.integers(n)), int(rng.integers(0, 4))) for _ in range(nt)]
        ph = int(rng.integers(4))
        coeff = complex(rng.normal(), rng.normal())
        op.add(coeff, ka.SparsePauliString(terms, n, ka.PauliPhase(ph)))
        dense += coeff * phases[ph] * pauli_matrix(terms, n)
        single = orc.apply_pauli_string(a, n, [q for q, _ in terms], [k for _, k in terms])
        np.testing.assert_allclose(single, pauli_matrix(terms, n) @ a, atol=1e-14)
    want = np.vdot(a, dense @ a)
    assert abs(_expect(op, a, n) - want) < 1e-12


# ---- density matrix --------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_dm_matches_pure_state_projector(n):
    """simulate_density_matrix_test.cpp:24-302: DM evolution of |0><0| equals |psi><psi|."""
    rng = np.random.default_rng(40 + n)
    c = random_gate_circuit(n, 40, rng)
    ops, _, _ = c.flatten()
    psi, _ = orc.sv_simulate(basis("0" * n), n, ops)
    rho0 = orc.sv_to_dm(basis("0" * n), n)
    rho, _ = orc.dm_simulate(rho0, n, ops)
    np.testing.assert_allclose(rho, orc.sv_to_dm(psi, n), atol=1e-12)
    N = 1 << n
    m = rho.reshape(N, N).T
    np.testing.assert_allclose(m, np.outer(psi, psi.conj()), atol=1e-12)
    assert abs(orc.dm_trace(rho, n) - 1.0) < 1e-12


def test_dm_measurement_weights_are_block_frobenius_norms():
    """measure_density_matrix.cpp:21-51: w_b = sum |rho|^2 over the (b, b) block = p_b^2 for a
    pure state, so the outcome draws with p0^2 / (p0^2 + p1^2)."""
    n = 1
    a = np.array([math.sqrt(0.8), math.sqrt(0.2)], complex)
    rho = orc.sv_to_dm(a, n)
    ones = 0
    trials = 400
    c = ka.QuantumCircuit(1)
    c.add_m_gate(0)
    ops, preds, _ = c.flatten()
    for seed in range(trials):
        out, creg = orc.dm_simulate(rho, n, ops, preds, seed=seed, n_bits=1)
        ones += int(creg[0])
        if creg[0] == 1:
            # kept block scaled by sqrt(1 / w1) = 1 / 0.2
            np.testing.assert_allclose(out.reshape(2, 2).T, [[0, 0], [0, 1.0]], atol=1e-12)
    expected = 0.04 / (0.64 + 0.04)
    assert abs(ones / trials - expected) < 4 * math.sqrt(expected * (1 - expected) / trials)


def test_dm_trace_of_square_closed_forms():
    """trace_of_square (density_matrix.cpp:164-168): 1 for a pure state, 2^-n for the maximally mixed
    state, the product of the factors' values for rho_A (x) rho_B, and numpy's trace(M @ M) for a random
    matrix that is neither Hermitian nor of trace 1 (so a dropped conjugate or a transposed operand shows)."""
    rng = np.random.default_rng(17)

    def rand_rho(k):
        a = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        m = a @ a.conj().T
        return m / np.trace(m)

    for n in (1, 3, 5):
        psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        psi /= np.linalg.norm(psi)
        assert abs(orc.dm_trace_of_square(orc.sv_to_dm(psi, n), n) - 1.0) < 1e-13
        mixed = np.eye(1 << n) / (1 << n)
        assert abs(orc.dm_trace_of_square(mixed.T.reshape(-1), n) - 2.0 ** -n) < 1e-15
    ra, rb = rand_rho(2), rand_rho(2)
    full = np.kron(rb, ra)
    want = orc.dm_trace_of_square(ra.T.reshape(-1), 2) * orc.dm_trace_of_square(rb.T.reshape(-1), 2)
    assert abs(orc.dm_trace_of_square(full.T.reshape(-1), 4) - want) < 1e-14
    assert abs(want - np.trace(ra @ ra) * np.trace(rb @ rb)) < 1e-14
    m = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
    assert abs(orc.dm_trace_of_square(m.T.reshape(-1), 4) - np.trace(m @ m)) < 1e-11
    assert abs(np.trace(m @ m) - np.sum(np.abs(m) ** 2)) > 1.0   # the non-Hermitian case is not sum |m|^2


def test_dm_partial_trace_of_tensor_product():
    """density_matrix_test.cpp:253-473: tracing out one factor of rho_A (x) rho_B leaves the other."""
    rng = np.random.default_rng(5)

    def rand_rho(k):
        a = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        m = a @ a.conj().T
        return m / np.trace(m)

    ra, rb = rand_rho(2), rand_rho(1)   # qubits 0,1 in A (low bits), qubit 2 in B
    full = np.kron(rb, ra)
    vec = full.T.reshape(-1)
    red = orc.dm_partial_trace(vec, 3, [2]).reshape(4, 4).T
    np.testing.assert_allclose(red, ra, atol=1e-14)
    red = orc.dm_partial_trace(vec, 3, [0, 1]).reshape(2, 2).T
    np.testing.assert_allclose(red, rb, atol=1e-14)
    assert abs(np.trace(orc.dm_partial_trace(vec, 3, [0, 2]).reshape(2, 2)) - 1.0) < 1e-14


def test_dm_control_flow_matches_statevector_branch():
    """With M and if-else the DM follows the same branch as the SV run for the same seed when the
    outcome is deterministic."""
    n = 3
    c = ka.QuantumCircuit(n)
    c.add_x_gate(0)
    c.add_h_gate(1)
    c.add_m_gate(0, 0)
    a_sub = ka.QuantumCircuit(n)
    a_sub.add_x_gate(2)
    b_sub = ka.QuantumCircuit(n)
    b_sub.add_rz_gate(2, math.pi / 4)
    c.add_if_else_statement(0, a_sub, b_sub)
    ops, preds, _ = c.flatten()
    psi, creg = orc.sv_simulate(basis("000"), n, ops, preds, seed=1, n_bits=3)
    rho, creg2 = orc.dm_simulate(orc.sv_to_dm(basis("000"), n), n, ops, preds, seed=1, n_bits=3)
    assert list(creg) == list(creg2) == [1, -1, -1]
    np.testing.assert_allclose(rho, orc.sv_to_dm(psi, n), atol=1e-13)
