"""Per-shot re-simulation (perform_measurements_as_counts_marginal(circuit, state, ...),
measurements.cpp:178-208) through kq_sample_circuit_shots.

Every shot runs the circuit afresh (independent mid-circuit outcomes) and draws one index from its
final state.  The expected distribution is computed independently here: the oracle runs each
unitary segment, numpy projects and renormalises for each outcome history, and the final
distribution is the history-weighted mixture of the oracle's probabilities.  Counts are compared
with a chi-square test; seeded samplers are compared index-for-index with the seeded sampler over
each history's exact probabilities.  libkq shares one simulated branch per outcome history, so the
number of branches is bounded by the number of histories whatever the shot count."""
from __future__ import annotations

import math

import numpy as np
import pytest

import kettle_amd as ka
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

N = 4


def seg(build) -> np.ndarray:
    c = ka.QuantumCircuit(N)
    build(c)
    ops, _, _ = c.flatten()
    return ops


def run(psi, ops):
    out, _ = orc.sv_simulate(psi, N, ops)
    return out


def project(psi, q, b):
    idx = np.arange(psi.size)
    keep = ((idx >> q) & 1) == b
    p = float(np.sum(np.abs(psi[keep]) ** 2))
    return p, np.where(keep, psi, 0) / math.sqrt(p) if p > 0 else psi * 0


SEG_A = lambda c: (c.add_h_gate(0), c.add_ry_gate(1, 1.1), c.add_cx_gate(0, 2), c.add_rx_gate(3, 0.7))
IF_0 = lambda c: (c.add_x_gate(1), c.add_rz_gate(2, 0.4))
ELSE_0 = lambda c: c.add_h_gate(3)
SEG_B = lambda c: (c.add_cry_gate(2, 3, 0.9), c.add_h_gate(1))
IF_1 = lambda c: c.add_cx_gate(3, 0)
SEG_C = lambda c: c.add_ry_gate(0, 0.5)


def branching_circuit() -> ka.QuantumCircuit:
    def sub(build):
        s = ka.QuantumCircuit(N)
        build(s)
        return s

    c = ka.QuantumCircuit(N)
    SEG_A(c)
    c.add_m_gate(0, 0)
    c.add_if_else_statement(0, sub(IF_0), sub(ELSE_0))
    SEG_B(c)
    c.add_m_gate(1, 1)
    c.add_if_statement(1, sub(IF_1))
    SEG_C(c)
    return c


def histories(noise=None):
    """(probability, final probabilities) for every outcome history (b0, b1)."""
    psi = run(orc.zero_state(N), seg(SEG_A))
    out = []
    for b0 in (0, 1):
        p0, s0 = project(psi, 0, b0)
        if p0 == 0:
            continue
        s0 = run(s0, seg(IF_0 if b0 else ELSE_0))
        s0 = run(s0, seg(SEG_B))
        for b1 in (0, 1):
            p1, s1 = project(s0, 1, b1)
            if p1 == 0:
                continue
            if b1:
                s1 = run(s1, seg(IF_1))
            s1 = run(s1, seg(SEG_C))
            out.append((p0 * p1, orc.probabilities(s1, N, noise)))
    return out


def chi2_ok(idx, expected_probs, n_states):
    from scipy.stats import chisquare

    counts = np.bincount(np.asarray(idx, dtype=np.int64), minlength=n_states).astype(np.float64)
    exp = expected_probs * len(idx)
    keep = exp > 0
    assert np.all(counts[~keep] == 0), "an index of probability zero was drawn"
    return chisquare(counts[keep], exp[keep]).pvalue


@pytest.mark.parametrize("cache", [8 << 30, 0], ids=["cached", "rebuilt"])
def test_branching_circuit_distribution(cache):
    c = branching_circuit()
    hist = histories()
    expected = sum(w * p for w, p in hist)
    idx, n_branches = ka.perform_measurements_as_memory_circuit(c, ka.Statevector(N), 20000, max_cached_bytes=cache)
    assert n_branches <= 1 + 2 + 4
    assert chi2_ok(idx, expected, 1 << N) > 1e-4


def test_seeded_sampler_per_history():
    """A seeded sampler gives every shot of one history the same index: the one kettle's
    ProbabilitySampler_(probs, seed) draws first for that history's probabilities."""
    c = branching_circuit()
    hist = histories()
    leaf_index = [int(ka.perform_measurements_as_memory(p, 1, seed=11)[0]) for _, p in hist]
    idx, _ = ka.perform_measurements_as_memory_circuit(c, ka.Statevector(N), 8000, seed=11)
    assert set(idx.tolist()) <= set(leaf_index)
    # the histories' weights, summed over histories that share an index
    expected = np.zeros(1 << N)
    for (w, _), i in zip(hist, leaf_index):
        expected[i] += w
    assert chi2_ok(idx, expected, 1 << N) > 1e-4


def test_no_measurement_is_one_branch_with_noise():
    c = ka.QuantumCircuit(N)
    SEG_A(c)
    SEG_B(c)
    noise = ka.QuantumNoise(N)
    noise.set(2, 0.1)
    noise.set(0, 0.25)
    sv = ka.Statevector(N)
    idx, n_branches = ka.perform_measurements_as_memory_circuit(c, sv, 20000, noise=noise)
    assert n_branches == 1
    psi = run(run(orc.zero_state(N), seg(SEG_A)), seg(SEG_B))
    assert chi2_ok(idx, orc.probabilities(psi, N, noise.array()), 1 << N) > 1e-4
    # seeded: every shot draws the same index, the seeded sampler's first draw
    seeded, _ = ka.perform_measurements_as_memory_circuit(c, sv, 50, seed=5)
    assert len(set(seeded.tolist())) == 1
    assert int(seeded[0]) == int(ka.perform_measurements_as_memory(orc.probabilities(psi, N), 1, seed=5)[0])
    # the initial state is left as it was
    np.testing.assert_array_equal(sv.download(), orc.zero_state(N))


def test_marginal_counts_overload_and_errors():
    c = branching_circuit()
    counts = ka.perform_measurements_as_counts_marginal(c, ka.Statevector(N), 4000, [1, 3])
    assert sum(counts.values()) == 4000
    assert all(len(k) == N and k[1] == "x" and k[3] == "x" for k in counts)
    with pytest.raises(RuntimeError, match="different number of qubits"):
        ka.perform_measurements_as_memory_circuit(c, ka.Statevector(N + 1), 10)
    with pytest.raises(RuntimeError, match="marginal qubit index out of range"):
        ka.perform_measurements_as_counts_marginal(c, ka.Statevector(N), 10, [N])


def test_many_histories_share_branches_at_16_qubits():
    """Eight mid-circuit measurements on a 16-qubit state: at most 2^8 leaf histories however many
    shots, and the measured bits follow the exact product distribution (each measured qubit is in
    |+> when measured)."""
    n = 16
    c = ka.QuantumCircuit(n)
    for q in range(n):
        c.add_h_gate(q)
    for q in range(8):
        c.add_m_gate(q, q)
    idx, n_branches = ka.perform_measurements_as_memory_circuit(c, ka.Statevector(n), 30000)
    assert n_branches <= sum(2 ** k for k in range(9))
    low = np.asarray(idx, dtype=np.int64) & 0xFF
    assert chi2_ok(low, np.full(256, 1 / 256), 256) > 1e-4
