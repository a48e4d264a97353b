"""Batched energy + adjoint gradient (kq_batch_*) against the single-instance adjoint path and the
oracle.  Tolerances: complex128 energies and gradients within 1e-10 of the single-instance path
and energies within 1e-12 of the oracle; complex64 energies within 1e-6 of the FP64 oracle; graph
replays bit-identical to the eager run."""
from __future__ import annotations

import math

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd import workloads
from kettle_amd.batch import EnergyGradientBatch
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def oracle_energy(inst) -> float:
    amps, _ = orc.sv_simulate(orc.zero_state(inst.n), inst.n, inst.ops())
    return orc.pauli_expectation(amps, inst.n, inst.hamiltonian.packed()).real


def single(inst, precision):
    c, ids = inst.circuit()
    return ka.energy_and_gradient(c, inst.hamiltonian, ids, precision=precision)


def make_batch(insts, precision, forward_precision=None):
    b = EnergyGradientBatch(precision, forward_precision)
    for inst in insts:
        b.add_ops(inst.n, inst.ops(), len(inst.theta), inst.hamiltonian)
    return b


def small_instances(count: int, seed: int):
    return workloads.batched_vqe(count, n_range=(2, 13), depth_range=(1, 5), strings_range=(4, 24), seed=seed)


@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_batch_matches_single_instance_and_oracle(precision):
    insts = small_instances(24, 1)
    b = make_batch(insts, precision)
    r = b.run(use_graph=False)
    for i, inst in enumerate(insts):
        one = single(inst, precision)
        e_ref = oracle_energy(inst)
        if precision == ka.KQ_C128:
            assert abs(r.energies[i] - e_ref) <= 1e-12, i
            assert abs(r.energies[i] - one.energy) <= 1e-12, i
            assert np.max(np.abs(r.gradients[i] - one.gradient), initial=0.0) <= 1e-10, i
            assert r.drift[i] <= 1e-12, i
        else:
            assert abs(r.energies[i] - e_ref) <= 1e-6, i
            assert np.max(np.abs(r.gradients[i] - one.gradient), initial=0.0) <= 1e-4, i
            assert r.drift[i] <= 1e-5, i
    b.close()


def test_mixed_precision_batch_energies_are_fp64():
    """complex128 forward sweep, complex64 reverse sweep: energies FP64-exact, gradients at complex64
    accuracy."""
    insts = small_instances(12, 11)
    b = make_batch(insts, ka.KQ_C64, ka.KQ_C128)
    r = b.run(use_graph=False)
    for i, inst in enumerate(insts):
        one = single(inst, ka.KQ_C128)
        assert abs(r.energies[i] - oracle_energy(inst)) <= 1e-12, i
        assert np.max(np.abs(r.gradients[i] - one.gradient), initial=0.0) <= 1e-4, i
        assert r.drift[i] <= 1e-5, i
    rep = b.run(use_graph=True)
    assert np.array_equal(rep.energies, r.energies)
    b.close()
    with pytest.raises(ka.KqError):
        EnergyGradientBatch(ka.KQ_C128, ka.KQ_C64)


def test_batch_gradient_against_parameter_shift():
    insts = small_instances(6, 7)
    b = make_batch(insts, ka.KQ_C128)
    r = b.run(use_graph=False)
    for i, inst in enumerate(insts):
        ops = inst.ops()
        n = inst.n
        packed = inst.hamiltonian.packed()
        g = np.zeros(len(inst.theta))
        for j in np.nonzero(ops["aux"] >= 0)[0]:
            plus, minus = ops.copy(), ops.copy()
            plus["angle"][j] += math.pi / 2
            minus["angle"][j] -= math.pi / 2
            ep = orc.pauli_expectation(orc.sv_simulate(orc.zero_state(n), n, plus)[0], n, packed).real
            em = orc.pauli_expectation(orc.sv_simulate(orc.zero_state(n), n, minus)[0], n, packed).real
            g[ops[j]["aux"]] += 0.5 * (ep - em)
        assert np.max(np.abs(r.gradients[i] - g), initial=0.0) <= 1e-10, i


def test_graph_replay_and_parameter_updates():
    insts = small_instances(16, 3)
    b = make_batch(insts, ka.KQ_C128)
    eager = b.run(use_graph=False)
    for _ in range(3):
        rep = b.run(use_graph=True)
        assert np.array_equal(rep.energies, eager.energies)
        for a, c in zip(rep.gradients, eager.gradients):
            assert np.array_equal(a, c)
    # new angles for two instances: the replayed graph sees the re-lowered programs
    rng = np.random.default_rng(5)
    for i in (0, 9):
        values = rng.uniform(-math.pi, math.pi, size=len(insts[i].theta))
        b.set_parameters(i, values)
        insts[i].theta = values
    rep = b.run(use_graph=True)
    for i in (0, 9, 4):
        one = single(insts[i], ka.KQ_C128)
        assert abs(rep.energies[i] - one.energy) <= 1e-12
        assert np.max(np.abs(rep.gradients[i] - one.gradient)) <= 1e-10
    # every instance at once (an optimiser step): only the changed programs are written again
    values = [rng.uniform(-math.pi, math.pi, size=len(inst.theta)) for inst in insts]
    b.set_all_parameters(values)
    for inst, v in zip(insts, values):
        inst.theta = v
    rep = b.run(use_graph=True)
    for i in (1, 9, 15):
        one = single(insts[i], ka.KQ_C128)
        assert abs(rep.energies[i] - one.energy) <= 1e-12
        assert np.max(np.abs(rep.gradients[i] - one.gradient)) <= 1e-10
    with pytest.raises(ValueError):
        b.set_all_parameters(values[:-1])
    assert b.info()["launches"] > 0
    b.close()


def test_drift_triggered_recompute():
    """drift_tol = 0 forces the complex128 single-state recompute of every complex64 instance that
    drifted at all: its energy and gradient then match the complex128 path to 1e-12 / 1e-10; a loose
    tolerance recomputes nothing and leaves the batched results as they were."""
    insts = small_instances(16, 21)
    b = make_batch(insts, ka.KQ_C64)
    plain = b.run(use_graph=False)
    assert b.recomputed() == (0, 0)
    b.set_drift_recompute(0.0)
    for use_graph in (False, True):
        r = b.run(use_graph=use_graph)
        drifted = np.nonzero(r.drift > 0.0)[0]
        assert len(drifted) >= len(insts) // 2
        assert b.recomputed()[0] == len(drifted)
        for i in drifted:
            one = single(insts[i], ka.KQ_C128)
            assert abs(r.energies[i] - one.energy) <= 1e-12, i
            assert np.max(np.abs(r.gradients[i] - one.gradient), initial=0.0) <= 1e-10, i
        for i in np.nonzero(r.drift == 0.0)[0]:
            assert r.energies[i] == plain.energies[i], i
    assert b.recomputed()[1] == 2 * len(drifted)
    b.set_drift_recompute(1.0)
    r = b.run()
    assert b.recomputed()[0] == 0
    assert np.array_equal(r.energies, plain.energies)
    b.set_drift_recompute(None)
    b.run()
    assert b.recomputed()[0] == 0
    with pytest.raises(ka.KqError):
        b.set_drift_recompute(float("nan"))
    b.close()


def test_batch_errors():
    b = EnergyGradientBatch(ka.KQ_C64)
    with pytest.raises(ka.KqError, match="no instances"):
        b.run()
    c = ka.QuantumCircuit(3)
    c.add_rx_gate(0, 0.1, True)
    h = ka.PauliOperator(n_qubits=3)
    h.add(1.0, ka.SparsePauliString([(0, 3)], 3))
    b.add(c, h)
    with pytest.raises(ValueError):
        b.set_parameters(0, [0.1, 0.2])
    bad = ka.QuantumCircuit(2)
    bad.add_m_gate(0)
    with pytest.raises(RuntimeError):
        b.add(bad, h)
    r = b.run()
    assert abs(r.energies[0] - math.cos(0.1)) <= 1e-6
    assert abs(r.gradients[0][0] + math.sin(0.1)) <= 1e-5
    b.close()


def test_batched_vqe_4096_instances_sample_check():
    """The full 4096-instance uneven batch (8..22 qubits, complex128 forward, complex64 reverse
    sweep): every energy finite, the drift small, and a sample of instances against the complex128
    single-instance path."""
    insts = workloads.batched_vqe(4096)
    b = make_batch(insts, ka.KQ_C64, ka.KQ_C128)
    r = b.run(use_graph=False)
    assert np.all(np.isfinite(r.energies))
    assert np.max(r.drift) <= 1e-3
    for i in (0, 1, 2047, 4095):
        one = single(insts[i], ka.KQ_C128)
        assert abs(r.energies[i] - one.energy) <= 1e-10, i
        assert np.max(np.abs(r.gradients[i] - one.gradient)) <= 1e-4 * max(1.0, np.max(np.abs(one.gradient))), i
    r2 = b.run(use_graph=True)
    assert np.array_equal(r2.energies, r.energies)
    b.close()
