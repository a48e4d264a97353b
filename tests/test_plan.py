"""The fusion planner on the CPU (kq_debug_plan: host-only, no device).

The planner reorders gates (a gate may run before earlier gates it commutes with) and groups them
into passes.  These tests replay the planned order on the oracle and require the same state as the
circuit order, for statevectors and density matrices, over tile sizes, stage limits and cost caps;
kq_debug_plan also lowers every pass and checks that each layout addresses every amplitude exactly
once.  Tolerance 1e-12 (reordering commuting gates changes rounding only)."""
from __future__ import annotations

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd.circuit import Gate
from kat_cases import gate_matrix, random_gate_circuit
from oracle import oracle as orc

CONTROLLED = {g for g in Gate if Gate.CH <= g <= Gate.CP} | {Gate.CU}


@pytest.fixture(autouse=True)
def _fusion_default():
    yield
    ka.set_fusion(-1, 4096, 12, 1.0e9)


def op_matrix(op) -> np.ndarray:
    g = Gate(int(op["gate"]))
    if g in (Gate.U, Gate.CU):
        u = op["u"]
        return np.array([[u[0] + 1j * u[1], u[2] + 1j * u[3]], [u[4] + 1j * u[5], u[6] + 1j * u[7]]])
    base = Gate[g.name[1:]] if g in CONTROLLED else g
    return gate_matrix(base, float(op["angle"]))


def as_u_op(m: np.ndarray, target: int, control: int | None) -> tuple:
    u = np.array([m[0, 0].real, m[0, 0].imag, m[0, 1].real, m[0, 1].imag,
                  m[1, 0].real, m[1, 0].imag, m[1, 1].real, m[1, 1].imag])
    if control is None:
        return (int(Gate.U), target, 0, -1, 0.0, u)
    return (int(Gate.CU), control, target, -1, 0.0, u)


def records(ops, n: int, density: bool) -> list[tuple]:
    """Each op as the U / CU records the device runs (two per op for a density matrix)."""
    out = []
    for op in ops:
        g = Gate(int(op["gate"]))
        ctrl = g in CONTROLLED
        t = int(op["q1"] if ctrl else op["q0"])
        c = int(op["q0"]) if ctrl else None
        m = op_matrix(op)
        out.append(as_u_op(m, t, c))
        if density:
            out.append(as_u_op(np.conj(m), t + n, None if c is None else c + n))
    return out


def to_ops(recs) -> np.ndarray:
    arr = np.zeros(len(recs), dtype=ka._lib.OP_DTYPE)
    for i, r in enumerate(recs):
        arr[i] = r
    return arr


def random_state(nq: int, rng) -> np.ndarray:
    a = rng.normal(size=1 << nq) + 1j * rng.normal(size=1 << nq)
    return a / np.linalg.norm(a)


def check_plan(circuit: ka.QuantumCircuit, density: bool, seed: int) -> ka.FusionPlan:
    n = circuit.n_qubits()
    ops, _, _ = circuit.flatten()
    plan = ka.debug_plan(ops, n, ka.KQ_C128, density)
    n_rec = len(ops) * (2 if density else 1)
    assert sorted(plan.order.tolist()) == list(range(n_rec))
    assert np.all(np.diff(plan.pass_of.astype(np.int64)) >= 0)
    recs = records(ops, n, density)
    nv = 2 * n if density else n
    start = random_state(nv, np.random.default_rng(seed))
    in_order, _ = orc.sv_simulate(start, nv, to_ops(recs))
    planned, _ = orc.sv_simulate(start, nv, to_ops([recs[i] for i in plan.order]))
    assert np.max(np.abs(in_order - planned)) <= 1e-12
    if density:
        # the record replay is the density-matrix update itself
        rho, _ = orc.dm_simulate(start, n, ops)
        assert np.max(np.abs(rho - in_order)) <= 1e-12
    return plan


@pytest.mark.parametrize("n", [4, 5, 7, 9, 12, 14])
@pytest.mark.parametrize("k,stages,cost", [(-1, 6, 1e9), (5, 2, 1e9), (8, 3, 4.0), (12, 8, 12.0), (4, 1, 1e9)])
def test_planned_order_statevector(n, k, stages, cost):
    ka.set_fusion(k, 4096, stages, cost)
    c = random_gate_circuit(n, 12 * n, np.random.default_rng(100 + n))
    plan = check_plan(c, False, n)
    assert np.all(plan.pass_cost[plan.pass_kind != 2] <= max(cost, 1.0) + 1e-12)


@pytest.mark.parametrize("n", [2, 3, 5, 6])
@pytest.mark.parametrize("k", [-1, 6])
def test_planned_order_density_matrix(n, k):
    ka.set_fusion(k, 4096, 4, 1e9)
    c = random_gate_circuit(n, 10 * n, np.random.default_rng(200 + n))
    check_plan(c, True, n)


def test_diagonal_runs_do_not_claim_tile_slots():
    """A long run of controlled phases on high qubits fuses into the tile of the low qubits."""
    n = 16
    c = ka.QuantumCircuit(n)
    for q in range(6):
        c.add_h_gate(q)
    for a in range(n):
        for b in range(a + 1, n):
            c.add_cp_gate(a, b, 0.1 * (a + b))
    for q in range(6):
        c.add_h_gate(q)
    ka.set_fusion(-1, 4096, 6, 1e9)
    plan = check_plan(c, False, 1)
    assert len(plan.pass_kind) == 1


def test_30q_workload_plan_shape():
    """The 30-qubit QFT + layers circuit: every pass a tile pass within the cost cap."""
    from kettle_amd import workloads

    ka.set_fusion(-1, 4096, 6, 16.0)
    ops, _, _ = workloads.qft_plus_layers(30, 20).flatten()
    plan = ka.debug_plan(ops, 30)
    assert sorted(plan.order.tolist()) == list(range(len(ops)))
    assert np.all(plan.pass_cost <= 16.0 + 1e-12)
    assert 20 <= len(plan.pass_kind) <= 80


def test_debug_plan_rejects_non_gates():
    c = ka.QuantumCircuit(2)
    c.add_m_gate(0)
    ops, _, _ = c.flatten()
    with pytest.raises(ka.KqError):
        ka.debug_plan(ops, 2)


@pytest.mark.parametrize("k", [4, 6, 8, 12])
@pytest.mark.parametrize("n", [5, 9, 13, 16])
def test_qubit_moves_and_restore_on_the_host(k, n):
    """Passes planned with qubit moves, replayed on the host on the index bits holding each qubit,
    the moves applied as each store would, then the restore passes: the oracle's state."""
    import ctypes

    from kettle_amd import _lib
    ka.set_fusion(k)
    c = random_gate_circuit(n, 150, np.random.default_rng(31 * k + n))
    ops, _, _ = c.flatten()
    amps = np.zeros(1 << n, np.complex128)
    amps[0] = 1
    n_passes, n_restore = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(_lib.load().kq_debug_check_relabel(n, _lib.ptr(ops), len(ops), _lib.ptr(amps), ctypes.byref(n_passes),
                                                   ctypes.byref(n_restore)))
    start = np.zeros(1 << n, np.complex128)
    start[0] = 1
    ref, _ = orc.sv_simulate(start, n, ops)
    assert np.max(np.abs(amps - ref)) <= 1e-12


def exchange_circuit(n, rng):
    """Random layers with SWAP gates (three CX each) between them, near misses that are not
    exchanges (CX(a, b) CX(b, a) CX(b, a); an exchange split by another gate), then a QFT (whose
    swap network is SWAP gates)."""
    from kettle_amd.workloads import add_brick_ladder, add_random_1q_layer
    c = ka.QuantumCircuit(n)
    for layer in range(10):
        add_random_1q_layer(c, rng)
        add_brick_ladder(c, rng, layer % 2)
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        c.add_swap_gate(a, b)
        c.add_rz_gate(a, float(rng.uniform(0, 6)))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        c.add_cx_gate(a, b)
        c.add_cx_gate(b, a)
        c.add_cx_gate(b, a)
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        c.add_cx_gate(a, b)
        c.add_cx_gate(b, a)
        c.add_ry_gate(a, float(rng.uniform(0, 6)))
        c.add_cx_gate(a, b)
    c.add_qft_gate(list(range(n)))
    return c


@pytest.mark.parametrize("k", [6, 12])
@pytest.mark.parametrize("n", [5, 9, 14])
def test_exchanges_as_relabels_on_the_host(k, n):
    """SWAP gates dropped and the gates after them relabelled, the rest planned with qubit moves and
    replayed on the host, then the restore passes: the oracle's state (which runs every CX)."""
    import ctypes

    from kettle_amd import _lib
    ka.set_fusion(k)
    ops, _, _ = exchange_circuit(n, np.random.default_rng(17 * k + n)).flatten()
    amps = np.zeros(1 << n, np.complex128)
    amps[0] = 1
    n_passes, n_restore = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(_lib.load().kq_debug_check_relabel(n, _lib.ptr(ops), len(ops), _lib.ptr(amps), ctypes.byref(n_passes),
                                                   ctypes.byref(n_restore)))
    start = np.zeros(1 << n, np.complex128)
    start[0] = 1
    ref, _ = orc.sv_simulate(start, n, ops)
    assert np.max(np.abs(amps - ref)) <= 1e-12


def test_qubit_moves_cut_the_30q_pass_count():
    from kettle_amd import workloads
    c = workloads.qft_plus_layers(n=30, layers=20)
    ops, _, _ = c.flatten()
    try:
        ka.set_fusion_relabel(False)
        fixed = len(ka.debug_plan(ops, 30).pass_cost)
        ka.set_fusion_relabel(True)
        moved = len(ka.debug_plan(ops, 30).pass_cost)
    finally:
        ka.set_fusion_relabel(True)
    assert moved < fixed
