"""Synthetic workloads: the benchmark configurations as kettle circuits, plus ``n_local``.

Random draws use numpy's PCG64 seeded as stated (seed 0 unless given), so the GPU path, the oracle
and the CPU baseline all run the identical gate list.  Haar-random U(2) matrices follow kettle's
``generate_random_unitary2x2`` (gates/random_u_gates.cpp:13-36) with numpy uniforms.
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass

import numpy as np

from ._lib import OP_DTYPE

from .api import PackedPauliOperator, PauliOperator, SparsePauliString
from .circuit import Gate, Matrix2X2, QuantumCircuit


def random_unitary2x2(rng: np.random.Generator) -> Matrix2X2:
    """U(2) sample by kettle's construction: alpha, phi, chi uniform, psi = asin(sqrt(u))."""
    alpha = 2.0 * math.pi * rng.random()
    phi = 2.0 * math.pi * rng.random()
    chi = 2.0 * math.pi * rng.random()
    psi = math.asin(math.sqrt(rng.random()))
    g = complex(math.cos(alpha), math.sin(alpha))
    ep = complex(math.cos(phi), math.sin(phi))
    ec = complex(math.cos(chi), math.sin(chi))
    return Matrix2X2(g * math.cos(psi) * ep, g * math.sin(psi) * ec, -g * math.sin(psi) / ec, g * math.cos(psi) / ep)


ONE_Q_CHOICES = ("H", "SX", "T", "RX", "RY", "RZ", "U")
TWO_Q_CHOICES = ("CX", "CZ", "CRZ")


def add_random_1q_layer(c: QuantumCircuit, rng: np.random.Generator, qubits=None) -> None:
    for q in range(c.n_qubits()) if qubits is None else qubits:
        kind = ONE_Q_CHOICES[rng.integers(len(ONE_Q_CHOICES))]
        if kind == "H":
            c.add_h_gate(q)
        elif kind == "SX":
            c.add_sx_gate(q)
        elif kind == "T":
            c.add_t_gate(q)
        elif kind in ("RX", "RY", "RZ"):
            getattr(c, f"add_{kind.lower()}_gate")(q, float(rng.uniform(0.0, 2.0 * math.pi)))
        else:
            c.add_u_gate(random_unitary2x2(rng), q)


def add_brick_ladder(c: QuantumCircuit, rng: np.random.Generator, offset: int, kinds=TWO_Q_CHOICES) -> None:
    n = c.n_qubits()
    for i in range(offset, n - 1, 2):
        kind = kinds[rng.integers(len(kinds))]
        if kind == "CX":
            c.add_cx_gate(i, i + 1)
        elif kind == "CZ":
            c.add_cz_gate(i, i + 1)
        else:
            c.add_crz_gate(i, i + 1, float(rng.uniform(0.0, 2.0 * math.pi)))


def random_circuit(n: int, depth: int, seed: int = 0) -> QuantumCircuit:
    """Depth-``depth`` random circuit: a random 1-qubit layer then a CX/CZ/CRZ brick ladder on
    (i, i+1) pairs with the offset alternating 0/1."""
    rng = np.random.default_rng(seed)
    c = QuantumCircuit(n)
    for layer in range(depth):
        add_random_1q_layer(c, rng)
        add_brick_ladder(c, rng, layer % 2)
    return c


def random_pauli_packed(n: int, n_strings: int, rng: np.random.Generator) -> PackedPauliOperator:
    """Each site I/X/Y/Z with p = .7/.1/.1/.1; real weights N(0, 1) (all sites drawn first, then
    all weights)."""
    draws = rng.choice(4, size=(n_strings, n), p=[0.7, 0.1, 0.1, 0.1])
    weights = rng.normal(size=n_strings)
    rows, cols = np.nonzero(draws)
    offsets = np.concatenate([[0], np.cumsum(np.count_nonzero(draws, axis=1))])
    coeffs = np.stack([weights, np.zeros(n_strings)], axis=1).reshape(-1)
    return PackedPauliOperator(n, offsets, cols, draws[rows, cols], np.zeros(n_strings), coeffs)


def random_pauli_operator(n: int, n_strings: int, rng: np.random.Generator) -> PauliOperator:
    """``random_pauli_packed`` as a PauliOperator."""
    return random_pauli_packed(n, n_strings, rng).to_operator()


@dataclass
class PlainSVWorkload:
    circuit: QuantumCircuit
    hamiltonian: PauliOperator
    shots: int


def plain_statevector(n: int = 20, depth: int = 40, n_strings: int = 64, seed: int = 0) -> PlainSVWorkload:
    """Plain statevector workload: random circuit, 4096 shots, <H> over random Pauli strings."""
    c = random_circuit(n, depth, seed)
    rng = np.random.default_rng(seed + 1)
    return PlainSVWorkload(c, random_pauli_operator(n, n_strings, rng), 4096)


def qft_plus_layers(n: int = 30, layers: int = 20, seed: int = 0) -> QuantumCircuit:
    """QFT on all qubits (H + CP ladder + swap network), then ``layers`` random 1-qubit layers
    each followed by a CX brick ladder."""
    rng = np.random.default_rng(seed)
    c = QuantumCircuit(n)
    c.add_qft_gate(list(range(n)))
    for layer in range(layers):
        add_random_1q_layer(c, rng)
        add_brick_ladder(c, rng, layer % 2, kinds=("CX",))
    return c


def density_matrix_circuit(n: int = 12, depth: int = 30, seed: int = 0, measure_every: int = 10,
                           distinct_bits: bool = False) -> QuantumCircuit:
    """Random circuit as the plain workload with, after every ``measure_every`` layers, M on qubit 0
    into bit 0 and an if-else on that bit: X(0) when 1, RZ(0, pi/4) when 0.  With ``distinct_bits`` the
    k-th measurement writes (and its if-else reads) bit k, so the register keeps the whole outcome
    history; the gates are the same."""
    rng = np.random.default_rng(seed)
    c = QuantumCircuit(n)
    k = 0
    for layer in range(depth):
        add_random_1q_layer(c, rng)
        add_brick_ladder(c, rng, layer % 2)
        if (layer + 1) % measure_every == 0:
            bit = k if distinct_bits else 0
            c.add_m_gate(0, bit)
            if_sub = QuantumCircuit(n)
            if_sub.add_x_gate(0)
            else_sub = QuantumCircuit(n)
            else_sub.add_rz_gate(0, math.pi / 4)
            c.add_if_else_statement(bit, if_sub, else_sub)
            k += 1
    return c


def density_matrix_segments(n: int = 12, depth: int = 30, seed: int = 0, measure_every: int = 10) -> list[QuantumCircuit]:
    """The gate runs of ``density_matrix_circuit`` between its measurements (same gates, same seed), for
    replaying a recorded outcome history on a statevector."""
    rng = np.random.default_rng(seed)
    segs = [QuantumCircuit(n)]
    for layer in range(depth):
        add_random_1q_layer(segs[-1], rng)
        add_brick_ladder(segs[-1], rng, layer % 2)
        if (layer + 1) % measure_every == 0:
            segs.append(QuantumCircuit(n))
    return segs


class NLocalEntanglement(enum.IntEnum):
    FULL = 0
    LINEAR = 1


def n_local(n_qubits: int, rotation_blocks, entanglement_blocks, entanglement_kind: NLocalEntanglement,
            n_repetitions: int, skip_last_rotation_layer: bool = False):
    """kettle's ``n_local`` (optimize/n_local.cpp:275-305) for primitive rotation and
    1-control-1-target entanglement gates; angle gates are parameterised with initial value 0.0.
    Returns (circuit, parameter ids in creation order)."""
    c = QuantumCircuit(n_qubits)
    ids: list[str] = []

    def rotations():
        for g in rotation_blocks:
            g = Gate(g)
            if g in (Gate.RX, Gate.RY, Gate.RZ, Gate.P):
                for q in range(n_qubits):
                    ids.append(getattr(c, f"add_{g.name.lower()}_gate")(q, 0.0, True))
            elif g in (Gate.H, Gate.X, Gate.Y, Gate.Z, Gate.S, Gate.SDAG, Gate.T, Gate.TDAG, Gate.SX, Gate.SXDAG):
                for q in range(n_qubits):
                    getattr(c, f"add_{g.name.lower()}_gate")(q)
            elif g in (Gate.CRX, Gate.CRY, Gate.CRZ, Gate.CP):
                for i in range(n_qubits // 2):
                    ids.append(getattr(c, f"add_{g.name.lower()}_gate")(2 * i, 2 * i + 1, 0.0, True))
            else:
                for i in range(n_qubits // 2):
                    getattr(c, f"add_{g.name.lower()}_gate")(2 * i, 2 * i + 1)

    def entangle():
        for g in entanglement_blocks:
            g = Gate(g)
            if entanglement_kind == NLocalEntanglement.LINEAR:
                pairs = [(i, i + 1) for i in range(n_qubits - 1)]
            else:
                pairs = [(i, j) for i in range(n_qubits - 1) for j in range(i + 1, n_qubits)]
            for a, b in pairs:
                if g in (Gate.CRX, Gate.CRY, Gate.CRZ, Gate.CP):
                    ids.append(getattr(c, f"add_{g.name.lower()}_gate")(a, b, 0.0, True))
                else:
                    getattr(c, f"add_{g.name.lower()}_gate")(a, b)

    for _ in range(n_repetitions):
        rotations()
        entangle()
    if not skip_last_rotation_layer:
        rotations()
    return c, ids


def tfim_plus_random(n: int, n_random: int, rng: np.random.Generator, J: float = 1.0, h: float = 0.7) -> PauliOperator:
    """H = -J sum_i Z_i Z_{i+1 mod n} - h sum_i X_i + ``n_random`` random weighted strings."""
    op = PauliOperator(n_qubits=n)
    for i in range(n):
        op.add(-J, SparsePauliString([(i, 3), ((i + 1) % n, 3)], n))
    for i in range(n):
        op.add(-h, SparsePauliString([(i, 1)], n))
    rand = random_pauli_operator(n, n_random, rng)
    for c, s in rand.weighted_pauli_strings():
        op.add(c, s)
    return op


@dataclass
class VQEWorkload:
    circuit: QuantumCircuit
    params: list[str]
    hamiltonian: PauliOperator
    theta: np.ndarray


def vqe(n: int = 24, layers: int = 20, n_random: int = 200, seed: int = 0) -> VQEWorkload:
    """Hardware-efficient ansatz n_local(n, {RY, RZ}, {CX}, LINEAR, layers, skip last) with
    parameters uniform in [-pi, pi]; TFIM chain (J = 1, h = 0.7, periodic) + random strings."""
    c, ids = n_local(n, [Gate.RY, Gate.RZ], [Gate.CX], NLocalEntanglement.LINEAR, layers, True)
    rng = np.random.default_rng(seed)
    theta = rng.uniform(-math.pi, math.pi, size=len(ids))
    for pid, t in zip(ids, theta):
        c.set_parameter_value(pid, float(t))
    return VQEWorkload(c, ids, tfim_plus_random(n, n_random, rng), theta)


@dataclass
class BatchedVQEInstance:
    n: int
    depth: int
    theta: np.ndarray                 # parameter values in n_local creation order
    hamiltonian: PackedPauliOperator

    def ops(self) -> np.ndarray:
        """The n_local(n, {RY, RZ}, {CX}, LINEAR, depth, skip last) gate array with parameter slots."""
        return n_local_ry_rz_cx_ops(self.n, self.depth, self.theta)

    def circuit(self) -> tuple[QuantumCircuit, list[str]]:
        """The same instance as a parameterised QuantumCircuit (kettle's n_local)."""
        c, ids = n_local(self.n, [Gate.RY, Gate.RZ], [Gate.CX], NLocalEntanglement.LINEAR, self.depth, True)
        for pid, t in zip(ids, self.theta):
            c.set_parameter_value(pid, float(t))
        return c, ids


def n_local_ry_rz_cx_ops(n: int, depth: int, theta: np.ndarray) -> np.ndarray:
    """Op array of n_local(n, {RY, RZ}, {CX}, LINEAR, depth, skip last rotation layer): per
    repetition RY on every qubit, RZ on every qubit (parameter slots in that order), then CX(q, q+1)."""
    per = 3 * n - 1
    ops = np.zeros(depth * per, dtype=OP_DTYPE)
    q = np.arange(n, dtype=np.int32)
    for r in range(depth):
        blk = ops[r * per:(r + 1) * per]
        blk["gate"][:n] = int(Gate.RY)
        blk["gate"][n:2 * n] = int(Gate.RZ)
        blk["gate"][2 * n:] = int(Gate.CX)
        blk["q0"][:n] = q
        blk["q0"][n:2 * n] = q
        blk["q0"][2 * n:] = q[:-1]
        blk["q1"][2 * n:] = q[1:]
        slots = 2 * n * r + np.arange(2 * n)
        blk["aux"][:2 * n] = slots
        blk["aux"][2 * n:] = -1
        blk["angle"][:2 * n] = theta[slots]
    return ops


def batched_vqe(n_instances: int = 4096, n_range=(8, 22), depth_range=(4, 16), strings_range=(32, 256),
                seed: int = 0) -> list[BatchedVQEInstance]:
    """Uneven VQE instances: qubits uniform in n_range, ansatz depth uniform in depth_range of
    n_local(n, {RY, RZ}, {CX}, LINEAR, depth, skip last), parameters uniform in [-pi, pi], and a
    Hamiltonian of U{strings_range} random weighted Pauli strings (I/X/Y/Z with p = .7/.1/.1/.1,
    weights N(0, 1)).  Instance i draws from its own generator seeded 0x9E3779B9 * i + seed."""
    out = []
    for i in range(n_instances):
        rng = np.random.default_rng((0x9E3779B9 * i + seed) & 0xFFFFFFFFFFFFFFFF)
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        n_strings = int(rng.integers(strings_range[0], strings_range[1] + 1))
        theta = rng.uniform(-math.pi, math.pi, size=2 * n * depth)
        out.append(BatchedVQEInstance(n, depth, theta, random_pauli_packed(n, n_strings, rng)))
    return out
