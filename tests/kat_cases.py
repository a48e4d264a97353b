"""Known-answer cases restated from the reference's tests, plus an independent dense-matrix
simulator used to pin the oracle.

Sources (closed forms written as math here, not copied code):
  test/source/simulation/simulate_test.cpp:188-887   X/H/RX/RY/RZ/CX/CZ/H+CX/CRX/CRZ/CP answers
  test/source/simulation/simulate_test.cpp:890-1212  U / CU "mimic" equivalences
Bitstrings are little endian (character i = qubit i), as kettle's Statevector(string).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable

import numpy as np

import kettle_amd as ka
from kettle_amd.circuit import Gate

R = 1.0 / math.sqrt(2.0)
COS_PI_8 = math.cos(math.pi / 8)
SIN_PI_8 = math.sin(math.pi / 8)
COS_PI_16 = math.cos(math.pi / 16)
SIN_PI_16 = math.sin(math.pi / 16)
ANGLES = [0.0, math.pi / 6, math.pi / 4, math.pi / 3, math.pi / 2, math.pi / 1.5, 0.99 * math.pi, math.pi]


def basis(bitstring: str) -> np.ndarray:
    n = len(bitstring)
    a = np.zeros(1 << n, np.complex128)
    a[sum(1 << i for i, ch in enumerate(bitstring) if ch == "1")] = 1.0
    return a


@dataclass
class Kat:
    name: str
    n: int
    build: Callable[[ka.QuantumCircuit], None]
    initial: str
    expected: np.ndarray


def _kats() -> list[Kat]:
    out: list[Kat] = []
    add = out.append
    # X (simulate_test.cpp:188-247)
    add(Kat("X(0) 1q", 1, lambda c: c.add_x_gate(0), "0", np.array([0, 1], complex)))
    add(Kat("X(1) 2q", 2, lambda c: c.add_x_gate(1), "00", np.array([0, 0, 1, 0], complex)))
    add(Kat("X(0)X(1) 2q", 2, lambda c: (c.add_x_gate(0), c.add_x_gate(1)), "00", np.array([0, 0, 0, 1], complex)))
    # H (:249-328)
    add(Kat("H(0) 1q", 1, lambda c: c.add_h_gate(0), "0", np.array([R, R], complex)))
    add(Kat("H(0) 2q", 2, lambda c: c.add_h_gate(0), "00", np.array([R, R, 0, 0], complex)))
    add(Kat("H(1) 2q", 2, lambda c: c.add_h_gate(1), "00", np.array([R, 0, R, 0], complex)))
    add(Kat("H(0)H(1) 2q", 2, lambda c: (c.add_h_gate(0), c.add_h_gate(1)), "00", np.full(4, 0.5, complex)))
    # RX (:330-399)
    add(Kat("RX(pi/4,0) 1q", 1, lambda c: c.add_rx_gate(0, math.pi / 4), "0", np.array([COS_PI_8, -1j * SIN_PI_8])))
    add(Kat("RX(pi/4,0) 2q", 2, lambda c: c.add_rx_gate(0, math.pi / 4), "00",
            np.array([COS_PI_8, -1j * SIN_PI_8, 0, 0])))
    add(Kat("RX(pi/4,0)RX(pi/8,1)", 2, lambda c: (c.add_rx_gate(0, math.pi / 4), c.add_rx_gate(1, math.pi / 8)), "00",
            np.array([COS_PI_8 * COS_PI_16, -1j * SIN_PI_8 * COS_PI_16, -1j * COS_PI_8 * SIN_PI_16,
                      -SIN_PI_8 * SIN_PI_16])))
    # RY (:401-443): |0> -> cos(t/2)|0> + sin(t/2)|1>
    for t in ANGLES:
        add(Kat(f"RY({t:.3f},0)", 1, lambda c, t=t: c.add_ry_gate(0, t), "0",
                np.array([math.cos(t / 2), math.sin(t / 2)], complex)))
    # RZ (:445-519): |0> -> exp(-i t/2)|0>; on |+>: exp(-it/2)/sqrt2, exp(it/2)/sqrt2
    for t in ANGLES:
        add(Kat(f"RZ({t:.3f},0)", 1, lambda c, t=t: c.add_rz_gate(0, t), "0",
                np.array([complex(math.cos(t / 2), -math.sin(t / 2)), 0])))
        add(Kat(f"H RZ({t:.3f},1) 2q", 2, lambda c, t=t: (c.add_h_gate(1), c.add_rz_gate(1, t)), "00",
                np.array([R * complex(math.cos(t / 2), -math.sin(t / 2)), 0,
                          R * complex(math.cos(t / 2), math.sin(t / 2)), 0])))
    # CX (:521-607)
    for init, exp in [("00", "00"), ("10", "11"), ("01", "01"), ("11", "10")]:
        add(Kat(f"CX(0,1)|{init}>", 2, lambda c: c.add_cx_gate(0, 1), init, basis(exp)))
    for init, exp in [("00", "00"), ("10", "10"), ("01", "11"), ("11", "01")]:
        add(Kat(f"CX(1,0)|{init}>", 2, lambda c: c.add_cx_gate(1, 0), init, basis(exp)))
    for init, exp in [("100", "110"), ("110", "100"), ("101", "111"), ("001", "001")]:
        add(Kat(f"CX(0,1)|{init}> 3q", 3, lambda c: c.add_cx_gate(0, 1), init, basis(exp)))
    for init, exp in [("100", "101"), ("101", "100"), ("110", "111"), ("010", "010")]:
        add(Kat(f"CX(0,2)|{init}> 3q", 3, lambda c: c.add_cx_gate(0, 2), init, basis(exp)))
    # CZ (:609-695): phase -1 on control = target = 1
    for init in ["00", "10", "01", "11"]:
        sign = -1.0 if init == "11" else 1.0
        add(Kat(f"CZ(0,1)|{init}>", 2, lambda c: c.add_cz_gate(0, 1), init, sign * basis(init)))
        add(Kat(f"CZ(1,0)|{init}>", 2, lambda c: c.add_cz_gate(1, 0), init, sign * basis(init)))
    # H then CX: Bell state (:697-739)
    add(Kat("H(0) CX(0,1)", 2, lambda c: (c.add_h_gate(0), c.add_cx_gate(0, 1)), "00", np.array([R, 0, 0, R], complex)))
    # swap by three CX then X(0) (:719-739): |00> -X(0)-> |10> -swap-> |01>
    add(Kat("X(0) SWAP(0,1)", 2, lambda c: (c.add_x_gate(0), c.add_swap_gate(0, 1)), "00", basis("01")))
    # CRX (:741-798)
    for t in ANGLES:
        c_, s_ = math.cos(t / 2), math.sin(t / 2)
        add(Kat(f"CRX({t:.3f})|10>", 2, lambda c, t=t: c.add_crx_gate(0, 1, t), "10", np.array([0, c_, 0, -1j * s_])))
        add(Kat(f"CRX({t:.3f})|11>", 2, lambda c, t=t: c.add_crx_gate(0, 1, t), "11", np.array([0, -1j * s_, 0, c_])))
        add(Kat(f"CRX({t:.3f})|01>", 2, lambda c, t=t: c.add_crx_gate(0, 1, t), "01", basis("01")))
        add(Kat(f"H CRX({t:.3f})", 2, lambda c, t=t: (c.add_h_gate(0), c.add_crx_gate(0, 1, t)), "00",
                np.array([R, R * c_, 0, -1j * R * s_])))
        # CRZ (:800-853)
        add(Kat(f"H CRZ({t:.3f})", 2, lambda c, t=t: (c.add_h_gate(0), c.add_crz_gate(0, 1, t)), "00",
                np.array([R, R * complex(c_, -s_), 0, 0])))
        add(Kat(f"HH CRZ({t:.3f})", 2, lambda c, t=t: (c.add_h_gate(0), c.add_h_gate(1), c.add_crz_gate(0, 1, t)),
                "00", np.array([0.5, 0.5 * complex(c_, -s_), 0.5, 0.5 * complex(c_, s_)])))
        # CP (:855-888)
        add(Kat(f"CP({t:.3f})|11>", 2, lambda c, t=t: c.add_cp_gate(0, 1, t), "11",
                np.array([0, 0, 0, complex(math.cos(t), math.sin(t))])))
        add(Kat(f"CP({t:.3f})|10>", 2, lambda c, t=t: c.add_cp_gate(0, 1, t), "10", basis("10")))
    return out


KATS = _kats()


# U / CU mimic equivalences (simulate_test.cpp:890-1212): the U gate with a named gate's matrix
# must reproduce that gate.  Matrices are the textbook ones of gates/common_u_gates.cpp:10-172.
def gate_matrix(g: Gate, theta: float = 0.0) -> np.ndarray:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    table = {
        Gate.H: [[R, R], [R, -R]],
        Gate.X: [[0, 1], [1, 0]],
        Gate.Y: [[0, -1j], [1j, 0]],
        Gate.Z: [[1, 0], [0, -1]],
        Gate.S: [[1, 0], [0, 1j]],
        Gate.SDAG: [[1, 0], [0, -1j]],
        Gate.T: [[1, 0], [0, complex(R, R)]],
        Gate.TDAG: [[1, 0], [0, complex(R, -R)]],
        Gate.SX: [[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]],
        Gate.SXDAG: [[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]],
        Gate.RX: [[c, -1j * s], [-1j * s, c]],
        Gate.RY: [[c, -s], [s, c]],
        Gate.RZ: [[complex(c, -s), 0], [0, complex(c, s)]],
        Gate.P: [[1, 0], [0, complex(math.cos(theta), math.sin(theta))]],
    }
    base = {Gate.CH: Gate.H, Gate.CX: Gate.X, Gate.CY: Gate.Y, Gate.CZ: Gate.Z, Gate.CS: Gate.S,
            Gate.CSDAG: Gate.SDAG, Gate.CT: Gate.T, Gate.CTDAG: Gate.TDAG, Gate.CSX: Gate.SX,
            Gate.CSXDAG: Gate.SXDAG, Gate.CRX: Gate.RX, Gate.CRY: Gate.RY, Gate.CRZ: Gate.RZ, Gate.CP: Gate.P}
    return np.array(table[base.get(g, g)], dtype=np.complex128)


def to_m2(m: np.ndarray) -> ka.Matrix2X2:
    return ka.Matrix2X2(complex(m[0, 0]), complex(m[0, 1]), complex(m[1, 0]), complex(m[1, 1]))


# ---- independent dense simulator (brute-force Kronecker products), for pinning the oracle --------

def embed_1q(m: np.ndarray, t: int, n: int) -> np.ndarray:
    """2^n x 2^n operator of m on qubit t, qubit q = bit q of the index."""
    full = np.array([[1.0 + 0j]])
    for q in reversed(range(n)):
        full = np.kron(full, m if q == t else np.eye(2))
    return full


def embed_controlled(m: np.ndarray, c: int, t: int, n: int) -> np.ndarray:
    p1 = np.array([[0, 0], [0, 1]], complex)
    p0 = np.array([[1, 0], [0, 0]], complex)
    a = np.array([[1.0 + 0j]])
    b = np.array([[1.0 + 0j]])
    for q in reversed(range(n)):
        a = np.kron(a, p0 if q == c else np.eye(2))
        b = np.kron(b, p1 if q == c else (m if q == t else np.eye(2)))
    return a + b


def dense_unitary(ops, n: int) -> np.ndarray:
    """Product of the flattened straight-line gate list as one dense matrix."""
    u = np.eye(1 << n, dtype=np.complex128)
    for op in ops:
        g = Gate(int(op["gate"]))
        if g in (Gate.U, Gate.CU):
            m = np.array(op["u"][0::2] + 1j * op["u"][1::2]).reshape(2, 2)
        else:
            m = gate_matrix(g, float(op["angle"]))
        if g in (Gate.U,) or int(g) <= int(Gate.P):
            full = embed_1q(m, int(op["q0"]), n)
        else:
            full = embed_controlled(m, int(op["q0"]), int(op["q1"]), n)
        u = full @ u
    return u


def pauli_matrix(terms, n: int) -> np.ndarray:
    """Dense operator T_m ... T_1 of a term list applied in order (simulate_pauli.cpp:66-85)."""
    mats = {1: np.array([[0, 1], [1, 0]], complex), 2: np.array([[0, -1j], [1j, 0]], complex),
            3: np.array([[1, 0], [0, -1]], complex)}
    op = np.eye(1 << n, dtype=np.complex128)
    for q, k in terms:
        if int(k) == 0:
            continue
        op = embed_1q(mats[int(k)], int(q), n) @ op
    return op


ALL_GATES = [g for g in Gate if g != Gate.M]


def random_gate_circuit(n: int, n_gates: int, rng: np.random.Generator) -> ka.QuantumCircuit:
    """Every primitive gate kind on random qubits (controlled gates need n >= 2)."""
    from kettle_amd.workloads import random_unitary2x2

    c = ka.QuantumCircuit(n)
    gates = [g for g in ALL_GATES if n >= 2 or not (Gate.CH <= g <= Gate.CP or g == Gate.CU)]
    # start from a superposition so every gate acts non-trivially
    for q in range(n):
        c.add_h_gate(q)
    for _ in range(n_gates):
        g = gates[rng.integers(len(gates))]
        name = g.name.lower()
        if g == Gate.U:
            c.add_u_gate(random_unitary2x2(rng), int(rng.integers(n)))
        elif g == Gate.CU:
            a, b = rng.choice(n, size=2, replace=False)
            c.add_cu_gate(random_unitary2x2(rng), int(a), int(b))
        elif Gate.CH <= g <= Gate.CSXDAG:
            a, b = rng.choice(n, size=2, replace=False)
            getattr(c, f"add_{name}_gate")(int(a), int(b))
        elif Gate.CRX <= g <= Gate.CP:
            a, b = rng.choice(n, size=2, replace=False)
            getattr(c, f"add_{name}_gate")(int(a), int(b), float(rng.uniform(0, 2 * math.pi)))
        elif Gate.RX <= g <= Gate.P:
            getattr(c, f"add_{name}_gate")(int(rng.integers(n)), float(rng.uniform(0, 2 * math.pi)))
        else:
            getattr(c, f"add_{name}_gate")(int(rng.integers(n)))
    return c
