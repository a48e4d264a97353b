"""Hamiltonian text reader (read_pauli_operator), CPU only.  Known answer from the reference's
test/source/io/read_pauli_operator_test.cpp; the energy of a read operator is checked against
the oracle's expectation of the same strings built by hand."""
from __future__ import annotations

import io

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd.io import read_pauli_operator
from oracle import oracle as orc

KAT = (" 1.100000000000e+01    0.000000000000e+00   :                                                         \n"
       "-1.875000000000e-01    0.000000000000e+00   :   (0, X)   (1, X)   (2, X)   (3, X)   (4, X)   (5, X)   \n"
       " 1.875000000000e-01    0.000000000000e+00   :   (0, X)   (1, X)   (2, X)   (3, Z)                     \n")


def test_known_answer():
    op = read_pauli_operator(io.StringIO(KAT), 6)
    assert op.size() == 3 and op.n_qubits() == 6
    (c0, s0), (c1, s1), (c2, s2) = op.weighted_pauli_strings()
    assert (c0, c1, c2) == (11.0, -0.1875, 0.1875)
    assert s0.terms() == []
    assert s1.terms() == [(q, ka.PauliTerm.X) for q in range(6)]
    assert s2.terms() == [(0, ka.PauliTerm.X), (1, ka.PauliTerm.X), (2, ka.PauliTerm.X), (3, ka.PauliTerm.Z)]


def test_errors_and_complex_weights(tmp_path):
    op = read_pauli_operator(io.StringIO("0.5 -0.25 : (2, Y) (0, Z)\n"), 3)
    (c, s), = op.weighted_pauli_strings()
    assert c == complex(0.5, -0.25) and s.terms() == [(2, ka.PauliTerm.Y), (0, ka.PauliTerm.Z)]
    with pytest.raises(RuntimeError, match="already present"):
        read_pauli_operator(io.StringIO("1 0 : (1, X) (1, Z)\n"), 2)
    with pytest.raises(RuntimeError, match="beyond qubit range"):
        read_pauli_operator(io.StringIO("1 0 : (4, X)\n"), 2)
    with pytest.raises(KeyError):
        read_pauli_operator(io.StringIO("1 0 : (0, Q)\n"), 2)
    with pytest.raises(OSError, match="unable to open"):
        read_pauli_operator(tmp_path / "missing.txt", 2)
    path = tmp_path / "h.txt"
    path.write_text(KAT)
    assert read_pauli_operator(path, 6).size() == 3


def test_read_operator_energy_on_the_oracle():
    op = read_pauli_operator(io.StringIO(KAT), 6)
    rng = np.random.default_rng(3)
    psi = rng.normal(size=64) + 1j * rng.normal(size=64)
    psi /= np.linalg.norm(psi)
    e = orc.pauli_expectation(psi, 6, op.packed())
    xs = np.arange(64)
    flip_all = psi[xs ^ 0b111111]
    flip3 = psi[xs ^ 0b000111] * np.where((xs >> 3) & 1, -1.0, 1.0)
    expected = 11.0 - 0.1875 * np.vdot(psi, flip_all) + 0.1875 * np.vdot(psi, flip3)
    assert abs(e - expected) <= 1e-12
