"""Hamiltonian text reader: kettle's ``read_pauli_operator``
(io/read_pauli_operator.cpp:29-69 in the reference).

One weighted Pauli string per line, ``<re> <im> : (q, P) (q, P) ...`` with P in {X, Y, Z}; a line
with no terms is the identity string.
"""
from __future__ import annotations

import io
import os
import re
from typing import TextIO

from .api import PauliOperator, PauliTerm, SparsePauliString

_TERMS = {"X": PauliTerm.X, "Y": PauliTerm.Y, "Z": PauliTerm.Z}
_TERM_RE = re.compile(r"\(\s*(\d+)\s*,\s*(\S)\s*\)")


def read_pauli_operator(source: str | os.PathLike | TextIO, n_qubits: int) -> PauliOperator:
    """``source``: a path, or an open text stream."""
    if isinstance(source, (str, os.PathLike)):
        try:
            with open(source, encoding="utf-8") as f:
                return _read(f, n_qubits)
        except FileNotFoundError as e:
            raise OSError(f"ERROR: unable to open: {os.fspath(source)}") from e
    return _read(source, n_qubits)


def _read(stream: TextIO | io.StringIO, n_qubits: int) -> PauliOperator:
    op = PauliOperator(n_qubits=n_qubits)
    for line in stream:
        if not line.strip():
            continue
        head, _, tail = line.partition(":")
        re_part, im_part = head.split()[:2]
        s = SparsePauliString([], n_qubits)
        for q, t in _TERM_RE.findall(tail):
            if t not in _TERMS:
                raise KeyError(f"ERROR: unknown Pauli term '{t}'")
            s.add(int(q), _TERMS[t])
        op.add(complex(float(re_part), float(im_part)), s)
    return op
