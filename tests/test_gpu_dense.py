"""Dense 4-qubit block (kq_debug_dense_block): the VALU form and the MFMA form of psi[r, :] <- U psi[r, :]
against the oracle.  U is built column by column from a kettle gate list on qubits 0..3 (the oracle's
simulate of each basis state), so applying it to a random n-qubit state must equal the oracle's
simulate of the same gates on that state; a Haar-random U is checked against numpy's row product.
Tolerances: complex128 within 1e-12, complex64 (FP32 arithmetic on a normalised state) within 2e-6;
the two forms within the same bounds of each other."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd import _lib, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = {ka.KQ_C128: 1e-12, ka.KQ_C64: 2e-6}


def gate_list_unitary(ops) -> np.ndarray:
    u = np.empty((16, 16), dtype=np.complex128)
    for j in range(16):
        e = np.zeros(16, dtype=np.complex128)
        e[j] = 1.0
        u[:, j], _ = orc.sv_simulate(e, 4, ops)
    return u


def apply_dense(state: ka.Statevector, u: np.ndarray, use_mfma: bool) -> float:
    flat = np.ascontiguousarray(u, dtype=np.complex128).view(np.float64).reshape(-1)
    ms = ctypes.c_double()
    _lib.check(_lib.load().kq_debug_dense_block(state.handle, _lib.ptr(flat), int(use_mfma), ctypes.byref(ms)))
    return ms.value


def mfma16(precision, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """One 16x16x4 matrix instruction on lane operands a[l], b[l]; returns [lane, element] of the result."""
    d = np.zeros(256)
    _lib.check(_lib.load().kq_debug_mfma16_layout(precision, _lib.ptr(np.ascontiguousarray(a, dtype=np.float64)),
                                                   _lib.ptr(np.ascontiguousarray(b, dtype=np.float64)), _lib.ptr(d)))
    return d.reshape(64, 4)


@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_mfma16_lane_layout(precision):
    """The layout the MFMA form relies on: A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k, and
    D[i][j] in lane j + 16 g, element e, with (g, e) = (i % 4, i // 4) for the f64 instruction and
    (i // 4, i % 4) for the f32 one.  Checked with a random integer product (exact in both precisions)
    against numpy, and with one-hot operands, so that a transposed or permuted operand fails."""
    rng = np.random.default_rng(5)
    A = rng.integers(-8, 9, size=(16, 4)).astype(np.float64)
    B = rng.integers(-8, 9, size=(4, 16)).astype(np.float64)
    lanes = np.arange(64)
    a = A[lanes % 16, lanes // 16]
    b = B[lanes // 16, lanes % 16]
    d = mfma16(precision, a, b)
    i, j = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    g, e = (i % 4, i // 4) if precision == ka.KQ_C128 else (i // 4, i % 4)
    np.testing.assert_array_equal(d[j + 16 * g, e], A @ B)
    for p in (0, 1, 5, 16, 37, 63):            # one A entry: only row p % 16 holds B's row p // 16
        onehot = np.zeros(64)
        onehot[p] = 1.0
        d = mfma16(precision, onehot, b)
        want = np.zeros((16, 16))
        want[p % 16] = B[p // 16]
        np.testing.assert_array_equal(d[j + 16 * g, e], want)


def random_state(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
@pytest.mark.parametrize("use_mfma", [False, True], ids=["valu", "mfma"])
@pytest.mark.parametrize("n", [8, 14])
def test_dense_block_equals_kettle_gates(precision, use_mfma, n):
    ops, _, _ = workloads.random_circuit(4, depth=8, seed=n + 3).flatten()
    u = gate_list_unitary(ops)
    assert np.allclose(u.conj().T @ u, np.eye(16), atol=1e-12)
    psi = random_state(n, n)
    want, _ = orc.sv_simulate(psi.copy(), n, ops)
    st = ka.Statevector(psi, precision)
    apply_dense(st, u, use_mfma)
    got = st.download()
    assert np.max(np.abs(got - want)) <= TOL[precision]


@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_dense_block_forms_agree_on_haar_unitary(precision):
    rng = np.random.default_rng(7)
    q, r = np.linalg.qr(rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16)))
    u = q * (np.diag(r) / np.abs(np.diag(r)))
    n = 20
    psi = random_state(n, 1)
    want = (psi.reshape(-1, 16) @ u.T).reshape(-1)
    out = []
    for use_mfma in (False, True):
        st = ka.Statevector(psi, precision)
        apply_dense(st, u, use_mfma)
        out.append(st.download())
        assert np.max(np.abs(out[-1] - want)) <= TOL[precision], use_mfma
    assert np.max(np.abs(out[0] - out[1])) <= TOL[precision]
    # a qubit-relabelled state is put back in order first: apply after a gate pass
    st = ka.Statevector(psi, precision)
    ka.apply_gates(st, workloads.random_circuit(n, depth=2, seed=4).flatten()[0])
    mid = st.download()
    apply_dense(st, u, True)
    assert np.max(np.abs(st.download() - (mid.reshape(-1, 16) @ u.T).reshape(-1))) <= TOL[precision]


def test_dense_block_errors():
    st = ka.Statevector(6, ka.KQ_C128)
    u = np.eye(16, dtype=np.complex128)
    with pytest.raises(ka.KqError):
        apply_dense(st, u, False)
    dm = ka.DensityMatrix(4)
    with pytest.raises(ka.KqError):
        apply_dense(dm, u, True)
