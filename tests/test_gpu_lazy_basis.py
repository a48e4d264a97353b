"""Pending basis states and known zeros: kq_state_set_basis (and the |0...0> of a new state) writes
nothing until the state's next operation.  A gate run whose first pass is a compiled tile pass builds
that pass's tiles from the basis index (no read of the buffer); every other operation writes the basis
vector first.  Until a qubit has been in a pass's tile its index bit keeps the basis value wherever the
state is non-zero, so later passes skip (in place) or zero-fill (out of place) the tiles whose outside
bits differ.  Checked against the oracle run from the same basis state (tolerances as the other parity
tests: complex128 1e-12, complex64 2e-5 max-abs)."""
from __future__ import annotations

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd import _lib
from kettle_amd.workloads import random_unitary2x2
from kat_cases import random_gate_circuit
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def basis_vec(n: int, b: int) -> np.ndarray:
    v = np.zeros(1 << n, dtype=np.complex128)
    v[b] = 1.0
    return v


def run_bytes(fn) -> float:
    """Algorithmic bytes the pass profiler counts over fn() (reads and writes of every pass)."""
    lib = _lib.load()
    stats = np.zeros(len(_lib.PASS_CLASSES), dtype=_lib.PASS_STATS_DTYPE)
    _lib.check(lib.kq_profile_begin())
    fn()
    _lib.check(lib.kq_synchronize())
    _lib.check(lib.kq_profile_end(_lib.ptr(stats)))
    return float(stats["bytes"].sum()), int(stats["launches"].sum())


@pytest.mark.parametrize("n,precision,tol", [(18, ka.KQ_C128, 1e-12), (20, ka.KQ_C64, 2e-5), (21, ka.KQ_C128, 1e-12)])
def test_compiled_run_from_pending_basis(n, precision, tol):
    rng = np.random.default_rng(7 + n)
    ops, _, _ = random_gate_circuit(n, 80, rng).flatten()
    sv = ka.Statevector(n, precision)
    ab = 16 if precision == ka.KQ_C128 else 8
    for b in (0, int(rng.integers(1 << n)), (1 << n) - 1):
        print(f"n={n} precision={precision} basis={b}", flush=True)
        sv.upload(basis_vec(n, b))
        eager_bytes, eager_launches = run_bytes(lambda: ka.apply_gates(sv, ops))
        sv.set_basis(b)
        nbytes, launches = run_bytes(lambda: ka.apply_gates(sv, ops))
        # the same passes; the first only writes, later ones skip known-zero tiles
        assert launches == eager_launches >= 1
        assert nbytes <= eager_bytes - ab * (1 << n) + 1.0
        want = orc.sv_simulate(basis_vec(n, b), n, ops)[0]
        np.testing.assert_allclose(sv.download(), want, atol=tol, rtol=0)


def test_new_state_is_zero_and_reads_write_the_pending_basis():
    n = 19
    sv = ka.Statevector(n)
    np.testing.assert_array_equal(sv.download(), basis_vec(n, 0))
    b = 0x5A5A5
    sv.set_basis(b)
    np.testing.assert_array_equal(ka.calculate_probabilities_raw(sv), np.abs(basis_vec(n, b)) ** 2)
    sv.set_basis(3)
    sv.set_basis(b)                       # the later reset wins
    np.testing.assert_array_equal(sv.download(), basis_vec(n, b))
    # a run of no gates leaves the reset pending; the read then writes it
    sv.set_basis(11)
    ka.apply_gates(sv, np.zeros(0, dtype=_lib.OP_DTYPE))
    np.testing.assert_array_equal(sv.download(), basis_vec(n, 11))


def test_overwrites_drop_the_pending_basis():
    n = 18
    rng = np.random.default_rng(3)
    x = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    x /= np.linalg.norm(x)
    sv = ka.Statevector(n)
    sv.set_basis(77)
    sv.upload(x)
    np.testing.assert_allclose(sv.download(), x, atol=0, rtol=0)
    src = ka.Statevector(n)
    src.upload(x)
    dst = ka.Statevector(n)
    dst.set_basis(5)
    _lib.check(_lib.load().kq_state_copy(dst.handle, src.handle))
    np.testing.assert_allclose(dst.download(), x, atol=0, rtol=0)
    # a pending source is written before it is copied
    src.set_basis(9)
    _lib.check(_lib.load().kq_state_copy(dst.handle, src.handle))
    np.testing.assert_array_equal(dst.download(), basis_vec(n, 9))


def test_relabel_only_run_on_pending_basis():
    """A run of SWAP gates only (qubit relabels, no pass) keeps the reset pending; the read writes it and
    then puts every qubit back on its own bit."""
    n = 18
    c = ka.QuantumCircuit(n)
    c.add_swap_gate(0, 17)
    c.add_swap_gate(3, 9)
    ops, _, _ = c.flatten()
    sv = ka.Statevector(n)
    b = (1 << 0) | (1 << 3) | (1 << 12)
    sv.set_basis(b)
    ka.apply_gates(sv, ops)
    want = orc.sv_simulate(basis_vec(n, b), n, ops)[0]
    np.testing.assert_array_equal(sv.download(), want)


@pytest.mark.parametrize("jit", [0, 1], ids=["interpreted", "compiled"])
def test_interpreted_and_small_runs_from_pending_basis(jit):
    ka.set_jit(jit)
    try:
        for n in (3, 9, 18):
            rng = np.random.default_rng(100 + n)
            ops, _, _ = random_gate_circuit(n, 40, rng).flatten()
            b = int(rng.integers(1 << n))
            sv = ka.Statevector(n)
            sv.set_basis(b)
            ka.apply_gates(sv, ops)
            want = orc.sv_simulate(basis_vec(n, b), n, ops)[0]
            np.testing.assert_allclose(sv.download(), want, atol=1e-12, rtol=0, err_msg=f"n={n}")
    finally:
        ka.set_jit(1)


def test_density_matrix_from_pending_basis():
    n = 10                                   # vec(rho) of 20 index bits: compiled passes
    rng = np.random.default_rng(42)
    c = random_gate_circuit(n, 30, rng)
    ops, _, _ = c.flatten()
    b = int(rng.integers(1 << n))
    N = 1 << n
    dm = ka.DensityMatrix(n)
    dm.set_basis(b)
    ka.simulate(c, dm)
    rho0 = np.zeros(N * N, dtype=np.complex128)
    rho0[b + b * N] = 1.0
    want = orc.dm_simulate(rho0, n, ops)[0]
    np.testing.assert_allclose(dm.download(), want, atol=1e-12, rtol=0)


@pytest.mark.parametrize("n,groups", [(20, [(0, 6), (14, 20)]), (22, [(3, 9), (16, 22)]), (21, [(0, 21)])])
def test_runs_that_leave_qubits_untouched(n, groups):
    """Gates on some qubits only, from a basis state with every bit set (the untouched qubits keep their
    1): runs of several passes with qubit moves, where later passes (the last one out of place) start
    from known outside bits; then a second run on the result, and a read in between."""
    rng = np.random.default_rng(n)
    c = ka.QuantumCircuit(n)
    qs = [q for lo, hi in groups for q in range(lo, hi)]
    for layer in range(6):
        for q in qs:
            c.add_u_gate(random_unitary2x2(rng), q)
        for a, b in zip(qs[layer % 2::2], qs[layer % 2 + 1::2]):
            c.add_cx_gate(a, b)
    ops, _, _ = c.flatten()
    b = (1 << n) - 1
    sv = ka.Statevector(n)
    sv.set_basis(b)
    nbytes, launches = run_bytes(lambda: ka.apply_gates(sv, ops))
    assert launches >= 1
    assert nbytes < 2 * launches * 16 * (1 << n)
    want = orc.sv_simulate(basis_vec(n, b), n, ops)[0]
    # a second run continues from the known zeros the first left
    ka.apply_gates(sv, ops)
    want = orc.sv_simulate(want, n, ops)[0]
    np.testing.assert_allclose(sv.download(), want, atol=1e-12, rtol=0)
    # after a read the zeros are no longer assumed: a third run reads everything
    ka.apply_gates(sv, ops)
    want = orc.sv_simulate(want, n, ops)[0]
    np.testing.assert_allclose(sv.download(), want, atol=1e-12, rtol=0)
