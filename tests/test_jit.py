"""Run-time compiled tile passes: generated kernels against the interpreted tile kernel and the oracle.

The generated pass kernel holds a pass's structure as straight-line code and reads only the gate
coefficients from memory; these tests check that it gives the oracle's amplitudes (complex128 within
1e-12, complex64 within 2e-5), that a structure repeated with other angles reuses the compiled kernel,
and (host only) that the generated sources of every gate kind compile for gfx950."""
from __future__ import annotations

import math

import numpy as np
import pytest

import kettle_amd as ka
from kat_cases import basis, random_gate_circuit
from oracle import oracle as orc

TOL128 = 1e-12
TOL64 = 2e-5


@pytest.fixture(autouse=True)
def _restore():
    ka.set_fusion(-1, 4096, 12, 1.0e9)
    yield
    ka.set_fusion(-1, 4096, 12, 1.0e9)
    ka.set_jit(1)


def ops_of(circuit):
    ops, _, _ = circuit.flatten()
    return ops


# ---- host only ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_generated_sources_compile(precision):
    """Every gate kind (controls on thread, register and block bits; merged diagonal runs)."""
    rng = np.random.default_rng(7 + precision)
    ops = ops_of(random_gate_circuit(12, 120, rng))
    assert ka.debug_jit_compile(ops, 12, precision) >= 1


@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_generated_adjoint_sources_compile(precision):
    """The run-time adjoint reverse-pass kernels (two vectors, derivative reductions) of a parameterised
    circuit with every gate class, controls on register, thread and block bits and merged diagonal runs."""
    import ctypes
    from kettle_amd import _lib, workloads
    rng = np.random.default_rng(40 + precision)
    c = random_gate_circuit(14, 120, rng)
    w = workloads.vqe(n=14, layers=3, n_random=10, seed=2)
    for ops in (ops_of(c), w.circuit.flatten(w.params)[0]):
        nk, nf = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(_lib.load().kq_debug_adjoint_jit_compile(14, precision, _lib.ptr(ops), len(ops), 160 * 1024,
                                                            ctypes.byref(nk), ctypes.byref(nf)))
        assert nk.value >= 1 and nf.value == 0


def test_generated_source_is_straight_line():
    rng = np.random.default_rng(3)
    ops = ops_of(random_gate_circuit(12, 80, rng))
    src, n_passes = ka.debug_jit_source(ops, 12, 0)
    assert n_passes >= 1
    assert 'extern "C" __global__' in src and "kq_pass" in src
    assert "switch" not in src.split("using R =")[1]   # no gate dispatch in the kernel body
    with pytest.raises(ka.KqError, match="out of range"):
        ka.debug_jit_source(ops, 12, n_passes)


def test_set_jit_rejects_bad_mode():
    with pytest.raises(ka.KqError, match="jit mode"):
        ka.set_jit(3)


# ---- device ----------------------------------------------------------------------------------------------

def run_sv(circuit, precision=ka.KQ_C128):
    sv = ka.Statevector(circuit.n_qubits(), precision)
    ka.simulate(circuit, sv)
    return sv.download()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [5, 8, 12, 15])
def test_jit_random_complex128(n):
    ka.set_jit(2)
    before = ka.jit_info()
    c = random_gate_circuit(n, 200, np.random.default_rng(4000 + n))
    out = run_sv(c)
    info = ka.jit_info()
    assert info["failures"] == before["failures"]
    assert info["compiled"] + info["cache_hits"] > before["compiled"] + before["cache_hits"]
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c))
    assert np.max(np.abs(out - ref)) <= TOL128


@pytest.mark.gpu
@pytest.mark.parametrize("n", [6, 13])
def test_jit_random_complex64(n):
    ka.set_jit(2)
    before = ka.jit_info()
    c = random_gate_circuit(n, 150, np.random.default_rng(5000 + n))
    out = run_sv(c, ka.KQ_C64)
    assert ka.jit_info()["failures"] == before["failures"]
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c))
    assert np.max(np.abs(out - ref)) <= TOL64


@pytest.mark.gpu
@pytest.mark.parametrize("k", [4, 8, 12])
def test_jit_matches_interpreter_bitwise_structure(k):
    """Same plan, generated vs interpreted: equal to rounding (both FP64, same operation order)."""
    n = 14
    c = random_gate_circuit(n, 300, np.random.default_rng(6000 + k))
    ka.set_fusion(k)
    ka.set_jit(0)
    interp = run_sv(c)
    ka.set_jit(2)
    jit = run_sv(c)
    assert np.max(np.abs(jit - interp)) <= 1e-13


@pytest.mark.gpu
def test_jit_reuses_kernels_for_new_angles():
    n = 12
    ka.set_jit(2)

    def layer_circuit(theta):
        c = ka.QuantumCircuit(n)
        for q in range(n):
            c.add_h_gate(q)
        for layer in range(3):
            for q in range(n):
                c.add_ry_gate(q, float(theta[layer, q]))
                c.add_rz_gate(q, float(theta[layer, q] * 0.5))
            for q in range(n - 1):
                c.add_cx_gate(q, q + 1)
            c.add_cp_gate(0, n - 1, float(theta[layer, 0]))
        return c

    rng = np.random.default_rng(11)
    c1 = layer_circuit(rng.uniform(0, 2 * math.pi, size=(3, n)))
    run_sv(c1)
    after_first = ka.jit_info()
    c2 = layer_circuit(rng.uniform(0, 2 * math.pi, size=(3, n)))
    out = run_sv(c2)
    after_second = ka.jit_info()
    assert after_second["compiled"] == after_first["compiled"]
    assert after_second["cache_hits"] > after_first["cache_hits"]
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c2))
    assert np.max(np.abs(out - ref)) <= TOL128


@pytest.mark.gpu
def test_jit_density_matrix():
    n = 6
    ka.set_jit(2)
    c = random_gate_circuit(n, 60, np.random.default_rng(77))
    rho0 = orc.sv_to_dm(basis("0" * n), n)
    ref, _ = orc.dm_simulate(rho0, n, ops_of(c))
    dm = ka.DensityMatrix(n)
    ka.simulate(c, dm)
    assert np.max(np.abs(dm.download() - ref)) <= TOL128


@pytest.mark.gpu
def test_jit_mode_one_leaves_small_states_interpreted():
    ka.set_jit(1)
    before = ka.jit_info()
    run_sv(random_gate_circuit(10, 50, np.random.default_rng(5)))
    after = ka.jit_info()
    assert after["compiled"] == before["compiled"] and after["cache_hits"] == before["cache_hits"]


@pytest.mark.gpu
@pytest.mark.parametrize("tuning", [(0, 1, True), (0, 2, False), (4, 0, True)],
                         ids=["persistent1", "persistent2-plain", "waves4"])
def test_jit_tunings_match_the_oracle(tuning):
    """Persistent blocks walking several tiles each (21 qubits: 512 tiles of 2^12 amplitudes), and a
    register cap, against the oracle."""
    n = 21
    c = random_gate_circuit(n, 120, np.random.default_rng(8000 + tuning[0] + tuning[1]))
    ka.set_jit(2)
    ka.set_jit_tuning(*tuning)
    try:
        before = ka.jit_info()
        out = run_sv(c)
        assert ka.jit_info()["failures"] == before["failures"]
    finally:
        ka.set_jit_tuning()
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c))
    assert np.max(np.abs(out - ref)) <= TOL128


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2], ids=["loads", "loads-stores"])
def test_jit_nontemporal_modes_match_the_oracle(mode):
    """Nontemporal loads (and stores) forced on a 21-qubit state (32 MiB, where the automatic mode leaves
    them off), against the oracle."""
    n = 21
    c = random_gate_circuit(n, 120, np.random.default_rng(8100 + mode))
    ka.set_jit(2)
    ka.set_jit_nontemporal(mode)
    try:
        out = run_sv(c)
    finally:
        ka.set_jit_nontemporal(3)
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c))
    assert np.max(np.abs(out - ref)) <= TOL128


def test_set_jit_nontemporal_rejects_bad_mode():
    from kettle_amd import _lib
    assert _lib.load().kq_set_jit_nontemporal(4) == _lib.KQ_ERR_INVALID
    assert _lib.load().kq_set_jit_nontemporal(-1) == _lib.KQ_ERR_INVALID
    ka.set_jit_nontemporal(3)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [ka.KQ_C128, ka.KQ_C64], ids=["c128", "c64"])
def test_jit_flipped_forms_20q(precision):
    """Six low qubits per tile and up to 16 stages on the layered circuit: one-qubit gates (dense, real,
    diagonal) and merged diagonal runs that meet controlled X gates still pending on their register bit
    run as X G X under the X gates' condition, against the oracle."""
    from kettle_amd import workloads
    n = 20
    c = workloads.qft_plus_layers(n=n, layers=12)
    ops = ops_of(c)
    ka.set_fusion_low_qubits(6)
    ka.set_fusion(-1, 4096, 16, 1.0e9)
    try:
        src = "".join(ka.debug_jit_source(ops, n, i)[0] for i in range(ka.debug_jit_source(ops, n, 0)[1]))
        assert "flip_coef<R>(coef_at" in src and "flip_coef<R>(Coef<R>" in src
        ka.set_jit(2)
        sv = ka.Statevector(n, precision)
        ka.apply_gates(sv, ops)
        out = sv.download()
    finally:
        ka.set_fusion_low_qubits(5)
        ka.set_fusion_autotune(True)
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops)
    assert np.max(np.abs(out - ref)) <= (TOL128 if precision == ka.KQ_C128 else TOL64)


@pytest.mark.gpu
@pytest.mark.parametrize("normalise", [True, False], ids=["normalised", "plain"])
def test_jit_layered_workload_20q(normalise):
    """QFT + random one-qubit layers with CX brick ladders at 20 qubits: controlled X gates with
    controls on thread and block bits (deferred to the next write as address XORs) and every
    normalised one-qubit form, against the oracle."""
    from kettle_amd import workloads
    n = 20
    c = workloads.qft_plus_layers(n=n, layers=6)
    ka.set_jit(2)
    ka.set_jit_tuning(0, 0, normalise)
    try:
        out = run_sv(c)
    finally:
        ka.set_jit_tuning()
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c))
    assert np.max(np.abs(out - ref)) <= TOL128


@pytest.mark.gpu
def test_jit_normalised_forms_at_extreme_angles():
    """Angles where a pivot entry is exactly 0 or where |cos| = |sin| (pivot choice ties), and U
    matrices with a zero diagonal, against the oracle (complex128 and complex64)."""
    from kettle_amd.circuit import Matrix2X2
    n = 13
    c = ka.QuantumCircuit(n)
    for q in range(n):
        c.add_h_gate(q)
    angles = [0.0, math.pi, -math.pi, math.pi / 2, -math.pi / 2, 3 * math.pi / 2, 2 * math.pi, 1e-300, 0.3]
    for i, theta in enumerate(angles * 3):
        q = i % n
        c.add_rx_gate(q, theta)
        c.add_ry_gate((q + 1) % n, theta)
        c.add_rz_gate((q + 2) % n, theta)
        c.add_p_gate((q + 3) % n, theta)
        c.add_cx_gate((q + 5) % n, (q + 4) % n)
    c.add_u_gate(Matrix2X2(0, 1j, 1j, 0), 3)
    c.add_u_gate(Matrix2X2(0, -1, 1, 0), 7)
    c.add_u_gate(Matrix2X2(1, 0, 0, -1), 9)
    for q in range(n):
        c.add_sx_gate(q)
        c.add_sxdag_gate((q + 6) % n)
    ka.set_jit(2)
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c))
    assert np.max(np.abs(run_sv(c) - ref)) <= TOL128
    assert np.max(np.abs(run_sv(c, ka.KQ_C64) - ref)) <= TOL64


@pytest.mark.gpu
@pytest.mark.parametrize("low", [2, 3, 4])
@pytest.mark.parametrize("jit", [0, 2], ids=["interpreted", "compiled"])
def test_low_qubit_counts_match_the_oracle(low, jit):
    """Tiles holding fewer of the lowest qubits than the default five (shorter contiguous chunks),
    through both pass kernels, against the oracle."""
    from kettle_amd import workloads
    n = 19
    c = workloads.qft_plus_layers(n=n, layers=4)
    ka.set_jit(jit)
    ka.set_fusion_low_qubits(low)
    try:
        out = run_sv(c)
    finally:
        ka.set_fusion_low_qubits(5)
    ref, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c))
    assert np.max(np.abs(out - ref)) <= TOL128


@pytest.mark.gpu
def test_plan_tuner_matches_the_oracle_and_caches_its_choice():
    """The plan tuner (switched off by the fixtures' explicit settings, on here) picks tile settings for a
    20-qubit circuit with a QFT and rotation layers; the state matches the oracle, and the same gate
    structure with other angles reuses the cached choice (no new candidate plans)."""
    n = 20

    def circuit(theta):
        c = ka.QuantumCircuit(n)
        c.add_qft_gate(list(range(n)))
        for layer in range(4):
            for q in range(n):
                c.add_ry_gate(q, float(theta[layer, q]))
                c.add_rz_gate(q, float(theta[layer, q] * 0.7))
            for q in range(layer % 2, n - 1, 2):
                c.add_cx_gate(q, q + 1)
        return c

    rng = np.random.default_rng(123)
    ka.set_jit(1)
    ka.set_fusion_autotune(True)
    try:
        c1 = circuit(rng.uniform(0, 2 * math.pi, size=(4, n)))
        out1 = run_sv(c1)
        first = ka.fusion_autotune_info()
        assert first["on"] and first["plans_tried"] > 0
        assert 4 <= first["low_qubits"] <= 7 and 12 <= first["max_stages"] <= 28
        ref1, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c1))
        assert np.max(np.abs(out1 - ref1)) <= TOL128
        c2 = circuit(rng.uniform(0, 2 * math.pi, size=(4, n)))
        out2 = run_sv(c2)
        again = ka.fusion_autotune_info()
        assert again["plans_tried"] == first["plans_tried"]
        ref2, _ = orc.sv_simulate(basis("0" * n), n, ops_of(c2))
        assert np.max(np.abs(out2 - ref2)) <= TOL128
    finally:
        ka.set_fusion_autotune(False)


def test_kernels_that_spill_under_the_register_cap_compile_without_it():
    """Host only: under a cap of four waves per SIMD (128 VGPRs) the passes of the 28-qubit QFT + layers
    circuit with 6 low qubits and up to 28 stages spill (private memory traffic on every tile); those
    compile again without the cap, and every kernel still compiles."""
    from kettle_amd import workloads
    ops = ops_of(workloads.qft_plus_layers(n=28, layers=20))
    before = ka.jit_info()["uncapped"]
    ka.set_fusion(-1, 4096, 28, 1.0e9)
    ka.set_fusion_low_qubits(6)
    ka.set_jit_tuning(4, 0, True)
    try:
        assert ka.debug_jit_compile(ops, 28, ka.KQ_C128) >= 1
    finally:
        ka.set_fusion_low_qubits(5)
        ka.set_jit_tuning()
    assert ka.jit_info()["uncapped"] > before
