"""Statevector snapshots (save_statevector / load_statevector) and the numpy text reader.

Known answers from the reference's test/source/io/numpy_statevector_test.cpp:12-49 and
statevector_test.cpp:11-33 (save a simulated 3-qubit state, load it back, compare); line format
from io/statevector.cpp:45-92 ("%.14f", a space before non-negative values, BIG lists amplitudes
in bit-reversed index order)."""
from __future__ import annotations

import io
import math

import numpy as np
import pytest

import kettle_amd as ka
from kettle_amd.io import (Endian, bit_reversal, parse_numpy_statevector, parse_statevector_text,
                           statevector_text)
from oracle import oracle as orc

NUMPY_KAT = ("3                                                    \n"
             " (4.950737714883371443e-02+4.950737714883371443e-02j)\n"
             " (9.901475429766742886e-02+9.901475429766742886e-02j)\n"
             " (1.485221314465011433e-01+1.485221314465011433e-01j)\n"
             " (1.980295085953348577e-01+1.980295085953348577e-01j)\n"
             " (2.475368857441685444e-01+2.475368857441685444e-01j)\n"
             " (2.970442628930022866e-01+2.970442628930022866e-01j)\n"
             " (3.465516400418360288e-01+3.465516400418360288e-01j)\n"
             " (3.960590171906697154e-01+3.960590171906697154e-01j)\n")


def test_numpy_known_answer():
    amps = parse_numpy_statevector(NUMPY_KAT)
    want = np.array([1.1 * i * (1 + 1j) for i in range(1, 9)])
    want /= np.linalg.norm(want)
    np.testing.assert_allclose(amps, want, atol=1e-15, rtol=0)


def test_numpy_negative_parts_and_big_endian():
    text = "2\n (1.0e+00-0.0e+00j)\n (-2.5e-01+5.0e-01j)\n (3.0e-01-7.0e-01j)\n (0.0e+00+1.0e+00j)\n"
    little = parse_numpy_statevector(text)
    np.testing.assert_array_equal(little, [1.0, -0.25 + 0.5j, 0.3 - 0.7j, 1.0j])
    big = parse_numpy_statevector(text, Endian.BIG)
    np.testing.assert_array_equal(big, little[[0, 2, 1, 3]])


def test_bit_reversal_is_an_involution():
    for n in range(1, 9):
        r = bit_reversal(n)
        assert np.array_equal(r[r], np.arange(1 << n))
    assert list(bit_reversal(3)) == [0, 4, 2, 6, 1, 5, 3, 7]


def test_snapshot_format_and_round_trip():
    rng = np.random.default_rng(4)
    amps = rng.normal(size=16) + 1j * rng.normal(size=16)
    amps[3] = -0.0 + 0.25j
    amps /= np.linalg.norm(amps)
    text = statevector_text(amps)
    lines = text.splitlines()
    assert lines[0] == "ENDIANNESS: LITTLE" and lines[1] == "NUMBER OF STATES: 16"
    z = amps[0]
    assert lines[2] == (" " if z.real >= 0 else "") + f"{z.real:.14f}" + "   " + (" " if z.imag >= 0 else "") + f"{z.imag:.14f}"
    np.testing.assert_allclose(parse_statevector_text(text), amps, atol=1e-14, rtol=0)
    big = statevector_text(amps, Endian.BIG)
    assert big.splitlines()[0] == "ENDIANNESS: BIG"
    assert big.splitlines()[2 + 1] == text.splitlines()[2 + 8]        # line 1 of BIG = amplitude 0b1000
    np.testing.assert_allclose(parse_statevector_text(big), amps, atol=1e-14, rtol=0)


def test_errors(tmp_path):
    with pytest.raises(RuntimeError, match="string_to_endian"):
        parse_statevector_text("ENDIANNESS: MIDDLE\nNUMBER OF STATES: 2\n 1.0 0.0\n 0.0 0.0\n")
    with pytest.raises(OSError, match="unable to open file to load statevector"):
        ka.load_statevector(tmp_path / "missing.txt")
    with pytest.raises(OSError, match="unable to open file for statevector"):
        ka.read_numpy_statevector(tmp_path / "missing.txt")


@pytest.mark.gpu
@pytest.mark.parametrize("endian", [Endian.LITTLE, Endian.BIG])
def test_device_state_round_trip(tmp_path, endian):
    """statevector_test.cpp:11-33: the simulated state saved and loaded back."""
    c = ka.QuantumCircuit(3)
    c.add_h_gate([0, 1, 2])
    c.add_x_gate([0])
    c.add_y_gate([1, 2])
    c.add_rx_gate([(0, math.pi / 4), (2, math.pi / 2)])
    sv = ka.Statevector(3)
    ka.simulate(c, sv)
    path = tmp_path / "state.txt"
    ka.save_statevector(path, sv, endian)
    back = ka.load_statevector(path)
    np.testing.assert_allclose(back.download(), sv.download(), atol=1e-13, rtol=0)
    ops, _, _ = c.flatten()
    ref, _ = orc.sv_simulate(orc.zero_state(3), 3, ops)
    np.testing.assert_allclose(back.download(), ref, atol=1e-13, rtol=0)
    numpy_text = "3\n" + "".join(f" ({z.real:.18e}{z.imag:+.18e}j)\n" for z in ref)
    np.testing.assert_allclose(ka.read_numpy_statevector(io.StringIO(numpy_text)).download(), ref, atol=1e-16, rtol=0)
