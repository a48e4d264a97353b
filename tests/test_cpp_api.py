"""The kettle-compatible C++20 API (include/kettle, libket.so over libkq's C-ABI).

The CPU test checks that libket.so and the C++ test executable are built and resolve every
libkq symbol they need; the GPU test runs the executable, whose cases mirror the reference's
own unit tests (simulate_test.cpp, control_flow_test.cpp, measure_test.cpp,
pauli_operator_test.cpp, density_matrix_test.cpp, n_local_test.cpp).
"""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIBKET = ROOT / "kettle_amd" / "lib" / "libket.so"
EXE = ROOT / "tests" / "cpp" / "ket_api_test"
HOST_EXE = ROOT / "tests" / "cpp" / "ket_host_test"


def _require_built() -> None:
    if not LIBKET.exists() or not EXE.exists() or not HOST_EXE.exists():
        pytest.fail("libket.so / tests/cpp executables missing: run __graft_entry__.build() first")


def test_libket_links_against_libkq():
    _require_built()
    out = subprocess.run(["ldd", str(EXE)], capture_output=True, text=True, check=True).stdout
    assert re.search(r"libket\.so => .*kettle_amd/lib/libket\.so", out), out
    assert re.search(r"libkq\.so => .*kettle_amd/lib/libkq\.so", out), out
    assert "not found" not in out, out
    undefined = subprocess.run(["nm", "-D", "--undefined-only", str(LIBKET)], capture_output=True, text=True,
                               check=True).stdout
    needed = set(re.findall(r" U (kq_\w+)", undefined))
    exported = subprocess.run(["nm", "-D", "--defined-only", str(LIBKET.with_name("libkq.so"))], capture_output=True,
                              text=True, check=True).stdout
    assert needed and needed <= set(re.findall(r" T (kq_\w+)", exported))


def test_libket_exports_simulate():
    _require_built()
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", str(LIBKET)], capture_output=True, text=True,
                         check=True).stdout
    for name in ["simulate", "expectation_value", "perform_measurements_as_counts_marginal", "partial_trace",
                 "energy_and_gradient", "n_local"]:
        assert re.search(rf" T ket::{name}(\[abi:\w+\])?\(", out), name


def test_cpp_host_cases():
    """Host-only C++ cases (readers, builders, Pauli bookkeeping) run without a device."""
    _require_built()
    proc = subprocess.run([str(HOST_EXE)], capture_output=True, text=True, timeout=120)
    print(proc.stdout[-4000:])
    failed = [line for line in proc.stdout.splitlines() if line.startswith("FAIL")]
    assert proc.returncode == 0 and not failed, "\n".join(failed) or proc.stderr


@pytest.mark.gpu
def test_cpp_api_cases():
    _require_built()
    proc = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=300)
    print(proc.stdout[-4000:])
    print(proc.stderr[-2000:])
    failed = [line for line in proc.stdout.splitlines() if line.startswith("FAIL")]
    assert proc.returncode == 0 and not failed, "\n".join(failed) or proc.stderr
