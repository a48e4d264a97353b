"""Build libkq (HIP, gfx950) and the CPU oracle in-tree.

    python -m kettle_amd.build            # both libraries
    python -m kettle_amd.build --kq       # libkq only

libkq.so lands in kettle_amd/lib/, the oracle in oracle/liboracle.so.  Every HIP translation unit
is compiled for gfx950 only, with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "kettle_amd" / "csrc"
LIB_DIR = ROOT / "kettle_amd" / "lib"
BUILD_DIR = ROOT / "kettle_amd" / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

KQ_SOURCES = ["kq_kernels.hip", "kq_tile.hip", "kq_plan.cpp", "kq_api.cpp", "kq_run.cpp", "kq_measure.cpp", "kq_pauli.cpp",
              "kq_adjoint.cpp", "kq_debug.cpp", "kq_batch.cpp", "kq_jit.cpp", "kq_dense.hip"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-I", str(ROOT / "include"),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    # hipcc's host pass does not know the packed-fp32 feature; drop that one known warning
    noise = [line for line in proc.stderr.splitlines() if "packed-fp32-ops" not in line and line.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


# per-source extra flags.  The tile kernel dispatches gates to many specialised, structurally
# identical code paths; LLVM's common-code sinking would merge them into one block that addresses
# the per-thread amplitude slots through a phi of pointers, which forces the slots into scratch.
EXTRA_FLAGS = {"kq_tile.hip": ["-mllvm", "-simplifycfg-sink-common=false"]}


def embed_device_header() -> Path:
    """The device templates as a C string for the run-time compiler (kq_jit.cpp)."""
    text = (CSRC / "kq_tile_device.hpp").read_text().replace("#pragma once\n", "")
    out = BUILD_DIR / "kq_tile_device.inc"
    out.write_text('R"KQDEV(' + text + ')KQDEV"\n')
    return out


def build_kq(jobs: int = 8) -> Path:
    BUILD_DIR.mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    embed_device_header()
    objs = []
    cmds = []
    for src in KQ_SOURCES:
        obj = BUILD_DIR / (src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC, *HIP_FLAGS, "-I", str(BUILD_DIR), *EXTRA_FLAGS.get(src, []), *lang, "-c", str(CSRC / src),
                     "-o", str(obj)])
        objs.append(str(obj))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_run, c) for c in cmds]:
            fut.result()
    out = LIB_DIR / "libkq.so"
    tmp = LIB_DIR / "libkq.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs, "-L/opt/rocm/lib", "-lhiprtc",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, out)
    return out


KET_SOURCES = ["ket_common.cpp", "ket_state.cpp", "ket_circuit.cpp", "ket_simulate.cpp", "ket_measure.cpp",
               "ket_pauli.cpp", "ket_optimize.cpp", "ket_io.cpp"]
HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-Wall", "-Wextra", "-I", str(ROOT / "include")]


def build_ket(jobs: int = 4) -> Path:
    """The C++20 kettle-compatible host library (ket:: API) over libkq's C-ABI."""
    BUILD_DIR.mkdir(parents=True, exist_ok=True)
    host = ROOT / "kettle_amd" / "host"
    objs = []
    cmds = []
    for src in KET_SOURCES:
        obj = BUILD_DIR / (src + ".o")
        cmds.append(["g++", *HOST_FLAGS, "-c", str(host / src), "-o", str(obj)])
        objs.append(str(obj))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_run, c) for c in cmds]:
            fut.result()
    out = LIB_DIR / "libket.so"
    tmp = LIB_DIR / "libket.so.tmp"
    _run(["g++", "-shared", "-o", str(tmp), *objs, "-L", str(LIB_DIR), "-lkq", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


CPP_TESTS = ["ket_api_test", "ket_host_test"]


def build_cpp_tests() -> list[Path]:
    """The C++ API test executables (tests/cpp), linked against libket and libkq: ket_api_test
    runs on the GPU, ket_host_test covers the host-only parts (builders, readers, validation)."""
    outs = []
    for name in CPP_TESTS:
        src = ROOT / "tests" / "cpp" / f"{name}.cpp"
        out = ROOT / "tests" / "cpp" / name
        _run(["g++", *HOST_FLAGS, str(src), "-o", str(out), "-L", str(LIB_DIR), "-lket", "-lkq",
              f"-Wl,-rpath,{LIB_DIR}"])
        outs.append(out)
    return outs


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--kq", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    args = ap.parse_args()
    both = not (args.kq or args.oracle)
    if both or args.oracle:
        print(build_oracle())
    if both or args.kq:
        print(build_kq())


if __name__ == "__main__":
    main()
