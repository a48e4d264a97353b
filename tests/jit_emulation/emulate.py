"""Host execution of the run-time generated tile kernels (test infrastructure).

libkq's `kq_debug_pass` returns, pass by pass, exactly what a statevector would run: the generated
kernel source with its coefficient buffer, or a single sweep gate.  This module compiles those kernel
sources with g++ against a small host runtime (one OS thread per GPU thread, a std::barrier for the
workgroup barrier, the tile array shared by the workgroup's threads) and runs them on a host state, so
the generated code itself is checked against the oracle without a GPU."""
from __future__ import annotations

import ctypes
import hashlib
import re
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

import kettle_amd as ka
from kettle_amd import _lib

PRELUDE = r"""
#include <cstdint>
#include <cmath>
#include <barrier>
struct double2 { double x, y; };
struct float2 { float x, y; };
struct EmuDim { unsigned long long x; };
extern thread_local EmuDim threadIdx;
extern EmuDim blockIdx;
extern EmuDim gridDim;
extern std::barrier<>* emu_barrier;
#define __device__
#define __forceinline__ inline
#define __global__
#define __launch_bounds__(...)
#define __shared__ static
inline void __syncthreads() { emu_barrier->arrive_and_wait(); }
"""

RUNTIME = r"""
#include <barrier>
#include <cstdint>
#include <thread>
#include <vector>
struct EmuDim { unsigned long long x; };
thread_local EmuDim threadIdx;
EmuDim blockIdx;
EmuDim gridDim;
std::barrier<>* emu_barrier = nullptr;
template <typename F>
static void emu_run(F call, uint32_t blocks, uint32_t threads)
{
    gridDim.x = blocks;
    for (uint32_t b = 0; b < blocks; ++b) {
        blockIdx.x = b;
        std::barrier<> bar(threads);
        emu_barrier = &bar;
        std::vector<std::thread> pool;
        for (uint32_t t = 0; t < threads; ++t) {
            pool.emplace_back([=] { threadIdx.x = t; call(); });
        }
        for (auto& th : pool) {
            th.join();
        }
    }
}
extern "C" void emu_launch(void (*fn)(void*, const double*, uint32_t), void* a, const double* cf, uint32_t blocks,
                           uint32_t threads)
{
    emu_run([=] { fn(a, cf, blocks); }, blocks, threads);
}
// out-of-place passes: a fourth argument, the destination buffer
extern "C" void emu_launch_to(void (*fn)(void*, const double*, uint32_t, void*), void* a, const double* cf, uint32_t blocks,
                              uint32_t threads, void* dst)
{
    emu_run([=] { fn(a, cf, blocks, dst); }, blocks, threads);
}
// passes of a run from a basis state: one (the basis index) or two (known-bit mask and values) more arguments
extern "C" void emu_launch_x1(void (*fn)(void*, const double*, uint32_t, uint64_t), void* a, const double* cf,
                              uint32_t blocks, uint32_t threads, uint64_t x)
{
    emu_run([=] { fn(a, cf, blocks, x); }, blocks, threads);
}
extern "C" void emu_launch_x2(void (*fn)(void*, const double*, uint32_t, uint64_t, uint64_t), void* a, const double* cf,
                              uint32_t blocks, uint32_t threads, uint64_t x, uint64_t y)
{
    emu_run([=] { fn(a, cf, blocks, x, y); }, blocks, threads);
}
extern "C" void emu_launch_to_x1(void (*fn)(void*, const double*, uint32_t, void*, uint64_t), void* a, const double* cf,
                                 uint32_t blocks, uint32_t threads, void* dst, uint64_t x)
{
    emu_run([=] { fn(a, cf, blocks, dst, x); }, blocks, threads);
}
extern "C" void emu_launch_to_x2(void (*fn)(void*, const double*, uint32_t, void*, uint64_t, uint64_t), void* a,
                                 const double* cf, uint32_t blocks, uint32_t threads, void* dst, uint64_t x, uint64_t y)
{
    emu_run([=] { fn(a, cf, blocks, dst, x, y); }, blocks, threads);
}
"""


def host_source(src: str, name: str) -> str:
    """The generated kernel as host C++: fixed-width typedefs from <cstdint>, the workgroup barrier
    as the runtime's std::barrier, no GPU attributes."""
    body = "\n".join(line for line in src.split("\n") if not line.startswith("typedef "))
    body = re.sub(r"__device__ __forceinline__ void lds_barrier\(\)\s*\{[^}]*\}",
                  "inline void lds_barrier() { emu_barrier->arrive_and_wait(); }", body)
    body = re.sub(r"__builtin_amdgcn_s_setprio\([^)]*\)", "(void) 0", body)
    body = body.replace("__attribute__((always_inline))", "")
    body = re.sub(r"__attribute__\(\(amdgpu_waves_per_eu\([^)]*\)\)\)", "", body)
    body = body.replace("void kq_pass(", f"void {name}(")
    body = body.replace("#pragma once", "")
    return PRELUDE + body


def _passes(ops: np.ndarray, n: int, precision: int):
    c = ctypes
    lib = _lib.load()
    out = []
    i = 0
    while True:
        ln, nc, npas = c.c_size_t(), c.c_size_t(), c.c_size_t()
        info = np.zeros(5, dtype=np.uint64)
        args = (n, precision, _lib.ptr(ops), len(ops), 160 * 1024, i)
        _lib.check(lib.kq_debug_pass(*args, None, 0, c.byref(ln), None, 0, c.byref(nc), _lib.ptr(info), c.byref(npas)))
        buf = c.create_string_buffer(ln.value + 1)
        coefs = np.zeros(max(nc.value, 1), dtype=np.float64)
        _lib.check(lib.kq_debug_pass(*args, buf, ln.value + 1, c.byref(ln), _lib.ptr(coefs), coefs.size, c.byref(nc),
                                     _lib.ptr(info), c.byref(npas)))
        out.append({"sweep": int(info[0]) == 1, "to_other": int(info[0]) == 2, "src": buf.value.decode(), "coefs": coefs[:nc.value].copy(),
                    "blocks": int(info[1]), "threads": int(info[2]), "target": int(info[3]), "control": int(info[4]) - 1})
        i += 1
        if i >= npas.value:
            return out


_BUILD = Path(tempfile.gettempdir()) / "kq_jit_emulation"


def _compile(sources: list[tuple[str, str]]) -> ctypes.CDLL:
    """sources: (function name, host source); one shared library (cached by content)."""
    _BUILD.mkdir(exist_ok=True)
    key = hashlib.sha1((RUNTIME + "".join(s for _, s in sources)).encode()).hexdigest()[:16]
    lib_path = _BUILD / f"emu_{key}.so"
    if not lib_path.exists():
        files = []
        rt = _BUILD / "runtime.cpp"
        rt.write_text(RUNTIME)
        files.append(rt)
        for name, src in sources:
            f = _BUILD / f"{key}_{name}.cpp"
            f.write_text(src)
            files.append(f)

        def cc(f):
            o = f.with_suffix(".o")
            r = subprocess.run(["g++", "-std=c++20", "-O1", "-fPIC", "-w", "-c", str(f), "-o", str(o)],
                               capture_output=True, text=True)
            if r.returncode:
                raise RuntimeError(f"host build of {f.name} failed:\n{r.stderr[:4000]}")
            return o

        with ThreadPoolExecutor(8) as ex:
            objs = list(ex.map(cc, files))
        r = subprocess.run(["g++", "-shared", "-o", str(lib_path), *map(str, objs), "-lpthread"], capture_output=True,
                           text=True)
        if r.returncode:
            raise RuntimeError(r.stderr[:4000])
    return ctypes.CDLL(str(lib_path))


def _sweep(psi: np.ndarray, m: np.ndarray, target: int, control: int) -> None:
    u = m[0::2] + 1j * m[1::2]
    idx = np.arange(psi.size)
    sel = ((idx >> target) & 1) == 0
    if control >= 0:
        sel &= ((idx >> control) & 1) == 1
    i0 = idx[sel]
    i1 = i0 | (1 << target)
    a, b = psi[i0].copy(), psi[i1].copy()
    psi[i0] = u[0] * a + u[1] * b
    psi[i1] = u[2] * a + u[3] * b


def run(ops: np.ndarray, n: int, initial: np.ndarray, precision: int = ka.KQ_C128) -> np.ndarray:
    """The state after the gate list, as the generated kernels compute it (host emulation)."""
    passes = _passes(np.ascontiguousarray(ops), n, precision)
    tiles = [(f"kq_pass_{i}", host_source(p["src"], f"kq_pass_{i}")) for i, p in enumerate(passes) if not p["sweep"]]
    lib = _compile(tiles)
    fn_t = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32)
    launch = lib.emu_launch
    launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    launch_to = lib.emu_launch_to
    launch_to.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                          ctypes.c_void_p]
    c64 = precision == ka.KQ_C64
    psi = np.ascontiguousarray(initial, dtype=np.complex64 if c64 else np.complex128).copy()
    for i, p in enumerate(passes):
        if p["sweep"]:
            m = p["coefs"]
            if c64:
                # the sweep record keeps FP64 coefficients; apply in complex64 like the device
                work = psi.astype(np.complex128)
                _sweep(work, m, p["target"], p["control"])
                psi = work.astype(np.complex64)
            else:
                _sweep(psi, m, p["target"], p["control"])
            continue
        fn = getattr(lib, f"kq_pass_{i}")
        cf = np.ascontiguousarray(p["coefs"] if p["coefs"].size else np.zeros(1))
        if p["to_other"]:
            # stores into a second buffer, which then holds the state (every entry is written once)
            dst = np.full_like(psi, np.nan)
            launch_to(ctypes.cast(fn, ctypes.c_void_p), psi.ctypes.data, cf.ctypes.data, p["blocks"], p["threads"],
                      dst.ctypes.data)
            psi = dst
        else:
            launch(ctypes.cast(fn, ctypes.c_void_p), psi.ctypes.data, cf.ctypes.data, p["blocks"], p["threads"])
    return psi.astype(np.complex128)


def _passes_from_basis(ops: np.ndarray, n: int, precision: int, basis: int):
    c = ctypes
    lib = _lib.load()
    out = []
    i = 0
    while True:
        ln, nc, npas = c.c_size_t(), c.c_size_t(), c.c_size_t()
        info = np.zeros(8, dtype=np.uint64)
        args = (n, precision, _lib.ptr(ops), len(ops), 160 * 1024, i, basis)
        _lib.check(lib.kq_debug_pass_from_basis(*args, None, 0, c.byref(ln), None, 0, c.byref(nc), _lib.ptr(info),
                                                c.byref(npas)))
        buf = c.create_string_buffer(ln.value + 1)
        coefs = np.zeros(max(nc.value, 1), dtype=np.float64)
        _lib.check(lib.kq_debug_pass_from_basis(*args, buf, ln.value + 1, c.byref(ln), _lib.ptr(coefs), coefs.size,
                                                c.byref(nc), _lib.ptr(info), c.byref(npas)))
        out.append({"sweep": int(info[0]) == 1, "to_other": int(info[0]) == 2, "src": buf.value.decode(),
                    "coefs": coefs[:nc.value].copy(), "blocks": int(info[1]), "threads": int(info[2]),
                    "target": int(info[3]), "control": int(info[4]) - 1, "input": int(info[5]),
                    "zmask": int(info[6]), "zval": int(info[7])})
        i += 1
        if i >= npas.value:
            return out


def run_from_basis(ops: np.ndarray, n: int, basis: int, precision: int = ka.KQ_C128):
    """The state after the gate list from the basis state |basis>, as the generated kernels of a run from a
    pending reset compute it: the buffer starts as NaN (a pass that read it would spread NaN), the first
    pass builds the state, later passes skip known-zero tiles.  Returns (state, input mode per tile pass)."""
    passes = _passes_from_basis(np.ascontiguousarray(ops), n, precision, basis)
    tiles = [(f"kq_pass_{i}", host_source(p["src"], f"kq_pass_{i}")) for i, p in enumerate(passes) if not p["sweep"]]
    lib = _compile(tiles)
    vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
    shims = {
        (False, 0): (lib.emu_launch, []), (True, 0): (lib.emu_launch_to, [vp]),
        (False, 1): (lib.emu_launch_x1, [u64]), (True, 1): (lib.emu_launch_to_x1, [vp, u64]),
        (False, 2): (lib.emu_launch_x2, [u64, u64]), (True, 2): (lib.emu_launch_to_x2, [vp, u64, u64]),
    }
    for f, extra in shims.values():
        f.argtypes = [vp, vp, vp, u32, u32] + extra
    c64 = precision == ka.KQ_C64
    psi = np.full(1 << n, np.nan + 1j * np.nan, dtype=np.complex64 if c64 else np.complex128)
    modes = []
    for i, p in enumerate(passes):
        if p["sweep"]:
            raise RuntimeError("run_from_basis: sweep passes are not emulated here")
        fn = ctypes.cast(getattr(lib, f"kq_pass_{i}"), vp)
        cf = np.ascontiguousarray(p["coefs"] if p["coefs"].size else np.zeros(1))
        extra = [] if p["input"] == 0 else ([p["zval"]] if p["input"] == 1 else [p["zmask"], p["zval"]])
        launch, _ = shims[(p["to_other"], len(extra))]
        modes.append(p["input"])
        if p["to_other"]:
            dst = np.full_like(psi, np.nan)
            launch(fn, psi.ctypes.data, cf.ctypes.data, p["blocks"], p["threads"], dst.ctypes.data, *extra)
            psi = dst
        else:
            launch(fn, psi.ctypes.data, cf.ctypes.data, p["blocks"], p["threads"], *extra)
    return psi.astype(np.complex128), modes
