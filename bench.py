"""Benchmark of kettle's simulate() hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-vqe] [--no-cpu]

Workload (one step): simulate() of the 30-qubit complex128 circuit "QFT on all qubits + 20 layers
of random 1-qubit gates each followed by a CX brick ladder" on |0...0> (state resident in HBM;
the reset to |0...0> is part of the step).  value = gates/s (kettle primitive gates after the QFT
and swap expansion, divided by the step time).

The JSON line also carries
  roofline      the dominant kernel's (the fused tile pass) achieved HBM GB/s against 8 TB/s,
                from HIP events on the stream libkq launches on, over the timed steps;
  passes        achieved GB/s per gate class (low-index LDS tile, high-index tile, sweep);
  cpu_baseline  the oracle (kettle's loops restated, 1 core, in a child process pinned to one CPU) per
                kettle gate type on a bounded sample, extrapolated to the 30-qubit circuit, as gates/s;
  plain_sv_20q  the 20-qubit plain statevector workload (simulate, 4096 shots, <H> over 64 strings);
  density_matrix the 14-qubit density-matrix workload (simulate with mid-circuit M, trace, partial trace,
                8192 shots);
  vqe           energy+gradient evaluations/s at 24 qubits, 960 parameters, 248 Pauli strings:
                adjoint complex64 vs the plain per-parameter FP64 GPU path (parameter shift,
                2P + 1 energies; timed on a bounded number of energies and scaled);
  batch         the same for 4096 uneven instances (8..22 qubits) in one batch per run.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# HBM bytes per tile-pass launch over algorithmic bytes, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
# this 30-qubit workload (kq_pass, the 14 launches of one step from |0...0>), FETCH_SIZE scaled by 2.0 and
# WRITE_SIZE by 1.0 (the factors a plain copy kernel of known size reads): measured HBM bytes of the step's
# launches / their algorithmic bytes (the pass that builds the state writes it, the passes over known-zero
# tiles move the other tiles only, the rest read and write the state), the excess being coefficient tables
TRAFFIC_PER_ALGORITHMIC_BYTE = 1.0004


def log(msg: str) -> None:
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def c64_section(n: int, layers: int, steps: int) -> dict:
    """The same circuit family at n qubits in complex64 (the large-statevector workload's second size:
    33 qubits = 64 GiB), timed like the headline (reset to |0...0> + gates per step, one warm-up
    step), and on the resident dense state."""
    import kettle_amd as ka
    from kettle_amd import _lib, workloads

    ops, _, _ = workloads.qft_plus_layers(n=n, layers=layers).flatten()
    sv = ka.Statevector(n, ka.KQ_C64)
    lib = _lib.load()

    def step():
        sv.set_basis(0)
        ka.apply_gates(sv, ops)

    step()
    ka.apply_gates(sv, ops)   # compiles the dense first pass
    _lib.check(lib.kq_synchronize())
    _, n_passes, _ = ka.plan_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    _lib.check(lib.kq_synchronize())
    dt = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(2):
        ka.apply_gates(sv, ops)
    _lib.check(lib.kq_synchronize())
    dt_dense = (time.perf_counter() - t0) / 2
    tuner = ka.fusion_autotune_info()
    del sv
    return {"workload": f"{n}q complex64 QFT + {layers} random 1q layers with CX brick ladders",
            "gates": len(ops), "value": round(len(ops) / dt, 1), "unit": "gates/s", "ms_per_step": round(dt * 1e3, 2),
            "dense_resident_state_ms_per_step": round(dt_dense * 1e3, 2),
            "steps": steps, "passes_per_step": n_passes, "low_qubits_per_tile": tuner["low_qubits"],
            "max_stages_per_pass": tuner["max_stages"],
            "GB_moved_per_pass": round(2 * 8 * (1 << n) / 1e9, 1)}


def cpu_baseline(n: int, sample_qubits: int, reps: int, per_class: int) -> dict:
    """The oracle (kettle's loops restated, one thread) timed per kettle gate type on a bounded sample,
    in a child process pinned to one core (oracle/cpu_baseline.py), extrapolated to the n-qubit circuit."""
    import shutil
    import subprocess

    cfg = json.dumps({"qubits": n, "sample_qubits": sample_qubits, "reps": reps, "per_class": per_class})
    cpu = min(os.sched_getaffinity(0))
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", cfg]
    if shutil.which("taskset"):
        cmd = ["taskset", "-c", str(cpu)] + cmd
    log(f"cpu baseline: per gate type at {sample_qubits} qubits, {reps} runs each, one core ({' '.join(cmd[:3])})")
    proc = subprocess.run(cmd, capture_output=True, text=True, cwd=str(ROOT), timeout=600)
    if proc.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + proc.stderr[-2000:])
    return json.loads(proc.stdout.strip().splitlines()[-1])


def plain_sv_section(n: int, reps: int) -> dict:
    """The plain statevector workload (20 qubits complex128, depth 40): simulate from |0...0>, 4096-shot
    counts and <H> over 64 Pauli strings, each timed after a warm-up.  The 16 MiB state stays in the
    256 MB Infinity Cache, so the gate rate is quoted in GB/s of algorithmic traffic, not as HBM %."""
    import kettle_amd as ka
    from kettle_amd import _lib, workloads

    lib = _lib.load()
    w = workloads.plain_statevector(n=n)
    ops, _, _ = w.circuit.flatten()
    sv = ka.Statevector(n, ka.KQ_C128)

    def timed(fn):
        fn()
        _lib.check(lib.kq_synchronize())
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        _lib.check(lib.kq_synchronize())
        return (time.perf_counter() - t0) / reps

    def sim():
        sv.set_basis(0)
        ka.apply_gates(sv, ops)

    stats = np.zeros(len(_lib.PASS_CLASSES), dtype=_lib.PASS_STATS_DTYPE)
    sim()
    _lib.check(lib.kq_profile_begin())
    t_sim = timed(sim)
    _lib.check(lib.kq_profile_end(_lib.ptr(stats)))
    _, n_passes, _ = ka.plan_stats()
    kbytes = float(stats["bytes"].sum())
    kms = float(stats["ms"].sum())
    t_shots = timed(lambda: ka.perform_measurements_as_memory(sv, w.shots, seed=0))
    t_exp = timed(lambda: ka.expectation_value(w.hamiltonian, sv))
    return {"workload": f"{n}q complex128 random circuit depth 40 ({len(ops)} gates), {w.shots} shots, "
                        f"<H> over {w.hamiltonian.size()} Pauli strings",
            "simulate_ms": round(t_sim * 1e3, 3), "gates_per_s": round(len(ops) / t_sim, 1),
            "passes": n_passes, "pass_GB_per_s": round(kbytes / (kms * 1e-3) / 1e9, 1) if kms else None,
            "shots_ms": round(t_shots * 1e3, 3), "shots_per_s": round(w.shots / t_shots, 1),
            "expectation_ms": round(t_exp * 1e3, 3),
            "note": "16 MiB state: Infinity-Cache resident, GB/s of algorithmic bytes, not HBM %"}


def density_matrix_section(n: int, reps: int) -> dict:
    """The density-matrix workload at n qubits (complex128, depth 30, M on qubit 0 with an if-else
    every 10 layers): simulate, trace, 2-qubit partial trace and 8192 shots from the diagonal."""
    import kettle_amd as ka
    from kettle_amd import _lib, workloads

    lib = _lib.load()
    c = workloads.density_matrix_circuit(n=n, depth=30, seed=n)
    ops, _, _ = c.flatten()
    n_gates = int(np.sum(ops["gate"] < int(ka.Gate.M)))
    dm = ka.DensityMatrix(n)
    stats = np.zeros(len(_lib.PASS_CLASSES), dtype=_lib.PASS_STATS_DTYPE)

    def sim():
        dm.set_basis(0)
        ka.simulate(c, dm, 5)

    sim()
    _lib.check(lib.kq_synchronize())
    _lib.check(lib.kq_profile_begin())
    t0 = time.perf_counter()
    for _ in range(reps):
        sim()
    _lib.check(lib.kq_synchronize())
    t_sim = (time.perf_counter() - t0) / reps
    _lib.check(lib.kq_profile_end(_lib.ptr(stats)))
    kbytes, kms, launches = float(stats["bytes"].sum()), float(stats["ms"].sum()), int(stats["launches"].sum())
    t0 = time.perf_counter()
    tr = dm.trace()
    pt = ka.partial_trace(dm, list(range(2, n)))
    t_read = time.perf_counter() - t0
    t0 = time.perf_counter()
    ka.perform_measurements_as_counts_marginal(ka.calculate_probabilities_raw(dm), 8192, seed=0)
    t_shots = time.perf_counter() - t0
    return {"workload": f"{n}q complex128 density matrix, depth 30 ({n_gates} gates) with 3 mid-circuit M + if-else",
            "vec_rho_bytes": 16 << (2 * n), "simulate_ms": round(t_sim * 1e3, 3),
            "gates_per_s": round(n_gates / t_sim, 1), "pass_launches_per_simulate": launches // reps,
            "pass_GB_per_s": round(kbytes / (kms * 1e-3) / 1e9, 1) if kms else None,
            "pass_frac_of_hbm_peak": round(kbytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if kms else None,
            "trace": [tr.real, tr.imag], "trace_and_partial_trace_ms": round(t_read * 1e3, 3),
            "shots_8192_ms": round(t_shots * 1e3, 3), "partial_trace_shape": list(pt.shape)}


def vqe_section(n: int, layers: int, n_random: int, energies_sample: int) -> dict:
    import kettle_amd as ka
    from kettle_amd import workloads

    w = workloads.vqe(n=n, layers=layers, n_random=n_random)
    P = len(w.params)
    # adjoint, complex64 state with FP64 reductions
    st = ka.Statevector(n, ka.KQ_C64)
    r = ka.energy_and_gradient(w.circuit, w.hamiltonian, w.params, state=st)   # warm-up
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        st.set_basis(0)
        r = ka.energy_and_gradient(w.circuit, w.hamiltonian, w.params, state=st)
    t_adj = (time.perf_counter() - t0) / reps
    # one more evaluation with every pass and the H psi launch bracketed by HIP events
    from kettle_amd import _lib
    lib = _lib.load()
    stats = np.zeros(len(_lib.PASS_CLASSES), dtype=_lib.PASS_STATS_DTYPE)
    st.set_basis(0)
    _lib.check(lib.kq_profile_begin())
    ka.energy_and_gradient(w.circuit, w.hamiltonian, w.params, state=st)
    _lib.check(lib.kq_profile_end(_lib.ptr(stats)))
    parts = {}
    for i, cls in enumerate(_lib.PASS_CLASSES):
        s = stats[i]
        if s["launches"]:
            parts[cls] = {"launches": int(s["launches"]), "ms": round(float(s["ms"]), 3),
                          "GB_per_s": round(float(s["bytes"]) / (float(s["ms"]) * 1e-3) / 1e9, 1)}
    fw = stats[0].copy()
    for f in ("ms", "bytes"):
        fw[f] = stats[0][f] + stats[1][f]
    rv = stats[3]
    dom_rev = float(rv["ms"]) >= float(fw["ms"])
    dom = rv if dom_rev else fw
    dom_gbs = float(dom["bytes"]) / (float(dom["ms"]) * 1e-3) / 1e9 if dom["ms"] else 0.0
    roofline = {"bound": "hbm", "kernel": "kq_adjoint_pass (reverse sweep)" if dom_rev else "kq_pass (forward sweep)",
                "achieved": round(dom_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
                "traffic": None, "bytes_per_launch": (4 if dom_rev else 2) * 8 * (1 << n),
                "classes": parts, "profiled_eval_ms": round(float(stats["ms"].sum()), 3)}
    # the FP64 reference energy on the device (complex128 forward + expectation)
    sv = ka.Statevector(n, ka.KQ_C128)
    ka.simulate(w.circuit, sv)
    e64 = ka.expectation_value(w.hamiltonian, sv).real
    # plain per-parameter FP64 GPU path: one energy = simulate + expectation (complex128)
    ops, _, _ = w.circuit.flatten(w.params)
    t0 = time.perf_counter()
    for i in range(energies_sample):
        shifted = ops.copy()
        shifted["angle"][i % len(ops)] += np.pi / 2
        sv.set_basis(0)
        ka.apply_gates(sv, shifted)
        ka.expectation_value(w.hamiltonian, sv)
    t_energy = (time.perf_counter() - t0) / max(energies_sample, 1)
    t_plain = (2 * P + 1) * t_energy
    if energies_sample == 0:
        return {"config": f"{n}q n_local(RY,RZ; CX linear) x{layers}, {P} params, {w.hamiltonian.size()} Pauli strings",
                "adjoint_c64_evals_per_s": round(1.0 / t_adj, 4), "energy_err_vs_fp64": abs(r.energy - e64),
                "roofline": roofline}
    return {
        "config": f"{n}q n_local(RY,RZ; CX linear) x{layers}, {P} params, {w.hamiltonian.size()} Pauli strings",
        "adjoint_c64_evals_per_s": round(1.0 / t_adj, 4),
        "plain_fp64_param_shift_evals_per_s": round(1.0 / t_plain, 6),
        "speedup": round(t_plain / t_adj, 2),
        "energy_err_vs_fp64": abs(r.energy - e64),
        "plain_timing": f"{energies_sample} complex128 energies timed ({t_energy * 1e3:.2f} ms each), x (2P+1)",
        "checkpoint_restores": r.restores,
        "roofline": roofline,
    }


def batch_section(n_instances: int, reps: int, plain_sample: int) -> dict:
    """Batched uneven VQE: every instance's energy + full gradient per run (complex64 states, FP64
    reductions), replayed as a captured hipGraph; against the plain per-instance FP64 GPU path
    (parameter shift: 2P + 1 complex128 simulate + expectation per instance), timed on a sample."""
    import kettle_amd as ka
    from kettle_amd import _lib, workloads
    from kettle_amd.batch import EnergyGradientBatch

    insts = workloads.batched_vqe(n_instances)

    def timed(forward_precision):
        t0 = time.perf_counter()
        b = EnergyGradientBatch(ka.KQ_C64, forward_precision)
        for inst in insts:
            b.add_ops(inst.n, inst.ops(), len(inst.theta), inst.hamiltonian)
        r = b.run(use_graph=False)            # plans, lowers, uploads, runs eagerly
        b.run(use_graph=True)                 # captures the graph
        log(f"batch: {n_instances} instances, forward precision {forward_precision}: setup + first runs "
            f"{time.perf_counter() - t0:.1f} s")
        thetas = [inst.theta for inst in insts]
        t0 = time.perf_counter()
        for i in range(reps):
            b.set_all_parameters(thetas)   # an optimiser step: every instance's programs are lowered again
            r = b.run(use_graph=True)
        t = (time.perf_counter() - t0) / reps
        info = b.info()
        info.update(b.plan_info())
        b.close()
        return r, t, info

    r64, t64, _ = timed(ka.KQ_C64)
    r, t_run, info = timed(ka.KQ_C128)
    achieved = info["algorithmic_bytes"] / t_run / 1e9
    # plain per-instance FP64 path on a sample, scaled to the whole batch
    idx = np.unique(np.linspace(0, n_instances - 1, plain_sample).astype(int))
    t_plain_sample = 0.0
    worst = 0.0
    worst64 = 0.0
    jit0 = ka.jit_info()
    t_warm = 0.0
    for i in idx:
        inst = insts[i]
        ops = inst.ops()
        sv = ka.Statevector(inst.n, ka.KQ_C128)
        # one untimed energy first: compiles this structure's pass kernels and plans it
        t1 = time.perf_counter()
        sv.set_basis(0)
        ka.apply_gates(sv, ops)
        ka.expectation_value(inst.hamiltonian, sv)
        _lib.check(_lib.load().kq_synchronize())
        t_warm += time.perf_counter() - t1
        t1 = time.perf_counter()
        for _ in range(2):
            sv.set_basis(0)
            ka.apply_gates(sv, ops)
            e = ka.expectation_value(inst.hamiltonian, sv).real
        t_e = (time.perf_counter() - t1) / 2
        t_plain_sample += (2 * len(inst.theta) + 1) * t_e
        worst = max(worst, abs(r.energies[i] - e))
        worst64 = max(worst64, abs(r64.energies[i] - e))
    t_plain = t_plain_sample * n_instances / len(idx)
    jit1 = ka.jit_info()
    return {
        "config": f"{n_instances} instances, qubits U[8,22], n_local(RY,RZ; CX linear) depth U[4,16], "
                  f"U[32,256] Pauli strings; complex128 forward sweep (FP64 energies), complex64 reverse "
                  f"sweep with FP64 reductions, hipGraph replay",
        "batched_adjoint_c64_evals_per_s": round(n_instances / t_run, 2),
        "ms_per_batch": round(t_run * 1e3, 2),
        "plain_fp64_param_shift_evals_per_s": round(n_instances / t_plain, 4),
        "speedup": round(t_plain / t_run, 1),
        "energy_err_vs_fp64_sample": worst,
        "all_c64_evals_per_s": round(n_instances / t64, 2),
        "all_c64_energy_err_vs_fp64_sample": worst64,
        "all_c64_note": "complex64 forward sweep too: informational, its energy error is that of FP32 amplitudes "
                        "and is not claimed against the 1e-6 bar (the reported mode runs the forward sweep in complex128)",
        "roofline": {"bound": "hbm", "kernel": "whole batch run (hipGraph: tile-batch passes, Pauli, sums)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_run": round(info["algorithmic_bytes"]),
                     "forward_passes": info["forward_passes"], "reverse_passes": info["reverse_passes"],
                     "note": "algorithmic bytes of every launch over the whole replay's wall time (host "
                             "read-back included): a lower bound on the passes' own rate"},
        "max_inverse_drift": float(np.max(r.drift)),
        "launches_per_batch": info["launches"],
        "device_bytes": info["device_bytes"],
        "plain_timing": f"{len(idx)} sampled instances, 2 complex128 energies each after one untimed warm-up energy "
                        f"(kernel compile and planning), x (2P+1), scaled",
        "plain_warmup_s": round(t_warm, 2),
        "plain_warmup_compile_s": round(jit1["compile_seconds"] - jit0["compile_seconds"], 2),
    }


def dense_section(n: int, reps: int) -> dict:
    """A/B of a dense 4-qubit block (16 x 16 complex U on qubits 0..3 of an n-qubit state, in place):
    the VALU form against the MFMA form (kq_debug_dense_block), complex128 and complex64, each launch
    timed with HIP events after one warm-up launch.  128 flops per amplitude; bytes = one read + one
    write of the state."""
    import kettle_amd as ka
    from kettle_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(0)
    q, r = np.linalg.qr(rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16)))
    u = np.ascontiguousarray(q * (np.diag(r) / np.abs(np.diag(r)))).view(np.float64).reshape(-1)
    out = {"workload": f"{n}q state, dense 16x16 complex block on qubits 0..3, in place", "flops_per_amp": 128}
    for prec, name, ab in ((ka.KQ_C128, "c128", 16), (ka.KQ_C64, "c64", 8)):
        st = ka.Statevector(n, prec)
        res = {}
        for form, mfma in (("valu", 0), ("mfma", 1)):
            ms = ctypes.c_double()
            _lib.check(lib.kq_debug_dense_block(st.handle, _lib.ptr(u), mfma, ctypes.byref(ms)))
            t = []
            for _ in range(reps):
                _lib.check(lib.kq_debug_dense_block(st.handle, _lib.ptr(u), mfma, ctypes.byref(ms)))
                t.append(ms.value)
            avg = float(np.mean(t))
            res[form] = {"avg_ms": round(avg, 4), "min_ms": round(min(t), 4),
                         "GB_per_s": round(2 * ab * (1 << n) / (avg * 1e-3) / 1e9, 1),
                         "TFLOP_per_s": round(128 * (1 << n) / (avg * 1e-3) / 1e12, 2)}
        res["mfma_speedup"] = round(res["valu"]["avg_ms"] / res["mfma"]["avg_ms"], 3)
        out[name] = res
        del st
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=20)
    ap.add_argument("--no-vqe", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample-qubits", type=int, default=25,
                    help="CPU baseline: qubits of the per-gate-type sample (scaled to --qubits)")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--cpu-per-class", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true", help="skip the 20-qubit plain statevector section")
    ap.add_argument("--no-dm", action="store_true", help="skip the density-matrix section")
    ap.add_argument("--dm-qubits", type=int, default=14)
    ap.add_argument("--vqe-qubits", type=int, default=24)
    ap.add_argument("--vqe-energies", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--c64-qubits", type=int, default=33,
                    help="also time the circuit family in complex64 at this many qubits (0: skip)")
    ap.add_argument("--c64-steps", type=int, default=3)
    ap.add_argument("--no-stream-probe", action="store_true",
                    help="skip the plain-stream HBM reference (read / write / in place / copy GB/s)")
    ap.add_argument("--batch-instances", type=int, default=4096)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--jit", type=int, default=1,
                    help="run-time compiled tile passes: 0 off, 1 states of >= 18 qubits, 2 all")
    ap.add_argument("--max-stages", type=int, default=0,
                    help="register stages (LDS remaps + 1) per tile pass; 0 keeps the library default")
    ap.add_argument("--low-qubits", type=int, default=0,
                    help="qubits 0..c-1 in every tile pass (contiguous chunk of 2^c amplitudes per access); "
                         "0 keeps the library default")
    ap.add_argument("--jit-tuning", default="3,0,1",
                    help="min waves per SIMD, persistent blocks per CU, pivot normalisation (0/1) of the "
                         "run-time compiled passes")
    ap.add_argument("--jit-priority", type=int, default=1,
                    help="wave priority around a tile's HBM traffic in the run-time compiled passes: 0 off, "
                         "1 loads, 2 loads and stores")
    ap.add_argument("--only", choices=["c64", "plain", "dm", "vqe", "batch", "dense"], default=None,
                    help="run one side section alone and print its JSON (profiling runs)")
    ap.add_argument("--jit-nontemporal", type=int, default=3,
                    help="nontemporal hint on the run-time compiled passes' HBM traffic: 0 off, 1 loads, 2 loads and "
                         "stores, 3 loads and stores of vectors above 256 MiB")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="only run N plain copies of a --qubits complex128 state (HBM counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the single-GPU shard is the measured unit); rerun with --gpus 1", file=sys.stderr)
        return 2

    import kettle_amd as ka
    from kettle_amd import _lib, workloads

    lib = _lib.load()
    name = ctypes.create_string_buffer(256)
    ncu = ctypes.c_int()
    mem = ctypes.c_size_t()
    _lib.check(lib.kq_device_info(name, 256, ctypes.byref(ncu), ctypes.byref(mem)))
    log(f"device {name.value.decode()} CUs={ncu.value} mem={mem.value / 2**30:.0f} GiB")

    n = args.qubits
    if args.only:
        ka.set_jit_priority(args.jit_priority)
        ka.set_jit_nontemporal(args.jit_nontemporal)
        only = {"c64": lambda: c64_section(args.c64_qubits, args.layers, args.c64_steps),
                "plain": lambda: plain_sv_section(20, 5),
                "dm": lambda: density_matrix_section(args.dm_qubits, 2),
                "vqe": lambda: vqe_section(args.vqe_qubits, 20, 200, args.vqe_energies),
                "batch": lambda: batch_section(args.batch_instances, args.batch_reps, 32),
                "dense": lambda: dense_section(args.qubits, 5)}[args.only]
        log(f"{args.only} section only")
        print(json.dumps({args.only: only()}))
        return 0
    if args.calib_copy:
        # a plain copy kernel of known size: reads and writes 16 * 2^n bytes per launch
        src = ka.Statevector(n, ka.KQ_C128)
        dst = ka.Statevector(n, ka.KQ_C128)
        for _ in range(args.calib_copy):
            _lib.check(lib.kq_state_copy(dst.handle, src.handle))
        _lib.check(lib.kq_synchronize())
        print(json.dumps({"calib_copy": args.calib_copy, "bytes_read_per_launch": 16 << n,
                          "bytes_written_per_launch": 16 << n}))
        return 0
    circuit = workloads.qft_plus_layers(n=n, layers=args.layers)
    ops, _, _ = circuit.flatten()
    n_gates = len(ops)
    log(f"workload: {n}q complex128 QFT + {args.layers} layers, {n_gates} gates")

    sv = ka.Statevector(n, ka.KQ_C128)

    # A step is kettle's workflow: the state reset to |0...0>, then the circuit (Statevector(n) +
    # simulate()).  The reset is left pending and the first pass builds its tiles from the basis index:
    # it reads nothing, and only the workgroup whose tile holds the basis amplitude runs its gates (every
    # other tile is zero and stays zero).  The same circuit on the resident dense state (the previous
    # run's output: every pass reads and writes all 2^n amplitudes) is timed apart below.
    def step():
        sv.set_basis(0)
        ka.apply_gates(sv, ops)

    def step_dense():
        ka.apply_gates(sv, ops)

    ka.set_jit(args.jit)
    if args.low_qubits > 0:
        ka.set_fusion_low_qubits(args.low_qubits)
    if args.max_stages > 0:
        ka.set_fusion(-1, 0, args.max_stages, 0.0)
    waves, per_cu, normalise = (int(x) for x in args.jit_tuning.split(","))
    ka.set_jit_tuning(waves, per_cu, bool(normalise))
    ka.set_jit_priority(args.jit_priority)
    ka.set_jit_nontemporal(args.jit_nontemporal)
    step_dense()   # from the new state's pending |0...0>; then the dense first pass below compiles
    step_dense()
    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    jit_warm = ka.jit_info()
    tuner = ka.fusion_autotune_info()
    _, n_passes, n_tiles = ka.plan_stats()
    _lib.check(lib.kq_synchronize())
    stats = np.zeros(len(_lib.PASS_CLASSES), dtype=_lib.PASS_STATS_DTYPE)
    _lib.check(lib.kq_profile_begin())
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i + 1}/{args.steps} done")
    _lib.check(lib.kq_synchronize())
    elapsed = time.perf_counter() - t0
    _lib.check(lib.kq_profile_end(_lib.ptr(stats)))

    ms_per_step = elapsed * 1e3 / args.steps
    gates_per_s = n_gates / (elapsed / args.steps)
    jit_end = ka.jit_info()
    # a step leaves qubits on the index bits the last pass chose; reading the state first puts every
    # qubit back (restore passes): one more step (after one that compiles the restore passes) timed
    # with that read-out included
    step()
    sv.device_ptr()
    t1 = time.perf_counter()
    step()
    sv.device_ptr()
    _lib.check(lib.kq_synchronize())
    ms_with_restore = (time.perf_counter() - t1) * 1e3
    # the circuit on the resident dense state: every pass reads and writes the whole vector
    step_dense()
    _lib.check(lib.kq_synchronize())
    dense_stats = np.zeros(len(_lib.PASS_CLASSES), dtype=_lib.PASS_STATS_DTYPE)
    _lib.check(lib.kq_profile_begin())
    t1 = time.perf_counter()
    for _ in range(3):
        step_dense()
    _lib.check(lib.kq_synchronize())
    ms_dense = (time.perf_counter() - t1) * 1e3 / 3
    _lib.check(lib.kq_profile_end(_lib.ptr(dense_stats)))
    dense_tile_ms = float(dense_stats["ms"][0] + dense_stats["ms"][1])
    dense_tile_bytes = float(dense_stats["bytes"][0] + dense_stats["bytes"][1])
    dense_launches = int(dense_stats["launches"][0] + dense_stats["launches"][1])
    # (every dense pass reads and writes the whole state; the timed steps' first passes do not)
    fastest = min(float(dense_stats[c]["min_ms"]) for c in (0, 1) if dense_stats[c]["launches"]) if dense_launches else 0.0
    passes = {}
    for i, cls in enumerate(_lib.PASS_CLASSES):
        s = stats[i]
        if s["launches"]:
            passes[cls] = {
                "launches_per_step": int(s["launches"]) // args.steps,
                "gates_per_step": int(s["gates"]) // args.steps,
                "avg_ms": round(float(s["ms"]) / int(s["launches"]), 4),
                "GB_per_s": round(float(s["bytes"]) / (float(s["ms"]) * 1e-3) / 1e9, 1),
            }
    tile = stats[0].copy()
    for f in ("ms", "bytes"):
        tile[f] = stats[0][f] + stats[1][f]
    tile["launches"] = stats[0]["launches"] + stats[1]["launches"]
    kernel_ms = float(stats["ms"].sum()) / args.steps
    achieved = float(tile["bytes"]) / (float(tile["ms"]) * 1e-3) / 1e9 if tile["ms"] else 0.0
    jit_used = jit_end["cache_hits"] > jit_warm["cache_hits"]
    roofline = {
        "bound": "hbm",
        "kernel": "kq_pass (run-time compiled tile pass, double)" if jit_used else "k_tile<double>",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        # algorithmic bytes per launch, averaged over the step's launches: 13 passes read and write 2^n
        # amplitudes, the first (built from the basis index) only writes them
        "traffic": round(TRAFFIC_PER_ALGORITHMIC_BYTE * float(tile["bytes"]) / max(int(tile["launches"]), 1)),
        "bytes_per_launch": round(float(tile["bytes"]) / max(int(tile["launches"]), 1)),
        "avg_launch_ms": round(float(tile["ms"]) / max(int(tile["launches"]), 1), 4),
    }
    out = {
        "metric": "gates/s & HBM % at 30q c128; VQE energy+grad/s, adjoint vs plain GPU",
        "value": round(gates_per_s, 3),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": f"{n}q complex128 QFT + {args.layers} random 1q layers with CX brick ladders",
            "qubits": n,
            "gates": n_gates,
            "passes_per_step": n_passes,
            "low_qubits_per_tile": tuner["low_qubits"] if tuner["on"] else (args.low_qubits or 5),
            "max_stages_per_pass": tuner["max_stages"] if tuner["on"] else (args.max_stages or 12),
            "tile_settings": ("plan tuner (cost model, host-only, chosen in the warm-up call; "
                              f"{tuner['plans_tried']} candidate plans)") if tuner["on"] else "fixed by flags",
            "tile_passes_per_step": n_tiles,
            "kernel_ms_per_step": round(kernel_ms, 3),
            "ms_per_step_with_state_restored": round(ms_with_restore, 3),
            "step": ("kq_state_set_basis(0) + kq_apply_gates (kettle's Statevector(n) + simulate()): the reset is "
                     "left pending, the first pass reads nothing and runs its gates only in the tile holding the "
                     "basis amplitude, the other 13 read and write the whole state"),
            "dense_resident_state": {
                "ms_per_step": round(ms_dense, 3),
                "gates_per_s": round(n_gates / (ms_dense * 1e-3), 1),
                "kq_pass_GB_per_s": round(dense_tile_bytes / (dense_tile_ms * 1e-3) / 1e9, 1) if dense_tile_ms else None,
                "kq_pass_avg_launch_ms": round(dense_tile_ms / max(dense_launches, 1), 4),
                "note": "the same circuit on the previous run's output (no reset): all 14 passes read and write 2^n "
                        "amplitudes of dense data",
            },
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "passes": passes,
        "jit": {"mode": args.jit, "tuning": args.jit_tuning, "kernels_compiled": jit_end["compiled"], "failures": jit_end["failures"],
                "compile_s_outside_timed_region": round(jit_warm["compile_seconds"], 2),
                "compiles_inside_timed_region": jit_end["compiled"] - jit_warm["compiled"]},
    }
    del sv
    if not args.no_stream_probe:
        log("hbm stream probe")
        det = (ctypes.c_double * 24)()
        _lib.check(lib.kq_stream_probe_shapes(n, det))
        modes = ("read", "write", "in_place", "copy")
        shapes = ("persistent_2_per_cu", "persistent_8_per_cu", "block_per_4096", "block_per_4096_nontemporal",
                  "block_per_4096_prio", "block_per_4096_nontemporal_prio")
        best = [max(det[4 * sh + m] for sh in range(6)) for m in range(4)]
        probe = {m: round(best[i], 1) for i, m in enumerate(modes)}
        probe.update({"unit": "GB/s", "elements": f"2^{n} x 16 B", "best_of": "6 launch shapes",
                      "shapes": {sh: [round(det[4 * j + m], 1) for m in range(4)] for j, sh in enumerate(shapes)}})
        out["hbm_stream_probe"] = probe
        # the tile pass reads and writes every amplitude once, like the in-place stream
        roofline["frac_of_in_place_stream"] = round(achieved / max(best[2], 1e-9), 4)
        # fastest single pass launch of the dense steps (a read and a write of the whole state) against the probe
        roofline["fastest_launch_GB_per_s"] = round(2 * 16 * (1 << n) / (fastest * 1e-3) / 1e9, 1) if fastest > 0 else None
    if args.c64_qubits > 0:
        log(f"complex64 section: {args.c64_qubits} qubits")
        out["complex64_large"] = c64_section(args.c64_qubits, args.layers, args.c64_steps)
    if not args.no_plain:
        log("plain statevector section: 20 qubits")
        out["plain_sv_20q"] = plain_sv_section(20, 5)
    if not args.no_dm:
        log(f"density matrix section: {args.dm_qubits} qubits")
        out["density_matrix"] = density_matrix_section(args.dm_qubits, 2)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_sample_qubits, args.cpu_reps, args.cpu_per_class)
    if not args.no_vqe:
        log("vqe section")
        out["vqe"] = vqe_section(args.vqe_qubits, 20, 200, args.vqe_energies)
    if not args.no_batch:
        log("batched vqe section")
        out["batch"] = batch_section(args.batch_instances, args.batch_reps, 32)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
