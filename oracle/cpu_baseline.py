"""CPU baseline of the benchmark (TEST INFRASTRUCTURE: run only by bench.py's cpu_baseline leg, as a
child process pinned to one core): kettle's gate loop restated (the oracle, one thread, kettle's release
flags) timed per gate class on a bounded sample, then extrapolated to the whole circuit.

    taskset -c 0 python -m oracle.cpu_baseline <workload json>

Input (stdin or argv[1], JSON): {"qubits": n, "sample_qubits": m, "reps": r, "per_class": k}.  The
workload circuit is the benchmark's (kettle_amd.workloads.qft_plus_layers(n, 20)); for every kettle gate
type in it, k gates of that type spread through the circuit are each run r times at m qubits (one state,
so the allocation is outside the timing), the median of the r runs is taken, the per-type mean of those
medians is scaled by 2^(n - m) (the loops stream over every amplitude pair, so their time is linear in
the state size once it is far out of cache), and the circuit time is sum over types of count x time.
Prints one JSON object."""
from __future__ import annotations

import json
import os
import platform
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def host_info() -> dict:
    model = ""
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    libstdcxx = ""
    try:
        out = subprocess.run(["g++", "--version"], capture_output=True, text=True, timeout=10).stdout
        libstdcxx = out.splitlines()[0] if out else ""
    except (OSError, subprocess.SubprocessError):
        pass
    mem = 0
    try:
        mem = os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES")
    except (ValueError, OSError):
        pass
    return {"cpu_model": model or platform.processor() or platform.machine(), "nproc": os.cpu_count(),
            "affinity": sorted(os.sched_getaffinity(0)), "host_ram_gib": round(mem / 2**30, 1),
            "libstdcxx_of": libstdcxx}


def main() -> None:
    cfg = json.loads(sys.argv[1] if len(sys.argv) > 1 else sys.stdin.read())
    from kettle_amd import workloads
    from kettle_amd.circuit import Gate
    from oracle import oracle as orc
    n, m, reps, k = int(cfg["qubits"]), int(cfg["sample_qubits"]), int(cfg["reps"]), int(cfg["per_class"])
    ops, _, _ = workloads.qft_plus_layers(n=n, layers=20).flatten()
    kinds = sorted(set(int(g) for g in ops["gate"]))
    sample, owner = [], []
    for kind in kinds:
        idx = np.nonzero(ops["gate"] == kind)[0]
        pick = idx[np.unique(np.linspace(0, len(idx) - 1, min(k, len(idx))).astype(int))]
        for i in pick:
            op = ops[i].copy()
            # the same gate at m qubits: qubits >= m folded onto the top four of the sample
            fold = lambda q: int(q) if q < m else m - 1 - (int(q) - m) % 4   # noqa: E731
            controlled = int(Gate.CH) <= int(op["gate"]) <= int(Gate.CP) or int(op["gate"]) == int(Gate.CU)
            op["q0"] = fold(op["q0"])
            if controlled:
                op["q1"] = fold(op["q1"])
                if op["q0"] == op["q1"]:
                    op["q0"] = (int(op["q1"]) + 1) % m
            sample.append(op)
            owner.append(kind)
    sample = np.array(sample, dtype=ops.dtype)
    secs = orc.time_gate_reps(m, sample, reps)
    med = np.median(secs, axis=1)
    scale = 2.0 ** (n - m)
    per_type, total = {}, 0.0
    for kind in kinds:
        t = float(np.mean(med[np.array(owner) == kind])) * scale
        count = int(np.sum(ops["gate"] == kind))
        per_type[Gate(kind).name] = {"count": count, "est_s_per_gate": round(t, 4)}
        total += count * t
    out = {"value": round(len(ops) / total, 6), "unit": "gates/s", "cores": 1, "kind": "port",
           "sample": (f"{len(sample)} gates ({k} per kettle gate type, {len(kinds)} types) of the {n}-qubit circuit "
                      f"run {reps}x each at {m} qubits on one core, median per gate, x 2^{n - m} to {n} qubits, "
                      f"summed over the circuit's {len(ops)} gates by type; oracle gate loop only, "
                      f"{float(secs.sum()):.1f} s timed"),
           "per_type": per_type, "est_seconds_per_step": round(total, 1), **host_info()}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
