"""Fusion autotuner: picks the tile size k, stages per pass and per-pass arithmetic cap.

    python -m kettle_amd.autotune --qubits 30 --precision c128 [--layers 20] [--update]

For each candidate setting it times simulate() of the benchmark circuit (QFT + random layers with
CX brick ladders) on the device and keeps the fastest.  Results are stored per (precision, qubit
count) in kettle_amd/tuning.json; ``apply_tuning(precision, n)`` sets libkq's fusion parameters
from that table (nearest tuned qubit count).  It is an offline, measured sweep; libkq's own plan
tuner (kq_set_fusion_autotune, on by default) chooses tile settings per gate structure from a cost
model without running anything, and an explicit setting from this table switches that tuner off.
"""
from __future__ import annotations

import argparse
import itertools
import json
import time
from pathlib import Path

TUNING_PATH = Path(__file__).resolve().parent / "tuning.json"
_applied: tuple | None = None
_table: dict | None = None


def _load_table() -> dict:
    global _table
    if _table is None:
        _table = json.loads(TUNING_PATH.read_text()) if TUNING_PATH.exists() else {}
    return _table


def lookup(precision: int, n_qubits: int) -> dict | None:
    """The tuned setting for the nearest tuned qubit count of this precision (or None)."""
    from ._lib import KQ_C64

    key = "c64" if precision == KQ_C64 else "c128"
    entries = _load_table().get(key, {})
    if not entries:
        return None
    best = min(entries, key=lambda s: abs(int(s) - n_qubits))
    return entries[best]


def apply_tuning(precision: int, n_qubits: int) -> None:
    """Set libkq's fusion parameters for a state of this precision and size."""
    global _applied
    from .api import set_fusion

    t = lookup(precision, n_qubits)
    if t is None:
        return
    cfg = (t["k"], t["max_stages"], t["max_cost"])
    if cfg != _applied:
        set_fusion(t["k"], 4096, t["max_stages"], t["max_cost"])
        _applied = cfg


def reset_applied() -> None:
    global _applied
    _applied = None


def tune(n: int, precision: int, layers: int, ks, stages, costs, reps: int = 2, log=print) -> dict:
    import kettle_amd as ka
    from . import _lib, workloads

    circuit = workloads.qft_plus_layers(n=n, layers=layers)
    ops, _, _ = circuit.flatten()
    sv = ka.Statevector(n, precision)
    lib = _lib.load()
    results = []
    for k, st, cost in itertools.product(ks, stages, costs):
        ka.set_fusion(k, 4096, st, cost)
        sv.set_basis(0)
        ka.apply_gates(sv, ops)   # warm-up and plan
        _lib.check(lib.kq_synchronize())
        t0 = time.perf_counter()
        for _ in range(reps):
            sv.set_basis(0)
            ka.apply_gates(sv, ops)
        _lib.check(lib.kq_synchronize())
        dt = (time.perf_counter() - t0) / reps
        _, passes, _ = ka.plan_stats()
        rec = {"k": k, "max_stages": st, "max_cost": cost, "ms": round(dt * 1e3, 3),
               "gates_per_s": round(len(ops) / dt, 1), "passes": passes}
        results.append(rec)
        log(json.dumps(rec))
    best = min(results, key=lambda r: r["ms"])
    return {"best": best, "all": results}


def main() -> None:
    from ._lib import KQ_C64, KQ_C128

    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--precision", choices=["c128", "c64"], default="c128")
    ap.add_argument("--layers", type=int, default=20)
    ap.add_argument("--k", type=int, nargs="*", default=[10, 11, 12])
    ap.add_argument("--stages", type=int, nargs="*", default=[3, 4, 6, 8])
    ap.add_argument("--cost", type=float, nargs="*", default=[4, 6, 8, 12, 16, 1e9])
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--update", action="store_true", help="write the best setting into tuning.json")
    ap.add_argument("--out", type=str, default="")
    args = ap.parse_args()
    prec = KQ_C64 if args.precision == "c64" else KQ_C128
    res = tune(args.qubits, prec, args.layers, args.k, args.stages, args.cost, args.reps)
    print(json.dumps({"best": res["best"]}))
    if args.out:
        Path(args.out).write_text(json.dumps(res, indent=1))
    if args.update:
        table = _load_table()
        best = res["best"]
        table.setdefault(args.precision, {})[str(args.qubits)] = {
            "k": best["k"], "max_stages": best["max_stages"], "max_cost": best["max_cost"],
            "measured_gates_per_s": best["gates_per_s"],
        }
        TUNING_PATH.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
