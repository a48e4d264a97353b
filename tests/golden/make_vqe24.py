"""Writes vqe24_oracle.json: the oracle's FP64 energy of the 24-qubit, 960-parameter, 248-string VQE
workload (kettle_amd.workloads.vqe(24, 20, 200, seed 0)) and its exact parameter-shift gradient at a
few parameters, for the full-size adjoint checks on the device (the oracle is too slow to run there).

    python tests/golden/make_vqe24.py [--jobs 8]

Each oracle energy is a single-threaded simulate + expectation of 2^24 amplitudes (minutes); the
energies run in parallel processes."""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

N, LAYERS, N_RANDOM, SEED = 24, 20, 200, 0
SHIFT_PARAMS = [0, 317, 640, 959]


def _energy(task):
    import numpy as np
    from kettle_amd import workloads
    from oracle import oracle as orc
    op_index, delta = task
    w = workloads.vqe(n=N, layers=LAYERS, n_random=N_RANDOM, seed=SEED)
    ops, _, _ = w.circuit.flatten(w.params)
    if op_index >= 0:
        ops = ops.copy()
        ops["angle"][op_index] += delta
    t0 = time.time()
    amps, _ = orc.sv_simulate(orc.zero_state(N), N, ops)
    e = orc.pauli_expectation(amps, N, w.hamiltonian.packed()).real
    return task, float(e), time.time() - t0


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    from kettle_amd import workloads
    w = workloads.vqe(n=N, layers=LAYERS, n_random=N_RANDOM, seed=SEED)
    ops, _, _ = w.circuit.flatten(w.params)
    # each of these parameters drives exactly one RY / RZ gate in n_local
    op_of = {int(ops[i]["aux"]): i for i in range(len(ops)) if ops[i]["aux"] >= 0}
    tasks = [(-1, 0.0)] + [(op_of[p], s) for p in SHIFT_PARAMS for s in (math.pi / 2, -math.pi / 2)]
    with ProcessPoolExecutor(args.jobs) as ex:
        results = {t: (e, dt) for t, e, dt in ex.map(_energy, tasks)}
    energy = results[(-1, 0.0)][0]
    shift = {str(p): 0.5 * (results[(op_of[p], math.pi / 2)][0] - results[(op_of[p], -math.pi / 2)][0])
             for p in SHIFT_PARAMS}
    out = {
        "workload": f"workloads.vqe(n={N}, layers={LAYERS}, n_random={N_RANDOM}, seed={SEED})",
        "n_params": len(w.params),
        "n_strings": w.hamiltonian.size(),
        "energy": energy,
        "parameter_shift": shift,
        "oracle_seconds_per_energy": round(max(dt for _, dt in results.values()), 1),
    }
    (HERE / "vqe24_oracle.json").write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
