#!/usr/bin/env python3
"""Gate-sweep benchmark of the MI355X dense state-vector engine.

One step = the whole 30-qubit complex128 random-layer sweep (an H layer, then
40 layers of RX|RY|RZ on every qubit + a CNOT brickwork: 2390 gates), applied
unfused, one kernel per gate, to a 16 GiB state resident in HBM.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]

Prints ONE JSON line: whole-job algorithmic HBM throughput (TB/s), the
roofline of the dominant kernel (per-launch HIP-event durations on the
state's stream, over the timed region) and the reference's own C on the
host cores, timed on a bounded prefix of the same sweep.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEED = 20261202


def kernel_of(t: int) -> str:
    if t in (4, 7, 8, 1, 2, 17):
        return "k_pair"          # dense 2x2 on pairs (H, RX, RY, X, Y, U3)
    if t == 9:
        return "k_diag"          # RZ: both halves rescaled, contiguous stream
    if t in (10, 11, 22, 19, 20, 21, 23):
        return "k_ctrl_pair"     # controlled pairs (CNOT ...)
    return "other"


def cpu_baseline(n: int, gates: np.ndarray, budget_s: float) -> dict | None:
    """The reference's C (oracle/_ref) on the host: a prefix of the same sweep,
    gate by gate, until `budget_s` of CPU work; state init excluded."""
    try:
        from oracle.oracle import RefLib, gates_array
        from moonlab_amd.workloads import algorithmic_bytes
        ref = RefLib()
    except (FileNotFoundError, OSError) as e:
        print(f"[bench] cpu baseline unavailable: {e}", file=sys.stderr, flush=True)
        return None
    s = ref.L.ref_state_new(n)
    if not s:
        return None
    try:
        ga = gates_array([tuple(g) for g in gates])
        done = 0
        t0 = time.perf_counter()
        while done < len(ga) and time.perf_counter() - t0 < budget_s:
            rc = ref.L.ref_apply_gate(s, ga[done:done + 1].ctypes.data)
            if rc:
                raise RuntimeError(f"reference gate {done} failed ({rc})")
            done += 1
        dt = time.perf_counter() - t0
    finally:
        ref.free(s)
    b = algorithmic_bytes(gates[:done], n)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(b / dt / 1e12, 6), "unit": "TB/s", "cores": cores, "kind": "reference",
            "sample": f"first {done} of {len(gates)} gates of the same {n}-qubit sweep "
                      f"({b / 1e9:.1f} GB algorithmic, {dt:.1f} s), tsotchke/moonlab C "
                      f"(gates.c stride-block loops, OpenMP) built by oracle/Makefile",
            "gates_per_s": round(done / dt, 3)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="short run for profiler counters: one device copy of the state "
                         "(known bytes, the calibration kernel) and the first layer of the sweep")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: this build runs on a single GPU "
              "(multi-GPU sharding is a design in DESIGN.md, not a run).", file=sys.stderr)
        return 2

    import moonlab_amd as m
    from moonlab_amd import workloads

    n = args.qubits
    gates = workloads.random_layer_circuit(n, args.depth, SEED, h_layer=True)
    G = len(gates)
    bytes_step = workloads.algorithmic_bytes(gates, n)
    kinds = [kernel_of(int(t)) for t in gates["type"]]
    D = 1 << n
    per_launch = {"k_pair": 2 * D * 16, "k_diag": 2 * D * 16, "k_ctrl_pair": 2 * (D // 2) * 16}
    print(f"[bench] n={n} gates/step={G} algorithmic={bytes_step / 1e12:.3f} TB/step", file=sys.stderr,
          flush=True)

    if args.probe:
        a = m.StateVector(n, m.C128)
        b = m.StateVector(n, m.C128)
        a.apply(gates[:n])
        b.copy_from(a)          # k_convert<double, double>: reads and writes 2^n x 16 B
        b.synchronize()
        b.apply(gates[n:2 * n + (n - 1)])   # first layer: n rotations + n - 1 CNOTs
        b.synchronize()
        print(json.dumps({"probe": True, "qubits": n, "copy_bytes_each_way": D * 16,
                          "gates": int(2 * n + n - 1)}), flush=True)
        return 0

    st = m.StateVector(n, m.C128)
    L = m.lib()
    h = st.handle
    rows = [np.ascontiguousarray(gates[i:i + 1]) for i in range(G)]
    ptrs = [r.ctypes.data for r in rows]

    def sweep(mark: bool) -> None:
        for p in ptrs:
            rc = L.moonlab_hip_apply_gate(h, p)
            if rc:
                raise RuntimeError(m.last_error())
            if mark:
                L.moonlab_hip_timer_mark(h)

    for w in range(args.warmup):
        sweep(False)
        st.synchronize()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)

    K = args.steps
    st.timer_begin(K * G + 1)
    st.synchronize()
    t0 = time.perf_counter()
    st.timer_mark()
    for k in range(K):
        sweep(True)
        print(f"[bench] step {k + 1}/{K} enqueued", file=sys.stderr, flush=True)
    st.synchronize()
    t1 = time.perf_counter()
    gaps = st.timer_elapsed(K * G)
    wall = t1 - t0
    _, nrm = st.expect_z_all()
    st.free()

    dur = {}
    for i, ms in enumerate(gaps):
        dur.setdefault(kinds[i % G], []).append(float(ms))
    stats = {k: {"launches": len(v), "avg_ms": float(np.mean(v)),
                 "total_ms": float(np.sum(v))} for k, v in dur.items()}
    dominant = max((k for k in stats if k in per_launch), key=lambda k: stats[k]["total_ms"])
    avg_s = stats[dominant]["avg_ms"] * 1e-3
    achieved = per_launch[dominant] / avg_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dominant}_{n}q.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")

    value = bytes_step * K / wall / 1e12
    cpu = None if args.no_cpu else cpu_baseline(n, gates, args.cpu_seconds)
    out = {
        "metric": "HBM TB/s on 30q gate sweep",
        "value": round(value, 4),
        "unit": "TB/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n}-qubit complex128 random-layer gate sweep: H layer + depth "
                               f"{args.depth} of (RX|RY|RZ per qubit, CNOT brickwork), {G} gates, "
                               f"one kernel per gate, seed {SEED}",
                   "qubits": n, "depth": args.depth, "gates_per_step": G,
                   "algorithmic_bytes_per_step": bytes_step},
        "gates_per_s": round(G * K / wall, 2),
        "kernels": {k: {"launches": v["launches"], "avg_ms": round(v["avg_ms"], 4),
                        "GB_per_s": round(per_launch[k] / (v["avg_ms"] * 1e-3) / 1e9, 1)
                        if k in per_launch else None} for k, v in stats.items()},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "norm_error": abs(nrm - 1.0),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
