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

T_START = time.perf_counter()
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


def max_amp_error(m, n: int, gates: np.ndarray, ref_amps) -> float:
    """max |a_gpu - a_ref| over all 2^n amplitudes, the GPU state re-run on
    the same gate prefix and read back in 2^24-amplitude chunks."""
    worst = 0.0
    with m.StateVector(n, m.C128) as st:
        st.apply(gates)
        chunk = 1 << min(n, 24)
        for off in range(0, 1 << n, chunk):
            a = st.read(off, chunk)
            worst = max(worst, float(np.max(np.abs(a - ref_amps[off:off + chunk]))))
    return worst


def host_info() -> dict:
    """CPU model, logical CPUs, the CPUs this process may use and host RAM."""
    model, mem = None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal"):
                    mem = int(line.split()[1]) * 1024
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)),
            "host_ram_GiB": round(mem / 2 ** 30, 1) if mem else None}


def family_counts(gates: np.ndarray) -> dict:
    names = {4: "H", 7: "RX", 8: "RY", 9: "RZ", 10: "CNOT"}
    out = {}
    for t in gates["type"]:
        k = names.get(int(t), str(int(t)))
        out[k] = out.get(k, 0) + 1
    return out


def cpu_baseline(n: int, gates: np.ndarray, prefix: int, m=None) -> dict | None:
    """The reference's C (oracle/_ref) on the host: the first `prefix` gates
    of the same sweep (the H layer and the first full layer: rotations and
    CNOTs), gate by gate; state init excluded.  With the engine module `m`,
    the GPU's state after the same prefix is compared with the reference's
    amplitude by amplitude over the whole vector."""
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
        ga = gates_array([tuple(g) for g in gates[:prefix]])
        done = 0
        t0 = last = time.perf_counter()
        while done < len(ga):
            rc = ref.L.ref_apply_gate(s, ga[done:done + 1].ctypes.data)
            if rc:
                raise RuntimeError(f"reference gate {done} failed ({rc})")
            done += 1
            if time.perf_counter() - last > 10.0:
                last = time.perf_counter()
                print(f"[bench] sweep cpu baseline: {done}/{len(ga)} gates", file=sys.stderr, flush=True)
        dt = time.perf_counter() - t0
        err = None
        if m is not None and done > 0:
            D = 1 << n
            ptr = ref.L.ref_state_amplitudes(s)
            ref_amps = np.ctypeslib.as_array(ptr, shape=(2 * D,)).view(np.complex128)
            err = max_amp_error(m, n, gates[:done], ref_amps)
    finally:
        ref.free(s)
    b = algorithmic_bytes(gates[:done], n)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(b / dt / 1e12, 6), "unit": "TB/s", "cores": cores, "kind": "reference",
            "sample": f"first {done} of {len(gates)} gates of the same {n}-qubit sweep (the H layer and the "
                      f"first full layer: {b / 1e9:.1f} GB algorithmic, {dt:.1f} s), tsotchke/moonlab C "
                      f"(gates.c stride-block loops, OpenMP on {cores} threads) built by oracle/Makefile",
            "gates_per_s": round(done / dt, 3),
            "host": host_info(),
            "gates_by_family": family_counts(gates[:done]),
            "max_abs_amp_error_gpu_vs_reference": err,
            "amp_error_over": f"all 2^{n} amplitudes after the {done}-gate prefix"}


def batch_cpu_baseline(inst, budget_s: float) -> dict | None:
    """The reference's C on the host: the VQE energy loop (gate_ry, gate_rz,
    gate_cz, then measurement_correlation_zz / measurement_expectation_x) over
    the batch in its own order until `budget_s` of CPU work."""
    try:
        from oracle.oracle import RefLib
        from moonlab_amd.workloads import hea_gate_list
        ref = RefLib()
    except (FileNotFoundError, OSError) as e:
        print(f"[bench] batch cpu baseline unavailable: {e}", file=sys.stderr, flush=True)
        return None
    done = 0
    energies = []
    t0 = time.perf_counter()
    for n, L, ang, h in inst:
        if time.perf_counter() - t0 > budget_s:
            break
        s = ref.state(n, hea_gate_list(n, L, ang))
        e = -sum(ref.L.measurement_correlation_zz(s, q, q + 1) for q in range(n - 1))
        e -= h * sum(ref.L.measurement_expectation_x(s, q) for q in range(n))
        ref.free(s)
        energies.append(e)
        done += 1
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(done / dt, 3), "unit": "circuits/s", "cores": cores, "kind": "reference",
            "sample": f"first {done} of {len(inst)} instances in batch order ({dt:.1f} s), "
                      "tsotchke/moonlab C gate_* + measurement_* built by oracle/Makefile",
            "energies": energies}


FP64_PEAK_TFLOPS = 78.6


def hea_flops(inst) -> float:
    """Algorithmic FP64 flops of the HEA/TFIM batch: each RY and RZ a dense
    complex 2x2 on 2^(n-1) pairs (4 complex multiply-adds, 28 flops per pair),
    CZ sign flips counted as free, and the observables: |a|^2 (3 flops) plus
    n - 1 ZZ accumulations per amplitude, and 2 Re(a_i* a_j) (4 flops) per pair
    for each <X_q>."""
    f = 0.0
    for n, L, _, _ in inst:
        D = 1 << n
        f += 2 * L * n * (D // 2) * 28 + D * 3 + D * (n - 1) + n * (D // 2) * 4
    return f


def batch_plain_gpu(m, inst, sample: int, batched_e) -> dict:
    """Plain per-circuit dispatch of the same instances on the same GPU: one
    device state per qubit count, reset, one kernel per gate, then the TFIM
    terms through the single-state reductions (one per term), circuit after
    circuit, on the first `sample` instances in batch order."""
    from moonlab_amd import workloads
    states = {}
    got = []
    for n, L, ang, h in inst[:8]:                     # warm the code objects and states
        st = states.setdefault(n, m.StateVector(n, m.C128))
        st.reset().apply(workloads.hea_gate_list(n, L, ang))
        st.synchronize()
    t0 = time.perf_counter()
    for n, L, ang, h in inst[:sample]:
        st = states.setdefault(n, m.StateVector(n, m.C128))
        st.reset().apply(workloads.hea_gate_list(n, L, ang))
        e = -sum(st.correlation_zz(q, q + 1) for q in range(n - 1)) - h * sum(st.expect_x(q) for q in range(n))
        got.append(e)
    dt = time.perf_counter() - t0
    for st in states.values():
        st.free()
    return {"circuits_per_s": round(sample / dt, 1), "sample": f"first {sample} of {len(inst)} instances in batch "
            "order, one circuit at a time: state reset, one kernel per gate, one reduction per TFIM term",
            "seconds": round(dt, 3),
            "max_abs_energy_diff_vs_batched": float(np.max(np.abs(np.array(got) - batched_e[:sample])))}


def run_batch_workload(m, count: int, steps: int, warmup: int, cpu_seconds: float, with_cpu: bool,
                       plain_sample: int = 512) -> dict:
    from moonlab_amd import workloads
    inst = workloads.hea_batch(count, SEED)
    b = m.Batch()
    circuits = [{"n": n, "gates": workloads.hea_gate_list(n, L, ang), "observables": m.OBS_ZZ | m.OBS_X}
                for n, L, ang, h in inst]
    t = time.perf_counter()
    b.run(circuits, autotune=True)            # picks workgroups/CU and tile bits on this batch
    tuned = dict(b.last_stats)
    print(f"[bench] batch autotune {time.perf_counter() - t:.1f}s -> {tuned}", file=sys.stderr, flush=True)
    # the instances are packed once (HeaPack), as a parameter loop keeps
    # them; each timed step writes the angle vector into the pack (the
    # optimizer's update) and evaluates every energy
    pack = m.HeaPack(inst)
    theta = pack.angles.copy()
    for _ in range(warmup):
        b.hea_tfim(pack)
    walls, devs = [], []
    e = None
    for k in range(steps):
        t0 = time.perf_counter()
        pack.angles[:] = theta
        e = b.hea_tfim(pack)
        walls.append(time.perf_counter() - t0)
        devs.append(b.last_stats["device_seconds"])
        print(f"[bench] batch step {k + 1}/{steps}: {walls[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    wall = float(np.sum(walls))
    out = {"workload": f"{count} hardware-efficient-ansatz instances, n ~ U{{8..18}}, layers ~ U{{4..12}}, "
                       f"RY+RZ per qubit + CZ ladder, TFIM energy (h ~ U[0.5,1.5]), complex128, seed {SEED}",
           "step": "write the angle vector into the packed instances (HeaPack), then every energy",
           "circuits_per_s": round(count * steps / wall, 1),
           "ms_per_batch": round(wall / steps * 1e3, 3),
           "device_ms_per_batch": round(float(np.mean(devs)) * 1e3, 3),
           "launches_per_batch": 1,
           "autotune": {"tile_bits": tuned["tile_bits"], "workgroups": tuned["workgroups"]},
           "hbm_passes_per_batch": tuned["total_passes"]}
    flops = hea_flops(inst)
    dev = float(np.mean(devs))
    out["roofline"] = {"bound": "fp64 issue (circuits resident in LDS; see profiles/k_batch_hea4096.md)",
                       "kernel": "k_batch", "achieved": round(flops / dev / 1e12, 2), "peak": FP64_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": round(flops / dev / 1e12 / FP64_PEAK_TFLOPS, 4),
                       "algorithmic_flops_per_batch": flops}
    if plain_sample > 0:
        out["plain_per_circuit_gpu"] = batch_plain_gpu(m, inst, min(plain_sample, count), e)
        out["batched_vs_plain_gpu"] = round(out["circuits_per_s"] / out["plain_per_circuit_gpu"]["circuits_per_s"], 1)
    if with_cpu:
        cpu = batch_cpu_baseline(inst, cpu_seconds)
        if cpu:
            got = np.array(cpu.pop("energies"))
            out["max_abs_energy_error_vs_reference"] = float(np.max(np.abs(e[:len(got)] - got))) if len(got) else None
            out["cpu_baseline"] = cpu
    return out


def serving_cpu_baseline(reqs, budget_s: float, max_qubits: int = 18) -> dict | None:
    """The reference's request path on the host: moonlab_qgtl_execute (build,
    simulate, 1024 shots) for the requests of at most `max_qubits` qubits in
    arrival order until `budget_s` (one larger request alone can take the
    host minutes: the reference clones the state for every shot)."""
    try:
        from oracle.oracle import RefLib
        ref = RefLib()
    except (FileNotFoundError, OSError) as e:
        print(f"[bench] serving cpu baseline unavailable: {e}", file=sys.stderr, flush=True)
        return None
    small = [r for r in reqs if r["n"] <= max_qubits]
    done = 0
    t0 = last = time.perf_counter()
    for r in small:
        if time.perf_counter() - t0 > budget_s:
            break
        ref.qgtl_execute(r["n"], r["gates"], r["shots"], r["shot_seed"])
        done += 1
        if time.perf_counter() - last > 5.0:
            last = time.perf_counter()
            print(f"[bench] serving cpu baseline: {done} requests", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(done / dt, 3), "unit": "circuits/s", "cores": cores, "kind": "reference",
            "sample": f"first {done} of the {len(small)} requests with n <= {max_qubits} (of {len(reqs)}) in "
                      f"arrival order ({dt:.1f} s), each through tsotchke/moonlab's moonlab_qgtl_execute "
                      "(complex128, 1024 shots) built by oracle/Makefile"}


def complex64_accuracy(m, reqs, done, max_qubits: int = 22) -> dict:
    """Served complex64 results against FP64: <Z_q> of each served request vs
    the engine's complex128 run of the same circuit, and the fidelity of the
    complex64 state with the complex128 state (mixed-precision inner product)."""
    dz, fid, k = 0.0, 1.0, 0
    for r, d in zip(reqs, done):
        if r["precision"] != m.C64 or r["n"] > max_qubits or d.status != 0:
            continue
        with m.StateVector(r["n"], m.C128) as s128, m.StateVector(r["n"], m.C64) as s64:
            s128.apply(r["gates"])
            z, _ = s128.expect_z_all()
            s64.apply(r["gates"])
            ov = s64.inner(s128)
            fid = min(fid, abs(ov) ** 2 / (s64.norm() * s128.norm()))
        dz = max(dz, float(np.max(np.abs(np.asarray(d.expect_z) - np.asarray(z)))))
        k += 1
    return {"requests_checked": k, "max_qubits": max_qubits, "max_abs_dz_vs_fp64": dz,
            "min_fidelity_vs_fp64": fid}


def run_serving_workload(m, count: int, rate: float, sequential: bool) -> dict:
    """Poisson arrivals submitted in real time; latency = finish - arrival."""
    import threading
    from moonlab_amd import workloads
    reqs = workloads.serving_requests(count, SEED + 5, rate)
    srv = m.Server(sequential=sequential)
    try:
        # warm the engines (state pool, code objects) outside the measured stream
        for r in reqs[:4]:
            srv.wait(srv.submit(r["n"] if r["n"] <= 16 else 12, r["gates"][:8] if r["n"] > 16 else r["gates"],
                                shots=16, seed=1, precision=r["precision"]))
        done = [None] * len(reqs)
        tickets = [None] * len(reqs)
        ready = threading.Semaphore(0)

        def collect():
            for i in range(len(reqs)):
                ready.acquire()
                done[i] = srv.wait(tickets[i])
                if (i + 1) % 50 == 0:
                    print(f"[bench] serving ({'plain' if sequential else 'served'}) completed {i + 1}/{len(reqs)}",
                          file=sys.stderr, flush=True)

        th = threading.Thread(target=collect)
        th.start()
        t0 = srv.now() + 0.05
        for i, r in enumerate(reqs):
            delay = t0 + r["arrival"] - srv.now()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(r["n"], r["gates"], shots=r["shots"], seed=r["shot_seed"],
                                    precision=r["precision"])
            ready.release()
            if (i + 1) % 100 == 0:
                print(f"[bench] serving ({'plain' if sequential else 'served'}) submitted {i + 1}/{len(reqs)}",
                      file=sys.stderr, flush=True)
        th.join()
        st = srv.stats()
    finally:
        srv.close()
    accuracy = None if sequential else complex64_accuracy(m, reqs, done)
    lat = np.array([d.finish_time - d.submit_time for d in done])
    first = min(d.submit_time for d in done)
    last = max(d.finish_time for d in done)
    bad = sum(1 for d in done if d.status != 0)
    by_path = {}
    for name, p in (("batched", 0), ("large", 1), ("plain", 2)):
        sel = [d for d in done if d.path == p]
        if sel:
            w = np.array([d.start_time - d.submit_time for d in sel])
            sv = np.array([d.finish_time - d.start_time for d in sel])
            by_path[name] = {"requests": len(sel), "wait_p50_ms": round(float(np.percentile(w, 50)) * 1e3, 2),
                             "service_p50_ms": round(float(np.percentile(sv, 50)) * 1e3, 2),
                             "service_sum_s": round(float(np.sum(sv)), 3)}
    return {"circuits_per_s": round(len(done) / (last - first), 2),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
            "by_path": by_path,
            "complex64_vs_fp64": accuracy,
            "makespan_s": round(last - first, 3), "failed": bad,
            "batches": int(st["batches"]), "largest_batch": int(st["largest_batch"]),
            "peak_admitted_GiB": round(st["peak_bytes_admitted"] / 2 ** 30, 3)}


def fusion_cpu_baseline(n: int, gates: np.ndarray, budget_s: float) -> dict | None:
    """The reference's C (complex128, its only precision) on a prefix of the
    same n-qubit circuit, gate by gate, until `budget_s` (state init excluded);
    skipped when the host lacks memory for the state."""
    try:
        import psutil
        from oracle.oracle import RefLib, gates_array
        ref = RefLib()
    except (ImportError, FileNotFoundError, OSError) as e:
        print(f"[bench] fusion cpu baseline unavailable: {e}", file=sys.stderr, flush=True)
        return None
    need = (1 << n) * 16
    if psutil.virtual_memory().available < 2 * need:
        print(f"[bench] fusion cpu baseline skipped: {need / 2**30:.0f} GiB state does not fit the host",
              file=sys.stderr, flush=True)
        return None
    print(f"[bench] fusion cpu baseline: allocating the {need / 2**30:.0f} GiB reference state", file=sys.stderr,
          flush=True)
    t_init = time.perf_counter()
    s = ref.L.ref_state_new(n)
    if not s:
        return None
    t_init = time.perf_counter() - t_init
    try:
        ga = gates_array([tuple(g) for g in gates])
        done = 0
        t0 = time.perf_counter()
        last = t0
        while done < len(ga) and time.perf_counter() - t0 < budget_s:
            if ref.L.ref_apply_gate(s, ga[done:done + 1].ctypes.data):
                raise RuntimeError(f"reference gate {done} failed")
            done += 1
            if time.perf_counter() - last > 5.0:
                last = time.perf_counter()
                print(f"[bench] fusion cpu baseline: {done} gates", file=sys.stderr, flush=True)
        dt = time.perf_counter() - t0
    finally:
        ref.free(s)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(done / dt, 4), "unit": "gates/s", "cores": cores, "kind": "reference",
            "sample": f"first {done} of {len(gates)} gates of the same {n}-qubit circuit in complex128 "
                      f"({dt:.1f} s; state init {t_init:.1f} s excluded), tsotchke/moonlab C built by oracle/Makefile"}


def run_blocked_workload(m, n: int, depth: int) -> dict:
    """The large-request engine: the n-qubit random-layer circuit (the
    serving stream's circuit family; H layer first) run as cache-blocked
    register-phase passes (k_reg_pass), beside the same circuit as fused
    <= 5-qubit MFMA blocks and as one kernel per gate, per precision.  The
    complex128 register-engine state is checked against the per-gate state
    (fidelity, and max |amplitude| difference over 64 random runs of 4096
    amplitudes)."""
    from moonlab_amd import workloads
    g = workloads.random_layer_circuit(n, depth, SEED + 7, h_layer=True)
    D = 1 << n
    out = {}
    with m.StateVector(12, m.C64) as w:         # load the code objects outside the timing
        w.apply(g[:40] if n == 12 else workloads.random_layer_circuit(12, 2, 1), blocked=True)
        w.synchronize()

    def timed(st, **kw):
        st.timer_begin(2)
        st.timer_mark()
        st.apply(g, **kw)
        st.timer_mark()
        return float(st.timer_elapsed(1)[0]) * 1e-3

    for prec, name in ((m.C128, "c128"), (m.C64, "c64")):
        b = 8 if prec == m.C64 else 16
        passes = m.blocked_pass_count(g, n, prec)
        m.set_blocked_lookahead(0)
        greedy = m.blocked_pass_count(g, n, prec)   # the planner without lookahead, for comparison
        m.set_blocked_lookahead()
        res = {"passes": passes, "passes_per_layer": round(passes / depth, 2), "passes_greedy_planner": greedy}
        st = m.StateVector(n, prec)
        t_reg = timed(st, blocked=True)
        res["register_seconds"] = round(t_reg, 4)
        res["register_ms_per_pass"] = round(t_reg / passes * 1e3, 3)
        res["register_TB_per_s_per_pass"] = round(passes * 2 * D * b / t_reg / 1e12, 3)
        if prec == m.C128:
            ref = m.StateVector(n, prec)
            t_gate = timed(ref)
            res["per_gate_seconds"] = round(t_gate, 4)
            _, n1 = st.expect_z_all()
            _, n2 = ref.expect_z_all()
            res["fidelity_vs_per_gate"] = abs(st.inner(ref)) ** 2 / (n1 * n2)
            rng = np.random.default_rng(SEED)
            run = min(D, 1 << 12)
            err = 0.0
            for off in rng.integers(0, D // run, size=64) * run:   # 64 random runs of 4096 amplitudes
                a = st.read(int(off), run)
                c = ref.read(int(off), run)
                err = max(err, float(np.max(np.abs(a - c))))
            res["max_abs_amp_diff_vs_per_gate_sampled"] = err
            ref.free()
        st.free()
        st = m.StateVector(n, prec)
        t_fused = timed(st, fuse=5)
        st.free()
        res["fused_seconds"] = round(t_fused, 4)
        res["speedup_vs_fused"] = round(t_fused / t_reg, 3)
        if "per_gate_seconds" in res:
            res["speedup_vs_per_gate"] = round(res["per_gate_seconds"] / t_reg, 3)
        print(f"[bench] blocked n={n} {name}: {passes} passes, register {t_reg:.3f} s, fused {t_fused:.3f} s",
              file=sys.stderr, flush=True)
        out[name] = res
    return out


def run_fusion_workload(m, depth: int, cases) -> dict:
    """Precision + fusion: depth-`depth` random SU(2) + CZ brickwork circuits,
    fused into <= 5-qubit blocks (composed in FP64) and contracted on the
    matrix cores (FP32 MFMA for complex64, FP64 MFMA for complex128).  The
    FP64 comparison state is the cache-blocked complex128 path, which shares
    neither the fusion planner nor the MFMA kernel."""
    from moonlab_amd import workloads
    out = {}
    with m.StateVector(12, m.C64) as w:         # load the code objects outside the timing
        w.apply(workloads.su2_cz_circuit(12, 2, 1), fuse=5)
        w.synchronize()
    for n, prec in cases:
        g = workloads.su2_cz_circuit(n, depth, SEED + n)
        plan = m.fusion_plan(g, 5)
        ks = [len(b.qubits) for b in plan]
        D = 1 << n
        amp_b = 8 if prec == m.C64 else 16
        dense_flops = sum(8 * (1 << k) * D for k in ks if k >= 3)
        passes = len(plan)
        print(f"[bench] fusion n={n} {'c64' if prec == m.C64 else 'c128'}: {len(g)} gates -> {passes} blocks",
              file=sys.stderr, flush=True)
        st = m.StateVector(n, prec)
        st.timer_begin(2)
        st.timer_mark()
        st.apply(g, fuse=5)
        st.timer_mark()
        t_fused = float(st.timer_elapsed(1)[0]) * 1e-3
        z, nrm = st.expect_z_all()
        print(f"[bench] fusion n={n}: fused run {t_fused:.2f} s", file=sys.stderr, flush=True)
        fid = None
        try:
            ref = m.StateVector(n, m.C128)
        except m.EngineError:
            ref = None
        if ref is None:                           # both states do not fit: compare <Z> only
            st.free()
            st = None
            ref = m.StateVector(n, m.C128)
        ref.timer_begin(2)
        ref.timer_mark()
        ref.apply(g, blocked=True)
        ref.timer_mark()
        t_ref = float(ref.timer_elapsed(1)[0]) * 1e-3
        z_ref, nrm_ref = ref.expect_z_all()
        if st is not None:
            ov = st.inner(ref)
            fid = abs(ov) ** 2 / (nrm * nrm_ref)
            st.free()
        ref.free()
        print(f"[bench] fusion n={n}: FP64 comparison run {t_ref:.2f} s", file=sys.stderr, flush=True)
        peak = 157.3 if prec == m.C64 else 78.6
        tflops = dense_flops / t_fused / 1e12
        out[f"{n}q_{'c64' if prec == m.C64 else 'c128'}"] = {
            "gates": len(g), "fused_blocks": passes,
            "block_qubits_hist": {str(k): ks.count(k) for k in sorted(set(ks))},
            "device_seconds": round(t_fused, 4),
            "ms_per_block": round(t_fused / passes * 1e3, 3),
            "state_TB_per_s": round(passes * 2 * D * amp_b / t_fused / 1e12, 3),
            "mfma_TFLOP_per_s": round(tflops, 2),
            "mfma_peak_TFLOP_per_s": peak,
            "mfma_frac": round(tflops / peak, 4),
            "fidelity_vs_fp64": fid,
            "max_abs_dz_vs_fp64": float(np.max(np.abs(np.asarray(z) - np.asarray(z_ref)))),
            "norm": nrm,
            "fp64_comparison": "cache-blocked complex128 path",
            "fp64_comparison_seconds": round(t_ref, 3),
            "gates_per_s": round(len(g) / t_fused, 2),
        }
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=None,
                    help="layers of the random-layer circuit (default 40 for the sweep, 10 for --workload blocked)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-count", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=100.0)
    ap.add_argument("--serve-sweep", default="200,400",
                    help="other arrival rates (per s) at which the served leg runs again")
    ap.add_argument("--no-fusion", action="store_true", help="leave out the fusion leg of the default line")
    ap.add_argument("--batch-plain-sample", type=int, default=512,
                    help="instances of the batch run one at a time for the plain per-circuit GPU leg")
    ap.add_argument("--time-budget", type=float, default=540.0,
                    help="seconds the whole default run may take: legs past it are skipped and named")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--workload", choices=("sweep", "fusion", "blocked", "batch", "serving"), default="sweep",
                    help="sweep: the 30-qubit gate sweep (with the batch and serving legs); "
                         "fusion: 32q complex64 and 33q complex128 SU(2)+CZ depth 60 on fused MFMA blocks; "
                         "blocked: cache-blocked register-phase passes; "
                         "batch: the batch leg alone; serving: the served stream alone (both for profiler runs)")
    ap.add_argument("--fusion-depth", type=int, default=60)
    ap.add_argument("--fusion-cases", default="32:c64,33:c128")
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

    if args.workload == "blocked":
        t0 = time.perf_counter()
        depth = args.depth if args.depth is not None else 10
        res = run_blocked_workload(m, args.qubits, depth)
        wall = time.perf_counter() - t0
        c128 = res["c128"]
        print(json.dumps({
            "metric": "circuits/s, complex128, cache-blocked register-phase passes (large-request path)",
            "value": round(1.0 / c128["register_seconds"], 3), "unit": "circuits/s", "n_gpus": 1, "steps": 1,
            "warmup": 1, "ms_per_step": round(wall * 1e3, 1), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{args.qubits}-qubit random-layer circuit (H layer, depth {depth} of "
                                   "RX|RY|RZ + CNOT brickwork), cache-blocked register-phase passes",
                       "qubits": args.qubits, "depth": depth},
            "roofline": {"bound": "hbm", "kernel": "k_reg_pass<double>",
                         "achieved": round(c128["register_TB_per_s_per_pass"] * 1e3, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(c128["register_TB_per_s_per_pass"] * 1e3 / HBM_PEAK_GBS, 4),
                         "traffic": None},
            "cases": res}), flush=True)
        return 0

    if args.workload == "serving":
        t0 = time.perf_counter()
        sv = run_serving_workload(m, args.serve_count, args.serve_rate, False)
        wall = time.perf_counter() - t0
        print(json.dumps({
            "metric": "circuits/s, served request stream (co-batching + cache-blocked large requests)",
            "value": sv["circuits_per_s"], "unit": "circuits/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": round(sv["makespan_s"] * 1e3, 1), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64/f32", "data": "synthetic",
            "config": {"workload": f"{args.serve_count} Poisson requests at {args.serve_rate:g}/s"},
            "served": sv, "bench_seconds": round(wall, 1)}), flush=True)
        return 0

    if args.workload == "batch":
        t0 = time.perf_counter()
        res = run_batch_workload(m, args.batch_count, args.batch_steps, 1, 10.0, not args.no_cpu,
                                 args.batch_plain_sample)
        wall = time.perf_counter() - t0
        print(json.dumps({
            "metric": "circuits/s, persistent batch launch (HEA + TFIM energy, complex128)",
            "value": res["circuits_per_s"], "unit": "circuits/s", "n_gpus": 1, "steps": args.batch_steps,
            "warmup": 1, "ms_per_step": res["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": res["workload"]}, "batch": res, "bench_seconds": round(wall, 1)}), flush=True)
        return 0

    if args.workload == "fusion":
        cases = [(int(c.split(":")[0]), m.C64 if c.split(":")[1] == "c64" else m.C128)
                 for c in args.fusion_cases.split(",")]
        t0 = time.perf_counter()
        res = run_fusion_workload(m, args.fusion_depth, cases)
        wall = time.perf_counter() - t0
        cpu = None
        if not args.no_cpu:
            n0 = cases[0][0]
            cpu = fusion_cpu_baseline(n0, workloads.su2_cz_circuit(n0, args.fusion_depth, SEED + n0), 20.0)
        first = next(iter(res.values()))
        print(json.dumps({
            "metric": "fidelity and <Z> error vs FP64, MFMA TFLOP/s on fused SU(2)+CZ circuits",
            "value": first["mfma_TFLOP_per_s"], "unit": "TFLOP/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": round(wall * 1e3, 1), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32/f64", "data": "synthetic",
            "config": {"workload": f"random SU(2) per qubit + CZ brickwork, depth {args.fusion_depth}, "
                                   "fused into <= 5-qubit blocks (FP64 composition) on MFMA",
                       "cases": args.fusion_cases},
            "cases": res, "cpu_baseline": cpu}), flush=True)
        return 0

    n = args.qubits
    if args.depth is None:
        args.depth = 40
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
    skipped = {}

    def fits(leg: str, est_s: float) -> bool:
        """Run a leg only if its estimated time still fits the run's budget."""
        left = args.time_budget - (time.perf_counter() - T_START)
        if est_s <= left:
            return True
        skipped[leg] = f"estimated {est_s:.0f} s, {left:.0f} s of the {args.time_budget:.0f} s budget left"
        print(f"[bench] skipping {leg}: {skipped[leg]}", file=sys.stderr, flush=True)
        return False

    cpu = None
    if not args.no_cpu and fits("sweep cpu_baseline", 75.0):
        cpu = cpu_baseline(n, gates, 3 * n - 1, m)   # the H layer and the first full layer
    fusion = None
    if not args.no_fusion and n == 30 and fits("fusion", 110.0):
        cases = [(int(c.split(":")[0]), m.C64 if c.split(":")[1] == "c64" else m.C128)
                 for c in args.fusion_cases.split(",")]
        fusion = {"workload": f"random SU(2) per qubit + CZ brickwork, depth {args.fusion_depth}, fused into <= 5-qubit "
                              "blocks composed in FP64, contracted on MFMA (FP32 for complex64, FP64 for complex128)",
                  "cases": run_fusion_workload(m, args.fusion_depth, cases)}
        if not args.no_cpu and fits("fusion cpu_baseline", 45.0):
            n0 = cases[0][0]
            fusion["cpu_baseline"] = fusion_cpu_baseline(n0, workloads.su2_cz_circuit(n0, args.fusion_depth, SEED + n0),
                                                         10.0)
    batch = None
    if not args.no_batch and fits("batch", 40.0):
        batch = run_batch_workload(m, args.batch_count, args.batch_steps, 1, 10.0, not args.no_cpu,
                                   args.batch_plain_sample)
    serving = None
    if not args.no_serve and fits("serving", 80.0):
        served = run_serving_workload(m, args.serve_count, args.serve_rate, False)
        plain = run_serving_workload(m, args.serve_count, args.serve_rate, True)
        cpu_serve = None
        if not args.no_cpu:
            cpu_serve = serving_cpu_baseline(workloads.serving_requests(args.serve_count, SEED + 5, args.serve_rate),
                                             10.0)
        # the same request stream at other arrival rates, served only: plain
        # dispatch runs one request at a time, so its circuits/s does not
        # depend on the arrival rate once arrivals outpace it (~9/s)
        by_rate = {}
        for r in [float(x) for x in args.serve_sweep.split(",") if x.strip()]:
            if r == args.serve_rate or not fits(f"serving at {r:g}/s", 12.0):
                continue
            sv = run_serving_workload(m, args.serve_count, r, False)
            by_rate[f"{r:g}"] = {"circuits_per_s": sv["circuits_per_s"], "p50_ms": sv["p50_ms"],
                                 "p99_ms": sv["p99_ms"], "makespan_s": sv["makespan_s"], "failed": sv["failed"],
                                 "ratio_vs_plain": round(sv["circuits_per_s"] / plain["circuits_per_s"], 2)}
        serving = {"workload": f"{args.serve_count} Poisson requests at {args.serve_rate:g}/s: random-layer "
                               "circuits, n in 10..30 with P(n) ~ 2^(-(n-10)/4), depth ~ U{10..50}, "
                               "complex128/complex64 50/50, 1024 shots + <Z_q> each, 256 GiB HBM budget",
                   "served": served, "plain_sequential": plain,
                   "circuits_per_s_ratio": round(served["circuits_per_s"] / plain["circuits_per_s"], 2),
                   "served_at_other_rates": by_rate,
                   "cpu_baseline": cpu_serve}
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
        "fusion": fusion,
        "batch": batch,
        "serving": serving,
        "skipped_for_time": skipped or None,
        "bench_seconds": round(time.perf_counter() - T_START, 1),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
