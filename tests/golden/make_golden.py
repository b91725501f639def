"""Write the golden fixtures in tests/golden/ from the reference's own C.

Runs in the build container only: it needs oracle/_ref/libmoonlab_ref.so,
built by `make -C oracle ref` from the reference sources (tsotchke/moonlab
v1.2.0).  Inputs are the reference's 75-circuit oracle corpus
(tests/oracle/corpus/circuit_corpus.json, seed 20260717) and seeded circuits
from moonlab_amd.workloads; outputs are what the reference computes for them.

    python tests/golden/make_golden.py [--ref-corpus PATH]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib, gates_array  # noqa: E402
from oracle import oracle as O  # noqa: E402
from moonlab_amd import workloads  # noqa: E402

NAME_TO_CODE = {"h": O.H, "x": O.X, "y": O.Y, "z": O.Z, "s": O.S, "sdg": O.SDG, "t": O.T,
                "tdg": O.TDG, "rx": O.RX, "ry": O.RY, "rz": O.RZ, "cnot": O.CNOT, "cz": O.CZ,
                "swap": O.SWAP}


class JobResults(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_int), ("total_shots", ctypes.c_int),
                ("outcomes", ctypes.POINTER(ctypes.c_uint64)), ("num_workers_used", ctypes.c_int),
                ("worker_seconds", ctypes.POINTER(ctypes.c_double))]


class QOpts(ctypes.Structure):
    _fields_ = [("num_shots", ctypes.c_int), ("rng_seed", ctypes.c_uint64),
                ("return_probabilities", ctypes.c_int)]


class QRes(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_int), ("num_shots", ctypes.c_int),
                ("outcomes", ctypes.POINTER(ctypes.c_uint64)),
                ("probabilities", ctypes.POINTER(ctypes.c_double))]


def ref_job(L, n, gates, shots, workers, seed):
    """Run a job through the reference's moonlab_scheduler_run."""
    j = L.moonlab_job_create(n)
    for g in gates:
        t, tgt, ctl, th = int(g[0]), int(g[1]), int(g[2]), float(g[4])
        p = (ctypes.c_double * 1)(th) if t in (O.RX, O.RY, O.RZ) else None
        assert L.moonlab_job_add_gate(j, t, tgt, ctl, p) == 0
    L.moonlab_job_set_num_shots(j, shots)
    L.moonlab_job_set_num_workers(j, workers)
    L.moonlab_job_set_rng_seed(j, seed)
    need = L.moonlab_job_to_json(j, None, 0)
    buf = ctypes.create_string_buffer(need + 1)
    L.moonlab_job_to_json(j, buf, need + 1)
    r = JobResults()
    rc = L.moonlab_scheduler_run(j, ctypes.byref(r))
    assert rc == 0, rc
    out = np.array([r.outcomes[i] for i in range(r.total_shots)], dtype=np.uint64)
    used = r.num_workers_used
    L.moonlab_job_results_free(ctypes.byref(r))
    L.moonlab_job_free(j)
    return out, used, buf.value.decode()


def ref_qgtl(L, n, gates, shots, seed, probs):
    c = L.moonlab_qgtl_circuit_create(n)
    for g in gates:
        t, tgt, ctl, th = int(g[0]), int(g[1]), int(g[2]), float(g[4])
        p = (ctypes.c_double * 1)(th) if t in (O.RX, O.RY, O.RZ) else None
        assert L.moonlab_qgtl_add_gate(c, t, tgt, ctl, p) == 0
    need = ctypes.c_size_t(0)
    L.moonlab_qgtl_circuit_serialize(c, None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value + 1)
    L.moonlab_qgtl_circuit_serialize(c, buf, need.value + 1, None)
    o = QOpts(shots, seed, 1 if probs else 0)
    r = QRes()
    assert L.moonlab_qgtl_execute(c, ctypes.byref(o), ctypes.byref(r)) == 0
    outc = np.array([r.outcomes[i] for i in range(shots)], dtype=np.uint64) if shots else None
    pr = np.ctypeslib.as_array(r.probabilities, (1 << n,)).copy() if probs else None
    L.moonlab_qgtl_results_free(ctypes.byref(r))
    L.moonlab_qgtl_circuit_free(c)
    return outc, pr, buf.value.decode()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref-corpus", default="/root/reference/tests/oracle/corpus/circuit_corpus.json")
    args = ap.parse_args()
    ref = RefLib()
    L = ref.L
    vp, ip = ctypes.c_void_p, ctypes.c_int
    L.moonlab_job_create.restype = vp
    L.moonlab_job_create.argtypes = [ip]
    for f in ("moonlab_job_free",):
        getattr(L, f).argtypes = [vp]
    L.moonlab_job_add_gate.argtypes = [vp, ip, ip, ip, ctypes.POINTER(ctypes.c_double)]
    L.moonlab_job_set_num_shots.argtypes = [vp, ip]
    L.moonlab_job_set_num_workers.argtypes = [vp, ip]
    L.moonlab_job_set_rng_seed.argtypes = [vp, ctypes.c_uint64]
    L.moonlab_job_to_json.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    L.moonlab_scheduler_run.argtypes = [vp, ctypes.POINTER(JobResults)]
    L.moonlab_job_results_free.argtypes = [ctypes.POINTER(JobResults)]
    L.moonlab_qgtl_circuit_create.restype = vp
    L.moonlab_qgtl_circuit_create.argtypes = [ip]
    L.moonlab_qgtl_add_gate.argtypes = [vp, ip, ip, ip, ctypes.POINTER(ctypes.c_double)]
    L.moonlab_qgtl_circuit_serialize.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t,
                                                 ctypes.POINTER(ctypes.c_size_t)]
    L.moonlab_qgtl_execute.argtypes = [vp, ctypes.POINTER(QOpts), ctypes.POINTER(QRes)]
    L.moonlab_qgtl_results_free.argtypes = [ctypes.POINTER(QRes)]
    L.moonlab_qgtl_circuit_free.argtypes = [vp]

    # 1. the reference's oracle corpus: circuits (inputs) + amplitudes and <Z> (outputs)
    corpus = json.load(open(args.ref_corpus))["circuits"]
    blob = {}
    index = []
    for c in corpus:
        n = c["num_qubits"]
        rows = []
        for g in c["gates"]:
            q = g["q"]
            t = NAME_TO_CODE[g["g"]]
            rows.append((t, q[0], -1, -1, g["p"]) if len(q) == 1 else (t, q[1], q[0], -1, g["p"]))
        ga = gates_array(rows)
        amps, z = ref.run(n, ga)
        key = c["id"]
        blob[key + ":gates"] = ga
        blob[key + ":amps"] = amps
        blob[key + ":z"] = z
        index.append((key, n))
    np.savez_compressed(os.path.join(HERE, "corpus_ref.npz"), **blob)
    with open(os.path.join(HERE, "corpus_index.json"), "w") as f:
        json.dump({"source": "tests/oracle/corpus/circuit_corpus.json (seed 20260717)",
                   "circuits": index}, f, indent=1)

    # 2. the 20-qubit, depth-20 random-layer circuit (seed 20261202)
    n, seed = 20, 20261202
    ga = workloads.random_layer_circuit(n, 20, seed)
    amps, z, p = ref.run(n, ga, want_probs=True)
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(1 << n, size=4096, replace=False)).astype(np.int64)
    np.savez_compressed(os.path.join(HERE, "layers20_ref.npz"), seed=np.uint64(seed), depth=20,
                        gates=ga, z=z, sample_idx=idx, sample_amps=amps[idx],
                        norm=np.sum(p), prob_top=np.sort(p)[-64:])

    # 3. shot outcomes through the reference job path and circuit path
    bell = [(O.H, 0, -1, -1, 0.0), (O.CNOT, 1, 0, -1, 0.0)]
    ghz = [(O.H, 0, -1, -1, 0.0), (O.CNOT, 1, 0, -1, 0.0), (O.CNOT, 2, 1, -1, 0.0)]
    rnd10 = workloads.random_layer_circuit(10, 6, 99)
    shots = {}
    for name, n, gl, nshot, w, sd in (("bell", 2, bell, 1024, 4, 0xdeadbeef),
                                       ("ghz", 3, ghz, 3000, 3, 0xcafe),
                                       ("rnd10", 10, rnd10, 2000, 7, 0x1234567),
                                       ("rnd10_w1", 10, rnd10, 999, 1, 42)):
        out, used, js = ref_job(L, n, gates_array(gl), nshot, w, sd)
        shots[name] = {"n": n, "shots": nshot, "workers": w, "seed": sd, "used": used,
                       "outcomes": out.tolist(), "json": js}
    outc, pr, text = ref_qgtl(L, 10, rnd10, 1500, 0xfeedface, True)
    shots["qgtl_rnd10"] = {"n": 10, "shots": 1500, "seed": 0xfeedface, "outcomes": outc.tolist(),
                           "probabilities": pr.tolist(), "text": text}
    np.save(os.path.join(HERE, "rnd10_gates.npy"), rnd10)
    with open(os.path.join(HERE, "shots_ref.json"), "w") as f:
        json.dump(shots, f)

    # 4. hardware-efficient-ansatz TFIM energies (RY, RZ per qubit + CZ ladder)
    rng = np.random.default_rng(11)
    rows = []
    for _ in range(24):
        n = int(rng.integers(2, 11))
        layers = int(rng.integers(1, 6))
        h = float(rng.uniform(0.5, 1.5))
        ang = rng.uniform(0.0, 2.0 * np.pi, size=(layers, n, 2))
        gl = []
        for l in range(layers):
            for q in range(n):
                gl.append((O.RY, q, -1, -1, float(ang[l, q, 0])))
                gl.append((O.RZ, q, -1, -1, float(ang[l, q, 1])))
            for q in range(n - 1):
                gl.append((O.CZ, q + 1, q, -1, 0.0))
        s = ref.state(n, gates_array(gl))
        e = -sum(L.measurement_correlation_zz(s, q, q + 1) for q in range(n - 1))
        e -= h * sum(L.measurement_expectation_x(s, q) for q in range(n))
        ref.free(s)
        rows.append({"n": n, "layers": layers, "h": h, "angles": ang.ravel().tolist(), "energy": e})
    with open(os.path.join(HERE, "hea_tfim_ref.json"), "w") as f:
        json.dump(rows, f)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
