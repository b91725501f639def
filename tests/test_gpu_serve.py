"""Serving loop: parity, invariance alone vs in the mix, admission.

Tolerances: complex128 <Z_q> within 1e-12 of the oracle and shot outcomes
exact (the request's xorshift64 draws, the collapse outcome rule);
complex64 <Z_q> within 2e-5.  A request's results are bit-identical alone and
in a mix (both precisions: the same kernels run with the same per-request
reduction order).
"""
import numpy as np
import pytest

from moonlab_amd import workloads

pytestmark = pytest.mark.gpu


def _submit_all(srv, reqs):
    return [srv.submit(r["n"], r["gates"], shots=r["shots"], seed=r["shot_seed"], precision=r["precision"])
            for r in reqs]


@pytest.mark.parametrize("large_path", ["fused", "blocked"])
def test_served_results_match_oracle(eng, oracle, large_path):
    reqs = workloads.serving_requests(40, 7, 1000.0, n_range=(6, 14), depth_range=(4, 12), shots=256)
    for r in reqs:
        r["precision"] = eng.C128
    with eng.Server(batch_max_qubits=12, large_path=large_path) as srv:
        tickets = _submit_all(srv, reqs)
        res = [srv.wait(t) for t in tickets]
    paths = set()
    for r, x in zip(reqs, res):
        a = oracle.run(r["n"], r["gates"])
        assert x.status == 0
        assert np.max(np.abs(x.expect_z - oracle.expect_z_all(r["n"], a))) < 1e-12
        assert x.outcomes.tolist() == oracle.qgtl_shots(r["n"], a, r["shots"], r["shot_seed"]).tolist()
        assert x.submit_time <= x.start_time <= x.finish_time
        paths.add(x.path)
    assert paths == {0, 1}   # both the co-batched and the single-state path were exercised


def test_results_identical_alone_and_in_mix(eng):
    reqs = workloads.serving_requests(60, 11, 1000.0, n_range=(8, 20), depth_range=(5, 20), shots=512)
    with eng.Server(batch_max_qubits=14) as srv:
        mixed = [srv.wait(t) for t in _submit_all(srv, reqs)]
        alone = [srv.wait(srv.submit(r["n"], r["gates"], shots=r["shots"], seed=r["shot_seed"],
                                     precision=r["precision"])) for r in reqs]
    for r, a, b in zip(reqs, mixed, alone):
        assert a.outcomes.tobytes() == b.outcomes.tobytes(), r["n"]
        assert a.expect_z.tobytes() == b.expect_z.tobytes(), r["n"]


def test_complex64_requests(eng, oracle):
    reqs = workloads.serving_requests(20, 13, 1000.0, n_range=(8, 18), depth_range=(5, 15), shots=64)
    with eng.Server(batch_max_qubits=14) as srv:
        for r in reqs:
            r["precision"] = eng.C64
        res = [srv.wait(t) for t in _submit_all(srv, reqs)]
    for r, x in zip(reqs, res):
        a = oracle.run(r["n"], r["gates"])
        assert np.max(np.abs(x.expect_z - oracle.expect_z_all(r["n"], a))) < 2e-5


def test_admission_by_footprint(eng):
    n = 20
    budget = 3 * (1 << n) * 16          # three 20-qubit complex128 states
    reqs = workloads.serving_requests(12, 17, 1000.0, n_range=(n, n), depth_range=(3, 5), shots=16)
    for r in reqs:
        r["precision"] = eng.C128
    with eng.Server(hbm_budget_bytes=budget, batch_max_qubits=12, large_workers=6) as srv:
        res = [srv.wait(t) for t in _submit_all(srv, reqs)]
        st = srv.stats()
        assert all(x.status == 0 for x in res)
        assert st["peak_bytes_admitted"] <= budget
        assert st["large_requests"] == len(reqs)
        with pytest.raises(eng.SchedulerError):
            srv.submit(22, [(4, 0)], shots=1)      # 64 MiB > budget: never admissible


def test_sequential_baseline_agrees(eng):
    reqs = workloads.serving_requests(16, 19, 1000.0, n_range=(8, 16), depth_range=(4, 10), shots=128)
    for r in reqs:
        r["precision"] = eng.C128
    with eng.Server(batch_max_qubits=12) as srv:
        served = [srv.wait(t) for t in _submit_all(srv, reqs)]
    with eng.Server(sequential=True) as srv:
        seq = [srv.wait(t) for t in _submit_all(srv, reqs)]
    for a, b in zip(served, seq):
        assert b.path == 2
        assert np.max(np.abs(a.expect_z - b.expect_z)) < 1e-12
        assert a.outcomes.tolist() == b.outcomes.tolist()


def _frame_circuits(k, seed):
    """Random circuits over the circuit API's gate set, as (n, gate tuples, text)."""
    from moonlab_amd import Circuit
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        n = int(rng.integers(3, 15))
        c, gl = Circuit(n), []
        for _ in range(int(rng.integers(5, 40))):
            t = int(rng.integers(0, 14))
            tg = int(rng.integers(0, n))
            ctl = int((tg + rng.integers(1, n)) % n) if t >= 10 else -1
            th = float(rng.uniform(-np.pi, np.pi)) if t in (7, 8, 9) else None
            c.add_gate(t, tg, ctl, th)
            gl.append((t, tg, ctl, -1, th if th is not None else 0.0))
        out.append((n, gl, c.serialize().encode()))
    return out


def _decode(resp, dtype):
    head, _, body = resp.partition(b"\n")
    verb, count = head.split(b" ")
    return verb, np.frombuffer(body, dtype=dtype).copy(), int(count)


def test_control_frames_match_reference_vectors(eng, oracle):
    """CIRCUIT and SHOTS frames through the server against the circuit API's
    golden vectors (moonlab_qgtl_execute, seed fixed): probabilities within
    1e-13 (the blocked path folds gate runs into FP64 2x2s), outcomes exact;
    then 48 frames from 8 threads at once, each identical to its run alone and
    to the oracle."""
    import json
    import os
    from concurrent.futures import ThreadPoolExecutor
    q = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "shots_ref.json")))["qgtl_rnd10"]
    text = q["text"].encode()
    with eng.Server() as srv:
        assert eng.handle_frame(b"HEALTH\n", srv) == (eng.CONTROL_OK, b"OK alive\n")
        rc, resp = eng.handle_frame(b"CIRCUIT %d\n" % len(text) + text, srv)
        verb, p, cnt = _decode(resp, np.float64)
        assert rc == eng.CONTROL_OK and verb == b"OK" and cnt == 1 << 10 == len(p)
        assert np.max(np.abs(p - np.array(q["probabilities"]))) < 1e-13
        rc, resp = eng.handle_frame(b"SHOTS %d %d\n" % (q["shots"], len(text)) + text, srv, seed=q["seed"])
        verb, o, cnt = _decode(resp, np.uint64)
        assert rc == eng.CONTROL_OK and verb == b"SAMPLES" and cnt == q["shots"]
        assert o.tolist() == q["outcomes"]

        circ = _frame_circuits(24, 11)
        frames = []
        for i, (n, gl, t) in enumerate(circ):
            frames.append((b"CIRCUIT %d\n" % len(t) + t, 0))
            frames.append((b"SHOTS 300 %d\n" % len(t) + t, 1000 + i))
        alone = [eng.handle_frame(f, srv, seed=s) for f, s in frames]
        with ThreadPoolExecutor(8) as ex:
            mixed = list(ex.map(lambda fs: eng.handle_frame(fs[0], srv, seed=fs[1]), frames))
    assert mixed == alone
    for i, (n, gl, _) in enumerate(circ):
        a = oracle.run(n, gl)
        _, p, _ = _decode(alone[2 * i][1], np.float64)
        assert np.max(np.abs(p - oracle.probabilities(n, a))) < 1e-13
        _, o, _ = _decode(alone[2 * i + 1][1], np.uint64)
        assert o.tolist() == oracle.qgtl_shots(n, a, 300, 1000 + i).tolist()


def test_extension_gates_through_circuit_api_and_frames(eng, oracle):
    """Controlled and multi-qubit gates added with add_gate_ex run through
    moonlab_qgtl_execute and through a CIRCUIT frame: probabilities against
    the oracle's gate table (1e-13); seeded SHOTS frames equal the circuit
    API's outcomes for the same seed."""
    rng = np.random.default_rng(12)
    n = 9
    c = eng.Circuit(n)
    gl = []
    for _ in range(80):
        t = int(rng.integers(0, 25))
        qs = [int(x) for x in rng.permutation(n)[:3]]
        if t <= 13:
            th = float(rng.uniform(-3, 3)) if t in (7, 8, 9) else None
            ctl = qs[1] if t >= 10 else -1
            c.add_gate(t, qs[0], ctl, th)
            gl.append((t, qs[0], ctl, -1, th or 0.0))
        else:
            nq = 3 if t in (23, 24) else (2 if t >= 18 else 1)
            npar = 3 if t == 17 else (1 if t in (16, 18, 19, 20, 21) else 0)
            ps = [float(x) for x in rng.uniform(-3, 3, npar)]
            c.add_gate_ex(t, qs[:nq], ps)
            full = qs[:nq] + [-1] * (3 - nq)
            pp = ps + [0.0] * (3 - npar)
            gl.append((t, full[0], full[1], full[2], pp[0], pp[1], pp[2]))
    want = oracle.probabilities(n, oracle.run(n, gl))
    r = c.execute(num_shots=64, rng_seed=99, return_probabilities=True)
    assert np.max(np.abs(r.probabilities - want)) < 1e-13
    text = c.serialize().encode()
    with eng.Server() as srv:
        rc, resp = eng.handle_frame(b"CIRCUIT %d\n" % len(text) + text, srv)
        assert rc == eng.CONTROL_OK
        p = np.frombuffer(resp.partition(b"\n")[2], dtype=np.float64)
        assert np.max(np.abs(p - want)) < 1e-13
        rc, resp = eng.handle_frame(b"SHOTS 64 %d\n" % len(text) + text, srv, seed=99)
        assert rc == eng.CONTROL_OK
        assert np.frombuffer(resp.partition(b"\n")[2], dtype=np.uint64).tolist() == r.outcomes.tolist()


def test_bad_request_fails_alone(eng, oracle):
    """A request with a gate the engine rejects (qubit out of range, repeated
    qubit, unknown code) fails its own submit; the valid requests submitted
    around it, which it would have been co-batched with, all complete and
    match the oracle."""
    reqs = workloads.serving_requests(12, 23, 1000.0, n_range=(6, 10), depth_range=(3, 6), shots=64)
    bad_gates = [[(4, 0), (10, 12, 3)],          # CNOT target past n = 10
                 [(4, 0), (10, 2, 2)],           # control == target
                 [(4, 0), (25, 0)],              # unknown code
                 [(4, 0), (23, 1, 0, 1)]]        # Toffoli with a repeated qubit
    with eng.Server(batch_max_qubits=12) as srv:
        tickets = []
        for i, r in enumerate(reqs):
            tickets.append(srv.submit(r["n"], r["gates"], shots=r["shots"], seed=r["shot_seed"], precision=eng.C128))
            if i < len(bad_gates):
                with pytest.raises(eng.SchedulerError):
                    srv.submit(10, bad_gates[i], shots=8, seed=1, precision=eng.C128)
        res = [srv.wait(t) for t in tickets]
    for r, x in zip(reqs, res):
        assert x.status == 0
        a = oracle.run(r["n"], r["gates"])
        assert np.max(np.abs(x.expect_z - oracle.expect_z_all(r["n"], a))) < 1e-12
        assert x.outcomes.tolist() == oracle.qgtl_shots(r["n"], a, r["shots"], r["shot_seed"]).tolist()


def test_frames_attribute_tenant_and_fire_hook(eng):
    """moonlab_mi355x_server_handle_frame_ex (control_plane.c:699-731): the
    tenant is the request context while the frame runs, the completion hook
    fires once per completed CIRCUIT / SHOTS frame with job NULL, backend
    "control_plane", num_qubits and total_shots (0 for CIRCUIT), never for a
    rejected frame or HEALTH, and the context is cleared afterwards."""
    L = eng.lib()
    seen = []

    @eng.COMPLETION_HOOK_FN
    def hook(job, res, backend, ctx):
        seen.append((bool(job), backend.decode(), res.contents.num_qubits, res.contents.total_shots,
                     L.moonlab_scheduler_current_tenant_id()))

    text = eng.Circuit(3).add_gate(eng.GateType.H, 0).add_gate(eng.GateType.CNOT, 1, 0).serialize().encode()
    assert L.moonlab_scheduler_set_completion_hook(hook, None) == 0
    try:
        with eng.Server() as srv:
            rc, _ = eng.handle_frame(b"CIRCUIT %d\n" % len(text) + text, srv, tenant="acme-corp")
            assert rc == eng.CONTROL_OK
            rc, _ = eng.handle_frame(b"SHOTS 100 %d\n" % len(text) + text, srv, seed=5, tenant="globex")
            assert rc == eng.CONTROL_OK
            rc, _ = eng.handle_frame(b"SHOTS 10 %d\n" % len(text) + text, srv, seed=5)
            assert rc == eng.CONTROL_OK
            assert eng.handle_frame(b"HEALTH\n", srv, tenant="acme-corp")[0] == eng.CONTROL_OK
            bad = b"NUM_QUBITS 2\nFOO 0\n"
            assert eng.handle_frame(b"CIRCUIT %d\n" % len(bad) + bad, srv, tenant="acme-corp")[0] == \
                eng.CONTROL_REJECTED
        assert seen == [(False, "control_plane", 3, 0, b"acme-corp"),
                        (False, "control_plane", 3, 100, b"globex"),
                        (False, "control_plane", 3, 10, None)]
        assert L.moonlab_scheduler_current_tenant_id() is None
    finally:
        L.moonlab_scheduler_set_completion_hook(eng.COMPLETION_HOOK_FN(), None)
