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
