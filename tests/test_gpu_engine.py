"""HIP engine parity against the CPU oracle and the reference's goldens.

Tolerances: complex128 amplitudes and FP64 reductions within 1e-12 of the
oracle (absolute, per amplitude / per value); complex64 states within
5e-6 per amplitude and fidelity >= 1 - 1e-6 (FP32 storage, O(gates * 2^-24)
error).  Shot outcomes are compared exactly (same uniform streams, same
outcome rule) at n <= 20.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_corpus, random_gate_list, random_state

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL64 = 5e-6


def _bytes_equal(a, b):
    return a.tobytes() == b.tobytes()


def test_every_gate_code_vs_oracle(eng, oracle):
    rng = np.random.default_rng(1)
    n = 9
    for t in list(range(1, 25)):
        for trial in range(3):
            a0 = random_state(n, rng)
            qs = rng.permutation(n)[:3]
            g = [(t, int(qs[0]), int(qs[1]), int(qs[2]), float(rng.uniform(-7, 7)),
                  float(rng.uniform(-7, 7)), float(rng.uniform(-7, 7)))]
            want = oracle.apply(n, a0.copy(), g)
            with eng.StateVector(n) as s:
                s.load(a0).apply(g)
                got = s.amplitudes()
            assert np.max(np.abs(got - want)) < TOL, (t, trial)
            with eng.StateVector(n, eng.C64) as s:
                s.load(a0).apply(g)
                got = s.amplitudes()
            assert np.max(np.abs(got - want)) < TOL64, (t, trial, "c64")


def test_every_target_position_and_tuning(eng, oracle):
    """All target/control positions (incl. 0, 1, n-1) and every launch shape
    give bit-identical results."""
    rng = np.random.default_rng(2)
    n = 12
    a0 = random_state(n, rng)
    gl = []
    for q in range(n):
        gl += [(4, q), (7, q, -1, -1, 0.3 + q), (9, q, -1, -1, 1.1 * q), (6, q)]
    for c in range(n):
        for t in (0, 1, n - 1, (c + 5) % n):
            if t != c:
                gl += [(10, t, c), (12, t, c), (13, t, c), (11, t, c)]
    want = oracle.apply(n, a0.copy(), gl)
    ref_bytes = None
    for ept in (1, 2, 4, 8):
        for block in (64, 256, 1024):
            eng.set_gate_tuning(ept, block)
            with eng.StateVector(n) as s:
                got = s.load(a0).apply(gl).amplitudes()
            assert np.max(np.abs(got - want)) < TOL, (ept, block)
            if ref_bytes is None:
                ref_bytes = got
            assert _bytes_equal(got, ref_bytes), (ept, block)
    eng.set_gate_tuning(4, 256)
    # streaming (non-temporal) and cached access give the same bits
    eng.set_streaming_threshold(0)
    try:
        with eng.StateVector(n) as s:
            got = s.load(a0).apply(gl).amplitudes()
    finally:
        eng.set_streaming_threshold(512 << 20)
    assert _bytes_equal(got, ref_bytes)


def test_state_copy_and_precision_conversion(eng, oracle):
    rng = np.random.default_rng(21)
    n = 13
    a0 = random_state(n, rng)
    with eng.StateVector(n) as s, eng.StateVector(n) as t, eng.StateVector(n, eng.C64) as h:
        s.load(a0)
        t.copy_from(s)
        assert _bytes_equal(t.amplitudes(), a0)
        h.copy_from(s)
        assert np.max(np.abs(h.amplitudes() - a0)) < 1e-7
        t.copy_from(h)
        assert _bytes_equal(t.amplitudes(), a0.astype(np.complex64).astype(np.complex128))


def test_pauli_expectation_vs_oracle(eng, oracle):
    """moonlab_hip_expect_pauli against the restatement of
    vqe_measure_pauli_expectation, every letter on every qubit position."""
    rng = np.random.default_rng(32)
    for n in (1, 2, 5, 9, 13):
        gl = random_gate_list(n, 60, rng)
        a = oracle.run(n, gl)
        strings = ["".join(rng.choice(list("IXYZ"), size=n)) for _ in range(10)]
        strings += ["Y" * n, "X" * n, "Z" * n, "I" * n]
        with eng.StateVector(n) as s, eng.StateVector(n, eng.C64) as h:
            s.load(a)
            h.load(a)
            for p in strings:
                want = oracle.expect_pauli(n, a, p)
                assert abs(s.expect_pauli(p) - want) < TOL, p
                assert abs(h.expect_pauli(p) - want) < 1e-6, p
            with pytest.raises(eng.EngineError):
                s.expect_pauli("I" * n + "X")


def test_mixed_precision_inner_product(eng):
    """<a|b> between a complex64 and a complex128 state: FP64 products of the
    stored values, within 1e-12 of numpy's."""
    rng = np.random.default_rng(22)
    n = 14
    a0, b0 = random_state(n, rng), random_state(n, rng)
    with eng.StateVector(n, eng.C64) as a, eng.StateVector(n) as b:
        a.load(a0)
        b.load(b0)
        a32 = a.amplitudes()
        assert abs(a.inner(b) - np.vdot(a32, b0)) < 1e-12
        assert abs(b.inner(a) - np.vdot(b0, a32)) < 1e-12
        b.load(a0)
        assert abs(abs(a.inner(b)) ** 2 - 1.0) < 1e-6


def test_matrix_entry_points_vs_oracle(eng, oracle):
    import scipy.linalg
    rng = np.random.default_rng(3)
    n = 8

    def unitary(k):
        h = rng.normal(size=(k, k)) + 1j * rng.normal(size=(k, k))
        return scipy.linalg.expm(1j * (h + h.conj().T))

    for _ in range(6):
        a0 = random_state(n, rng)
        q0, q1, q2 = [int(x) for x in rng.permutation(n)[:3]]
        m2, m4 = unitary(2), unitary(4)
        want = oracle.apply_matrix_1q(n, a0.copy(), q0, m2)
        with eng.StateVector(n) as s:
            assert np.max(np.abs(s.load(a0).apply_1q(q0, m2).amplitudes() - want)) < TOL
        # moonlab_hip_apply_2q basis |q1 q0>, q0 low == oracle's (qubit1 = q0)
        want = oracle.apply_matrix_2q(n, a0.copy(), q0, q1, m4)
        with eng.StateVector(n) as s:
            assert np.max(np.abs(s.load(a0).apply_2q(q0, q1, m4).amplitudes() - want)) < TOL
        # multi-control X / Z against Toffoli / CZ compositions
        want = oracle.apply(n, a0.copy(), [(23, q2, q0, q1)])
        with eng.StateVector(n) as s:
            got = s.load(a0).apply_mcx((1 << q0) | (1 << q1), q2).amplitudes()
        assert np.max(np.abs(got - want)) < TOL
        want = oracle.apply(n, a0.copy(), [(12, q1, q0)])
        with eng.StateVector(n) as s:
            got = s.load(a0).apply_mcz((1 << q0) | (1 << q1)).amplitudes()
        assert np.max(np.abs(got - want)) < TOL
        # controlled 2x2 == controlled-RY through the gate table
        th = float(rng.uniform(-3, 3))
        ry = np.array([[np.cos(th / 2), -np.sin(th / 2)], [np.sin(th / 2), np.cos(th / 2)]])
        want = oracle.apply(n, a0.copy(), [(20, q1, q0, -1, th)])
        with eng.StateVector(n) as s:
            got = s.load(a0).apply_controlled_1q(1 << q0, q1, ry).amplitudes()
        assert np.max(np.abs(got - want)) < TOL


def test_reference_corpus(eng):
    worst = 0.0
    for key, n, g, amps, z in load_corpus():
        with eng.StateVector(n) as s:
            got = s.apply(g).amplitudes()
            zz, nrm = s.expect_z_all()
        worst = max(worst, np.max(np.abs(got - amps)))
        assert np.max(np.abs(zz - z)) < TOL, key
        assert abs(nrm - 1.0) < TOL
    assert worst < TOL


def test_random_full_table_circuits(eng, oracle):
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 5, 11, 12, 13, 16):
        gl = random_gate_list(n, 120, rng)
        want = oracle.run(n, gl)
        with eng.StateVector(n) as s:
            got = s.apply(gl).amplitudes()
        assert np.max(np.abs(got - want)) < TOL, n


def test_layers20_against_reference_and_oracle(eng, oracle):
    """The 20-qubit, depth-20 random-layer circuit (780 gates)."""
    f = np.load(os.path.join(GOLDEN, "layers20_ref.npz"))
    n = 20
    g = f["gates"]
    want = oracle.run(n, g)
    with eng.StateVector(n) as s:
        got = s.apply(g).amplitudes()
        z, nrm = s.expect_z_all()
        p = s.probabilities()
    assert np.max(np.abs(got - want)) < TOL
    assert np.max(np.abs(got[f["sample_idx"]] - f["sample_amps"])) < TOL
    assert np.max(np.abs(z - f["z"])) < TOL
    assert np.max(np.abs(z - oracle.expect_z_all(n, want))) < TOL
    assert np.max(np.abs(p - oracle.probabilities(n, want))) < 1e-15
    assert abs(nrm - 1.0) < TOL


def test_reductions_vs_oracle(eng, oracle):
    rng = np.random.default_rng(6)
    for n in (1, 2, 3, 7, 11, 12, 13, 15):
        a = random_state(n, rng)
        a *= 1.0 + 0.1 * rng.uniform()   # unnormalised on purpose
        with eng.StateVector(n) as s:
            s.load(a)
            z, nrm = s.expect_z_all()
            assert np.max(np.abs(z - oracle.expect_z_all(n, a))) < TOL, n
            assert abs(nrm - oracle.norm_sq(n, a)) < TOL
            for q in range(n):
                assert abs(s.prob_one(q) - oracle.L.svo_probability_one(n, a.view(np.float64).ctypes.data_as(
                    __import__("ctypes").POINTER(__import__("ctypes").c_double)), q)) < TOL
                assert abs(s.expect_x(q) - oracle.expect_x(n, a, q)) < TOL
                assert abs(s.expect_y(q) - oracle.expect_y(n, a, q)) < TOL
            for q in range(n - 1):
                assert abs(s.correlation_zz(q, q + 1) - oracle.correlation_zz(n, a, q, q + 1)) < TOL
            b = random_state(n, rng)
            with eng.StateVector(n) as t:
                t.load(b)
                ip = s.inner(t)
            assert abs(ip - np.vdot(a, b)) < TOL
            s.normalize()
            assert abs(s.norm() - 1.0) < TOL


def test_reductions_are_bit_reproducible(eng):
    rng = np.random.default_rng(8)
    n = 18
    a = random_state(n, rng)
    with eng.StateVector(n) as s:
        s.load(a)
        z1, n1 = s.expect_z_all()
        z2, n2 = s.expect_z_all()
    assert _bytes_equal(z1, z2) and n1 == n2


def test_sampler_matches_collapse_rule(eng, oracle):
    rng = np.random.default_rng(9)
    for n in (1, 3, 8, 11, 12, 13, 16):
        a = random_state(n, rng)
        if n >= 3:
            a[: (1 << n) // 8] = 0.0    # a zero stretch
            a[-3:] = 0.0                 # trailing zeros: the fallback must skip them
        a /= np.linalg.norm(a)
        u = np.concatenate([rng.uniform(size=400), [0.0, 1e-300, 1.0 - 2 ** -53, 1.0, 2.0]])
        want = np.array([oracle.measure_all_outcome(n, a, x) for x in u], dtype=np.uint64)
        with eng.StateVector(n) as s:
            got = s.load(a).sample(u)
        assert np.array_equal(got, want), n


def test_job_shots_bit_exact_vs_reference(eng):
    shots = json.load(open(os.path.join(GOLDEN, "shots_ref.json")))
    rnd10 = np.load(os.path.join(GOLDEN, "rnd10_gates.npy"))
    circuits = {"bell": [(4, 0, -1, 0.0), (10, 1, 0, 0.0)],
                "ghz": [(4, 0, -1, 0.0), (10, 1, 0, 0.0), (10, 2, 1, 0.0)],
                "rnd10": [(int(g["type"]), int(g["target"]), int(g["control"]), float(g["theta"])) for g in rnd10]}
    circuits["rnd10_w1"] = circuits["rnd10"]
    eng.register_mi355x_backend()
    for backend in (None, "mi355x"):
        for name, gl in circuits.items():
            s = shots[name]
            j = eng.Job(s["n"]).set_num_shots(s["shots"]).set_num_workers(s["workers"]).set_rng_seed(s["seed"])
            for t, tg, c, th in gl:
                j.add_gate(t, tg, c, th if t in (7, 8, 9) else None)
            j.set_backend(backend)
            r = j.execute()
            assert r.outcomes == s["outcomes"], (backend, name)
            assert r.num_workers_used == s["used"]
            assert len(r.worker_seconds) == s["used"]
    q = shots["qgtl_rnd10"]
    c = eng.Circuit(10)
    for t, tg, ctl, th in circuits["rnd10"]:
        c.add_gate(t, tg, ctl, th if t in (7, 8, 9) else None)
    r = c.execute(num_shots=q["shots"], rng_seed=q["seed"], return_probabilities=True)
    assert r.outcomes.tolist() == q["outcomes"]
    assert np.max(np.abs(r.probabilities - np.array(q["probabilities"]))) < 1e-15


def test_reference_unit_cases(eng):
    """tests/unit/test_qgtl_backend.c and test_scheduler.c known answers."""
    c = eng.Circuit(2).add_gate(4, 0).add_gate(10, 1, 0)
    p = c.execute(return_probabilities=True).probabilities
    assert abs(p[0] - 0.5) < 1e-9 and abs(p[3] - 0.5) < 1e-9 and p[1] < 1e-9 and p[2] < 1e-9
    r = c.execute(num_shots=1024, rng_seed=0xdeadbeef)
    assert set(r.outcomes.tolist()) <= {0, 3}
    n00 = int(np.sum(r.outcomes == 0))
    assert 512 - 80 < n00 < 512 + 80
    g = eng.Circuit(2)
    for t, tg, ctl in [(4, 0, -1), (4, 1, -1), (12, 1, 0), (4, 0, -1), (4, 1, -1), (3, 0, -1), (3, 1, -1),
                       (12, 1, 0), (4, 0, -1), (4, 1, -1)]:
        g.add_gate(t, tg, ctl)
    assert g.execute(return_probabilities=True).probabilities[3] > 0.99
    ry = eng.Circuit(1).add_gate(8, 0, theta=np.pi / 2).execute(return_probabilities=True).probabilities
    assert abs(ry[0] - 0.5) < 1e-9
    j = eng.Job(3).add_gate(4, 0).add_gate(10, 1, 0).add_gate(10, 2, 1)
    r = j.set_num_shots(3000).set_num_workers(3).set_rng_seed(0xCAFE).execute()
    assert set(r.outcomes) <= {0, 7} and len(r.outcomes) == 3000


def test_execute_ex_outputs(eng, oracle):
    rng = np.random.default_rng(10)
    n = 14
    from moonlab_amd import workloads
    g = workloads.random_layer_circuit(n, 8, 77)
    j = eng.Job(n).set_num_shots(500).set_num_workers(3).set_rng_seed(0xABCDEF)
    for row in g:
        t = int(row["type"])
        j.add_gate(t, int(row["target"]), int(row["control"]), float(row["theta"]) if t in (7, 8, 9) else None)
    want = oracle.run(n, g)
    r = j.execute_ex(amplitudes=True, probabilities=True, expect_z=True)
    assert np.max(np.abs(r.amplitudes - want)) < TOL
    assert np.max(np.abs(r.expect_z - oracle.expect_z_all(n, want))) < TOL
    assert np.max(np.abs(r.probabilities - np.abs(want) ** 2)) < 1e-15
    assert r.outcomes.tolist() == oracle.job_shots(n, want, 500, 3, 0xABCDEF).tolist()
    r64 = j.execute_ex(precision=eng.C64, amplitudes=True, expect_z=True)
    fid = abs(np.vdot(want, r64.amplitudes)) ** 2
    assert fid > 1 - 1e-6
    assert np.max(np.abs(r64.expect_z - oracle.expect_z_all(n, want))) < 1e-5
    rf = j.execute_ex(amplitudes=True, fusion_max_qubits=4)
    assert np.max(np.abs(rf.amplitudes - want)) < TOL


def _dense_ref(a0, n, qs, u):
    k = len(qs)
    t = a0.reshape([2] * n)                      # axis i = qubit n-1-i
    axes = [n - 1 - q for q in qs[::-1]]         # local bit j = qs[j]
    moved = np.moveaxis(t, axes, list(range(k))).reshape(1 << k, -1)
    return np.moveaxis((u @ moved).reshape([2] * n), list(range(k)), axes).reshape(-1)


def test_dense_blocks_every_tile_width(eng):
    """The matrix-core block kernel at every tile width it picks (64 groups
    per tile on states just large enough, 128..512 groups on larger ones),
    both precisions, low / scattered / high qubit sets, against numpy."""
    import scipy.linalg
    rng = np.random.default_rng(13)
    for n in (9, 10, 11, 12, 13, 15):
        for k in (3, 4, 5):
            if n < k + 6:
                continue
            h = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
            u = scipy.linalg.expm(1j * (h + h.conj().T))
            for qs in (list(range(k)), sorted(int(q) for q in rng.permutation(n)[:k]), list(range(n - k, n))):
                a0 = random_state(n, rng)
                want = _dense_ref(a0, n, qs, u)
                with eng.StateVector(n) as s:
                    got = s.load(a0).apply_dense(qs, u).amplitudes()
                assert np.max(np.abs(got - want)) < TOL, (n, qs)
                with eng.StateVector(n, eng.C64) as s:
                    got = s.load(a0).apply_dense(qs, u).amplitudes()
                assert np.max(np.abs(got - want)) < 2e-6, (n, qs, "c64")


def test_fused_application_vs_unfused(eng, oracle):
    rng = np.random.default_rng(12)
    from moonlab_amd import workloads
    for n, k in ((6, 3), (10, 4), (14, 5), (16, 5), (13, 2), (9, 1), (12, 3), (15, 4)):
        g = workloads.su2_cz_circuit(n, 6, 1000 + n)
        want = oracle.run(n, g)
        with eng.StateVector(n) as s:
            got = s.apply(g, fuse=k).amplitudes()
        assert np.max(np.abs(got - want)) < TOL, (n, k)
        with eng.StateVector(n, eng.C64) as s:
            got = s.apply(g, fuse=k).amplitudes()
        assert abs(np.vdot(want, got)) ** 2 > 1 - 1e-6, (n, k)
    # a general dense block at scattered qubits
    import scipy.linalg
    n = 11
    for qs in ([0, 1, 2, 3, 4], [1, 4, 6, 9, 10], [0, 5, 10], [2, 3, 7, 8], [6, 7, 8, 9, 10], [10, 8, 6][::-1]):
        k = len(qs)
        h = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        u = scipy.linalg.expm(1j * (h + h.conj().T))
        a0 = random_state(n, rng)
        # reference: brute force in numpy on the reshaped state
        t = a0.reshape([2] * n)  # axis i = qubit n-1-i
        axes = [n - 1 - q for q in qs[::-1]]
        moved = np.moveaxis(t, axes, list(range(k))).reshape(1 << k, -1)
        want = np.moveaxis((u @ moved).reshape([2] * n), list(range(k)), axes).reshape(-1)
        with eng.StateVector(n) as s:
            got = s.load(a0).apply_dense(qs, u).amplitudes()
        assert np.max(np.abs(got - want)) < TOL, qs
        with eng.StateVector(n, eng.C64) as s:   # FP32 matrix cores
            got = s.load(a0).apply_dense(qs, u).amplitudes()
        assert np.max(np.abs(got - want)) < 2e-6, qs
    # VALU path: states too small for a 2^(k+6) tile
    n = 8
    h = rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32))
    u = scipy.linalg.expm(1j * (h + h.conj().T))
    a0 = random_state(n, rng)
    t = a0.reshape([2] * n)
    qs = [0, 2, 3, 5, 7]
    axes = [n - 1 - q for q in qs[::-1]]
    moved = np.moveaxis(t, axes, list(range(5))).reshape(32, -1)
    want = np.moveaxis((u @ moved).reshape([2] * n), list(range(5)), axes).reshape(-1)
    with eng.StateVector(n) as s:
        got = s.load(a0).apply_dense(qs, u).amplitudes()
    assert np.max(np.abs(got - want)) < TOL


def test_hea_tfim_energy_against_reference(eng):
    rows = json.load(open(os.path.join(GOLDEN, "hea_tfim_ref.json")))
    for r in rows:
        n, L, h = r["n"], r["layers"], r["h"]
        ang = np.array(r["angles"]).reshape(L, n, 2)
        gl = []
        for l in range(L):
            for q in range(n):
                gl.append((8, q, -1, -1, float(ang[l, q, 0])))
                gl.append((9, q, -1, -1, float(ang[l, q, 1])))
            for q in range(n - 1):
                gl.append((12, q + 1, q))
        with eng.StateVector(n) as s:
            s.apply(gl)
            e = -sum(s.correlation_zz(q, q + 1) for q in range(n - 1)) - h * sum(s.expect_x(q) for q in range(n))
        assert abs(e - r["energy"]) < TOL


def test_large_state_properties(eng):
    """At 26 qubits (1 GiB c128): a random-layer circuit followed by its
    inverse returns |0..0>; the norm stays 1 throughout."""
    from moonlab_amd import workloads
    n = 26
    g = workloads.random_layer_circuit(n, 4, 5, h_layer=True)
    inv = []
    for row in g[::-1]:
        t = int(row["type"])
        th = -float(row["theta"]) if t in (7, 8, 9) else 0.0
        inv.append((t, int(row["target"]), int(row["control"]), -1, th))
    with eng.StateVector(n) as s:
        s.apply(g)
        _, nrm = s.expect_z_all()
        assert abs(nrm - 1.0) < 1e-12
        s.apply(inv)
        head = s.read(0, 4)
        z, nrm = s.expect_z_all()
    assert abs(head[0] - 1.0) < 1e-12 and np.max(np.abs(head[1:])) < 1e-12
    assert np.max(np.abs(z - 1.0)) < 1e-12


def test_cache_blocked_application_vs_oracle(eng, oracle):
    """Cache-blocked passes (12-qubit LDS tiles) over the whole gate table."""
    rng = np.random.default_rng(41)
    for n in (1, 3, 9, 12, 13, 15, 18):
        gl = random_gate_list(n, 150, rng)
        want = oracle.run(n, gl)
        with eng.StateVector(n) as s:
            got = s.apply(gl, blocked=True).amplitudes()
        assert np.max(np.abs(got - want)) < TOL, n
        with eng.StateVector(n, eng.C64) as s:
            got = s.apply(gl, blocked=True).amplitudes()
        assert abs(np.vdot(want, got)) ** 2 > 1 - 1e-6, n
    from moonlab_amd import workloads
    n = 20
    g = workloads.random_layer_circuit(n, 20, 20261202)
    want = oracle.run(n, g)
    r = None
    j = eng.Job(n).set_num_shots(300).set_rng_seed(77)
    for row in g:
        t = int(row["type"])
        j.add_gate(t, int(row["target"]), int(row["control"]), float(row["theta"]) if t in (7, 8, 9) else None)
    r = j.execute_ex(amplitudes=True, expect_z=True, blocked=True)
    assert np.max(np.abs(r.amplitudes - want)) < TOL
    assert r.outcomes.tolist() == oracle.job_shots(n, want, 300, 1, 77).tolist()


@pytest.mark.parametrize("engine,tile_bits", [("register", 12), ("register", 13), ("register16", 12),
                                               ("register8", 13), ("lds", 11), ("lds", 12)])
def test_blocked_engines_vs_oracle(eng, oracle, engine, tile_bits):
    """Each blocked engine and tile (register phases of 8 or 16 values per
    thread on 12- and 13-qubit tiles, LDS ops on 11 and 12) over the whole
    gate table: complex128 within
    1e-12 of the oracle, complex64 by fidelity; 24 qubits gives every
    workgroup several tiles."""
    rng = np.random.default_rng(4100 + tile_bits)
    try:
        eng.set_blocked_engine(engine)
        eng.set_blocked_tuning(tile_bits, 0)
        for n, ng in ((12, 200), (13, 200), (14, 200), (17, 160), (24, 60)):
            gl = random_gate_list(n, ng, rng)
            want = oracle.run(n, gl)
            with eng.StateVector(n) as s:
                got = s.apply(gl, blocked=True).amplitudes()
            assert np.max(np.abs(got - want)) < TOL, (engine, tile_bits, n)
            with eng.StateVector(n, eng.C64) as s:
                got = s.apply(gl, blocked=True).amplitudes()
            assert abs(np.vdot(want, got)) ** 2 > 1 - 1e-6, (engine, tile_bits, n)
    finally:
        eng.set_blocked_engine("register")
        eng.set_blocked_tuning(-1, 512)


@pytest.mark.gpu
def test_blocked_random_layers_vs_oracle(eng, oracle):
    """The serving stream's circuit family (rotation per qubit + CNOT
    brickwork) on the default register-phase engine: its phases are mostly
    paired rotations (ARM_ROT2) and CNOT exchanges.  complex128 within 1e-12
    of the oracle.  complex64: the direction of the state by fidelity after
    normalising, and the norm separately -- the gate matrices rounded to
    float32 are unitary only to ~1e-7 (c^2 + s^2 != 1, 2 * fl(1/sqrt 2)^2 < 1),
    so the norm drifts by ~1e-6 over a few hundred gates, by design of the
    precision, while the direction stays within ~1e-12."""
    from moonlab_amd import workloads
    for n, depth in ((14, 12), (20, 8), (23, 6)):
        gl = workloads.random_layer_circuit(n, depth, 700 + n, h_layer=True)
        want = oracle.run(n, gl)
        with eng.StateVector(n) as s:
            got = s.apply(gl, blocked=True).amplitudes()
        assert np.max(np.abs(got - want)) < TOL, n
        with eng.StateVector(n, eng.C64) as s:
            got = s.apply(gl, blocked=True).amplitudes()
        nrm = np.vdot(got, got).real
        assert abs(nrm - 1) < 2e-5, (n, nrm)
        assert abs(np.vdot(want, got)) ** 2 / nrm > 1 - 1e-8, n


@pytest.mark.gpu
def test_plugin_runs_in_reference_scheduler(eng):
    """Registered into the reference's own scheduler (compiled reference C),
    a job set to "mi355x" returns the same outcomes as the reference's
    "simulator" backend for the same job and seed (scheduler.c:201-236)."""
    import json
    import subprocess
    import sys
    from conftest import ROOT
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libmoonlab_ref.so")):
        pytest.skip("oracle/_ref not built")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "plugin_host.py"), "run"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["rc"] == 0 and out["found"]
    assert out["simulator"]["rc"] == 0 and out["mi355x"]["rc"] == 0
    assert len(out["mi355x"]["outcomes"]) == 3000
    assert out["mi355x"]["outcomes"] == out["simulator"]["outcomes"]


def test_concurrent_jobs_match_sequential(eng):
    """moonlab_scheduler_run from many threads at once (the reference's
    tests/concurrency/conc_scheduler.c pattern): each job's outcomes equal
    its outcomes run alone, for both precisions of execute_ex too."""
    from concurrent.futures import ThreadPoolExecutor
    from moonlab_amd import workloads
    jobs = []
    for i in range(12):
        n = 8 + i % 9
        j = eng.Job(n).set_num_shots(300).set_num_workers(1 + i % 3).set_rng_seed(7000 + i)
        for row in workloads.random_layer_circuit(n, 6, 900 + i, h_layer=True):
            t = int(row["type"])
            j.add_gate(t, int(row["target"]), int(row["control"]), float(row["theta"]) if t in (7, 8, 9) else None)
        jobs.append(j)
    alone = [j.execute().outcomes for j in jobs]
    alone_z = [j.execute_ex(precision=eng.C64, expect_z=True).expect_z for j in jobs]
    with ThreadPoolExecutor(8) as ex:
        mixed = list(ex.map(lambda j: j.execute().outcomes, jobs * 3))
        mixed_z = list(ex.map(lambda j: j.execute_ex(precision=eng.C64, expect_z=True).expect_z, jobs * 2))
    assert mixed == alone * 3
    for a, b in zip(mixed_z, alone_z * 2):
        assert _bytes_equal(a, b)


def _inverse(g):
    """The random-layer circuit's inverse: reversed, rotation angles negated
    (CNOT is its own inverse)."""
    out = []
    for row in g[::-1]:
        t = int(row["type"])
        th = -float(row["theta"]) if t in (7, 8, 9) else 0.0
        out.append((t, int(row["target"]), int(row["control"]), -1, th))
    return out


@pytest.mark.parametrize("n,prec,path,tol", [
    (30, "c128", "gates", 1e-12),     # the 16 GiB sweep state, one kernel per gate
    (30, "c128", "blocked", 1e-12),   # register-phase passes
    (30, "c64", "blocked", 2e-5),
    (32, "c64", "fused", 2e-5),       # 32 GiB, FP32 MFMA blocks
    (33, "c128", "fused", 1e-12),     # 128 GiB, FP64 MFMA blocks
    (33, "c128", "blocked", 1e-12),
])
def test_full_size_round_trips(eng, n, prec, path, tol):
    """At the benchmark sizes, through size-independent properties: the norm
    stays 1 after a random-layer circuit, and the circuit followed by its
    inverse returns |0..0> (amplitude 0 within tol of 1, every <Z_q> = 1)."""
    from moonlab_amd import workloads
    g = workloads.random_layer_circuit(n, 2, 17, h_layer=True)
    p = eng.C128 if prec == "c128" else eng.C64
    kw = {"blocked": True} if path == "blocked" else {"fuse": 5} if path == "fused" else {}
    print(f"round trip n={n} {prec} {path}", flush=True)
    if n >= 32:   # room for the state: drop the job paths' parked states
        eng.lib().moonlab_mi355x_set_state_cache_bytes(0)
        eng.lib().moonlab_mi355x_set_state_cache_bytes(64 << 30)
    with eng.StateVector(n, p) as s:
        s.apply(g, **kw)
        _, nrm = s.expect_z_all()
        assert abs(nrm - 1.0) < (1e-10 if prec == "c128" else 1e-5)
        s.apply(_inverse(g), **kw)
        head = s.read(0, 2)
        z, nrm = s.expect_z_all()
        shots = s.sample(np.linspace(0.0, 0.999, 64)) if prec == "c128" else np.zeros(1)
    assert abs(head[0] - 1.0) < tol and abs(head[1]) < tol
    assert not np.any(shots)   # |0..0>: the sampler's CDF over the whole state puts every draw on 0
    assert np.max(np.abs(z - 1.0)) < 4 * tol


@pytest.mark.parametrize("n", [33, 34])
def test_gate_families_at_33_and_34_qubits(eng, n):
    """Every one-shot gate kernel on complex128 states of 128 and 256 GiB (34
    qubits is the largest one MI355X holds; launches of 2^32 work-items and
    more): a product of one gate of each family followed by its inverse
    returns |0..0>."""
    fwd = [(4, 0), (4, 32), (4, 17), (8, 32, -1, -1, 0.7), (7, 1, -1, -1, 0.4),
           (9, 32, -1, -1, 1.1),                 # diagonal on every element
           (5, 32), (6, 2),                      # phase on the |1> half
           (10, 0, 32), (10, 32, 0),             # controlled pairs
           (12, 17, 32),                         # CZ: masked elements
           (13, 0, 32),                          # SWAP
           (18, 1, 32, -1, 0.3),                 # CPHASE
           (20, 32, 17, -1, 0.9),                # CRY
           (23, 5, 32, 0),                       # Toffoli
           (24, 32, 1, 2),                       # Fredkin
           (17, 32, -1, -1, 0.3, 0.2, 0.1)]      # U3
    inv = {5: 14, 6: 15}                          # S -> S^dagger, T -> T^dagger
    back = []
    for g in fwd[::-1]:
        t = g[0]
        if t in (7, 8, 9, 18, 20):
            back.append((t, g[1], g[2], g[3], -g[4]))
        elif t == 17:   # U3(theta, phi, lam)^-1 = U3(-theta, -lam, -phi)
            back.append((17, g[1], -1, -1, -g[4], -g[6], -g[5]))
        else:
            back.append((inv.get(t, t),) + tuple(g[1:]))
    eng.lib().moonlab_mi355x_set_state_cache_bytes(0)
    eng.lib().moonlab_mi355x_set_state_cache_bytes(64 << 30)
    with eng.StateVector(n) as s:
        s.apply(fwd)
        _, nrm = s.expect_z_all()
        assert abs(nrm - 1.0) < 1e-12
        s.apply(back)
        head = s.read(0, 2)
        z, _ = s.expect_z_all()
    assert abs(head[0] - 1.0) < 1e-12 and abs(head[1]) < 1e-12
    assert np.max(np.abs(z - 1.0)) < 1e-12


def _state_of(s):
    return s.read(0, 1 << s.n).astype(np.complex128)


@pytest.mark.parametrize("n", [5, 14, 20])
def test_collapsing_measurements_vs_oracle(eng, oracle, n):
    """measure_qubit / measure_partial / measure_all against the oracle's
    restatement of measurement_single_qubit / _partial / _all_qubits
    (complex128): the same outcome, the same zeroed amplitudes, the kept ones
    within 1e-12; normalized and unnormalized states, draws at and past the
    total probability, several qubit orders."""
    rng = np.random.default_rng(40 + n)
    base = random_state(n, rng)
    qsets = [[0], [n - 1], [2, 0], [n - 1, 1, 3], [4, n - 2, 0, 2, 1]]
    for a in (base, base * 1.7):
        for u in (0.0, 0.3, 0.77, 1.0 - 2 ** -53, 1.0, 1.5):
            for q in (0, n // 2, n - 1):
                want_a = a.copy()
                want = oracle.measure_qubit(n, want_a, q, u)
                with eng.StateVector(n) as s:
                    got = s.load(a).measure_qubit(q, u)
                    got_a = _state_of(s)
                assert got == want, (q, u)
                assert np.array_equal(got_a == 0, want_a == 0) and np.max(np.abs(got_a - want_a)) < 1e-12
            for qs in qsets:
                want_a = a.copy()
                want = oracle.measure_partial(n, want_a, qs, u)
                with eng.StateVector(n) as s:
                    got = s.load(a).measure_partial(qs, u)
                    got_a = _state_of(s)
                assert got == want, (qs, u)
                assert np.array_equal(got_a == 0, want_a == 0) and np.max(np.abs(got_a - want_a)) < 1e-12
            want_a = a.copy()
            want = oracle.measure_all(n, want_a, u)
            with eng.StateVector(n) as s:
                got = s.load(a).measure_all(u)
                got_a = _state_of(s)
            assert got == want and np.array_equal(got_a, want_a), u


def test_collapsing_measurement_edge_cases(eng):
    """A zero-weight single-qubit outcome gives -1 and leaves the state zeroed
    (measurement.c:219-221); repeated and out-of-range qubits are refused;
    complex64 states collapse within 1e-6 and renormalize to 1."""
    rng = np.random.default_rng(5)
    n = 9
    a = random_state(n, rng)
    a[np.arange(1 << n) & 4 == 0] = 0.0     # qubit 2 is surely 1
    a /= np.linalg.norm(a)
    with eng.StateVector(n) as s:
        assert s.load(a).measure_qubit(2, 1.5) == -1
        assert not np.any(_state_of(s))
        with pytest.raises(eng.EngineError):
            s.measure_partial([1, 1], 0.5)
        with pytest.raises(eng.EngineError):
            s.measure_qubit(n, 0.5)
    b = random_state(n, rng)
    with eng.StateVector(n, eng.C64) as s:
        r = s.load(b).measure_partial([3, 0], 0.6)
        got = _state_of(s)
        _, nrm = s.expect_z_all()
    keep = ((np.arange(1 << n) >> 3) & 1) | (((np.arange(1 << n) >> 0) & 1) << 1)
    assert np.all(got[keep != r] == 0) and abs(nrm - 1.0) < 1e-5
