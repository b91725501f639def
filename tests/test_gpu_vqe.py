"""Variational workloads on the batch engine: Pauli-sum observables in the
persistent launch, energies and parameter-shift gradients over Pauli-sum
Hamiltonians, the quantum geometric tensor (metric and Berry curvature) and
natural-gradient directions -- against the oracle (oracle/sv_oracle.c,
pinned to the reference's vqe_compute_* in tests/test_oracle_cpu.py).

Tolerances: <P_k>, energies, metric and curvature within 1e-12 (FP64; the
device sums run in another order than the oracle's sequential loops);
gradients within 1e-11 (differences of two energies); natural-gradient
directions within 1e-9 relative (a solve of a regularized metric whose
condition number reaches ~1e3)."""
import numpy as np
import pytest

from moonlab_amd import workloads

pytestmark = pytest.mark.gpu


def _hamiltonian(n, rng, count):
    terms = [(float(rng.uniform(-1, 1)), "".join(rng.choice(list("IXYZ"), size=n))) for _ in range(count)]
    # strings whose X/Y bits reach the top qubits (partners in other tiles for n > 12)
    terms.append((0.3, "X" * n))
    terms.append((-0.2, "Y" + "I" * (n - 2) + "Y" if n > 1 else "Y"))
    terms.append((0.4, "I" * n))
    return terms


def _instance(n, L, rng, entangler, terms=True):
    return {"n": n, "layers": L, "angles": rng.uniform(0, 2 * np.pi, 2 * n * L),
            "terms": _hamiltonian(n, rng, 5) if terms else [], "entangler": entangler,
            "hf": int(rng.integers(0, 1 << n)), "constant": float(rng.uniform(-1, 1))}


def test_batch_pauli_terms_vs_oracle(eng, oracle):
    """<P_k> of random strings on random circuits, resident (n <= 12) and
    cache-blocked (n 13..16) states, both precisions; the weighted sum; and
    each circuit's values bit-identical alone and in the mix."""
    rng = np.random.default_rng(81)
    circuits = []
    for i, n in enumerate((3, 7, 12, 13, 15, 16)):
        g = workloads.random_layer_circuit(n, 4, 100 + i, h_layer=True)
        circuits.append({"n": n, "gates": g, "precision": eng.C128 if i % 2 == 0 else eng.C64,
                         "observables": eng.OBS_Z, "pauli": _hamiltonian(n, rng, 9)})
    b = eng.Batch()
    mixed = b.run(circuits)
    for c, r in zip(circuits, mixed):
        n = c["n"]
        a = oracle.run(n, c["gates"])
        want = np.array([oracle.expect_pauli(n, a, s) for _, s in c["pauli"]])
        tol = 1e-12 if c["precision"] == eng.C128 else 2e-5
        assert np.max(np.abs(r.pauli_expect - want)) < tol, n
        assert abs(r.pauli_energy - sum(w * p for (w, _), p in zip(c["pauli"], r.pauli_expect))) < 1e-13
        alone = b.run([c])[0]
        assert alone.pauli_expect.tobytes() == r.pauli_expect.tobytes()
    with pytest.raises(eng.SchedulerError):
        b.run([{"n": 3, "gates": [], "pauli": [(1.0, "XQZ")]}])
    with pytest.raises(eng.SchedulerError):
        b.run([{"n": 3, "gates": [], "pauli": [(1.0, "XZ")]}])


@pytest.mark.parametrize("entangler", [1, 0])
def test_vqe_energies_and_gradients_vs_oracle(eng, oracle, entangler):
    rng = np.random.default_rng(83 + entangler)
    inst = [_instance(n, L, rng, entangler) for n, L in ((1, 1), (2, 3), (5, 2), (8, 1), (13, 1))]
    b = eng.Batch()
    e = b.vqe_energies(inst)
    e2, grads = b.vqe_gradients(inst)
    for v, ei, e2i, g in zip(inst, e, e2, grads):
        want = oracle.vqe_energy(v["n"], v["layers"], v["angles"], v["terms"], entangler, v["hf"], v["constant"])
        assert abs(ei - want) < 1e-12
        assert e2i == ei
        gw = oracle.vqe_gradient(v["n"], v["layers"], v["angles"], v["terms"], entangler, v["hf"])
        assert np.max(np.abs(g - gw)) < 1e-11


def test_vqe_qgt_and_berry_vs_oracle(eng, oracle):
    rng = np.random.default_rng(89)
    inst = [_instance(n, L, rng, ent, terms=False) for n, L, ent in
            ((1, 1, 1), (3, 2, 1), (6, 2, 0), (9, 1, 1), (13, 1, 0))]
    b = eng.Batch()
    got = b.vqe_qgt(inst)
    for v, (g, f) in zip(inst, got):
        gw, fw = oracle.vqe_qgt(v["n"], v["layers"], v["angles"], v["entangler"], v["hf"])
        assert np.max(np.abs(g - gw)) < 1e-12, v["n"]
        assert np.max(np.abs(f - fw)) < 1e-12, v["n"]
        assert np.array_equal(g, g.T) and np.array_equal(f, -f.T)


def test_vqe_natural_gradient_vs_oracle(eng, oracle):
    rng = np.random.default_rng(97)
    inst = [_instance(n, L, rng, 1) for n, L in ((2, 2), (4, 2), (7, 1))]
    b = eng.Batch()
    _, grads = b.vqe_gradients(inst)
    dirs = b.vqe_natural_gradient(inst, grads, 1e-3)
    for v, g, (d, status) in zip(inst, grads, dirs):
        assert status == 0
        gw, _ = oracle.vqe_qgt(v["n"], v["layers"], v["angles"], 1, v["hf"])
        want = oracle.solve(gw + 1e-3 * np.eye(gw.shape[0]), g)
        assert np.max(np.abs(d - want)) <= 1e-9 * max(1.0, np.max(np.abs(want)))
    # RY(0), RZ(0) on |0>: the RZ generator Z leaves |0> alone, so the metric
    # is exactly diag(1/4, 0); without regularization it is singular: status
    # -1 and the gradient handed back
    one = [{"n": 1, "layers": 1, "angles": [0.0, 0.0], "terms": [(1.0, "X")], "entangler": 1}]
    (d, status), = b.vqe_natural_gradient(one, [np.array([0.7, -0.1])], 0.0)
    assert status == -1 and d.tolist() == [0.7, -0.1]
