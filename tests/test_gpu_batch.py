"""Persistent batch engine vs the CPU oracle and the reference's goldens.

Tolerances: complex128 observables and energies within 1e-12 absolute;
complex64 within 2e-5; shot outcomes exact (same draws, same outcome rule);
a circuit's results are bit-identical alone and inside any mix.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, random_gate_list

pytestmark = pytest.mark.gpu

TOL = 1e-12


def hea_gates(n, L, ang):
    g = []
    for l in range(L):
        for q in range(n):
            g.append((8, q, -1, -1, float(ang[l, q, 0])))
            g.append((9, q, -1, -1, float(ang[l, q, 1])))
        for q in range(n - 1):
            g.append((12, q + 1, q))
    return g


def test_hea_tfim_golden(eng):
    rows = json.load(open(os.path.join(GOLDEN, "hea_tfim_ref.json")))
    b = eng.Batch()
    inst = [(r["n"], r["layers"], np.array(r["angles"]).reshape(r["layers"], r["n"], 2), r["h"]) for r in rows]
    e = b.hea_tfim(inst)
    want = np.array([r["energy"] for r in rows])
    assert np.max(np.abs(e - want)) < TOL


def test_hea_pack_matches_instance_list(eng):
    """A HeaPack gives the energies and gradients of the same instances
    passed as a list, bit for bit, and angles written into it in place take
    effect on the next call."""
    from moonlab_amd import workloads
    inst = workloads.hea_batch(48, 31)
    rng = np.random.default_rng(32)
    inst2 = [(n, L, rng.uniform(0, 2 * np.pi, (L, n, 2)), h) for (n, L, _, h) in inst]
    b = eng.Batch()
    pack = eng.HeaPack(inst)
    assert np.array_equal(b.hea_tfim(pack), b.hea_tfim(inst))
    pack.angles[:] = np.concatenate([np.asarray(a, dtype=np.float64).reshape(-1) for (_, _, a, _) in inst2])
    assert np.array_equal(b.hea_tfim(pack), b.hea_tfim(inst2))
    small = inst[:6]
    e1, g1 = b.hea_tfim_gradients(eng.HeaPack(small))
    e2, g2 = b.hea_tfim_gradients(small)
    assert np.array_equal(e1, e2) and all(np.array_equal(x, y) for x, y in zip(g1, g2))


def test_hea_tfim_cache_blocked_vs_oracle(eng, oracle):
    """n = 13..18 runs from HBM slots through LDS tiles; n <= 12 stays in LDS."""
    rng = np.random.default_rng(31)
    inst = []
    for n in (8, 12, 13, 14, 15, 16, 17, 18):
        L = int(rng.integers(4, 13))
        inst.append((n, L, rng.uniform(0, 2 * np.pi, size=(L, n, 2)), float(rng.uniform(0.5, 1.5))))
    b = eng.Batch()
    e = b.hea_tfim(inst)
    want = np.array([oracle.hea_tfim_energy(n, L, ang, h) for n, L, ang, h in inst])
    assert np.max(np.abs(e - want)) < TOL, np.abs(e - want)
    for tb in (10, 11):
        b.set_config(tile_bits=tb)
        e2 = b.hea_tfim(inst)
        assert np.max(np.abs(e2 - want)) < TOL, tb
    b.set_config(workgroups_per_cu=1, tile_bits=12)
    assert np.max(np.abs(b.hea_tfim(inst) - want)) < TOL


def _circuits(rng, sizes, ngates=60, shots=64, prec=0):
    out = []
    for n in sizes:
        gl = random_gate_list(n, ngates, rng)
        out.append({"n": n, "gates": gl, "precision": prec, "observables": 7,
                    "draws": rng.uniform(size=shots)})
    return out


def test_generic_batch_vs_oracle(eng, oracle):
    rng = np.random.default_rng(32)
    circuits = _circuits(rng, [1, 2, 3, 5, 8, 11, 12, 13, 14, 16], ngates=80, shots=200)
    # edge draws for the shot rule
    circuits[5]["draws"] = np.concatenate([circuits[5]["draws"], [0.0, 1.0 - 2 ** -53, 1.0, 3.0]])
    res = eng.Batch().run(circuits)
    for c, r in zip(circuits, res):
        n = c["n"]
        a = oracle.run(n, c["gates"])
        assert abs(r.norm - oracle.norm_sq(n, a)) < TOL
        assert np.max(np.abs(r.expect_z - oracle.expect_z_all(n, a))) < TOL, n
        if n > 1:
            zz = [oracle.correlation_zz(n, a, q, q + 1) for q in range(n - 1)]
            assert np.max(np.abs(r.zz_chain - zz)) < TOL, n
        xs = [oracle.expect_x(n, a, q) for q in range(n)]
        assert np.max(np.abs(r.expect_x - xs)) < TOL, n
        want = [oracle.measure_all_outcome(n, a, u) for u in c["draws"]]
        assert r.outcomes.tolist() == want, n


def test_results_identical_alone_and_in_mix(eng):
    rng = np.random.default_rng(33)
    mix = _circuits(rng, [4, 9, 12, 14, 6, 15, 10], ngates=50, shots=100)
    mix += _circuits(rng, [7, 13, 11], ngates=40, shots=50, prec=1)
    b = eng.Batch()
    together = b.run(mix)
    for c, r in zip(mix, together):
        alone = b.run([c])[0]
        for f in ("expect_z", "zz_chain", "expect_x", "outcomes"):
            x, y = getattr(r, f), getattr(alone, f)
            assert x.tobytes() == y.tobytes(), (c["n"], f)
        assert r.norm == alone.norm


def test_complex64_batch(eng, oracle):
    rng = np.random.default_rng(34)
    circuits = _circuits(rng, [6, 10, 12, 14], ngates=60, shots=0, prec=1)
    for c in circuits:
        c["draws"] = np.zeros(0)
    res = eng.Batch().run(circuits)
    for c, r in zip(circuits, res):
        a = oracle.run(c["n"], c["gates"])
        assert np.max(np.abs(r.expect_z - oracle.expect_z_all(c["n"], a))) < 2e-5
        assert abs(r.norm - 1.0) < 2e-5


def test_autotune_keeps_results(eng, oracle):
    rng = np.random.default_rng(35)
    inst = []
    for _ in range(64):
        n = int(rng.integers(8, 15))
        L = int(rng.integers(4, 9))
        inst.append((n, L, rng.uniform(0, 2 * np.pi, size=(L, n, 2)), float(rng.uniform(0.5, 1.5))))
    circuits = [{"n": n, "gates": hea_gates(n, L, ang), "observables": 6} for n, L, ang, h in inst]
    b = eng.Batch()
    res = b.run(circuits, autotune=True)
    assert b.last_stats["tile_bits"] in (10, 11, 12)
    for (n, L, ang, h), r in zip(inst[:8], res[:8]):
        e = -np.sum(r.zz_chain) - h * np.sum(r.expect_x)
        assert abs(e - oracle.hea_tfim_energy(n, L, ang, h)) < TOL


def test_hea_tfim_gradients_vs_oracle(eng, oracle):
    """Batched parameter-shift gradients (one launch for every shifted
    circuit) against the oracle's shift rule, complex128 and complex64."""
    rng = np.random.default_rng(41)
    inst = []
    for _ in range(6):
        n, L = int(rng.integers(2, 9)), int(rng.integers(1, 4))
        inst.append((n, L, rng.uniform(0, 2 * np.pi, size=(L, n, 2)), float(rng.uniform(0.5, 1.5))))
    b = eng.Batch()
    try:
        e, g = b.hea_tfim_gradients(inst)
        e64, g64 = b.hea_tfim_gradients(inst, precision=eng.C64)
    finally:
        b.free()
    for i, (n, L, ang, h) in enumerate(inst):
        assert abs(e[i] - oracle.hea_tfim_energy(n, L, ang, h)) < 1e-12
        want = oracle.hea_tfim_gradient(n, L, ang, h)
        assert np.max(np.abs(g[i] - want)) < 1e-12
        assert np.max(np.abs(g64[i] - want)) < 1e-5


def _bytes_equal(a, b):
    return np.asarray(a).tobytes() == np.asarray(b).tobytes()


def test_graph_replays_match_direct_launches(eng, oracle):
    """A VQE-style loop (same ansatz shapes, new angles each call) captures
    the batch step on its second call and replays it after; every call,
    with unrelated eager GPU work and synchronisation in between, gives the
    same bits as the same batch launched directly, and matches the oracle."""
    from moonlab_amd import workloads
    rng = np.random.default_rng(51)
    shapes = [(int(rng.integers(4, 15)), int(rng.integers(1, 5))) for _ in range(24)]
    g_on, g_off = eng.Batch(), eng.Batch()
    g_off.set_graphs(False)
    replayed = []
    try:
        for it in range(5):
            inst = [(n, L, rng.uniform(0, 2 * np.pi, size=(L, n, 2)), float(rng.uniform(0.5, 1.5)))
                    for n, L in shapes]
            e_on = g_on.hea_tfim(inst)
            replayed.append(g_on.last_stats["graph_replayed"])
            e_off = g_off.hea_tfim(inst)
            assert _bytes_equal(e_on, e_off), it
            for i in (0, 7, 23):
                n, L, ang, h = inst[i]
                assert abs(e_on[i] - oracle.hea_tfim_energy(n, L, ang, h)) < 1e-12
            with eng.StateVector(14) as s:     # unrelated eager work between replays
                s.apply(workloads.random_layer_circuit(14, 3, it)).synchronize()
    finally:
        g_on.free()
        g_off.free()
    assert replayed == [False, False, True, True, True]


def test_large_batch_matches_small_batches(eng, oracle):
    """160 mixed circuits in one run (host planning split over threads) give
    the same bits as the same circuits in runs of 40 (planned on one thread),
    and match the oracle."""
    rng = np.random.default_rng(36)
    sizes = [int(x) for x in rng.integers(1, 17, size=160)]
    circuits = _circuits(rng, sizes, ngates=40, shots=20)
    b = eng.Batch()
    try:
        whole = b.run(circuits)
        parts = []
        for k in range(0, 160, 40):
            parts += b.run(circuits[k:k + 40])
    finally:
        b.free()
    for c, x, y in zip(circuits, whole, parts):
        for f in ("expect_z", "zz_chain", "expect_x", "outcomes"):
            assert getattr(x, f).tobytes() == getattr(y, f).tobytes(), (c["n"], f)
    for c, r in list(zip(circuits, whole))[::16]:
        a = oracle.run(c["n"], c["gates"])
        assert np.max(np.abs(r.expect_z - oracle.expect_z_all(c["n"], a))) < TOL


def test_small_tiles_reorder_full_gate_table(eng, oracle):
    """10-qubit tiles on 11..16-qubit circuits of every gate code: many
    passes, ops moved across pass boundaries past ops on other qubits."""
    rng = np.random.default_rng(37)
    circuits = _circuits(rng, [11, 12, 13, 14, 15, 16], ngates=120, shots=50)
    b = eng.Batch()
    try:
        b.set_config(tile_bits=10)
        res = b.run(circuits)
    finally:
        b.free()
    for c, r in zip(circuits, res):
        a = oracle.run(c["n"], c["gates"])
        assert abs(r.norm - 1.0) < TOL
        assert np.max(np.abs(r.expect_z - oracle.expect_z_all(c["n"], a))) < TOL, c["n"]
        xs = [oracle.expect_x(c["n"], a, q) for q in range(c["n"])]
        assert np.max(np.abs(r.expect_x - xs)) < TOL, c["n"]
        assert r.outcomes.tolist() == [oracle.measure_all_outcome(c["n"], a, u) for u in c["draws"]]
