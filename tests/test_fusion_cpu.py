"""Fusion planner and FP64 block composition, on the host (no device).

The engine's plan (moonlab_hip_fusion_plan) is executed here block by block
with numpy and compared with the oracle's gate-by-gate run; the reference's
own fusion test asks fused == unfused within 1e-10 (tests/unit/test_fusion.c:
143-177), this one asks 1e-12.
"""
import numpy as np
import pytest

from conftest import random_gate_list, random_state


def apply_block(a, n, qubits, m):
    k = len(qubits)
    t = a.reshape([2] * n)
    axes = [n - 1 - q for q in qubits[::-1]]          # local bit j = qubits[j]
    moved = np.moveaxis(t, axes, list(range(k))).reshape(1 << k, -1)
    return np.moveaxis((m @ moved).reshape([2] * n), list(range(k)), axes).reshape(-1)


def gate_width(row):
    t = int(row[0])
    return 3 if t in (23, 24) else (2 if t in (10, 11, 12, 13) or t >= 18 else 1)


@pytest.mark.parametrize("kmax", [1, 2, 3, 4, 5])
def test_plan_reproduces_circuit(kmax, oracle):
    import moonlab_amd as m
    rng = np.random.default_rng(100 + kmax)
    for n in (3, 5, 7, 9):
        gl = random_gate_list(n, 90, rng)
        a0 = random_state(n, rng)
        want = oracle.apply(n, a0.copy(), gl)
        blocks = m.fusion_plan(gl, kmax)
        assert sum(b.num_gates for b in blocks) == len(gl)
        got = a0.copy()
        for b in blocks:
            assert len(b.qubits) <= max(kmax, 3)
            u = b.matrix
            assert np.max(np.abs(u @ u.conj().T - np.eye(len(u)))) < 1e-12
            got = apply_block(got, n, b.qubits, u)
        assert np.max(np.abs(got - want)) < 1e-12, (n, kmax)


def test_plan_wider_gate_than_kmax_stands_alone():
    import moonlab_amd as m
    gl = [(23, 0, 1, 2, 0, 0, 0), (0, 0, -1, -1, 0, 0, 0), (10, 3, 1, -1, 0, 0, 0)]
    blocks = m.fusion_plan(gl, 1)
    assert [b.qubits for b in blocks][0] == [0, 1, 2]
    assert sum(b.num_gates for b in blocks) == 3


def test_plan_blocks_commute_across_disjoint_qubits():
    """Disjoint gates do not close a block: one SU(2) + CZ brickwork layer on
    8 qubits fits two 4-qubit blocks and one straddling 2-qubit block."""
    import moonlab_amd as m
    from moonlab_amd import workloads
    gl = workloads.su2_cz_circuit(8, 1, 5)
    blocks = m.fusion_plan(gl, 4)
    assert len(blocks) <= 4
    assert sum(b.num_gates for b in blocks) == len(gl)


def test_plan_pass_count_on_su2_cz_depth60():
    """32 qubits, depth 60: the open-block planner keeps one pass per 4
    qubits per layer (480; an in-order planner needs 1320); windows of 5
    neighbouring qubits run several layers in their light cone (316)."""
    import moonlab_amd as m
    from moonlab_amd import workloads
    gl = workloads.su2_cz_circuit(32, 60, 1)
    blocks = m.fusion_plan(gl, 5)
    assert len(blocks) <= 330
    assert sum(b.num_gates for b in blocks) == len(gl)


@pytest.mark.parametrize("kmax", [3, 4, 5])
def test_window_plans_reproduce_brickwork_circuits(kmax, oracle):
    """Nearest-neighbour SU(2) + CZ brickwork on 10 and 11 qubits, where the
    window planner wins: the blocks, applied in order, give the circuit's
    state (1e-12), and each block is unitary."""
    import moonlab_amd as m
    from moonlab_amd import workloads
    rng = np.random.default_rng(7 + kmax)
    for n in (10, 11):
        gl = workloads.su2_cz_circuit(n, 12, 3 + n)
        a0 = random_state(n, rng)
        want = oracle.apply(n, a0.copy(), gl)
        blocks = m.fusion_plan(gl, kmax)
        assert len(blocks) < 12 * n / kmax * 1.5
        assert sum(b.num_gates for b in blocks) == len(gl)
        got = a0.copy()
        for b in blocks:
            assert len(b.qubits) <= kmax
            u = b.matrix
            assert np.max(np.abs(u @ u.conj().T - np.eye(len(u)))) < 1e-12
            got = apply_block(got, n, b.qubits, u)
        assert np.max(np.abs(got - want)) < 1e-12, (n, kmax)


def test_plan_rejects_bad_input():
    import moonlab_amd as m
    with pytest.raises(m.EngineError):
        m.fusion_plan([(0, 64, -1, -1, 0, 0, 0)], 5)
    with pytest.raises(m.EngineError):
        m.fusion_plan([(10, 3, 3, -1, 0, 0, 0)], 5)
    with pytest.raises(m.EngineError):
        m.fusion_plan([(0, 0, -1, -1, 0, 0, 0)], 6)
