"""Pinning the CPU oracle (oracle/sv_oracle.c) before it is trusted.

Each check is chosen so a plausible mistake in the oracle (a dropped term, a
wrong sign, a swapped qubit role, a transposed matrix, a wrong PRNG constant)
fails it:
  * golden fixtures written by the reference's own C (tests/golden/),
  * the compiled reference itself on random circuits of the full gate table,
  * closed forms (Bell, GHZ, Grover-2, rotations of |0>),
  * a brute-force dense unitary built from textbook matrices with Kronecker
    products and matrix exponentials (n <= 5).
"""
import json
import os

import ctypes

import numpy as np
import pytest
import scipy.linalg

from conftest import GOLDEN, load_corpus, random_gate_list
from oracle import oracle as O
from oracle.oracle import gates_array

# ---------------------------------------------------------------------------
# golden fixtures from the reference's C
# ---------------------------------------------------------------------------


def test_oracle_matches_reference_corpus(oracle):
    worst = 0.0
    for key, n, g, amps, z in load_corpus():
        a = oracle.run(n, g)
        worst = max(worst, np.max(np.abs(a - amps)))
        assert np.max(np.abs(oracle.expect_z_all(n, a) - z)) < 1e-13, key
    assert worst < 1e-13


def test_oracle_matches_reference_layers20(oracle):
    f = np.load(os.path.join(GOLDEN, "layers20_ref.npz"))
    n = 20
    a = oracle.run(n, f["gates"])
    assert len(f["gates"]) == 780
    assert np.max(np.abs(a[f["sample_idx"]] - f["sample_amps"])) < 1e-13
    assert np.max(np.abs(oracle.expect_z_all(n, a) - f["z"])) < 1e-12
    assert abs(oracle.norm_sq(n, a) - 1.0) < 1e-12


def test_oracle_probabilities_match_reference(oracle, ref):
    """svo_probabilities (measurement_probability_distribution,
    measurement.c:140-161): the 64 largest probabilities and their sum against
    the reference's own vector for the 20-qubit depth-20 circuit
    (layers20_ref.npz), then elementwise against the compiled reference's
    distribution on random circuits of the whole gate table.  The reference
    squares cabs(a) (hypot); |a|^2 differs from it by a few ulp at most."""
    f = np.load(os.path.join(GOLDEN, "layers20_ref.npz"))
    n = 20
    p = oracle.probabilities(n, oracle.run(n, f["gates"]))
    top = np.sort(p)[-64:]
    assert np.max(np.abs(top - f["prob_top"]) / f["prob_top"]) < 1e-12
    assert abs(np.sum(p) - f["norm"]) < 1e-12
    rng = np.random.default_rng(41)
    for n in (1, 3, 7, 11):
        g = gates_array(random_gate_list(n, 40, rng, allow3=False) if n > 1 else [(4, 0), (8, 0, -1, -1, 0.7)])
        g = g[g["type"] <= 13]                   # the reference harness runs the 14 job-path codes
        a, _, pr = ref.run(n, g, want_probs=True)
        po = oracle.probabilities(n, a)
        assert np.max(np.abs(po - pr)) <= 4 * np.finfo(float).eps * max(1.0, np.max(pr)), n
        assert abs(np.sum(po) - 1.0) < 1e-13


def test_oracle_fidelity_matches_reference(oracle, ref):
    """svo_fidelity against the reference's quantum_state_fidelity
    (state.c:289-306) on pairs of random circuit states, and the closed forms
    |<psi|psi>|^2 = 1, |<0..0|+..+>|^2 = 2^-n, orthogonal states 0.  A missing
    conjugate (sum a_i b_i) fails the random pairs."""
    L = ref.L
    L.quantum_state_fidelity.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.quantum_state_fidelity.restype = ctypes.c_double
    rng = np.random.default_rng(43)
    def circuit(n):
        # a generic complex product state (RY, RZ, RX of random angles on
        # every qubit), then random gates of the 14 job-path codes
        rows = [(t, q, -1, -1, float(rng.uniform(0, 6.3))) for q in range(n) for t in (O.RY, O.RZ, O.RX)]
        g = gates_array(rows + [tuple(x) for x in random_gate_list(n, 30, rng, allow3=False)])
        return g[g["type"] <= 13]

    for n in (2, 5, 9):
        g1, g2 = circuit(n), circuit(n)
        s1, s2 = ref.state(n, g1), ref.state(n, g2)
        try:
            want = L.quantum_state_fidelity(s1, s2)
            a1 = np.ctypeslib.as_array(L.ref_state_amplitudes(s1), shape=(2 << n,)).view(np.complex128).copy()
            a2 = np.ctypeslib.as_array(L.ref_state_amplitudes(s2), shape=(2 << n,)).view(np.complex128).copy()
        finally:
            ref.free(s1)
            ref.free(s2)
        got = oracle.fidelity(n, a1, a2)
        assert abs(got - want) < 1e-14, n
        assert abs(got - abs(np.vdot(a1, a2)) ** 2) < 1e-14
        assert abs(got - abs(np.sum(a1 * a2)) ** 2) > 1e-6      # the conjugate matters here
        assert abs(oracle.fidelity(n, a1, a1) - 1.0) < 1e-13
    n = 6
    zero = oracle.zero_state(n)
    plus = oracle.run(n, [(O.H, q) for q in range(n)])
    assert abs(oracle.fidelity(n, zero, plus) - 2.0 ** -n) < 1e-15
    one = oracle.run(n, [(O.X, 0)])
    assert oracle.fidelity(n, zero, one) == 0.0


def test_oracle_job_shots_match_reference():
    """Per-worker seeds, slicing and the collapse outcome rule, bit-exact."""
    orc = O.Oracle()
    shots = json.load(open(os.path.join(GOLDEN, "shots_ref.json")))
    rnd10 = np.load(os.path.join(GOLDEN, "rnd10_gates.npy"))
    circuits = {"bell": [(O.H, 0), (O.CNOT, 1, 0)],
                "ghz": [(O.H, 0), (O.CNOT, 1, 0), (O.CNOT, 2, 1)],
                "rnd10": rnd10, "rnd10_w1": rnd10}
    for name, gl in circuits.items():
        s = shots[name]
        a = orc.run(s["n"], gl)
        got = orc.job_shots(s["n"], a, s["shots"], s["workers"], s["seed"])
        assert got.tolist() == s["outcomes"], name
    q = shots["qgtl_rnd10"]
    a = orc.run(10, rnd10)
    assert orc.qgtl_shots(10, a, q["shots"], q["seed"]).tolist() == q["outcomes"]
    assert np.max(np.abs(np.abs(a) ** 2 - np.array(q["probabilities"]))) < 1e-15


def test_oracle_hea_energy_matches_reference(oracle):
    rows = json.load(open(os.path.join(GOLDEN, "hea_tfim_ref.json")))
    for r in rows:
        e = oracle.hea_tfim_energy(r["n"], r["layers"], np.array(r["angles"]), r["h"])
        assert abs(e - r["energy"]) < 1e-12


def test_oracle_vs_compiled_reference_full_gate_table(oracle, ref):
    rng = np.random.default_rng(5)
    for _ in range(25):
        n = int(rng.integers(3, 9))
        gl = random_gate_list(n, 80, rng)
        a = oracle.run(n, gl)
        ra, rz = ref.run(n, gates_array(gl))
        assert np.max(np.abs(a - ra)) < 1e-13
        assert np.max(np.abs(oracle.expect_z_all(n, a) - rz)) < 1e-13
        s = ref.state(n, gates_array(gl))
        try:
            for q in range(n):
                assert abs(oracle.expect_x(n, a, q) - ref.L.measurement_expectation_x(s, q)) < 1e-13
                assert abs(oracle.expect_y(n, a, q) - ref.L.measurement_expectation_y(s, q)) < 1e-13
            for q in range(n - 1):
                assert abs(oracle.correlation_zz(n, a, q, q + 1)
                           - ref.L.measurement_correlation_zz(s, q, q + 1)) < 1e-13
        finally:
            ref.free(s)


def test_oracle_collapse_rule_matches_reference(oracle, ref):
    """measurement_all_qubits' outcome for edge draws, incl. u >= total."""
    rng = np.random.default_rng(3)
    n = 6
    gl = random_gate_list(n, 30, rng) + [(O.CNOT, 2, 1)]
    a = oracle.run(n, gl)
    a[5:9] = 0.0  # zero-probability stretch
    a[-4:] = 0.0  # trailing zeros: the fallback must find the last nonzero
    for u in [0.0, 1e-300, 0.3, 0.5, 0.999999, 1.0 - 2 ** -53, 1.0, 1.5]:
        s = ref.L.ref_state_new(n)
        amp = ref.L.ref_state_amplitudes(s)
        for i in range(1 << n):
            amp[2 * i] = a[i].real
            amp[2 * i + 1] = a[i].imag
        expect = ref.L.measurement_all_qubits(s, u)
        ref.free(s)
        assert oracle.measure_all_outcome(n, a, u) == expect, u


def _ref_state_from(ref, n, a):
    s = ref.L.ref_state_new(n)
    amp = ref.L.ref_state_amplitudes(s)
    for i in range(1 << n):
        amp[2 * i] = a[i].real
        amp[2 * i + 1] = a[i].imag
    return s


def _ref_amplitudes(ref, s, n):
    amp = ref.L.ref_state_amplitudes(s)
    return np.array([complex(amp[2 * i], amp[2 * i + 1]) for i in range(1 << n)])


def test_oracle_collapsing_measurements_match_reference(oracle, ref):
    """measurement_single_qubit / measurement_partial / measurement_all_qubits
    of the compiled reference against the oracle on the same states: the
    outcome exactly, the zeroed amplitudes exactly, the renormalized ones
    within 1e-15 (the reference build contracts its norm sum into FMAs), for
    draws at and past the total probability, an unnormalized state, a
    zero-weight outcome (single qubit: -1 below the 1e-15 floor) and several
    qubit orders."""
    rng = np.random.default_rng(31)
    n = 7
    base = oracle.run(n, random_gate_list(n, 40, rng))
    states = [base, base * 1.7]                      # normalized and not
    z = base.copy()
    z[np.arange(1 << n) & 8 == 0] = 0.0              # qubit 3 is surely 1
    z /= np.linalg.norm(z)
    states.append(z)
    draws = [0.0, 0.2, 0.5, 0.77, 0.999999, 1.0 - 2 ** -53, 1.0, 1.5]
    for a in states:
        for u in draws:
            for q in (0, 3, 6):
                s = _ref_state_from(ref, n, a)
                want = ref.L.measurement_single_qubit(s, q, u)
                ra = _ref_amplitudes(ref, s, n)
                ref.free(s)
                b = a.copy()
                assert oracle.measure_qubit(n, b, q, u) == want, (q, u)
                assert np.array_equal(b == 0, ra == 0) and np.max(np.abs(b - ra)) < 1e-15, (q, u)
            for qs in ([2], [0, 5], [6, 1, 3], [4, 2, 0, 6]):
                s = _ref_state_from(ref, n, a)
                want = ref.L.measurement_partial(s, (ctypes.c_int * len(qs))(*qs), len(qs), u)
                ra = _ref_amplitudes(ref, s, n)
                ref.free(s)
                b = a.copy()
                assert oracle.measure_partial(n, b, qs, u) == want, (qs, u)
                assert np.array_equal(b == 0, ra == 0) and np.max(np.abs(b - ra)) < 1e-15, (qs, u)
            s = _ref_state_from(ref, n, a)
            want = ref.L.measurement_all_qubits(s, u)
            ra = _ref_amplitudes(ref, s, n)
            ref.free(s)
            b = a.copy()
            assert oracle.measure_all(n, b, u) == want and np.array_equal(b, ra), u
    # a draw past P(1) = 1 picks the zero-weight outcome 0: -1 and, as the
    # reference leaves it, the state zeroed
    b = z.copy()
    assert oracle.measure_qubit(n, b, 3, 1.5) == -1 and not np.any(b)


# ---------------------------------------------------------------------------
# closed forms (tests/unit/test_qgtl_backend.c:30-182, t_rotations.c)
# ---------------------------------------------------------------------------


def test_closed_forms(oracle):
    a = oracle.run(2, [(O.H, 0), (O.CNOT, 1, 0)])
    assert np.allclose(np.abs(a) ** 2, [0.5, 0, 0, 0.5], atol=1e-15)
    a = oracle.run(3, [(O.H, 0), (O.CNOT, 1, 0), (O.CNOT, 2, 1)])
    p = np.abs(a) ** 2
    assert abs(p[0] - 0.5) < 1e-15 and abs(p[7] - 0.5) < 1e-15
    grover = [(O.H, 0), (O.H, 1), (O.CZ, 1, 0), (O.H, 0), (O.H, 1), (O.Z, 0), (O.Z, 1),
              (O.CZ, 1, 0), (O.H, 0), (O.H, 1)]
    assert abs(np.abs(oracle.run(2, grover)[3]) ** 2 - 1.0) < 1e-14
    for th in [0.0, 1e-300, 1e-16, np.pi - 1e-15, 2 * np.pi, 1e6, -1e6, 0.7]:
        c, s = np.cos(th / 2), np.sin(th / 2)
        assert np.allclose(oracle.run(1, [(O.RX, 0, -1, -1, th)]), [c, -1j * s], atol=1e-15)
        assert np.allclose(oracle.run(1, [(O.RY, 0, -1, -1, th)]), [c, s], atol=1e-15)
        a = oracle.run(1, [(O.H, 0), (O.RZ, 0, -1, -1, th)])
        assert np.allclose(a, np.array([np.exp(-0.5j * th), np.exp(0.5j * th)]) / np.sqrt(2), atol=1e-15)
    n = 7
    a = oracle.run(n, [(O.H, q) for q in range(n)])
    assert np.allclose(a, 2 ** (-n / 2), atol=1e-15)
    assert np.allclose(oracle.expect_z_all(n, a), 0.0, atol=1e-15)


def test_job_path_cy_is_controlled_minus_y(oracle):
    """CY through the job path is S; CNOT; S-dagger = C(-Y)
    (moonlab_qgtl_backend.c:138-144), unlike gate_cy."""
    rng = np.random.default_rng(0)
    a0 = rng.normal(size=4) + 1j * rng.normal(size=4)
    a_job = oracle.apply(2, a0.copy(), [(O.CY, 1, 0)])
    a_true = oracle.apply(2, a0.copy(), [(O.CY_TRUE, 1, 0)])
    # control = qubit 0: indices 1 and 3 form the pair
    assert np.allclose(a_job[[1, 3]], -a_true[[1, 3]], atol=1e-15)
    assert np.allclose(a_job[[0, 2]], a0[[0, 2]], atol=0)


# ---------------------------------------------------------------------------
# brute force: full 2^n x 2^n unitaries from textbook matrices
# ---------------------------------------------------------------------------

X = np.array([[0, 1], [1, 0]], complex)
Y = np.array([[0, -1j], [1j, 0]], complex)
Zm = np.array([[1, 0], [0, -1]], complex)
P0 = np.diag([1, 0]).astype(complex)
P1 = np.diag([0, 1]).astype(complex)
Hm = np.array([[1, 1], [1, -1]], complex) / np.sqrt(2)


def embed(ops, n):
    m = np.array([[1.0 + 0j]])
    for q in range(n - 1, -1, -1):  # qubit 0 is the rightmost (least significant) factor
        m = np.kron(m, ops.get(q, np.eye(2)))
    return m


def textbook(g, n):
    t, tg, c, e, th, ph, la = g
    rot = lambda P: scipy.linalg.expm(-0.5j * th * P)
    ph1 = lambda x: np.diag([1, np.exp(1j * x)])
    one = {1: X, 2: Y, 3: Zm, 4: Hm, 5: ph1(np.pi / 2), 6: ph1(np.pi / 4), 7: rot(X), 8: rot(Y),
           9: rot(Zm), 14: ph1(-np.pi / 2), 15: ph1(-np.pi / 4), 16: ph1(th)}
    if t in one:
        return embed({tg: one[t]}, n)
    if t == 17:
        u = np.array([[np.cos(th / 2), -np.exp(1j * la) * np.sin(th / 2)],
                      [np.exp(1j * ph) * np.sin(th / 2), np.exp(1j * (ph + la)) * np.cos(th / 2)]])
        return embed({tg: u}, n)
    ctrl = lambda U: embed({c: P0}, n) + embed({c: P1, tg: U}, n)
    two = {10: X, 11: -Y, 12: Zm, 18: ph1(th), 19: rot(X), 20: rot(Y), 21: rot(Zm), 22: Y}
    if t in two:
        return ctrl(two[t])
    swap = lambda a, b, extra: 0.5 * sum(embed({**extra, a: S, b: S}, n) for S in (np.eye(2), X, Y, Zm))
    if t == 13:
        return swap(c, tg, {})
    if t == 23:
        return np.eye(1 << n) - embed({c: P1, e: P1}, n) + embed({c: P1, e: P1, tg: X}, n)
    if t == 24:
        return embed({c: P0}, n) + swap(tg, e, {c: P1})
    raise ValueError(t)


def test_oracle_vs_brute_force_unitaries(oracle):
    rng = np.random.default_rng(17)
    for _ in range(30):
        n = int(rng.integers(1, 6))
        gl = random_gate_list(n, 12, rng)
        ga = gates_array(gl)
        u = np.eye(1 << n, dtype=complex)
        for g in ga:
            u = textbook(tuple(g), n) @ u
        a0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        a = oracle.apply(n, a0.copy(), ga)
        assert np.max(np.abs(a - u @ a0)) < 1e-12


def test_oracle_generic_matrices_and_compose(oracle):
    rng = np.random.default_rng(23)
    n = 4
    for _ in range(10):
        m2 = scipy.linalg.expm(1j * (lambda h: h + h.conj().T)(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))))
        m4 = scipy.linalg.expm(1j * (lambda h: h + h.conj().T)(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))))
        q1, q2 = [int(x) for x in rng.permutation(n)[:2]]
        a0 = rng.normal(size=16) + 1j * rng.normal(size=16)
        a = oracle.apply_matrix_1q(n, a0.copy(), q1, m2)
        assert np.allclose(a, embed({q1: m2}, n) @ a0, atol=1e-13)
        # 4x4 basis |q2 q1>, q1 low: the full operator is a sum over 2x2 blocks
        full = sum(m4[(r2 << 1) | r1, (c2 << 1) | c1]
                   * embed({q1: np.outer(np.eye(2)[r1], np.eye(2)[c1]),
                            q2: np.outer(np.eye(2)[r2], np.eye(2)[c2])}, n)
                   for r1 in range(2) for r2 in range(2) for c1 in range(2) for c2 in range(2))
        a = oracle.apply_matrix_2q(n, a0.copy(), q1, q2, m4)
        assert np.allclose(a, full @ a0, atol=1e-13)
    # composition of a block = product of textbook unitaries
    gl = random_gate_list(3, 15, rng)
    u = np.eye(8, dtype=complex)
    for g in gates_array(gl):
        u = textbook(tuple(g), 3) @ u
    assert np.allclose(oracle.compose_unitary([0, 1, 2], gl), u, atol=1e-12)


def test_splitmix_constant_pinned(oracle):
    # scheduler.c:145-146 multiplies by 0x94d49bb133111ebb; the job-shot goldens
    # above already depend on it, this pins the value directly.
    x = 0x0123456789ABCDEF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
    x = ((x ^ (x >> 27)) * 0x94D49BB133111EBB) & ((1 << 64) - 1)
    assert oracle.splitmix64(0x0123456789ABCDEF) == x ^ (x >> 31)


def _pauli_dense(pauli):
    """Kronecker product with qubit 0 as the least significant factor."""
    P = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1.0, -1.0])}
    m = np.array([[1.0 + 0j]])
    for c in pauli:               # pauli[q] acts on qubit q
        m = np.kron(P[c], m)
    return m


def test_oracle_pauli_expectation_pinned(oracle, ref):
    """The restatement of vqe_measure_pauli_expectation against brute force
    (dense Kronecker products) and, for one- and two-letter strings, the
    reference's own measurement_expectation_{x,y,z} / correlation_zz."""
    rng = np.random.default_rng(31)
    for n in (1, 2, 3, 5):
        for _ in range(12):
            gl = random_gate_list(n, 25, rng)
            a = oracle.run(n, gl)
            pauli = "".join(rng.choice(list("IXYZ"), size=n))
            want = float(np.real(np.vdot(a, _pauli_dense(pauli) @ a)))
            assert abs(oracle.expect_pauli(n, a, pauli) - want) < 1e-12, pauli
    n = 6
    gl = random_gate_list(n, 60, rng)
    a = oracle.run(n, gl)
    s = ref.state(n, gates_array([tuple(g) for g in gl]))
    try:
        for q in range(n):
            for c, f in (("X", "measurement_expectation_x"), ("Y", "measurement_expectation_y"),
                         ("Z", "measurement_expectation_z")):
                p = "I" * q + c + "I" * (n - q - 1)
                assert abs(oracle.expect_pauli(n, a, p) - getattr(ref.L, f)(s, q)) < 1e-12, p
        for q in range(n - 1):
            p = "I" * q + "ZZ" + "I" * (n - q - 2)
            assert abs(oracle.expect_pauli(n, a, p) - ref.L.measurement_correlation_zz(s, q, q + 1)) < 1e-12
    finally:
        ref.free(s)


def test_oracle_parameter_shift_gradient_pinned(oracle):
    """The two-term shift rule on the HEA/TFIM energy equals the central
    finite difference (a dropped factor, sign or index fails this)."""
    rng = np.random.default_rng(33)
    for n, L in ((2, 1), (3, 2), (5, 2)):
        ang = rng.uniform(0, 2 * np.pi, size=(L, n, 2))
        h = float(rng.uniform(0.5, 1.5))
        g = oracle.hea_tfim_gradient(n, L, ang, h)
        eps = 1e-5
        fd = np.zeros_like(ang)
        for idx in np.ndindex(ang.shape):
            a = ang.copy()
            a[idx] += eps
            ep = oracle.hea_tfim_energy(n, L, a, h)
            a[idx] -= 2 * eps
            em = oracle.hea_tfim_energy(n, L, a, h)
            fd[idx] = (ep - em) / (2 * eps)
        assert np.max(np.abs(g - fd)) < 1e-8


# ---------------------------------------------------------------------------
# variational energies, gradients, geometric tensor (vqe.c, vqe_qng.c)
# ---------------------------------------------------------------------------

def _random_hamiltonian(n, rng, count):
    terms = []
    for _ in range(count):
        s = "".join(rng.choice(list("IXYZ"), size=n))
        terms.append((float(rng.uniform(-1, 1)), s))
    terms.append((0.5, "I" * n))
    return terms


def test_oracle_vqe_matches_reference_solver(oracle, ref):
    """svo_vqe_energy / _gradient / _qgt and the natural-gradient solve
    against the compiled reference's vqe_compute_energy, vqe_compute_gradient,
    vqe_compute_qgt, vqe_compute_berry_curvature and
    vqe_natural_gradient_direction on its hardware-efficient ansatz (CNOT
    ladder), with Hartree-Fock bits and a constant."""
    rng = np.random.default_rng(61)
    for n, L in ((1, 1), (2, 2), (3, 1), (4, 2), (5, 3)):
        terms = _random_hamiltonian(n, rng, 6)
        hf = int(rng.integers(0, 1 << n))
        ang = rng.uniform(0, 2 * np.pi, 2 * n * L)
        v = ref.vqe(n, L, terms, hf=hf, constant=0.25)
        assert abs(oracle.vqe_energy(n, L, ang, terms, 1, hf, 0.25) - v.energy(ang)) < 1e-13
        g = oracle.vqe_gradient(n, L, ang, terms, 1, hf)
        assert np.max(np.abs(g - v.gradient(ang))) < 1e-12
        qg, qf = oracle.vqe_qgt(n, L, ang, 1, hf)
        assert np.max(np.abs(qg - v.qgt(ang))) < 1e-14
        assert np.max(np.abs(qf - v.berry(ang))) < 1e-14
        A = qg + 1e-3 * np.eye(qg.shape[0])
        ng = oracle.solve(A, g)
        want = v.natural_gradient(ang, g, 1e-3)
        assert np.max(np.abs(ng - want)) <= 1e-9 * max(1.0, np.max(np.abs(want))), (n, L)
        assert np.max(np.abs(ng - np.linalg.solve(A, g))) <= 1e-9 * max(1.0, np.max(np.abs(want)))


def test_oracle_qgt_closed_form_and_finite_differences(oracle):
    """One qubit, |psi> = RZ(phi) RY(theta)|0>: metric diag(1/4, sin^2(theta)/4)
    and Berry curvature F_theta,phi = -sin(theta)/2 (the -2 Im convention).
    Then the metric and curvature of a 3-qubit CZ-ladder ansatz against
    central differences of the prepared state (<d_i psi|d_j psi> -
    <d_i psi|psi><psi|d_j psi> from differenced states)."""
    for theta, phi in ((0.3, 1.1), (2.0, -0.7), (np.pi / 2, 0.0)):
        g, f = oracle.vqe_qgt(1, 1, [theta, phi], 0, 0)
        assert np.allclose(g, [[0.25, 0.0], [0.0, np.sin(theta) ** 2 / 4]], atol=1e-15)
        assert np.allclose(f, [[0.0, -np.sin(theta) / 2], [np.sin(theta) / 2, 0.0]], atol=1e-15)
    rng = np.random.default_rng(67)
    n, L = 3, 2
    ang = rng.uniform(0, 2 * np.pi, 2 * n * L)
    psi = oracle.vqe_state(n, L, ang, 0, 0b010)
    d = []
    h = 1e-5
    for i in range(ang.size):
        a1, a2 = ang.copy(), ang.copy()
        a1[i] += h
        a2[i] -= h
        d.append((oracle.vqe_state(n, L, a1, 0, 0b010) - oracle.vqe_state(n, L, a2, 0, 0b010)) / (2 * h))
    d = np.array(d)
    ov = d.conj() @ psi
    Q = d.conj() @ d.T - np.outer(ov, ov.conj())
    g, f = oracle.vqe_qgt(n, L, ang, 0, 0b010)
    assert np.max(np.abs(g - Q.real)) < 1e-8
    assert np.max(np.abs(f - (-2 * Q.imag))) < 1e-8
    assert np.max(np.abs(f)) > 1e-2                 # the RZ layers make the curvature nonzero


def test_oracle_solve_vs_numpy(oracle):
    rng = np.random.default_rng(71)
    for m in (1, 2, 7, 30):
        A = rng.normal(size=(m, m)) + m * np.eye(m)
        b = rng.normal(size=m)
        assert np.max(np.abs(oracle.solve(A, b) - np.linalg.solve(A, b))) < 1e-12
    with pytest.raises(np.linalg.LinAlgError):
        oracle.solve(np.zeros((3, 3)), np.ones(3))
