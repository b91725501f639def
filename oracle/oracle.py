"""ctypes access to the two CPU checkers.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package
``moonlab_amd`` (which refuses to run without its HIP library).

* :class:`Oracle` wraps ``oracle/libsv_oracle.so``, this repository's C
  restatement of the reference's dense state-vector semantics.
* :class:`RefLib` wraps ``oracle/_ref/libmoonlab_ref.so``, the reference's own
  C compiled from its sources by ``oracle/Makefile`` (``make -C oracle ref``).
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libsv_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmoonlab_ref.so")

# Gate codes shared with oracle/sv_oracle.h (0..13 = moonlab_qgtl_gate_t).
I, X, Y, Z, H, S, T, RX, RY, RZ, CNOT, CY, CZ, SWAP = range(14)
SDG, TDG, PHASE, U3, CPHASE, CRX, CRY, CRZ, CY_TRUE, TOFFOLI, FREDKIN = range(14, 25)

GATE_DTYPE = np.dtype([("type", "<i4"), ("target", "<i4"), ("control", "<i4"),
                       ("extra", "<i4"), ("theta", "<f8"), ("phi", "<f8"),
                       ("lam", "<f8")])


def gates_array(gates: Iterable[Sequence]) -> np.ndarray:
    """Pack (type, target, control, extra, theta, phi, lam) tuples; missing
    trailing fields default to -1 / 0.0."""
    rows = []
    for g in gates:
        g = tuple(g)
        t = g[0]
        tgt = g[1]
        ctl = g[2] if len(g) > 2 else -1
        ext = g[3] if len(g) > 3 else -1
        th = g[4] if len(g) > 4 else 0.0
        ph = g[5] if len(g) > 5 else 0.0
        la = g[6] if len(g) > 6 else 0.0
        rows.append((t, tgt, ctl, ext, th, ph, la))
    return np.array(rows, dtype=GATE_DTYPE)


def _ptr(a: np.ndarray, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class Oracle:
    """The C restatement (oracle/sv_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO) -> None:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.svo_init_zero.argtypes = [ctypes.c_int, dp]
        L.svo_apply_gates.argtypes = [ctypes.c_int, dp, vp, ctypes.c_int]
        L.svo_apply_gates.restype = ctypes.c_int
        L.svo_apply_matrix_1q.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp]
        L.svo_apply_matrix_2q.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, dp]
        L.svo_probabilities.argtypes = [ctypes.c_int, dp, dp]
        L.svo_probability_one.argtypes = [ctypes.c_int, dp, ctypes.c_int]
        L.svo_probability_one.restype = ctypes.c_double
        L.svo_expect_z_all.argtypes = [ctypes.c_int, dp, dp]
        for f in ("svo_expect_x", "svo_expect_y"):
            getattr(L, f).argtypes = [ctypes.c_int, dp, ctypes.c_int]
            getattr(L, f).restype = ctypes.c_double
        L.svo_correlation_zz.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int]
        L.svo_correlation_zz.restype = ctypes.c_double
        L.svo_expect_pauli.argtypes = [ctypes.c_int, dp, ctypes.c_char_p]
        L.svo_expect_pauli.restype = ctypes.c_double
        L.svo_norm_sq.argtypes = [ctypes.c_int, dp]
        L.svo_norm_sq.restype = ctypes.c_double
        L.svo_fidelity.argtypes = [ctypes.c_int, dp, dp]
        L.svo_fidelity.restype = ctypes.c_double
        L.svo_measure_all_outcome.argtypes = [ctypes.c_int, dp, ctypes.c_double]
        L.svo_measure_all_outcome.restype = ctypes.c_uint64
        L.svo_measure_all.argtypes = [ctypes.c_int, dp, ctypes.c_double]
        L.svo_measure_all.restype = ctypes.c_uint64
        L.svo_measure_qubit.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_double]
        L.svo_measure_qubit.restype = ctypes.c_int
        L.svo_measure_partial.argtypes = [ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                          ctypes.c_double]
        L.svo_measure_partial.restype = ctypes.c_uint64
        L.svo_splitmix64.argtypes = [ctypes.c_uint64]
        L.svo_splitmix64.restype = ctypes.c_uint64
        L.svo_qgtl_shots.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_uint64, u64p]
        L.svo_job_shots.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_uint64, u64p]
        L.svo_job_shots.restype = ctypes.c_int
        L.svo_hea_tfim_energy.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, dp]
        L.svo_hea_tfim_energy.restype = ctypes.c_double
        L.svo_compose_unitary.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), vp,
                                          ctypes.c_int, dp]
        L.svo_compose_unitary.restype = ctypes.c_int
        cpp = ctypes.POINTER(ctypes.c_char_p)
        L.svo_vqe_state.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_uint64, dp]
        L.svo_vqe_energy.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                     cpp, dp, ctypes.c_double]
        L.svo_vqe_energy.restype = ctypes.c_double
        L.svo_vqe_gradient.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.c_int, cpp, dp, dp]
        L.svo_vqe_qgt.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_uint64, dp, dp]
        L.svo_vqe_qgt.restype = ctypes.c_int
        L.svo_solve.argtypes = [ctypes.c_int, dp, dp, dp]
        L.svo_solve.restype = ctypes.c_int
        self.L = L

    # -- state evolution -------------------------------------------------
    def zero_state(self, n: int) -> np.ndarray:
        a = np.zeros(1 << n, dtype=np.complex128)
        a[0] = 1.0
        return a

    def apply(self, n: int, amps: np.ndarray, gates) -> np.ndarray:
        g = gates if isinstance(gates, np.ndarray) else gates_array(gates)
        assert amps.dtype == np.complex128 and amps.flags.c_contiguous
        if len(g):
            rc = self.L.svo_apply_gates(n, _ptr(amps.view(np.float64)), g.ctypes.data, len(g))
            if rc:
                raise ValueError("oracle rejected a gate")
        return amps

    def run(self, n: int, gates) -> np.ndarray:
        return self.apply(n, self.zero_state(n), gates)

    def apply_matrix_1q(self, n, amps, q, m):
        m = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(2, 2))
        self.L.svo_apply_matrix_1q(n, _ptr(amps.view(np.float64)), q, _ptr(m.view(np.float64)))
        return amps

    def apply_matrix_2q(self, n, amps, q1, q2, m):
        m = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(4, 4))
        self.L.svo_apply_matrix_2q(n, _ptr(amps.view(np.float64)), q1, q2,
                                   _ptr(m.view(np.float64)))
        return amps

    # -- reductions --------------------------------------------------------
    def probabilities(self, n, amps):
        p = np.empty(1 << n, dtype=np.float64)
        self.L.svo_probabilities(n, _ptr(amps.view(np.float64)), _ptr(p))
        return p

    def expect_z_all(self, n, amps):
        z = np.empty(n, dtype=np.float64)
        self.L.svo_expect_z_all(n, _ptr(amps.view(np.float64)), _ptr(z))
        return z

    def expect_x(self, n, amps, q):
        return self.L.svo_expect_x(n, _ptr(amps.view(np.float64)), q)

    def expect_y(self, n, amps, q):
        return self.L.svo_expect_y(n, _ptr(amps.view(np.float64)), q)

    def expect_pauli(self, n, amps, pauli: str):
        return self.L.svo_expect_pauli(n, _ptr(amps.view(np.float64)), pauli.encode())

    def correlation_zz(self, n, amps, q1, q2):
        return self.L.svo_correlation_zz(n, _ptr(amps.view(np.float64)), q1, q2)

    def norm_sq(self, n, amps):
        return self.L.svo_norm_sq(n, _ptr(amps.view(np.float64)))

    def fidelity(self, n, a, b):
        return self.L.svo_fidelity(n, _ptr(a.view(np.float64)), _ptr(b.view(np.float64)))

    def measure_all_outcome(self, n, amps, u):
        return int(self.L.svo_measure_all_outcome(n, _ptr(amps.view(np.float64)), u))

    # collapsing measurements: amps (complex128, C-contiguous) is updated in place
    def measure_all(self, n, amps, u):
        return int(self.L.svo_measure_all(n, _ptr(amps.view(np.float64)), u))

    def measure_qubit(self, n, amps, q, u):
        return int(self.L.svo_measure_qubit(n, _ptr(amps.view(np.float64)), q, u))

    def measure_partial(self, n, amps, qubits, u):
        qs = (ctypes.c_int * len(qubits))(*qubits)
        return int(self.L.svo_measure_partial(n, _ptr(amps.view(np.float64)), qs, len(qubits), u))

    def splitmix64(self, x):
        return int(self.L.svo_splitmix64(x))

    def qgtl_shots(self, n, amps, shots, seed):
        out = np.zeros(shots, dtype=np.uint64)
        self.L.svo_qgtl_shots(n, _ptr(amps.view(np.float64)), shots, seed,
                              _ptr(out, ctypes.c_uint64))
        return out

    def job_shots(self, n, amps, shots, workers, seed):
        out = np.zeros(shots, dtype=np.uint64)
        w = self.L.svo_job_shots(n, _ptr(amps.view(np.float64)), shots, workers, seed,
                                 _ptr(out, ctypes.c_uint64))
        if w < 0:
            raise ValueError("bad job-shot arguments")
        return out

    def hea_tfim_energy(self, n, layers, angles, h):
        angles = np.ascontiguousarray(angles, dtype=np.float64)
        assert angles.size == layers * n * 2
        scratch = np.empty(1 << n, dtype=np.complex128)
        return self.L.svo_hea_tfim_energy(n, layers, _ptr(angles), h,
                                          _ptr(scratch.view(np.float64)))

    def hea_tfim_gradient(self, n, layers, angles, h):
        """Two-term parameter shift over svo_hea_tfim_energy, the rule of
        vqe_compute_gradient for the hardware-efficient ansatz (vqe.c:2094-2130)."""
        a = np.ascontiguousarray(angles, dtype=np.float64).reshape(-1).copy()
        g = np.zeros_like(a)
        for k in range(a.size):
            a[k] += np.pi / 2
            ep = self.hea_tfim_energy(n, layers, a, h)
            a[k] -= np.pi
            em = self.hea_tfim_energy(n, layers, a, h)
            a[k] += np.pi / 2
            g[k] = 0.5 * (ep - em)
        return g.reshape(layers, n, 2)

    # -- variational: Pauli sums, gradients, geometric tensor ---------------
    @staticmethod
    def _terms(terms):
        strs = (ctypes.c_char_p * max(len(terms), 1))(*[t[1].encode() for t in terms])
        coef = np.ascontiguousarray([t[0] for t in terms] or [0.0], dtype=np.float64)
        return strs, coef

    def vqe_state(self, n, layers, angles, entangler=1, hf=0):
        a = np.ascontiguousarray(angles, dtype=np.float64).reshape(-1)
        out = np.empty(1 << n, dtype=np.complex128)
        self.L.svo_vqe_state(n, layers, _ptr(a), entangler, hf, _ptr(out.view(np.float64)))
        return out

    def vqe_energy(self, n, layers, angles, terms, entangler=1, hf=0, constant=0.0):
        """terms: (coefficient, pauli string) pairs; string[q] acts on qubit q."""
        a = np.ascontiguousarray(angles, dtype=np.float64).reshape(-1)
        strs, coef = self._terms(terms)
        return self.L.svo_vqe_energy(n, layers, _ptr(a), entangler, hf, len(terms), strs, _ptr(coef), constant)

    def vqe_gradient(self, n, layers, angles, terms, entangler=1, hf=0):
        a = np.ascontiguousarray(angles, dtype=np.float64).reshape(-1)
        strs, coef = self._terms(terms)
        g = np.zeros(a.size)
        self.L.svo_vqe_gradient(n, layers, _ptr(a), entangler, hf, len(terms), strs, _ptr(coef), _ptr(g))
        return g

    def vqe_qgt(self, n, layers, angles, entangler=1, hf=0):
        """(metric, Berry curvature), each P x P with P = 2 n layers."""
        a = np.ascontiguousarray(angles, dtype=np.float64).reshape(-1)
        P = 2 * n * layers
        g = np.zeros((P, P))
        f = np.zeros((P, P))
        if self.L.svo_vqe_qgt(n, layers, _ptr(a), entangler, hf, _ptr(g), _ptr(f)):
            raise MemoryError("svo_vqe_qgt failed")
        return g, f

    def solve(self, A, b):
        A = np.array(A, dtype=np.float64, order="C")
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(b.size)
        if self.L.svo_solve(b.size, _ptr(A), _ptr(b), _ptr(x)):
            raise np.linalg.LinAlgError("singular")
        return x

    def compose_unitary(self, qubits, gates):
        g = gates if isinstance(gates, np.ndarray) else gates_array(gates)
        k = len(qubits)
        q = (ctypes.c_int * k)(*qubits)
        mat = np.zeros((1 << k, 1 << k), dtype=np.complex128)
        rc = self.L.svo_compose_unitary(k, q, g.ctypes.data if len(g) else None, len(g),
                                        _ptr(mat.view(np.float64)))
        if rc:
            raise ValueError("compose_unitary failed")
        return mat


class QgtlOptions(ctypes.Structure):   # moonlab_qgtl_options_t (moonlab_qgtl_backend.h:90-95)
    _fields_ = [("num_shots", ctypes.c_int), ("rng_seed", ctypes.c_uint64),
                ("return_probabilities", ctypes.c_int)]


class QgtlResults(ctypes.Structure):   # moonlab_qgtl_results_t (moonlab_qgtl_backend.h:97-109)
    _fields_ = [("num_qubits", ctypes.c_int), ("num_shots", ctypes.c_int),
                ("outcomes", ctypes.POINTER(ctypes.c_uint64)),
                ("probabilities", ctypes.POINTER(ctypes.c_double))]


class RefLib:
    """The reference's own C (oracle/_ref/libmoonlab_ref.so)."""

    def __init__(self, path: str = REF_SO) -> None:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.ref_state_new.argtypes = [ctypes.c_int]
        L.ref_state_new.restype = vp
        L.ref_state_delete.argtypes = [vp]
        L.ref_state_amplitudes.argtypes = [vp]
        L.ref_state_amplitudes.restype = dp
        L.ref_apply_gate.argtypes = [vp, vp]
        L.ref_apply_gate.restype = ctypes.c_int
        L.ref_apply_gates.argtypes = [vp, vp, ctypes.c_int]
        L.ref_apply_gates.restype = ctypes.c_int
        L.ref_run.argtypes = [ctypes.c_int, vp, ctypes.c_int, dp, dp, dp]
        L.ref_run.restype = ctypes.c_int
        for f in ("measurement_expectation_x", "measurement_expectation_y",
                  "measurement_expectation_z", "measurement_probability_one"):
            getattr(L, f).argtypes = [vp, ctypes.c_int]
            getattr(L, f).restype = ctypes.c_double
        L.measurement_correlation_zz.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.measurement_correlation_zz.restype = ctypes.c_double
        L.measurement_all_qubits.argtypes = [vp, ctypes.c_double]
        L.measurement_all_qubits.restype = ctypes.c_uint64
        L.measurement_single_qubit.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.measurement_single_qubit.restype = ctypes.c_int
        L.measurement_partial.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_double]
        L.measurement_partial.restype = ctypes.c_uint64
        L.quantum_state_clone.argtypes = [vp, vp]
        L.quantum_state_clone.restype = ctypes.c_int
        ip = ctypes.c_int
        L.moonlab_qgtl_circuit_create.restype = vp
        L.moonlab_qgtl_circuit_create.argtypes = [ip]
        L.moonlab_qgtl_add_gate.argtypes = [vp, ip, ip, ip, dp]
        L.moonlab_qgtl_execute.argtypes = [vp, ctypes.POINTER(QgtlOptions), ctypes.POINTER(QgtlResults)]
        L.moonlab_qgtl_results_free.argtypes = [ctypes.POINTER(QgtlResults)]
        L.moonlab_qgtl_circuit_free.argtypes = [vp]
        self.L = L

    def qgtl_execute(self, n: int, gates, shots: int, seed: int) -> np.ndarray:
        """One request through the reference's circuit API (moonlab_qgtl_execute:
        build the circuit, simulate, draw `shots` outcomes); returns the outcomes."""
        L = self.L
        c = L.moonlab_qgtl_circuit_create(n)
        if not c:
            raise MemoryError("moonlab_qgtl_circuit_create failed")
        try:
            for g in gates:
                t, tgt, ctl, th = int(g[0]), int(g[1]), int(g[2]), float(g[4])
                p = (ctypes.c_double * 1)(th) if t in (RX, RY, RZ) else None
                if L.moonlab_qgtl_add_gate(c, t, tgt, ctl, p) != 0:
                    raise ValueError("reference rejected a gate")
            o = QgtlOptions(shots, seed, 0)
            r = QgtlResults()
            if L.moonlab_qgtl_execute(c, ctypes.byref(o), ctypes.byref(r)) != 0:
                raise ValueError("moonlab_qgtl_execute failed")
            out = np.array([r.outcomes[i] for i in range(shots)], dtype=np.uint64)
            L.moonlab_qgtl_results_free(ctypes.byref(r))
            return out
        finally:
            L.moonlab_qgtl_circuit_free(c)

    def run(self, n: int, gates, want_probs: bool = False):
        g = gates if isinstance(gates, np.ndarray) else gates_array(gates)
        amps = np.zeros(1 << n, dtype=np.complex128)
        z = np.zeros(n, dtype=np.float64)
        p = np.zeros(1 << n, dtype=np.float64) if want_probs else None
        rc = self.L.ref_run(n, g.ctypes.data if len(g) else None, len(g),
                            _ptr(amps.view(np.float64)), _ptr(z),
                            _ptr(p) if p is not None else None)
        if rc:
            raise ValueError(f"reference rejected the circuit (rc={rc})")
        return (amps, z, p) if want_probs else (amps, z)

    def state(self, n: int, gates):
        """A live reference quantum_state_t* after the gates (caller frees)."""
        g = gates if isinstance(gates, np.ndarray) else gates_array(gates)
        s = self.L.ref_state_new(n)
        if not s:
            raise MemoryError("ref_state_new failed")
        if len(g) and self.L.ref_apply_gates(s, g.ctypes.data, len(g)):
            self.L.ref_state_delete(s)
            raise ValueError("reference rejected the circuit")
        return s

    def free(self, s):
        self.L.ref_state_delete(s)

    # -- the reference's own VQE solver (vqe.c, vqe_qng.c) over its
    #    hardware-efficient ansatz (CNOT ladder) -------------------------------
    def vqe(self, n, layers, terms, hf=0, constant=0.0):
        L = self.L
        if not getattr(L, "_vqe_declared", False):
            dp = ctypes.POINTER(ctypes.c_double)
            vp = ctypes.c_void_p
            L.ref_vqe_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                      dp, ctypes.c_uint64, ctypes.c_double]
            L.ref_vqe_new.restype = vp
            L.ref_vqe_free.argtypes = [vp]
            L.ref_vqe_energy.argtypes = [vp, dp]
            L.ref_vqe_energy.restype = ctypes.c_double
            for f in ("ref_vqe_gradient", "ref_vqe_qgt", "ref_vqe_berry"):
                getattr(L, f).argtypes = [vp, dp, dp]
                getattr(L, f).restype = ctypes.c_int
            L.ref_vqe_natural_gradient.argtypes = [vp, dp, dp, ctypes.c_double, dp]
            L.ref_vqe_natural_gradient.restype = ctypes.c_int
            L._vqe_declared = True
        return RefVqe(L, n, layers, terms, hf, constant)


class RefVqe:
    """One reference vqe_solver_t (test infrastructure)."""

    def __init__(self, L, n, layers, terms, hf, constant):
        self.L, self.P = L, 2 * n * layers
        strs = (ctypes.c_char_p * len(terms))(*[t[1].encode() for t in terms])
        coef = np.ascontiguousarray([t[0] for t in terms], dtype=np.float64)
        self.h = L.ref_vqe_new(n, layers, len(terms), strs, _ptr(coef), hf, constant)
        if not self.h:
            raise ValueError("ref_vqe_new failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_vqe_free(self.h)
            self.h = None

    def _p(self, params):
        return np.ascontiguousarray(params, dtype=np.float64).reshape(-1)

    def energy(self, params):
        return self.L.ref_vqe_energy(self.h, _ptr(self._p(params)))

    def _mat(self, f, params, shape):
        out = np.zeros(shape)
        if f(self.h, _ptr(self._p(params)), _ptr(out)):
            raise ValueError("reference call failed")
        return out

    def gradient(self, params):
        return self._mat(self.L.ref_vqe_gradient, params, (self.P,))

    def qgt(self, params):
        return self._mat(self.L.ref_vqe_qgt, params, (self.P, self.P))

    def berry(self, params):
        return self._mat(self.L.ref_vqe_berry, params, (self.P, self.P))

    def natural_gradient(self, params, grad, reg):
        out = np.zeros(self.P)
        g = np.ascontiguousarray(grad, dtype=np.float64)
        if self.L.ref_vqe_natural_gradient(self.h, _ptr(self._p(params)), _ptr(g), reg, _ptr(out)):
            raise np.linalg.LinAlgError("singular")
        return out
