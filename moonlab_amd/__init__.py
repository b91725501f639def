"""moonlab_amd -- Python access to the MI355X dense state-vector engine.

Everything here is a thin ctypes layer over ``moonlab_amd/lib/libmoonlab_mi355x.so``
(built by ``python -m moonlab_amd.build``).  The job and circuit classes mirror
the reference's Python binding (``bindings/python/moonlab/scheduler.py``,
``bindings/python/moonlab/qgtl.py``): same method names, same status codes,
same exceptions.  There is no CPU execution path: without the library, or
without a visible GPU, calls that execute raise.
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libmoonlab_mi355x.so")

# ---- status codes (scheduler.h:36-42,187; moonlab_qgtl_backend.h:77-81) ----
MOONLAB_SCHED_OK = 0
MOONLAB_SCHED_BAD_ARG = -501
MOONLAB_SCHED_OOM = -502
MOONLAB_SCHED_INTERNAL = -503
MOONLAB_SCHED_BUFFER_TOO_SMALL = -504
MOONLAB_SCHED_NOT_BUILT = -505
MOONLAB_SCHED_BACKEND_NOT_FOUND = -506
MOONLAB_SCHED_BACKEND_BUSY = -507
MOONLAB_QGTL_OK = 0
MOONLAB_QGTL_BAD_ARG = -301
MOONLAB_QGTL_OOM = -302
MOONLAB_QGTL_UNSUPPORTED = -303
MOONLAB_QGTL_INTERNAL = -304

C128 = 0
C64 = 1


class GateType(enum.IntEnum):
    """moonlab_qgtl_gate_t (moonlab_qgtl_backend.h:57-72) plus the wider
    gate_* family the engine also runs (codes 14..24)."""
    I = 0
    X = 1
    Y = 2
    Z = 3
    H = 4
    S = 5
    T = 6
    RX = 7
    RY = 8
    RZ = 9
    CNOT = 10
    CY = 11
    CZ = 12
    SWAP = 13
    SDG = 14
    TDG = 15
    PHASE = 16
    U3 = 17
    CPHASE = 18
    CRX = 19
    CRY = 20
    CRZ = 21
    CY_TRUE = 22
    TOFFOLI = 23
    FREDKIN = 24


class SchedulerError(RuntimeError):
    """Any negative status from the job API."""


class QGTLError(RuntimeError):
    """Any negative status from the circuit API."""


class EngineError(RuntimeError):
    """Any negative status from the moonlab_hip_* layer."""


GATE_DTYPE = np.dtype([("type", "<i4"), ("target", "<i4"), ("control", "<i4"),
                       ("extra", "<i4"), ("theta", "<f8"), ("phi", "<f8"), ("lam", "<f8")])


def gate_array(gates) -> np.ndarray:
    """(type, target[, control[, extra[, theta[, phi[, lam]]]]]) tuples -> records."""
    if isinstance(gates, np.ndarray) and gates.dtype == GATE_DTYPE:
        return np.ascontiguousarray(gates)
    rows = []
    for g in gates:
        g = tuple(g)
        rows.append((int(g[0]), int(g[1]), int(g[2]) if len(g) > 2 else -1,
                     int(g[3]) if len(g) > 3 else -1, float(g[4]) if len(g) > 4 else 0.0,
                     float(g[5]) if len(g) > 5 else 0.0, float(g[6]) if len(g) > 6 else 0.0))
    return np.array(rows, dtype=GATE_DTYPE)


class _JobResults(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_int), ("total_shots", ctypes.c_int),
                ("outcomes", ctypes.POINTER(ctypes.c_uint64)),
                ("num_workers_used", ctypes.c_int),
                ("worker_seconds", ctypes.POINTER(ctypes.c_double))]


class _QgtlOpts(ctypes.Structure):
    _fields_ = [("num_shots", ctypes.c_int), ("rng_seed", ctypes.c_uint64),
                ("return_probabilities", ctypes.c_int)]


class _QgtlResults(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_int), ("num_shots", ctypes.c_int),
                ("outcomes", ctypes.POINTER(ctypes.c_uint64)),
                ("probabilities", ctypes.POINTER(ctypes.c_double))]


class _Backend(ctypes.Structure):
    pass


BACKEND_EXECUTE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p,
                                      ctypes.POINTER(_JobResults), ctypes.c_void_p)
COMPLETION_HOOK_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(_JobResults),
                                      ctypes.c_char_p, ctypes.c_void_p)
_Backend._fields_ = [("name", ctypes.c_char_p), ("execute", BACKEND_EXECUTE_FN),
                     ("ctx", ctypes.c_void_p), ("description", ctypes.c_char_p)]


class _ExecOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("return_amplitudes", ctypes.c_int),
                ("return_probabilities", ctypes.c_int), ("return_expect_z", ctypes.c_int),
                ("fusion_max_qubits", ctypes.c_int), ("skip_shots", ctypes.c_int),
                ("blocked", ctypes.c_int)]


class _ExecResults(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_int), ("precision", ctypes.c_int),
                ("total_shots", ctypes.c_int), ("num_workers_used", ctypes.c_int),
                ("outcomes", ctypes.POINTER(ctypes.c_uint64)),
                ("amplitudes", ctypes.POINTER(ctypes.c_double)),
                ("probabilities", ctypes.POINTER(ctypes.c_double)),
                ("expect_z", ctypes.POINTER(ctypes.c_double)),
                ("norm", ctypes.c_double), ("device_seconds", ctypes.c_double),
                ("wall_seconds", ctypes.c_double)]


class _BatchItem(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_int), ("precision", ctypes.c_int),
                ("gates", ctypes.c_void_p), ("num_gates", ctypes.c_int),
                ("observables", ctypes.c_int), ("num_shots", ctypes.c_int),
                ("draws", ctypes.POINTER(ctypes.c_double)),
                ("norm", ctypes.POINTER(ctypes.c_double)),
                ("expect_z", ctypes.POINTER(ctypes.c_double)),
                ("zz_chain", ctypes.POINTER(ctypes.c_double)),
                ("expect_x", ctypes.POINTER(ctypes.c_double)),
                ("outcomes", ctypes.POINTER(ctypes.c_uint64)),
                ("num_pauli_terms", ctypes.c_int),
                ("pauli_terms", ctypes.c_void_p),
                ("pauli_expect", ctypes.POINTER(ctypes.c_double)),
                ("pauli_energy", ctypes.POINTER(ctypes.c_double))]


class _PauliTerm(ctypes.Structure):   # moonlab_mi355x_pauli_term_t (pauli_term_t, vqe.h:101-105)
    _fields_ = [("coefficient", ctypes.c_double), ("pauli_string", ctypes.c_char_p)]


class _Vqe(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_int), ("num_layers", ctypes.c_int),
                ("angles", ctypes.POINTER(ctypes.c_double)), ("entangler", ctypes.c_int),
                ("hf_reference", ctypes.c_uint64), ("num_terms", ctypes.c_int),
                ("terms", ctypes.POINTER(_PauliTerm)), ("constant", ctypes.c_double)]


ENT_CZ, ENT_CNOT = 0, 1


def _pauli_terms(terms):
    """(coefficient, string) pairs -> a moonlab_mi355x_pauli_term_t array and
    the objects it points into."""
    arr = (_PauliTerm * max(len(terms), 1))()
    strs = [t[1].encode() for t in terms]
    for i, (c, _) in enumerate(terms):
        arr[i].coefficient = float(c)
        arr[i].pauli_string = strs[i]
    return arr, strs


class _BatchStats(ctypes.Structure):
    _fields_ = [("device_seconds", ctypes.c_double), ("wall_seconds", ctypes.c_double),
                ("workgroups", ctypes.c_int), ("tile_bits", ctypes.c_int),
                ("total_passes", ctypes.c_int), ("total_ops", ctypes.c_int),
                ("graph_replayed", ctypes.c_int), ("reserved", ctypes.c_int)]


class _Hea(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_int), ("num_layers", ctypes.c_int),
                ("angles", ctypes.POINTER(ctypes.c_double)), ("h", ctypes.c_double)]


_HEA_DTYPE = np.dtype([("num_qubits", np.int32), ("num_layers", np.int32), ("angles", np.uintp),
                       ("h", np.float64)])
assert _HEA_DTYPE.itemsize == ctypes.sizeof(_Hea)

OBS_Z, OBS_ZZ, OBS_X = 1, 2, 4


class _ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("batch_max_qubits", ctypes.c_int),
                ("max_batch", ctypes.c_int), ("batch_engines", ctypes.c_int),
                ("large_workers", ctypes.c_int), ("sequential", ctypes.c_int), ("large_path", ctypes.c_int)]


class _Request(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_int), ("precision", ctypes.c_int),
                ("gates", ctypes.c_void_p), ("num_gates", ctypes.c_int),
                ("num_shots", ctypes.c_int), ("rng_seed", ctypes.c_uint64),
                ("want_expect_z", ctypes.c_int), ("want_probabilities", ctypes.c_int)]


class _Response(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("num_qubits", ctypes.c_int), ("num_shots", ctypes.c_int),
                ("outcomes", ctypes.POINTER(ctypes.c_uint64)),
                ("expect_z", ctypes.POINTER(ctypes.c_double)), ("norm", ctypes.c_double),
                ("submit_time", ctypes.c_double), ("start_time", ctypes.c_double),
                ("finish_time", ctypes.c_double), ("path", ctypes.c_int),
                ("probabilities", ctypes.POINTER(ctypes.c_double))]


class _ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("batched_requests", ctypes.c_uint64),
                ("large_requests", ctypes.c_uint64), ("peak_bytes_admitted", ctypes.c_uint64),
                ("largest_batch", ctypes.c_int)]

_lib: Optional[ctypes.CDLL] = None


def _declare(L: ctypes.CDLL) -> None:
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    u32, u64 = ctypes.c_uint32, ctypes.c_uint64
    u64p = ctypes.POINTER(ctypes.c_uint64)
    sig = {
        # job API
        "moonlab_job_create": ([ip], vp),
        "moonlab_job_free": ([vp], None),
        "moonlab_job_add_gate": ([vp, ip, ip, ip, dp], ip),
        "moonlab_job_set_num_shots": ([vp, ip], ip),
        "moonlab_job_set_num_workers": ([vp, ip], ip),
        "moonlab_job_set_rng_seed": ([vp, u64], ip),
        "moonlab_job_num_qubits": ([vp], ip),
        "moonlab_job_num_gates": ([vp], ip),
        "moonlab_job_num_shots": ([vp], ip),
        "moonlab_job_num_workers": ([vp], ip),
        "moonlab_scheduler_run": ([vp, ctypes.POINTER(_JobResults)], ip),
        "moonlab_scheduler_run_mpi": ([vp, ctypes.POINTER(_JobResults), vp], ip),
        "moonlab_job_results_free": ([ctypes.POINTER(_JobResults)], None),
        "moonlab_job_to_json": ([vp, ctypes.c_char_p, ctypes.c_size_t], ip),
        "moonlab_register_backend": ([ctypes.POINTER(_Backend)], ip),
        "moonlab_unregister_backend": ([ctypes.c_char_p], ip),
        "moonlab_find_backend": ([ctypes.c_char_p], ctypes.POINTER(_Backend)),
        "moonlab_num_backends": ([], ip),
        "moonlab_list_backends": ([ctypes.POINTER(ctypes.c_char_p), ip], ip),
        "moonlab_job_set_backend": ([vp, ctypes.c_char_p], ip),
        "moonlab_job_backend": ([vp], ctypes.c_char_p),
        "moonlab_scheduler_set_completion_hook": ([COMPLETION_HOOK_FN, vp], ip),
        "moonlab_scheduler_set_request_context": ([ctypes.c_char_p, ctypes.c_char_p], None),
        "moonlab_scheduler_current_tenant_id": ([], ctypes.c_char_p),
        "moonlab_scheduler_current_request_id": ([], ctypes.c_char_p),
        "moonlab_scheduler_fire_completion_hook": ([vp, ctypes.POINTER(_JobResults),
                                                    ctypes.c_char_p], None),
        # circuit API
        "moonlab_qgtl_circuit_create": ([ip], vp),
        "moonlab_qgtl_circuit_free": ([vp], None),
        "moonlab_qgtl_add_gate": ([vp, ip, ip, ip, dp], ip),
        "moonlab_qgtl_execute": ([vp, ctypes.POINTER(_QgtlOpts), ctypes.POINTER(_QgtlResults)], ip),
        "moonlab_qgtl_results_free": ([ctypes.POINTER(_QgtlResults)], None),
        "moonlab_qgtl_circuit_num_qubits": ([vp], ip),
        "moonlab_qgtl_circuit_num_gates": ([vp], ip),
        "moonlab_mi355x_circuit_add_gate_ex": ([vp, ip, ctypes.POINTER(ctypes.c_int), ip,
                                                ctypes.POINTER(ctypes.c_double), ip], ip),
        "moonlab_qgtl_circuit_serialize": ([vp, ctypes.c_char_p, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_size_t)], ip),
        "moonlab_qgtl_circuit_deserialize": ([ctypes.c_char_p, ctypes.c_size_t,
                                              ctypes.POINTER(ctypes.c_int)], vp),
        "moonlab_qgtl_circuit_save": ([vp, ctypes.c_char_p], ip),
        "moonlab_qgtl_circuit_load": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)], vp),
        # MI355X additions
        "moonlab_mi355x_register_backend": ([], ip),
        "moonlab_mi355x_execute_ex": ([vp, ctypes.POINTER(_ExecOpts),
                                       ctypes.POINTER(_ExecResults)], ip),
        "moonlab_mi355x_results_free": ([ctypes.POINTER(_ExecResults)], None),
        "moonlab_mi355x_set_state_cache_bytes": ([u64], ip),
        "moonlab_mi355x_batch_create": ([], vp),
        "moonlab_mi355x_batch_free": ([vp], None),
        "moonlab_mi355x_batch_set_config": ([vp, ip, ip], ip),
        "moonlab_mi355x_batch_set_graphs": ([vp, ip], ip),
        "moonlab_mi355x_batch_run": ([vp, ctypes.POINTER(_BatchItem), ip, ctypes.POINTER(_BatchStats)], ip),
        "moonlab_mi355x_batch_autotune": ([vp, ctypes.POINTER(_BatchItem), ip, ctypes.POINTER(_BatchStats)], ip),
        "moonlab_mi355x_server_create": ([ctypes.POINTER(_ServerConfig)], vp),
        "moonlab_mi355x_server_free": ([vp], None),
        "moonlab_mi355x_server_now": ([vp], ctypes.c_double),
        "moonlab_mi355x_server_submit": ([vp, ctypes.POINTER(_Request), ctypes.POINTER(ctypes.c_uint64)], ip),
        "moonlab_mi355x_server_wait": ([vp, ctypes.c_uint64, ctypes.POINTER(_Response), ctypes.c_double], ip),
        "moonlab_mi355x_response_free": ([ctypes.POINTER(_Response)], None),
        "moonlab_mi355x_server_stats": ([vp, ctypes.POINTER(_ServerStats)], ip),
        "moonlab_mi355x_server_handle_frame": ([vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint64,
                                                ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t)],
                                               ip),
        "moonlab_mi355x_server_handle_frame_ex": ([vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint64,
                                                   ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p),
                                                   ctypes.POINTER(ctypes.c_size_t)], ip),
        "moonlab_mi355x_hea_tfim_energies": ([vp, ctypes.POINTER(_Hea), ip, ip, dp, ctypes.POINTER(_BatchStats)], ip),
        "moonlab_mi355x_hea_tfim_gradients": ([vp, ctypes.POINTER(_Hea), ip, ip, dp, dp,
                                               ctypes.POINTER(_BatchStats)], ip),
        "moonlab_mi355x_vqe_energies": ([vp, ctypes.POINTER(_Vqe), ip, ip, dp, ctypes.POINTER(_BatchStats)], ip),
        "moonlab_mi355x_vqe_gradients": ([vp, ctypes.POINTER(_Vqe), ip, ip, dp, dp, ctypes.POINTER(_BatchStats)],
                                         ip),
        "moonlab_mi355x_vqe_qgt": ([vp, ctypes.POINTER(_Vqe), ip, dp, dp, ctypes.POINTER(_BatchStats)], ip),
        "moonlab_mi355x_vqe_natural_gradient": ([vp, ctypes.POINTER(_Vqe), ip, dp, ctypes.c_double, dp,
                                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_BatchStats)], ip),
        # state-vector layer
        "moonlab_hip_device_count": ([], ip),
        "moonlab_hip_last_error": ([], ctypes.c_char_p),
        "moonlab_hip_set_gate_tuning": ([ip, ip], ip),
        "moonlab_hip_set_blocked_lookahead": ([ip], ip),
        "moonlab_hip_state_create": ([u32, ctypes.POINTER(vp)], ip),
        "moonlab_hip_state_create_ex": ([u32, ip, ctypes.POINTER(vp)], ip),
        "moonlab_hip_state_free": ([vp], None),
        "moonlab_hip_state_n_qubits": ([vp], u32),
        "moonlab_hip_state_dim": ([vp], u64),
        "moonlab_hip_state_precision": ([vp], ip),
        "moonlab_hip_state_reset": ([vp], ip),
        "moonlab_hip_apply_1q": ([vp, u32, dp], ip),
        "moonlab_hip_apply_cnot": ([vp, u32, u32], ip),
        "moonlab_hip_apply_2q": ([vp, u32, u32, dp], ip),
        "moonlab_hip_apply_x": ([vp, u32], ip),
        "moonlab_hip_apply_h": ([vp, u32], ip),
        "moonlab_hip_apply_mcx": ([vp, u64, u32], ip),
        "moonlab_hip_apply_mcz": ([vp, u64], ip),
        "moonlab_hip_apply_fredkin": ([vp, u32, u32, u32], ip),
        "moonlab_hip_apply_gate": ([vp, vp], ip),
        "moonlab_hip_apply_gates": ([vp, vp, ip], ip),
        "moonlab_hip_apply_diag_1q": ([vp, u32, dp], ip),
        "moonlab_hip_apply_controlled_1q": ([vp, u64, u32, dp], ip),
        "moonlab_hip_apply_dense": ([vp, ip, ctypes.POINTER(ctypes.c_int), dp], ip),
        "moonlab_hip_apply_gates_fused": ([vp, vp, ip, ip], ip),
        "moonlab_hip_fusion_plan": ([vp, ip, ip, u64p, ctypes.POINTER(ctypes.c_int), dp, ip], ip),
        "moonlab_hip_apply_gates_blocked": ([vp, vp, ip], ip),
        "moonlab_hip_expect_pauli": ([vp, u64, u64, dp], ip),
        "moonlab_hip_blocked_pass_count": ([vp, ip, ip, ip], ip),
        "moonlab_hip_set_blocked_tuning": ([ip, ip], ip),
        "moonlab_hip_set_blocked_engine": ([ip], ip),
        "moonlab_hip_prob_z": ([vp, u32, dp], ip),
        "moonlab_hip_norm": ([vp, dp], ip),
        "moonlab_hip_expect_z_all": ([vp, dp, dp], ip),
        "moonlab_hip_expect_z_all_noalloc": ([vp, u32, ip, vp, vp], ip),
        "moonlab_hip_expect_x": ([vp, u32, dp], ip),
        "moonlab_hip_expect_y": ([vp, u32, dp], ip),
        "moonlab_hip_correlation_zz": ([vp, u32, u32, dp], ip),
        "moonlab_hip_probabilities": ([vp, dp], ip),
        "moonlab_hip_inner_product": ([vp, vp, dp, dp], ip),
        "moonlab_hip_normalize": ([vp], ip),
        "moonlab_hip_measure_qubit": ([vp, ctypes.c_uint32, ctypes.c_double, ctypes.POINTER(ctypes.c_int)], ip),
        "moonlab_hip_measure_partial": ([vp, ctypes.POINTER(ctypes.c_int), ip, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_uint64)], ip),
        "moonlab_hip_measure_all": ([vp, ctypes.c_double, ctypes.POINTER(ctypes.c_uint64)], ip),
        "moonlab_hip_sample": ([vp, dp, ip, u64p], ip),
        "moonlab_hip_state_copy_to_host": ([vp, dp], ip),
        "moonlab_hip_state_copy_from_host": ([vp, dp], ip),
        "moonlab_hip_state_read": ([vp, u64, u64, vp], ip),
        "moonlab_hip_state_device_ptr": ([vp], vp),
        "moonlab_hip_state_copy": ([vp, vp], ip),
        "moonlab_hip_set_streaming_threshold": ([u64], ip),
        "moonlab_hip_state_stream": ([vp], vp),
        "moonlab_hip_synchronize": ([vp], ip),
        "moonlab_hip_timer_begin": ([vp, ip], ip),
        "moonlab_hip_timer_mark": ([vp], ip),
        "moonlab_hip_timer_elapsed": ([vp, ctypes.POINTER(ctypes.c_float), ip], ip),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def lib() -> ctypes.CDLL:
    """Load the engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m moonlab_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def device_count() -> int:
    return int(lib().moonlab_hip_device_count())


def last_error() -> str:
    return lib().moonlab_hip_last_error().decode()


def _check_sched(rc: int, what: str) -> None:
    if rc != MOONLAB_SCHED_OK:
        raise SchedulerError(f"{what}: status {rc}")


def _check_hip(rc: int, what: str) -> None:
    if rc != 0:
        raise EngineError(f"{what}: status {rc}: {last_error()}")


# ---------------------------------------------------------------------------
# Job API (mirrors bindings/python/moonlab/scheduler.py)
# ---------------------------------------------------------------------------
@dataclass
class JobResults:
    num_qubits: int
    total_shots: int
    outcomes: List[int]
    num_workers_used: int
    worker_seconds: List[float]


@dataclass
class ExecResults:
    num_qubits: int
    precision: int
    outcomes: np.ndarray
    amplitudes: Optional[np.ndarray]
    probabilities: Optional[np.ndarray]
    expect_z: Optional[np.ndarray]
    norm: float
    device_seconds: float
    wall_seconds: float
    num_workers_used: int


class Job:
    """A moonlab job (scheduler.h:45-210)."""

    def __init__(self, num_qubits: int) -> None:
        self._h = lib().moonlab_job_create(int(num_qubits))
        if not self._h:
            raise SchedulerError(f"moonlab_job_create({num_qubits}) failed")

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.moonlab_job_free(h)
            self._h = None

    @property
    def handle(self) -> int:
        return self._h

    @property
    def num_qubits(self) -> int:
        return lib().moonlab_job_num_qubits(self._h)

    @property
    def num_gates(self) -> int:
        return lib().moonlab_job_num_gates(self._h)

    @property
    def num_shots(self) -> int:
        return lib().moonlab_job_num_shots(self._h)

    @property
    def num_workers(self) -> int:
        return lib().moonlab_job_num_workers(self._h)

    def add_gate(self, gate: int, target: int, control: int = -1,
                 theta: Optional[float] = None) -> "Job":
        p = (ctypes.c_double * 1)(theta) if theta is not None else None
        _check_sched(lib().moonlab_job_add_gate(self._h, int(gate), target, control, p), "add_gate")
        return self

    def set_num_shots(self, n: int) -> "Job":
        _check_sched(lib().moonlab_job_set_num_shots(self._h, n), "set_num_shots")
        return self

    def set_num_workers(self, n: int) -> "Job":
        _check_sched(lib().moonlab_job_set_num_workers(self._h, n), "set_num_workers")
        return self

    def set_rng_seed(self, seed: int) -> "Job":
        _check_sched(lib().moonlab_job_set_rng_seed(self._h, seed), "set_rng_seed")
        return self

    def set_backend(self, name: Optional[str]) -> "Job":
        _check_sched(lib().moonlab_job_set_backend(self._h, name.encode() if name else None),
                     "set_backend")
        return self

    @property
    def backend(self) -> Optional[str]:
        b = lib().moonlab_job_backend(self._h)
        return b.decode() if b else None

    def run_raw(self):
        """moonlab_scheduler_run: (status, results or None)."""
        res = _JobResults()
        rc = lib().moonlab_scheduler_run(self._h, ctypes.byref(res))
        if rc != MOONLAB_SCHED_OK:
            return rc, None
        out = JobResults(res.num_qubits, res.total_shots,
                         [res.outcomes[i] for i in range(res.total_shots)],
                         res.num_workers_used,
                         [res.worker_seconds[i] for i in range(res.num_workers_used)])
        lib().moonlab_job_results_free(ctypes.byref(res))
        return rc, out

    def execute(self) -> JobResults:
        rc, out = self.run_raw()
        _check_sched(rc, "moonlab_scheduler_run")
        return out

    def execute_ex(self, precision: int = C128, amplitudes: bool = False,
                   probabilities: bool = False, expect_z: bool = False,
                   fusion_max_qubits: int = 0, skip_shots: bool = False,
                   blocked: bool = False) -> ExecResults:
        o = _ExecOpts(precision, int(amplitudes), int(probabilities), int(expect_z),
                      fusion_max_qubits, int(skip_shots), int(blocked))
        r = _ExecResults()
        _check_sched(lib().moonlab_mi355x_execute_ex(self._h, ctypes.byref(o), ctypes.byref(r)),
                     "moonlab_mi355x_execute_ex")
        n = r.num_qubits
        dim = 1 << n
        try:
            outc = np.ctypeslib.as_array(r.outcomes, (r.total_shots,)).copy() \
                if r.total_shots else np.zeros(0, np.uint64)
            amps = np.ctypeslib.as_array(r.amplitudes, (2 * dim,)).copy().view(np.complex128) \
                if r.amplitudes else None
            probs = np.ctypeslib.as_array(r.probabilities, (dim,)).copy() if r.probabilities else None
            z = np.ctypeslib.as_array(r.expect_z, (n,)).copy() if r.expect_z else None
            return ExecResults(n, r.precision, outc, amps, probs, z, r.norm, r.device_seconds,
                               r.wall_seconds, r.num_workers_used)
        finally:
            lib().moonlab_mi355x_results_free(ctypes.byref(r))

    def to_json(self) -> str:
        need = lib().moonlab_job_to_json(self._h, None, 0)
        if need < 0:
            raise SchedulerError(f"to_json: status {need}")
        buf = ctypes.create_string_buffer(need + 1)
        lib().moonlab_job_to_json(self._h, buf, need + 1)
        return buf.value.decode()


def list_backends() -> List[str]:
    n = lib().moonlab_num_backends()
    arr = (ctypes.c_char_p * max(n, 1))()
    got = lib().moonlab_list_backends(arr, n)
    return [arr[i].decode() for i in range(got)]


def register_mi355x_backend() -> None:
    _check_sched(lib().moonlab_mi355x_register_backend(), "moonlab_mi355x_register_backend")


# ---------------------------------------------------------------------------
# Circuit API (mirrors bindings/python/moonlab/qgtl.py)
# ---------------------------------------------------------------------------
@dataclass
class CircuitResults:
    num_qubits: int
    outcomes: np.ndarray
    probabilities: Optional[np.ndarray]


class Circuit:
    def __init__(self, num_qubits: int = 0, _handle=None) -> None:
        self._h = _handle if _handle is not None else lib().moonlab_qgtl_circuit_create(int(num_qubits))
        if not self._h:
            raise QGTLError(f"moonlab_qgtl_circuit_create({num_qubits}) failed")

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.moonlab_qgtl_circuit_free(h)
            self._h = None

    @property
    def num_qubits(self) -> int:
        return lib().moonlab_qgtl_circuit_num_qubits(self._h)

    @property
    def num_gates(self) -> int:
        return lib().moonlab_qgtl_circuit_num_gates(self._h)

    def add_gate_raw(self, gate: int, target: int, control: int = -1,
                     theta: Optional[float] = None) -> int:
        p = (ctypes.c_double * 1)(theta) if theta is not None else None
        return lib().moonlab_qgtl_add_gate(self._h, int(gate), target, control, p)

    def add_gate(self, gate: int, target: int, control: int = -1,
                 theta: Optional[float] = None) -> "Circuit":
        rc = self.add_gate_raw(gate, target, control, theta)
        if rc:
            raise QGTLError(f"add_gate: status {rc}")
        return self

    def add_gate_ex(self, gate: int, qubits, params=()) -> "Circuit":
        """A gate of codes 14..24 (moonlab_mi355x_circuit_add_gate_ex): qubits
        in target, control, extra order; params theta (or theta, phi, lambda)."""
        qs = (ctypes.c_int * len(qubits))(*[int(q) for q in qubits])
        ps = (ctypes.c_double * max(len(params), 1))(*[float(x) for x in params])
        rc = lib().moonlab_mi355x_circuit_add_gate_ex(self._h, int(gate), qs, len(qubits), ps, len(params))
        if rc:
            raise QGTLError(f"add_gate_ex: status {rc}")
        return self

    def execute(self, num_shots: int = 0, rng_seed: int = 0,
                return_probabilities: bool = False) -> CircuitResults:
        o = _QgtlOpts(num_shots, rng_seed, int(return_probabilities))
        r = _QgtlResults()
        rc = lib().moonlab_qgtl_execute(self._h, ctypes.byref(o), ctypes.byref(r))
        if rc:
            raise QGTLError(f"moonlab_qgtl_execute: status {rc}")
        try:
            n = r.num_qubits
            outc = np.ctypeslib.as_array(r.outcomes, (r.num_shots,)).copy() \
                if r.outcomes else np.zeros(0, np.uint64)
            probs = np.ctypeslib.as_array(r.probabilities, (1 << n,)).copy() \
                if r.probabilities else None
            return CircuitResults(n, outc, probs)
        finally:
            lib().moonlab_qgtl_results_free(ctypes.byref(r))

    def serialize(self) -> str:
        need = ctypes.c_size_t(0)
        rc = lib().moonlab_qgtl_circuit_serialize(self._h, None, 0, ctypes.byref(need))
        if rc:
            raise QGTLError(f"serialize: status {rc}")
        buf = ctypes.create_string_buffer(need.value + 1)
        rc = lib().moonlab_qgtl_circuit_serialize(self._h, buf, need.value + 1, None)
        if rc:
            raise QGTLError(f"serialize: status {rc}")
        return buf.value.decode()

    @classmethod
    def deserialize(cls, text: str) -> "Circuit":
        st = ctypes.c_int(0)
        data = text.encode()
        h = lib().moonlab_qgtl_circuit_deserialize(data, len(data), ctypes.byref(st))
        if not h:
            raise QGTLError(f"deserialize: status {st.value}")
        return cls(_handle=h)


# ---------------------------------------------------------------------------
# State-vector layer (moonlab_hip_*)
# ---------------------------------------------------------------------------
class StateVector:
    """A device-resident state of 2^n amplitudes (complex128 or complex64)."""

    def __init__(self, num_qubits: int, precision: int = C128) -> None:
        h = ctypes.c_void_p()
        _check_hip(lib().moonlab_hip_state_create_ex(num_qubits, precision, ctypes.byref(h)),
                   "state_create")
        self._h = h.value
        self.n = num_qubits
        self.precision = precision

    def free(self) -> None:
        if getattr(self, "_h", None) and _lib is not None:
            _lib.moonlab_hip_state_free(self._h)
            self._h = None

    def __del__(self) -> None:
        self.free()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    @property
    def handle(self):
        return self._h

    def reset(self) -> "StateVector":
        _check_hip(lib().moonlab_hip_state_reset(self._h), "reset")
        return self

    def apply(self, gates, fuse: int = 0, blocked: bool = False) -> "StateVector":
        g = gate_array(gates)
        if len(g):
            if blocked:
                _check_hip(lib().moonlab_hip_apply_gates_blocked(self._h, g.ctypes.data, len(g)),
                           "apply_gates_blocked")
            elif fuse:
                _check_hip(lib().moonlab_hip_apply_gates_fused(self._h, g.ctypes.data, len(g), fuse),
                           "apply_gates_fused")
            else:
                _check_hip(lib().moonlab_hip_apply_gates(self._h, g.ctypes.data, len(g)), "apply_gates")
        return self

    def apply_1q(self, q: int, m) -> "StateVector":
        m = np.ascontiguousarray(np.asarray(m, np.complex128).reshape(2, 2)).view(np.float64)
        _check_hip(lib().moonlab_hip_apply_1q(self._h, q, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "apply_1q")
        return self

    def apply_2q(self, q0: int, q1: int, m) -> "StateVector":
        m = np.ascontiguousarray(np.asarray(m, np.complex128).reshape(4, 4)).view(np.float64)
        _check_hip(lib().moonlab_hip_apply_2q(self._h, q0, q1,
                                              m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "apply_2q")
        return self

    def apply_dense(self, qubits: Sequence[int], m) -> "StateVector":
        k = len(qubits)
        m = np.ascontiguousarray(np.asarray(m, np.complex128).reshape(1 << k, 1 << k)).view(np.float64)
        q = (ctypes.c_int * k)(*qubits)
        _check_hip(lib().moonlab_hip_apply_dense(self._h, k, q,
                                                 m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "apply_dense")
        return self

    def apply_mcx(self, control_mask: int, target: int) -> "StateVector":
        _check_hip(lib().moonlab_hip_apply_mcx(self._h, control_mask, target), "apply_mcx")
        return self

    def apply_mcz(self, mask: int) -> "StateVector":
        _check_hip(lib().moonlab_hip_apply_mcz(self._h, mask), "apply_mcz")
        return self

    def apply_controlled_1q(self, control_mask: int, target: int, m) -> "StateVector":
        m = np.ascontiguousarray(np.asarray(m, np.complex128).reshape(2, 2)).view(np.float64)
        _check_hip(lib().moonlab_hip_apply_controlled_1q(
            self._h, control_mask, target, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
            "apply_controlled_1q")
        return self

    def copy_from(self, other: "StateVector") -> "StateVector":
        _check_hip(lib().moonlab_hip_state_copy(self._h, other._h), "state_copy")
        return self

    def synchronize(self) -> None:
        _check_hip(lib().moonlab_hip_synchronize(self._h), "synchronize")

    def amplitudes(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=np.complex128)
        _check_hip(lib().moonlab_hip_state_copy_to_host(
            self._h, out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "copy_to_host")
        return out

    def load(self, amps: np.ndarray) -> "StateVector":
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        assert a.size == 1 << self.n
        _check_hip(lib().moonlab_hip_state_copy_from_host(
            self._h, a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "copy_from_host")
        return self

    def read(self, offset: int, count: int) -> np.ndarray:
        dt = np.complex64 if self.precision == C64 else np.complex128
        out = np.empty(count, dtype=dt)
        _check_hip(lib().moonlab_hip_state_read(self._h, offset, count, out.ctypes.data), "state_read")
        return out

    def expect_z_all(self):
        z = np.empty(self.n, dtype=np.float64)
        nrm = ctypes.c_double(0.0)
        _check_hip(lib().moonlab_hip_expect_z_all(
            self._h, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(nrm)), "expect_z_all")
        return z, nrm.value

    def prob_one(self, q: int) -> float:
        v = ctypes.c_double()
        _check_hip(lib().moonlab_hip_prob_z(self._h, q, ctypes.byref(v)), "prob_z")
        return v.value

    def norm(self) -> float:
        v = ctypes.c_double()
        _check_hip(lib().moonlab_hip_norm(self._h, ctypes.byref(v)), "norm")
        return v.value

    def expect_x(self, q: int) -> float:
        v = ctypes.c_double()
        _check_hip(lib().moonlab_hip_expect_x(self._h, q, ctypes.byref(v)), "expect_x")
        return v.value

    def expect_y(self, q: int) -> float:
        v = ctypes.c_double()
        _check_hip(lib().moonlab_hip_expect_y(self._h, q, ctypes.byref(v)), "expect_y")
        return v.value

    def expect_pauli(self, pauli: str) -> float:
        """<P> for a Pauli string, pauli[q] in 'IXYZ' acting on qubit q (the
        reference's pauli_term_t convention, vqe.c:2185-2199)."""
        x = z = 0
        for q, c in enumerate(pauli.upper()):
            if c not in "IXYZ":
                raise ValueError(f"bad Pauli letter {c!r}")
            if c in "XY":
                x |= 1 << q
            if c in "YZ":
                z |= 1 << q
        out = ctypes.c_double()
        _check_hip(lib().moonlab_hip_expect_pauli(self._h, x, z, ctypes.byref(out)), "expect_pauli")
        return out.value

    def correlation_zz(self, q1: int, q2: int) -> float:
        v = ctypes.c_double()
        _check_hip(lib().moonlab_hip_correlation_zz(self._h, q1, q2, ctypes.byref(v)), "correlation_zz")
        return v.value

    def probabilities(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=np.float64)
        _check_hip(lib().moonlab_hip_probabilities(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "probabilities")
        return out

    def inner(self, other: "StateVector") -> complex:
        re, im = ctypes.c_double(), ctypes.c_double()
        _check_hip(lib().moonlab_hip_inner_product(self._h, other._h, ctypes.byref(re), ctypes.byref(im)),
                   "inner_product")
        return complex(re.value, im.value)

    def normalize(self) -> "StateVector":
        _check_hip(lib().moonlab_hip_normalize(self._h), "normalize")
        return self

    def measure_qubit(self, qubit: int, u: float) -> int:
        """Collapse one qubit (measurement_single_qubit): 0 or 1, or -1 when
        the drawn outcome carries no weight."""
        r = ctypes.c_int(0)
        _check_hip(lib().moonlab_hip_measure_qubit(self._h, qubit, u, ctypes.byref(r)), "measure_qubit")
        return r.value

    def measure_partial(self, qubits, u: float) -> int:
        """Collapse the given qubits (measurement_partial); bit j of the
        result is qubits[j]."""
        qs = (ctypes.c_int * len(qubits))(*[int(q) for q in qubits])
        r = ctypes.c_uint64(0)
        _check_hip(lib().moonlab_hip_measure_partial(self._h, qs, len(qubits), u, ctypes.byref(r)),
                   "measure_partial")
        return r.value

    def measure_all(self, u: float) -> int:
        """Collapse every qubit (measurement_all_qubits)."""
        r = ctypes.c_uint64(0)
        _check_hip(lib().moonlab_hip_measure_all(self._h, u, ctypes.byref(r)), "measure_all")
        return r.value

    def sample(self, u) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(u.size, dtype=np.uint64)
        _check_hip(lib().moonlab_hip_sample(self._h, u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            u.size, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))),
                   "sample")
        return out

    # stream timing
    def timer_begin(self, capacity: int) -> None:
        _check_hip(lib().moonlab_hip_timer_begin(self._h, capacity), "timer_begin")

    def timer_mark(self) -> None:
        _check_hip(lib().moonlab_hip_timer_mark(self._h), "timer_mark")

    def timer_elapsed(self, max_gaps: int) -> np.ndarray:
        out = np.zeros(max_gaps, dtype=np.float32)
        n = lib().moonlab_hip_timer_elapsed(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                            max_gaps)
        if n < 0:
            raise EngineError(f"timer_elapsed failed: {last_error()}")
        return out[:n]


@dataclass
class FusedBlock:
    qubits: List[int]
    num_gates: int
    matrix: np.ndarray     # 2^k x 2^k complex128, local bit j = qubits[j]


def fusion_plan(gates, kmax: int = 5) -> List[FusedBlock]:
    """The blocks StateVector.apply(gates, fuse=kmax) executes, in order,
    with their FP64 unitaries (host only, no device needed)."""
    g = gate_array(gates)
    L = lib()
    nb = L.moonlab_hip_fusion_plan(g.ctypes.data, len(g), kmax, None, None, None, 0)
    if nb < 0:
        raise EngineError(f"fusion_plan failed ({nb}): {last_error()}")
    masks = np.zeros(max(nb, 1), dtype=np.uint64)
    counts = np.zeros(max(nb, 1), dtype=np.int32)
    mats = np.zeros((max(nb, 1), 2048), dtype=np.float64)
    L.moonlab_hip_fusion_plan(g.ctypes.data, len(g), kmax,
                              masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                              counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              mats.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nb)
    out = []
    for i in range(nb):
        qs = [q for q in range(64) if (int(masks[i]) >> q) & 1]
        d = 1 << len(qs)
        m = mats[i, :2 * d * d].view(np.complex128).reshape(d, d).copy()
        out.append(FusedBlock(qs, int(counts[i]), m))
    return out


def blocked_pass_count(gates, n: int, precision: int = C128) -> int:
    """HBM passes StateVector.apply(gates, blocked=True) makes on n qubits."""
    g = gate_array(gates)
    k = lib().moonlab_hip_blocked_pass_count(g.ctypes.data, len(g), n, precision)
    if k < 0:
        raise EngineError(f"blocked_pass_count failed ({k})")
    return k


def set_blocked_tuning(tile_bits: int = 0, threads: int = 0) -> None:
    """Tile (11, 12 or 13 qubits; -1: by precision, the default) and LDS-engine
    workgroup size (256 or 512) of the cache-blocked path; 0 keeps a value."""
    _check_hip(lib().moonlab_hip_set_blocked_tuning(tile_bits, threads), "set_blocked_tuning")


def set_blocked_engine(engine: str) -> None:
    """Engine of the cache-blocked path for states of >= 12 qubits:
    "register" (each thread keeps 16 tile amplitudes in registers; default;
    "register16" is the same), "register8" (8 per thread, twice the waves)
    or "lds" (one LDS sweep per op or register segment)."""
    code = {"register": 0, "lds": 1, "register16": 2, "register8": 3}[engine]
    _check_hip(lib().moonlab_hip_set_blocked_engine(code), "set_blocked_engine")


def set_streaming_threshold(nbytes: int) -> None:
    _check_hip(lib().moonlab_hip_set_streaming_threshold(nbytes), "set_streaming_threshold")


def set_blocked_lookahead(windows: int = 8) -> None:
    """Seed ops per pass and phase of the blocked planner's lookahead (0: greedy)."""
    _check_hip(lib().moonlab_hip_set_blocked_lookahead(windows), "set_blocked_lookahead")


def set_gate_tuning(pairs_per_thread: int = 0, block_threads: int = 0) -> None:
    _check_hip(lib().moonlab_hip_set_gate_tuning(pairs_per_thread, block_threads), "set_gate_tuning")


# ---------------------------------------------------------------------------
# Batched small circuits (one persistent launch)
# ---------------------------------------------------------------------------
@dataclass
class BatchResult:
    norm: float
    expect_z: Optional[np.ndarray]
    zz_chain: Optional[np.ndarray]
    expect_x: Optional[np.ndarray]
    outcomes: Optional[np.ndarray]
    pauli_expect: Optional[np.ndarray] = None
    pauli_energy: Optional[float] = None


class HeaPack:
    """Hardware-efficient-ansatz instances packed once in the layout of
    moonlab_mi355x_hea_t: one flat float64 angle buffer (`angles`; instance
    i's [layers, n, 2] angles start at offsets[i]) and the descriptor array
    pointing into it.  A parameter loop writes new angles into `angles` in
    place (`pack.angles[:] = theta`) and passes the pack to
    Batch.hea_tfim / hea_tfim_gradients again, instead of a list of
    instances that each call packs anew (8 ms for 4096 instances)."""

    def __init__(self, instances) -> None:
        self._ptr, (self._arr, self._flat) = Batch._hea_array(instances)
        self.count = len(instances)
        self.shapes = [(int(n), int(L)) for (n, L, _, _) in instances]
        sizes = np.array([2 * n * L for n, L in self.shapes], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]) if self.count else np.zeros(0, np.int64)

    @property
    def angles(self) -> np.ndarray:
        """The flat angle buffer the descriptors point into (write in place)."""
        return self._flat


class Batch:
    """moonlab_mi355x_batch_*: many circuits in one persistent launch."""

    def __init__(self) -> None:
        self._h = lib().moonlab_mi355x_batch_create()
        if not self._h:
            raise EngineError("moonlab_mi355x_batch_create failed (no HIP device?)")
        self.last_stats = None

    def free(self) -> None:
        if getattr(self, "_h", None) and _lib is not None:
            _lib.moonlab_mi355x_batch_free(self._h)
            self._h = None

    def __del__(self) -> None:
        self.free()

    def set_graphs(self, enable: bool) -> None:
        _check_sched(lib().moonlab_mi355x_batch_set_graphs(self._h, int(enable)), "batch_set_graphs")

    def set_config(self, workgroups_per_cu: int = 0, tile_bits: int = 0) -> None:
        _check_sched(lib().moonlab_mi355x_batch_set_config(self._h, workgroups_per_cu, tile_bits), "set_config")

    @staticmethod
    def _stats(st):
        return {"device_seconds": st.device_seconds, "wall_seconds": st.wall_seconds,
                "workgroups": st.workgroups, "tile_bits": st.tile_bits,
                "total_passes": st.total_passes, "total_ops": st.total_ops,
                "graph_replayed": bool(st.graph_replayed)}

    def run(self, circuits, autotune: bool = False) -> List[BatchResult]:
        """circuits: dicts with n, gates, precision (C128), observables (OBS_* bits),
        draws (uniforms for shots, optional)."""
        keep = []
        items = (_BatchItem * len(circuits))()
        outs = []
        for i, c in enumerate(circuits):
            n = int(c["n"])
            g = gate_array(c.get("gates", []))
            obs = int(c.get("observables", OBS_Z))
            draws = np.ascontiguousarray(c.get("draws", np.zeros(0)), dtype=np.float64)
            norm = np.zeros(1)
            z = np.zeros(n)
            zz = np.zeros(max(n - 1, 1))
            x = np.zeros(n)
            oc = np.zeros(max(draws.size, 1), dtype=np.uint64)
            keep += [g, draws, norm, z, zz, x, oc]
            dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            it = items[i]
            it.num_qubits = n
            it.precision = int(c.get("precision", C128))
            it.gates = g.ctypes.data if len(g) else None
            it.num_gates = len(g)
            it.observables = obs
            it.num_shots = int(draws.size)
            it.draws = dp(draws) if draws.size else None
            it.norm = dp(norm)
            it.expect_z = dp(z) if obs & OBS_Z else None
            it.zz_chain = dp(zz) if obs & OBS_ZZ else None
            it.expect_x = dp(x) if obs & OBS_X else None
            it.outcomes = oc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)) if draws.size else None
            terms = c.get("pauli", [])
            pe = np.zeros(max(len(terms), 1))
            pen = np.zeros(1)
            if terms:
                parr, pstr = _pauli_terms(terms)
                keep += [parr, pstr, pe, pen]
                it.num_pauli_terms = len(terms)
                it.pauli_terms = ctypes.cast(parr, ctypes.c_void_p)
                it.pauli_expect = dp(pe)
                it.pauli_energy = dp(pen)
            outs.append((norm, z if obs & OBS_Z else None, zz[:max(n - 1, 0)] if obs & OBS_ZZ else None,
                         x if obs & OBS_X else None, oc[:draws.size] if draws.size else None,
                         pe[:len(terms)] if terms else None, pen if terms else None))
        st = _BatchStats()
        fn = lib().moonlab_mi355x_batch_autotune if autotune else lib().moonlab_mi355x_batch_run
        _check_sched(fn(self._h, items, len(circuits), ctypes.byref(st)), "batch_run")
        self.last_stats = self._stats(st)
        return [BatchResult(float(o[0][0]), o[1], o[2], o[3], o[4], o[5],
                            float(o[6][0]) if o[6] is not None else None) for o in outs]

    @staticmethod
    def _vqe_array(instances):
        """instances: dicts with n, layers, angles ([layers, n, 2] or flat),
        terms ((coefficient, string) pairs), entangler (ENT_CNOT default: the
        reference ansatz), hf (Hartree-Fock bits), constant."""
        k = len(instances)
        arr = (_Vqe * max(k, 1))()
        keep = []
        for i, v in enumerate(instances):
            n, L = int(v["n"]), int(v["layers"])
            a = np.ascontiguousarray(v["angles"], dtype=np.float64).reshape(-1)
            if a.size != 2 * n * L:
                raise ValueError("angles must hold layers x qubits x 2 values")
            terms = v.get("terms", [])
            parr, pstr = _pauli_terms(terms)
            keep += [a, parr, pstr]
            arr[i].num_qubits = n
            arr[i].num_layers = L
            arr[i].angles = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            arr[i].entangler = int(v.get("entangler", ENT_CNOT))
            arr[i].hf_reference = int(v.get("hf", 0))
            arr[i].num_terms = len(terms)
            arr[i].terms = ctypes.cast(parr, ctypes.POINTER(_PauliTerm))
            arr[i].constant = float(v.get("constant", 0.0))
        return arr, keep

    def vqe_energies(self, instances, precision: int = C128) -> np.ndarray:
        arr, keep = self._vqe_array(instances)
        e = np.zeros(max(len(instances), 1))
        st = _BatchStats()
        _check_sched(lib().moonlab_mi355x_vqe_energies(self._h, arr, len(instances), precision,
                                                       e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                       ctypes.byref(st)), "vqe_energies")
        self.last_stats = self._stats(st)
        del keep
        return e[:len(instances)]

    def vqe_gradients(self, instances, precision: int = C128):
        """(energies, [gradient of P = 2 n L values per instance])."""
        arr, keep = self._vqe_array(instances)
        sizes = [2 * int(v["n"]) * int(v["layers"]) for v in instances]
        e = np.zeros(max(len(instances), 1))
        g = np.zeros(max(sum(sizes), 1))
        st = _BatchStats()
        dp = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        _check_sched(lib().moonlab_mi355x_vqe_gradients(self._h, arr, len(instances), precision, dp(e), dp(g),
                                                        ctypes.byref(st)), "vqe_gradients")
        self.last_stats = self._stats(st)
        del keep
        offs = np.cumsum([0] + sizes)
        return e[:len(instances)], [g[offs[i]:offs[i + 1]].copy() for i in range(len(instances))]

    def vqe_qgt(self, instances):
        """[(metric, Berry curvature)] per instance, each P x P, complex128 circuits."""
        arr, keep = self._vqe_array(instances)
        sizes = [(2 * int(v["n"]) * int(v["layers"])) ** 2 for v in instances]
        q = np.zeros(max(sum(sizes), 1))
        f = np.zeros(max(sum(sizes), 1))
        st = _BatchStats()
        dp = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        _check_sched(lib().moonlab_mi355x_vqe_qgt(self._h, arr, len(instances), dp(q), dp(f), ctypes.byref(st)),
                     "vqe_qgt")
        self.last_stats = self._stats(st)
        del keep
        offs = np.cumsum([0] + sizes)
        out = []
        for i, v in enumerate(instances):
            P = 2 * int(v["n"]) * int(v["layers"])
            out.append((q[offs[i]:offs[i + 1]].reshape(P, P).copy(), f[offs[i]:offs[i + 1]].reshape(P, P).copy()))
        return out

    def vqe_natural_gradient(self, instances, gradients, regularization: float = 1e-3):
        """[(direction, status)] per instance: (g + eps I)^-1 gradient, status
        -1 (direction = gradient) where the metric is singular."""
        arr, keep = self._vqe_array(instances)
        g = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in gradients])
                                 if len(gradients) else np.zeros(1))
        d = np.zeros_like(g)
        stt = np.zeros(max(len(instances), 1), dtype=np.int32)
        st = _BatchStats()
        dp = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        _check_sched(lib().moonlab_mi355x_vqe_natural_gradient(
            self._h, arr, len(instances), dp(g), regularization, dp(d),
            stt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(st)), "vqe_natural_gradient")
        self.last_stats = self._stats(st)
        del keep
        sizes = [2 * int(v["n"]) * int(v["layers"]) for v in instances]
        offs = np.cumsum([0] + sizes)
        return [(d[offs[i]:offs[i + 1]].copy(), int(stt[i])) for i in range(len(instances))]

    @staticmethod
    def _hea_array(instances):
        """The moonlab_mi355x_hea_t array, built column-wise: one contiguous
        angle buffer and one structured array (the layout of _Hea)."""
        k = len(instances)
        flats = [np.asarray(x[2], dtype=np.float64).reshape(-1) for x in instances]
        ns = np.fromiter((x[0] for x in instances), np.int32, k)
        ls = np.fromiter((x[1] for x in instances), np.int32, k)
        sizes = 2 * ns.astype(np.int64) * ls
        got = np.fromiter((f.size for f in flats), np.int64, k)
        if np.any(got != sizes):
            raise ValueError("angles must hold layers x qubits x 2 values per instance")
        flat = np.concatenate(flats) if k else np.zeros(1)
        arr = np.zeros(max(k, 1), _HEA_DTYPE)
        arr["num_qubits"][:k] = ns
        arr["num_layers"][:k] = ls
        arr["angles"][:k] = flat.ctypes.data + 8 * (np.cumsum(sizes) - sizes)
        arr["h"][:k] = np.fromiter((x[3] for x in instances), np.float64, k)
        return arr.ctypes.data_as(ctypes.POINTER(_Hea)), (arr, flat)

    @staticmethod
    def _hea_args(instances):
        """(descriptor pointer, keep-alive, count, [(n, layers)]) for a list
        of instances or a HeaPack."""
        if isinstance(instances, HeaPack):
            return instances._ptr, instances, instances.count, instances.shapes
        arr, keep = Batch._hea_array(instances)
        return arr, keep, len(instances), [(int(n), int(L)) for (n, L, _, _) in instances]

    def hea_tfim_gradients(self, instances, precision: int = C128):
        """Energies and parameter-shift gradients ([layers, n, 2] per instance),
        every shifted circuit of every instance in one batch run.  instances:
        (n, layers, angles, h) tuples or a HeaPack."""
        arr, keep, count, shapes = self._hea_args(instances)
        e = np.zeros(count)
        sizes = [L * n * 2 for (n, L) in shapes]
        g = np.zeros(max(sum(sizes), 1))
        st = _BatchStats()
        _check_sched(lib().moonlab_mi355x_hea_tfim_gradients(
            self._h, arr, count, precision, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(st)), "hea_tfim_gradients")
        self.last_stats = self._stats(st)
        out, off = [], 0
        for (n, L), k in zip(shapes, sizes):
            out.append(g[off:off + k].reshape(L, n, 2).copy())
            off += k
        del keep
        return e, out

    def hea_tfim(self, instances, precision: int = C128) -> np.ndarray:
        """instances: (n, layers, angles[layers, n, 2], h) tuples, or a
        HeaPack of them -> energies."""
        arr, keep, count, _ = self._hea_args(instances)
        e = np.zeros(count)
        st = _BatchStats()
        _check_sched(lib().moonlab_mi355x_hea_tfim_energies(
            self._h, arr, count, precision, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            ctypes.byref(st)), "hea_tfim_energies")
        self.last_stats = self._stats(st)
        del keep
        return e


# ---------------------------------------------------------------------------
# Serving loop
# ---------------------------------------------------------------------------
@dataclass
class Response:
    status: int
    outcomes: Optional[np.ndarray]
    expect_z: Optional[np.ndarray]
    norm: float
    submit_time: float
    start_time: float
    finish_time: float
    path: int
    probabilities: Optional[np.ndarray] = None


# Status codes of the control plane's wire protocol (control_plane.h:53-62).
CONTROL_OK, CONTROL_BAD_ARG, CONTROL_OOM, CONTROL_IO_ERROR, CONTROL_PROTOCOL, CONTROL_REJECTED = \
    0, -401, -402, -403, -404, -405

_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]
_libc.free.restype = None


def handle_frame(frame: bytes, server: Optional["Server"] = None, seed: int = 0,
                 tenant: Optional[str] = None) -> Tuple[int, bytes]:
    """One control-plane request frame (CIRCUIT / SHOTS / HEALTH) through
    moonlab_mi355x_server_handle_frame (with a tenant: _ex, which sets the
    request context and fires the completion hook): returns (status,
    response frame).  Without a server only HEALTH and frames the parser
    rejects are answered."""
    out = ctypes.c_void_p()
    n = ctypes.c_size_t(0)
    h = server._h if server is not None else None
    if tenant is None:
        rc = lib().moonlab_mi355x_server_handle_frame(h, frame, len(frame), seed, ctypes.byref(out),
                                                      ctypes.byref(n))
    else:
        rc = lib().moonlab_mi355x_server_handle_frame_ex(h, frame, len(frame), seed, tenant.encode(),
                                                         ctypes.byref(out), ctypes.byref(n))
    try:
        data = ctypes.string_at(out.value, n.value) if out.value else b""
    finally:
        _libc.free(out)
    return rc, data


class Server:
    """moonlab_mi355x_server_*: continuous batching with HBM admission."""

    def __init__(self, hbm_budget_bytes: int = 0, batch_max_qubits: int = 0, max_batch: int = 0,
                 batch_engines: int = 0, large_workers: int = 0, sequential: bool = False,
                 large_path: str = "blocked") -> None:
        if large_path not in ("fused", "blocked"):
            raise ValueError("large_path is 'blocked' or 'fused'")
        cfg = _ServerConfig(hbm_budget_bytes, batch_max_qubits, max_batch, batch_engines, large_workers,
                            int(sequential), 1 if large_path == "fused" else 0)
        self._h = lib().moonlab_mi355x_server_create(ctypes.byref(cfg))
        if not self._h:
            raise EngineError("moonlab_mi355x_server_create failed (no HIP device?)")
        self._keep = {}

    def close(self) -> None:
        if getattr(self, "_h", None) and _lib is not None:
            _lib.moonlab_mi355x_server_free(self._h)
            self._h = None

    def __del__(self) -> None:
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def now(self) -> float:
        return lib().moonlab_mi355x_server_now(self._h)

    def submit(self, n: int, gates, shots: int = 1024, seed: int = 1, precision: int = C128,
               expect_z: bool = True, probabilities: bool = False) -> int:
        g = gate_array(gates)
        req = _Request(n, precision, g.ctypes.data if len(g) else None, len(g), shots, seed, int(expect_z),
                       int(probabilities))
        t = ctypes.c_uint64(0)
        _check_sched(lib().moonlab_mi355x_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)), "submit")
        return t.value

    def wait(self, ticket: int, timeout: float = -1.0) -> Response:
        r = _Response()
        rc = lib().moonlab_mi355x_server_wait(self._h, ticket, ctypes.byref(r), timeout)
        try:
            if rc != 0:
                raise SchedulerError(f"server_wait: status {rc}")
            outc = np.ctypeslib.as_array(r.outcomes, (r.num_shots,)).copy() if r.num_shots else None
            z = np.ctypeslib.as_array(r.expect_z, (r.num_qubits,)).copy() if r.expect_z else None
            p = np.ctypeslib.as_array(r.probabilities, (1 << r.num_qubits,)).copy() if r.probabilities else None
            return Response(r.status, outc, z, r.norm, r.submit_time, r.start_time, r.finish_time, r.path, p)
        finally:
            lib().moonlab_mi355x_response_free(ctypes.byref(r))

    def stats(self) -> dict:
        st = _ServerStats()
        _check_sched(lib().moonlab_mi355x_server_stats(self._h, ctypes.byref(st)), "server_stats")
        return {f: getattr(st, f) for f, _ in _ServerStats._fields_}
