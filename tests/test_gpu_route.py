"""The reference's own gate and measurement calls on the MI355X through the
GPU seam (libmoonlab_mi355x_route.so linked into the reference build in place
of its CPU stub, oracle/Makefile ref_routed): quantum_state_create_gpu, the
reference's gate_* on the GPU-backed state (routed through qsim_gpu_route_*),
its measurement_* reads (which sync the host mirror), against the oracle.

Tolerance: amplitudes and <Z_q> within 1e-12 (FP64 on both sides)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, random_gate_list
from oracle.oracle import gates_array

pytestmark = pytest.mark.gpu

ROUTED = os.path.join(ROOT, "oracle", "_ref", "libmoonlab_ref_routed.so")


def _lib():
    if not os.path.exists(ROUTED):
        pytest.skip("oracle/_ref/libmoonlab_ref_routed.so not built (make -C oracle ref_routed)")
    L = ctypes.CDLL(ROUTED)
    vp = ctypes.c_void_p
    L.quantum_state_create_gpu.argtypes = [ctypes.c_size_t, ctypes.POINTER(vp)]
    L.quantum_state_create_gpu.restype = ctypes.c_int
    L.quantum_state_sync_to_host.argtypes = [vp]
    L.quantum_state_destroy.argtypes = [vp]
    L.ref_apply_gates.argtypes = [vp, vp, ctypes.c_int]
    L.ref_state_amplitudes.argtypes = [vp]
    L.ref_state_amplitudes.restype = ctypes.POINTER(ctypes.c_double)
    L.measurement_expectation_z.argtypes = [vp, ctypes.c_int]
    L.measurement_expectation_z.restype = ctypes.c_double
    return L


def test_reference_gates_run_on_the_gpu_through_the_seam(eng, oracle):
    L = _lib()
    rng = np.random.default_rng(107)
    for n in (3, 9, 14):
        # the reference's gate table (codes 0..24 of ref_extra.c), CY on the job path
        g = gates_array(random_gate_list(n, 60, rng))
        p = ctypes.c_void_p()
        assert L.quantum_state_create_gpu(n, ctypes.byref(p)) == 0 and p.value
        try:
            assert L.ref_apply_gates(p, g.ctypes.data, len(g)) == 0
            want = oracle.run(n, g)
            z = np.array([L.measurement_expectation_z(p, q) for q in range(n)])   # syncs the mirror
            assert np.max(np.abs(z - oracle.expect_z_all(n, want))) < 1e-12
            assert L.quantum_state_sync_to_host(p) == 0
            amps = np.ctypeslib.as_array(L.ref_state_amplitudes(p), shape=(2 << n,)).view(np.complex128)
            assert np.max(np.abs(amps - want)) < 1e-12, n
        finally:
            L.quantum_state_destroy(p)
