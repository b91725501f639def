"""moonlab_hip_expect_z_all_noalloc: the norm and <Z_q> sums of
moonlab_hip_expect_z_all from a fixed device array instead of an allocated
workspace, called through the C ABI on raw device addresses.  Its output must
equal the allocating entry point's bit for bit (same kernels, grid and
partial layout), on consecutive calls whose second input is smaller (the
array keeps the previous call's partials), in both precisions and on both
reduction shapes (n < 12: one block; n >= 12: tiles)."""
import numpy as np
import pytest

from moonlab_amd import workloads

pytestmark = pytest.mark.gpu

HIP_ERR_BAD_ARG = -6      # MOONLAB_HIP_ERR_BAD_ARG (include/moonlab_hip_statevec.h)


def _noalloc(eng, st, out):
    """Run the no-alloc entry point on st's amplitudes and stream, writing
    into out's device memory; returns the n + 1 doubles."""
    L = eng.lib()
    rc = L.moonlab_hip_expect_z_all_noalloc(L.moonlab_hip_state_device_ptr(st.handle), st.n, st.precision,
                                            L.moonlab_hip_state_stream(st.handle),
                                            L.moonlab_hip_state_device_ptr(out.handle))
    assert rc == 0, eng.last_error()
    st.synchronize()
    return out.read(0, (st.n + 2) // 2).view(np.float64)[:st.n + 1].copy()


def _state(eng, n, prec, seed):
    st = eng.StateVector(n, prec)
    st.apply(workloads.random_layer_circuit(n, 3, seed, h_layer=True))
    return st


def test_noalloc_matches_allocating_entry_point(eng):
    with eng.StateVector(6, eng.C128) as out:          # 64 complex128 = 128 doubles of output room
        for n, prec, seed in ((22, eng.C128, 1), (14, eng.C64, 2), (9, eng.C128, 3), (13, eng.C128, 4),
                              (5, eng.C64, 5)):
            with _state(eng, n, prec, seed) as st:
                got = _noalloc(eng, st, out)
                z, nrm = st.expect_z_all()
                assert got[0] == nrm, (n, prec)
                assert got[1:].tobytes() == np.asarray(z, dtype=np.float64).tobytes(), (n, prec)
                again = _noalloc(eng, st, out)
                assert again.tobytes() == got.tobytes()


def test_noalloc_rejects_sizes_beyond_its_array(eng):
    L = eng.lib()
    with eng.StateVector(4, eng.C128) as st:
        p = L.moonlab_hip_state_device_ptr(st.handle)
        assert L.moonlab_hip_expect_z_all_noalloc(p, 35, eng.C128, None, p) == HIP_ERR_BAD_ARG
        assert L.moonlab_hip_expect_z_all_noalloc(p, 36, eng.C64, None, p) == HIP_ERR_BAD_ARG
        assert "out of range" in eng.last_error()
        assert L.moonlab_hip_expect_z_all_noalloc(p, 0, eng.C128, None, p) == HIP_ERR_BAD_ARG
        assert L.moonlab_hip_expect_z_all_noalloc(None, 4, eng.C128, None, p) == HIP_ERR_BAD_ARG
        assert L.moonlab_hip_expect_z_all_noalloc(p, 4, 7, None, p) == HIP_ERR_BAD_ARG
