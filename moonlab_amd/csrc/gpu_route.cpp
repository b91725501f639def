// gpu_route.cpp -- the reference's GPU seam on the MI355X engine
// (libmoonlab_mi355x_route.so; declarations and contract in
// include/moonlab_gpu_route.h).
//
// The reference's CUDA build satisfies this seam in
// src/quantum/state_gpu.cu:31-198 and src/backends/cuda_statevec.cu (managed
// memory, one kernel per routed gate); its CPU build in
// src/quantum/gpu_route_cpu_unavailable.c (every entry -1).  Here each entry
// forwards to the engine's moonlab_hip_* layer: plain device memory, one
// stream per state, the engine's gate kernels.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/moonlab_gpu_route.h"
#include "../../include/moonlab_hip_statevec.h"

namespace {

moonlab_hip_state_t *dev(const moonlab_route_state_t *s)
{
    return s ? (moonlab_hip_state_t *)s->gpu_state : nullptr;
}

int ok(moonlab_hip_status_t st) { return st == MOONLAB_HIP_OK ? 0 : -1; }

}  // namespace

extern "C" {

int quantum_state_create_gpu(size_t num_qubits, moonlab_route_state_t **out_state)
{
    if (!out_state) return MOONLAB_ROUTE_QS_INVALID_PARAM;
    *out_state = nullptr;
    if (num_qubits < 1 || num_qubits > MOONLAB_HIP_MAX_QUBITS_C128) return MOONLAB_ROUTE_QS_INVALID_PARAM;
    if (moonlab_hip_device_count() <= 0) return MOONLAB_ROUTE_QS_NOT_SUPPORTED;
    moonlab_route_state_t *s = (moonlab_route_state_t *)calloc(1, sizeof(*s));
    if (!s) return MOONLAB_ROUTE_QS_OUT_OF_MEMORY;
    s->num_qubits = num_qubits;
    s->state_dim = (size_t)1 << num_qubits;
    // the host mirror the state layer reads after a sync, freed by the
    // reference's quantum_state_free (owns_memory)
    s->amplitudes = (double *)calloc(s->state_dim, 2 * sizeof(double));
    if (!s->amplitudes) {
        free(s);
        return MOONLAB_ROUTE_QS_OUT_OF_MEMORY;
    }
    s->amplitudes[0] = 1.0;
    s->owns_memory = 1;
    moonlab_hip_state_t *h = nullptr;
    const moonlab_hip_status_t st = moonlab_hip_state_create((uint32_t)num_qubits, &h);
    if (st != MOONLAB_HIP_OK) {
        free(s->amplitudes);
        free(s);
        return st == MOONLAB_HIP_ERR_OUT_OF_MEMORY ? MOONLAB_ROUTE_QS_OUT_OF_MEMORY : MOONLAB_ROUTE_QS_NOT_SUPPORTED;
    }
    s->gpu_state = h;
    s->gpu_backend = MOONLAB_ROUTE_GPU_BACKEND_ROCM;
    *out_state = s;
    return MOONLAB_ROUTE_QS_SUCCESS;
}

int quantum_state_sync_to_host(moonlab_route_state_t *state)
{
    if (!state) return MOONLAB_ROUTE_QS_INVALID_PARAM;
    if (!state->gpu_state) return MOONLAB_ROUTE_QS_SUCCESS;
    return moonlab_hip_state_copy_to_host(dev(state), state->amplitudes) == MOONLAB_HIP_OK
               ? MOONLAB_ROUTE_QS_SUCCESS
               : MOONLAB_ROUTE_QS_DRIVER;
}

int quantum_state_sync_from_host(moonlab_route_state_t *state)
{
    if (!state) return MOONLAB_ROUTE_QS_INVALID_PARAM;
    if (!state->gpu_state) return MOONLAB_ROUTE_QS_SUCCESS;
    return moonlab_hip_state_copy_from_host(dev(state), state->amplitudes) == MOONLAB_HIP_OK
               ? MOONLAB_ROUTE_QS_SUCCESS
               : MOONLAB_ROUTE_QS_DRIVER;
}

int qsim_gpu_route_hadamard(moonlab_route_state_t *state, int qubit)
{
    if (!dev(state) || qubit < 0) return -1;
    return ok(moonlab_hip_apply_h(dev(state), (uint32_t)qubit));
}

int qsim_gpu_route_pauli_x(moonlab_route_state_t *state, int qubit)
{
    if (!dev(state) || qubit < 0) return -1;
    return ok(moonlab_hip_apply_x(dev(state), (uint32_t)qubit));
}

int qsim_gpu_route_cnot(moonlab_route_state_t *state, int control, int target)
{
    if (!dev(state) || control < 0 || target < 0) return -1;
    return ok(moonlab_hip_apply_cnot(dev(state), (uint32_t)control, (uint32_t)target));
}

int qsim_gpu_route_apply_1q_matrix(moonlab_route_state_t *state, int qubit, const double m[8])
{
    if (!dev(state) || !m || qubit < 0) return -1;
    return ok(moonlab_hip_apply_1q(dev(state), (uint32_t)qubit, m));
}

int qsim_gpu_route_apply_2q_matrix(moonlab_route_state_t *state, int q0, int q1, const double m[32])
{
    if (!dev(state) || !m || q0 < 0 || q1 < 0) return -1;
    return ok(moonlab_hip_apply_2q(dev(state), (uint32_t)q0, (uint32_t)q1, m));
}

int qsim_gpu_route_mcx(moonlab_route_state_t *state, uint64_t control_mask, int target)
{
    if (!dev(state) || target < 0) return -1;
    return ok(moonlab_hip_apply_mcx(dev(state), control_mask, (uint32_t)target));
}

int qsim_gpu_route_mcz(moonlab_route_state_t *state, uint64_t all_mask)
{
    if (!dev(state)) return -1;
    return ok(moonlab_hip_apply_mcz(dev(state), all_mask));
}

int qsim_gpu_route_fredkin(moonlab_route_state_t *state, int control, int t1, int t2)
{
    if (!dev(state) || control < 0 || t1 < 0 || t2 < 0) return -1;
    return ok(moonlab_hip_apply_fredkin(dev(state), (uint32_t)control, (uint32_t)t1, (uint32_t)t2));
}

void moonlab_cuda_state_free(void *gpu_state) { moonlab_hip_state_free((moonlab_hip_state_t *)gpu_state); }

int moonlab_cuda_runtime_probe_discrete(void) { return moonlab_hip_device_count() > 0 ? 1 : 0; }

int moonlab_cuda_state_create(uint32_t n_qubits, void **out_state)
{
    if (!out_state) return -1;
    moonlab_hip_state_t *h = nullptr;
    const int rc = ok(moonlab_hip_state_create(n_qubits, &h));
    *out_state = rc == 0 ? h : nullptr;
    return rc;
}

int moonlab_cuda_apply_1q(void *state, uint32_t target, const double m[8])
{
    if (!state || !m) return -1;
    return ok(moonlab_hip_apply_1q((moonlab_hip_state_t *)state, target, m));
}

int moonlab_cuda_apply_cnot(void *state, uint32_t control, uint32_t target)
{
    if (!state) return -1;
    return ok(moonlab_hip_apply_cnot((moonlab_hip_state_t *)state, control, target));
}

int moonlab_cuda_state_copy_to_host(const void *state, double *out)
{
    if (!state || !out) return -1;
    return ok(moonlab_hip_state_copy_to_host((const moonlab_hip_state_t *)state, out));
}

int moonlab_cuda_state_copy_from_host(void *state, const double *in)
{
    if (!state || !in) return -1;
    return ok(moonlab_hip_state_copy_from_host((moonlab_hip_state_t *)state, in));
}

int moonlab_cuda_select_device_for_rank(int local_rank, int *device_id, int *device_count)
{
    const int count = moonlab_hip_device_count();
    if (device_count) *device_count = count;
    if (count <= 0 || local_rank < 0) {
        if (device_id) *device_id = -1;
        return -1;
    }
    const int d = local_rank % count;
    if (device_id) *device_id = d;
    return hipSetDevice(d) == hipSuccess ? 0 : -1;
}

}  // extern "C"
