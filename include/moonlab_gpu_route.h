/*
 * moonlab_gpu_route.h -- the reference's GPU seam, implemented on the MI355X
 * engine (libmoonlab_mi355x_route.so).
 *
 * The reference's gate primitives (src/quantum/gates.c:43-79) look up
 * qsim_gpu_route_* and run a gate on the device when a quantum_state_t
 * carries GPU backing (gpu_state != NULL); its state layer
 * (src/quantum/state.c:20-31, 83-110, 640-672) calls
 * quantum_state_sync_to_host / _from_host around every host read and
 * moonlab_cuda_state_free on destruction.  A CUDA build defines these in
 * src/quantum/state_gpu.cu:31-198 and src/backends/cuda_statevec.cu; a
 * CPU build links src/quantum/gpu_route_cpu_unavailable.c, whose every
 * entry returns -1.  This library defines the same symbols with the same
 * meaning on gfx950: link it into a reference build in place of
 * gpu_route_cpu_unavailable.c, and quantum_state_create_gpu returns a state
 * whose gates, measurements and reads run on the MI355X through the
 * reference's own gate_* / measurement_* calls.
 *
 * gpu_state holds a moonlab_hip_state_t * (complex128) and gpu_backend is
 * 2 (the value state.h:126-128 reserves for a ROCm backend).  The host
 * `amplitudes` mirror is stale until quantum_state_sync_to_host().
 */
#ifndef MOONLAB_GPU_ROUTE_H
#define MOONLAB_GPU_ROUTE_H

#include <stddef.h>
#include <stdint.h>

#ifndef MOONLAB_API
#define MOONLAB_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's quantum_state_t (src/quantum/state.h:102-131): the seam
 * reads and writes it, so its field order and types are the contract. */
typedef struct {
    size_t    num_qubits;
    size_t    state_dim;
    double   *amplitudes;            /* state_dim interleaved (re, im) pairs */
    uint64_t *measurement_outcomes;
    size_t    num_measurements;
    size_t    max_measurements;
    int       owns_memory;
    void     *gpu_state;
    int       gpu_backend;
} moonlab_route_state_t;

/* qs_error_t values the seam returns (src/quantum/state.h:78-88). */
#define MOONLAB_ROUTE_QS_SUCCESS          0
#define MOONLAB_ROUTE_QS_OUT_OF_MEMORY   (-4)
#define MOONLAB_ROUTE_QS_INVALID_PARAM   (-6)
#define MOONLAB_ROUTE_QS_NOT_SUPPORTED   (-7)
#define MOONLAB_ROUTE_QS_DRIVER          (-8)
#define MOONLAB_ROUTE_GPU_BACKEND_ROCM    2

/* Replaces quantum_state_create_gpu (state_gpu.cu:31-72): |0..0> on the
 * device, 1..34 qubits, with a host mirror of the same size.
 * NOT_SUPPORTED with *out = NULL when no GPU is visible. */
MOONLAB_API int quantum_state_create_gpu(size_t num_qubits, moonlab_route_state_t **out_state);
/* Replace quantum_state_sync_to_host / _from_host (state_gpu.cu:74-98):
 * success with nothing to do on a state without GPU backing,
 * INVALID_PARAM on NULL. */
MOONLAB_API int quantum_state_sync_to_host(moonlab_route_state_t *state);
MOONLAB_API int quantum_state_sync_from_host(moonlab_route_state_t *state);

/* The gate routes gates.c calls (state_gpu.cu:108-198): 0 when the gate ran
 * on the device, -1 without GPU backing or on a device error.  Matrices are
 * row-major (re, im) pairs; the 2-qubit basis is |q1 q0>. */
MOONLAB_API int qsim_gpu_route_hadamard(moonlab_route_state_t *state, int qubit);
MOONLAB_API int qsim_gpu_route_pauli_x(moonlab_route_state_t *state, int qubit);
MOONLAB_API int qsim_gpu_route_cnot(moonlab_route_state_t *state, int control, int target);
MOONLAB_API int qsim_gpu_route_apply_1q_matrix(moonlab_route_state_t *state, int qubit, const double m[8]);
MOONLAB_API int qsim_gpu_route_apply_2q_matrix(moonlab_route_state_t *state, int q0, int q1, const double m[32]);
MOONLAB_API int qsim_gpu_route_mcx(moonlab_route_state_t *state, uint64_t control_mask, int target);
MOONLAB_API int qsim_gpu_route_mcz(moonlab_route_state_t *state, uint64_t all_mask);
MOONLAB_API int qsim_gpu_route_fredkin(moonlab_route_state_t *state, int control, int t1, int t2);

/* The moonlab_cuda_* entries the reference's state layer and
 * gpu_route_cpu_unavailable.c:24-75 name, forwarding to moonlab_hip_*:
 * state handles are moonlab_hip_state_t *; return 0 or -1. */
MOONLAB_API void moonlab_cuda_state_free(void *gpu_state);
MOONLAB_API int  moonlab_cuda_runtime_probe_discrete(void);   /* 1: a discrete GPU is visible */
MOONLAB_API int  moonlab_cuda_state_create(uint32_t n_qubits, void **out_state);
MOONLAB_API int  moonlab_cuda_apply_1q(void *state, uint32_t target, const double m[8]);
MOONLAB_API int  moonlab_cuda_apply_cnot(void *state, uint32_t control, uint32_t target);
MOONLAB_API int  moonlab_cuda_state_copy_to_host(const void *state, double *out);
MOONLAB_API int  moonlab_cuda_state_copy_from_host(void *state, const double *in);
/* One process per GPU: device local_rank mod the device count. */
MOONLAB_API int  moonlab_cuda_select_device_for_rank(int local_rank, int *device_id, int *device_count);

#ifdef __cplusplus
}
#endif
#endif
