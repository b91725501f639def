/*
 * moonlab_scheduler.h -- job API + backend registry C ABI (drop-in boundary).
 *
 * Same names, types, status codes and semantics as the reference's
 * src/distributed/scheduler.h (lines cited per entry).  A job whose backend
 * name resolves to an MI355X engine entry ("mi355x", and the default
 * "simulator" in this library) runs every gate as a HIP kernel on the GPU.
 */
#ifndef MOONLAB_MI355X_SCHEDULER_H
#define MOONLAB_MI355X_SCHEDULER_H

#include <stddef.h>
#include <stdint.h>

#include "moonlab_qgtl_backend.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (scheduler.h:36-42, 187). */
#define MOONLAB_SCHED_OK                ( 0)
#define MOONLAB_SCHED_BAD_ARG           (-501)
#define MOONLAB_SCHED_OOM               (-502)
#define MOONLAB_SCHED_INTERNAL          (-503)
#define MOONLAB_SCHED_BUFFER_TOO_SMALL  (-504)
#define MOONLAB_SCHED_NOT_BUILT         (-505)
#define MOONLAB_SCHED_BACKEND_NOT_FOUND (-506)
#define MOONLAB_SCHED_BACKEND_BUSY      (-507)

typedef struct moonlab_job moonlab_job_t;

/* Results record (scheduler.h:51-57). */
typedef struct {
    int       num_qubits;
    int       total_shots;
    uint64_t *outcomes;
    int       num_workers_used;
    double   *worker_seconds;
} moonlab_job_results_t;

/* Replaces moonlab_job_create (scheduler.h:66-67): 1..32 qubits else NULL. */
MOONLAB_API moonlab_job_t *moonlab_job_create(int num_qubits);
/* Replaces moonlab_job_free (scheduler.h:73). */
MOONLAB_API void moonlab_job_free(moonlab_job_t *job);
/* Replaces moonlab_job_add_gate (scheduler.h:78-82); like the reference it
 * checks only job != NULL and params for RX/RY/RZ (scheduler.c:99-121). */
MOONLAB_API int moonlab_job_add_gate(moonlab_job_t *job, moonlab_qgtl_gate_t type,
                                     int target, int control, const double *params);
/* Replace the setters (scheduler.h:87-103). */
MOONLAB_API int moonlab_job_set_num_shots(moonlab_job_t *job, int num_shots);
MOONLAB_API int moonlab_job_set_num_workers(moonlab_job_t *job, int num_workers);
MOONLAB_API int moonlab_job_set_rng_seed(moonlab_job_t *job, uint64_t seed);
/* Replace the getters (scheduler.h:112-137). */
MOONLAB_API int moonlab_job_num_qubits(const moonlab_job_t *job);
MOONLAB_API int moonlab_job_num_gates(const moonlab_job_t *job);
MOONLAB_API int moonlab_job_num_shots(const moonlab_job_t *job);
MOONLAB_API int moonlab_job_num_workers(const moonlab_job_t *job);
/* Replaces moonlab_scheduler_run (scheduler.h:156-158). */
MOONLAB_API int moonlab_scheduler_run(moonlab_job_t *job, moonlab_job_results_t *out);
/* Replaces moonlab_job_results_free (scheduler.h:164-165). */
MOONLAB_API void moonlab_job_results_free(moonlab_job_results_t *r);
/* Replaces moonlab_scheduler_run_mpi (scheduler.h:208-211): one GPU per
 * process here and no MPI transport, so it returns MOONLAB_SCHED_NOT_BUILT
 * exactly as the reference does in a non-MPI build. */
MOONLAB_API int moonlab_scheduler_run_mpi(moonlab_job_t *job, moonlab_job_results_t *out,
                                          void *ctx_opaque);

/* Backend plug-in surface (scheduler.h:235-252). */
typedef int (*moonlab_backend_execute_fn)(const moonlab_job_t *job,
                                          moonlab_job_results_t *out, void *ctx);
typedef struct {
    const char                *name;
    moonlab_backend_execute_fn execute;
    void                      *ctx;
    const char                *description;
} moonlab_backend_t;

/* Replace the registry entry points (scheduler.h:262-314). */
MOONLAB_API int moonlab_register_backend(const moonlab_backend_t *backend);
MOONLAB_API int moonlab_unregister_backend(const char *name);
MOONLAB_API const moonlab_backend_t *moonlab_find_backend(const char *name);
MOONLAB_API int moonlab_num_backends(void);
MOONLAB_API int moonlab_list_backends(const char **out_names, int max);
MOONLAB_API int moonlab_job_set_backend(moonlab_job_t *job, const char *backend_name);
MOONLAB_API const char *moonlab_job_backend(const moonlab_job_t *job);

/* Completion hook and request context (scheduler.h:339-463). */
typedef void (*moonlab_completion_hook_fn)(const moonlab_job_t *job,
                                           const moonlab_job_results_t *out,
                                           const char *backend_name, void *ctx);
MOONLAB_API int moonlab_scheduler_set_completion_hook(moonlab_completion_hook_fn hook, void *ctx);
/* Replaces moonlab_job_to_json (scheduler.h:379-381); byte-identical output. */
MOONLAB_API int moonlab_job_to_json(const moonlab_job_t *job, char *buf, size_t bufsize);
MOONLAB_API void moonlab_scheduler_set_request_context(const char *tenant_id,
                                                       const char *request_id);
MOONLAB_API const char *moonlab_scheduler_current_tenant_id(void);
MOONLAB_API const char *moonlab_scheduler_current_request_id(void);
MOONLAB_API void moonlab_scheduler_fire_completion_hook(const moonlab_job_t *job,
                                                        const moonlab_job_results_t *results,
                                                        const char *backend_name);

#ifdef __cplusplus
}
#endif
#endif
