/*
 * moonlab_mi355x.h -- MI355X engine entry points beside the job API.
 *
 * The reference's job results carry only shot outcomes
 * (src/distributed/scheduler.h:51-57).  These additions hand back what the
 * north-star workloads also need -- amplitudes, probabilities, <Z_q> for
 * every qubit, a complex64 option and gate fusion -- without changing any
 * reference signature.  The execute_ex shape follows moonlab_qgtl_execute's
 * options/results pair (src/applications/moonlab_qgtl_backend.h:90-109).
 */
#ifndef MOONLAB_MI355X_H
#define MOONLAB_MI355X_H

#include <stddef.h>
#include <stdint.h>

#include "moonlab_hip_statevec.h"
#include "moonlab_scheduler.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MOONLAB_MI355X_BACKEND_NAME "mi355x"

/* Register the MI355X engine under the name "mi355x" through
 * moonlab_register_backend (scheduler.h:262), so that
 * moonlab_job_set_backend(job, "mi355x") routes a job to the GPU.
 * Idempotent (same-name registration replaces, scheduler.c:290-297). */
MOONLAB_API int moonlab_mi355x_register_backend(void);

/* The backend's execute callback (moonlab_backend_execute_fn,
 * scheduler.h:235-238).  Reads the job only through the public job API
 * (moonlab_job_num_* and moonlab_job_to_json), so it also runs under the
 * reference's own scheduler.c when registered there. */
MOONLAB_API int moonlab_mi355x_backend_execute(const moonlab_job_t *job,
                                               moonlab_job_results_t *out, void *ctx);

typedef struct {
    int precision;             /* MOONLAB_HIP_C128 (default) or MOONLAB_HIP_C64 */
    int return_amplitudes;     /* 1: fill `amplitudes` (2 * 2^n doubles, interleaved) */
    int return_probabilities;  /* 1: fill `probabilities` (2^n doubles) */
    int return_expect_z;       /* 1: fill `expect_z` (n doubles) and `norm` */
    int fusion_max_qubits;     /* 0: one kernel per gate; 1..5: fuse into k-qubit blocks */
    int skip_shots;            /* 1: ignore the job's shot count (state outputs only) */
} moonlab_mi355x_exec_options_t;

typedef struct {
    int       num_qubits;
    int       precision;
    int       total_shots;
    int       num_workers_used;
    uint64_t *outcomes;        /* total_shots, same streams as moonlab_scheduler_run */
    double   *amplitudes;      /* or NULL */
    double   *probabilities;   /* or NULL */
    double   *expect_z;        /* or NULL */
    double    norm;            /* sum |a|^2 when expect_z was requested */
    double    device_seconds;  /* gate + reduction + sampling time on the GPU stream */
    double    wall_seconds;    /* whole call, host view */
} moonlab_mi355x_results_t;

/* Run a job on the GPU with extra outputs.  Returns MOONLAB_SCHED_* codes. */
MOONLAB_API int moonlab_mi355x_execute_ex(const moonlab_job_t *job,
                                          const moonlab_mi355x_exec_options_t *opts,
                                          moonlab_mi355x_results_t *out);
MOONLAB_API void moonlab_mi355x_results_free(moonlab_mi355x_results_t *r);

/* Device states released by finished jobs are kept for reuse, keyed by
 * (qubits, precision), up to this many bytes (default 64 GiB).  0 frees the
 * cache. */
MOONLAB_API int moonlab_mi355x_set_state_cache_bytes(uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
