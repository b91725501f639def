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
    int blocked;               /* 1: cache-blocked passes (moonlab_hip_apply_gates_blocked) */
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

/* ---- Batched small circuits: one persistent launch -------------------------
 *
 * Thousands of independent circuits of uneven size run in ONE kernel launch:
 * each GPU workgroup takes the next circuit from a device work queue and
 * simulates it to completion (states up to 12 qubits entirely in LDS; larger
 * ones cache-blocked through LDS from a per-workgroup HBM slot), then writes
 * the requested observables and shots.  Results for a circuit do not depend
 * on what else is in the batch (fixed per-circuit reduction order).
 * The workload shape is the reference's VQE energy loop
 * (src/algorithms/vqe.c:1287-1350) over many parameter sets at once. */

typedef struct moonlab_mi355x_batch moonlab_mi355x_batch_t;

/* One Hamiltonian term, the reference's pauli_term_t (vqe.h:101-105):
 * pauli_string[q] in "IXYZ" is the operator on qubit q, exactly num_qubits
 * characters; Y = i X Z. */
typedef struct {
    double      coefficient;
    const char *pauli_string;
} moonlab_mi355x_pauli_term_t;

#define MOONLAB_MI355X_OBS_Z  1   /* <Z_q> for every qubit (and the norm) */
#define MOONLAB_MI355X_OBS_ZZ 2   /* <Z_q Z_q+1>, q = 0..n-2 */
#define MOONLAB_MI355X_OBS_X  4   /* <X_q> for every qubit */
#define MOONLAB_MI355X_BATCH_MAX_QUBITS 24

typedef struct {
    int                       num_qubits;   /* 1..24 */
    int                       precision;    /* MOONLAB_HIP_C128 or MOONLAB_HIP_C64 */
    const moonlab_hip_gate_t *gates;        /* codes of moonlab_hip_statevec.h */
    int                       num_gates;
    int                       observables;  /* MOONLAB_MI355X_OBS_* bits */
    int                       num_shots;    /* 0 = none */
    const double             *draws;        /* num_shots uniform draws in [0, 1) */
    /* outputs, caller-owned; NULL when not requested */
    double                   *norm;         /* 1 value */
    double                   *expect_z;     /* n values */
    double                   *zz_chain;     /* n - 1 values */
    double                   *expect_x;     /* n values */
    uint64_t                 *outcomes;     /* num_shots values, the collapse rule of
                                               measurement_all_qubits (measurement.c:263-307) */
    /* Pauli-sum observables (vqe_compute_energy's terms, vqe.c:1287-1350):
     * <P_k> for each string, measured exactly on the final state */
    int                                num_pauli_terms;
    const moonlab_mi355x_pauli_term_t *pauli_terms;
    double                            *pauli_expect;   /* num_pauli_terms values, or NULL */
    double                            *pauli_energy;   /* sum_k c_k <P_k> in term order, or NULL */
} moonlab_mi355x_batch_item_t;

typedef struct {
    double device_seconds;   /* the persistent launch, HIP events on the batch stream */
    double wall_seconds;     /* the whole call: planning, upload, launch, download */
    int    workgroups;
    int    tile_bits;
    int    total_passes;     /* HBM sweeps planned for the circuits above tile size */
    int    total_ops;
    int    graph_replayed;   /* 1: the run replayed the HIP graph captured for this batch shape */
    int    reserved;
} moonlab_mi355x_batch_stats_t;

MOONLAB_API moonlab_mi355x_batch_t *moonlab_mi355x_batch_create(void);
MOONLAB_API void moonlab_mi355x_batch_free(moonlab_mi355x_batch_t *b);
/* Workgroups per CU (1 or 2) and tile bits (8..12); 0 keeps a value. */
MOONLAB_API int moonlab_mi355x_batch_set_config(moonlab_mi355x_batch_t *b, int workgroups_per_cu,
                                                int tile_bits);
/* Time every configuration on these items and keep the fastest; the
 * items' outputs are written by the last timed run. */
MOONLAB_API int moonlab_mi355x_batch_autotune(moonlab_mi355x_batch_t *b,
                                              moonlab_mi355x_batch_item_t *items, int count,
                                              moonlab_mi355x_batch_stats_t *stats);
MOONLAB_API int moonlab_mi355x_batch_run(moonlab_mi355x_batch_t *b,
                                         moonlab_mi355x_batch_item_t *items, int count,
                                         moonlab_mi355x_batch_stats_t *stats);
/* HIP graphs (default on): a run's whole device step -- program upload,
 * queue reset, the persistent launch, results download -- is captured when a
 * batch shape (circuit sizes, passes, op and shot counts, launch
 * configuration) repeats on the next run, and replayed with one
 * hipGraphLaunch while it keeps repeating, as in a VQE loop that re-evaluates
 * the same ansatz with new angles; a shape seen once runs without capture.
 * device_seconds covers upload + launch + download. */
MOONLAB_API int moonlab_mi355x_batch_set_graphs(moonlab_mi355x_batch_t *b, int enable);

/* Hardware-efficient ansatz + open-chain transverse-field Ising energy, the
 * batched form of vqe_compute_energy: per layer RY, RZ on every qubit then
 * CZ(q, q+1) for q = 0..n-2, from |0..0>; E = -sum <Z_q Z_q+1> - h sum <X_q>. */
typedef struct {
    int           num_qubits;
    int           num_layers;
    const double *angles;     /* layers * n * 2: [layer][qubit][ry, rz] */
    double        h;
} moonlab_mi355x_hea_t;

MOONLAB_API int moonlab_mi355x_hea_tfim_energies(moonlab_mi355x_batch_t *b,
                                                 const moonlab_mi355x_hea_t *instances, int count,
                                                 int precision, double *energies,
                                                 moonlab_mi355x_batch_stats_t *stats);

/* Energies and parameter-shift gradients of the same ansatz for many
 * instances: each instance adds its base point and the +-pi/2 shift of each
 * of its 2 n L angles to ONE batch run, and dE/dtheta_k =
 * [E(theta + pi/2 e_k) - E(theta - pi/2 e_k)] / 2, exact for RY / RZ --
 * the rule of vqe_compute_gradient's fallback (vqe.c:2094-2130); its adjoint
 * fast path computes the same exact gradient.  gradients holds sum_i 2 n_i L_i
 * values, instance after instance, in the angles' [layer][qubit][ry, rz] order. */
MOONLAB_API int moonlab_mi355x_hea_tfim_gradients(moonlab_mi355x_batch_t *b,
                                                  const moonlab_mi355x_hea_t *instances, int count,
                                                  int precision, double *energies, double *gradients,
                                                  moonlab_mi355x_batch_stats_t *stats);

/* ---- Variational workloads over Pauli-sum Hamiltonians ---------------------
 *
 * The reference's VQE solver (src/algorithms/vqe.h:545-823) evaluates one
 * parameter set at a time on the CPU.  These calls take many instances and
 * run every circuit they need -- base points, +-pi/2 shifts, the replays with
 * a parameter's generator inserted -- in ONE persistent batch launch.
 *
 * An instance is the reference's hardware-efficient ansatz
 * (vqe.c:764-781): X on the Hartree-Fock bits (vqe.c:1302-1305), then per
 * layer RY(theta[p]), RZ(theta[p+1]) on every qubit and an entangling ladder
 * over (q, q+1) -- CNOT with control q (MOONLAB_MI355X_ENT_CNOT, the
 * reference's) or CZ (MOONLAB_MI355X_ENT_CZ); angles are [layer][qubit][ry, rz],
 * P = 2 n L of them.  E = sum_k c_k <P_k> + constant (vqe_compute_energy,
 * vqe.c:1287-1350; the constant is nuclear_repulsion). */
#define MOONLAB_MI355X_ENT_CZ   0
#define MOONLAB_MI355X_ENT_CNOT 1

typedef struct {
    int                                num_qubits;     /* 1..24 */
    int                                num_layers;
    const double                      *angles;         /* 2 n L values */
    int                                entangler;      /* MOONLAB_MI355X_ENT_* */
    uint64_t                           hf_reference;   /* bit q = 1: X on qubit q first */
    int                                num_terms;
    const moonlab_mi355x_pauli_term_t *terms;          /* strings of num_qubits characters */
    double                             constant;
} moonlab_mi355x_vqe_t;

/* Energies, one per instance. */
MOONLAB_API int moonlab_mi355x_vqe_energies(moonlab_mi355x_batch_t *b, const moonlab_mi355x_vqe_t *instances,
                                            int count, int precision, double *energies,
                                            moonlab_mi355x_batch_stats_t *stats);
/* Energies and parameter-shift gradients (vqe_compute_gradient, vqe.c:2094-2130):
 * gradients holds sum_i P_i values, instance after instance. */
MOONLAB_API int moonlab_mi355x_vqe_gradients(moonlab_mi355x_batch_t *b, const moonlab_mi355x_vqe_t *instances,
                                             int count, int precision, double *energies, double *gradients,
                                             moonlab_mi355x_batch_stats_t *stats);
/* Quantum geometric tensor Q_ij = <d_i psi|d_j psi> - <d_i psi|psi><psi|d_j psi>
 * with exact derivatives (the generator inserted after its rotation,
 * vqe_qng.c:66-170), complex128: qgt = Re Q (vqe_compute_qgt) and berry =
 * -2 Im Q (vqe_compute_berry_curvature), each P_i x P_i row-major, instance
 * after instance; either may be NULL.  The P_i + 1 states of an instance
 * are held on the device at once (MOONLAB_SCHED_OOM past 16 GiB). */
MOONLAB_API int moonlab_mi355x_vqe_qgt(moonlab_mi355x_batch_t *b, const moonlab_mi355x_vqe_t *instances, int count,
                                       double *qgt, double *berry, moonlab_mi355x_batch_stats_t *stats);
/* Natural-gradient directions (g + regularization I)^-1 gradient per
 * instance (vqe_natural_gradient_direction, vqe_qng.c:264-282), gradients /
 * directions laid out as in moonlab_mi355x_vqe_gradients.  status (may be
 * NULL) gets 0, or -1 where the metric is singular -- that instance's
 * direction is then its gradient, the fallback the reference asks callers
 * to take. */
MOONLAB_API int moonlab_mi355x_vqe_natural_gradient(moonlab_mi355x_batch_t *b, const moonlab_mi355x_vqe_t *instances,
                                                    int count, const double *gradients, double regularization,
                                                    double *directions, int *status,
                                                    moonlab_mi355x_batch_stats_t *stats);

/* ---- Serving: continuous batching with HBM-footprint admission --------------
 *
 * A request loop in the spirit of the control plane's CIRCUIT / SHOTS verbs
 * (src/control/control_plane.c:587-748), which runs one synchronous CPU
 * simulation per request.  Here requests are queued and admitted while their
 * state footprint (2^n x 16 or 8 bytes) fits the HBM budget; small requests
 * (at most batch_max_qubits) are co-batched into persistent batch launches,
 * each batch taking everything admitted since the previous one was launched
 * (two batch engines alternate, so a batch is built while another runs);
 * larger requests run as fused blocks (or cache-blocked passes) on the large workers' streams.  A request's
 * results are the same alone and in any mix. */

typedef struct moonlab_mi355x_server moonlab_mi355x_server_t;

typedef struct {
    uint64_t hbm_budget_bytes;   /* admission budget, default 256 GiB */
    int      batch_max_qubits;   /* co-batch requests up to this size (default 16, max 24) */
    int      max_batch;          /* requests per batch launch (default 1024) */
    int      batch_engines;      /* concurrent batch launches (default 2) */
    int      large_workers;      /* concurrent large requests, each on its own stream (default 2) */
    int      sequential;         /* 1: plain dispatch -- one request at a time, one kernel per gate */
    int      large_path;         /* large requests: 0 cache-blocked register-phase passes
                                    (default), 1 fused <= 5-qubit blocks on MFMA */
} moonlab_mi355x_server_config_t;

typedef struct {
    int                       num_qubits;    /* 1..34 */
    int                       precision;     /* MOONLAB_HIP_C128 or MOONLAB_HIP_C64 */
    const moonlab_hip_gate_t *gates;         /* copied at submit */
    int                       num_gates;
    int                       num_shots;     /* draws: moonlab_qgtl_execute's xorshift64 stream */
    uint64_t                  rng_seed;      /* 0 = clock-based seed, as the reference */
    int                       want_expect_z; /* 1: <Z_q> for every qubit and the norm */
    int                       want_probabilities; /* 1: the 2^n probabilities (single-state path) */
} moonlab_mi355x_request_t;

typedef struct {
    int       status;          /* MOONLAB_SCHED_OK or a negative code */
    int       num_qubits;
    int       num_shots;
    uint64_t *outcomes;        /* malloc'd, num_shots */
    double   *expect_z;        /* malloc'd, num_qubits, or NULL */
    double    norm;
    double    submit_time;     /* seconds on the server clock (moonlab_mi355x_server_now) */
    double    start_time;
    double    finish_time;
    int       path;            /* 0 = co-batched, 1 = single-state, 2 = sequential */
    double   *probabilities;   /* malloc'd, 2^num_qubits, or NULL */
} moonlab_mi355x_response_t;

typedef struct {
    uint64_t submitted, completed, batches, batched_requests, large_requests;
    uint64_t peak_bytes_admitted;
    int      largest_batch;
} moonlab_mi355x_server_stats_t;

MOONLAB_API moonlab_mi355x_server_t *moonlab_mi355x_server_create(const moonlab_mi355x_server_config_t *cfg);
MOONLAB_API void moonlab_mi355x_server_free(moonlab_mi355x_server_t *s);
MOONLAB_API double moonlab_mi355x_server_now(const moonlab_mi355x_server_t *s);
MOONLAB_API int moonlab_mi355x_server_submit(moonlab_mi355x_server_t *s, const moonlab_mi355x_request_t *req,
                                             uint64_t *ticket);
/* Wait for a ticket's response (ownership passes to the caller; free with
 * moonlab_mi355x_response_free).  MOONLAB_SCHED_BACKEND_BUSY on timeout. */
MOONLAB_API int moonlab_mi355x_server_wait(moonlab_mi355x_server_t *s, uint64_t ticket,
                                           moonlab_mi355x_response_t *out, double timeout_s);
MOONLAB_API void moonlab_mi355x_response_free(moonlab_mi355x_response_t *r);
MOONLAB_API int moonlab_mi355x_server_stats(const moonlab_mi355x_server_t *s, moonlab_mi355x_server_stats_t *st);

/* ---- Controlled and multi-qubit gates through the circuit API -------------
 *
 * Addition: records one gate of codes 14..24 of the state-vector gate table
 * (moonlab_hip_gate_type_t: S^dagger, T^dagger, PHASE, U3, CPHASE, CRX, CRY,
 * CRZ, true CY, Toffoli, Fredkin; gates.c:341-1048) in a circuit, which
 * moonlab_qgtl_execute and the control-plane frames then run on the engine.
 * qubits[] in the order of moonlab_hip_gate_t's target, control and extra
 * fields (Toffoli: target, control, control 2; Fredkin: target, control,
 * target 2); params: theta (PHASE, CPHASE, CRX, CRY, CRZ) or theta, phi,
 * lambda (U3).  Returns MOONLAB_QGTL_BAD_ARG for another code, a wrong qubit
 * or parameter count, or a repeated or out-of-range qubit.  The text format
 * writes them as SDG, TDG, P, U3, CPHASE, CRX, CRY, CRZ, CYT, CCX and CSWAP
 * lines (qubits, then parameters as %.17g); the reference's parser rejects
 * those lines, so only circuits without them stay readable by it.  Gates of
 * these codes added through moonlab_qgtl_add_gate keep the reference's
 * behaviour: moonlab_qgtl_execute returns MOONLAB_QGTL_UNSUPPORTED. */
MOONLAB_API int moonlab_mi355x_circuit_add_gate_ex(moonlab_qgtl_circuit_t *c, int type, const int *qubits,
                                                   int num_qubits, const double *params, int num_params);

/* ---- The control plane's wire format as the serving front door ------------
 *
 * Status codes of src/control/control_plane.h:53-62. */
#ifndef MOONLAB_CONTROL_OK
#define MOONLAB_CONTROL_OK            ( 0)
#define MOONLAB_CONTROL_BAD_ARG       (-401)
#define MOONLAB_CONTROL_OOM           (-402)
#define MOONLAB_CONTROL_IO_ERROR      (-403)
#define MOONLAB_CONTROL_PROTOCOL      (-404)
#define MOONLAB_CONTROL_REJECTED      (-405)
#endif

/* Addition: one request frame of the control plane's wire protocol
 * (control_plane.h:9-33) through the serving loop.  Request frames:
 *   "CIRCUIT <bytes>\n" + moonlab-circuit v1 text  -> "OK <2^n>\n" + 2^n
 *                                                     little-endian doubles
 *   "SHOTS <shots> <bytes>\n" + the same text      -> "SAMPLES <shots>\n" +
 *                                                     shots uint64 outcomes
 *   "HEALTH\n"                                      -> "OK alive\n"
 * and "ERR <status> <message>\n" on failure, with the reference's limits
 * (4 MB body, 2^20 shots, MOONLAB_CONTROL_MAX_QUBITS qubits, default 24,
 * at most 30) and codes (control_plane.c:294-748).  The circuit runs in
 * complex128 through the server: concurrent calls, e.g. one per connection
 * thread of the reference's transport, are co-batched.  seed 0 draws the
 * shots from a clock-based seed as the control plane does; any other value
 * seeds the circuit API's xorshift64 stream, so the outcomes equal
 * moonlab_qgtl_execute's for that seed.  *out_response (malloc'd, free with
 * free()) holds the response frame to send, error frames included; returns
 * the MOONLAB_CONTROL_* status.  srv may be NULL for HEALTH and frames the
 * parser rejects. */
MOONLAB_API int moonlab_mi355x_server_handle_frame(moonlab_mi355x_server_t *srv, const void *frame, size_t len,
                                                   uint64_t seed, void **out_response, size_t *out_len);

/* The same with the connection's tenant (control_plane.c:699-731; NULL or ""
 * for none): while the request runs the tenant is the calling thread's
 * request context (moonlab_scheduler_set_request_context), and after a
 * completed CIRCUIT or SHOTS request the completion hook fires with
 * job = NULL, backend "control_plane", num_qubits and total_shots (0 for
 * CIRCUIT); the context is cleared before returning.
 * moonlab_mi355x_server_handle_frame is this call with tenant_id = NULL. */
MOONLAB_API int moonlab_mi355x_server_handle_frame_ex(moonlab_mi355x_server_t *srv, const void *frame, size_t len,
                                                      uint64_t seed, const char *tenant_id, void **out_response,
                                                      size_t *out_len);

#ifdef __cplusplus
}
#endif
#endif
