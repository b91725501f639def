/*
 * moonlab_hip_statevec.h -- MI355X (gfx950) dense state-vector engine, C ABI.
 *
 * This is the device layer the reference reaches through its CUDA backend
 * (src/backends/cuda_statevec.h) and its weak GPU-route seam
 * (src/quantum/gates.c:43-79).  Every moonlab_cuda_* entry point has a
 * moonlab_hip_* counterpart with the same argument meaning, layout and
 * status values (cited per entry); the rest are additions the job path,
 * the batch engine and the serving loop need (reductions for all qubits in
 * one pass, shot sampling, complex64 states, the wider gate set).
 *
 * Layout: 2^n interleaved (re, im) amplitudes in HBM, qubit q = bit q of the
 * basis index.  Gate matrices are row-major (re, im) pairs: 2x2 = 8 doubles,
 * 4x4 = 32 doubles in the basis |q1 q0> with q0 the LOW bit
 * (cuda_statevec.h:104-115).
 *
 * Every state owns one HIP stream (default priority).  Gate calls enqueue
 * and return; reductions, copies and sampling synchronise that stream.
 * Distinct states may be driven from distinct threads concurrently; one
 * state must not be used from two threads at once.
 */
#ifndef MOONLAB_HIP_STATEVEC_H
#define MOONLAB_HIP_STATEVEC_H

#include <stddef.h>
#include <stdint.h>

#ifndef MOONLAB_API
#define MOONLAB_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct moonlab_hip_state moonlab_hip_state_t;

/* Status values (cuda_statevec.h:40-46), plus BAD_ARG. */
typedef enum {
    MOONLAB_HIP_OK                = 0,
    MOONLAB_HIP_ERR_OUT_OF_MEMORY = -1,
    MOONLAB_HIP_ERR_NO_DEVICE     = -2,
    MOONLAB_HIP_ERR_DRIVER        = -3,
    MOONLAB_HIP_ERR_INVALID_QUBIT = -4,
    MOONLAB_HIP_ERR_NUM_QUBITS    = -5,
    MOONLAB_HIP_ERR_BAD_ARG       = -6
} moonlab_hip_status_t;

/* Amplitude precision. */
enum {
    MOONLAB_HIP_C128 = 0,   /* complex128: 16 B per amplitude */
    MOONLAB_HIP_C64  = 1    /* complex64:   8 B per amplitude */
};

/* Largest state one MI355X holds: 2^34 x 16 B = 256 GiB of its 288 GB. */
#define MOONLAB_HIP_MAX_QUBITS_C128 34
#define MOONLAB_HIP_MAX_QUBITS_C64  35

/* Gate codes.  0..13 are moonlab_qgtl_gate_t with the job path's meaning
 * (CY = S; CNOT; S-dagger as moonlab_qgtl_backend.c:138-144 applies it);
 * 14.. are the wider gate_* family of src/quantum/gates.h. */
enum {
    MOONLAB_HIP_GATE_I = 0, MOONLAB_HIP_GATE_X = 1, MOONLAB_HIP_GATE_Y = 2,
    MOONLAB_HIP_GATE_Z = 3, MOONLAB_HIP_GATE_H = 4, MOONLAB_HIP_GATE_S = 5,
    MOONLAB_HIP_GATE_T = 6, MOONLAB_HIP_GATE_RX = 7, MOONLAB_HIP_GATE_RY = 8,
    MOONLAB_HIP_GATE_RZ = 9, MOONLAB_HIP_GATE_CNOT = 10, MOONLAB_HIP_GATE_CY = 11,
    MOONLAB_HIP_GATE_CZ = 12, MOONLAB_HIP_GATE_SWAP = 13,
    MOONLAB_HIP_GATE_SDG = 14,      /* gate_s_dagger  (gates.c:341) */
    MOONLAB_HIP_GATE_TDG = 15,      /* gate_t_dagger  (gates.c:377) */
    MOONLAB_HIP_GATE_PHASE = 16,    /* gate_phase     (gates.c:381), theta */
    MOONLAB_HIP_GATE_U3 = 17,       /* gate_u3        (gates.c:520), theta phi lam */
    MOONLAB_HIP_GATE_CPHASE = 18,   /* gate_cphase    (gates.c:737), theta */
    MOONLAB_HIP_GATE_CRX = 19,      /* gate_crx       (gates.c:800), theta */
    MOONLAB_HIP_GATE_CRY = 20,      /* gate_cry       (gates.c:833), theta */
    MOONLAB_HIP_GATE_CRZ = 21,      /* gate_crz       (gates.c:866), theta */
    MOONLAB_HIP_GATE_CY_TRUE = 22,  /* gate_cy        (gates.c:670): controlled-Y */
    MOONLAB_HIP_GATE_TOFFOLI = 23,  /* gate_toffoli   (gates.c:905): control, extra = control 2 */
    MOONLAB_HIP_GATE_FREDKIN = 24   /* gate_fredkin   (gates.c:934): control, target, extra = target 2 */
};

/* One gate of a gate list.  `control` is also SWAP's first qubit. */
typedef struct {
    int32_t type;
    int32_t target;
    int32_t control;
    int32_t extra;
    double  theta;
    double  phi;
    double  lam;
} moonlab_hip_gate_t;

/* ---- State lifetime (cuda_statevec.h:61-81) --------------------------- */

/* Replaces moonlab_cuda_state_create (cuda_statevec.h:61-62): complex128,
 * initialised to |0...0>.  n_qubits 1..34. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_state_create(uint32_t n_qubits,
                                                          moonlab_hip_state_t **out_state);
/* Addition: choose complex128 or complex64 (1..35 qubits). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_state_create_ex(uint32_t n_qubits, int precision,
                                                             moonlab_hip_state_t **out_state);
/* Replaces moonlab_cuda_state_free (cuda_statevec.h:67).  Safe on NULL. */
MOONLAB_API void moonlab_hip_state_free(moonlab_hip_state_t *state);
/* Replace moonlab_cuda_state_n_qubits / _dim (cuda_statevec.h:72-81). */
MOONLAB_API uint32_t moonlab_hip_state_n_qubits(const moonlab_hip_state_t *state);
MOONLAB_API uint64_t moonlab_hip_state_dim(const moonlab_hip_state_t *state);
/* Additions. */
MOONLAB_API int moonlab_hip_state_precision(const moonlab_hip_state_t *state);
MOONLAB_API moonlab_hip_status_t moonlab_hip_state_reset(moonlab_hip_state_t *state);

/* ---- Gates (cuda_statevec.h:84-214) ----------------------------------- */

/* Replaces moonlab_cuda_apply_1q (cuda_statevec.h:84-97). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_1q(moonlab_hip_state_t *state,
                                                      uint32_t target, const double m[8]);
/* Replaces moonlab_cuda_apply_cnot (cuda_statevec.h:99-103). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_cnot(moonlab_hip_state_t *state,
                                                        uint32_t control, uint32_t target);
/* Replaces moonlab_cuda_apply_2q (cuda_statevec.h:104-118): basis |q1 q0>. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_2q(moonlab_hip_state_t *state, uint32_t q0,
                                                      uint32_t q1, const double m[32]);
/* Replace moonlab_cuda_apply_x / _h (cuda_statevec.h:138-150). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_x(moonlab_hip_state_t *state, uint32_t target);
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_h(moonlab_hip_state_t *state, uint32_t target);
/* Replaces moonlab_cuda_apply_mcx (cuda_statevec.h:180-190). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_mcx(moonlab_hip_state_t *state,
                                                       uint64_t control_mask, uint32_t target);
/* Replaces moonlab_cuda_apply_mcz (cuda_statevec.h:192-200). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_mcz(moonlab_hip_state_t *state,
                                                       uint64_t all_mask);
/* Replaces moonlab_cuda_apply_fredkin (cuda_statevec.h:202-210). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_fredkin(moonlab_hip_state_t *state,
                                                           uint32_t control, uint32_t t1,
                                                           uint32_t t2);
/* Additions: any gate of the code table; a gate list; diag(d0, d1) on one
 * qubit (d = {re0, im0, re1, im1}); a 2x2 on `target` applied where every
 * bit of control_mask is 1. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_gate(moonlab_hip_state_t *state,
                                                        const moonlab_hip_gate_t *gate);
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_gates(moonlab_hip_state_t *state,
                                                         const moonlab_hip_gate_t *gates,
                                                         int num_gates);
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_diag_1q(moonlab_hip_state_t *state,
                                                           uint32_t target, const double d[4]);
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_controlled_1q(moonlab_hip_state_t *state,
                                                                 uint64_t control_mask,
                                                                 uint32_t target,
                                                                 const double m[8]);

/* Addition: apply a dense 2^k x 2^k matrix (row-major (re, im) doubles,
 * local basis bit j = qubits[j], qubits strictly ascending, k = 1..5): the
 * fused-block primitive.  complex64 states contract on FP32 matrix cores. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_dense(moonlab_hip_state_t *state, int k,
                                                         const int *qubits, const double *mat);
/* Addition: apply a gate list fused into blocks of at most kmax (1..5)
 * qubits; each block's unitary is composed in FP64 on the host. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_gates_fused(moonlab_hip_state_t *state,
                                                               const moonlab_hip_gate_t *gates,
                                                               int num_gates, int kmax);

/* Addition: the fusion plan moonlab_hip_apply_gates_fused would execute,
 * computed on the host (no device needed).  Block i acts on the qubits set in
 * masks[i] (its local basis bit j is the j-th lowest of them) and holds
 * counts[i] gates; when mats is not NULL its FP64 unitary (2^k x 2^k,
 * row-major (re, im) doubles) is written at mats + 2048 i.  Qubits must lie
 * in 0..63.  Returns the block count (only the first max_blocks are written)
 * or a negative moonlab_hip_status_t value. */
MOONLAB_API int moonlab_hip_fusion_plan(const moonlab_hip_gate_t *gates, int num_gates, int kmax,
                                        uint64_t *masks, int *counts, double *mats, int max_blocks);

/* Addition: apply a gate list cache-blocked: single-qubit runs fold into one
 * FP64 2x2 per qubit, then ops are packed into passes whose non-diagonal
 * targets fit a 12-qubit tile set; each pass streams the state through LDS
 * in 4096-amplitude tiles once and applies all of its ops there. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_apply_gates_blocked(moonlab_hip_state_t *state,
                                                                 const moonlab_hip_gate_t *gates,
                                                                 int num_gates);
/* Addition: how many passes over the state moonlab_hip_apply_gates_blocked
 * makes for this gate list on an n-qubit state of the given precision (host
 * only, no device needed), or a negative moonlab_hip_status_t value. */
MOONLAB_API int moonlab_hip_blocked_pass_count(const moonlab_hip_gate_t *gates, int num_gates, int n,
                                               int precision);
/* Addition: tile (11, 12 or 13 qubits; the register-phase engine uses 12 or
 * 13 -- by default 12 for complex128 and 13 for complex64 -- the LDS op engine
 * 11 or 12) and the LDS op engine's workgroup size (256 or 512) of the
 * blocked path; 0 keeps a value, tile_bits -1 restores the default by
 * precision.  Process-wide, like the gate tuning.  Different tiles may
 * reorder commuting gates, so results agree to rounding; one setting gives
 * the same result on every call. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_set_blocked_tuning(int tile_bits, int threads);
/* Addition: engine of the blocked path for states of at least 12 qubits:
 * 0 = register phases (default: each thread keeps 16 tile amplitudes in
 * registers, LDS only changes which tile qubits those are), 1 = LDS ops (one
 * LDS sweep per op or register segment), 2 = register phases (same as 0),
 * 3 = register phases with 8 amplitudes per thread (twice the waves, three
 * register qubits per phase).  Smaller states use the LDS engine. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_set_blocked_engine(int engine);

/* Addition: how the blocked path's planner picks each pass's tile set and
 * each phase's register set.  windows > 0 (default 8): the first `windows`
 * pending ops seed candidate windows of neighbouring qubits (three per seed:
 * starting at, centred on and ending at the op's target) and the set taking
 * the most ops wins (about 2.8x fewer passes than the greedy choice on
 * 30-qubit random layers of depth 40); 0: both sets grow greedily in
 * circuit order.  Process-wide, like the other tuning calls. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_set_blocked_lookahead(int windows);

/* ---- Reductions ---------------------------------------------------------- */

/* Replaces moonlab_cuda_prob_z (cuda_statevec.h:120-128): P(qubit = 1). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_prob_z(const moonlab_hip_state_t *state,
                                                    uint32_t target, double *out_prob);
/* Replaces moonlab_cuda_norm (cuda_statevec.h:130-136): sum |a|^2. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_norm(const moonlab_hip_state_t *state,
                                                  double *out_norm);
/* Addition: <Z_q> = 1 - 2 P(q = 1) for every qubit, and sum |a|^2, from ONE
 * pass over the state, accumulated in FP64 (either output may be NULL). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_expect_z_all(const moonlab_hip_state_t *state,
                                                          double *out_z, double *out_norm);
/* Addition with no counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the sums of moonlab_hip_expect_z_all on
 * a state the caller owns, enqueued on `stream` (NULL = the default stream)
 * with no allocation, no host copy and no synchronization.  `amplitudes` is
 * the device address of 2^n_qubits complex values (precision
 * MOONLAB_HIP_C128 or MOONLAB_HIP_C64) and `out` the device address of
 * n_qubits + 1 doubles, both on the current device: out[0] = sum |a|^2,
 * out[1 + q] = <Z_q> = 1 - 2 P(q = 1), bit-identical to
 * moonlab_hip_expect_z_all.  The block partials live in a fixed 576 KiB
 * device array per GPU (2048 blocks x 36 doubles), so n_qubits is limited to
 * 34 (complex128) / 35 (complex64); larger values return
 * MOONLAB_HIP_ERR_BAD_ARG.  Calls that overlap in time on one GPU share that
 * array and race: serialize them (one stream, or events). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_expect_z_all_noalloc(const void *amplitudes, uint32_t n_qubits,
                                                                  int precision, void *stream, double *out);
/* Additions: measurement_expectation_x/_y and measurement_correlation_zz
 * (src/quantum/measurement.c:442-534), FP64 accumulation. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_expect_x(const moonlab_hip_state_t *state,
                                                      uint32_t q, double *out);
MOONLAB_API moonlab_hip_status_t moonlab_hip_expect_y(const moonlab_hip_state_t *state,
                                                      uint32_t q, double *out);
MOONLAB_API moonlab_hip_status_t moonlab_hip_correlation_zz(const moonlab_hip_state_t *state,
                                                            uint32_t q1, uint32_t q2,
                                                            double *out);
/* Addition: |a_i|^2 for all i into a host buffer of 2^n doubles. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_probabilities(const moonlab_hip_state_t *state,
                                                           double *out);
/* Addition: <P> for the Pauli string with X on the qubits of x_mask only,
 * Z on z_mask only and Y on both, in one pass (FP64 sums).  Replaces the
 * exact path of vqe_measure_pauli_expectation (vqe.c:2156-2222), which clones
 * the state, rotates it (H, S-dagger H) and sums parities. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_expect_pauli(const moonlab_hip_state_t *state,
                                                          uint64_t x_mask, uint64_t z_mask,
                                                          double *out);
/* Addition: <a|b> (same n; either precision, products and sums in FP64, so a
 * complex64 run can be compared with a complex128 one without a copy). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_inner_product(const moonlab_hip_state_t *a,
                                                           const moonlab_hip_state_t *b,
                                                           double *out_re, double *out_im);
/* Addition: scale the state by 1/sqrt(sum |a|^2) (FP64 norm). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_normalize(moonlab_hip_state_t *state);

/* ---- Sampling -------------------------------------------------------------- */

/* Addition: for each uniform draw u[k] in [0, 1), the outcome rule of
 * measurement_all_qubits (src/quantum/measurement.c:263-307): the first
 * basis index i with u < sum_{j <= i} |a_j|^2, else the last index of
 * nonzero probability.  The state is not collapsed. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_sample(const moonlab_hip_state_t *state,
                                                    const double *u, int num_samples,
                                                    uint64_t *out);

/* ---- Collapsing measurements ------------------------------------------------ */

/* Replaces measurement_single_qubit (src/quantum/measurement.c:181-231): the
 * outcome is 1 when u < P(1) (the unnormalized sum over the qubit's |1> half),
 * else 0; the other half is zeroed and the kept half scaled by 1 / sqrt(its
 * norm), FP64.  When the kept norm is at most 1e-15 (MEASUREMENT_SMALL_NORM)
 * *outcome is -1 and the state is left zeroed but not rescaled, as the
 * reference leaves it. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_measure_qubit(moonlab_hip_state_t *state, uint32_t qubit,
                                                           double u, int *outcome);

/* Replaces measurement_partial (src/quantum/measurement.c:320-410): outcome
 * bit j is qubit qubits[j]; the 2^k outcome probabilities (FP64, one device
 * pass), the first outcome whose cumulative probability exceeds u (else the
 * last outcome of nonzero probability), then collapse and renormalization
 * (skipped at or below the 1e-15 norm floor).  k in 1..20 distinct qubits
 * (the reference also accepts repeats and larger k). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_measure_partial(moonlab_hip_state_t *state, const int *qubits,
                                                             int k, double u, uint64_t *outcome);

/* Replaces measurement_all_qubits (src/quantum/measurement.c:263-307): the
 * outcome of moonlab_hip_sample for u, then the state becomes that basis
 * state with amplitude exactly 1. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_measure_all(moonlab_hip_state_t *state, double u,
                                                         uint64_t *outcome);

/* ---- Host transfer (cuda_statevec.h:152-178) ------------------------------- */

/* Replaces moonlab_cuda_state_copy_to_host (cuda_statevec.h:152-160): 2 * dim
 * doubles, interleaved; complex64 states are widened to doubles. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_state_copy_to_host(const moonlab_hip_state_t *state,
                                                                double *out);
/* Replaces moonlab_cuda_state_copy_from_host (cuda_statevec.h:162-178). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_state_copy_from_host(moonlab_hip_state_t *state,
                                                                  const double *in);
/* Addition: raw copy of amplitudes [offset, offset + count) in the state's own
 * precision (16 B or 8 B each) into `out`. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_state_read(const moonlab_hip_state_t *state,
                                                        uint64_t offset, uint64_t count,
                                                        void *out);
/* Addition: device-side clone (quantum_state_clone, src/quantum/state.c:
 * 164-191, without a host round trip): dst takes src's amplitudes, converted
 * when the precisions differ.  Same qubit count required. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_state_copy(moonlab_hip_state_t *dst,
                                                        const moonlab_hip_state_t *src);
/* Addition: device pointer and stream (hipStream_t) of the state. */
MOONLAB_API void *moonlab_hip_state_device_ptr(moonlab_hip_state_t *state);
MOONLAB_API void *moonlab_hip_state_stream(moonlab_hip_state_t *state);

/* Replaces moonlab_cuda_synchronize (cuda_statevec.h:212-218). */
MOONLAB_API moonlab_hip_status_t moonlab_hip_synchronize(moonlab_hip_state_t *state);

/* ---- Timing on the state's stream ----------------------------------------- */

/* Addition: allocate `capacity` HIP events on the state's stream; each mark
 * records the next one; elapsed returns the (count - 1) gaps in ms after
 * synchronising.  Used to time each kernel launch of a sweep. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_timer_begin(moonlab_hip_state_t *state,
                                                         int capacity);
MOONLAB_API moonlab_hip_status_t moonlab_hip_timer_mark(moonlab_hip_state_t *state);
MOONLAB_API int moonlab_hip_timer_elapsed(moonlab_hip_state_t *state, float *gaps_ms,
                                          int max_gaps);

/* ---- Device / tuning -------------------------------------------------------- */

/* Addition: number of visible HIP devices (0 when none). */
MOONLAB_API int moonlab_hip_device_count(void);
/* Addition: last error text of the calling thread ("" when none). */
MOONLAB_API const char *moonlab_hip_last_error(void);
/* Addition: launch shape of the streaming gate kernels: amplitude pairs per
 * thread (1, 2, 4, 8) and threads per block (64..1024).  0 keeps a value. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_set_gate_tuning(int pairs_per_thread,
                                                             int block_threads);

/* Addition: states of at least `bytes` use streaming (non-temporal) loads
 * and stores in the gate kernels; smaller ones keep cached access so
 * consecutive gates can hit L2 / MALL.  Default 512 MiB. */
MOONLAB_API moonlab_hip_status_t moonlab_hip_set_streaming_threshold(uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
