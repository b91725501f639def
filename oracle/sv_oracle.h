/*
 * sv_oracle.h -- CPU restatement of Moonlab's dense state-vector semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * It is never linked into, called by, or used as a fallback for the product
 * library (moonlab_amd/libmoonlab_mi355x.so).
 *
 * State layout matches the reference: 2^n interleaved (re, im) doubles,
 * qubit q = bit q of the basis index (src/quantum/state.h:102-131).
 */
#ifndef SV_ORACLE_H
#define SV_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate codes.  0..13 equal moonlab_qgtl_gate_t
 * (src/applications/moonlab_qgtl_backend.h:57-72) with the JOB-PATH meaning
 * of each tag (CY = S; CNOT; S-dagger, moonlab_qgtl_backend.c:138-144).
 * 14.. are the wider gate_* family of src/quantum/gates.h. */
enum {
    SVO_I = 0, SVO_X = 1, SVO_Y = 2, SVO_Z = 3, SVO_H = 4, SVO_S = 5, SVO_T = 6,
    SVO_RX = 7, SVO_RY = 8, SVO_RZ = 9, SVO_CNOT = 10, SVO_CY = 11, SVO_CZ = 12,
    SVO_SWAP = 13,
    SVO_SDG = 14, SVO_TDG = 15, SVO_PHASE = 16, SVO_U3 = 17, SVO_CPHASE = 18,
    SVO_CRX = 19, SVO_CRY = 20, SVO_CRZ = 21, SVO_CY_TRUE = 22,
    SVO_TOFFOLI = 23, SVO_FREDKIN = 24
};

/* One gate.  `control` is the control (or SWAP's qubit1); `extra` is the
 * second control of TOFFOLI or the second target of FREDKIN. */
typedef struct {
    int32_t type;
    int32_t target;
    int32_t control;
    int32_t extra;
    double  theta;
    double  phi;
    double  lam;
} svo_gate_t;

/* Initialise |0...0> into amps (2 * 2^n doubles). */
void svo_init_zero(int n, double *amps);

/* Apply one gate in place.  Returns 0, or -1 on an invalid qubit/type. */
int svo_apply_gate(int n, double *amps, const svo_gate_t *g);

/* Apply a gate list in place; returns the first error or 0. */
int svo_apply_gates(int n, double *amps, const svo_gate_t *g, int num_gates);

/* Generic matrices (src/quantum/gates.c:1354-1486).  m2 = 2x2 row-major
 * (re, im) pairs; m4 = 4x4 row-major in the basis |q2 q1> with qubit1 the
 * LOW bit (CPU convention, gates.c:1417-1456). */
void svo_apply_matrix_1q(int n, double *amps, int q, const double m2[8]);
void svo_apply_matrix_2q(int n, double *amps, int q1, int q2, const double m4[32]);

/* Reductions (src/quantum/measurement.c). */
void   svo_probabilities(int n, const double *amps, double *p_out);   /* cabs^2 */
double svo_probability_one(int n, const double *amps, int q);
void   svo_expect_z_all(int n, const double *amps, double *z_out);
double svo_expect_x(int n, const double *amps, int q);
double svo_expect_y(int n, const double *amps, int q);
double svo_correlation_zz(int n, const double *amps, int q1, int q2);
/* <P> for a Pauli string (pauli[q] in "IXYZ" acts on qubit q), the exact
 * path of vqe_measure_pauli_expectation (vqe.c:2156-2222). */
double svo_expect_pauli(int n, const double *amps, const char *pauli);
double svo_norm_sq(int n, const double *amps);
double svo_fidelity(int n, const double *a, const double *b);

/* Collapse-free "measure all" outcome for a uniform draw u
 * (measurement.c:263-307 outcome rule, without the collapse write). */
uint64_t svo_measure_all_outcome(int n, const double *amps, double u);

/* Collapsing measurements: measurement_all_qubits (-> |outcome>),
 * measurement_single_qubit (outcome or -1 below the norm floor) and
 * measurement_partial (outcome bits in the order of `qubits`),
 * measurement.c:181-231, 263-307, 320-410. */
uint64_t svo_measure_all(int n, double *amps, double u);
int svo_measure_qubit(int n, double *amps, int q, double u);
uint64_t svo_measure_partial(int n, double *amps, const int *qubits, int k, double u);

/* PRNGs of the job path. */
uint64_t svo_splitmix64(uint64_t x);
uint64_t svo_xorshift64(uint64_t *s);
double   svo_u_from_xorshift(uint64_t *s);

/* moonlab_qgtl_execute's shot loop for one seed (moonlab_qgtl_backend.c:218-243). */
void svo_qgtl_shots(int n, const double *amps, int num_shots, uint64_t seed,
                    uint64_t *out);

/* moonlab_scheduler_run's "simulator" fan-out (scheduler.c:194-239):
 * worker slices and per-worker seeds, outcomes in slice order.  seed != 0. */
int svo_job_shots(int n, const double *amps, int total_shots, int num_workers,
                  uint64_t seed, uint64_t *out);

/* Hardware-efficient ansatz (RY, RZ per qubit then CZ(q, q+1) ladder, per
 * layer) from |0..0>, then the open-chain TFIM energy
 *   E = -sum_i <Z_i Z_{i+1}> - h sum_i <X_i>
 * using measurement_correlation_zz / measurement_expectation_x semantics.
 * angles: layers * n * 2 doubles, laid out [layer][qubit][ry, rz]. */
double svo_hea_tfim_energy(int n, int layers, const double *angles, double h,
                           double *amps_scratch);

/* Variational energies over a Pauli-sum Hamiltonian (vqe.c:1287-1350): the
 * Hartree-Fock bits set by X (vqe.c:1302-1305), then per layer RY, RZ on
 * every qubit and the entangling ladder over (q, q+1) -- entangler 1: CNOT
 * with control q (the reference ansatz, vqe.c:764-781), 0: CZ -- then
 * E = sum_k c_k <P_k> (svo_expect_pauli, term order) + constant.
 * angles: layers * n * 2 doubles, [layer][qubit][ry, rz].  strings[k][q] in
 * "IXYZ" acts on qubit q. */
void   svo_vqe_state(int n, int layers, const double *angles, int entangler, uint64_t hf, double *amps);
double svo_vqe_energy(int n, int layers, const double *angles, int entangler, uint64_t hf, int nterms,
                      const char *const *strings, const double *coeffs, double constant);
/* Parameter-shift gradient, dE/dtheta_k = [E(theta + pi/2 e_k) - E(theta -
 * pi/2 e_k)] / 2 (vqe.c:2094-2130). */
void   svo_vqe_gradient(int n, int layers, const double *angles, int entangler, uint64_t hf, int nterms,
                        const char *const *strings, const double *coeffs, double *grad);
/* The quantum geometric tensor of the ansatz state (vqe_qng.c:66-170): exact
 * derivatives d_i psi = (-i/2) U_after P_i U_upto |HF> (P_i = Y after RY, Z
 * after RZ), ov_i = <psi|d_i psi>, Q_ij = <d_i psi|d_j psi> - <d_i psi|psi>
 * <psi|d_j psi>; qgt = Re Q (vqe_qng.c:172-205), berry = -2 Im Q (:207-245),
 * each P x P row-major (P = 2 n layers), either may be NULL. */
int    svo_vqe_qgt(int n, int layers, const double *angles, int entangler, uint64_t hf, double *qgt,
                   double *berry);
/* Solve (A) x = b for an m x m row-major A (overwritten): Gaussian
 * elimination with partial pivoting; -1 when a pivot vanishes. */
int    svo_solve(int m, double *A, const double *b, double *x);

/* FP64 composition of a list of gates acting only on `qubits` (k <= 5,
 * ascending) into a 2^k x 2^k row-major complex matrix in the local basis
 * (local bit j = qubits[j]).  Brute force: applies the gates column by
 * column to basis vectors of a k-qubit register.  Returns 0 or -1. */
int svo_compose_unitary(int k, const int *qubits, const svo_gate_t *g,
                        int num_gates, double *mat_out);

#ifdef __cplusplus
}
#endif
#endif
