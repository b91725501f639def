/*
 * ref_extra.c -- thin glue linked INTO oracle/_ref/libmoonlab_ref.so beside the
 * reference's own sources, so ctypes can drive the reference's gate_* and
 * measurement_* functions by gate code.  TEST INFRASTRUCTURE ONLY.
 *
 * Gate codes are the oracle's (oracle/sv_oracle.h); codes 0..13 keep the job
 * path's meaning (CY = S; CNOT; S-dagger, moonlab_qgtl_backend.c:138-144).
 */
#include "quantum/state.h"
#include "quantum/gates.h"
#include "quantum/measurement.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t type, target, control, extra;
    double theta, phi, lam;
} ref_gate_t;

quantum_state_t *ref_state_new(int n)
{
    quantum_state_t *s = (quantum_state_t *)calloc(1, sizeof(*s));
    if (!s) return NULL;
    if (quantum_state_init(s, (size_t)n) != QS_SUCCESS) { free(s); return NULL; }
    return s;
}

void ref_state_delete(quantum_state_t *s)
{
    if (!s) return;
    quantum_state_free(s);
    free(s);
}

double *ref_state_amplitudes(quantum_state_t *s) { return (double *)s->amplitudes; }

int ref_apply_gate(quantum_state_t *s, const ref_gate_t *g)
{
    const int t = g->target, c = g->control, e = g->extra;
    switch (g->type) {
    case 0: return 0;
    case 1: return gate_pauli_x(s, t);
    case 2: return gate_pauli_y(s, t);
    case 3: return gate_pauli_z(s, t);
    case 4: return gate_hadamard(s, t);
    case 5: return gate_s(s, t);
    case 6: return gate_t(s, t);
    case 7: return gate_rx(s, t, g->theta);
    case 8: return gate_ry(s, t, g->theta);
    case 9: return gate_rz(s, t, g->theta);
    case 10: return gate_cnot(s, c, t);
    case 11: {
        int rc = gate_s(s, t);
        if (rc == 0) rc = gate_cnot(s, c, t);
        if (rc == 0) rc = gate_s_dagger(s, t);
        return rc;
    }
    case 12: return gate_cz(s, c, t);
    case 13: return gate_swap(s, c, t);
    case 14: return gate_s_dagger(s, t);
    case 15: return gate_t_dagger(s, t);
    case 16: return gate_phase(s, t, g->theta);
    case 17: return gate_u3(s, t, g->theta, g->phi, g->lam);
    case 18: return gate_cphase(s, c, t, g->theta);
    case 19: return gate_crx(s, c, t, g->theta);
    case 20: return gate_cry(s, c, t, g->theta);
    case 21: return gate_crz(s, c, t, g->theta);
    case 22: return gate_cy(s, c, t);
    case 23: return gate_toffoli(s, c, e, t);
    case 24: return gate_fredkin(s, c, t, e);
    default: return -1;
    }
}

int ref_apply_gates(quantum_state_t *s, const ref_gate_t *g, int ng)
{
    for (int i = 0; i < ng; i++) {
        const int rc = ref_apply_gate(s, &g[i]);
        if (rc) return rc;
    }
    return 0;
}

/* Run a gate list from |0..0> and hand back amplitudes, <Z_q> for every qubit
 * and the probability vector (either output may be NULL). */
int ref_run(int n, const ref_gate_t *g, int ng, double *amps_out, double *z_out, double *p_out)
{
    quantum_state_t *s = ref_state_new(n);
    if (!s) return -1;
    int rc = ref_apply_gates(s, g, ng);
    if (rc == 0) {
        const size_t dim = (size_t)1 << n;
        if (amps_out) memcpy(amps_out, s->amplitudes, dim * 2 * sizeof(double));
        if (z_out)
            for (int q = 0; q < n; q++) z_out[q] = measurement_expectation_z(s, q);
        if (p_out) measurement_probability_distribution(s, p_out);
    }
    ref_state_delete(s);
    return rc;
}

/* ---- VQE glue: the reference's solver over its hardware-efficient ansatz
 * (vqe.c:764-781: RY, RZ per qubit, then CNOT(q, q+1)) and a Pauli-sum
 * Hamiltonian given as strings (pauli_string[q] acts on qubit q), with the
 * Hartree-Fock bits and the constant (nuclear repulsion) of
 * pauli_hamiltonian_t.  Drives vqe_compute_energy / _gradient / _qgt /
 * _berry_curvature / vqe_natural_gradient_direction (vqe.c, vqe_qng.c). */
#include "algorithms/vqe.h"

typedef struct {
    pauli_hamiltonian_t *h;
    vqe_ansatz_t *a;
    vqe_optimizer_t *o;
    quantum_entropy_ctx_t ent;
    vqe_solver_t *s;
} ref_vqe_t;

static int ref_vqe_zero_bytes(void *user, uint8_t *buf, size_t n)
{
    (void)user;
    memset(buf, 0, n);
    return 0;
}

ref_vqe_t *ref_vqe_new(int n, int layers, int nterms, const char *const *strings, const double *coeffs,
                       uint64_t hf_reference, double constant)
{
    ref_vqe_t *v = (ref_vqe_t *)calloc(1, sizeof(*v));
    if (!v) return NULL;
    v->h = pauli_hamiltonian_create((size_t)n, (size_t)nterms);
    v->a = vqe_create_hardware_efficient_ansatz((size_t)n, (size_t)layers);
    v->o = vqe_optimizer_create(VQE_OPTIMIZER_GRADIENT_DESCENT);
    if (!v->h || !v->a || !v->o) goto fail;
    for (int k = 0; k < nterms; k++)
        if (pauli_hamiltonian_add_term(v->h, coeffs[k], strings[k], (size_t)k) != 0) goto fail;
    v->h->hf_reference = hf_reference;
    v->h->nuclear_repulsion = constant;
    quantum_entropy_init(&v->ent, ref_vqe_zero_bytes, NULL);
    v->s = vqe_solver_create(v->h, v->a, v->o, &v->ent);
    if (!v->s) goto fail;
    return v;
fail:
    if (v->h) pauli_hamiltonian_free(v->h);
    if (v->a) vqe_ansatz_free(v->a);
    if (v->o) vqe_optimizer_free(v->o);
    free(v);
    return NULL;
}

void ref_vqe_free(ref_vqe_t *v)
{
    if (!v) return;
    vqe_solver_free(v->s);
    pauli_hamiltonian_free(v->h);
    vqe_ansatz_free(v->a);
    vqe_optimizer_free(v->o);
    free(v);
}

int ref_vqe_num_parameters(const ref_vqe_t *v) { return (int)v->a->num_parameters; }
double ref_vqe_energy(ref_vqe_t *v, const double *p) { return vqe_compute_energy(v->s, p); }
int ref_vqe_gradient(ref_vqe_t *v, const double *p, double *g) { return vqe_compute_gradient(v->s, p, g); }
int ref_vqe_qgt(ref_vqe_t *v, const double *p, double *out) { return vqe_compute_qgt(v->s, p, out); }
int ref_vqe_berry(ref_vqe_t *v, const double *p, double *out) { return vqe_compute_berry_curvature(v->s, p, out); }
int ref_vqe_natural_gradient(ref_vqe_t *v, const double *p, const double *g, double reg, double *out)
{
    return vqe_natural_gradient_direction(v->s, p, g, reg, out);
}
