/*
 * sv_oracle.c -- CPU restatement of Moonlab's dense state-vector path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py; never by the product library.
 *
 * Every function restates the arithmetic of the reference function it cites
 * (file:line in tsotchke/moonlab v1.2.0), in the same per-element operation
 * order, so results agree with the reference's own complex128 CPU path to a
 * few ulp per gate.  Loops over disjoint amplitude blocks are OpenMP-parallel
 * (as in the reference, which threads the same loops); reductions run
 * sequentially in index order as the reference's do.
 *
 * Parity pinning (tests/test_oracle_cpu.py): every gate and reduction here is
 * checked against (1) the reference C compiled from its own sources
 * (oracle/_ref) over the reference's 75-circuit corpus and random circuits,
 * (2) committed golden fixtures written by that build, (3) closed forms
 * (Bell/GHZ/Grover-2/rotations on |0>), and (4) a brute-force dense unitary
 * for n <= 6.
 */
#include "sv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define RE(a, i) ((a)[2 * (size_t)(i)])
#define IM(a, i) ((a)[2 * (size_t)(i) + 1])

/* The reference threads a block loop only from 2^18 amplitudes on
 * (gates.c:121); below that a parallel region costs more than it saves. */
#define SVO_PAR_DIM (1ULL << 18)

static int qubit_ok(int n, int q) { return q >= 0 && q < n; }

void svo_init_zero(int n, double *amps)
{
    /* quantum_state_init: zero everything, amplitude[0] = 1 (state.c:39-81). */
    const uint64_t dim = 1ULL << n;
    memset(amps, 0, (size_t)dim * 2 * sizeof(double));
    amps[0] = 1.0;
}

/* ---------------------------------------------------------------------------
 * Per-pair kernels.  Each takes the pair (i0, i1 = i0 | 1<<q).
 * ------------------------------------------------------------------------- */

/* Loop every pair (i0 with bit q clear, i1 = i0 + 2^q) -- the stride-block
 * traversal of gates.c:250-266 flattened to a pair index. */
#define FOR_EACH_PAIR(n, q, ...)                                              \
    do {                                                                       \
        const uint64_t _dim = 1ULL << (n);                                     \
        const uint64_t _stride = 1ULL << (q);                                  \
        const int64_t _npairs = (int64_t)(_dim >> 1);                          \
        _Pragma("omp parallel for schedule(static) if (_dim >= SVO_PAR_DIM)") \
        for (int64_t _p = 0; _p < _npairs; _p++) {                             \
            const uint64_t p = (uint64_t)_p;                                   \
            const uint64_t i0 = ((p >> (q)) << ((q) + 1)) | (p & (_stride - 1)); \
            const uint64_t i1 = i0 + _stride;                                  \
            (void)i1;                                                          \
            __VA_ARGS__                                                        \
        }                                                                      \
    } while (0)

/* Multiply amplitude k by the complex phase (pr, pi):
 * (ar*pr - ai*pi, ar*pi + ai*pr)  (simd_apply_phase, simd_ops.c:926-1010). */
static inline void mul_phase(double *a, uint64_t k, double pr, double pi)
{
    const double ar = RE(a, k), ai = IM(a, k);
    RE(a, k) = ar * pr - ai * pi;
    IM(a, k) = ar * pi + ai * pr;
}

static void g_x(int n, double *a, int q)
{
    /* swap the two halves of every block (gate_pauli_x, gates.c:123-154). */
    FOR_EACH_PAIR(n, q, {
        const double r = RE(a, i0), m = IM(a, i0);
        RE(a, i0) = RE(a, i1); IM(a, i0) = IM(a, i1);
        RE(a, i1) = r;         IM(a, i1) = m;
    });
}

static void g_y(int n, double *a, int q)
{
    /* swap, then new a0 = -i * old a1, new a1 = +i * old a0
     * (gate_pauli_y, gates.c:156-192; simd_multiply_by_i simd_ops.c:823-873). */
    FOR_EACH_PAIR(n, q, {
        const double r0 = RE(a, i0), m0 = IM(a, i0);
        const double r1 = RE(a, i1), m1 = IM(a, i1);
        RE(a, i0) = m1;  IM(a, i0) = -r1;
        RE(a, i1) = -m0; IM(a, i1) = r0;
    });
}

static void g_z(int n, double *a, int q)
{
    /* negate the |1> half (gate_pauli_z, gates.c:194-224). */
    FOR_EACH_PAIR(n, q, {
        RE(a, i1) = -RE(a, i1); IM(a, i1) = -IM(a, i1);
    });
}

static void g_h(int n, double *a, int q)
{
    /* (a0 + a1) * INV_SQRT2, (a0 - a1) * INV_SQRT2 per component
     * (gate_hadamard gates.c:226-272; simd_hadamard_pair simd_ops.c:1026-1103). */
    const double r2 = 0.70710678118654752440;
    FOR_EACH_PAIR(n, q, {
        const double r0 = RE(a, i0), m0 = IM(a, i0);
        const double r1 = RE(a, i1), m1 = IM(a, i1);
        RE(a, i0) = (r0 + r1) * r2; IM(a, i0) = (m0 + m1) * r2;
        RE(a, i1) = (r0 - r1) * r2; IM(a, i1) = (m0 - m1) * r2;
    });
}

static void g_mul_i_on_one(int n, double *a, int q, int negate)
{
    /* S: |1> half times +i; S-dagger: times -i (gates.c:309-371). */
    FOR_EACH_PAIR(n, q, {
        const double r = RE(a, i1), m = IM(a, i1);
        if (negate) { RE(a, i1) = m;  IM(a, i1) = -r; }
        else        { RE(a, i1) = -m; IM(a, i1) = r;  }
    });
}

static void g_phase_on_one(int n, double *a, int q, double pr, double pi)
{
    /* qs_phase_on_one (gates.c:276-307): |1> half times a fixed phase. */
    FOR_EACH_PAIR(n, q, { mul_phase(a, i1, pr, pi); });
}

static void g_rx(int n, double *a, int q, double theta)
{
    /* a0' = c a0 - i s a1, a1' = c a1 - i s a0 with c, s = cos, sin(theta/2)
     * (gate_rx, gates.c:385-427), expanded into real arithmetic. */
    const double c = cos(theta / 2.0), s = sin(theta / 2.0);
    FOR_EACH_PAIR(n, q, {
        const double r0 = RE(a, i0), m0 = IM(a, i0);
        const double r1 = RE(a, i1), m1 = IM(a, i1);
        RE(a, i0) = c * r0 + s * m1; IM(a, i0) = c * m0 - s * r1;
        RE(a, i1) = c * r1 + s * m0; IM(a, i1) = c * m1 - s * r0;
    });
}

static void g_ry(int n, double *a, int q, double theta)
{
    /* a0' = c a0 - s a1, a1' = s a0 + c a1 (gate_ry, gates.c:429-479). */
    const double c = cos(theta / 2.0), s = sin(theta / 2.0);
    FOR_EACH_PAIR(n, q, {
        const double r0 = RE(a, i0), m0 = IM(a, i0);
        const double r1 = RE(a, i1), m1 = IM(a, i1);
        RE(a, i0) = c * r0 - s * r1; IM(a, i0) = c * m0 - s * m1;
        RE(a, i1) = s * r0 + c * r1; IM(a, i1) = s * m0 + c * m1;
    });
}

static void g_rz(int n, double *a, int q, double theta)
{
    /* |0> half times e^{-i theta/2}, |1> half times e^{+i theta/2}
     * (gate_rz, gates.c:481-518). */
    const double c = cos(theta / 2.0), s = sin(theta / 2.0);
    FOR_EACH_PAIR(n, q, {
        mul_phase(a, i0, c, -s);
        mul_phase(a, i1, c, s);
    });
}

void svo_apply_matrix_1q(int n, double *a, int q, const double m[8])
{
    /* new a0 = m00 a0 + m01 a1, new a1 = m10 a0 + m11 a1
     * (apply_single_qubit_gate, gates.c:1354-1393). */
    FOR_EACH_PAIR(n, q, {
        const double r0 = RE(a, i0), m0 = IM(a, i0);
        const double r1 = RE(a, i1), m1 = IM(a, i1);
        RE(a, i0) = (m[0] * r0 - m[1] * m0) + (m[2] * r1 - m[3] * m1);
        IM(a, i0) = (m[0] * m0 + m[1] * r0) + (m[2] * m1 + m[3] * r1);
        RE(a, i1) = (m[4] * r0 - m[5] * m0) + (m[6] * r1 - m[7] * m1);
        IM(a, i1) = (m[4] * m0 + m[5] * r0) + (m[6] * m1 + m[7] * r1);
    });
}

static void g_u3(int n, double *a, int q, double th, double ph, double la)
{
    /* U3 = [[c, -e^{i lam} s], [e^{i phi} s, e^{i(phi+lam)} c]]
     * (gate_u3, gates.c:520-565). */
    const double c = cos(th / 2.0), s = sin(th / 2.0);
    const double m[8] = {
        c, 0.0,
        -cos(la) * s, -sin(la) * s,
        cos(ph) * s, sin(ph) * s,
        cos(ph + la) * c, sin(ph + la) * c,
    };
    svo_apply_matrix_1q(n, a, q, m);
}

/* Quad loop: every index with bits lo, hi clear (lo < hi). */
#define FOR_EACH_QUAD(n, lo, hi, ...)                                         \
    do {                                                                       \
        const uint64_t _dim = 1ULL << (n);                                     \
        const int64_t _nq = (int64_t)(_dim >> 2);                              \
        _Pragma("omp parallel for schedule(static) if (_dim >= SVO_PAR_DIM)") \
        for (int64_t _p = 0; _p < _nq; _p++) {                                 \
            uint64_t b = (uint64_t)_p;                                         \
            b = ((b >> (lo)) << ((lo) + 1)) | (b & ((1ULL << (lo)) - 1));      \
            b = ((b >> (hi)) << ((hi) + 1)) | (b & ((1ULL << (hi)) - 1));      \
            __VA_ARGS__                                                        \
        }                                                                      \
    } while (0)

static void g_cnot(int n, double *a, int c, int t)
{
    /* swap the target pair inside the control = 1 subspace
     * (gate_cnot, gates.c:571-639). */
    const int lo = c < t ? c : t, hi = c < t ? t : c;
    const uint64_t cb = 1ULL << c, tb = 1ULL << t;
    FOR_EACH_QUAD(n, lo, hi, {
        const uint64_t i0 = b | cb, i1 = b | cb | tb;
        const double r = RE(a, i0), m = IM(a, i0);
        RE(a, i0) = RE(a, i1); IM(a, i0) = IM(a, i1);
        RE(a, i1) = r;         IM(a, i1) = m;
    });
}

static void g_cz(int n, double *a, int c, int t)
{
    /* negate where both bits are 1 (gate_cz, gates.c:641-668). */
    const int lo = c < t ? c : t, hi = c < t ? t : c;
    const uint64_t mk = (1ULL << c) | (1ULL << t);
    FOR_EACH_QUAD(n, lo, hi, {
        RE(a, b | mk) = -RE(a, b | mk); IM(a, b | mk) = -IM(a, b | mk);
    });
}

static void g_swap(int n, double *a, int q1, int q2)
{
    /* exchange |..1..0..> and |..0..1..> (gate_swap, gates.c:699-735). */
    if (q1 == q2) return;
    const int lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
    const uint64_t b1 = 1ULL << q1, b2 = 1ULL << q2;
    FOR_EACH_QUAD(n, lo, hi, {
        const uint64_t i0 = b | b1, i1 = b | b2;
        const double r = RE(a, i0), m = IM(a, i0);
        RE(a, i0) = RE(a, i1); IM(a, i0) = IM(a, i1);
        RE(a, i1) = r;         IM(a, i1) = m;
    });
}

static void g_cphase(int n, double *a, int c, int t, double theta)
{
    /* amplitude *= e^{i theta} where both bits are 1 (gate_cphase, gates.c:737-765). */
    const int lo = c < t ? c : t, hi = c < t ? t : c;
    const uint64_t mk = (1ULL << c) | (1ULL << t);
    const double pr = cos(theta), pi = sin(theta);
    FOR_EACH_QUAD(n, lo, hi, { mul_phase(a, b | mk, pr, pi); });
}

/* Controlled 2x2 on the target inside the control = 1 subspace
 * (gate_cy / gate_crx / gate_cry, gates.c:670-697, 800-864). */
static void g_controlled_1q(int n, double *a, int c, int t, const double m[8])
{
    const int lo = c < t ? c : t, hi = c < t ? t : c;
    const uint64_t cb = 1ULL << c, tb = 1ULL << t;
    FOR_EACH_QUAD(n, lo, hi, {
        const uint64_t i0 = b | cb, i1 = b | cb | tb;
        const double r0 = RE(a, i0), m0 = IM(a, i0);
        const double r1 = RE(a, i1), m1 = IM(a, i1);
        RE(a, i0) = (m[0] * r0 - m[1] * m0) + (m[2] * r1 - m[3] * m1);
        IM(a, i0) = (m[0] * m0 + m[1] * r0) + (m[2] * m1 + m[3] * r1);
        RE(a, i1) = (m[4] * r0 - m[5] * m0) + (m[6] * r1 - m[7] * m1);
        IM(a, i1) = (m[4] * m0 + m[5] * r0) + (m[6] * m1 + m[7] * r1);
    });
}

static void g_crz(int n, double *a, int c, int t, double theta)
{
    /* control = 1: |0>_t times e^{-i theta/2}, |1>_t times e^{+i theta/2}
     * (gate_crz, gates.c:866-899). */
    const int lo = c < t ? c : t, hi = c < t ? t : c;
    const uint64_t cb = 1ULL << c, tb = 1ULL << t;
    const double co = cos(theta / 2.0), si = sin(theta / 2.0);
    FOR_EACH_QUAD(n, lo, hi, {
        mul_phase(a, b | cb, co, -si);
        mul_phase(a, b | cb | tb, co, si);
    });
}

static void g_toffoli(int n, double *a, int c1, int c2, int t)
{
    /* flip the target where both controls are 1 (gate_toffoli, gates.c:905-932). */
    const uint64_t dim = 1ULL << n;
    const uint64_t cm = (1ULL << c1) | (1ULL << c2), tb = 1ULL << t;
    #pragma omp parallel for schedule(static) if (dim >= SVO_PAR_DIM)
    for (int64_t ii = 0; ii < (int64_t)dim; ii++) {
        const uint64_t i = (uint64_t)ii;
        if ((i & cm) == cm && !(i & tb)) {
            const uint64_t j = i | tb;
            const double r = RE(a, i), m = IM(a, i);
            RE(a, i) = RE(a, j); IM(a, i) = IM(a, j);
            RE(a, j) = r;        IM(a, j) = m;
        }
    }
}

static void g_fredkin(int n, double *a, int c, int t1, int t2)
{
    /* swap t1/t2 where the control is 1 (gate_fredkin, gates.c:934-967). */
    const uint64_t dim = 1ULL << n;
    const uint64_t cb = 1ULL << c, b1 = 1ULL << t1, b2 = 1ULL << t2;
    #pragma omp parallel for schedule(static) if (dim >= SVO_PAR_DIM)
    for (int64_t ii = 0; ii < (int64_t)dim; ii++) {
        const uint64_t i = (uint64_t)ii;
        if ((i & cb) && (i & b1) && !(i & b2)) {
            const uint64_t j = (i ^ b1) | b2;
            const double r = RE(a, i), m = IM(a, i);
            RE(a, i) = RE(a, j); IM(a, i) = IM(a, j);
            RE(a, j) = r;        IM(a, j) = m;
        }
    }
}

void svo_apply_matrix_2q(int n, double *a, int q1, int q2, const double m[32])
{
    /* 4x4 in basis |q2 q1> with qubit1 the LOW bit; rows computed as
     * sum over columns in order (apply_two_qubit_gate, gates.c:1395-1486). */
    const int lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
    const uint64_t b1 = 1ULL << q1, b2 = 1ULL << q2;
    FOR_EACH_QUAD(n, lo, hi, {
        const uint64_t idx[4] = { b, b | b1, b | b2, b | b1 | b2 };
        double vr[4], vi[4];
        for (int k = 0; k < 4; k++) { vr[k] = RE(a, idx[k]); vi[k] = IM(a, idx[k]); }
        for (int r = 0; r < 4; r++) {
            double sr = 0.0, si = 0.0;
            for (int col = 0; col < 4; col++) {
                const double mr = m[r * 8 + col * 2], mi = m[r * 8 + col * 2 + 1];
                sr += mr * vr[col] - mi * vi[col];
                si += mr * vi[col] + mi * vr[col];
            }
            RE(a, idx[r]) = sr; IM(a, idx[r]) = si;
        }
    });
}

int svo_apply_gate(int n, double *a, const svo_gate_t *g)
{
    const int t = g->target, c = g->control, e = g->extra;
    if (!qubit_ok(n, t)) return -1;
    switch (g->type) {
    case SVO_I: return 0;
    case SVO_X: g_x(n, a, t); return 0;
    case SVO_Y: g_y(n, a, t); return 0;
    case SVO_Z: g_z(n, a, t); return 0;
    case SVO_H: g_h(n, a, t); return 0;
    case SVO_S: g_mul_i_on_one(n, a, t, 0); return 0;
    case SVO_SDG: g_mul_i_on_one(n, a, t, 1); return 0;
    case SVO_T: g_phase_on_one(n, a, t, cos(M_PI / 4.0), sin(M_PI / 4.0)); return 0;
    case SVO_TDG: g_phase_on_one(n, a, t, cos(-M_PI / 4.0), sin(-M_PI / 4.0)); return 0;
    case SVO_PHASE: g_phase_on_one(n, a, t, cos(g->theta), sin(g->theta)); return 0;
    case SVO_RX: g_rx(n, a, t, g->theta); return 0;
    case SVO_RY: g_ry(n, a, t, g->theta); return 0;
    case SVO_RZ: g_rz(n, a, t, g->theta); return 0;
    case SVO_U3: g_u3(n, a, t, g->theta, g->phi, g->lam); return 0;
    default: break;
    }
    if (!qubit_ok(n, c) || c == t) return -1;
    switch (g->type) {
    case SVO_CNOT: g_cnot(n, a, c, t); return 0;
    case SVO_CZ: g_cz(n, a, c, t); return 0;
    case SVO_SWAP: g_swap(n, a, c, t); return 0;
    case SVO_CY:
        /* job path: S(target); CNOT(control, target); S-dagger(target)
         * (moonlab_qgtl_backend.c:138-144). */
        g_mul_i_on_one(n, a, t, 0);
        g_cnot(n, a, c, t);
        g_mul_i_on_one(n, a, t, 1);
        return 0;
    case SVO_CY_TRUE: {
        static const double Y[8] = { 0, 0, 0, -1, 0, 1, 0, 0 };
        g_controlled_1q(n, a, c, t, Y);
        return 0;
    }
    case SVO_CPHASE: g_cphase(n, a, c, t, g->theta); return 0;
    case SVO_CRX: {
        const double co = cos(g->theta / 2.0), si = sin(g->theta / 2.0);
        const double m[8] = { co, 0, 0, -si, 0, -si, co, 0 };
        g_controlled_1q(n, a, c, t, m);
        return 0;
    }
    case SVO_CRY: {
        const double co = cos(g->theta / 2.0), si = sin(g->theta / 2.0);
        const double m[8] = { co, 0, -si, 0, si, 0, co, 0 };
        g_controlled_1q(n, a, c, t, m);
        return 0;
    }
    case SVO_CRZ: g_crz(n, a, c, t, g->theta); return 0;
    default: break;
    }
    if (!qubit_ok(n, e) || e == t || e == c) return -1;
    switch (g->type) {
    case SVO_TOFFOLI: g_toffoli(n, a, c, e, t); return 0;
    case SVO_FREDKIN: g_fredkin(n, a, c, t, e); return 0;
    default: return -1;
    }
}

int svo_apply_gates(int n, double *a, const svo_gate_t *g, int num_gates)
{
    for (int i = 0; i < num_gates; i++) {
        const int rc = svo_apply_gate(n, a, &g[i]);
        if (rc) return rc;
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Reductions (src/quantum/measurement.c), sequential in index order.
 * ------------------------------------------------------------------------- */

static inline double prob_cabs(const double *a, uint64_t i)
{
    /* cabs(a) * cabs(a) -- hypot-based magnitude squared (measurement.c:101). */
    const double m = hypot(RE(a, i), IM(a, i));
    return m * m;
}

void svo_probabilities(int n, const double *a, double *p)
{
    /* measurement_probability_distribution (measurement.c:140-161). */
    const uint64_t dim = 1ULL << n;
    #pragma omp parallel for schedule(static) if (dim >= SVO_PAR_DIM)
    for (int64_t i = 0; i < (int64_t)dim; i++) p[i] = prob_cabs(a, (uint64_t)i);
}

double svo_probability_one(int n, const double *a, int q)
{
    /* measurement_probability_one (measurement.c:73-107). */
    const uint64_t dim = 1ULL << n, mk = 1ULL << q;
    double prob = 0.0;
    for (uint64_t i = 0; i < dim; i++)
        if (i & mk) prob += prob_cabs(a, i);
    return prob;
}

void svo_expect_z_all(int n, const double *a, double *z)
{
    /* <Z_q> = 1 - 2 P(1) (measurement.c:428-431), one sequential sum per qubit. */
    #pragma omp parallel for schedule(static)
    for (int q = 0; q < n; q++) z[q] = 1.0 - 2.0 * svo_probability_one(n, a, q);
}

double svo_expect_x(int n, const double *a, int q)
{
    /* sum over pairs j = i ^ 2^q > i of Re(conj(a_i) a_j + conj(a_j) a_i)
     * (measurement_expectation_x, measurement.c:442-464). */
    const uint64_t dim = 1ULL << n, mk = 1ULL << q;
    double e = 0.0;
    for (uint64_t i = 0; i < dim; i++) {
        const uint64_t j = i ^ mk;
        if (j > i) {
            const double x = RE(a, i) * RE(a, j) + IM(a, i) * IM(a, j);
            e += x + x;
        }
    }
    return e;
}

double svo_expect_y(int n, const double *a, int q)
{
    /* sum over pairs of Im(conj(a_i) a_j - conj(a_j) a_i), i has bit q clear
     * (measurement_expectation_y, measurement.c:473-497). */
    const uint64_t dim = 1ULL << n, mk = 1ULL << q;
    double e = 0.0;
    for (uint64_t i = 0; i < dim; i++) {
        const uint64_t j = i ^ mk;
        if (j > i) {
            const double y = RE(a, i) * IM(a, j) - IM(a, i) * RE(a, j);
            e += y + y;
        }
    }
    return e;
}

double svo_expect_pauli(int n, const double *a, const char *pauli)
{
    /* vqe_measure_pauli_expectation's exact path (vqe.c:2156-2222): rotate
     * a copy into the measurement basis -- H for X, S-dagger then H for Y
     * (:2185-2199) -- then sum |psi'(x)|^2 (cabs squared) times the parity
     * of x over the non-identity positions (:2214-2226). */
    const uint64_t dim = 1ULL << n;
    double *t = (double *)malloc(dim * 2 * sizeof(double));
    if (!t) return 0.0;
    memcpy(t, a, dim * 2 * sizeof(double));
    uint64_t mask = 0;
    for (int q = 0; q < n && pauli[q]; q++) {
        svo_gate_t g = {0, q, -1, -1, 0.0, 0.0, 0.0};
        switch (pauli[q]) {
        case 'X':
            g.type = SVO_H;
            svo_apply_gate(n, t, &g);
            break;
        case 'Y':
            g.type = SVO_SDG;
            svo_apply_gate(n, t, &g);
            g.type = SVO_H;
            svo_apply_gate(n, t, &g);
            break;
        default: break;
        }
        if (pauli[q] == 'X' || pauli[q] == 'Y' || pauli[q] == 'Z') mask |= 1ULL << q;
    }
    double e = 0.0;
    for (uint64_t x = 0; x < dim; x++) {
        const double p = prob_cabs(t, x);
        e += (__builtin_popcountll(x & mask) & 1) ? -p : p;
    }
    free(t);
    return e;
}

double svo_correlation_zz(int n, const double *a, int q1, int q2)
{
    /* P00 + P11 - P01 - P10 from four sequential sums
     * (measurement_correlation_zz, measurement.c:509-534). */
    const uint64_t dim = 1ULL << n, m1 = 1ULL << q1, m2 = 1ULL << q2;
    double p00 = 0.0, p01 = 0.0, p10 = 0.0, p11 = 0.0;
    for (uint64_t i = 0; i < dim; i++) {
        const double p = prob_cabs(a, i);
        const int b1 = (i & m1) != 0, b2 = (i & m2) != 0;
        if (!b1 && !b2) p00 += p;
        else if (!b1 && b2) p01 += p;
        else if (b1 && !b2) p10 += p;
        else p11 += p;
    }
    return p00 + p11 - p01 - p10;
}

double svo_norm_sq(int n, const double *a)
{
    const uint64_t dim = 1ULL << n;
    double s = 0.0;
    for (uint64_t i = 0; i < dim; i++) s += RE(a, i) * RE(a, i) + IM(a, i) * IM(a, i);
    return s;
}

double svo_fidelity(int n, const double *a, const double *b)
{
    /* |<a|b>|^2 (quantum_state_fidelity, state.c:289-306). */
    const uint64_t dim = 1ULL << n;
    double sr = 0.0, si = 0.0;
    for (uint64_t i = 0; i < dim; i++) {
        sr += RE(a, i) * RE(b, i) + IM(a, i) * IM(b, i);
        si += RE(a, i) * IM(b, i) - IM(a, i) * RE(b, i);
    }
    return sr * sr + si * si;
}

uint64_t svo_measure_all_outcome(int n, const double *a, double u)
{
    /* First i with u < cumulative; if none, the last index with nonzero
     * probability, else 0 (measurement_all_qubits, measurement.c:263-307). */
    const uint64_t dim = 1ULL << n;
    double cum = 0.0;
    uint64_t last_nz = 0;
    int have_nz = 0;
    for (uint64_t i = 0; i < dim; i++) {
        const double p = prob_cabs(a, i);
        if (p > 0.0) { last_nz = i; have_nz = 1; }
        cum += p;
        if (u < cum) return i;
    }
    return have_nz ? last_nz : 0;
}

/* Collapsing measurements (measurement.c:181-231, 263-307, 320-410).  Norm
 * floor MEASUREMENT_SMALL_NORM = 1e-15 (measurement.c:33). */
#define SVO_SMALL_NORM 1e-15

uint64_t svo_measure_all(int n, double *a, double u)
{
    /* measurement_all_qubits: the outcome rule above, then |outcome> with
     * amplitude exactly 1 (measurement.c:300-302). */
    const uint64_t r = svo_measure_all_outcome(n, a, u);
    const uint64_t dim = 1ULL << n;
    for (uint64_t i = 0; i < dim; i++) {
        RE(a, i) = i == r ? 1.0 : 0.0;
        IM(a, i) = 0.0;
    }
    return r;
}

int svo_measure_qubit(int n, double *a, int q, double u)
{
    /* measurement_single_qubit: outcome = u < P(1) (P(1) un-normalized,
     * measurement.c:197-200); zero the other half while summing the kept
     * half's cabs^2 in index order; below the norm floor return -1 with the
     * state left as the reference leaves it (zeroed, not rescaled); else
     * scale by 1 / sqrt(norm) (measurement.c:201-228). */
    if (!qubit_ok(n, q)) return -1;
    const uint64_t dim = 1ULL << n, mk = 1ULL << q;
    const int r = u < svo_probability_one(n, a, q) ? 1 : 0;
    double nf = 0.0;
    for (uint64_t i = 0; i < dim; i++) {
        if (((i & mk) ? 1 : 0) != r) {
            RE(a, i) = 0.0;
            IM(a, i) = 0.0;
        } else {
            nf += prob_cabs(a, i);
        }
    }
    if (nf <= SVO_SMALL_NORM) return -1;
    const double inv = 1.0 / sqrt(nf);
    for (uint64_t i = 0; i < dim; i++) {
        RE(a, i) *= inv;
        IM(a, i) *= inv;
    }
    return r;
}

uint64_t svo_measure_partial(int n, double *a, const int *qubits, int k, double u)
{
    /* measurement_partial: outcome bit j = qubit qubits[j]; outcome
     * probabilities accumulated in index order, the cumulative rule with the
     * last-nonzero fallback, then zero / renormalize (renormalization skipped
     * at or below the floor); 0 for invalid arguments (measurement.c:320-410). */
    if (k <= 0 || k > 30) return 0;
    for (int j = 0; j < k; j++)
        if (!qubit_ok(n, qubits[j])) return 0;
    const uint64_t dim = 1ULL << n, no = 1ULL << k;
    double *pr = (double *)calloc(no, sizeof(double));
    if (!pr) return 0;
    for (uint64_t i = 0; i < dim; i++) {
        uint64_t o = 0;
        for (int j = 0; j < k; j++)
            if (i & (1ULL << qubits[j])) o |= 1ULL << j;
        pr[o] += prob_cabs(a, i);
    }
    double cum = 0.0;
    uint64_t r = 0, last_nz = 0;
    int found = 0, have_nz = 0;
    for (uint64_t o = 0; o < no; o++) {
        if (pr[o] > 0.0) { last_nz = o; have_nz = 1; }
        cum += pr[o];
        if (u < cum) { r = o; found = 1; break; }
    }
    if (!found) r = have_nz ? last_nz : 0;
    double nf = 0.0;
    for (uint64_t i = 0; i < dim; i++) {
        uint64_t o = 0;
        for (int j = 0; j < k; j++)
            if (i & (1ULL << qubits[j])) o |= 1ULL << j;
        if (o != r) {
            RE(a, i) = 0.0;
            IM(a, i) = 0.0;
        } else {
            nf += prob_cabs(a, i);
        }
    }
    if (nf > SVO_SMALL_NORM) {
        const double inv = 1.0 / sqrt(nf);
        for (uint64_t i = 0; i < dim; i++) {
            RE(a, i) *= inv;
            IM(a, i) *= inv;
        }
    }
    free(pr);
    return r;
}

/* ---------------------------------------------------------------------------
 * Job-path PRNGs and shot protocol.
 * ------------------------------------------------------------------------- */

uint64_t svo_splitmix64(uint64_t x)
{
    /* The scheduler's seed mixer, constants exactly as scheduler.c:143-148
     * (its second multiplier is 0x94d49bb133111ebb, which differs from the
     * textbook SplitMix64 finaliser; shot parity needs the reference's). */
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d49bb133111ebbULL;
    return x ^ (x >> 31);
}

uint64_t svo_xorshift64(uint64_t *s)
{
    /* xorshift64 13/7/17 (moonlab_qgtl_backend.c:159-167). */
    uint64_t x = *s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *s = x;
    return x;
}

double svo_u_from_xorshift(uint64_t *s)
{
    /* top 53 bits / 2^53 (moonlab_qgtl_backend.c:238-239). */
    const uint64_t r = svo_xorshift64(s);
    return (double)(r >> 11) / (double)(1ULL << 53);
}

void svo_qgtl_shots(int n, const double *a, int num_shots, uint64_t seed, uint64_t *out)
{
    /* one u per shot, outcome by the collapse rule on an unchanged state
     * (the per-shot clone in moonlab_qgtl_backend.c:227-242 leaves the
     * post-circuit state untouched between shots). */
    uint64_t s = seed;
    for (int k = 0; k < num_shots; k++) out[k] = svo_measure_all_outcome(n, a, svo_u_from_xorshift(&s));
}

int svo_job_shots(int n, const double *a, int total, int workers, uint64_t seed, uint64_t *out)
{
    /* worker count clamp, contiguous slices and per-worker seeds
     * (scheduler.c:194-239, 152-189). */
    if (total <= 0 || workers < 1 || seed == 0) return -1;
    if (workers > total) workers = total;
    for (int w = 0; w < workers; w++) {
        const int start = (int)((long long)w * total / workers);
        const int end = (int)((long long)(w + 1) * total / workers);
        if (end - start <= 0) continue;
        const uint64_t sw = svo_splitmix64(seed ^ ((uint64_t)w << 32));
        svo_qgtl_shots(n, a, end - start, sw, out + start);
    }
    return workers;
}

/* ---------------------------------------------------------------------------
 * Batched-VQE workload: hardware-efficient ansatz + TFIM energy.
 * ------------------------------------------------------------------------- */

double svo_hea_tfim_energy(int n, int layers, const double *ang, double h, double *amps)
{
    /* Layer = RY, RZ on every qubit then CZ(q, q+1) for q = 0..n-2 (the
     * RY/RZ-per-qubit + entangling-ladder shape of vqe.c:764-781 with CZ as
     * the entangler); energy from measurement_correlation_zz and
     * measurement_expectation_x (measurement.c:442-534). */
    svo_init_zero(n, amps);
    for (int l = 0; l < layers; l++) {
        for (int q = 0; q < n; q++) {
            g_ry(n, amps, q, ang[((size_t)l * n + q) * 2 + 0]);
            g_rz(n, amps, q, ang[((size_t)l * n + q) * 2 + 1]);
        }
        for (int q = 0; q + 1 < n; q++) g_cz(n, amps, q, q + 1);
    }
    double e = 0.0;
    for (int q = 0; q + 1 < n; q++) e -= svo_correlation_zz(n, amps, q, q + 1);
    double sx = 0.0;
    for (int q = 0; q < n; q++) sx += svo_expect_x(n, amps, q);
    return e - h * sx;
}

/* ---------------------------------------------------------------------------
 * Variational energies, gradients and the quantum geometric tensor.
 * ------------------------------------------------------------------------- */

static void vqe_ansatz(int n, int layers, const double *ang, int entangler, uint64_t hf, double *amps,
                       long insert_at, int insert_gen)
{
    /* vqe_compute_energy's state preparation (vqe.c:1302-1318) over the
     * hardware-efficient ansatz (vqe.c:764-781); with insert_at >= 0 the
     * generator of parameter insert_at (Y after RY, Z after RZ) is applied
     * right after its rotation, qng_hea_state_derivatives (vqe_qng.c:66-103). */
    svo_init_zero(n, amps);
    for (int q = 0; q < n; q++)
        if ((hf >> q) & 1ULL) g_x(n, amps, q);
    long p = 0;
    for (int l = 0; l < layers; l++) {
        for (int q = 0; q < n; q++) {
            g_ry(n, amps, q, ang[p]);
            if (p == insert_at && insert_gen) g_y(n, amps, q);
            p++;
            g_rz(n, amps, q, ang[p]);
            if (p == insert_at && insert_gen) g_z(n, amps, q);
            p++;
        }
        for (int q = 0; q + 1 < n; q++) {
            if (entangler == 1) g_cnot(n, amps, q, q + 1);
            else g_cz(n, amps, q, q + 1);
        }
    }
}

void svo_vqe_state(int n, int layers, const double *ang, int entangler, uint64_t hf, double *amps)
{
    vqe_ansatz(n, layers, ang, entangler, hf, amps, -1, 0);
}

double svo_vqe_energy(int n, int layers, const double *ang, int entangler, uint64_t hf, int nterms,
                      const char *const *strings, const double *coeffs, double constant)
{
    const size_t dim = (size_t)1 << n;
    double *a = (double *)malloc(dim * 2 * sizeof(double));
    if (!a) return NAN;
    vqe_ansatz(n, layers, ang, entangler, hf, a, -1, 0);
    double e = 0.0;   /* term order, then the constant (vqe.c:1326-1345) */
    for (int k = 0; k < nterms; k++) e += coeffs[k] * svo_expect_pauli(n, a, strings[k]);
    e += constant;
    free(a);
    return e;
}

void svo_vqe_gradient(int n, int layers, const double *ang, int entangler, uint64_t hf, int nterms,
                      const char *const *strings, const double *coeffs, double *grad)
{
    const int P = 2 * n * layers;
    double *t = (double *)malloc((size_t)(P > 0 ? P : 1) * sizeof(double));
    if (!t) return;
    for (int k = 0; k < P; k++) {
        memcpy(t, ang, (size_t)P * sizeof(double));
        t[k] = ang[k] + M_PI / 2.0;
        const double ep = svo_vqe_energy(n, layers, t, entangler, hf, nterms, strings, coeffs, 0.0);
        t[k] = ang[k] - M_PI / 2.0;
        const double em = svo_vqe_energy(n, layers, t, entangler, hf, nterms, strings, coeffs, 0.0);
        grad[k] = 0.5 * (ep - em);
    }
    free(t);
}

int svo_vqe_qgt(int n, int layers, const double *ang, int entangler, uint64_t hf, double *qgt, double *berry)
{
    const int P = 2 * n * layers;
    const size_t dim = (size_t)1 << n;
    double *psi = (double *)malloc(dim * 2 * sizeof(double));
    double *d = (double *)malloc((size_t)(P > 0 ? P : 1) * dim * 2 * sizeof(double));
    double *ov = (double *)malloc((size_t)(P > 0 ? P : 1) * 2 * sizeof(double));
    if (!psi || !d || !ov) {
        free(psi); free(d); free(ov);
        return -1;
    }
    vqe_ansatz(n, layers, ang, entangler, hf, psi, -1, 0);
    for (int i = 0; i < P; i++) {
        double *di = d + (size_t)i * dim * 2;
        vqe_ansatz(n, layers, ang, entangler, hf, di, i, 1);
        for (size_t x = 0; x < dim; x++) {   /* (-i/2) (re + i im) = (im / 2) - i (re / 2) */
            const double re = di[2 * x], im = di[2 * x + 1];
            di[2 * x] = 0.5 * im;
            di[2 * x + 1] = -0.5 * re;
        }
        double sr = 0.0, si = 0.0;   /* <psi|d_i psi> */
        for (size_t x = 0; x < dim; x++) {
            sr += psi[2 * x] * di[2 * x] + psi[2 * x + 1] * di[2 * x + 1];
            si += psi[2 * x] * di[2 * x + 1] - psi[2 * x + 1] * di[2 * x];
        }
        ov[2 * i] = sr;
        ov[2 * i + 1] = si;
    }
    for (int i = 0; i < P; i++) {
        const double *di = d + (size_t)i * dim * 2;
        for (int j = i; j < P; j++) {
            const double *dj = d + (size_t)j * dim * 2;
            double ar = 0.0, ai = 0.0;   /* <d_i psi|d_j psi> */
            for (size_t x = 0; x < dim; x++) {
                ar += di[2 * x] * dj[2 * x] + di[2 * x + 1] * dj[2 * x + 1];
                ai += di[2 * x] * dj[2 * x + 1] - di[2 * x + 1] * dj[2 * x];
            }
            /* conj(ov_i) ov_j */
            const double br = ov[2 * i] * ov[2 * j] + ov[2 * i + 1] * ov[2 * j + 1];
            const double bi = ov[2 * i] * ov[2 * j + 1] - ov[2 * i + 1] * ov[2 * j];
            if (qgt) {
                qgt[(size_t)i * P + j] = ar - br;
                qgt[(size_t)j * P + i] = ar - br;
            }
            if (berry) {
                const double f = i == j ? 0.0 : -2.0 * (ai - bi);
                berry[(size_t)i * P + j] = f;
                berry[(size_t)j * P + i] = -f;
            }
        }
    }
    free(psi); free(d); free(ov);
    return 0;
}

int svo_solve(int m, double *A, const double *b, double *x)
{
    double *r = (double *)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    if (!r) return -1;
    memcpy(r, b, (size_t)m * sizeof(double));
    for (int c = 0; c < m; c++) {
        int piv = c;
        for (int i = c + 1; i < m; i++)
            if (fabs(A[(size_t)i * m + c]) > fabs(A[(size_t)piv * m + c])) piv = i;
        if (!(fabs(A[(size_t)piv * m + c]) > 0.0)) {
            free(r);
            return -1;
        }
        if (piv != c) {
            for (int k = 0; k < m; k++) {
                const double t = A[(size_t)c * m + k];
                A[(size_t)c * m + k] = A[(size_t)piv * m + k];
                A[(size_t)piv * m + k] = t;
            }
            const double t = r[c];
            r[c] = r[piv];
            r[piv] = t;
        }
        for (int i = c + 1; i < m; i++) {
            const double f = A[(size_t)i * m + c] / A[(size_t)c * m + c];
            for (int k = c; k < m; k++) A[(size_t)i * m + k] -= f * A[(size_t)c * m + k];
            r[i] -= f * r[c];
        }
    }
    for (int i = m - 1; i >= 0; i--) {
        double s = r[i];
        for (int k = i + 1; k < m; k++) s -= A[(size_t)i * m + k] * x[k];
        x[i] = s / A[(size_t)i * m + i];
    }
    free(r);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Fused-block composition (brute force, FP64).
 * ------------------------------------------------------------------------- */

int svo_compose_unitary(int k, const int *qubits, const svo_gate_t *g, int ng, double *mat)
{
    if (k < 1 || k > 6) return -1;
    const int dim = 1 << k;
    int local[64];
    for (int i = 0; i < 64; i++) local[i] = -1;
    for (int j = 0; j < k; j++) {
        if (qubits[j] < 0 || qubits[j] >= 64) return -1;
        local[qubits[j]] = j;
    }
    double *v = (double *)malloc((size_t)dim * 2 * sizeof(double));
    if (!v) return -1;
    for (int col = 0; col < dim; col++) {
        memset(v, 0, (size_t)dim * 2 * sizeof(double));
        v[2 * col] = 1.0;
        for (int i = 0; i < ng; i++) {
            svo_gate_t lg = g[i];
            lg.target = local[g[i].target];
            if (lg.target < 0) { free(v); return -1; }
            if (g[i].type >= SVO_CNOT && g[i].type != SVO_SDG && g[i].type != SVO_TDG &&
                g[i].type != SVO_PHASE && g[i].type != SVO_U3) {
                lg.control = local[g[i].control];
                if (lg.control < 0) { free(v); return -1; }
                if (g[i].type == SVO_TOFFOLI || g[i].type == SVO_FREDKIN) {
                    lg.extra = local[g[i].extra];
                    if (lg.extra < 0) { free(v); return -1; }
                }
            }
            if (svo_apply_gate(k, v, &lg)) { free(v); return -1; }
        }
        for (int r = 0; r < dim; r++) {
            mat[((size_t)r * dim + col) * 2] = v[2 * r];
            mat[((size_t)r * dim + col) * 2 + 1] = v[2 * r + 1];
        }
    }
    free(v);
    return 0;
}
