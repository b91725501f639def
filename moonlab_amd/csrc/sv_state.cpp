// sv_state.cpp -- the moonlab_hip_* C ABI: device state lifetime, gate
// dispatch, reductions, sampling, host transfer and stream timing.
//
// Replaces the reference's moonlab_cuda_* layer (src/backends/cuda_statevec.cu)
// with plain device memory (no managed memory, no host mirror) and one HIP
// stream per state.  Reductions finish on the device and return only n + 1
// doubles to the host.
#include "sv_internal.h"

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

using namespace mlx;

struct moonlab_hip_state {
    int         n;
    int         prec;
    uint64_t    dim;
    size_t      bytes;
    void       *amps;
    hipStream_t stream;
    int         device;
    // reduction / sampling workspace, allocated on first use
    ReduceWs    ws;
    double     *h_out;       // pinned, 64 doubles
    double     *d_u;
    uint64_t   *d_outcomes;
    int         shot_cap;
    // stream timing
    std::vector<hipEvent_t> events;
    int         next_event;
};

namespace {

thread_local std::string t_last_error;

moonlab_hip_status_t fail(moonlab_hip_status_t st, const char *what, hipError_t e = hipSuccess)
{
    char buf[256];
    if (e != hipSuccess)
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else
        snprintf(buf, sizeof buf, "%s", what);
    t_last_error = buf;
    return st;
}

moonlab_hip_status_t drv(hipError_t e, const char *what)
{
    if (e == hipSuccess) return MOONLAB_HIP_OK;
    if (e == hipErrorOutOfMemory) return fail(MOONLAB_HIP_ERR_OUT_OF_MEMORY, what, e);
    return fail(MOONLAB_HIP_ERR_DRIVER, what, e);
}

SvDev dev(const moonlab_hip_state_t *s)
{
    return SvDev{s->amps, s->n, s->prec, s->stream};
}

bool qubit_ok(const moonlab_hip_state_t *s, int64_t q) { return q >= 0 && q < s->n; }

size_t amp_bytes(int prec) { return prec == MOONLAB_HIP_C64 ? 8 : 16; }

moonlab_hip_status_t ensure_ws(moonlab_hip_state_t *s)
{
    if (s->ws.partial) return MOONLAB_HIP_OK;
    ReduceWs &w = s->ws;
    reduce_geometry(s->n, &w.tile_bits, &w.ntiles, &w.blocks);
    const size_t nvals = (size_t)s->n + 1;
    const int pblocks = w.blocks > 2048 ? w.blocks : 2048;   // also serves the X/Y/ZZ/inner sums
    hipError_t e = hipMalloc((void **)&w.partial, (size_t)pblocks * (nvals > 2 ? nvals : 2) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&w.out, 64 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&w.tile_sums, w.ntiles * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&w.tile_incl, w.ntiles * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->h_out, 64 * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        hipFree(w.partial); hipFree(w.out); hipFree(w.tile_sums); hipFree(w.tile_incl);
        if (s->h_out) hipHostFree(s->h_out);
        w.partial = w.out = w.tile_sums = w.tile_incl = nullptr;
        s->h_out = nullptr;
        return drv(e, "reduction workspace");
    }
    return MOONLAB_HIP_OK;
}

// Run the all-qubit reduction and copy n + 1 doubles back.
moonlab_hip_status_t reduce_all(const moonlab_hip_state_t *cs, bool tiles)
{
    moonlab_hip_state_t *s = const_cast<moonlab_hip_state_t *>(cs);
    moonlab_hip_status_t st = ensure_ws(s);
    if (st) return st;
    hipError_t e = launch_reduce_z(dev(s), s->ws, tiles);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->h_out, s->ws.out, (s->n + 1) * sizeof(double), hipMemcpyDeviceToHost,
                           s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return drv(e, "reduce");
}

moonlab_hip_status_t fetch_small(const moonlab_hip_state_t *s, int count)
{
    hipError_t e = hipMemcpyAsync(s->h_out, s->ws.out, count * sizeof(double), hipMemcpyDeviceToHost,
                                  s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return drv(e, "reduce");
}

moonlab_hip_status_t check_gate(const moonlab_hip_state_t *s, const moonlab_hip_gate_t &g)
{
    if (g.type < 0 || g.type > MOONLAB_HIP_GATE_FREDKIN)
        return fail(MOONLAB_HIP_ERR_BAD_ARG, "unknown gate code");
    if (g.type == MOONLAB_HIP_GATE_I) return qubit_ok(s, g.target) ? MOONLAB_HIP_OK
                                                                  : fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "qubit out of range");
    if (!qubit_ok(s, g.target)) return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "target out of range");
    const bool two = (g.type >= MOONLAB_HIP_GATE_CNOT && g.type <= MOONLAB_HIP_GATE_SWAP) ||
                     g.type >= MOONLAB_HIP_GATE_CPHASE;
    if (two) {
        if (!qubit_ok(s, g.control) || g.control == g.target)
            return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "control out of range or equal to target");
    }
    if (g.type == MOONLAB_HIP_GATE_TOFFOLI || g.type == MOONLAB_HIP_GATE_FREDKIN) {
        if (!qubit_ok(s, g.extra) || g.extra == g.target || g.extra == g.control)
            return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "third qubit out of range or repeated");
    }
    return MOONLAB_HIP_OK;
}

}  // namespace

namespace mlx {

moonlab_hip_status_t set_last_error(moonlab_hip_status_t st, const char *what, hipError_t e)
{
    return fail(st, what, e);
}

hipStream_t state_bind_stream(moonlab_hip_state_t *s, hipStream_t st)
{
    hipStream_t old = s->stream;
    if (old != st) {
        hipStreamSynchronize(old);
        s->stream = st;
    }
    return old;
}

}  // namespace mlx

extern "C" {

int moonlab_hip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *moonlab_hip_last_error(void) { return t_last_error.c_str(); }

moonlab_hip_status_t moonlab_hip_set_gate_tuning(int ept, int block)
{
    GateTuning &t = gate_tuning();
    if (ept) {
        if (ept != 1 && ept != 2 && ept != 4 && ept != 8) return fail(MOONLAB_HIP_ERR_BAD_ARG, "ept");
        t.ept = ept;
    }
    if (block) {
        if (block < 64 || block > 1024 || (block & 63)) return fail(MOONLAB_HIP_ERR_BAD_ARG, "block");
        t.block = block;
    }
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_set_blocked_tuning(int tile_bits, int threads)
{
    GateTuning &t = gate_tuning();
    if (tile_bits) {
        if (tile_bits != -1 && (tile_bits < 11 || tile_bits > 13))
            return fail(MOONLAB_HIP_ERR_BAD_ARG, "tile_bits is 11, 12, 13 or -1 (by precision)");
        t.blocked_tile_bits = tile_bits == -1 ? 0 : tile_bits;
    }
    if (threads) {
        if (threads != 256 && threads != 512) return fail(MOONLAB_HIP_ERR_BAD_ARG, "threads is 256 or 512");
        t.blocked_threads = threads;
    }
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_set_blocked_lookahead(int windows)
{
    if (windows < 0 || windows > 1024) return fail(MOONLAB_HIP_ERR_BAD_ARG, "lookahead windows 0..1024");
    gate_tuning().blocked_lookahead = windows;
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_set_blocked_engine(int engine)
{
    if (engine < 0 || engine > 3) return fail(MOONLAB_HIP_ERR_BAD_ARG, "engine is 0, 1, 2 or 3");
    gate_tuning().blocked_engine = engine;
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_set_streaming_threshold(uint64_t bytes)
{
    gate_tuning().nt_min_bytes = bytes;
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_state_create_ex(uint32_t n, int prec, moonlab_hip_state_t **out)
{
    if (!out) return fail(MOONLAB_HIP_ERR_BAD_ARG, "out_state is NULL");
    *out = nullptr;
    if (prec != MOONLAB_HIP_C128 && prec != MOONLAB_HIP_C64)
        return fail(MOONLAB_HIP_ERR_BAD_ARG, "precision");
    const uint32_t cap = prec == MOONLAB_HIP_C64 ? MOONLAB_HIP_MAX_QUBITS_C64 : MOONLAB_HIP_MAX_QUBITS_C128;
    if (n < 1 || n > cap) return fail(MOONLAB_HIP_ERR_NUM_QUBITS, "num_qubits out of range");
    if (moonlab_hip_device_count() <= 0)
        return fail(MOONLAB_HIP_ERR_NO_DEVICE, "no HIP device visible (the MI355X engine needs a GPU)");
    moonlab_hip_state_t *s = new (std::nothrow) moonlab_hip_state_t();
    if (!s) return fail(MOONLAB_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    s->n = (int)n;
    s->prec = prec;
    s->dim = 1ull << n;
    s->bytes = (size_t)s->dim * amp_bytes(prec);
    hipGetDevice(&s->device);
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return drv(e, "hipStreamCreate");
    }
    e = hipMalloc(&s->amps, s->bytes);
    if (e != hipSuccess) {
        hipStreamDestroy(s->stream);
        delete s;
        return drv(e, "hipMalloc(state)");
    }
    e = launch_init_zero(dev(s));
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        hipFree(s->amps);
        hipStreamDestroy(s->stream);
        delete s;
        return drv(e, "init |0..0>");
    }
    *out = s;
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_state_create(uint32_t n, moonlab_hip_state_t **out)
{
    return moonlab_hip_state_create_ex(n, MOONLAB_HIP_C128, out);
}

void moonlab_hip_state_free(moonlab_hip_state_t *s)
{
    if (!s) return;
    hipStreamSynchronize(s->stream);
    for (hipEvent_t ev : s->events) hipEventDestroy(ev);
    hipFree(s->ws.partial);
    hipFree(s->ws.out);
    hipFree(s->ws.tile_sums);
    hipFree(s->ws.tile_incl);
    hipFree(s->d_u);
    hipFree(s->d_outcomes);
    if (s->h_out) hipHostFree(s->h_out);
    hipFree(s->amps);
    hipStreamDestroy(s->stream);
    delete s;
}

uint32_t moonlab_hip_state_n_qubits(const moonlab_hip_state_t *s) { return s ? (uint32_t)s->n : 0; }
uint64_t moonlab_hip_state_dim(const moonlab_hip_state_t *s) { return s ? s->dim : 0; }
int moonlab_hip_state_precision(const moonlab_hip_state_t *s) { return s ? s->prec : -1; }
void *moonlab_hip_state_device_ptr(moonlab_hip_state_t *s) { return s ? s->amps : nullptr; }
void *moonlab_hip_state_stream(moonlab_hip_state_t *s) { return s ? (void *)s->stream : nullptr; }

moonlab_hip_status_t moonlab_hip_state_reset(moonlab_hip_state_t *s)
{
    if (!s) return fail(MOONLAB_HIP_ERR_BAD_ARG, "state is NULL");
    return drv(launch_init_zero(dev(s)), "reset");
}

moonlab_hip_status_t moonlab_hip_apply_gate(moonlab_hip_state_t *s, const moonlab_hip_gate_t *g)
{
    if (!s || !g) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    moonlab_hip_status_t st = check_gate(s, *g);
    if (st) return st;
    return drv(launch_gate(dev(s), *g), "gate launch");
}

moonlab_hip_status_t moonlab_hip_apply_gates(moonlab_hip_state_t *s, const moonlab_hip_gate_t *g, int count)
{
    if (!s || (count > 0 && !g) || count < 0) return fail(MOONLAB_HIP_ERR_BAD_ARG, "bad gate list");
    for (int i = 0; i < count; i++) {
        moonlab_hip_status_t st = check_gate(s, g[i]);
        if (st) return st;
    }
    const SvDev d = dev(s);
    for (int i = 0; i < count; i++) {
        hipError_t e = launch_gate(d, g[i]);
        if (e != hipSuccess) return drv(e, "gate launch");
    }
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_apply_1q(moonlab_hip_state_t *s, uint32_t t, const double m[8])
{
    if (!s || !m) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (!qubit_ok(s, t)) return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "target out of range");
    return drv(launch_matrix_1q(dev(s), (int)t, m), "apply_1q");
}

moonlab_hip_status_t moonlab_hip_apply_cnot(moonlab_hip_state_t *s, uint32_t c, uint32_t t)
{
    if (!s) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (!qubit_ok(s, c) || !qubit_ok(s, t) || c == t)
        return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "cnot qubits");
    return drv(launch_mcx(dev(s), 1ull << c, (int)t), "apply_cnot");
}

moonlab_hip_status_t moonlab_hip_apply_2q(moonlab_hip_state_t *s, uint32_t q0, uint32_t q1, const double m[32])
{
    if (!s || !m) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (!qubit_ok(s, q0) || !qubit_ok(s, q1) || q0 == q1)
        return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "2q qubits");
    return drv(launch_matrix_2q(dev(s), (int)q0, (int)q1, m), "apply_2q");
}

moonlab_hip_status_t moonlab_hip_apply_x(moonlab_hip_state_t *s, uint32_t t)
{
    moonlab_hip_gate_t g = {MOONLAB_HIP_GATE_X, (int32_t)t, -1, -1, 0, 0, 0};
    return moonlab_hip_apply_gate(s, &g);
}

moonlab_hip_status_t moonlab_hip_apply_h(moonlab_hip_state_t *s, uint32_t t)
{
    moonlab_hip_gate_t g = {MOONLAB_HIP_GATE_H, (int32_t)t, -1, -1, 0, 0, 0};
    return moonlab_hip_apply_gate(s, &g);
}

moonlab_hip_status_t moonlab_hip_apply_mcx(moonlab_hip_state_t *s, uint64_t cmask, uint32_t t)
{
    if (!s) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (!qubit_ok(s, t) || (cmask >> s->n) || (cmask & (1ull << t)) || __builtin_popcountll(cmask) > 6)
        return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "mcx mask");
    if (cmask == 0) return drv(launch_gate(dev(s), moonlab_hip_gate_t{MOONLAB_HIP_GATE_X, (int32_t)t, -1, -1, 0, 0, 0}), "x");
    return drv(launch_mcx(dev(s), cmask, (int)t), "apply_mcx");
}

moonlab_hip_status_t moonlab_hip_apply_mcz(moonlab_hip_state_t *s, uint64_t mask)
{
    if (!s) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (mask == 0 || (mask >> s->n) || __builtin_popcountll(mask) > 7)
        return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "mcz mask");
    return drv(launch_mcz(dev(s), mask), "apply_mcz");
}

moonlab_hip_status_t moonlab_hip_apply_fredkin(moonlab_hip_state_t *s, uint32_t c, uint32_t t1, uint32_t t2)
{
    moonlab_hip_gate_t g = {MOONLAB_HIP_GATE_FREDKIN, (int32_t)t1, (int32_t)c, (int32_t)t2, 0, 0, 0};
    return moonlab_hip_apply_gate(s, &g);
}

moonlab_hip_status_t moonlab_hip_apply_diag_1q(moonlab_hip_state_t *s, uint32_t t, const double d[4])
{
    if (!s || !d) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (!qubit_ok(s, t)) return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "target out of range");
    return drv(launch_diag_1q(dev(s), (int)t, d), "apply_diag_1q");
}

moonlab_hip_status_t moonlab_hip_apply_controlled_1q(moonlab_hip_state_t *s, uint64_t cmask, uint32_t t,
                                                     const double m[8])
{
    if (!s || !m) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (!qubit_ok(s, t) || (cmask >> s->n) || (cmask & (1ull << t)) || __builtin_popcountll(cmask) > 6)
        return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "controlled_1q mask");
    if (cmask == 0) return drv(launch_matrix_1q(dev(s), (int)t, m), "apply_1q");
    return drv(launch_controlled_1q(dev(s), cmask, (int)t, m), "apply_controlled_1q");
}

moonlab_hip_status_t moonlab_hip_expect_z_all(const moonlab_hip_state_t *s, double *z, double *norm)
{
    if (!s) return fail(MOONLAB_HIP_ERR_BAD_ARG, "state is NULL");
    moonlab_hip_status_t st = reduce_all(s, false);
    if (st) return st;
    if (norm) *norm = s->h_out[0];
    if (z)
        for (int q = 0; q < s->n; q++) z[q] = 1.0 - 2.0 * s->h_out[1 + q];
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_prob_z(const moonlab_hip_state_t *s, uint32_t t, double *out)
{
    if (!s || !out) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (!qubit_ok(s, t)) return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "target out of range");
    moonlab_hip_status_t st = reduce_all(s, false);
    if (st) return st;
    *out = s->h_out[1 + t];
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_norm(const moonlab_hip_state_t *s, double *out)
{
    if (!s || !out) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    moonlab_hip_status_t st = reduce_all(s, false);
    if (st) return st;
    *out = s->h_out[0];
    return MOONLAB_HIP_OK;
}

static moonlab_hip_status_t xy(const moonlab_hip_state_t *cs, uint32_t q, int which, double *out)
{
    moonlab_hip_state_t *s = const_cast<moonlab_hip_state_t *>(cs);
    if (!s || !out) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (!qubit_ok(s, q)) return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "qubit out of range");
    moonlab_hip_status_t st = ensure_ws(s);
    if (st) return st;
    hipError_t e = launch_xy(dev(s), (int)q, s->ws.partial, 2048, s->ws.out);
    if (e != hipSuccess) return drv(e, "expect_xy");
    st = fetch_small(s, 2);
    if (st) return st;
    const double v = s->h_out[which];
    *out = v + v;
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_expect_x(const moonlab_hip_state_t *s, uint32_t q, double *out)
{
    return xy(s, q, 0, out);
}

moonlab_hip_status_t moonlab_hip_expect_y(const moonlab_hip_state_t *s, uint32_t q, double *out)
{
    return xy(s, q, 1, out);
}

moonlab_hip_status_t moonlab_hip_correlation_zz(const moonlab_hip_state_t *cs, uint32_t q1, uint32_t q2,
                                                double *out)
{
    moonlab_hip_state_t *s = const_cast<moonlab_hip_state_t *>(cs);
    if (!s || !out) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (!qubit_ok(s, q1) || !qubit_ok(s, q2)) return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "qubit out of range");
    moonlab_hip_status_t st = ensure_ws(s);
    if (st) return st;
    const uint64_t mask = (1ull << q1) ^ (1ull << q2);
    hipError_t e = launch_zmask(dev(s), mask, s->ws.partial, 2048, s->ws.out);
    if (e != hipSuccess) return drv(e, "correlation_zz");
    st = fetch_small(s, 1);
    if (st) return st;
    *out = s->h_out[0];
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_expect_pauli(const moonlab_hip_state_t *cs, uint64_t x_mask, uint64_t z_mask,
                                              double *out)
{
    moonlab_hip_state_t *s = const_cast<moonlab_hip_state_t *>(cs);
    if (!s || !out) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    const uint64_t all = s->n >= 64 ? ~0ull : (1ull << s->n) - 1;
    if ((x_mask | z_mask) & ~all) return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "Pauli mask names a qubit out of range");
    moonlab_hip_status_t st = ensure_ws(s);
    if (st) return st;
    hipError_t e = x_mask ? launch_pauli(dev(s), x_mask, z_mask, s->ws.partial, 2048, s->ws.out)
                          : launch_zmask(dev(s), z_mask, s->ws.partial, 2048, s->ws.out);
    if (e != hipSuccess) return drv(e, "expect_pauli");
    st = fetch_small(s, 1);
    if (st) return st;
    *out = s->h_out[0];
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_inner_product(const moonlab_hip_state_t *a, const moonlab_hip_state_t *b,
                                               double *re, double *im)
{
    if (!a || !b) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (a->n != b->n) return fail(MOONLAB_HIP_ERR_BAD_ARG, "shape mismatch");
    moonlab_hip_state_t *s = const_cast<moonlab_hip_state_t *>(a);
    moonlab_hip_status_t st = ensure_ws(s);
    if (st) return st;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess) e = launch_inner(dev(a), dev(b), s->ws.partial, 2048, s->ws.out);
    if (e != hipSuccess) return drv(e, "inner_product");
    st = fetch_small(s, 2);
    if (st) return st;
    if (re) *re = s->h_out[0];
    if (im) *im = s->h_out[1];
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_normalize(moonlab_hip_state_t *s)
{
    double nrm = 0.0;
    moonlab_hip_status_t st = moonlab_hip_norm(s, &nrm);
    if (st) return st;
    if (!(nrm > 0.0)) return fail(MOONLAB_HIP_ERR_BAD_ARG, "zero norm");
    return drv(launch_scale(dev(s), 1.0 / sqrt(nrm)), "normalize");
}

moonlab_hip_status_t moonlab_hip_probabilities(const moonlab_hip_state_t *s, double *out)
{
    if (!s || !out) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    double *d_p = nullptr;
    hipError_t e = hipMallocAsync((void **)&d_p, s->dim * sizeof(double), s->stream);
    if (e == hipSuccess) e = launch_probabilities(dev(s), d_p);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_p, s->dim * sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (d_p) hipFreeAsync(d_p, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return drv(e, "probabilities");
}

moonlab_hip_status_t moonlab_hip_sample(const moonlab_hip_state_t *cs, const double *u, int num, uint64_t *out)
{
    moonlab_hip_state_t *s = const_cast<moonlab_hip_state_t *>(cs);
    if (!s || (num > 0 && (!u || !out)) || num < 0) return fail(MOONLAB_HIP_ERR_BAD_ARG, "bad sample arguments");
    if (num == 0) return MOONLAB_HIP_OK;
    moonlab_hip_status_t st = ensure_ws(s);
    if (st) return st;
    if (num > s->shot_cap) {
        hipFree(s->d_u);
        hipFree(s->d_outcomes);
        s->d_u = nullptr;
        s->d_outcomes = nullptr;
        s->shot_cap = 0;
        hipError_t e = hipMalloc((void **)&s->d_u, (size_t)num * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&s->d_outcomes, (size_t)num * sizeof(uint64_t));
        if (e != hipSuccess) return drv(e, "sample buffers");
        s->shot_cap = num;
    }
    const SvDev d = dev(s);
    hipError_t e = launch_reduce_z(d, s->ws, true);
    if (e == hipSuccess) e = launch_tile_scan(s->ws, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_u, u, (size_t)num * sizeof(double), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = launch_sample(d, s->ws, s->d_u, num, s->d_outcomes);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, s->d_outcomes, (size_t)num * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return drv(e, "sample");
}

// Collapsing measurements (measurement.c:181-231, 263-307, 320-410): the
// outcome sums come from one device pass, the outcome rule runs on the host
// over 2^k doubles as the reference's does, and one more pass collapses.
// Below MEASUREMENT_SMALL_NORM (1e-15, measurement.c:33) the kept amplitudes
// are not rescaled.
moonlab_hip_status_t moonlab_hip_measure_qubit(moonlab_hip_state_t *s, uint32_t qubit, double u, int *outcome)
{
    if (!s || !outcome) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (qubit >= (uint32_t)s->n) return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "qubit out of range");
    const SvDev d = dev(s);
    const int q = (int)qubit;
    double p[2] = {0.0, 0.0};
    hipError_t e = launch_outcome_probs(d, &q, 1, p);
    if (e != hipSuccess) return drv(e, "measure_qubit");
    const int r = u < p[1] ? 1 : 0;   // P(1) unnormalized, as measurement_probability_one
    const uint64_t mask = 1ull << q, want = (uint64_t)r << q;
    if (p[r] <= 1e-15) {
        e = launch_collapse(d, mask, want, 1.0, 0);
        *outcome = -1;
    } else {
        e = launch_collapse(d, mask, want, 1.0 / sqrt(p[r]), 1);
        *outcome = r;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return drv(e, "measure_qubit");
}

moonlab_hip_status_t moonlab_hip_measure_partial(moonlab_hip_state_t *s, const int *qubits, int k, double u,
                                                 uint64_t *outcome)
{
    if (!s || !qubits || !outcome) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (k < 1 || k > 20 || k > s->n) return fail(MOONLAB_HIP_ERR_BAD_ARG, "measure 1..20 qubits");
    uint64_t mask = 0;
    for (int j = 0; j < k; j++) {
        if (qubits[j] < 0 || qubits[j] >= s->n) return fail(MOONLAB_HIP_ERR_INVALID_QUBIT, "qubit out of range");
        if (mask & (1ull << qubits[j])) return fail(MOONLAB_HIP_ERR_BAD_ARG, "repeated qubit");
        mask |= 1ull << qubits[j];
    }
    const SvDev d = dev(s);
    std::vector<double> p((size_t)1 << k);
    hipError_t e = launch_outcome_probs(d, qubits, k, p.data());
    if (e != hipSuccess) return drv(e, "measure_partial");
    // the cumulative rule with the last-nonzero fallback (measurement.c:357-376)
    double cum = 0.0;
    uint64_t r = 0, last_nz = 0;
    bool found = false, have_nz = false;
    for (uint64_t o = 0; o < p.size(); o++) {
        if (p[o] > 0.0) {
            last_nz = o;
            have_nz = true;
        }
        cum += p[o];
        if (u < cum) {
            r = o;
            found = true;
            break;
        }
    }
    if (!found) r = have_nz ? last_nz : 0;
    uint64_t want = 0;
    for (int j = 0; j < k; j++)
        if ((r >> j) & 1ull) want |= 1ull << qubits[j];
    const double nf = p[r];
    e = nf > 1e-15 ? launch_collapse(d, mask, want, 1.0 / sqrt(nf), 1) : launch_collapse(d, mask, want, 1.0, 0);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    *outcome = r;
    return drv(e, "measure_partial");
}

moonlab_hip_status_t moonlab_hip_measure_all(moonlab_hip_state_t *s, double u, uint64_t *outcome)
{
    if (!s || !outcome) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    uint64_t r = 0;
    const moonlab_hip_status_t st = moonlab_hip_sample(s, &u, 1, &r);
    if (st) return st;
    hipError_t e = launch_collapse(dev(s), s->dim - 1, r, 1.0, 2);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    *outcome = r;
    return drv(e, "measure_all");
}

moonlab_hip_status_t moonlab_hip_state_read(const moonlab_hip_state_t *s, uint64_t off, uint64_t count, void *out)
{
    if (!s || (count && !out)) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (off > s->dim || count > s->dim - off) return fail(MOONLAB_HIP_ERR_BAD_ARG, "range");
    const size_t b = amp_bytes(s->prec);
    hipError_t e = hipMemcpyAsync(out, (const char *)s->amps + off * b, count * b, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return drv(e, "state_read");
}

moonlab_hip_status_t moonlab_hip_state_copy_to_host(const moonlab_hip_state_t *s, double *out)
{
    if (!s || !out) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (s->prec == MOONLAB_HIP_C128) return moonlab_hip_state_read(s, 0, s->dim, out);
    // complex64: copy in chunks and widen on the host
    const uint64_t chunk = 1ull << 24;
    std::vector<float> buf;
    try {
        buf.resize((size_t)2 * (s->dim < chunk ? s->dim : chunk));
    } catch (...) {
        return fail(MOONLAB_HIP_ERR_OUT_OF_MEMORY, "host staging");
    }
    for (uint64_t off = 0; off < s->dim; off += chunk) {
        const uint64_t cnt = s->dim - off < chunk ? s->dim - off : chunk;
        moonlab_hip_status_t st = moonlab_hip_state_read(s, off, cnt, buf.data());
        if (st) return st;
        for (uint64_t i = 0; i < 2 * cnt; i++) out[2 * off + i] = (double)buf[i];
    }
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_state_copy_from_host(moonlab_hip_state_t *s, const double *in)
{
    if (!s || !in) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (s->prec == MOONLAB_HIP_C128) {
        hipError_t e = hipMemcpyAsync(s->amps, in, s->bytes, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        return drv(e, "copy_from_host");
    }
    const uint64_t chunk = 1ull << 24;
    std::vector<float> buf;
    try {
        buf.resize((size_t)2 * (s->dim < chunk ? s->dim : chunk));
    } catch (...) {
        return fail(MOONLAB_HIP_ERR_OUT_OF_MEMORY, "host staging");
    }
    for (uint64_t off = 0; off < s->dim; off += chunk) {
        const uint64_t cnt = s->dim - off < chunk ? s->dim - off : chunk;
        for (uint64_t i = 0; i < 2 * cnt; i++) buf[i] = (float)in[2 * off + i];
        hipError_t e = hipMemcpyAsync((char *)s->amps + off * 8, buf.data(), cnt * 8, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return drv(e, "copy_from_host");
    }
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_state_copy(moonlab_hip_state_t *dst, const moonlab_hip_state_t *src)
{
    if (!dst || !src) return fail(MOONLAB_HIP_ERR_BAD_ARG, "NULL argument");
    if (dst->n != src->n) return fail(MOONLAB_HIP_ERR_BAD_ARG, "qubit counts differ");
    if (dst == src) return MOONLAB_HIP_OK;
    hipError_t e = hipStreamSynchronize(src->stream);
    if (e == hipSuccess) e = launch_convert(src->amps, src->prec, dst->amps, dst->prec, dst->dim, dst->stream);
    return drv(e, "state_copy");
}

moonlab_hip_status_t moonlab_hip_synchronize(moonlab_hip_state_t *s)
{
    if (!s) return fail(MOONLAB_HIP_ERR_BAD_ARG, "state is NULL");
    return drv(hipStreamSynchronize(s->stream), "synchronize");
}

moonlab_hip_status_t moonlab_hip_timer_begin(moonlab_hip_state_t *s, int capacity)
{
    if (!s || capacity < 2) return fail(MOONLAB_HIP_ERR_BAD_ARG, "timer capacity");
    while ((int)s->events.size() < capacity) {
        hipEvent_t ev;
        hipError_t e = hipEventCreate(&ev);
        if (e != hipSuccess) return drv(e, "hipEventCreate");
        s->events.push_back(ev);
    }
    s->next_event = 0;
    return MOONLAB_HIP_OK;
}

moonlab_hip_status_t moonlab_hip_timer_mark(moonlab_hip_state_t *s)
{
    if (!s || s->next_event >= (int)s->events.size()) return fail(MOONLAB_HIP_ERR_BAD_ARG, "timer full");
    return drv(hipEventRecord(s->events[s->next_event++], s->stream), "hipEventRecord");
}

int moonlab_hip_timer_elapsed(moonlab_hip_state_t *s, float *gaps, int max_gaps)
{
    if (!s || !gaps) return -1;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return -1;
    int n = s->next_event - 1;
    if (n > max_gaps) n = max_gaps;
    for (int i = 0; i < n; i++)
        if (hipEventElapsedTime(&gaps[i], s->events[i], s->events[i + 1]) != hipSuccess) return -1;
    return n < 0 ? 0 : n;
}

}  // extern "C"
