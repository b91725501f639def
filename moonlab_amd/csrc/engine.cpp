// engine.cpp -- circuit execution on the MI355X state-vector engine, the
// "mi355x" scheduler backend and the extended execute entry point.
//
// A job runs ONCE on the device and every worker's shots are drawn from that
// one state: the reference re-simulates the whole circuit per worker and
// clones the state per shot (scheduler.c:152-189, moonlab_qgtl_backend.c:
// 218-243); both are deterministic, so one simulation with the same per-worker
// uniform streams gives the same outcomes.
#include "engine.h"
#include "sv_internal.h"

#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "../../include/moonlab_mi355x.h"
#include "fusion.h"

namespace mlx {

bool gate_takes_param(int t)
{
    return t == MOONLAB_QGTL_GATE_RX || t == MOONLAB_QGTL_GATE_RY || t == MOONLAB_QGTL_GATE_RZ;
}

bool gate_is_two_qubit(int t)
{
    return t == MOONLAB_QGTL_GATE_CNOT || t == MOONLAB_QGTL_GATE_CY || t == MOONLAB_QGTL_GATE_CZ ||
           t == MOONLAB_QGTL_GATE_SWAP;
}

bool gate_rec_valid(const GateRec &g, int n)
{
    if (g.type < MOONLAB_QGTL_GATE_I || g.type > MOONLAB_QGTL_GATE_SWAP) return false;
    if (g.target < 0 || g.target >= n) return false;
    if (gate_is_two_qubit(g.type)) {
        if (g.control < 0 || g.control >= n || g.control == g.target) return false;
    }
    return true;
}

uint64_t splitmix64(uint64_t x)
{
    // Constants as scheduler.c:145-146 has them (the second multiplier is
    // not the textbook SplitMix64 one; outcome parity needs this one).
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d49bb133111ebbULL;
    return x ^ (x >> 31);
}

double xorshift_u(uint64_t *s)
{
    uint64_t x = *s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *s = x;
    return (double)(x >> 11) / (double)(1ULL << 53);
}

int fill_job_draws(uint64_t seed, int total, int workers, std::vector<double> &u)
{
    if (workers > total) workers = total;
    if (workers < 1) workers = 1;
    u.resize((size_t)total);
    for (int w = 0; w < workers; w++) {
        const int start = (int)((long long)w * total / workers);
        const int end = (int)((long long)(w + 1) * total / workers);
        if (end <= start) continue;
        uint64_t s = seed ? splitmix64(seed ^ ((uint64_t)w << 32))
                          : (uint64_t)(time(NULL) ^ ((uint64_t)w << 32) ^ 0x9e3779b97f4a7c15ULL);
        for (int k = start; k < end; k++) u[(size_t)k] = xorshift_u(&s);
    }
    return workers;
}

// ---------------------------------------------------------------------------
// Device-state pool: finished jobs park their state for the next job of the
// same (qubits, precision), so a steady stream of circuits pays no hipMalloc.
// ---------------------------------------------------------------------------
namespace {

struct Pool {
    std::mutex mu;
    std::multimap<std::pair<int, int>, moonlab_hip_state_t *> free_states;
    uint64_t cached_bytes = 0;
    uint64_t cap_bytes = 64ull << 30;
};

Pool &pool()
{
    static Pool *p = new Pool();   // never destroyed: states may outlive static teardown
    return *p;
}

uint64_t state_bytes(int n, int prec) { return (1ull << n) * (prec == MOONLAB_HIP_C64 ? 8ull : 16ull); }

int acquire_state(int n, int prec, moonlab_hip_state_t **out)
{
    {
        Pool &p = pool();
        std::lock_guard<std::mutex> lk(p.mu);
        auto it = p.free_states.find({n, prec});
        if (it != p.free_states.end()) {
            *out = it->second;
            p.free_states.erase(it);
            p.cached_bytes -= state_bytes(n, prec);
            return moonlab_hip_state_reset(*out);
        }
    }
    return moonlab_hip_state_create_ex((uint32_t)n, prec, out);
}

void release_state(moonlab_hip_state_t *s)
{
    if (!s) return;
    const int n = (int)moonlab_hip_state_n_qubits(s), prec = moonlab_hip_state_precision(s);
    const uint64_t b = state_bytes(n, prec);
    Pool &p = pool();
    {
        std::lock_guard<std::mutex> lk(p.mu);
        if (p.cached_bytes + b <= p.cap_bytes) {
            p.free_states.insert({{n, prec}, s});
            p.cached_bytes += b;
            return;
        }
    }
    moonlab_hip_state_free(s);
}

}  // namespace

int run_circuit(const RunRequest &rq, RunResult &rr)
{
    std::vector<moonlab_hip_gate_t> g((size_t)rq.num_gates);
    for (int i = 0; i < rq.num_gates; i++) {
        const GateRec &r = rq.gates[i];
        g[(size_t)i] = moonlab_hip_gate_t{r.type, r.target, r.control, r.extra, r.theta, r.phi, r.lam};
    }
    return run_circuit_gates(rq, g.data(), rq.num_gates, rr);
}

int run_circuit_gates(const RunRequest &rq, const moonlab_hip_gate_t *gl, int ng, RunResult &rr)
{
    moonlab_hip_state_t *s = nullptr;
    int st = acquire_state(rq.num_qubits, rq.precision, &s);
    if (st) return st;
    std::vector<moonlab_hip_gate_t> g(gl, gl + ng);
    // a caller-chosen stream (a serving worker's) replaces the pooled state's own for this run
    hipStream_t own = rq.stream ? state_bind_stream(s, (hipStream_t)rq.stream) : nullptr;
    hipStream_t stream = (hipStream_t)moonlab_hip_state_stream(s);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, stream);
    if (rq.blocked)
        st = moonlab_hip_apply_gates_blocked(s, g.data(), ng);
    else if (rq.fusion_max > 0)
        st = apply_fused(s, g.data(), ng, rq.fusion_max);
    else
        st = moonlab_hip_apply_gates(s, g.data(), ng);
    if (!st && rq.num_draws > 0) {
        rr.outcomes.resize((size_t)rq.num_draws);
        st = moonlab_hip_sample(s, rq.draws, rq.num_draws, rr.outcomes.data());
    }
    if (!st && rq.want_z) {
        rr.z.resize((size_t)rq.num_qubits);
        st = moonlab_hip_expect_z_all(s, rr.z.data(), &rr.norm);
    }
    hipEventRecord(e1, stream);
    if (!st && rq.want_probs) {
        rr.probs.resize((size_t)1 << rq.num_qubits);
        st = moonlab_hip_probabilities(s, rr.probs.data());
    }
    if (!st && rq.want_amps) {
        rr.amps.resize((size_t)2 << rq.num_qubits);
        st = moonlab_hip_state_copy_to_host(s, rr.amps.data());
    }
    if (!st) {
        hipStreamSynchronize(stream);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) rr.device_seconds = ms * 1e-3;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (own) state_bind_stream(s, own);   // waits for the caller's stream, then hands the state back
    if (st) {
        moonlab_hip_synchronize(s);
        moonlab_hip_state_free(s);   // do not pool a state left in an unknown condition
        return st;
    }
    release_state(s);
    return 0;
}

}  // namespace mlx

using namespace mlx;

namespace {

// Read the job through the public API only: counts from the getters, gates
// and seed from moonlab_job_to_json's "moonlab/job/v0.7.0" dump
// (scheduler.c:492-537).  The in-tree precedent for a backend re-parsing
// that dump is vendor_noise_backend.c:246-300.
int read_job(const moonlab_job_t *job, int *n, uint64_t *seed, std::vector<GateRec> &gates)
{
    *n = moonlab_job_num_qubits(job);
    const int probe = moonlab_job_to_json(job, nullptr, 0);
    if (*n < 1 || probe <= 0) return MOONLAB_SCHED_BAD_ARG;
    std::string js((size_t)probe + 1, '\0');
    if (moonlab_job_to_json(job, &js[0], js.size()) != probe) return MOONLAB_SCHED_INTERNAL;
    const char *p = strstr(js.c_str(), "\"rng_seed\": \"");
    *seed = p ? strtoull(p + 13, nullptr, 16) : 0;
    const char *cur = strstr(js.c_str(), "\"gates\": [");
    if (!cur) return MOONLAB_SCHED_INTERNAL;
    gates.clear();
    while ((cur = strstr(cur, "{ \"type\": ")) != nullptr) {
        GateRec g{0, -1, -1, 0.0, 0};
        char *endp = nullptr;
        g.type = (int)strtol(cur + 10, &endp, 10);
        const char *close = strchr(cur, '}');
        if (!close) return MOONLAB_SCHED_INTERNAL;
        const char *t = strstr(cur, "\"target\": ");
        if (t && t < close) g.target = (int)strtol(t + 10, nullptr, 10);
        const char *c = strstr(cur, "\"control\": ");
        if (c && c < close) g.control = (int)strtol(c + 11, nullptr, 10);
        const char *th = strstr(cur, "\"theta\": ");
        if (th && th < close) {
            g.theta = strtod(th + 9, nullptr);
            g.has_param = 1;
        }
        gates.push_back(g);
        cur = close + 1;
    }
    if ((int)gates.size() != moonlab_job_num_gates(job)) return MOONLAB_SCHED_INTERNAL;
    return MOONLAB_SCHED_OK;
}

int map_status(int st)
{
    if (st == MOONLAB_HIP_ERR_OUT_OF_MEMORY) return MOONLAB_SCHED_OOM;
    if (st == MOONLAB_HIP_ERR_BAD_ARG || st == MOONLAB_HIP_ERR_INVALID_QUBIT ||
        st == MOONLAB_HIP_ERR_NUM_QUBITS)
        return MOONLAB_SCHED_INTERNAL;
    return MOONLAB_SCHED_INTERNAL;
}

int execute_common(const moonlab_job_t *job, const moonlab_mi355x_exec_options_t *opts, RunResult &rr,
                   int *shots_out, int *workers_out)
{
    int n = 0;
    uint64_t seed = 0;
    std::vector<GateRec> gates;
    int rc = read_job(job, &n, &seed, gates);
    if (rc) return rc;
    // Bad indices surface at run time as INTERNAL, as the reference's
    // worker rebuild reports them (scheduler.c:159-166).
    for (const GateRec &g : gates)
        if (!gate_rec_valid(g, n)) return MOONLAB_SCHED_INTERNAL;
    const int shots = opts && opts->skip_shots ? 0 : moonlab_job_num_shots(job);
    int workers = moonlab_job_num_workers(job);
    std::vector<double> u;
    if (shots > 0) workers = fill_job_draws(seed, shots, workers, u);
    else workers = workers < 1 ? 1 : workers;
    RunRequest rq{};
    rq.num_qubits = n;
    rq.precision = opts ? opts->precision : MOONLAB_HIP_C128;
    rq.gates = gates.data();
    rq.num_gates = (int)gates.size();
    rq.draws = u.empty() ? nullptr : u.data();
    rq.num_draws = shots;
    rq.want_amps = opts && opts->return_amplitudes;
    rq.want_probs = opts && opts->return_probabilities;
    rq.want_z = opts && opts->return_expect_z;
    rq.fusion_max = opts ? opts->fusion_max_qubits : 0;
    rq.blocked = opts && opts->blocked;
    const int st = run_circuit(rq, rr);
    if (st) {
        fprintf(stderr, "moonlab-mi355x: %s\n", moonlab_hip_last_error());
        return map_status(st);
    }
    *shots_out = shots;
    *workers_out = workers;
    return MOONLAB_SCHED_OK;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int moonlab_mi355x_backend_execute(const moonlab_job_t *job, moonlab_job_results_t *out, void *ctx)
{
    (void)ctx;
    if (!job || !out) return MOONLAB_SCHED_BAD_ARG;
    if (moonlab_job_num_shots(job) <= 0) return MOONLAB_SCHED_BAD_ARG;
    const double t0 = now_s();
    RunResult rr;
    int shots = 0, workers = 1;
    const int rc = execute_common(job, nullptr, rr, &shots, &workers);
    if (rc) return rc;
    const double wall = now_s() - t0;
    out->num_qubits = moonlab_job_num_qubits(job);
    out->total_shots = shots;
    out->num_workers_used = workers;
    out->outcomes = (uint64_t *)malloc((size_t)shots * sizeof(uint64_t));
    out->worker_seconds = (double *)calloc((size_t)workers, sizeof(double));
    if (!out->outcomes || !out->worker_seconds) {
        free(out->outcomes);
        free(out->worker_seconds);
        out->outcomes = nullptr;
        out->worker_seconds = nullptr;
        return MOONLAB_SCHED_OOM;
    }
    memcpy(out->outcomes, rr.outcomes.data(), (size_t)shots * sizeof(uint64_t));
    // One device run serves every worker slice; each slot carries the
    // slice's share of the wall time.
    for (int w = 0; w < workers; w++) {
        const int start = (int)((long long)w * shots / workers);
        const int end = (int)((long long)(w + 1) * shots / workers);
        out->worker_seconds[w] = wall * (double)(end - start) / (double)shots;
    }
    return MOONLAB_SCHED_OK;
}

int moonlab_mi355x_register_backend(void)
{
    static const moonlab_backend_t be = {
        MOONLAB_MI355X_BACKEND_NAME, moonlab_mi355x_backend_execute, nullptr,
        "AMD Instinct MI355X dense state-vector engine (hand-written gfx950 HIP kernels)."};
    return moonlab_register_backend(&be);
}

int moonlab_mi355x_execute_ex(const moonlab_job_t *job, const moonlab_mi355x_exec_options_t *opts,
                              moonlab_mi355x_results_t *out)
{
    if (!job || !opts || !out) return MOONLAB_SCHED_BAD_ARG;
    if (opts->precision != MOONLAB_HIP_C128 && opts->precision != MOONLAB_HIP_C64) return MOONLAB_SCHED_BAD_ARG;
    if (opts->fusion_max_qubits < 0 || opts->fusion_max_qubits > 5) return MOONLAB_SCHED_BAD_ARG;
    memset(out, 0, sizeof(*out));
    const double t0 = now_s();
    RunResult rr;
    int shots = 0, workers = 1;
    const int rc = execute_common(job, opts, rr, &shots, &workers);
    if (rc) return rc;
    out->num_qubits = moonlab_job_num_qubits(job);
    out->precision = opts->precision;
    out->total_shots = shots;
    out->num_workers_used = workers;
    out->norm = rr.norm;
    out->device_seconds = rr.device_seconds;
    auto dup = [](const auto &v, auto **dst) -> bool {
        using T = typename std::decay_t<decltype(v)>::value_type;
        if (v.empty()) return true;
        *dst = (T *)malloc(v.size() * sizeof(T));
        if (!*dst) return false;
        memcpy(*dst, v.data(), v.size() * sizeof(T));
        return true;
    };
    if (!dup(rr.outcomes, &out->outcomes) || !dup(rr.amps, &out->amplitudes) ||
        !dup(rr.probs, &out->probabilities) || !dup(rr.z, &out->expect_z)) {
        moonlab_mi355x_results_free(out);
        return MOONLAB_SCHED_OOM;
    }
    out->wall_seconds = now_s() - t0;
    return MOONLAB_SCHED_OK;
}

void moonlab_mi355x_results_free(moonlab_mi355x_results_t *r)
{
    if (!r) return;
    free(r->outcomes);
    free(r->amplitudes);
    free(r->probabilities);
    free(r->expect_z);
    r->outcomes = nullptr;
    r->amplitudes = r->probabilities = r->expect_z = nullptr;
}

int moonlab_mi355x_set_state_cache_bytes(uint64_t bytes)
{
    Pool &p = pool();
    std::vector<moonlab_hip_state_t *> drop;
    {
        std::lock_guard<std::mutex> lk(p.mu);
        p.cap_bytes = bytes;
        while (p.cached_bytes > p.cap_bytes && !p.free_states.empty()) {
            auto it = p.free_states.begin();
            p.cached_bytes -= state_bytes((int)moonlab_hip_state_n_qubits(it->second),
                                          moonlab_hip_state_precision(it->second));
            drop.push_back(it->second);
            p.free_states.erase(it);
        }
    }
    for (auto *s : drop) moonlab_hip_state_free(s);
    return MOONLAB_SCHED_OK;
}

}  // extern "C"
