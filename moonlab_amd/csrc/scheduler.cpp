// scheduler.cpp -- job API, backend registry, completion hook and request
// context with the reference's semantics (src/distributed/scheduler.c), so a
// caller of moonlab_scheduler_run links against this library unchanged.
//
// Differences from the reference, all deliberate:
//   * the default "simulator" entry runs the MI355X engine (there is no CPU
//     execution path in this library); it reads jobs exactly as the "mi355x"
//     entry does;
//   * moonlab_scheduler_run_mpi returns MOONLAB_SCHED_NOT_BUILT, as the
//     reference does without MPI (scheduler.h:183-211).
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <new>
#include <vector>

#include "../../include/moonlab_mi355x.h"
#include "engine.h"

using mlx::GateRec;

struct moonlab_job {
    int                  num_qubits;
    int                  num_shots;
    int                  num_workers;
    uint64_t             rng_seed;
    std::vector<GateRec> gates;
    char                *backend_name;   // heap copy, nullptr = "simulator"
};

namespace {

constexpr int kMaxBackends = 16;   // registry size (scheduler.c:252)

moonlab_backend_t g_backends[kMaxBackends];
int g_num_backends = 0;
pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
pthread_once_t g_once = PTHREAD_ONCE_INIT;
moonlab_completion_hook_fn g_hook = nullptr;
void *g_hook_ctx = nullptr;
thread_local const char *t_tenant = nullptr;
thread_local const char *t_request = nullptr;

void register_default(void)
{
    g_backends[0].name = "simulator";
    g_backends[0].execute = moonlab_mi355x_backend_execute;
    g_backends[0].ctx = nullptr;
    g_backends[0].description =
        "Default backend: the MI355X dense state-vector engine (HIP kernels on gfx950).";
    g_num_backends = 1;
}

void ensure_default(void) { pthread_once(&g_once, register_default); }

}  // namespace

// Process-wide count of completion-hook fires (scheduler.c:30).
std::atomic<uint64_t> g_mi355x_hook_fires{0};

extern "C" {

moonlab_job_t *moonlab_job_create(int n)
{
    if (n < 1 || n > 32) return nullptr;
    moonlab_job_t *j = new (std::nothrow) moonlab_job_t();
    if (!j) return nullptr;
    j->num_qubits = n;
    j->num_workers = 1;
    return j;
}

void moonlab_job_free(moonlab_job_t *j)
{
    if (!j) return;
    free(j->backend_name);
    delete j;
}

int moonlab_job_num_qubits(const moonlab_job_t *j) { return j ? j->num_qubits : MOONLAB_SCHED_BAD_ARG; }
int moonlab_job_num_gates(const moonlab_job_t *j) { return j ? (int)j->gates.size() : MOONLAB_SCHED_BAD_ARG; }
int moonlab_job_num_shots(const moonlab_job_t *j) { return j ? j->num_shots : MOONLAB_SCHED_BAD_ARG; }
int moonlab_job_num_workers(const moonlab_job_t *j) { return j ? j->num_workers : MOONLAB_SCHED_BAD_ARG; }

int moonlab_job_add_gate(moonlab_job_t *j, moonlab_qgtl_gate_t type, int target, int control,
                         const double *params)
{
    // Only NULL job and missing theta are rejected at add time; qubit
    // indices are checked when the job runs (scheduler.c:99-121, 159-166).
    if (!j) return MOONLAB_SCHED_BAD_ARG;
    const bool p = mlx::gate_takes_param((int)type);
    if (p && !params) return MOONLAB_SCHED_BAD_ARG;
    try {
        j->gates.push_back(GateRec{(int)type, target, control, p ? params[0] : 0.0, p ? 1 : 0});
    } catch (...) {
        return MOONLAB_SCHED_OOM;
    }
    return MOONLAB_SCHED_OK;
}

int moonlab_job_set_num_shots(moonlab_job_t *j, int n)
{
    if (!j || n < 0) return MOONLAB_SCHED_BAD_ARG;
    j->num_shots = n;
    return MOONLAB_SCHED_OK;
}

int moonlab_job_set_num_workers(moonlab_job_t *j, int n)
{
    if (!j || n < 1) return MOONLAB_SCHED_BAD_ARG;
    j->num_workers = n;
    return MOONLAB_SCHED_OK;
}

int moonlab_job_set_rng_seed(moonlab_job_t *j, uint64_t seed)
{
    if (!j) return MOONLAB_SCHED_BAD_ARG;
    j->rng_seed = seed;
    return MOONLAB_SCHED_OK;
}

int moonlab_register_backend(const moonlab_backend_t *b)
{
    if (!b || !b->name || !b->execute) return MOONLAB_SCHED_BAD_ARG;
    ensure_default();
    pthread_mutex_lock(&g_lock);
    for (int i = 0; i < g_num_backends; i++) {
        if (strcmp(g_backends[i].name, b->name) == 0) {   // same name replaces
            g_backends[i] = *b;
            pthread_mutex_unlock(&g_lock);
            return MOONLAB_SCHED_OK;
        }
    }
    if (g_num_backends >= kMaxBackends) {
        pthread_mutex_unlock(&g_lock);
        return MOONLAB_SCHED_OOM;
    }
    g_backends[g_num_backends++] = *b;
    pthread_mutex_unlock(&g_lock);
    return MOONLAB_SCHED_OK;
}

int moonlab_unregister_backend(const char *name)
{
    if (!name) return MOONLAB_SCHED_BAD_ARG;
    ensure_default();
    pthread_mutex_lock(&g_lock);
    for (int i = 0; i < g_num_backends; i++) {
        if (strcmp(g_backends[i].name, name) == 0) {
            for (int k = i; k + 1 < g_num_backends; k++) g_backends[k] = g_backends[k + 1];
            g_num_backends--;
            pthread_mutex_unlock(&g_lock);
            return MOONLAB_SCHED_OK;
        }
    }
    pthread_mutex_unlock(&g_lock);
    return MOONLAB_SCHED_BACKEND_NOT_FOUND;
}

const moonlab_backend_t *moonlab_find_backend(const char *name)
{
    if (!name) return nullptr;
    ensure_default();
    pthread_mutex_lock(&g_lock);
    const moonlab_backend_t *hit = nullptr;
    for (int i = 0; i < g_num_backends; i++)
        if (strcmp(g_backends[i].name, name) == 0) {
            hit = &g_backends[i];
            break;
        }
    pthread_mutex_unlock(&g_lock);
    return hit;
}

int moonlab_num_backends(void)
{
    ensure_default();
    pthread_mutex_lock(&g_lock);
    const int n = g_num_backends;
    pthread_mutex_unlock(&g_lock);
    return n;
}

int moonlab_list_backends(const char **out, int max)
{
    if (!out || max <= 0) return 0;
    ensure_default();
    pthread_mutex_lock(&g_lock);
    const int n = g_num_backends < max ? g_num_backends : max;
    for (int i = 0; i < n; i++) out[i] = g_backends[i].name;
    pthread_mutex_unlock(&g_lock);
    return n;
}

int moonlab_job_set_backend(moonlab_job_t *j, const char *name)
{
    if (!j) return MOONLAB_SCHED_BAD_ARG;
    free(j->backend_name);
    j->backend_name = nullptr;
    if (!name) return MOONLAB_SCHED_OK;
    const size_t len = strlen(name);
    j->backend_name = (char *)malloc(len + 1);
    if (!j->backend_name) return MOONLAB_SCHED_OOM;
    memcpy(j->backend_name, name, len + 1);
    return MOONLAB_SCHED_OK;
}

const char *moonlab_job_backend(const moonlab_job_t *j) { return j ? j->backend_name : nullptr; }

int moonlab_scheduler_set_completion_hook(moonlab_completion_hook_fn hook, void *ctx)
{
    ensure_default();
    pthread_mutex_lock(&g_lock);
    g_hook = hook;
    g_hook_ctx = ctx;
    pthread_mutex_unlock(&g_lock);
    return MOONLAB_SCHED_OK;
}

void moonlab_scheduler_set_request_context(const char *tenant, const char *request)
{
    t_tenant = tenant;
    t_request = request;
}

const char *moonlab_scheduler_current_tenant_id(void) { return t_tenant; }
const char *moonlab_scheduler_current_request_id(void) { return t_request; }

void moonlab_scheduler_fire_completion_hook(const moonlab_job_t *job, const moonlab_job_results_t *res,
                                            const char *backend_name)
{
    pthread_mutex_lock(&g_lock);
    moonlab_completion_hook_fn hook = g_hook;
    void *ctx = g_hook_ctx;
    pthread_mutex_unlock(&g_lock);
    if (hook) {
        hook(job, res, backend_name, ctx);
        g_mi355x_hook_fires.fetch_add(1);
    }
}

int moonlab_scheduler_run(moonlab_job_t *j, moonlab_job_results_t *out)
{
    if (!j || !out) return MOONLAB_SCHED_BAD_ARG;
    if (j->num_shots <= 0) return MOONLAB_SCHED_BAD_ARG;
    memset(out, 0, sizeof(*out));
    // An unknown name is an error, never a silent fallback (scheduler.c:447-456).
    const char *bname = j->backend_name ? j->backend_name : "simulator";
    const moonlab_backend_t *be = moonlab_find_backend(bname);
    if (!be) return MOONLAB_SCHED_BACKEND_NOT_FOUND;
    const int rc = be->execute(j, out, be->ctx);
    if (rc != MOONLAB_SCHED_OK) {
        moonlab_job_results_free(out);
        return rc;
    }
    pthread_mutex_lock(&g_lock);
    moonlab_completion_hook_fn hook = g_hook;
    void *ctx = g_hook_ctx;
    pthread_mutex_unlock(&g_lock);
    if (hook) {
        hook(j, out, bname, ctx);
        g_mi355x_hook_fires.fetch_add(1);
    }
    return MOONLAB_SCHED_OK;
}

void moonlab_job_results_free(moonlab_job_results_t *r)
{
    if (!r) return;
    free(r->outcomes);
    free(r->worker_seconds);
    r->outcomes = nullptr;
    r->worker_seconds = nullptr;
}

int moonlab_scheduler_run_mpi(moonlab_job_t *job, moonlab_job_results_t *out, void *ctx)
{
    (void)job;
    (void)out;
    (void)ctx;
    return MOONLAB_SCHED_NOT_BUILT;
}

// Byte-identical to the reference dump (scheduler.c:492-537): the same keys,
// spacing, "%.17g" theta and "0x%016llx" seed, and the same size-probe rule
// (returns the length it would write; truncates when bufsize is short).
int moonlab_job_to_json(const moonlab_job_t *j, char *buf, size_t bufsize)
{
    if (!j) return MOONLAB_SCHED_BAD_ARG;
    int total = 0;
    auto put = [&](const char *fmt, auto... args) -> bool {
        char *dst = buf ? buf + total : nullptr;
        const size_t room = buf ? (bufsize > (size_t)total ? bufsize - (size_t)total : 0) : 0;
        const int need = snprintf(dst, room, fmt, args...);
        if (need < 0) return false;
        total += need;
        return true;
    };
    bool ok = put("{\n") && put("  \"schema\": \"moonlab/job/v0.7.0\",\n") &&
              put("  \"num_qubits\": %d,\n", j->num_qubits) && put("  \"num_shots\": %d,\n", j->num_shots) &&
              put("  \"num_workers\": %d,\n", j->num_workers) &&
              put("  \"rng_seed\": \"0x%016llx\",\n", (unsigned long long)j->rng_seed) && put("  \"gates\": [");
    const int ng = (int)j->gates.size();
    for (int g = 0; ok && g < ng; g++) {
        const GateRec &r = j->gates[(size_t)g];
        ok = put("\n    { \"type\": %d, \"target\": %d", r.type, r.target);
        if (ok && mlx::gate_is_two_qubit(r.type)) ok = put(", \"control\": %d", r.control);
        if (ok && r.has_param) ok = put(", \"theta\": %.17g", r.theta);
        if (ok) ok = put(" }%s", g + 1 < ng ? "," : "");
    }
    if (ok) ok = put("\n  ]\n") && put("}\n");
    return ok ? total : MOONLAB_SCHED_INTERNAL;
}

}  // extern "C"
