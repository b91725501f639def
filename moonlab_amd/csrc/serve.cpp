// serve.cpp -- continuous-batching request loop over the MI355X engine.
//
// The reference's serving path (src/control/control_plane.c:587-748) parses a
// request and runs moonlab_qgtl_execute synchronously on the connection's
// thread: one full CPU simulation per request, no batching, no residency.
// Here:
//   * submit() copies the request, draws its shot uniforms with the circuit
//     API's xorshift64 stream (moonlab_qgtl_backend.c:151-167, 226-240) and
//     queues it;
//   * admission: a request leaves the queue only while its state footprint
//     (2^n x 16 or 8 B) fits the HBM budget;
//   * small requests are co-batched: each batch engine thread takes every
//     admitted small request that arrived since it last launched (up to
//     max_batch) into ONE persistent batch launch; with two engines, one batch
//     is built and launched while the other runs, so arrivals never wait for
//     a whole batch to drain before getting on the GPU;
//   * large requests run as cache-blocked passes (register-phase tiles; or
//     fused <= 5-qubit blocks on the matrix cores, large_path = 1), each large
//     worker on one stream of its own for its whole life, overlapping with the
//     batches on the device;
//   * sequential mode is the plain baseline: one request at a time, one
//     kernel per gate.
// Results never depend on co-tenants: the batch engine's per-circuit planning
// and reduction order, and the single-state kernels, are independent of what
// else runs.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/moonlab_mi355x.h"
#include "engine.h"
#include "fusion.h"

namespace {

struct Req {
    uint64_t                        id = 0;
    int                             n = 0, prec = 0, shots = 0;
    bool                            want_z = false, want_probs = false;
    std::vector<moonlab_hip_gate_t> gates;
    std::vector<double>             draws;
    uint64_t                        footprint = 0;
    double                          t_submit = 0, t_start = 0, t_finish = 0;
    int                             status = 0, path = 0;
    bool                            done = false;
    std::vector<uint64_t>           outcomes;
    std::vector<double>             z;
    std::vector<double>             probs;
    double                          norm = 0;
};

double clock_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct moonlab_mi355x_server {
    moonlab_mi355x_server_config_t   cfg{};
    mutable std::mutex               mu;
    std::condition_variable          cv_work, cv_done;
    std::deque<Req *>                small_q, large_q, seq_q;
    std::unordered_map<uint64_t, Req *> reqs;
    uint64_t                         next_id = 1;
    uint64_t                         admitted_bytes = 0;
    bool                             stopping = false;
    std::vector<std::thread>         threads;
    std::vector<moonlab_mi355x_batch_t *> engines;
    std::vector<hipStream_t>         large_streams;   // one per large worker
    moonlab_mi355x_server_stats_t    stats{};
    double                           t0 = 0;
};

namespace {

void finish(moonlab_mi355x_server_t *s, Req *r)
{
    // caller holds s->mu
    r->t_finish = clock_s() - s->t0;
    r->done = true;
    s->admitted_bytes -= r->footprint;
    s->stats.completed++;
}

// Take admitted requests from the front of q while the budget allows.
std::vector<Req *> admit(moonlab_mi355x_server_t *s, std::deque<Req *> &q, int max_take)
{
    std::vector<Req *> got;
    while (!q.empty() && (int)got.size() < max_take) {
        Req *r = q.front();
        if (s->admitted_bytes + r->footprint > s->cfg.hbm_budget_bytes && s->admitted_bytes > 0) break;
        q.pop_front();
        s->admitted_bytes += r->footprint;
        if (s->admitted_bytes > s->stats.peak_bytes_admitted) s->stats.peak_bytes_admitted = s->admitted_bytes;
        got.push_back(r);
    }
    return got;
}

void batch_thread(moonlab_mi355x_server_t *s, moonlab_mi355x_batch_t *eng)
{
    std::vector<moonlab_mi355x_batch_item_t> items;
    for (;;) {
        std::vector<Req *> take;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stopping || !s->small_q.empty(); });
            if (s->stopping && s->small_q.empty()) return;
            take = admit(s, s->small_q, s->cfg.max_batch);
            if (take.empty()) {
                s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
                continue;
            }
            const double now = clock_s() - s->t0;
            for (Req *r : take) r->t_start = now;
            s->stats.batches++;
            s->stats.batched_requests += take.size();
            if ((int)take.size() > s->stats.largest_batch) s->stats.largest_batch = (int)take.size();
        }
        items.assign(take.size(), moonlab_mi355x_batch_item_t{});
        for (size_t i = 0; i < take.size(); i++) {
            Req *r = take[i];
            r->outcomes.resize((size_t)r->shots);
            r->z.resize((size_t)r->n);
            moonlab_mi355x_batch_item_t &it = items[i];
            it.num_qubits = r->n;
            it.precision = r->prec;
            it.gates = r->gates.data();
            it.num_gates = (int)r->gates.size();
            it.observables = r->want_z ? MOONLAB_MI355X_OBS_Z : 0;
            it.num_shots = r->shots;
            it.draws = r->draws.data();
            it.norm = &r->norm;
            it.expect_z = r->want_z ? r->z.data() : nullptr;
            it.outcomes = r->shots ? r->outcomes.data() : nullptr;
        }
        const int rc = moonlab_mi355x_batch_run(eng, items.data(), (int)items.size(), nullptr);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Req *r : take) {
                r->status = rc;
                r->path = 0;
                finish(s, r);
            }
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

// mode 0: one kernel per gate; 1: cache-blocked passes; 2: fused <= 5-qubit blocks
void run_single(Req *r, int mode, hipStream_t stream)
{
    mlx::RunRequest rq{};
    rq.num_qubits = r->n;
    rq.precision = r->prec;
    rq.draws = r->draws.empty() ? nullptr : r->draws.data();
    rq.num_draws = r->shots;
    rq.want_z = r->want_z;
    rq.want_probs = r->want_probs;
    rq.blocked = mode == 1;
    rq.fusion_max = mode == 2 ? 5 : 0;
    rq.stream = stream;
    mlx::RunResult rr;
    r->status = mlx::run_circuit_gates(rq, r->gates.data(), (int)r->gates.size(), rr);
    if (r->status == 0) {
        r->outcomes = std::move(rr.outcomes);
        r->z = std::move(rr.z);
        r->probs = std::move(rr.probs);
        r->norm = rr.norm;
    } else {
        r->status = MOONLAB_SCHED_INTERNAL;
    }
}

void large_thread(moonlab_mi355x_server_t *s, hipStream_t stream)
{
    // one stream per worker for its whole life: a request's kernels queue
    // behind its own worker's only, not behind every pooled state's stream
    for (;;) {
        Req *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                if (s->stopping && s->large_q.empty()) return true;
                if (s->large_q.empty()) return false;
                const Req *h = s->large_q.front();
                return s->admitted_bytes == 0 || s->admitted_bytes + h->footprint <= s->cfg.hbm_budget_bytes;
            });
            if (s->large_q.empty()) break;
            std::vector<Req *> got = admit(s, s->large_q, 1);
            if (got.empty()) continue;
            r = got[0];
            r->t_start = clock_s() - s->t0;
            s->stats.large_requests++;
        }
        run_single(r, s->cfg.large_path == 1 ? 2 : 1, stream);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            r->path = 1;
            finish(s, r);
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

void sequential_thread(moonlab_mi355x_server_t *s)
{
    for (;;) {
        Req *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stopping || !s->seq_q.empty(); });
            if (s->seq_q.empty()) return;
            r = s->seq_q.front();
            s->seq_q.pop_front();
            s->admitted_bytes += r->footprint;
            if (s->admitted_bytes > s->stats.peak_bytes_admitted) s->stats.peak_bytes_admitted = s->admitted_bytes;
            r->t_start = clock_s() - s->t0;
        }
        run_single(r, 0, nullptr);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            r->path = 2;
            finish(s, r);
        }
        s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

moonlab_mi355x_server_t *moonlab_mi355x_server_create(const moonlab_mi355x_server_config_t *cfg)
{
    if (moonlab_hip_device_count() <= 0) return nullptr;
    moonlab_mi355x_server_t *s = new (std::nothrow) moonlab_mi355x_server_t();
    if (!s) return nullptr;
    s->cfg.hbm_budget_bytes = 256ull << 30;
    s->cfg.batch_max_qubits = 16;
    s->cfg.max_batch = 1024;
    s->cfg.batch_engines = 2;
    s->cfg.large_workers = 2;
    if (cfg) {
        if (cfg->hbm_budget_bytes) s->cfg.hbm_budget_bytes = cfg->hbm_budget_bytes;
        if (cfg->batch_max_qubits > 0) s->cfg.batch_max_qubits = cfg->batch_max_qubits;
        if (cfg->max_batch > 0) s->cfg.max_batch = cfg->max_batch;
        if (cfg->batch_engines > 0) s->cfg.batch_engines = cfg->batch_engines;
        if (cfg->large_workers > 0) s->cfg.large_workers = cfg->large_workers;
        s->cfg.large_path = cfg->large_path == 1 ? 1 : 0;
        s->cfg.sequential = cfg->sequential;
    }
    if (s->cfg.batch_max_qubits > MOONLAB_MI355X_BATCH_MAX_QUBITS) s->cfg.batch_max_qubits = MOONLAB_MI355X_BATCH_MAX_QUBITS;
    s->t0 = clock_s();
    if (s->cfg.sequential) {
        s->threads.emplace_back(sequential_thread, s);
        return s;
    }
    // The engines' streams and the large workers' streams are created here,
    // one after another: HIP hands streams to its hardware queues in turn
    // (4 queues per process by default, GPU_MAX_HW_QUEUES), so with the
    // default two engines and two workers each stream gets a queue of its own
    // and a batch launch never waits behind a large request's passes in one
    // queue, whatever streams the process made before.
    for (int i = 0; i < s->cfg.batch_engines; i++) {
        moonlab_mi355x_batch_t *b = moonlab_mi355x_batch_create();
        if (!b) {
            moonlab_mi355x_server_free(s);
            return nullptr;
        }
        s->engines.push_back(b);
    }
    for (int i = 0; i < s->cfg.large_workers; i++) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            moonlab_mi355x_server_free(s);
            return nullptr;
        }
        s->large_streams.push_back(st);
    }
    for (moonlab_mi355x_batch_t *b : s->engines) s->threads.emplace_back(batch_thread, s, b);
    for (hipStream_t st : s->large_streams) s->threads.emplace_back(large_thread, s, st);
    return s;
}

void moonlab_mi355x_server_free(moonlab_mi355x_server_t *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    for (std::thread &t : s->threads) t.join();
    for (moonlab_mi355x_batch_t *b : s->engines) moonlab_mi355x_batch_free(b);
    for (hipStream_t st : s->large_streams) hipStreamDestroy(st);
    for (auto &kv : s->reqs) delete kv.second;
    delete s;
}

double moonlab_mi355x_server_now(const moonlab_mi355x_server_t *s) { return s ? clock_s() - s->t0 : 0.0; }

int moonlab_mi355x_server_submit(moonlab_mi355x_server_t *s, const moonlab_mi355x_request_t *q, uint64_t *ticket)
{
    if (!s || !q || !ticket) return MOONLAB_SCHED_BAD_ARG;
    const int cap = q->precision == MOONLAB_HIP_C64 ? MOONLAB_HIP_MAX_QUBITS_C64 : MOONLAB_HIP_MAX_QUBITS_C128;
    if (q->num_qubits < 1 || q->num_qubits > cap || q->num_gates < 0 || (q->num_gates > 0 && !q->gates) ||
        q->num_shots < 0 || (q->precision != MOONLAB_HIP_C128 && q->precision != MOONLAB_HIP_C64))
        return MOONLAB_SCHED_BAD_ARG;
    // A co-batch launch reports one status for all of its requests, so a
    // gate the engine would reject must fail its own submit, never the batch
    // of other tenants it would have joined.
    for (int i = 0; i < q->num_gates; i++)
        if (!mlx::gate_valid(q->gates[i], q->num_qubits)) return MOONLAB_SCHED_BAD_ARG;
    Req *r = new (std::nothrow) Req();
    if (!r) return MOONLAB_SCHED_OOM;
    r->n = q->num_qubits;
    r->prec = q->precision;
    r->shots = q->num_shots;
    r->want_z = q->want_expect_z != 0;
    r->want_probs = q->want_probabilities != 0;
    r->gates.assign(q->gates, q->gates + q->num_gates);
    uint64_t seed = q->rng_seed ? q->rng_seed : ((uint64_t)time(NULL) ^ 0x9e3779b97f4a7c15ULL);
    r->draws.resize((size_t)q->num_shots);
    for (int k = 0; k < q->num_shots; k++) r->draws[(size_t)k] = mlx::xorshift_u(&seed);
    r->footprint = (1ull << r->n) * (r->prec == MOONLAB_HIP_C64 ? 8ull : 16ull);
    if (r->footprint > s->cfg.hbm_budget_bytes) {
        delete r;
        return MOONLAB_SCHED_OOM;   // can never be admitted
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        r->t_submit = clock_s() - s->t0;
        s->reqs[r->id] = r;
        s->stats.submitted++;
        // the batch kernel returns observables and shots, not 2^n
        // probabilities: a probability request takes the single-state path
        if (s->cfg.sequential) s->seq_q.push_back(r);
        else if (r->n <= s->cfg.batch_max_qubits && !r->want_probs) s->small_q.push_back(r);
        else s->large_q.push_back(r);
        *ticket = r->id;
    }
    s->cv_work.notify_all();
    return MOONLAB_SCHED_OK;
}

int moonlab_mi355x_server_wait(moonlab_mi355x_server_t *s, uint64_t ticket, moonlab_mi355x_response_t *out,
                               double timeout_s)
{
    if (!s || !out) return MOONLAB_SCHED_BAD_ARG;
    memset(out, 0, sizeof(*out));
    Req *r = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return MOONLAB_SCHED_BAD_ARG;
        r = it->second;
        const auto pred = [&] { return r->done; };
        if (timeout_s < 0) s->cv_done.wait(lk, pred);
        else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
            return MOONLAB_SCHED_BACKEND_BUSY;
        s->reqs.erase(it);
    }
    out->status = r->status;
    out->num_qubits = r->n;
    out->num_shots = r->shots;
    out->norm = r->norm;
    out->submit_time = r->t_submit;
    out->start_time = r->t_start;
    out->finish_time = r->t_finish;
    out->path = r->path;
    if (r->status == 0) {
        if (r->shots) {
            out->outcomes = (uint64_t *)malloc((size_t)r->shots * sizeof(uint64_t));
            if (out->outcomes) memcpy(out->outcomes, r->outcomes.data(), (size_t)r->shots * sizeof(uint64_t));
        }
        if (r->want_z) {
            out->expect_z = (double *)malloc((size_t)r->n * sizeof(double));
            if (out->expect_z) memcpy(out->expect_z, r->z.data(), (size_t)r->n * sizeof(double));
        }
        if (r->want_probs) {
            out->probabilities = (double *)malloc(r->probs.size() * sizeof(double));
            if (out->probabilities) memcpy(out->probabilities, r->probs.data(), r->probs.size() * sizeof(double));
        }
        if ((r->shots && !out->outcomes) || (r->want_z && !out->expect_z) || (r->want_probs && !out->probabilities))
            out->status = MOONLAB_SCHED_OOM;
    }
    const int st = out->status;
    delete r;
    return st;
}

void moonlab_mi355x_response_free(moonlab_mi355x_response_t *r)
{
    if (!r) return;
    free(r->outcomes);
    free(r->expect_z);
    free(r->probabilities);
    r->outcomes = nullptr;
    r->expect_z = nullptr;
    r->probabilities = nullptr;
}

int moonlab_mi355x_server_stats(const moonlab_mi355x_server_t *s, moonlab_mi355x_server_stats_t *st)
{
    if (!s || !st) return MOONLAB_SCHED_BAD_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    return MOONLAB_SCHED_OK;
}

}  // extern "C"
