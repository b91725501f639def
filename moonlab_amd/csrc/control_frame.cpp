// control_frame.cpp -- the control plane's wire format as the front door of
// the continuous-batching server.
//
// The reference's control plane (src/control/control_plane.c:294-748) reads
// one request frame per connection, runs moonlab_qgtl_execute synchronously
// on the connection's thread and writes one response frame.  Here a frame
// goes through the serving loop instead: every connection thread that calls
// moonlab_mi355x_server_handle_frame submits to the same server, so requests
// from concurrent connections are co-batched and share the device.  The
// transport (sockets, TLS, AUTH, METRICS) stays the reference's; this file
// holds only the frame grammar, its limits and its error replies.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/moonlab_mi355x.h"
#include "../../include/moonlab_scheduler.h"
#include "engine.h"

namespace mlx {
const std::vector<GateRec> &qgtl_circuit_gates(const moonlab_qgtl_circuit_t *c);
}

namespace {

// Limits of control_plane.c:264-292.
constexpr long kMaxBody = 1L << 22;    // 4 MB circuit text
constexpr long kMaxShots = 1L << 20;   // 1 M samples
constexpr int kMaxQubitsDefault = 24, kMaxQubitsCeil = 30;
constexpr size_t kVerbLine = 96;       // the reference reads the verb line into 96 bytes

int max_qubits()
{
    int cap = kMaxQubitsDefault;
    if (const char *e = getenv("MOONLAB_CONTROL_MAX_QUBITS"))
        if (*e && atoi(e) > 0) cap = atoi(e);
    if (cap > kMaxQubitsCeil) cap = kMaxQubitsCeil;
    return cap;
}

int reply(std::string &out, int status, const char *msg)
{
    char hdr[128];
    snprintf(hdr, sizeof hdr, "ERR %d %s\n", status, msg);
    out = hdr;
    return status;
}

template <typename T>
void append_binary(std::string &out, const T *v, size_t n)
{
    out.append((const char *)v, n * sizeof(T));
}

int handle(moonlab_mi355x_server_t *srv, const char *f, size_t len, uint64_t seed, const char *tenant,
           std::string &out)
{
    const char *nl = (const char *)memchr(f, '\n', len < kVerbLine ? len : kVerbLine);
    if (!nl) return reply(out, MOONLAB_CONTROL_PROTOCOL, "no verb line");   // EOF before '\n', control_plane.c:176
    const std::string verb(f, (size_t)(nl - f));
    if (verb == "HEALTH") {
        out = "OK alive\n";
        return MOONLAB_CONTROL_OK;
    }
    bool shots_mode = false;
    long num_shots = 0, body_bytes = -1;
    if (verb.compare(0, 8, "CIRCUIT ") == 0) {
        if (sscanf(verb.c_str() + 8, "%ld", &body_bytes) != 1)
            return reply(out, MOONLAB_CONTROL_PROTOCOL, "expected CIRCUIT <N>");
    } else if (verb.compare(0, 6, "SHOTS ") == 0) {
        if (sscanf(verb.c_str() + 6, "%ld %ld", &num_shots, &body_bytes) != 2)
            return reply(out, MOONLAB_CONTROL_PROTOCOL, "expected SHOTS <shots> <bytes>");
        if (num_shots <= 0 || num_shots > kMaxShots) return reply(out, MOONLAB_CONTROL_BAD_ARG, "shots out of range");
        shots_mode = true;
    } else {
        return reply(out, MOONLAB_CONTROL_PROTOCOL, "unknown verb");
    }
    if (body_bytes <= 0 || body_bytes > kMaxBody) return reply(out, MOONLAB_CONTROL_BAD_ARG, "body bytes out of range");
    const char *body = nl + 1;
    if ((size_t)(f + len - body) < (size_t)body_bytes) return reply(out, MOONLAB_CONTROL_PROTOCOL, "short body");   // a short read, control_plane.c:160

    int st = 0;
    moonlab_qgtl_circuit_t *c = moonlab_qgtl_circuit_deserialize(body, (size_t)body_bytes, &st);
    if (!c) return reply(out, MOONLAB_CONTROL_REJECTED, "deserialize");
    const int n = moonlab_qgtl_circuit_num_qubits(c);
    if (n > max_qubits()) {
        moonlab_qgtl_circuit_free(c);
        return reply(out, MOONLAB_CONTROL_BAD_ARG, "qubit count exceeds server cap");
    }
    const std::vector<mlx::GateRec> &recs = mlx::qgtl_circuit_gates(c);
    std::vector<moonlab_hip_gate_t> gates(recs.size());
    for (size_t i = 0; i < recs.size(); i++)
        gates[i] = moonlab_hip_gate_t{recs[i].type, recs[i].target, recs[i].control, recs[i].extra,
                                      recs[i].theta,   recs[i].phi,    recs[i].lam};
    moonlab_qgtl_circuit_free(c);
    if (!srv) return reply(out, MOONLAB_CONTROL_REJECTED, "execute");

    moonlab_mi355x_request_t q{};
    q.num_qubits = n;
    q.precision = MOONLAB_HIP_C128;   // the circuit API's precision
    q.gates = gates.data();
    q.num_gates = (int)gates.size();
    q.num_shots = shots_mode ? (int)num_shots : 0;
    q.rng_seed = seed;                 // 0: clock-based, as control_plane.c:662-664
    q.want_probabilities = shots_mode ? 0 : 1;
    // Tenant attribution as control_plane.c:699-731: the tenant (NULL when
    // the connection has none) is the calling thread's request context while
    // the request runs, and a completed run fires the completion hook with
    // the fields billing / audit hooks read (num_qubits, total_shots).
    moonlab_scheduler_set_request_context(tenant && *tenant ? tenant : nullptr, nullptr);
    uint64_t ticket = 0;
    if (moonlab_mi355x_server_submit(srv, &q, &ticket) != MOONLAB_SCHED_OK) {
        moonlab_scheduler_set_request_context(nullptr, nullptr);
        return reply(out, MOONLAB_CONTROL_REJECTED, "execute");
    }
    moonlab_mi355x_response_t r;
    const int wr = moonlab_mi355x_server_wait(srv, ticket, &r, -1.0);
    if (wr != MOONLAB_SCHED_OK || r.status != MOONLAB_SCHED_OK) {
        moonlab_mi355x_response_free(&r);
        moonlab_scheduler_set_request_context(nullptr, nullptr);
        return reply(out, MOONLAB_CONTROL_REJECTED, "execute");
    }
    {
        moonlab_job_results_t hook_res{};
        hook_res.num_qubits = n;
        hook_res.total_shots = shots_mode ? r.num_shots : 0;
        moonlab_scheduler_fire_completion_hook(nullptr, &hook_res, "control_plane");
    }
    moonlab_scheduler_set_request_context(nullptr, nullptr);
    char hdr[64];
    if (shots_mode) {
        snprintf(hdr, sizeof hdr, "SAMPLES %d\n", r.num_shots);
        out = hdr;
        append_binary(out, r.outcomes, (size_t)r.num_shots);
    } else {
        const size_t dim = (size_t)1 << n;
        snprintf(hdr, sizeof hdr, "OK %zu\n", dim);
        out = hdr;
        append_binary(out, r.probabilities, dim);
    }
    moonlab_mi355x_response_free(&r);
    return MOONLAB_CONTROL_OK;
}

}  // namespace

extern "C" int moonlab_mi355x_server_handle_frame_ex(moonlab_mi355x_server_t *srv, const void *frame, size_t len,
                                                     uint64_t seed, const char *tenant_id, void **out_response,
                                                     size_t *out_len)
{
    if (!out_response || !out_len) return MOONLAB_CONTROL_BAD_ARG;
    *out_response = nullptr;
    *out_len = 0;
    std::string out;
    const int rc = frame ? handle(srv, (const char *)frame, len, seed, tenant_id, out)
                         : reply(out, MOONLAB_CONTROL_BAD_ARG, "no frame");
    void *buf = malloc(out.size());
    if (!buf) return MOONLAB_CONTROL_OOM;
    memcpy(buf, out.data(), out.size());
    *out_response = buf;
    *out_len = out.size();
    return rc;
}

extern "C" int moonlab_mi355x_server_handle_frame(moonlab_mi355x_server_t *srv, const void *frame, size_t len,
                                                  uint64_t seed, void **out_response, size_t *out_len)
{
    return moonlab_mi355x_server_handle_frame_ex(srv, frame, len, seed, nullptr, out_response, out_len);
}
