// engine.h -- internal interface of the circuit engine shared by the job API,
// the circuit API and the MI355X backend entry points.
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/moonlab_hip_statevec.h"

namespace mlx {

// One recorded gate of the job / circuit APIs (scheduler.c:40-46,
// moonlab_qgtl_backend.c:29-35): the tag, its qubits and theta.
struct GateRec {
    int    type;
    int    target;
    int    control;
    double theta;
    int    has_param;
    // gates added through moonlab_mi355x_circuit_add_gate_ex (codes 14..24 of
    // the state-vector gate table): second control or target, phi and lambda
    int    extra = -1;
    double phi = 0.0, lam = 0.0;
    int    ext = 0;
};

bool gate_takes_param(int type);   // RX, RY, RZ
bool gate_is_two_qubit(int type);  // CNOT, CY, CZ, SWAP

// Checks the circuit API applies at add time (moonlab_qgtl_backend.c:93-99).
bool gate_rec_valid(const GateRec &g, int num_qubits);

// splitmix64 seed mixer and xorshift64 draw of the job path
// (scheduler.c:143-148, moonlab_qgtl_backend.c:159-167, 238-239).
uint64_t splitmix64(uint64_t x);
double   xorshift_u(uint64_t *s);

// Per-worker shot slicing (scheduler.c:201-236): the uniform draws for all
// shots in slice order.  seed == 0 takes the reference's clock-based seed.
int fill_job_draws(uint64_t seed, int total_shots, int num_workers, std::vector<double> &u);

struct RunRequest {
    int                 num_qubits;
    int                 precision;
    const GateRec      *gates;
    int                 num_gates;
    const double       *draws;          // num_draws uniform values, or nullptr
    int                 num_draws;
    bool                want_amps;
    bool                want_probs;
    bool                want_z;
    int                 fusion_max;     // 0 = unfused
    bool                blocked;        // cache-blocked tile passes
    void               *stream = nullptr;   // run on this HIP stream instead of the state's own
};

struct RunResult {
    std::vector<uint64_t> outcomes;
    std::vector<double>   amps;         // 2 * 2^n (widened for complex64)
    std::vector<double>   probs;
    std::vector<double>   z;
    double                norm = 0.0;
    double                device_seconds = 0.0;
};

// Run a gate list from |0..0> on a pooled device state.  Returns a
// moonlab_hip_status_t value; the text is in moonlab_hip_last_error().
int run_circuit(const RunRequest &rq, RunResult &rr);
// The same with gates of the full code table (rq.gates is ignored).
int run_circuit_gates(const RunRequest &rq, const moonlab_hip_gate_t *gates, int num_gates, RunResult &rr);

}  // namespace mlx
