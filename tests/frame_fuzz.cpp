// frame_fuzz.cpp -- test infrastructure: host-only fuzz of the control-frame
// parser (control_frame.cpp) and the circuit text parser (qgtl_circuit.cpp),
// built with AddressSanitizer and UBSan by tests/test_abi_cpu.py.  Frames are
// mutated, truncated or random bytes in exact-size heap buffers, so any read
// past a frame is caught.  The device entry points are stubs that abort: with
// no server (srv = NULL) no frame may reach them.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include "moonlab_mi355x.h"
#include "moonlab_scheduler.h"
#include "engine.h"
namespace mlx {
bool gate_takes_param(int t) { return t == 7 || t == 8 || t == 9; }
bool gate_is_two_qubit(int t) { return t == 10 || t == 11 || t == 12 || t == 13; }
bool gate_rec_valid(const GateRec &, int) { return true; }
double xorshift_u(uint64_t *s) { *s ^= *s << 13; *s ^= *s >> 7; *s ^= *s << 17; return (double)(*s >> 11) / 9007199254740992.0; }
int run_circuit(const RunRequest &, RunResult &) { abort(); }
}
extern "C" {
int moonlab_mi355x_server_submit(moonlab_mi355x_server_t *, const moonlab_mi355x_request_t *, uint64_t *) { abort(); }
int moonlab_mi355x_server_wait(moonlab_mi355x_server_t *, uint64_t, moonlab_mi355x_response_t *, double) { abort(); }
void moonlab_mi355x_response_free(moonlab_mi355x_response_t *) { abort(); }
const char *moonlab_hip_last_error(void) { return ""; }
void moonlab_scheduler_set_request_context(const char *, const char *) { abort(); }
void moonlab_scheduler_fire_completion_hook(const moonlab_job_t *, const moonlab_job_results_t *, const char *) { abort(); }
}
int main() {
    std::mt19937_64 rng(7);
    const char *seeds[] = {"HEALTH\n", "CIRCUIT 40\nNUM_QUBITS 3\nH 0\nCNOT 1 0\nRZ 2 0.25\n#\n", "SHOTS 16 28\nNUM_QUBITS 2\nH 0\nCNOT 1 0\n\n",
                           "CIRCUIT 49\nNUM_QUBITS 3\nP 0 nan\nU3 1 0.5 inf 0.25\nCCX 2 0 1\n"};
    long n = 0;
    for (int i = 0; i < 200000; i++) {
        std::string f = seeds[i % 4];
        int k = rng() % 3;
        if (k == 0) { int c = 1 + rng() % 5; for (int j = 0; j < c && !f.empty(); j++) f[rng() % f.size()] = (char)(rng() % 256); }
        else if (k == 1) f.resize(rng() % (f.size() + 1));
        else { f.resize(rng() % 300); for (char &ch : f) ch = (char)(rng() % 256); }
        std::vector<char> buf(f.begin(), f.end());   // exact-size heap buffer: overreads are caught
        void *out = nullptr; size_t len = 0;
        moonlab_mi355x_server_handle_frame(nullptr, buf.empty() ? "" : buf.data(), buf.size(), 0, &out, &len);
        free(out);
        n++;
    }
    printf("fuzzed %ld frames\n", n);
}
