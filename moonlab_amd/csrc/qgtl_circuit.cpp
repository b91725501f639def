// qgtl_circuit.cpp -- circuit-ingestion API and the "# moonlab-circuit v1"
// text format, with the reference's validation and status rules
// (src/applications/moonlab_qgtl_backend.c).  moonlab_qgtl_execute runs the
// gate list on the MI355X engine; probabilities use re^2 + im^2 as the
// reference's execute does (moonlab_qgtl_backend.c:205-209); shots use the
// reference's xorshift64 stream and its collapse outcome rule.
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <new>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/moonlab_mi355x.h"
#include "engine.h"

using mlx::GateRec;

struct moonlab_qgtl_circuit {
    int                  num_qubits;
    std::vector<GateRec> gates;
};

namespace {

const char *kNames[14] = {"I", "X", "Y", "Z", "H", "S", "T", "RX", "RY", "RZ", "CNOT", "CY", "CZ", "SWAP"};

// Extension gates (moonlab_mi355x_circuit_add_gate_ex): codes 14..24 of the
// state-vector gate table, their text names, qubit count and parameters.
struct ExtGate {
    const char *name;
    int         nq, np;
};
const ExtGate kExt[11] = {{"SDG", 1, 0},    {"TDG", 1, 0},    {"P", 1, 1},    {"U3", 1, 3},
                          {"CPHASE", 2, 1}, {"CRX", 2, 1},    {"CRY", 2, 1},  {"CRZ", 2, 1},
                          {"CYT", 2, 0},    {"CCX", 3, 0},    {"CSWAP", 3, 0}};
constexpr int kExtFirst = MOONLAB_HIP_GATE_SDG;

int ext_from_name(const char *s)
{
    for (int i = 0; i < 11; i++)
        if (strcmp(s, kExt[i].name) == 0) return kExtFirst + i;
    return -1;
}

int gate_from_name(const char *s, int *out)
{
    for (int i = 0; i < 14; i++)
        if (strcmp(s, kNames[i]) == 0) {
            *out = i;
            return MOONLAB_QGTL_OK;
        }
    return MOONLAB_QGTL_BAD_ARG;
}

// Next non-blank, non-comment line, leading whitespace trimmed.
// 1 = line, 0 = end, -1 = line longer than the buffer.
int next_line(const char **p, const char *end, char *line, size_t cap)
{
    while (*p < end) {
        const char *start = *p, *q = *p;
        while (q < end && *q != '\n') q++;
        *p = q < end ? q + 1 : q;
        const char *ls = start;
        while (ls < q && isspace((unsigned char)*ls)) ls++;
        if (ls >= q || *ls == '#') continue;
        const size_t len = (size_t)(q - ls);
        if (len + 1 > cap) return -1;
        memcpy(line, ls, len);
        line[len] = '\0';
        return 1;
    }
    return 0;
}

}  // namespace

extern "C" {

moonlab_qgtl_circuit_t *moonlab_qgtl_circuit_create(int n)
{
    if (n < 1 || n > 32) return nullptr;
    moonlab_qgtl_circuit_t *c = new (std::nothrow) moonlab_qgtl_circuit_t();
    if (c) c->num_qubits = n;
    return c;
}

void moonlab_qgtl_circuit_free(moonlab_qgtl_circuit_t *c) { delete c; }

int moonlab_qgtl_circuit_num_qubits(const moonlab_qgtl_circuit_t *c)
{
    return c ? c->num_qubits : MOONLAB_QGTL_BAD_ARG;
}

int moonlab_qgtl_circuit_num_gates(const moonlab_qgtl_circuit_t *c)
{
    return c ? (int)c->gates.size() : MOONLAB_QGTL_BAD_ARG;
}

int moonlab_qgtl_add_gate(moonlab_qgtl_circuit_t *c, moonlab_qgtl_gate_t type, int target, int control,
                          const double *params)
{
    if (!c) return MOONLAB_QGTL_BAD_ARG;
    if (target < 0 || target >= c->num_qubits) return MOONLAB_QGTL_BAD_ARG;
    if (mlx::gate_is_two_qubit((int)type)) {
        if (control < 0 || control >= c->num_qubits || control == target) return MOONLAB_QGTL_BAD_ARG;
    }
    const bool p = mlx::gate_takes_param((int)type);
    if (p && !params) return MOONLAB_QGTL_BAD_ARG;
    try {
        c->gates.push_back(GateRec{(int)type, target, control, p ? params[0] : 0.0, p ? 1 : 0});
    } catch (...) {
        return MOONLAB_QGTL_OOM;
    }
    return MOONLAB_QGTL_OK;
}

int moonlab_mi355x_circuit_add_gate_ex(moonlab_qgtl_circuit_t *c, int type, const int *qubits, int num_qubits,
                                       const double *params, int num_params)
{
    if (!c || type < kExtFirst || type > MOONLAB_HIP_GATE_FREDKIN) return MOONLAB_QGTL_BAD_ARG;
    const ExtGate &e = kExt[type - kExtFirst];
    if (num_qubits != e.nq || !qubits || num_params != e.np || (e.np && !params)) return MOONLAB_QGTL_BAD_ARG;
    for (int j = 0; j < e.nq; j++) {
        if (qubits[j] < 0 || qubits[j] >= c->num_qubits) return MOONLAB_QGTL_BAD_ARG;
        for (int k = 0; k < j; k++)
            if (qubits[k] == qubits[j]) return MOONLAB_QGTL_BAD_ARG;
    }
    // the extension grammar is this library's own: angles must be finite
    // (a NaN angle would answer OK with an all-NaN state)
    for (int j = 0; j < e.np; j++)
        if (!std::isfinite(params[j])) return MOONLAB_QGTL_BAD_ARG;
    GateRec g{type, qubits[0], e.nq > 1 ? qubits[1] : -1, e.np ? params[0] : 0.0, e.np ? 1 : 0};
    g.extra = e.nq > 2 ? qubits[2] : -1;
    g.phi = e.np > 1 ? params[1] : 0.0;
    g.lam = e.np > 2 ? params[2] : 0.0;
    g.ext = 1;
    try {
        c->gates.push_back(g);
    } catch (...) {
        return MOONLAB_QGTL_OOM;
    }
    return MOONLAB_QGTL_OK;
}

int moonlab_qgtl_execute(moonlab_qgtl_circuit_t *c, const moonlab_qgtl_exec_options_t *opts,
                         moonlab_qgtl_results_t *out)
{
    if (!c || !opts || !out) return MOONLAB_QGTL_BAD_ARG;
    // codes outside the reference's 14 run only when added as extensions
    for (const GateRec &g : c->gates)
        if (g.type < 0 || (g.type > MOONLAB_QGTL_GATE_SWAP && !g.ext)) return MOONLAB_QGTL_UNSUPPORTED;
    std::vector<double> u;
    if (opts->num_shots > 0) {
        uint64_t seed = opts->rng_seed ? opts->rng_seed : ((uint64_t)time(NULL) ^ 0x9e3779b97f4a7c15ULL);
        u.resize((size_t)opts->num_shots);
        for (int k = 0; k < opts->num_shots; k++) u[(size_t)k] = mlx::xorshift_u(&seed);
    }
    mlx::RunRequest rq{};
    rq.num_qubits = c->num_qubits;
    rq.precision = MOONLAB_HIP_C128;
    rq.gates = c->gates.data();
    rq.num_gates = (int)c->gates.size();
    rq.draws = u.empty() ? nullptr : u.data();
    rq.num_draws = opts->num_shots > 0 ? opts->num_shots : 0;
    rq.want_probs = opts->return_probabilities != 0;
    mlx::RunResult rr;
    const int st = mlx::run_circuit(rq, rr);
    if (st) {
        fprintf(stderr, "moonlab-mi355x: %s\n", moonlab_hip_last_error());
        return st == MOONLAB_HIP_ERR_OUT_OF_MEMORY ? MOONLAB_QGTL_OOM : MOONLAB_QGTL_INTERNAL;
    }
    memset(out, 0, sizeof(*out));
    out->num_qubits = c->num_qubits;
    out->num_shots = opts->num_shots;
    if (opts->return_probabilities) {
        out->probabilities = (double *)malloc(rr.probs.size() * sizeof(double));
        if (!out->probabilities) return MOONLAB_QGTL_OOM;
        memcpy(out->probabilities, rr.probs.data(), rr.probs.size() * sizeof(double));
    }
    if (opts->num_shots > 0) {
        out->outcomes = (uint64_t *)malloc((size_t)opts->num_shots * sizeof(uint64_t));
        if (!out->outcomes) {
            free(out->probabilities);
            out->probabilities = nullptr;
            return MOONLAB_QGTL_OOM;
        }
        memcpy(out->outcomes, rr.outcomes.data(), (size_t)opts->num_shots * sizeof(uint64_t));
    }
    return MOONLAB_QGTL_OK;
}

void moonlab_qgtl_results_free(moonlab_qgtl_results_t *r)
{
    if (!r) return;
    free(r->outcomes);
    free(r->probabilities);
    r->outcomes = nullptr;
    r->probabilities = nullptr;
}

int moonlab_qgtl_circuit_serialize(const moonlab_qgtl_circuit_t *c, char *buf, size_t buf_size, size_t *out_written)
{
    if (!c) return MOONLAB_QGTL_BAD_ARG;
    if (buf_size > 0 && !buf) return MOONLAB_QGTL_BAD_ARG;
    size_t pos = 0;
    char tmp[96];
    // Append one chunk; on overflow report the size needed so far plus NUL
    // (moonlab_qgtl_backend.c:303-364).
    auto emit = [&](int n) -> bool {
        if (n < 0) return false;
        if (buf_size > 0) {
            if (pos + (size_t)n + 1 > buf_size) {
                if (out_written) *out_written = pos + (size_t)n + 1;
                return false;
            }
            memcpy(buf + pos, tmp, (size_t)n);
        }
        pos += (size_t)n;
        return true;
    };
    int n = snprintf(tmp, sizeof tmp, "# moonlab-circuit v1\n");
    if (n < 0) return MOONLAB_QGTL_INTERNAL;
    if (!emit(n)) return MOONLAB_QGTL_OOM;
    n = snprintf(tmp, sizeof tmp, "NUM_QUBITS %d\n", c->num_qubits);
    if (n < 0) return MOONLAB_QGTL_INTERNAL;
    if (!emit(n)) return MOONLAB_QGTL_OOM;
    for (const GateRec &g : c->gates) {
        if (g.ext) {   // extension gates: NAME qubits... params...
            const ExtGate &e = kExt[g.type - kExtFirst];
            const int qs[3] = {g.target, g.control, g.extra};
            const double ps[3] = {g.theta, g.phi, g.lam};
            n = snprintf(tmp, sizeof tmp, "%s", e.name);
            for (int j = 0; j < e.nq && n > 0; j++) n += snprintf(tmp + n, sizeof tmp - (size_t)n, " %d", qs[j]);
            for (int j = 0; j < e.np && n > 0; j++) n += snprintf(tmp + n, sizeof tmp - (size_t)n, " %.17g", ps[j]);
            if (n > 0) n += snprintf(tmp + n, sizeof tmp - (size_t)n, "\n");
            if (n <= 0 || (size_t)n >= sizeof tmp) return MOONLAB_QGTL_INTERNAL;
            if (!emit(n)) return MOONLAB_QGTL_OOM;
            continue;
        }
        if (g.type < 0 || g.type > 13) return MOONLAB_QGTL_INTERNAL;
        const char *name = kNames[g.type];
        if (mlx::gate_is_two_qubit(g.type)) n = snprintf(tmp, sizeof tmp, "%s %d %d\n", name, g.target, g.control);
        else if (g.has_param) n = snprintf(tmp, sizeof tmp, "%s %d %.17g\n", name, g.target, g.theta);
        else n = snprintf(tmp, sizeof tmp, "%s %d\n", name, g.target);
        if (n < 0) return MOONLAB_QGTL_INTERNAL;
        if (!emit(n)) return MOONLAB_QGTL_OOM;
    }
    if (buf_size > 0) {
        if (pos + 1 > buf_size) {
            if (out_written) *out_written = pos + 1;
            return MOONLAB_QGTL_OOM;
        }
        buf[pos] = '\0';
    }
    if (out_written) *out_written = pos;
    return MOONLAB_QGTL_OK;
}

moonlab_qgtl_circuit_t *moonlab_qgtl_circuit_deserialize(const char *buf, size_t buf_size, int *out_status)
{
    int status = MOONLAB_QGTL_BAD_ARG;
    moonlab_qgtl_circuit_t *c = nullptr;
    const char *p, *end;
    char line[256];
    bool got_header = false;
    if (!buf) goto done;
    if (buf_size == 0) buf_size = strlen(buf);
    else if (memchr(buf, '\0', buf_size)) goto done;   // NUL inside a bounded payload
    p = buf;
    end = buf + buf_size;
    for (;;) {
        const int ls = next_line(&p, end, line, sizeof line);
        if (ls == 0) break;
        if (ls < 0) { status = MOONLAB_QGTL_BAD_ARG; goto done; }
        if (!got_header) {
            int n = -1;
            if (sscanf(line, "NUM_QUBITS %d", &n) != 1 || n < 1 || n > 32) { status = MOONLAB_QGTL_BAD_ARG; goto done; }
            c = moonlab_qgtl_circuit_create(n);
            if (!c) { status = MOONLAB_QGTL_OOM; goto done; }
            got_header = true;
            continue;
        }
        char name[8] = {0};
        int consumed = 0, type = 0, target = -1, control = -1, rc;
        double theta = 0.0;
        if (sscanf(line, "%7s %n", name, &consumed) != 1) { status = MOONLAB_QGTL_BAD_ARG; goto done; }
        const char *rest = line + consumed;
        if (gate_from_name(name, &type) != MOONLAB_QGTL_OK) {
            // an extension gate: its qubits, then its parameters
            const int xt = ext_from_name(name);
            if (xt < 0) { status = MOONLAB_QGTL_BAD_ARG; goto done; }
            const ExtGate &e = kExt[xt - kExtFirst];
            int qs[3] = {-1, -1, -1};
            double ps[3] = {0.0, 0.0, 0.0};
            const char *p = rest;
            bool ok = true;
            for (int j = 0; j < e.nq && ok; j++) {
                int used = 0;
                ok = sscanf(p, "%d%n", &qs[j], &used) == 1;
                p += used;
            }
            for (int j = 0; j < e.np && ok; j++) {
                int used = 0;
                ok = sscanf(p, "%lf%n", &ps[j], &used) == 1;
                p += used;
            }
            if (!ok) { status = MOONLAB_QGTL_BAD_ARG; goto done; }
            rc = moonlab_mi355x_circuit_add_gate_ex(c, xt, qs, e.nq, ps, e.np);
            if (rc != MOONLAB_QGTL_OK) { status = rc; goto done; }
            continue;
        }
        if (mlx::gate_is_two_qubit(type)) {
            if (sscanf(rest, "%d %d", &target, &control) != 2) { status = MOONLAB_QGTL_BAD_ARG; goto done; }
            rc = moonlab_qgtl_add_gate(c, (moonlab_qgtl_gate_t)type, target, control, nullptr);
        } else if (mlx::gate_takes_param(type)) {
            if (sscanf(rest, "%d %lf", &target, &theta) != 2) { status = MOONLAB_QGTL_BAD_ARG; goto done; }
            rc = moonlab_qgtl_add_gate(c, (moonlab_qgtl_gate_t)type, target, 0, &theta);
        } else {
            if (sscanf(rest, "%d", &target) != 1) { status = MOONLAB_QGTL_BAD_ARG; goto done; }
            rc = moonlab_qgtl_add_gate(c, (moonlab_qgtl_gate_t)type, target, 0, nullptr);
        }
        if (rc != MOONLAB_QGTL_OK) { status = rc; goto done; }
    }
    if (!got_header) { status = MOONLAB_QGTL_BAD_ARG; goto done; }
    if (out_status) *out_status = MOONLAB_QGTL_OK;
    return c;
done:
    moonlab_qgtl_circuit_free(c);
    if (out_status) *out_status = status;
    return nullptr;
}

int moonlab_qgtl_circuit_save(const moonlab_qgtl_circuit_t *c, const char *path)
{
    if (!c || !path) return MOONLAB_QGTL_BAD_ARG;
    size_t need = 0;
    int rc = moonlab_qgtl_circuit_serialize(c, nullptr, 0, &need);
    if (rc != MOONLAB_QGTL_OK) return rc;
    std::string buf(need + 1, '\0');
    rc = moonlab_qgtl_circuit_serialize(c, &buf[0], need + 1, nullptr);
    if (rc != MOONLAB_QGTL_OK) return rc;
    FILE *fp = fopen(path, "w");
    if (!fp) return MOONLAB_QGTL_INTERNAL;
    const size_t wrote = fwrite(buf.data(), 1, need, fp);
    const bool ok = wrote == need && fclose(fp) == 0;
    return ok ? MOONLAB_QGTL_OK : MOONLAB_QGTL_INTERNAL;
}

moonlab_qgtl_circuit_t *moonlab_qgtl_circuit_load(const char *path, int *out_status)
{
    if (!path) {
        if (out_status) *out_status = MOONLAB_QGTL_BAD_ARG;
        return nullptr;
    }
    FILE *fp = fopen(path, "r");
    if (!fp) {
        if (out_status) *out_status = MOONLAB_QGTL_INTERNAL;
        return nullptr;
    }
    std::string data;
    char chunk[4096];
    size_t got;
    while ((got = fread(chunk, 1, sizeof chunk, fp)) > 0) data.append(chunk, got);
    fclose(fp);
    return moonlab_qgtl_circuit_deserialize(data.c_str(), data.size(), out_status);
}

}  // extern "C"

namespace mlx {

const std::vector<GateRec> &qgtl_circuit_gates(const moonlab_qgtl_circuit_t *c) { return c->gates; }

}  // namespace mlx
