// vqe.cpp -- variational workloads on the persistent batch engine.
//
// The reference evaluates one parameter set at a time on the CPU:
// vqe_compute_energy prepares |HF>, applies the ansatz and sums
// c_k <P_k> term by term, each term on a cloned and basis-rotated state
// (src/algorithms/vqe.c:1287-1350, 2156-2257); vqe_compute_gradient shifts
// one angle at a time (vqe.c:2094-2130); vqe_compute_qgt and
// vqe_compute_berry_curvature replay the ansatz once per parameter with the
// parameter's generator inserted and form every overlap on the host
// (src/algorithms/vqe_qng.c:66-245); vqe_natural_gradient_direction solves
// (g + eps I) x = grad (vqe_qng.c:247-282).
//
// Here every circuit of every instance -- base points, shifted points,
// generator-inserted replays -- runs in ONE persistent batch launch, the
// Pauli sums are measured exactly on the device from the final states, and
// the overlaps of the geometric tensor come from a device Gram kernel over
// the states the launch writes out.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/moonlab_mi355x.h"
#include "batch_device.h"
#include "batch_plan.h"

namespace mlx {
hipError_t launch_gram(const double2 *S, int K, uint64_t dim, double2 *G, hipStream_t st);
}
int         batch_run_states(moonlab_mi355x_batch_t *b, moonlab_mi355x_batch_item_t *items, int count,
                             moonlab_mi355x_batch_stats_t *stats, const mlx::BatchStateOut *out);
void       *batch_state_buffer(moonlab_mi355x_batch_t *b, size_t bytes);
hipStream_t batch_stream(moonlab_mi355x_batch_t *b);

namespace {

using mlx::PauliT;

// Device memory the geometric-tensor calls may hold at once for replayed
// states (instances are grouped to fit; one instance alone must fit).
constexpr uint64_t kQgtStateBytes = 16ull << 30;

int num_params(const moonlab_mi355x_vqe_t &v) { return 2 * v.num_qubits * v.num_layers; }

bool valid(const moonlab_mi355x_vqe_t &v, bool need_terms)
{
    if (v.num_qubits < 1 || v.num_qubits > MOONLAB_MI355X_BATCH_MAX_QUBITS || v.num_layers < 0) return false;
    if (v.num_layers > 0 && !v.angles) return false;
    if (v.entangler != MOONLAB_MI355X_ENT_CZ && v.entangler != MOONLAB_MI355X_ENT_CNOT) return false;
    if (v.num_qubits < 64 && (v.hf_reference >> v.num_qubits) != 0) return false;
    if (need_terms) {
        if (v.num_terms < 0 || (v.num_terms > 0 && !v.terms)) return false;
        for (int k = 0; k < v.num_terms; k++) {
            PauliT p;
            if (!mlx::parse_pauli(v.terms[k].pauli_string, v.num_qubits, p)) return false;
        }
    }
    return true;
}

// |HF> then the ansatz (vqe.c:1302-1318, 764-781): X on the Hartree-Fock
// bits; per layer RY(theta_p), RZ(theta_p+1) on every qubit, then the ladder
// over (q, q+1): CNOT with control q, or CZ.  With insert >= 0 the generator
// of parameter `insert` (Y after its RY, Z after its RZ) follows its
// rotation (vqe_qng.c:86-99).
void ansatz_gates(const moonlab_mi355x_vqe_t &v, const double *ang, int insert, std::vector<moonlab_hip_gate_t> &g)
{
    const int n = v.num_qubits;
    g.clear();
    for (int q = 0; q < n; q++)
        if ((v.hf_reference >> q) & 1ull) g.push_back(moonlab_hip_gate_t{MOONLAB_HIP_GATE_X, q, -1, -1, 0, 0, 0});
    int p = 0;
    for (int l = 0; l < v.num_layers; l++) {
        for (int q = 0; q < n; q++) {
            g.push_back(moonlab_hip_gate_t{MOONLAB_HIP_GATE_RY, q, -1, -1, ang[p], 0, 0});
            if (p == insert) g.push_back(moonlab_hip_gate_t{MOONLAB_HIP_GATE_Y, q, -1, -1, 0, 0, 0});
            p++;
            g.push_back(moonlab_hip_gate_t{MOONLAB_HIP_GATE_RZ, q, -1, -1, ang[p], 0, 0});
            if (p == insert) g.push_back(moonlab_hip_gate_t{MOONLAB_HIP_GATE_Z, q, -1, -1, 0, 0, 0});
            p++;
        }
        for (int q = 0; q + 1 < n; q++)
            g.push_back(moonlab_hip_gate_t{v.entangler == MOONLAB_MI355X_ENT_CNOT ? MOONLAB_HIP_GATE_CNOT
                                                                                  : MOONLAB_HIP_GATE_CZ,
                                           q + 1, q, -1, 0, 0, 0});
    }
}

// Energies of `count` parameter sets (one per entry of `angle_sets`) of
// instances `which[i]`: one batch run, Pauli sums on the device.
int energies_of(moonlab_mi355x_batch_t *b, const moonlab_mi355x_vqe_t *inst, const std::vector<int> &which,
                const std::vector<const double *> &angle_sets, int precision, double *out,
                moonlab_mi355x_batch_stats_t *stats)
{
    const size_t m = which.size();
    std::vector<std::vector<moonlab_hip_gate_t>> gl(m);
    std::vector<moonlab_mi355x_batch_item_t> items(m);
    for (size_t i = 0; i < m; i++) {
        const moonlab_mi355x_vqe_t &v = inst[which[i]];
        ansatz_gates(v, angle_sets[i], -1, gl[i]);
        moonlab_mi355x_batch_item_t &it = items[i];
        memset(&it, 0, sizeof it);
        it.num_qubits = v.num_qubits;
        it.precision = precision;
        it.gates = gl[i].data();
        it.num_gates = (int)gl[i].size();
        it.num_pauli_terms = v.num_terms;
        it.pauli_terms = v.terms;
        it.pauli_energy = out + i;
    }
    const int rc = batch_run_states(b, items.data(), (int)m, stats, nullptr);
    if (rc) return rc;
    for (size_t i = 0; i < m; i++) {
        if (inst[which[i]].num_terms == 0) out[i] = 0.0;
        out[i] += inst[which[i]].constant;
    }
    return MOONLAB_SCHED_OK;
}

// Gaussian elimination with partial pivoting on an m x m row-major A
// (overwritten); false when a pivot column is zero.
bool solve_dense(int m, std::vector<double> &A, const double *rhs, double *x)
{
    std::vector<double> r(rhs, rhs + m);
    for (int c = 0; c < m; c++) {
        int piv = c;
        for (int i = c + 1; i < m; i++)
            if (fabs(A[(size_t)i * m + c]) > fabs(A[(size_t)piv * m + c])) piv = i;
        if (!(fabs(A[(size_t)piv * m + c]) > 0.0)) return false;
        if (piv != c) {
            for (int k = 0; k < m; k++) std::swap(A[(size_t)c * m + k], A[(size_t)piv * m + k]);
            std::swap(r[(size_t)c], r[(size_t)piv]);
        }
        const double d = A[(size_t)c * m + c];
        for (int i = c + 1; i < m; i++) {
            const double f = A[(size_t)i * m + c] / d;
            if (f == 0.0) continue;
            for (int k = c; k < m; k++) A[(size_t)i * m + k] -= f * A[(size_t)c * m + k];
            r[(size_t)i] -= f * r[(size_t)c];
        }
    }
    for (int i = m - 1; i >= 0; i--) {
        double s = r[(size_t)i];
        for (int k = i + 1; k < m; k++) s -= A[(size_t)i * m + k] * x[k];
        x[i] = s / A[(size_t)i * m + i];
    }
    return true;
}

}  // namespace

extern "C" {

int moonlab_mi355x_vqe_energies(moonlab_mi355x_batch_t *b, const moonlab_mi355x_vqe_t *inst, int count,
                                int precision, double *energies, moonlab_mi355x_batch_stats_t *stats)
{
    if (!b || count < 0 || (count > 0 && (!inst || !energies))) return MOONLAB_SCHED_BAD_ARG;
    if (precision != MOONLAB_HIP_C128 && precision != MOONLAB_HIP_C64) return MOONLAB_SCHED_BAD_ARG;
    std::vector<int> which((size_t)count);
    std::vector<const double *> sets((size_t)count);
    for (int i = 0; i < count; i++) {
        if (!valid(inst[i], true)) return MOONLAB_SCHED_BAD_ARG;
        which[(size_t)i] = i;
        sets[(size_t)i] = inst[i].angles;
    }
    if (count == 0) return MOONLAB_SCHED_OK;
    return energies_of(b, inst, which, sets, precision, energies, stats);
}

int moonlab_mi355x_vqe_gradients(moonlab_mi355x_batch_t *b, const moonlab_mi355x_vqe_t *inst, int count,
                                 int precision, double *energies, double *gradients,
                                 moonlab_mi355x_batch_stats_t *stats)
{
    if (!b || count < 0 || (count > 0 && (!inst || !energies || !gradients))) return MOONLAB_SCHED_BAD_ARG;
    if (precision != MOONLAB_HIP_C128 && precision != MOONLAB_HIP_C64) return MOONLAB_SCHED_BAD_ARG;
    size_t total = 0, total_ang = 0;
    for (int i = 0; i < count; i++) {
        if (!valid(inst[i], true)) return MOONLAB_SCHED_BAD_ARG;
        const size_t P = (size_t)num_params(inst[i]);
        total += 1 + 2 * P;
        total_ang += (1 + 2 * P) * P;
    }
    if (count == 0) return MOONLAB_SCHED_OK;
    // each instance: its base point, then theta +- pi/2 e_k for every k
    std::vector<double> ang(total_ang + 1);
    std::vector<int> which(total);
    std::vector<const double *> sets(total);
    size_t ai = 0, ei = 0;
    for (int i = 0; i < count; i++) {
        const size_t P = (size_t)num_params(inst[i]);
        for (size_t k = 0; k < 1 + 2 * P; k++) {
            double *a = ang.data() + ai;
            if (P) memcpy(a, inst[i].angles, P * sizeof(double));
            if (k > 0) a[(k - 1) / 2] += ((k - 1) % 2 == 0 ? 1.0 : -1.0) * (M_PI / 2.0);
            which[ei] = i;
            sets[ei] = a;
            ai += P;
            ei++;
        }
    }
    std::vector<double> e(total);
    const int rc = energies_of(b, inst, which, sets, precision, e.data(), stats);
    if (rc) return rc;
    // dE/dtheta_k = [E(theta + pi/2 e_k) - E(theta - pi/2 e_k)] / 2, exact for
    // the two-eigenvalue generators of RY and RZ (vqe.c:2094-2130)
    ei = 0;
    size_t gi = 0;
    for (int i = 0; i < count; i++) {
        const size_t P = (size_t)num_params(inst[i]);
        energies[i] = e[ei];
        for (size_t k = 0; k < P; k++) gradients[gi + k] = 0.5 * (e[ei + 1 + 2 * k] - e[ei + 2 + 2 * k]);
        ei += 1 + 2 * P;
        gi += P;
    }
    return MOONLAB_SCHED_OK;
}

int moonlab_mi355x_vqe_qgt(moonlab_mi355x_batch_t *b, const moonlab_mi355x_vqe_t *inst, int count, double *qgt,
                           double *berry, moonlab_mi355x_batch_stats_t *stats)
{
    if (!b || count < 0 || (count > 0 && !inst) || (!qgt && !berry)) return MOONLAB_SCHED_BAD_ARG;
    for (int i = 0; i < count; i++)
        if (!valid(inst[i], false)) return MOONLAB_SCHED_BAD_ARG;
    hipStream_t st = batch_stream(b);
    size_t out_off = 0;   // into qgt / berry, instance after instance
    int i0 = 0;
    while (i0 < count) {
        // a group of instances whose K = P + 1 states fit the state budget
        uint64_t amps = 0;
        int i1 = i0;
        size_t max_k = 0;
        while (i1 < count) {
            const uint64_t K = (uint64_t)num_params(inst[i1]) + 1, dim = 1ull << inst[i1].num_qubits;
            if (amps + K * dim > kQgtStateBytes / 16) break;
            amps += K * dim;
            if (K > max_k) max_k = K;
            i1++;
        }
        if (i1 == i0) return MOONLAB_SCHED_OOM;   // one instance's replays exceed the budget
        const size_t gram_bytes = max_k * max_k * sizeof(double2);
        char *buf = (char *)batch_state_buffer(b, amps * sizeof(double2) + gram_bytes + 256);
        if (!buf) return MOONLAB_SCHED_OOM;
        double2 *d_states = (double2 *)buf;
        double2 *d_gram = (double2 *)(buf + (amps * sizeof(double2) + 255) / 256 * 256);
        // items: per instance its base state, then one replay per parameter
        std::vector<std::vector<moonlab_hip_gate_t>> gl;
        std::vector<moonlab_mi355x_batch_item_t> items;
        std::vector<int64_t> offs;
        std::vector<uint64_t> inst_off;
        uint64_t o = 0;
        for (int i = i0; i < i1; i++) {
            const int P = num_params(inst[i]);
            const uint64_t dim = 1ull << inst[i].num_qubits;
            inst_off.push_back(o);
            for (int k = -1; k < P; k++) {
                gl.emplace_back();
                ansatz_gates(inst[i], inst[i].angles, k, gl.back());
                offs.push_back((int64_t)o);
                o += dim;
            }
        }
        items.resize(gl.size());
        size_t c = 0;
        for (int i = i0; i < i1; i++) {
            const int P = num_params(inst[i]);
            for (int k = -1; k < P; k++, c++) {
                moonlab_mi355x_batch_item_t &it = items[c];
                memset(&it, 0, sizeof it);
                it.num_qubits = inst[i].num_qubits;
                it.precision = MOONLAB_HIP_C128;
                it.gates = gl[c].data();
                it.num_gates = (int)gl[c].size();
            }
        }
        const mlx::BatchStateOut so{d_states, offs.data()};
        int rc = batch_run_states(b, items.data(), (int)items.size(), stats, &so);
        if (rc) return rc;
        // overlaps: G = S^H S over [psi, phi_0 .. phi_P-1], phi_i the replay
        // with generator i, d_i psi = (-i/2) phi_i; then
        // Q_ij = (1/4) [<phi_i|phi_j> - <phi_i|psi><psi|phi_j>],
        // metric Re Q (vqe_qng.c:185-200), curvature -2 Im Q (:220-240)
        for (int i = i0; i < i1; i++) {
            const int P = num_params(inst[i]), K = P + 1;
            const uint64_t dim = 1ull << inst[i].num_qubits;
            hipError_t e = mlx::launch_gram(d_states + inst_off[(size_t)(i - i0)], K, dim, d_gram, st);
            std::vector<double2> G((size_t)K * K);
            if (e == hipSuccess)
                e = hipMemcpyAsync(G.data(), d_gram, G.size() * sizeof(double2), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) return MOONLAB_SCHED_INTERNAL;
            auto at = [&](int a, int bb) -> double2 {   // <S_a|S_b> from the upper triangle
                if (a <= bb) return G[(size_t)a * K + bb];
                const double2 t = G[(size_t)bb * K + a];
                return double2{t.x, -t.y};
            };
            for (int r = 0; r < P; r++)
                for (int s = r; s < P; s++) {
                    const double2 pp = at(r + 1, s + 1);   // <phi_r|phi_s>
                    const double2 rp = at(r + 1, 0);       // <phi_r|psi>
                    const double2 ps = at(0, s + 1);       // <psi|phi_s>
                    const double qr = 0.25 * (pp.x - (rp.x * ps.x - rp.y * ps.y));
                    const double qi = 0.25 * (pp.y - (rp.x * ps.y + rp.y * ps.x));
                    if (qgt) {
                        qgt[out_off + (size_t)r * P + s] = qr;
                        qgt[out_off + (size_t)s * P + r] = qr;
                    }
                    if (berry) {
                        const double f = r == s ? 0.0 : -2.0 * qi;
                        berry[out_off + (size_t)r * P + s] = f;
                        berry[out_off + (size_t)s * P + r] = -f;
                    }
                }
            out_off += (size_t)P * P;
        }
        i0 = i1;
    }
    return MOONLAB_SCHED_OK;
}

int moonlab_mi355x_vqe_natural_gradient(moonlab_mi355x_batch_t *b, const moonlab_mi355x_vqe_t *inst, int count,
                                        const double *gradients, double regularization, double *directions,
                                        int *status, moonlab_mi355x_batch_stats_t *stats)
{
    if (!b || count < 0 || (count > 0 && (!inst || !gradients || !directions))) return MOONLAB_SCHED_BAD_ARG;
    if (!std::isfinite(regularization)) return MOONLAB_SCHED_BAD_ARG;
    size_t total = 0;
    for (int i = 0; i < count; i++) {
        if (!valid(inst[i], false)) return MOONLAB_SCHED_BAD_ARG;
        total += (size_t)num_params(inst[i]) * (size_t)num_params(inst[i]);
    }
    std::vector<double> g(total + 1);
    const int rc = moonlab_mi355x_vqe_qgt(b, inst, count, g.data(), nullptr, stats);
    if (rc) return rc;
    size_t go = 0, vo = 0;
    for (int i = 0; i < count; i++) {
        const int P = num_params(inst[i]);
        std::vector<double> A(g.begin() + (long)go, g.begin() + (long)(go + (size_t)P * P));
        for (int k = 0; k < P; k++) A[(size_t)k * P + k] += regularization;   // Tikhonov shift
        const bool ok = solve_dense(P, A, gradients + vo, directions + vo);
        if (!ok) memcpy(directions + vo, gradients + vo, (size_t)P * sizeof(double));
        if (status) status[i] = ok ? 0 : -1;
        go += (size_t)P * P;
        vo += (size_t)P;
    }
    return MOONLAB_SCHED_OK;
}

}  // extern "C"
