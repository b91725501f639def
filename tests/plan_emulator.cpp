// plan_emulator.cpp -- test infrastructure: executes the cache-blocked pass
// plans of the batch planner (lower + plan_passes, register segments
// included) on a host state, following the op semantics of the device
// kernels (apply_op / apply_segment in batch.hip), and compares the result
// with the oracle's gate-by-gate run.  It checks the host planner -- tile
// sets, op reordering across passes, segment grouping -- independently of a
// GPU; the kernels themselves are checked on the GPU against the oracle.
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>
#include <climits>
#include "batch_device.h"
#include "batch_plan.h"
extern "C" {
#include "sv_oracle.h"
}
using cd = std::complex<double>;
static uint64_t ins0(uint64_t x, int q) { return ((x >> q) << (q + 1)) | (x & ((1ull << q) - 1)); }
static int bit_of(int32_t code, int l, uint64_t base) { if (code == INT32_MIN) return 0; return code >= 0 ? (l >> code) & 1 : (int)((base >> (-1 - code)) & 1ull); }
static void mat1(cd &a, cd &b, const mlx::BOp &op) {
    cd m0(op.m[0], op.m[1]), m1(op.m[2], op.m[3]), m2(op.m[4], op.m[5]), m3(op.m[6], op.m[7]);
    cd x = m0 * a + m1 * b, y = m2 * a + m3 * b; a = x; b = y; }
static void apply_op(std::vector<cd> &t, int T, const mlx::BOp &op, uint64_t base) {
    if ((base & op.cg) != op.cg) return;
    if (op.kind == mlx::BOP_MAT1) {
        for (int p = 0; p < T / 2; p++) { int i0 = (int)ins0(p, op.lq), i1 = i0 | (1 << op.lq); if ((i0 & (int)op.cl) != (int)op.cl) continue; mat1(t[i0], t[i1], op); }
    } else if (op.kind == mlx::BOP_XPERM) {
        for (int p = 0; p < T / 2; p++) { int i0 = (int)ins0(p, op.lq), i1 = i0 | (1 << op.lq); if ((i0 & (int)op.cl) != (int)op.cl) continue; std::swap(t[i0], t[i1]); }
    } else if (op.kind == mlx::BOP_DIAG) {
        for (int l = 0; l < T; l++) { if ((l & (int)op.cl) != (int)op.cl) continue; int k = (bit_of(op.lq2, l, base) << 1) | bit_of(op.lq, l, base); t[l] *= cd(op.m[2 * k], op.m[2 * k + 1]); }
    } else {
        int b1 = 1 << op.lq, b2 = 1 << op.lq2;
        for (int l = 0; l < T; l++) if ((l & b1) && !(l & b2) && (l & (int)op.cl) == (int)op.cl) std::swap(t[l], t[(l ^ b1) | b2]);
    }
}
static void apply_segment(std::vector<cd> &t, int T, const mlx::BOp *ops, int nops, uint32_t packed, int NR, uint64_t base) {
    int rq[3]; for (int i = 0; i < NR; i++) rq[i] = (packed >> (5 * i)) & 31;
    for (int i = 1; i < NR; i++) if (rq[i] <= rq[i - 1]) { printf("regs not ascending\n"); exit(1); }
    for (int g = 0; g < (T >> NR); g++) {
        int gb = g; for (int i = 0; i < NR; i++) gb = (int)ins0(gb, rq[i]);
        int idx[8]; cd v[8];
        for (int j = 0; j < (1 << NR); j++) { idx[j] = gb; for (int i = 0; i < NR; i++) if (j & (1 << i)) idx[j] |= 1 << rq[i]; v[j] = t[idx[j]]; }
        for (int k = 0; k < nops; k++) {
            const mlx::BOp &op = ops[k];
            if ((base & op.cg) != op.cg) continue;
            if (op.kind == mlx::BOP_DIAG) {
                for (int j = 0; j < (1 << NR); j++) { if ((idx[j] & (int)op.cl) != (int)op.cl) continue; int d = (bit_of(op.lq2, idx[j], base) << 1) | bit_of(op.lq, idx[j], base); v[j] *= cd(op.m[2 * d], op.m[2 * d + 1]); }
            } else if (op.kind == mlx::BOP_MAT1 || op.kind == mlx::BOP_XPERM) {
                int p = -1; for (int i = 0; i < NR; i++) if (rq[i] == op.lq) p = i;
                if (p < 0) { printf("target not in registers\n"); exit(1); }
                for (int j = 0; j < (1 << NR); j++) {
                    if (j & (1 << p)) continue;
                    if ((idx[j] & (int)op.cl) != (int)op.cl) continue;
                    if (op.kind == mlx::BOP_MAT1) mat1(v[j], v[j | (1 << p)], op);
                    else std::swap(v[j], v[j | (1 << p)]);
                }
            } else {
                int p1 = -1, p2 = -1; for (int i = 0; i < NR; i++) { if (rq[i] == op.lq) p1 = i; if (rq[i] == op.lq2) p2 = i; }
                if (p1 < 0 || p2 < 0) { printf("SWAP target not in registers\n"); exit(1); }
                for (int j = 0; j < (1 << NR); j++) { if (!(j & (1 << p1)) || (j & (1 << p2))) continue; if ((idx[j] & (int)op.cl) != (int)op.cl) continue; std::swap(v[j], v[(j ^ (1 << p1)) | (1 << p2)]); }
            }
        }
        for (int j = 0; j < (1 << NR); j++) t[idx[j]] = v[j];
    }
}
int main() {
    std::mt19937_64 rng(9);
    int segs_seen = 0, checked = 0;
    for (int trial = 0; trial < 60; trial++) {
        int n = 4 + rng() % 11;
        std::vector<moonlab_hip_gate_t> g;
        for (int k = 0; k < 150; k++) {
            int t = 1 + rng() % 24; int q[3]; for (int i = 0; i < 3; i++) { do { q[i] = rng() % n; } while ((i > 0 && q[i] == q[0]) || (i > 1 && q[i] == q[1])); }
            if (n < 3 && (t == 23 || t == 24)) t = 4;
            g.push_back({t, q[0], q[1], q[2], (double)(rng() % 1000) / 77.0, (double)(rng() % 1000) / 91.0, (double)(rng() % 1000) / 53.0});
        }
        std::vector<double> want(2ull << n, 0.0); want[0] = 1.0;
        svo_apply_gates(n, want.data(), (const svo_gate_t *)g.data(), (int)g.size());
        for (int tb : {3, 4, 6, 10, 12}) for (int sr : {2, 3}) {
            int tbb = tb < n ? tb : n;
            std::vector<mlx::GOp> gops; if (!mlx::lower(g.data(), (int)g.size(), n, gops)) { printf("lower failed\n"); return 1; }
            std::vector<mlx::BPass> passes; std::vector<mlx::BOp> ops;
            mlx::plan_passes(gops, n, tbb, sr, passes, ops);
            std::vector<cd> st(1ull << n, cd(0)); st[0] = 1.0;
            for (const mlx::BPass &P : passes) {
                int T = 1 << P.nbits; uint64_t ntiles = 1ull << (n - P.nbits);
                for (uint64_t ti = 0; ti < ntiles; ti++) {
                    uint64_t base = ti; for (int b = 0; b < P.nbits; b++) base = ins0(base, P.qubits[b]);
                    std::vector<cd> t(T);
                    auto addr = [&](int l) { uint64_t a = base; for (int b = 0; b < P.nbits; b++) if (l & (1 << b)) a |= 1ull << P.qubits[b]; return a; };
                    for (int l = 0; l < T; l++) t[l] = st[addr(l)];
                    for (int k = 0; k < P.num_ops;) {
                        const mlx::BOp &op = ops[P.first_op + k];
                        if (op.kind == mlx::BOP_SEG) { apply_segment(t, T, &ops[P.first_op + k + 1], op.lq, op.cl, op.lq2, base); k += 1 + op.lq; segs_seen++; }
                        else { apply_op(t, T, op, base); k++; }
                    }
                    for (int l = 0; l < T; l++) st[addr(l)] = t[l];
                }
            }
            double err = 0; for (uint64_t i = 0; i < (1ull << n); i++) err = std::max(err, std::abs(st[i] - cd(want[2 * i], want[2 * i + 1])));
            if (err > 1e-12) { printf("MISMATCH n=%d tb=%d sr=%d err=%g\n", n, tbb, sr, err); return 1; }
            checked++;
        }
    }
    printf("ok: %d plans checked, %d segment executions\n", checked, segs_seen);
}
