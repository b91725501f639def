// plan_emulator.cpp -- test infrastructure: executes the cache-blocked pass
// plans of the batch planner (lower + plan_passes, register segments
// included) on a host state, following the op semantics of the device
// kernels (apply_op / apply_segment in batch.hip), and compares the result
// with the oracle's gate-by-gate run.  Register-phase programs (k_reg_pass)
// run thread by thread with the kernel's LDS slots and layout changes
// (the layout helpers are the ones batch_device.h gives the kernel).  It checks the host planner -- tile
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

static long g_rot2_ops = 0;   // ARM_ROT2 ops run (paired rotations)
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

// k_reg_pass on one tile, step by step: 2^(TB-4) threads of 16 register
// values, LDS slots swizzled for BYTES-byte amplitudes, a layout change =
// every thread writes, then every thread reads.
template <int BYTES>
static void reg_tile(std::vector<cd> &tile, int TB, int RB, const mlx::BOp *ops, int nops, uint64_t base, int &relayouts) {
    const int NV = 1 << RB, TH = 1 << (TB - RB);
    std::vector<cd> lds(1 << TB);
    std::vector<int> written(1 << TB);
    std::vector<std::vector<cd>> v(TH, std::vector<cd>(NV));
    std::vector<int> tloc(TH);
    mlx::RegLayout io{}; for (int i = 0; i < RB; i++) io.r[i] = TB - RB + i;
    const uint32_t io_mask = ((1u << RB) - 1) << (TB - RB);
    for (int tid = 0; tid < TH; tid++) { tloc[tid] = tid; for (int j = 0; j < NV; j++) v[tid][j] = tile[tid + TH * j]; }
    // from_off / to_off: byte offsets of the register values' slots, as the
    // kernel gets them (computed for the IO layout, from the phase header
    // otherwise)
    auto relayout = [&](const uint32_t *from_off, const uint32_t *to_off, const mlx::RegLayout &to) {
        relayouts++;
        std::fill(written.begin(), written.end(), 0);
        for (int tid = 0; tid < TH; tid++) { uint32_t s0 = (uint32_t)mlx::lds_slot<BYTES>(tloc[tid]) * BYTES; for (int j = 0; j < NV; j++) { uint32_t b = s0 ^ from_off[j]; if (b % BYTES || b / BYTES >= (1u << TB)) { printf("slot out of range\n"); exit(1); } written[b / BYTES]++; lds[b / BYTES] = v[tid][j]; } }
        for (int x : written) if (x != 1) { printf("layout change does not write every slot once\n"); exit(1); }
        for (int tid = 0; tid < TH; tid++) { tloc[tid] = mlx::thread_base(to, tid, RB); uint32_t s1 = (uint32_t)mlx::lds_slot<BYTES>(tloc[tid]) * BYTES; for (int j = 0; j < NV; j++) v[tid][j] = lds[(s1 ^ to_off[j]) / BYTES]; }
    };
    uint32_t io_off[16], cur_off[16];
    for (int j = 0; j < NV; j++) cur_off[j] = io_off[j] = (uint32_t)mlx::lds_slot<BYTES>(j << (TB - RB)) * BYTES;
    mlx::RegLayout cur = io; uint32_t cur_mask = io_mask;
    for (int k = 0; k < nops;) {
        const mlx::BOp &h = ops[k];
        if (h.kind != mlx::BOP_PHASE || h.lq < 1) { printf("bad phase header\n"); exit(1); }
        const uint32_t mask = (uint32_t)h.lq2;
        mlx::RegLayout nl = mlx::unpack_layout(h.cl, RB);
        uint32_t m2 = 0; for (int i = 0; i < RB; i++) { m2 |= 1u << nl.r[i]; if (i && nl.r[i] <= nl.r[i - 1]) { printf("phase bits not ascending\n"); exit(1); } if (nl.r[i] >= TB) { printf("phase bit outside tile\n"); exit(1); } }
        if (m2 != mask) { printf("phase mask does not match its bits\n"); exit(1); }
        if (mask != cur_mask) {
            uint32_t nof[16]; memcpy(nof, h.m, sizeof nof);
            relayout(cur_off, nof, nl); memcpy(cur_off, nof, sizeof nof); cur = nl; cur_mask = mask;
        }
        for (int q = 1; q <= h.lq; q++) {
            const mlx::BOp &op = ops[k + q];
            if ((base & op.cg) != op.cg) continue;
            if (op.kind != mlx::BOP_DIAG && op.cl != 0) { printf("register-phase op with a thread-bit control\n"); exit(1); }
            const uint32_t jm = op.pad2;
            const int arm = op.pad;
            // the kernel's arms (sv_regpass.hip: reg_op), value by value
            for (int tid = 0; tid < TH; tid++) {
                std::vector<cd> &w = v[tid];
                if (arm < mlx::ARM_XPERM) {   // MAT1 arms
                    const int P = arm % 4, form = arm < mlx::ARM_MAT1 ? 0 : (arm - mlx::ARM_MAT1) / 4;
                    const bool masked = arm < mlx::ARM_MAT1;
                    if (op.kind != mlx::BOP_MAT1 || P != op.lq || P >= RB) { printf("MAT1 arm mismatch\n"); exit(1); }
                    for (int j = 0; j < NV; j++) {
                        if (j & (1 << P)) continue;
                        if (masked && ((jm >> j) & 1) != ((jm >> (j | (1 << P))) & 1)) { printf("control mask differs across the target\n"); exit(1); }
                        if (masked && !((jm >> j) & 1)) continue;
                        cd &a = w[j], &b = w[j | (1 << P)];
                        if (form == 1) { double m0 = op.m[0], m2 = op.m[1], m4 = op.m[2], m6 = op.m[3]; cd x = m0 * a + m2 * b, y = m4 * a + m6 * b; a = x; b = y; }
                        else if (form == 2) { cd u01(0, op.m[1]), u10(0, op.m[2]); cd x = op.m[0] * a + u01 * b, y = u10 * a + op.m[3] * b; a = x; b = y; }
                        else mat1(a, b, op);
                    }
                } else if (arm < mlx::ARM_CX) {   // XPERM on jm
                    const int P = arm - mlx::ARM_XPERM;
                    if (op.kind != mlx::BOP_XPERM || P != op.lq || P >= RB) { printf("XPERM arm mismatch\n"); exit(1); }
                    for (int j = 0; j < NV; j++) {
                        if (j & (1 << P)) continue;
                        if (((jm >> j) & 1) != ((jm >> (j | (1 << P))) & 1)) { printf("control mask differs across the target\n"); exit(1); }
                        if ((jm >> j) & 1) std::swap(w[j], w[j | (1 << P)]);
                    }
                } else if (arm < mlx::ARM_DIAG1) {   // CNOT, target P, control C
                    const int P = (arm - mlx::ARM_CX) / 4, C = (arm - mlx::ARM_CX) % 4;
                    if (op.kind != mlx::BOP_XPERM || P != op.lq || P == C || P >= RB || C >= RB) { printf("CNOT arm mismatch\n"); exit(1); }
                    for (int j = 0; j < NV; j++) if (!(j & (1 << P)) && (j & (1 << C))) std::swap(w[j], w[j | (1 << P)]);
                } else if (arm < mlx::ARM_DIAG11) {   // diag(d0, d1) on P
                    const int P = (arm - mlx::ARM_DIAG1) % 4; const bool only1 = arm >= mlx::ARM_DIAG1_ONLY1;
                    if (op.kind != mlx::BOP_DIAG || P >= RB) { printf("DIAG1 arm mismatch\n"); exit(1); }
                    for (int j = 0; j < NV; j++) {
                        if (j & (1 << P)) w[j] *= cd(op.m[2], op.m[3]);
                        else if (!only1) w[j] *= cd(op.m[0], op.m[1]);
                    }
                } else if (arm < mlx::ARM_DIAG) {   // diag(1, 1, 1, d3) on P < Q
                    const int P = (arm - mlx::ARM_DIAG11) / 4, Q = (arm - mlx::ARM_DIAG11) % 4;
                    if (op.kind != mlx::BOP_DIAG || P >= Q || Q >= RB) { printf("DIAG11 arm mismatch\n"); exit(1); }
                    for (int j = 0; j < NV; j++) if ((j & (1 << P)) && (j & (1 << Q))) w[j] *= cd(op.m[0], op.m[1]);
                } else if (arm == mlx::ARM_DIAG) {
                    if (op.kind != mlx::BOP_DIAG) { printf("DIAG arm on a non-diagonal op\n"); exit(1); }
                    auto gbit = [&](int32_t c) { return (c < 0 && c != INT32_MIN) ? (int)((base >> (-1 - c)) & 1) : 0; };
                    const int kg = gbit(op.lq) | (gbit(op.lq2) << 1);
                    for (int j = 0; j < NV; j++) {
                        if (!((jm >> j) & 1)) continue;
                        const int kk = kg | (int)((op.cl >> (2 * j)) & 3u);
                        w[j] *= cd(op.m[2 * kk], op.m[2 * kk + 1]);
                    }
                } else if (arm >= mlx::ARM_ROT2 && arm < mlx::ARM_ROT2 + 24) {   // two rotations, P < Q
                    static const int pp[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
                    const int pi = (arm - mlx::ARM_ROT2) / 4, F = ((arm - mlx::ARM_ROT2) >> 1) & 1, G = (arm - mlx::ARM_ROT2) & 1;
                    const int P = pp[pi][0], Q = pp[pi][1];
                    if (op.kind != mlx::BOP_MAT1 || P != op.lq || Q != op.lq2 || Q >= RB || mlx::rot2_pair(P, Q) != pi) { printf("ROT2 arm mismatch\n"); exit(1); }
                    auto rot = [&](int form, const double *m, cd &a, cd &b) {
                        if (form == 0) { cd x = m[0] * a + m[1] * b, y = m[2] * a + m[3] * b; a = x; b = y; }
                        else { cd x = m[0] * a + cd(0, m[1]) * b, y = cd(0, m[2]) * a + m[3] * b; a = x; b = y; }
                    };
                    if (tid == 0) g_rot2_ops++;
                    for (int j = 0; j < NV; j++) if (!(j & (1 << P))) rot(F, op.m, w[j], w[j | (1 << P)]);
                    for (int j = 0; j < NV; j++) if (!(j & (1 << Q))) rot(G, op.m + 4, w[j], w[j | (1 << Q)]);
                } else { printf("unknown arm %d\n", arm); exit(1); }
            }
        }
        k += 1 + h.lq;
    }
    if (cur_mask != io_mask) relayout(cur_off, io_off, io);
    for (int tid = 0; tid < TH; tid++) for (int j = 0; j < NV; j++) tile[tid + TH * j] = v[tid][j];
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
            mlx::plan_passes(gops, n, tbb, sr, passes, ops, 4, 16, mlx::kRegBits, (trial & 1) ? mlx::kBlockedLookahead : 0);
            for (const mlx::BPass &P : passes)
                if (!mlx::segments_fit_staging(&ops[P.first_op], P.num_ops)) { printf("segment exceeds the op staging buffer\n"); return 1; }
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
    // the staging bound: a long run of non-commuting ops on two qubits (RX,
    // CZ, RY, CZ, ...) fills register segments up to kOpBuf - 1 ops, never
    // more; a segment one op longer, or one that runs past its pass, is refused
    {
        const int n = 6;
        std::vector<moonlab_hip_gate_t> g;
        for (int k = 0; k < 200; k++) {
            g.push_back({7 + k % 2, k % 2, -1, -1, 0.1 + 0.01 * k, 0, 0});
            g.push_back({12, 1, 0, -1, 0, 0, 0});
        }
        std::vector<mlx::GOp> gops; if (!mlx::lower(g.data(), (int)g.size(), n, gops)) { printf("lower failed\n"); return 1; }
        std::vector<mlx::BPass> passes; std::vector<mlx::BOp> ops;
        mlx::plan_passes(gops, n, n, mlx::kSegRegsBatch, passes, ops);
        int longest = 0;
        for (const mlx::BPass &P : passes) {
            if (!mlx::segments_fit_staging(&ops[P.first_op], P.num_ops)) { printf("long run: segment exceeds the op staging buffer\n"); return 1; }
            for (int k = 0; k < P.num_ops; k++) if (ops[P.first_op + k].kind == mlx::BOP_SEG) longest = std::max(longest, (int)ops[P.first_op + k].lq);
        }
        if (longest != mlx::kOpBuf - 1) { printf("longest segment %d ops, expected %d\n", longest, mlx::kOpBuf - 1); return 1; }
        std::vector<mlx::BOp> bad(mlx::kOpBuf + 1);
        for (auto &b : bad) b.kind = mlx::BOP_MAT1;
        bad[0].kind = mlx::BOP_SEG; bad[0].lq = mlx::kOpBuf;
        if (mlx::segments_fit_staging(bad.data(), (int)bad.size())) { printf("oversized segment accepted\n"); return 1; }
        bad[0].lq = mlx::kOpBuf - 1;
        if (!mlx::segments_fit_staging(bad.data(), (int)bad.size())) { printf("largest segment refused\n"); return 1; }
        if (mlx::segments_fit_staging(bad.data(), mlx::kOpBuf - 1)) { printf("segment past its pass accepted\n"); return 1; }
    }
    // register-phase programs: 12..15 qubits, tiles of 12 and 13, both
    // amplitude widths of the LDS swizzle
    int reg_checked = 0, relayouts = 0;
    // trials 16..31 add CNOTs and Toffolis controlled from the top qubits,
    // which are mostly outside the tile set (exchanges on every register
    // value, on the tiles where the controls hold)
    for (int trial = 0; trial < 32; trial++) {
        int n = 12 + trial % 4;
        std::vector<moonlab_hip_gate_t> g;
        // a random product state first, so that every amplitude is nonzero
        // and no misplaced amplitude goes unseen
        for (int q = 0; q < n; q++)
            g.push_back({17, q, -1, -1, (double)(rng() % 1000) / 77.0, (double)(rng() % 1000) / 91.0, (double)(rng() % 1000) / 53.0});
        for (int k = 0; k < 120; k++) {
            int t = 1 + rng() % 24; int q[3]; for (int i = 0; i < 3; i++) { do { q[i] = rng() % n; } while ((i > 0 && q[i] == q[0]) || (i > 1 && q[i] == q[1])); }
            if (trial >= 16 && rng() % 2 && n >= 13) {
                t = rng() % 3 ? 10 : 23;
                q[1] = n - 1 - (int)(rng() % 2);
                q[2] = q[1] == n - 1 ? n - 2 : n - 1;
                do { q[0] = rng() % n; } while (q[0] == q[1] || q[0] == q[2]);
            }
            g.push_back({t, q[0], q[1], q[2], (double)(rng() % 1000) / 77.0, (double)(rng() % 1000) / 91.0, (double)(rng() % 1000) / 53.0});
        }
        std::vector<double> want(2ull << n, 0.0); want[0] = 1.0;
        svo_apply_gates(n, want.data(), (const svo_gate_t *)g.data(), (int)g.size());
        for (int tb : {12, 13}) {
            if (tb > n) continue;
            std::vector<mlx::GOp> gops; if (!mlx::lower(g.data(), (int)g.size(), n, gops)) { printf("lower failed\n"); return 1; }
            std::vector<mlx::BPass> passes; std::vector<mlx::BOp> ops;
            const int rb = 3 + ((trial >> 1) & 1);   // both register widths with both amplitude widths
            // lookahead windows on two trials in three
            mlx::plan_passes(gops, n, tb, mlx::kRegPhases, passes, ops, 3 + (trial & 1), (trial & 1) ? 8 : 16, rb,
                             trial % 3 ? mlx::kBlockedLookahead : 0);
            std::vector<cd> st(1ull << n, cd(0)); st[0] = 1.0;
            for (const mlx::BPass &P : passes) {
                if (P.nbits != tb) { printf("register pass with %d tile bits\n", P.nbits); return 1; }
                int T = 1 << P.nbits; uint64_t ntiles = 1ull << (n - P.nbits);
                for (uint64_t ti = 0; ti < ntiles; ti++) {
                    uint64_t base = ti; for (int b = 0; b < P.nbits; b++) base = ins0(base, P.qubits[b]);
                    std::vector<cd> t(T);
                    auto addr = [&](int l) { uint64_t a = base; for (int b = 0; b < P.nbits; b++) if (l & (1 << b)) a |= 1ull << P.qubits[b]; return a; };
                    for (int l = 0; l < T; l++) t[l] = st[addr(l)];
                    if (trial & 1) reg_tile<8>(t, P.nbits, rb, &ops[P.first_op], P.num_ops, base, relayouts);
                    else reg_tile<16>(t, P.nbits, rb, &ops[P.first_op], P.num_ops, base, relayouts);
                    for (int l = 0; l < T; l++) st[addr(l)] = t[l];
                }
            }
            double err = 0; for (uint64_t i = 0; i < (1ull << n); i++) err = std::max(err, std::abs(st[i] - cd(want[2 * i], want[2 * i + 1])));
            if (err > 1e-12) { printf("REGISTER PHASE MISMATCH n=%d tb=%d err=%g\n", n, tb, err); return 1; }
            reg_checked++;
        }
    }
    // nearest-neighbour layers on 18 qubits (rotations + CNOT brickwork), the
    // circuits whose passes the lookahead windows change the most
    for (int trial = 0; trial < 2; trial++) {
        const int n = 18, tb = 12;
        std::vector<moonlab_hip_gate_t> g;
        for (int q = 0; q < n; q++) g.push_back({4, q, -1, -1, 0, 0, 0});
        for (int d = 0; d < 16; d++) {
            for (int q = 0; q < n; q++)
                g.push_back({7 + (int)(rng() % 3), q, -1, -1, (double)(rng() % 1000) / 77.0, 0, 0});
            for (int q = 0; q + 1 < n; q += 2) g.push_back({10, q + 1, q, -1, 0, 0, 0});
            for (int q = 1; q + 1 < n; q += 2) g.push_back({10, q + 1, q, -1, 0, 0, 0});
        }
        std::vector<double> want(2ull << n, 0.0); want[0] = 1.0;
        svo_apply_gates(n, want.data(), (const svo_gate_t *)g.data(), (int)g.size());
        std::vector<mlx::GOp> gops; if (!mlx::lower(g.data(), (int)g.size(), n, gops)) { printf("lower failed\n"); return 1; }
        std::vector<mlx::BPass> passes, greedy; std::vector<mlx::BOp> ops, gops2;
        mlx::plan_passes(gops, n, tb, mlx::kRegPhases, greedy, gops2, 3, 16, 4, 0);
        mlx::plan_passes(gops, n, tb, mlx::kRegPhases, passes, ops, 3, 16, 4, mlx::kBlockedLookahead);
        if (passes.size() >= greedy.size()) { printf("lookahead planned %zu passes, greedy %zu\n", passes.size(), greedy.size()); return 1; }
        std::vector<cd> st(1ull << n, cd(0)); st[0] = 1.0;
        for (const mlx::BPass &P : passes) {
            int T = 1 << P.nbits; uint64_t ntiles = 1ull << (n - P.nbits);
            for (uint64_t ti = 0; ti < ntiles; ti++) {
                uint64_t base = ti; for (int b = 0; b < P.nbits; b++) base = ins0(base, P.qubits[b]);
                std::vector<cd> t(T);
                auto addr = [&](int l) { uint64_t a = base; for (int b = 0; b < P.nbits; b++) if (l & (1 << b)) a |= 1ull << P.qubits[b]; return a; };
                for (int l = 0; l < T; l++) t[l] = st[addr(l)];
                reg_tile<16>(t, P.nbits, 4, &ops[P.first_op], P.num_ops, base, relayouts);
                for (int l = 0; l < T; l++) st[addr(l)] = t[l];
            }
        }
        double err = 0; for (uint64_t i = 0; i < (1ull << n); i++) err = std::max(err, std::abs(st[i] - cd(want[2 * i], want[2 * i + 1])));
        if (err > 1e-12) { printf("LOOKAHEAD PLAN MISMATCH err=%g\n", err); return 1; }
        reg_checked++;
    }
    printf("ok: %d plans checked, %d segment executions, %d register-phase plans, %d layout changes, "
           "%ld paired rotations\n", checked, segs_seen, reg_checked, relayouts, g_rot2_ops);
}
