// batch.cpp -- host planner and C ABI of the persistent batch engine.
//
// Planning per circuit (independent of the rest of the batch, so a
// circuit's results are the same alone and in any mix):
//   1. single-qubit gates fold into one FP64 2x2 per qubit until a
//      multi-qubit gate touches that qubit (the reference's run-length 1q
//      fusion, src/optimization/fusion/fusion.c:247-292); a folded diagonal
//      becomes a DIAG op, anything else a MAT1;
//   2. every gate lowers to LDS ops: controlled gates to MAT1 with control
//      masks, CZ / CPHASE / CRZ to DIAG, SWAP / Fredkin to SWAP;
//   3. above the tile size, ops are packed into cache-blocked passes: a
//      pass's tile set holds at most `tile_bits` qubits and must contain the
//      target(s) of each non-diagonal op; controls and diagonal ops may sit
//      outside it (they read fixed bits of the tile's base index).
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <array>
#include <chrono>
#include <climits>
#include <complex>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/moonlab_mi355x.h"
#include "batch_device.h"
#include "batch_plan.h"
#include "fusion.h"
#include "sv_internal.h"

namespace mlx {

hipError_t launch_tile_pass(void *amps, int prec, int n, const BPass &P, const BOp *d_ops, int cus, int threads,
                            int tile_bits, hipStream_t st);
hipError_t launch_reg_pass(void *amps, int prec, int n, const BPass &P, const BOp *d_ops, int reg_bits, hipStream_t st);

hipError_t launch_batch(const BInst *insts, const int32_t *order, int ninst, int *counter, const BPass *passes,
                        const BOp *ops, double2 *scratch, uint64_t slot_amps, double *tile_scratch,
                        uint64_t tile_slot, const double *draws, double *obs, uint64_t *outcomes,
                        const PauliT *terms, void *state_out, int grid, hipStream_t st);

// A Pauli string over n qubits (s[q] in "IXYZ" acts on qubit q, strlen == n,
// pauli_term_t of vqe.h:101-105) as masks; false on any other character or
// length.
bool parse_pauli(const char *s, int n, PauliT &out)
{
    if (!s) return false;
    out = PauliT{};
    int len = 0;
    for (; s[len]; len++) {
        if (len >= n) return false;
        switch (s[len]) {
        case 'I': break;
        case 'X': out.x |= 1ull << len; break;
        case 'Z': out.z |= 1ull << len; break;
        case 'Y':
            out.x |= 1ull << len;
            out.z |= 1ull << len;
            out.ny++;
            break;
        default: return false;
        }
    }
    return len == n;
}

namespace {

bool is_identity(const cd u[4])
{
    return u[0] == cd(1.0) && u[1] == cd(0.0) && u[2] == cd(0.0) && u[3] == cd(1.0);
}

void emit_1q(std::vector<GOp> &out, int q, const cd u[4])
{
    if (is_identity(u)) return;
    GOp o{};
    if (u[1] == cd(0.0) && u[2] == cd(0.0)) {
        o.kind = BOP_DIAG;
        o.t1 = q;
        o.t2 = -1;
        o.m[0] = u[0].real(); o.m[1] = u[0].imag();
        o.m[2] = u[3].real(); o.m[3] = u[3].imag();
        o.m[4] = 1.0; o.m[6] = 1.0;
    } else {
        o.kind = BOP_MAT1;
        o.t1 = q;
        o.t2 = -1;
        for (int i = 0; i < 4; i++) {
            o.m[2 * i] = u[i].real();
            o.m[2 * i + 1] = u[i].imag();
        }
    }
    out.push_back(o);
}

void mat_op(std::vector<GOp> &out, int t, uint64_t ctrl, cd u00, cd u01, cd u10, cd u11)
{
    GOp o{};
    o.kind = BOP_MAT1;
    o.t1 = t;
    o.t2 = -1;
    o.ctrl = ctrl;
    const cd u[4] = {u00, u01, u10, u11};
    for (int i = 0; i < 4; i++) {
        o.m[2 * i] = u[i].real();
        o.m[2 * i + 1] = u[i].imag();
    }
    out.push_back(o);
}

void diag2_op(std::vector<GOp> &out, int qa, int qb, uint64_t ctrl, cd d0, cd d1, cd d2, cd d3)
{
    GOp o{};
    o.kind = BOP_DIAG;
    o.t1 = qa;
    o.t2 = qb;
    o.ctrl = ctrl;
    const cd d[4] = {d0, d1, d2, d3};
    for (int i = 0; i < 4; i++) {
        o.m[2 * i] = d[i].real();
        o.m[2 * i + 1] = d[i].imag();
    }
    out.push_back(o);
}

}  // namespace

// Steps 1 and 2: gate list -> global ops.
// X on t where every control in ctrl is 1: a pure exchange of pair values.
void xperm_op(std::vector<GOp> &out, int t, uint64_t ctrl)
{
    GOp o{};
    o.kind = BOP_XPERM;
    o.t1 = t;
    o.t2 = -1;
    o.ctrl = ctrl;
    out.push_back(o);
}

bool lower(const moonlab_hip_gate_t *g, int ng, int n, std::vector<GOp> &out)
{
    std::vector<std::array<cd, 4>> pend((size_t)n);
    std::vector<bool> has((size_t)n, false);
    auto flush = [&](int q) {
        if (has[(size_t)q]) emit_1q(out, q, pend[(size_t)q].data());
        has[(size_t)q] = false;
    };
    const cd I(0.0, 1.0);
    for (int i = 0; i < ng; i++) {
        const moonlab_hip_gate_t &x = g[i];
        if (x.type < 0 || x.type > MOONLAB_HIP_GATE_FREDKIN) return false;
        int qs[3];
        const int nq = gate_qubits(x, qs);
        for (int j = 0; j < nq; j++)
            if (qs[j] < 0 || qs[j] >= n || (j && qs[j] == qs[j - 1])) return false;
        cd u[4];
        if (gate_matrix_1q(x, u)) {
            const int q = x.target;
            if (!has[(size_t)q]) {
                pend[(size_t)q] = {u[0], u[1], u[2], u[3]};
                has[(size_t)q] = true;
            } else {   // new = u * pending
                const auto &p = pend[(size_t)q];
                pend[(size_t)q] = {u[0] * p[0] + u[1] * p[2], u[0] * p[1] + u[1] * p[3],
                                   u[2] * p[0] + u[3] * p[2], u[2] * p[1] + u[3] * p[3]};
            }
            continue;
        }
        for (int j = 0; j < nq; j++) flush(qs[j]);
        const int c = x.control, t = x.target, e = x.extra;
        const double co = cos(x.theta / 2.0), si = sin(x.theta / 2.0);
        switch (x.type) {
        case MOONLAB_HIP_GATE_CNOT: xperm_op(out, t, 1ull << c); break;
        case MOONLAB_HIP_GATE_CY: mat_op(out, t, 1ull << c, 0, I, -I, 0); break;
        case MOONLAB_HIP_GATE_CY_TRUE: mat_op(out, t, 1ull << c, 0, -I, I, 0); break;
        case MOONLAB_HIP_GATE_CRX: mat_op(out, t, 1ull << c, co, cd(0, -si), cd(0, -si), co); break;
        case MOONLAB_HIP_GATE_CRY: mat_op(out, t, 1ull << c, co, -si, si, co); break;
        case MOONLAB_HIP_GATE_TOFFOLI: xperm_op(out, t, (1ull << c) | (1ull << e)); break;
        case MOONLAB_HIP_GATE_CZ: diag2_op(out, c, t, 0, 1, 1, 1, -1); break;
        case MOONLAB_HIP_GATE_CPHASE: diag2_op(out, c, t, 0, 1, 1, 1, cd(cos(x.theta), sin(x.theta))); break;
        case MOONLAB_HIP_GATE_CRZ:   // index = bit(c) << 1 | bit(t)
            diag2_op(out, t, c, 0, 1, 1, cd(co, -si), cd(co, si));
            break;
        case MOONLAB_HIP_GATE_SWAP:
        case MOONLAB_HIP_GATE_FREDKIN: {
            GOp o{};
            o.kind = BOP_SWAP;
            o.t1 = x.type == MOONLAB_HIP_GATE_SWAP ? c : t;
            o.t2 = x.type == MOONLAB_HIP_GATE_SWAP ? t : e;
            o.ctrl = x.type == MOONLAB_HIP_GATE_FREDKIN ? 1ull << c : 0;
            out.push_back(o);
            break;
        }
        default: return false;
        }
    }
    for (int q = 0; q < n; q++) flush(q);
    return true;
}

namespace {

uint64_t needs_local(const GOp &o)
{
    if (o.kind == BOP_MAT1 || o.kind == BOP_XPERM) return 1ull << o.t1;
    if (o.kind == BOP_SWAP) return (1ull << o.t1) | (1ull << o.t2);
    return 0;
}

// Split a pass's tile-local ops into register segments (BOP_SEG): runs of at
// least two ops whose non-diagonal targets fit kSegRegs tile bits, padded
// with the highest unused tile bits (consecutive threads then touch
// consecutive LDS slots).  Shorter runs stay single ops.
void segment_ops(const std::vector<BOp> &in, int nbits, int seg_regs, std::vector<BOp> &out)
{
    if (seg_regs < 2 || nbits < seg_regs) {
        out.insert(out.end(), in.begin(), in.end());
        return;
    }
    auto need = [](const BOp &b) -> uint32_t {
        if (b.kind == BOP_MAT1 || b.kind == BOP_XPERM) return 1u << b.lq;
        if (b.kind == BOP_SWAP) return (1u << b.lq) | (1u << b.lq2);
        return 0u;
    };
    size_t i = 0;
    while (i < in.size()) {
        uint32_t regs = 0;
        size_t j = i;
        while (j < in.size() && (int)(j - i) < kOpBuf - 1 && __builtin_popcount(regs | need(in[j])) <= seg_regs)
            regs |= need(in[j++]);
        if (j - i < 2) {
            out.push_back(in[i++]);
            continue;
        }
        for (int q = nbits - 1; q >= 0 && __builtin_popcount(regs) < seg_regs; q--) regs |= 1u << q;
        BOp h{};
        h.kind = BOP_SEG;
        h.lq = (int32_t)(j - i);
        h.lq2 = seg_regs;
        for (int q = 0, k = 0; q < nbits; q++)
            if (regs & (1u << q)) h.cl |= (uint32_t)q << (5 * k++);
        out.push_back(h);
        out.insert(out.end(), in.begin() + (long)i, in.begin() + (long)j);
        i = j;
    }
}

// A pass's tile-local ops as a register-phase program for k_reg_pass.  Each
// phase has reg_bits register bits and takes, in order, every op whose
// non-diagonal target fits its register set and that shares no tile bit with
// an op deferred from the phase (ops on disjoint qubits commute); the
// deferred ops, still in order, make up the rest.  Every tile qubit an op
// touches goes in the register set, so ops never depend on thread bits.  The
// first phase tries the IO layout (the top reg_bits tile bits), which needs
// no LDS exchange; a register set is filled up with the highest free bits.
// SWAP lowers to three exchanges, X_b[a] X_a[b] X_b[a], its controls kept.
void plan_reg_phases(const std::vector<BOp> &in0, int nbits, int amp_bytes, int reg_bits, std::vector<BOp> &out,
                     int lookahead)
{
    const int nv = 1 << reg_bits;
    std::vector<BOp> in;
    in.reserve(in0.size());
    for (const BOp &b : in0) {
        if (b.kind != BOP_SWAP) {
            in.push_back(b);
            continue;
        }
        BOp x{};
        x.kind = BOP_XPERM;
        x.lq2 = INT32_MIN;
        x.cg = b.cg;
        BOp x1 = x, x2 = x;
        x1.lq = b.lq2;
        x1.cl = b.cl | (1u << b.lq);
        x2.lq = b.lq;
        x2.cl = b.cl | (1u << b.lq2);
        in.push_back(x1);
        in.push_back(x2);
        in.push_back(x1);
    }
    // every tile bit an op touches must be a register bit (at most 3: a
    // target and two controls, or two diagonal qubits)
    auto need = [](const BOp &b) -> uint32_t {
        uint32_t m = b.cl;
        if (b.kind == BOP_MAT1 || b.kind == BOP_XPERM) m |= 1u << b.lq;
        if (b.kind == BOP_DIAG) {
            if (b.lq >= 0) m |= 1u << b.lq;
            if (b.lq2 >= 0) m |= 1u << b.lq2;
        }
        return m;
    };
    auto touch = need;
    const uint32_t io = ((1u << reg_bits) - 1) << (nbits - reg_bits);
    std::vector<const BOp *> rest, later, cur;
    rest.reserve(in.size());
    for (const BOp &b : in) rest.push_back(&b);
    bool first = true;
    while (!rest.empty()) {
        uint32_t S = 0;
        auto scan = [&](bool fixed) {
            uint32_t blocked = 0;
            cur.clear();
            later.clear();
            for (const BOp *o : rest) {
                const uint32_t t = touch(*o), nd = need(*o);
                const bool fits = fixed ? (nd & ~S) == 0 : __builtin_popcount(S | nd) <= reg_bits;
                if (!(t & blocked) && fits) {
                    S |= nd;
                    cur.push_back(o);
                } else {
                    blocked |= t;
                    later.push_back(o);
                }
            }
        };
        if (first) {
            S = io;
            scan(true);
            if (cur.empty()) {
                S = 0;
                scan(false);
            }
        } else {
            scan(false);
        }
        // lookahead: register sets seeded by pending ops (the op's bits, then
        // the bits above its lowest one, then below) compete with the set
        // above; the one taking the most ops wins (the IO layout of the first
        // phase, which needs no layout change, on ties)
        if (lookahead > 0) {
            uint32_t bestS = S;
            std::vector<const BOp *> best_cur = cur, best_later = later;
            std::vector<uint32_t> tried;
            for (size_t i = 0; i < rest.size() && (int)tried.size() < lookahead; i++) {
                const uint32_t nd = need(*rest[i]);
                if (!nd || __builtin_popcount(nd) > reg_bits) continue;
                uint32_t S0 = nd;
                const int q = __builtin_ctz(nd);
                for (int k = q; k < nbits && __builtin_popcount(S0) < reg_bits; k++) S0 |= 1u << k;
                for (int k = q - 1; k >= 0 && __builtin_popcount(S0) < reg_bits; k--) S0 |= 1u << k;
                if (std::find(tried.begin(), tried.end(), S0) != tried.end()) continue;
                tried.push_back(S0);
                S = S0;
                scan(true);
                if (cur.size() > best_cur.size()) {
                    bestS = S;
                    best_cur.swap(cur);
                    best_later.swap(later);
                }
            }
            S = bestS;
            cur.swap(best_cur);
            later.swap(best_later);
        }
        first = false;
        for (int q = nbits - 1; q >= 0 && __builtin_popcount(S) < reg_bits; q--) S |= 1u << q;
        int r[kRegBits];
        for (int q = 0, k = 0; q < nbits; q++)
            if (S & (1u << q)) r[k++] = q;
        auto roff = [&](int j) {
            uint32_t o = 0;
            for (int i = 0; i < reg_bits; i++)
                if ((j >> i) & 1) o |= 1u << r[i];
            return o;
        };
        auto pos = [&](int q) {
            for (int i = 0; i < reg_bits; i++)
                if (r[i] == q) return i;
            return 15;
        };
        BOp h{};
        h.kind = BOP_PHASE;
        h.lq = (int32_t)cur.size();
        h.lq2 = (int32_t)S;
        for (int i = 0; i < reg_bits; i++) h.cl |= (uint32_t)r[i] << (5 * i);
        // byte offsets of the register values' LDS slots in this layout
        uint32_t slot_off[1 << kRegBits] = {};
        static_assert(sizeof slot_off <= sizeof h.m, "slot offsets fit the header's m[]");
        for (int j = 0; j < nv; j++)
            slot_off[j] = (uint32_t)((amp_bytes == 8 ? lds_slot<8>((int)roff(j)) : lds_slot<16>((int)roff(j))) *
                                     amp_bytes);
        memcpy(h.m, slot_off, sizeof slot_off);
        const size_t hpos = out.size();
        out.push_back(h);
        for (const BOp *o : cur) {
            BOp b = *o;
            const uint32_t creg = o->cl & S;
            uint32_t jm = 0;
            for (int j = 0; j < nv; j++)
                if ((roff(j) & creg) == creg) jm |= 1u << j;
            b.cl = 0;   // every local control is a register bit
            b.pad2 = jm;
            b.pad = 0;
            const bool all = jm == (nv == 32 ? ~0u : (1u << nv) - 1);
            auto one = [](const double *d) { return d[0] == 1.0 && d[1] == 0.0; };
            if (o->kind == BOP_DIAG) {
                // kmap: the local qubits' bits of each register value; with
                // no global qubit the factor is known per value, and values
                // whose factor is exactly 1 are left out of jm
                const int apos = o->lq >= 0 ? pos(o->lq) : -1, bpos = o->lq2 >= 0 ? pos(o->lq2) : -1;
                const bool global = (o->lq < 0 && o->lq != INT32_MIN) || (o->lq2 < 0 && o->lq2 != INT32_MIN);
                uint32_t kmap = 0;
                for (int j = 0; j < nv; j++) {
                    const uint32_t k = (apos >= 0 ? (j >> apos) & 1 : 0) | (bpos >= 0 ? ((j >> bpos) & 1) << 1 : 0);
                    kmap |= k << (2 * j);
                    if (!global && one(o->m + 2 * k)) b.pad2 &= ~(1u << j);
                }
                b.cl = kmap;
                b.pad = ARM_DIAG;
                if (!global && all && apos >= 0 && bpos < 0) {
                    b.pad = (one(o->m) ? ARM_DIAG1_ONLY1 : ARM_DIAG1) + apos;
                } else if (!global && all && apos >= 0 && bpos >= 0 && one(o->m) && one(o->m + 2) &&
                           one(o->m + 4)) {
                    b.pad = ARM_DIAG11 + 4 * std::min(apos, bpos) + std::max(apos, bpos);
                }
            } else {
                b.lq = pos(o->lq);   // the scan placed the target in S
                if (o->kind == BOP_MAT1) {
                    const double *m = o->m;
                    if (!all) b.pad = ARM_MAT1_MASKED + b.lq;
                    else if (m[1] == 0.0 && m[3] == 0.0 && m[5] == 0.0 && m[7] == 0.0) b.pad = ARM_MAT1_REAL + b.lq;
                    else if (m[1] == 0.0 && m[7] == 0.0 && m[2] == 0.0 && m[4] == 0.0) b.pad = ARM_MAT1_RX + b.lq;
                    else b.pad = ARM_MAT1 + b.lq;
                } else if (__builtin_popcount(creg) == 1) {   // XPERM with one control
                    b.pad = ARM_CX + 4 * b.lq + pos(__builtin_ctz(creg));
                } else {
                    b.pad = ARM_XPERM + b.lq;
                }
            }
            // the entries the arm reads go to m[0..3] (the register pass
            // prefetches only those; the general arms read all eight)
            double *m = b.m;
            if (b.pad >= ARM_MAT1_REAL && b.pad < ARM_MAT1_REAL + 4) {
                const double r[4] = {m[0], m[2], m[4], m[6]};   // u00, u01, u10, u11 (real)
                memcpy(m, r, sizeof r);
            } else if (b.pad >= ARM_MAT1_RX && b.pad < ARM_MAT1_RX + 4) {
                const double r[4] = {m[0], m[3], m[5], m[6]};   // Re u00, Im u01, Im u10, Re u11
                memcpy(m, r, sizeof r);
            } else if (b.pad >= ARM_DIAG11 && b.pad < ARM_DIAG) {
                const double r[2] = {m[6], m[7]};               // d3
                memcpy(m, r, sizeof r);
            }
            // two neighbouring real / RX-form rotations on different register
            // bits (they commute) become one ARM_ROT2 op: one dispatch, both
            // 2x2s in one pass over the register values
            auto rot_form = [](const BOp &x) {
                if (x.cg != 0) return -1;
                if (x.pad >= ARM_MAT1_REAL && x.pad < ARM_MAT1_REAL + 4) return 0;
                if (x.pad >= ARM_MAT1_RX && x.pad < ARM_MAT1_RX + 4) return 1;
                return -1;
            };
            BOp &prev = out.back();
            const int fb = rot_form(b);
            if (out.size() > hpos + 1 && fb >= 0 && rot_form(prev) >= 0 && prev.lq != b.lq) {
                const BOp &lo = prev.lq < b.lq ? prev : b, &hi = prev.lq < b.lq ? b : prev;
                BOp pr = lo;
                pr.pad = ARM_ROT2 + 4 * rot2_pair(lo.lq, hi.lq) + 2 * rot_form(lo) + rot_form(hi);
                pr.lq2 = hi.lq;
                memcpy(pr.m + 4, hi.m, 4 * sizeof(double));
                prev = pr;
                continue;
            }
            out.push_back(b);
        }
        out[hpos].lq = (int32_t)(out.size() - hpos - 1);
        rest.swap(later);
    }
}

// Every qubit an op reads or writes (targets, diagonal qubits, controls).
uint64_t touches(const GOp &o)
{
    uint64_t m = o.ctrl;
    if (o.t1 >= 0) m |= 1ull << o.t1;
    if (o.t2 >= 0) m |= 1ull << o.t2;
    return m;
}

}  // namespace

bool segments_fit_staging(const BOp *ops, int count)
{
    for (int k = 0; k < count;) {
        if (ops[k].kind == BOP_SEG) {
            const int lq = ops[k].lq;
            if (lq < 1 || 1 + lq > kOpBuf || k + 1 + lq > count) return false;
            k += 1 + lq;
        } else {
            k++;
        }
    }
    return true;
}

// Step 3: pack ops into passes and convert qubits to tile-local indices.
void plan_passes(const std::vector<GOp> &gops, int n, int tb, int seg_regs, std::vector<BPass> &passes,
                 std::vector<BOp> &ops, int anchor_bits, int amp_bytes, int reg_bits, int lookahead)
{
    const bool resident = n <= tb;
    // low qubits kept in every tile (runs of 2^anchor_bits contiguous
    // amplitudes); at most tb - 2, so any op (two non-diagonal targets at
    // most) fits an empty pass and every scan below schedules at least one op
    const int anchor = resident ? n : (tb - 2 < anchor_bits ? (tb - 2 > 0 ? tb - 2 : 0) : anchor_bits);
    uint64_t anchor_mask = anchor >= 64 ? ~0ull : ((1ull << anchor) - 1);
    std::vector<const GOp *> cur;
    uint64_t S = anchor_mask;
    auto close = [&]() {
        if (cur.empty()) return;
        // fill the tile set to exactly tb qubits with the lowest unused ones
        for (int q = 0; q < n && __builtin_popcountll(S) < (resident ? n : tb); q++) S |= 1ull << q;
        BPass P{};
        P.first_op = (int)ops.size();
        P.num_ops = (int)cur.size();
        P.nbits = 0;
        int local[64];
        for (int q = 0; q < 64; q++) local[q] = -1;
        for (int q = 0; q < n; q++)
            if (S & (1ull << q)) {
                local[q] = P.nbits;
                P.qubits[P.nbits++] = (int8_t)q;
            }
        std::vector<BOp> pass_ops;
        pass_ops.reserve(cur.size());
        for (const GOp *o : cur) {
            BOp b{};
            b.kind = o->kind;
            for (int q = 0; q < n; q++) {
                if (!(o->ctrl & (1ull << q))) continue;
                if (local[q] >= 0) b.cl |= 1u << local[q];
                else b.cg |= 1ull << q;
            }
            if (o->kind == BOP_DIAG) {
                b.lq = local[o->t1] >= 0 ? local[o->t1] : -1 - o->t1;
                b.lq2 = o->t2 < 0 ? INT32_MIN : (local[o->t2] >= 0 ? local[o->t2] : -1 - o->t2);
            } else {
                b.lq = local[o->t1];
                b.lq2 = o->kind == BOP_SWAP ? local[o->t2] : INT32_MIN;
            }
            memcpy(b.m, o->m, sizeof b.m);
            pass_ops.push_back(b);
        }
        if (seg_regs == kRegPhases) plan_reg_phases(pass_ops, P.nbits, amp_bytes, reg_bits, ops, lookahead);
        else segment_ops(pass_ops, P.nbits, seg_regs, ops);
        P.num_ops = (int)ops.size() - P.first_op;
        passes.push_back(P);
        cur.clear();
        S = anchor_mask;
    };
    if (resident) {
        for (const GOp &o : gops) cur.push_back(&o);
        close();
        return;
    }
    // Ops on disjoint qubits commute, so a pass takes, in circuit order, every
    // op whose non-diagonal targets fit the tile set and that shares no qubit
    // with an op already deferred from this pass; the deferred ops, still in
    // circuit order, make up the rest of the circuit for the next pass.
    //
    // With lookahead > 0 the tile set is chosen before the scan: besides the
    // set the scan grows from the anchor alone, windows seeded by the first
    // `lookahead` pending ops with a target outside the anchor are scanned
    // (the anchor, the seed's target, and the free qubits in a contiguous
    // run starting at, centred on or ending at that target), and the set
    // that takes the most ops wins (the first such, in seed order).  A window
    // of neighbouring qubits lets one pass run many layers of a nearest-
    // neighbour circuit inside it.
    std::vector<const GOp *> rest, later;
    rest.reserve(gops.size());
    for (const GOp &o : gops) rest.push_back(&o);
    // ops a tile set starting from S0 takes among rest[0, lim), growing while
    // it has free bits; with take / defer, the full split of rest
    auto scan = [&](uint64_t S0, size_t lim, std::vector<const GOp *> *take, std::vector<const GOp *> *defer) {
        uint64_t Sx = S0, blocked = 0;
        size_t taken = 0;
        for (size_t i = 0; i < lim; i++) {
            const GOp *o = rest[i];
            const uint64_t t = touches(*o), need = needs_local(*o);
            if (!(t & blocked) && __builtin_popcountll(Sx | need) <= tb) {
                Sx |= need;
                taken++;
                if (take) take->push_back(o);
            } else {
                blocked |= t;
                if (defer) defer->push_back(o);
            }
        }
        if (take) S = Sx;
        return taken;
    };
    // candidates are compared on the first kLookaheadOps pending ops, which
    // bounds planning time on very long circuits
    constexpr size_t kLookaheadOps = 4096;
    std::vector<uint64_t> tried;
    const int free_bits = tb - __builtin_popcountll(anchor_mask);
    while (!rest.empty()) {
        const size_t lim = rest.size() < kLookaheadOps ? rest.size() : kLookaheadOps;
        uint64_t bestS0 = anchor_mask;
        size_t best = lookahead > 0 ? scan(anchor_mask, lim, nullptr, nullptr) : 0;
        tried.clear();
        int seeds = 0;
        for (size_t i = 0; i < rest.size() && seeds < lookahead; i++) {
            const uint64_t nd = needs_local(*rest[i]) & ~anchor_mask;
            if (!nd) continue;
            seeds++;
            const int q = __builtin_ctzll(nd);
            // windows starting at, centred on and ending at the seed's target
            for (const int lo0 : {q, q - free_bits / 2, q - free_bits + 1}) {
                uint64_t S0 = anchor_mask | nd;
                const int lo = lo0 < 0 ? 0 : lo0;
                for (int k = lo; k < n && __builtin_popcountll(S0) < tb; k++) S0 |= 1ull << k;
                for (int k = lo - 1; k >= 0 && __builtin_popcountll(S0) < tb; k--) S0 |= 1ull << k;
                if (std::find(tried.begin(), tried.end(), S0) != tried.end()) continue;
                tried.push_back(S0);
                const size_t c = scan(S0, lim, nullptr, nullptr);
                if (c > best) {
                    best = c;
                    bestS0 = S0;
                }
            }
        }
        later.clear();
        scan(bestS0, rest.size(), &cur, &later);
        close();
        rest.swap(later);
    }
}

namespace {

// Tile bits and engine of the single-state blocked path: the register-phase
// engine for states of at least 12 qubits (16 register values per thread;
// 12-qubit tiles for complex128, 13-qubit tiles for complex64, unless the
// tuning fixes them), the LDS op engine otherwise or when selected (11- or
// 12-qubit tiles).  *anchor: the low qubits every tile
// keeps, 128 contiguous bytes (3 qubits at complex128, 4 at complex64; the
// LDS engine keeps 4); *reg_bits: register bits of the register engine.
int blocked_shape(int n, int prec, const GateTuning &t, bool *reg, int *anchor, int *reg_bits)
{
    *reg = t.blocked_engine != 1 && n >= 12;
    *reg_bits = kRegBits;
    if (*reg) {
        const bool c64 = prec == MOONLAB_HIP_C64;
        *anchor = c64 ? 4 : 3;
        *reg_bits = t.blocked_engine == 3 ? 3 : 4;
        const int want = t.blocked_tile_bits ? t.blocked_tile_bits : (c64 ? 13 : 12);
        return (want >= 13 && n >= 13) ? 13 : 12;
    }
    *anchor = 4;
    const int b = t.blocked_tile_bits == 0 || t.blocked_tile_bits > 12 ? 12 : t.blocked_tile_bits;
    return n < b ? n : b;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T> size_t align_up(size_t x) { return (x + alignof(T) - 1) / alignof(T) * alignof(T); }

}  // namespace
}  // namespace mlx

using namespace mlx;

struct moonlab_mi355x_batch {
    hipStream_t stream = nullptr;
    int         wg_per_cu = 2;
    int         tile_bits = kBatchMaxTileBits;
    int         num_cus = 0;
    // one device arena for the program (insts, order, passes, ops, draws) and one for results
    void       *d_prog = nullptr;
    size_t      prog_cap = 0;
    void       *h_prog = nullptr;   // pinned staging
    size_t      h_prog_cap = 0;
    void       *d_res = nullptr;
    size_t      res_cap = 0;
    void       *h_res = nullptr;
    size_t      h_res_cap = 0;
    int        *d_counter = nullptr;
    void       *d_scratch = nullptr;
    size_t      scratch_cap = 0;
    hipEvent_t  e0 = nullptr, e1 = nullptr;
    // HIP graph of the last batch shape (see moonlab_mi355x_batch_set_graphs)
    bool            use_graphs = true;
    hipGraphExec_t  gexec = nullptr;
    std::vector<uint64_t> gkey;       // shape of gexec
    std::vector<uint64_t> last_key;   // shape of the previous run
    // state and overlap buffers of the geometric-tensor calls (vqe.cpp)
    void       *d_states = nullptr;
    size_t      states_cap = 0;
};

namespace {

int grow_dev(void **p, size_t *cap, size_t need)
{
    if (need <= *cap) return 0;
    if (*p) hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t sz = need + need / 4 + 4096;
    if (hipMalloc(p, sz) != hipSuccess) return -1;
    *cap = sz;
    return 0;
}

int grow_host(void **p, size_t *cap, size_t need)
{
    if (need <= *cap) return 0;
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t sz = need + need / 4 + 4096;
    if (hipHostMalloc(p, sz, hipHostMallocDefault) != hipSuccess) return -1;
    *cap = sz;
    return 0;
}

// fn(t) for t = 0..nthreads-1 on host threads (t = 0 and any that could not
// be started on the calling thread)
template <typename F>
void on_threads(int nthreads, F &&fn)
{
    std::vector<std::thread> th;
    int started = 0;
    try {
        for (int t = 1; t < nthreads; t++, started++) th.emplace_back(fn, t);
    } catch (const std::system_error &) {
        // no more host threads: the calling thread runs what was not handed out
    }
    fn(0);
    for (int t = 1 + started; t < nthreads; t++) fn(t);
    for (std::thread &x : th) x.join();
}

int host_threads(int work_items)
{
    const unsigned hw = std::thread::hardware_concurrency();
    return work_items < 64 ? 1 : (int)std::min<unsigned>(16u, hw ? hw : 1u);
}

int run_batch(moonlab_mi355x_batch_t *b, moonlab_mi355x_batch_item_t *items, int count,
              moonlab_mi355x_batch_stats_t *stats, const BatchStateOut *state_out = nullptr)
{
    const double t0 = now_s();
    if (!b || count < 0 || (count > 0 && !items)) return MOONLAB_SCHED_BAD_ARG;
    if (count == 0) return MOONLAB_SCHED_OK;
    const int tbmax = b->tile_bits;
    std::vector<BInst> insts((size_t)count);
    std::vector<double> draws;
    std::vector<PauliT> pauli;
    for (int i = 0; i < count; i++)
        if (items[i].num_pauli_terms < 0 || (items[i].num_pauli_terms > 0 && !items[i].pauli_terms))
            return MOONLAB_SCHED_BAD_ARG;
    std::vector<std::pair<double, int>> cost((size_t)count);
    int64_t obs_total = 0, out_total = 0;
    int max_big = 0;

    // Lowering and pass planning are independent per circuit: chunks of
    // circuits are planned on host threads into chunk-local pass / op lists,
    // which are then concatenated in circuit order with their offsets shifted.
    struct Chunk {
        int                lo = 0, hi = 0;
        std::vector<BPass> passes;
        std::vector<BOp>   ops;
        bool               bad = false;
        bool               internal = false;
    };
    const int nthreads = host_threads(count);
    std::vector<Chunk> chunks((size_t)nthreads);
    auto plan_chunk = [&](Chunk &c) {
        std::vector<GOp> gops;
        for (int i = c.lo; i < c.hi; i++) {
            const moonlab_mi355x_batch_item_t &it = items[i];
            const int n = it.num_qubits;
            if (n < 1 || n > kBatchMaxQubits || (it.num_gates > 0 && !it.gates) || it.num_gates < 0 ||
                it.num_shots < 0 || (it.num_shots > 0 && (!it.draws || !it.outcomes)) ||
                (it.precision != MOONLAB_HIP_C128 && it.precision != MOONLAB_HIP_C64)) {
                c.bad = true;
                return;
            }
            gops.clear();
            if (!lower(it.gates, it.num_gates, n, gops)) {
                c.bad = true;
                return;
            }
            BInst &I = insts[(size_t)i];
            I.n = n;
            I.prec = it.precision;
            I.tile_bits = n < tbmax ? n : tbmax;
            I.first_pass = (int)c.passes.size();   // chunk-local (passes and ops) until placed
            plan_passes(gops, n, I.tile_bits, kSegRegsBatch, c.passes, c.ops);
            I.num_passes = (int)c.passes.size() - I.first_pass;
            for (int p = I.first_pass; p < I.first_pass + I.num_passes; p++) {
                const BPass &P = c.passes[(size_t)p];
                if (!segments_fit_staging(c.ops.data() + P.first_op, P.num_ops)) {
                    c.internal = true;   // a planner bug: the staged op loop could not advance
                    return;
                }
            }
            // longest first: sweeps x amplitudes (resident circuits: ops x tile)
            const double sweeps = n > I.tile_bits ? (double)(I.num_passes + 2) * 4.0 : 0.2 * (double)gops.size();
            cost[(size_t)i] = {-(sweeps * (double)(1ull << n)), i};
        }
    };
    for (int t = 0; t < nthreads; t++) {
        chunks[(size_t)t].lo = (int)((int64_t)count * t / nthreads);
        chunks[(size_t)t].hi = (int)((int64_t)count * (t + 1) / nthreads);
    }
    on_threads(nthreads, [&](int t) { plan_chunk(chunks[(size_t)t]); });
    // each chunk's passes and ops land at their offsets in the program arena
    // (copied below by the planning threads, straight into pinned memory)
    std::vector<size_t> pass_base((size_t)nthreads), op_base((size_t)nthreads);
    size_t total_passes = 0, total_ops = 0;
    for (int t = 0; t < nthreads; t++) {
        const Chunk &c = chunks[(size_t)t];
        if (c.bad) return MOONLAB_SCHED_BAD_ARG;
        if (c.internal) return MOONLAB_SCHED_INTERNAL;
        pass_base[(size_t)t] = total_passes;
        op_base[(size_t)t] = total_ops;
        total_passes += c.passes.size();
        total_ops += c.ops.size();
        for (int i = c.lo; i < c.hi; i++) insts[(size_t)i].first_pass += (int)pass_base[(size_t)t];
    }
    // observables, draws and outcome slots in circuit order
    for (int i = 0; i < count; i++) {
        const moonlab_mi355x_batch_item_t &it = items[i];
        BInst &I = insts[(size_t)i];
        const int n = I.n;
        I.flags = 0;
        if (it.observables & MOONLAB_MI355X_OBS_Z) I.flags |= BF_Z;
        if (it.observables & MOONLAB_MI355X_OBS_ZZ) I.flags |= BF_ZZ;
        if (it.observables & MOONLAB_MI355X_OBS_X) I.flags |= BF_X;
        I.num_shots = it.num_shots;
        I.obs_offset = obs_total;
        I.num_pauli = it.num_pauli_terms > 0 ? it.num_pauli_terms : 0;
        I.pauli_offset = (int64_t)pauli.size();
        for (int k = 0; k < I.num_pauli; k++) {
            PauliT pt;
            if (!parse_pauli(it.pauli_terms[k].pauli_string, n, pt)) return MOONLAB_SCHED_BAD_ARG;
            pauli.push_back(pt);
        }
        I.state_offset = state_out ? state_out->offsets[i] : -1;
        obs_total += obs_size(n) + I.num_pauli;
        I.draw_offset = (int64_t)draws.size();
        for (int sh = 0; sh < it.num_shots; sh++) draws.push_back(it.draws[sh]);
        I.out_offset = out_total;
        out_total += it.num_shots;
        if (n > I.tile_bits && n > max_big) max_big = n;
    }
    std::stable_sort(cost.begin(), cost.end());
    std::vector<int32_t> order((size_t)count);
    for (int i = 0; i < count; i++) order[(size_t)i] = cost[(size_t)i].second;

    // grid: as many workgroups as fit, bounded by the HBM slots large circuits need
    int grid = b->num_cus * b->wg_per_cu;
    if (grid > count) grid = count;
    uint64_t slot_amps = 0, tile_slot = 0;
    if (max_big) {
        slot_amps = 1ull << max_big;
        tile_slot = 2ull << (max_big - (tbmax < max_big ? tbmax : max_big));
        const uint64_t per_wg = slot_amps * 16 + tile_slot * 8;
        const uint64_t budget = 16ull << 30;
        const uint64_t fit = budget / per_wg;
        if ((uint64_t)grid > fit) grid = (int)(fit ? fit : 1);
    } else {
        slot_amps = 1;
        tile_slot = 2;
    }
    const size_t scratch_bytes = (size_t)grid * (slot_amps * 16 + tile_slot * 8) + 256;
    if (grow_dev(&b->d_scratch, &b->scratch_cap, scratch_bytes)) return MOONLAB_SCHED_OOM;

    // program arena layout
    size_t off = 0;
    auto place = [&](size_t bytes) { const size_t o = off; off = (off + bytes + 255) / 256 * 256; return o; };
    const size_t o_inst = place(insts.size() * sizeof(BInst));
    const size_t o_order = place(order.size() * sizeof(int32_t));
    const size_t o_pass = place((total_passes + 1) * sizeof(BPass));
    const size_t o_ops = place((total_ops + 1) * sizeof(BOp));
    const size_t o_draw = place((draws.size() + 1) * sizeof(double));
    const size_t o_pauli = place((pauli.size() + 1) * sizeof(PauliT));
    const size_t prog_bytes = off;
    off = 0;
    const size_t r_obs = place((size_t)(obs_total + 1) * sizeof(double));
    const size_t r_out = place((size_t)(out_total + 1) * sizeof(uint64_t));
    const size_t res_bytes = off;
    if (grow_dev(&b->d_prog, &b->prog_cap, prog_bytes) || grow_host(&b->h_prog, &b->h_prog_cap, prog_bytes) ||
        grow_dev(&b->d_res, &b->res_cap, res_bytes) || grow_host(&b->h_res, &b->h_res_cap, res_bytes))
        return MOONLAB_SCHED_OOM;
    char *hp = (char *)b->h_prog;
    memcpy(hp + o_inst, insts.data(), insts.size() * sizeof(BInst));
    memcpy(hp + o_order, order.data(), order.size() * sizeof(int32_t));
    {
        BPass *hpass = (BPass *)(hp + o_pass);
        BOp *hops = (BOp *)(hp + o_ops);
        auto place_chunk = [&](int t) {
            const Chunk &c = chunks[(size_t)t];
            BPass *dst = hpass + pass_base[(size_t)t];
            for (size_t k = 0; k < c.passes.size(); k++) {
                dst[k] = c.passes[k];
                dst[k].first_op += (int32_t)op_base[(size_t)t];
            }
            if (!c.ops.empty()) memcpy(hops + op_base[(size_t)t], c.ops.data(), c.ops.size() * sizeof(BOp));
        };
        on_threads(nthreads, place_chunk);
    }
    if (!draws.empty()) memcpy(hp + o_draw, draws.data(), draws.size() * sizeof(double));
    if (!pauli.empty()) memcpy(hp + o_pauli, pauli.data(), pauli.size() * sizeof(PauliT));
    void *d_state = state_out ? state_out->d_buf : nullptr;
    char *dp = (char *)b->d_prog;
    char *dr = (char *)b->d_res;
    double2 *scratch = (double2 *)b->d_scratch;
    double *tile_scratch = (double *)((char *)b->d_scratch + (size_t)grid * slot_amps * 16);

    // the device step: upload, queue reset + persistent launch, download
    auto step = [&]() -> hipError_t {
        hipError_t r = hipMemcpyAsync(dp, hp, prog_bytes, hipMemcpyHostToDevice, b->stream);
        if (r == hipSuccess)
            r = launch_batch((const BInst *)(dp + o_inst), (const int32_t *)(dp + o_order), count, b->d_counter,
                             (const BPass *)(dp + o_pass), (const BOp *)(dp + o_ops), scratch, slot_amps,
                             tile_scratch, tile_slot, (const double *)(dp + o_draw), (double *)(dr + r_obs),
                             (uint64_t *)(dr + r_out), (const PauliT *)(dp + o_pauli), d_state, grid, b->stream);
        if (r == hipSuccess) r = hipMemcpyAsync(b->h_res, dr, res_bytes, hipMemcpyDeviceToHost, b->stream);
        return r;
    };
    // everything the captured nodes bake in: buffers, sizes, offsets, launch shape
    const std::vector<uint64_t> key = {(uint64_t)(uintptr_t)dp, (uint64_t)(uintptr_t)dr,
                                       (uint64_t)(uintptr_t)hp, (uint64_t)(uintptr_t)b->h_res,
                                       (uint64_t)(uintptr_t)b->d_scratch, prog_bytes, res_bytes, o_inst, o_order,
                                       o_pass, o_ops, o_draw, r_obs, r_out, (uint64_t)count, (uint64_t)grid,
                                       slot_amps, tile_slot, o_pauli, (uint64_t)(uintptr_t)d_state};
    // A shape seen for the first time runs eagerly; its second run is captured
    // and launched as a graph; later runs of it replay that graph.
    bool replayed = false;
    hipError_t e = hipEventRecord(b->e0, b->stream);
    if (e == hipSuccess && b->use_graphs && b->gexec && b->gkey == key) {
        replayed = true;
        e = hipGraphLaunch(b->gexec, b->stream);
    } else if (e == hipSuccess && b->use_graphs && b->last_key == key) {
        if (b->gexec) hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
        b->gkey.clear();
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
        if (ce == hipSuccess) {
            const hipError_t se = step();
            ce = hipStreamEndCapture(b->stream, &g);
            if (se != hipSuccess) ce = se;
        }
        if (ce == hipSuccess) ce = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
        if (g) hipGraphDestroy(g);
        if (ce == hipSuccess) {
            b->gkey = key;
            e = hipGraphLaunch(b->gexec, b->stream);
        } else {
            // this runtime cannot capture the step: run it eagerly from now on
            fprintf(stderr, "moonlab-mi355x batch: graph capture failed (%s); launching directly\n",
                    hipGetErrorString(ce));
            (void)hipGetLastError();
            if (b->gexec) hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
            b->use_graphs = false;
            e = step();
        }
    } else if (e == hipSuccess) {
        e = step();
    }
    b->last_key = key;
    if (e == hipSuccess) e = hipEventRecord(b->e1, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        fprintf(stderr, "moonlab-mi355x batch: %s\n", hipGetErrorString(e));
        return MOONLAB_SCHED_INTERNAL;
    }
    const double *obs = (const double *)((char *)b->h_res + r_obs);
    const uint64_t *outc = (const uint64_t *)((char *)b->h_res + r_out);
    for (int i = 0; i < count; i++) {
        moonlab_mi355x_batch_item_t &it = items[i];
        const BInst &I = insts[(size_t)i];
        const int n = I.n;
        const double *o = obs + I.obs_offset;
        if (it.norm) *it.norm = o[0];
        if (it.expect_z)
            for (int q = 0; q < n; q++) it.expect_z[q] = 1.0 - 2.0 * o[1 + q];
        if (it.zz_chain)
            for (int q = 0; q + 1 < n; q++) it.zz_chain[q] = o[1 + n + q];
        if (it.expect_x)
            for (int q = 0; q < n; q++) it.expect_x[q] = o[1 + n + (n > 1 ? n - 1 : 0) + q];
        if (it.num_shots > 0) memcpy(it.outcomes, outc + I.out_offset, (size_t)it.num_shots * sizeof(uint64_t));
        if (I.num_pauli > 0) {
            const double *pv = o + obs_size(n);
            if (it.pauli_expect) memcpy(it.pauli_expect, pv, (size_t)I.num_pauli * sizeof(double));
            if (it.pauli_energy) {   // sum_k c_k <P_k> in term order (vqe.c:1326-1342)
                double e = 0.0;
                for (int k = 0; k < I.num_pauli; k++) e += it.pauli_terms[k].coefficient * pv[k];
                *it.pauli_energy = e;
            }
        }
    }
    if (stats) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, b->e0, b->e1);
        stats->device_seconds = ms * 1e-3;
        stats->wall_seconds = now_s() - t0;
        stats->workgroups = grid;
        stats->tile_bits = tbmax;
        int big_passes = 0;
        for (const BInst &I : insts)
            if (I.n > I.tile_bits) big_passes += I.num_passes;
        stats->total_passes = big_passes;
        stats->total_ops = (int)total_ops;
        stats->graph_replayed = replayed ? 1 : 0;
        stats->reserved = 0;
    }
    return MOONLAB_SCHED_OK;
}

}  // namespace

// Entry points of the batch engine for the variational calls (vqe.cpp).
int batch_run_states(moonlab_mi355x_batch_t *b, moonlab_mi355x_batch_item_t *items, int count,
                     moonlab_mi355x_batch_stats_t *stats, const BatchStateOut *out)
{
    return run_batch(b, items, count, stats, out);
}

void *batch_state_buffer(moonlab_mi355x_batch_t *b, size_t bytes)
{
    if (grow_dev(&b->d_states, &b->states_cap, bytes)) return nullptr;
    return b->d_states;
}

hipStream_t batch_stream(moonlab_mi355x_batch_t *b) { return b->stream; }

extern "C" {

moonlab_hip_status_t moonlab_hip_apply_gates_blocked(moonlab_hip_state_t *s, const moonlab_hip_gate_t *g,
                                                     int ng)
{
    if (!s || ng < 0 || (ng > 0 && !g)) return MOONLAB_HIP_ERR_BAD_ARG;
    if (ng == 0) return MOONLAB_HIP_OK;
    const int n = (int)moonlab_hip_state_n_qubits(s);
    std::vector<GOp> gops;
    if (!lower(g, ng, n, gops)) return MOONLAB_HIP_ERR_INVALID_QUBIT;
    std::vector<BPass> passes;
    std::vector<BOp> ops;
    const GateTuning tune = gate_tuning();
    const int prec = moonlab_hip_state_precision(s);
    bool reg = false;
    int anchor = 4, reg_bits = kRegBits;
    const int tb = blocked_shape(n, prec, tune, &reg, &anchor, &reg_bits);
    plan_passes(gops, n, tb, reg ? kRegPhases : kSegRegsPass, passes, ops, anchor,
                prec == MOONLAB_HIP_C64 ? 8 : 16, reg_bits, tune.blocked_lookahead);
    if (ops.empty()) return MOONLAB_HIP_OK;
    hipStream_t st = (hipStream_t)moonlab_hip_state_stream(s);
    static std::atomic<int> cached_cus{0};   // shared by every launching thread
    int cus = cached_cus.load(std::memory_order_relaxed);
    if (!cus) {
        int dev = 0;
        hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached_cus.store(cus, std::memory_order_relaxed);
    }
    // k_reg_pass fetches the head of the op after each op it runs (the next
    // phase's header after a phase's last op): one spare op after the last
    // keeps that fetch inside the buffer
    if (reg) ops.push_back(BOp{});
    BOp *d_ops = nullptr;
    hipError_t e = hipMallocAsync((void **)&d_ops, ops.size() * sizeof(BOp), st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ops, ops.data(), ops.size() * sizeof(BOp), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);   // `ops` is a local vector
    void *amps = moonlab_hip_state_device_ptr(s);
    const int lds_tile = tune.blocked_tile_bits == 0 || tune.blocked_tile_bits > 12 ? 12 : tune.blocked_tile_bits;
    for (size_t i = 0; e == hipSuccess && i < passes.size(); i++)
        e = reg ? launch_reg_pass(amps, prec, n, passes[i], d_ops, reg_bits, st)
                : launch_tile_pass(amps, prec, n, passes[i], d_ops, cus, tune.blocked_threads, lds_tile, st);
    if (d_ops) hipFreeAsync(d_ops, st);
    if (e == hipErrorOutOfMemory) return MOONLAB_HIP_ERR_OUT_OF_MEMORY;
    return e == hipSuccess ? MOONLAB_HIP_OK : MOONLAB_HIP_ERR_DRIVER;
}

int moonlab_hip_blocked_pass_count(const moonlab_hip_gate_t *g, int ng, int n, int precision)
{
    if (ng < 0 || (ng > 0 && !g) || n < 1 || n > 40) return MOONLAB_HIP_ERR_BAD_ARG;
    if (precision != MOONLAB_HIP_C128 && precision != MOONLAB_HIP_C64) return MOONLAB_HIP_ERR_BAD_ARG;
    std::vector<GOp> gops;
    if (!lower(g, ng, n, gops)) return MOONLAB_HIP_ERR_INVALID_QUBIT;
    std::vector<BPass> passes;
    std::vector<BOp> ops;
    bool reg = false;
    int anchor = 4, reg_bits = kRegBits;
    const int tb = blocked_shape(n, precision, gate_tuning(), &reg, &anchor, &reg_bits);
    plan_passes(gops, n, tb, reg ? kRegPhases : kSegRegsPass, passes, ops, anchor,
                precision == MOONLAB_HIP_C64 ? 8 : 16, reg_bits, gate_tuning().blocked_lookahead);
    return (int)passes.size();
}

moonlab_mi355x_batch_t *moonlab_mi355x_batch_create(void)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        fprintf(stderr, "moonlab-mi355x batch: no HIP device visible\n");
        return nullptr;
    }
    moonlab_mi355x_batch_t *b = new (std::nothrow) moonlab_mi355x_batch_t();
    if (!b) return nullptr;
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (b->num_cus <= 0) b->num_cus = 256;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&b->d_counter, 256) != hipSuccess || hipEventCreate(&b->e0) != hipSuccess ||
        hipEventCreate(&b->e1) != hipSuccess) {
        moonlab_mi355x_batch_free(b);
        return nullptr;
    }
    return b;
}

int moonlab_mi355x_batch_set_graphs(moonlab_mi355x_batch_t *b, int enable)
{
    if (!b) return MOONLAB_SCHED_BAD_ARG;
    b->use_graphs = enable != 0;
    if (!b->use_graphs && b->gexec) {
        hipStreamSynchronize(b->stream);
        hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
        b->gkey.clear();
    }
    return MOONLAB_SCHED_OK;
}

void moonlab_mi355x_batch_free(moonlab_mi355x_batch_t *b)
{
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    hipFree(b->d_prog);
    hipFree(b->d_res);
    hipFree(b->d_counter);
    hipFree(b->d_scratch);
    hipFree(b->d_states);
    if (b->h_prog) hipHostFree(b->h_prog);
    if (b->h_res) hipHostFree(b->h_res);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->e0) hipEventDestroy(b->e0);
    if (b->e1) hipEventDestroy(b->e1);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

int moonlab_mi355x_batch_set_config(moonlab_mi355x_batch_t *b, int wg_per_cu, int tile_bits)
{
    if (!b) return MOONLAB_SCHED_BAD_ARG;
    if (wg_per_cu) {
        if (wg_per_cu < 1 || wg_per_cu > 2) return MOONLAB_SCHED_BAD_ARG;
        b->wg_per_cu = wg_per_cu;
    }
    if (tile_bits) {
        if (tile_bits < 8 || tile_bits > kBatchMaxTileBits) return MOONLAB_SCHED_BAD_ARG;
        b->tile_bits = tile_bits;
    }
    return MOONLAB_SCHED_OK;
}

int moonlab_mi355x_batch_run(moonlab_mi355x_batch_t *b, moonlab_mi355x_batch_item_t *items, int count,
                             moonlab_mi355x_batch_stats_t *stats)
{
    return run_batch(b, items, count, stats);
}

int moonlab_mi355x_batch_autotune(moonlab_mi355x_batch_t *b, moonlab_mi355x_batch_item_t *items, int count,
                                  moonlab_mi355x_batch_stats_t *stats)
{
    if (!b) return MOONLAB_SCHED_BAD_ARG;
    double best = 1e300;
    int best_wg = b->wg_per_cu, best_tb = b->tile_bits;
    for (int wg = 1; wg <= 2; wg++)
        for (int tb = 10; tb <= kBatchMaxTileBits; tb++) {
            b->wg_per_cu = wg;
            b->tile_bits = tb;
            moonlab_mi355x_batch_stats_t st{};
            const int rc = run_batch(b, items, count, &st);
            if (rc) return rc;
            if (st.device_seconds < best) {
                best = st.device_seconds;
                best_wg = wg;
                best_tb = tb;
            }
        }
    b->wg_per_cu = best_wg;
    b->tile_bits = best_tb;
    return run_batch(b, items, count, stats);
}

int moonlab_mi355x_hea_tfim_energies(moonlab_mi355x_batch_t *b, const moonlab_mi355x_hea_t *inst, int count,
                                     int precision, double *energies, moonlab_mi355x_batch_stats_t *stats)
{
    if (!b || count < 0 || (count > 0 && (!inst || !energies))) return MOONLAB_SCHED_BAD_ARG;
    std::vector<std::vector<moonlab_hip_gate_t>> gl((size_t)count);
    std::vector<moonlab_mi355x_batch_item_t> items((size_t)count);
    std::vector<double> zz, xx;
    size_t zoff = 0, xoff = 0;
    for (int i = 0; i < count; i++) {
        zoff += (size_t)(inst[i].num_qubits > 1 ? inst[i].num_qubits - 1 : 1);
        xoff += (size_t)inst[i].num_qubits;
    }
    zz.assign(zoff, 0.0);
    xx.assign(xoff, 0.0);
    std::vector<size_t> zo((size_t)count), xo((size_t)count);
    zoff = xoff = 0;
    for (int i = 0; i < count; i++) {
        const int n = inst[i].num_qubits, L = inst[i].num_layers;
        if (n < 1 || n > kBatchMaxQubits || L < 0 || (L > 0 && !inst[i].angles)) return MOONLAB_SCHED_BAD_ARG;
        zo[(size_t)i] = zoff;
        xo[(size_t)i] = xoff;
        zoff += (size_t)(n > 1 ? n - 1 : 1);
        xoff += (size_t)n;
    }
    // the ansatz gate lists (vqe.c:764-781 layer order), built on host threads
    const int nthreads = host_threads(count);
    on_threads(nthreads, [&](int t) {
        const int lo = (int)((int64_t)count * t / nthreads), hi = (int)((int64_t)count * (t + 1) / nthreads);
        for (int i = lo; i < hi; i++) {
            const int n = inst[i].num_qubits, L = inst[i].num_layers;
            auto &g = gl[(size_t)i];
            g.reserve((size_t)L * (size_t)(3 * n - 1));
            for (int l = 0; l < L; l++) {
                for (int q = 0; q < n; q++) {
                    const double *a = inst[i].angles + ((size_t)l * n + q) * 2;
                    g.push_back(moonlab_hip_gate_t{MOONLAB_HIP_GATE_RY, q, -1, -1, a[0], 0, 0});
                    g.push_back(moonlab_hip_gate_t{MOONLAB_HIP_GATE_RZ, q, -1, -1, a[1], 0, 0});
                }
                for (int q = 0; q + 1 < n; q++)
                    g.push_back(moonlab_hip_gate_t{MOONLAB_HIP_GATE_CZ, q + 1, q, -1, 0, 0, 0});
            }
            moonlab_mi355x_batch_item_t &it = items[(size_t)i];
            memset(&it, 0, sizeof it);
            it.num_qubits = n;
            it.precision = precision;
            it.gates = g.data();
            it.num_gates = (int)g.size();
            it.observables = MOONLAB_MI355X_OBS_ZZ | MOONLAB_MI355X_OBS_X;
            it.zz_chain = zz.data() + zo[(size_t)i];
            it.expect_x = xx.data() + xo[(size_t)i];
        }
    });
    const int rc = run_batch(b, items.data(), count, stats);
    if (rc) return rc;
    // E = -sum_q <Z_q Z_q+1> - h sum_q <X_q>, summed in qubit order
    for (int i = 0; i < count; i++) {
        const int n = inst[i].num_qubits;
        double e = 0.0;
        for (int q = 0; q + 1 < n; q++) e -= items[(size_t)i].zz_chain[q];
        double sx = 0.0;
        for (int q = 0; q < n; q++) sx += items[(size_t)i].expect_x[q];
        energies[i] = e - inst[i].h * sx;
    }
    return MOONLAB_SCHED_OK;
}

int moonlab_mi355x_hea_tfim_gradients(moonlab_mi355x_batch_t *b, const moonlab_mi355x_hea_t *inst, int count,
                                      int precision, double *energies, double *gradients,
                                      moonlab_mi355x_batch_stats_t *stats)
{
    if (!b || count < 0 || (count > 0 && (!inst || !energies || !gradients))) return MOONLAB_SCHED_BAD_ARG;
    // instance i expands to its base point and +-pi/2 shifts of each of its
    // P = 2 n L angles, all evaluated in one batch run
    size_t total_inst = 0, total_angles = 0;
    for (int i = 0; i < count; i++) {
        const int n = inst[i].num_qubits, L = inst[i].num_layers;
        if (n < 1 || n > kBatchMaxQubits || L < 0 || (L > 0 && !inst[i].angles)) return MOONLAB_SCHED_BAD_ARG;
        const size_t P = (size_t)2 * n * L;
        total_inst += 1 + 2 * P;
        total_angles += (1 + 2 * P) * P;
    }
    std::vector<double> ang(total_angles);
    std::vector<moonlab_mi355x_hea_t> ex(total_inst);
    size_t ai = 0, ei = 0;
    for (int i = 0; i < count; i++) {
        const size_t P = (size_t)2 * inst[i].num_qubits * inst[i].num_layers;
        for (size_t k = 0; k < 1 + 2 * P; k++) {
            double *a = ang.data() + ai;
            if (P) memcpy(a, inst[i].angles, P * sizeof(double));
            if (k > 0) a[(k - 1) / 2] += ((k - 1) % 2 == 0 ? 1.0 : -1.0) * (M_PI / 2.0);
            ex[ei] = inst[i];
            ex[ei].angles = a;
            ai += P;
            ei++;
        }
    }
    std::vector<double> e(total_inst);
    const int rc = moonlab_mi355x_hea_tfim_energies(b, ex.data(), (int)total_inst, precision, e.data(), stats);
    if (rc) return rc;
    // dE/dtheta_k = [E(theta + pi/2 e_k) - E(theta - pi/2 e_k)] / 2, exact for
    // RY / RZ (two-eigenvalue generators, vqe.c:2094-2100, 2119-2130)
    ei = 0;
    size_t gi = 0;
    for (int i = 0; i < count; i++) {
        const size_t P = (size_t)2 * inst[i].num_qubits * inst[i].num_layers;
        energies[i] = e[ei];
        for (size_t k = 0; k < P; k++) gradients[gi + k] = 0.5 * (e[ei + 1 + 2 * k] - e[ei + 2 + 2 * k]);
        ei += 1 + 2 * P;
        gi += P;
    }
    return MOONLAB_SCHED_OK;
}

}  // extern "C"
