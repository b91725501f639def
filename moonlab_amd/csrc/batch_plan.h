// batch_plan.h -- gate-list lowering and cache-blocked pass planning shared by
// the persistent batch engine and the blocked single-state path.
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/moonlab_hip_statevec.h"
#include "batch_device.h"

namespace mlx {

// An op before tile assignment: global qubit indices.
struct GOp {
    int      kind;          // BOP_*
    int      t1, t2;        // targets (MAT1: t1; SWAP: t1, t2); DIAG: qa = t1, qb = t2 (-1 = none)
    uint64_t ctrl;          // control qubits (global bits)
    double   m[8];
};

// Gate list -> ops, folding each qubit's run of single-qubit gates into one
// FP64 2x2.  False on an invalid gate or qubit.
bool lower(const moonlab_hip_gate_t *g, int ng, int n, std::vector<GOp> &out);

// Pack ops into passes over tile sets of `tb` qubits (n <= tb: one resident
// pass over all n qubits) that always hold the lowest `anchor_bits` qubits,
// convert qubits to tile-local indices and group each pass's ops into
// register segments of `seg_regs` bits (BOP_SEG), or, for seg_regs =
// kRegPhases, into a register-phase program (BOP_PHASE) of `reg_bits`
// register bits per phase for amplitudes of `amp_bytes` bytes (16
// complex128, 8 complex64; the phase headers carry LDS slot offsets).
// lookahead > 0 picks each pass's tile set among windows of neighbouring
// qubits seeded by that many pending ops (three windows each), and each
// phase's register set among sets seeded by as many pending ops (the set
// taking the most ops); 0 grows both greedily in circuit order.
constexpr int kRegPhases = -1;
constexpr int kBlockedLookahead = 8;
void plan_passes(const std::vector<GOp> &gops, int n, int tb, int seg_regs, std::vector<BPass> &passes,
                 std::vector<BOp> &ops, int anchor_bits = 4, int amp_bytes = 16, int reg_bits = kRegBits,
                 int lookahead = 0);

// The batch kernel stages a pass's op descriptors in LDS kOpBuf at a time
// and applies a register segment only when its header and all of its ops sit
// in one staged chunk (run_pass_ops_staged, batch.hip).  True when every
// segment of ops[first, first + count) fits: header + ops <= kOpBuf and no
// segment runs past the pass's op range.  A plan that fails this would make
// the staged loop refuse to advance.
bool segments_fit_staging(const BOp *ops, int count);

// Pauli string (s[q] in "IXYZ" acts on qubit q, exactly n characters) as
// masks; false otherwise.
bool parse_pauli(const char *s, int n, PauliT &out);

// Final states of a batch run written to device memory: item i's state (its
// precision's amplitudes) lands at amplitude offset offsets[i] of d_buf, or
// nowhere when offsets[i] < 0.
struct BatchStateOut {
    void          *d_buf;
    const int64_t *offsets;
};

}  // namespace mlx
