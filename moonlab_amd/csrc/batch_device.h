// batch_device.h -- program format of the persistent batch kernel, shared by
// the host planner (batch.cpp) and the kernel (batch.hip).
#pragma once

#include <stdint.h>

namespace mlx {

// A batch op acts on a tile of 2^tile_bits amplitudes held in LDS.  Qubits of
// the pass's tile set are "local" (an LDS index bit); the others are fixed
// for the whole tile and read from the tile's base index.
enum BOpKind : int32_t {
    BOP_MAT1 = 0,   // 2x2 m[8] on local target lq, where controls are set
    BOP_DIAG = 1,   // diagonal: amplitude *= d[bit(qb) << 1 | bit(qa)], d = m[0..7]
    BOP_SWAP = 2,   // exchange local bits lq / lq2 where controls are set
    BOP_XPERM = 4,  // controlled X: exchange the pair across local bit lq where controls are set
    BOP_SEG = 3,    // header: the next lq ops act only on the lq2 (2 or 3) local bits packed
                    // 5 bits each (ascending) in cl, apart from diagonal and control bits;
                    // they run on 2^lq2 amplitudes held in registers per thread
};

// Register bits of a segment: the single-state tile pass holds 8 amplitudes
// per thread; the batch kernel, which also carries measurement state, 4.
// Op descriptors the batch kernel stages in LDS at a time; a register
// segment (header + its ops) never exceeds it.
constexpr int kOpBuf = 32;
constexpr int kSegRegsPass = 2;
constexpr int kSegRegsBatch = 2;

struct BOp {
    int32_t  kind;
    int32_t  lq;        // local target (MAT1, SWAP); qubit a of DIAG (local, or -1 - global)
    int32_t  lq2;       // second local target (SWAP); qubit b of DIAG (as lq; INT32_MIN = none)
    int32_t  pad;
    uint32_t cl;        // control bits that are local (mask over tile bits)
    uint32_t pad2;
    uint64_t cg;        // control bits that are global (mask over the full index)
    double   m[8];
};

// One cache-blocked pass: tile set (ascending qubits) and its op range.
struct BPass {
    int32_t first_op;
    int32_t num_ops;
    int32_t nbits;          // tile bits of this pass (== instance tile bits)
    int32_t kind;           // 0 = gate pass, 1 = measurement pass
    int8_t  qubits[24];     // tile set, ascending
};

// One batched circuit.
struct BInst {
    int32_t  n;
    int32_t  prec;          // 0 = complex128, 1 = complex64
    int32_t  tile_bits;     // min(n, max tile bits)
    int32_t  first_pass;
    int32_t  num_passes;    // gate passes
    int32_t  num_mpasses;   // measurement passes (follow the gate passes)
    int32_t  flags;         // BF_* below
    int32_t  num_shots;
    int64_t  obs_offset;    // into the FP64 observable buffer
    int64_t  draw_offset;   // into the uniform-draw buffer
    int64_t  out_offset;    // into the outcome buffer
};

enum : int32_t {
    BF_Z = 1,       // P(q = 1) for every qubit and the norm
    BF_ZZ = 2,      // <Z_q Z_q+1> for q = 0..n-2
    BF_X = 4,       // <X_q> for every qubit
};

// Observable block per instance, FP64: [norm, P1[n], ZZ[n-1], X[n]].
inline int64_t obs_size(int n) { return 1 + n + (n > 1 ? n - 1 : 0) + n; }

constexpr int kBatchMaxTileBits = 12;
constexpr int kBatchThreads = 1024;
constexpr int kBatchMaxQubits = 24;

}  // namespace mlx
