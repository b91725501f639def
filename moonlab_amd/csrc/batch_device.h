// batch_device.h -- program format of the persistent batch kernel, shared by
// the host planner (batch.cpp) and the kernel (batch.hip).
#pragma once

#include <hip/hip_runtime.h>
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
    BOP_PHASE = 5,  // register-phase program header (k_reg_pass): the next lq ops run on the
                    // layout whose kRegBits register bits are the tile bits of mask lq2
                    // (positions packed 5 bits each, ascending, in cl); m[] holds the
                    // 2^kRegBits uint32 byte offsets of the register values' LDS slots,
                    // lds_slot(reg_off(j)) x amplitude bytes; op fields below
};

// Register-phase programs (k_reg_pass): each thread holds 2^kRegBits tile
// amplitudes, value j at tile index tloc | roff(j), roff(j) = sum over i of
// bit i of j << r_i for the phase's register bits r_0 < .. < r_3.  Ops inside
// a phase reuse BOp with these meanings:
// Every tile qubit an op touches (target, controls, diagonal qubits) is a
// register bit of its phase, so an op acts alike on every thread:
//   pad = the kernel arm (RegArm below) the host picked for the op's kind,
//     form and register positions; cg = global controls (the op is skipped
//     on tiles where one is 0).
//   MAT1 / XPERM: lq = register position of the target (0..3); pad2 = 16-bit
//     mask of the j whose register bits satisfy the local controls (cl = 0).
//   DIAG: lq / lq2 as in the LDS engine (global codes give the factor index
//     bits kg); cl = kmap, bits 2j..2j+1 = the local qubits' bits of value j
//     (factor index kg | those bits); pad2 as above, less the values whose
//     factor is exactly 1 when no qubit is global.
//   SWAP does not occur (it is lowered to three exchanges).
// A program is planned for one register width rb (3 or 4 bits: 8 or 16
// values per thread); positions and masks above are over rb bits.
constexpr int kRegBits = 4;   // the widest register set

// Arms of the register-phase interpreter (P, Q, C = register positions).
enum RegArm : int32_t {
    ARM_MAT1_MASKED = 0,    // + P: 2x2 on the values jm selects
    ARM_MAT1 = 4,           // + P: general 2x2 on every value
    ARM_MAT1_REAL = 8,      // + P: real 2x2 (RY, H)
    ARM_MAT1_RX = 12,       // + P: real diagonal, imaginary off-diagonal (RX)
    ARM_XPERM = 16,         // + P: exchange across P on the values jm selects
    ARM_CX = 20,            // + 4 P + C: CNOT, target P, control C
    ARM_DIAG1 = 36,         // + P: diag(d0, d1) on P, no controls
    ARM_DIAG1_ONLY1 = 40,   // + P: diag(1, d1) on P, no controls
    ARM_DIAG11 = 44,        // + 4 P + Q (P < Q): diag(1, 1, 1, d3), no controls
    ARM_DIAG = 60,          // any diagonal (kmap, kg, jm)
    ARM_ROT2 = 64,          // + 4 rot2_pair(P, Q) + 2 F + G: a form-F 2x2 on P (m[0..3]) and a
                            // form-G 2x2 on Q (m[4..7]), P < Q, forms 0 real (RY, H) and 1 RX
};

// index of the register-bit pair P < Q (of 4 bits) among the ARM_ROT2 arms
__host__ __device__ constexpr int rot2_pair(int p, int q)
{
    return p == 0 ? q - 1 : p == 1 ? q + 1 : 5;
}
// The entries of m[] an arm reads: MAT1, MAT1_MASKED and DIAG all eight as
// built; MAT1_REAL m[0..3] = (u00, u01, u10, u11); MAT1_RX m[0..3] = (Re u00,
// Im u01, Im u10, Re u11); DIAG1 / DIAG1_ONLY1 m[0..3] = (d0, d1); DIAG11
// m[0..1] = d3; ROT2 the first rotation's four in m[0..3], the second's in
// m[4..7] (lq = P, lq2 = Q).

// Register bits of a phase, ascending tile bits.
struct RegLayout {
    int r[kRegBits];
};

__host__ __device__ inline RegLayout unpack_layout(uint32_t packed, int rb = kRegBits)
{
    RegLayout L{};
    for (int i = 0; i < rb; i++) L.r[i] = (int)((packed >> (5 * i)) & 31u);
    return L;
}

// tile-index bits of register value j
__host__ __device__ inline int reg_off(const RegLayout &L, int j, int rb = kRegBits)
{
    int o = 0;
    for (int i = 0; i < rb; i++)
        if ((j >> i) & 1) o |= 1 << L.r[i];
    return o;
}

// tile index of a thread's register value 0: the thread id spread over the
// non-register bits
__host__ __device__ inline int thread_base(const RegLayout &L, int tid, int rb = kRegBits)
{
    int x = tid;
    for (int i = 0; i < rb; i++) x = ((x >> L.r[i]) << (L.r[i] + 1)) | (x & ((1 << L.r[i]) - 1));
    return x;
}

// LDS slot of tile index x for BYTES-byte amplitudes.  The low bits are
// XORed with the folded higher bits, so lanes whose indices differ only in
// higher bits spread over the banks (8 slots of 16 bytes, or 16 of 8 bytes,
// per 128-byte row).  Linear over GF(2): slot(a | b) = slot(a) ^ slot(b) for
// disjoint a, b.
template <int BYTES>
__host__ __device__ inline int lds_slot(int x)
{
    if (BYTES == 16) return x ^ ((x >> 3 ^ x >> 6 ^ x >> 9 ^ x >> 12) & 7);
    return x ^ ((x >> 4 ^ x >> 8 ^ x >> 12) & 15);
}

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
    int64_t  pauli_offset;  // first Pauli term (PauliT array); results follow the standard observables
    int64_t  state_offset;  // amplitude offset of the final state in the state output buffer, -1: none
    int32_t  num_pauli;     // Pauli terms to measure
    int32_t  pad_;
};

// One Pauli string P = i^ny X^x Z^z (Y_q = i X_q Z_q): P|i> = i^ny
// (-1)^popcount(i & z) |i ^ x>, so <P> = sum_i Re[conj(a_(i^x)) a_i i^ny
// (-1)^popcount(i & z)].
struct PauliT {
    uint64_t x;
    uint64_t z;
    int32_t  ny;            // number of Y factors
    int32_t  pad;
};

enum : int32_t {
    BF_Z = 1,       // P(q = 1) for every qubit and the norm
    BF_ZZ = 2,      // <Z_q Z_q+1> for q = 0..n-2
    BF_X = 4,       // <X_q> for every qubit
};

// Observable block per instance, FP64: [norm, P1[n], ZZ[n-1], X[n]].
__host__ __device__ inline int64_t obs_size(int n) { return 1 + n + (n > 1 ? n - 1 : 0) + n; }

constexpr int kBatchMaxTileBits = 12;
constexpr int kBatchThreads = 1024;
constexpr int kBatchMaxQubits = 24;

}  // namespace mlx
