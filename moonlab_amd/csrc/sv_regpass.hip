// sv_regpass.hip -- register-phase passes of the cache-blocked single-state
// path (moonlab_hip_apply_gates_blocked, states of 12 qubits or more).
//
// A workgroup of 2^(TB - RB) threads streams tiles of 2^TB amplitudes and
// each thread keeps 2^RB of them in registers (RB = 3 or 4 register bits).  The ops of a phase
// act on the register values in place; LDS only carries a change of which
// tile bits are register bits (one write, one barrier, one read per change).
// The host planner (batch.cpp: plan_passes, plan_reg_phases) packs a circuit
// into passes and each pass into phases; the op format is in batch_device.h.
//
// Reference counterpart: the per-gate loops of src/quantum/gates.c:123-1048,
// run here as cache-blocked passes (the reference's CPU engine sweeps the
// whole state once per gate).
#include <hip/hip_runtime.h>

#include <limits.h>

#include <atomic>

#include "batch_device.h"
#include "sv_device.h"
#include "sv_internal.h"

namespace mlx {
namespace {

// tile_base for a tile set of exactly TB qubits
template <int TB>
__device__ __forceinline__ uint64_t tile_base_c(uint64_t t, const BPass &p)
{
#pragma unroll
    for (int b = 0; b < TB; b++) t = ins0(t, p.qubits[b]);
    return t;
}

// A layout change: every thread writes its register values at their tile
// indices (LDS slots), then reads the values of the new layout.  from / to:
// the byte offsets of the 2^RB register values' LDS slots in each layout
// (slot(reg_off(L, j)) x amplitude bytes; the planner stores them in the
// phase header), so the only per-value address work is one XOR with the
// thread's own slot.
template <typename R, int RB>
__device__ __forceinline__ void relayout(cplx_t<R> *t, cplx_t<R> (&v)[1 << RB], int &tloc, const uint32_t (&from)[1 << RB],
                                         const uint32_t (&to)[1 << RB], const RegLayout &tol)
{
    constexpr int NV = 1 << RB;
    constexpr int B = (int)sizeof(cplx_t<R>);
    char *tb = (char *)t;
    const uint32_t s0 = (uint32_t)lds_slot<B>(tloc) * B;
#pragma unroll
    for (int j = 0; j < NV; j++) *(cplx_t<R> *)(tb + (s0 ^ from[j])) = v[j];
    lds_barrier();
    tloc = thread_base(tol, (int)threadIdx.x, RB);
    const uint32_t s1 = (uint32_t)lds_slot<B>(tloc) * B;
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] = *(const cplx_t<R> *)(tb + (s1 ^ to[j]));
}

// 2x2 on register bit P for the register values j whose register-bit
// controls hold (bit j of jm; MASKED false: all, with no branch between the
// pairs, so their arithmetic interleaves).  FORM: 0 general (apply_op's
// BOP_MAT1 arithmetic), 1 real entries (RY, H), 2 real diagonal and
// imaginary off-diagonal (RX); forms 1 and 2 skip the zero products.
template <typename R, int RB, int P, bool MASKED, int FORM>
__device__ __forceinline__ void reg_mat1(cplx_t<R> (&v)[1 << RB], const BOp &op, uint32_t jm)
{
    constexpr int NV = 1 << RB;
    const R m0 = (R)op.m[0], m1 = (R)op.m[1], m2 = (R)op.m[2], m3 = (R)op.m[3];
    const R m4 = (R)op.m[4], m5 = (R)op.m[5], m6 = (R)op.m[6], m7 = (R)op.m[7];
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (j & (1 << P)) continue;
        if (MASKED && !((jm >> j) & 1u)) continue;
        const cplx_t<R> a = v[j], b = v[j | (1 << P)];
        cplx_t<R> x, y;
        if (FORM == 1) {
            x.x = fmar(m0, a.x, m2 * b.x);
            x.y = fmar(m0, a.y, m2 * b.y);
            y.x = fmar(m4, a.x, m6 * b.x);
            y.y = fmar(m4, a.y, m6 * b.y);
        } else if (FORM == 2) {   // u00 = m0, u01 = i m3, u10 = i m5, u11 = m6
            x.x = fmar(m0, a.x, -(m3 * b.y));
            x.y = fmar(m0, a.y, m3 * b.x);
            y.x = fmar(m6, b.x, -(m5 * a.y));
            y.y = fmar(m6, b.y, m5 * a.x);
        } else {
            x.x = fmar(m0, a.x, -(m1 * a.y)) + fmar(m2, b.x, -(m3 * b.y));
            x.y = fmar(m0, a.y, m1 * a.x) + fmar(m2, b.y, m3 * b.x);
            y.x = fmar(m4, a.x, -(m5 * a.y)) + fmar(m6, b.x, -(m7 * b.y));
            y.y = fmar(m4, a.y, m5 * a.x) + fmar(m6, b.y, m7 * b.x);
        }
        v[j] = x;
        v[j | (1 << P)] = y;
    }
}

// controlled X on register bit P for the values selected by jm (any
// controls)
template <typename R, int RB, int P>
__device__ __forceinline__ void reg_xperm(cplx_t<R> (&v)[1 << RB], uint32_t jm)
{
    constexpr int NV = 1 << RB;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (j & (1 << P)) continue;
        if (!((jm >> j) & 1u)) continue;
        const cplx_t<R> a = v[j];
        v[j] = v[j | (1 << P)];
        v[j | (1 << P)] = a;
    }
}

// CNOT with target register bit P and control register bit C: a fixed
// exchange of register values
template <typename R, int RB, int P, int C>
__device__ __forceinline__ void reg_cx(cplx_t<R> (&v)[1 << RB])
{
    constexpr int NV = 1 << RB;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if ((j & (1 << P)) || !(j & (1 << C))) continue;
        const cplx_t<R> a = v[j];
        v[j] = v[j | (1 << P)];
        v[j | (1 << P)] = a;
    }
}

// single-qubit diagonal on register bit P: value j *= d[bit P of j]
// (ONLY1: d0 is exactly 1 and those values are left alone)
template <typename R, int RB, int P, bool ONLY1>
__device__ __forceinline__ void reg_diag1(cplx_t<R> (&v)[1 << RB], const BOp &op)
{
    constexpr int NV = 1 << RB;
    const R d0r = (R)op.m[0], d0i = (R)op.m[1], d1r = (R)op.m[2], d1i = (R)op.m[3];
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (j & (1 << P)) v[j] = cmul(v[j], d1r, d1i);
        else if (!ONLY1) v[j] = cmul(v[j], d0r, d0i);
    }
}

// two-qubit diagonal diag(1, 1, 1, d3) on register bits P < Q (CZ, CPHASE)
template <typename R, int RB, int P, int Q>
__device__ __forceinline__ void reg_diag11(cplx_t<R> (&v)[1 << RB], const BOp &op)
{
    constexpr int NV = 1 << RB;
    const R d3r = (R)op.m[6], d3i = (R)op.m[7];
#pragma unroll
    for (int j = 0; j < NV; j++)
        if ((j & (1 << P)) && (j & (1 << Q))) v[j] = cmul(v[j], d3r, d3i);
}

// any diagonal op: value j *= d[k], k = kg | bits 2j..2j+1 of kmap (the
// local qubits' bits of register value j); kg holds the global qubits' bits
template <typename R, int RB>
__device__ __forceinline__ void reg_diag(cplx_t<R> (&v)[1 << RB], const BOp &op, uint32_t jm, uint32_t kmap, int kg)
{
    constexpr int NV = 1 << RB;
    MLX_DIAG4(R, op);
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (!((jm >> j) & 1u)) continue;
        const int k = kg | (int)((kmap >> (2 * j)) & 3u);
        v[j] = cmul(v[j], pick4(k, dr0, dr1, dr2, dr3), pick4(k, di0, di1, di2, di3));
    }
}

// bit of a global qubit code (-1 - q) of the tile's base index; 0 for none
__device__ __forceinline__ int global_bit(int32_t code, uint64_t base)
{
    return (code < 0 && code != INT32_MIN) ? (int)((base >> (-1 - code)) & 1ull) : 0;
}

// One op of a register phase, dispatched on its arm (BOp::pad, chosen by the
// host; the numbering is RegArm in batch_device.h).  The planner places every
// tile qubit an op touches in the phase's register bits, so the op is the
// same for every thread of the tile.
template <typename R, int RB>
__device__ __forceinline__ void reg_op(cplx_t<R> (&v)[1 << RB], const BOp &op, uint64_t base)
{
    constexpr int NV = 1 << RB;
    static_assert(RB == 3 || RB == 4, "register positions 0..RB-1");
    if ((base & op.cg) != op.cg) return;   // a global control is 0 for this whole tile
    const uint32_t jm = op.pad2;
    switch (op.pad) {
    case ARM_MAT1_MASKED + 0: reg_mat1<R, RB, 0, true, 0>(v, op, jm); break;
    case ARM_MAT1_MASKED + 1: reg_mat1<R, RB, 1, true, 0>(v, op, jm); break;
    case ARM_MAT1_MASKED + 2: reg_mat1<R, RB, 2, true, 0>(v, op, jm); break;
    case ARM_MAT1_MASKED + 3: if constexpr (RB > 3) reg_mat1<R, RB, 3, true, 0>(v, op, jm); break;
    case ARM_MAT1 + 0: reg_mat1<R, RB, 0, false, 0>(v, op, jm); break;
    case ARM_MAT1 + 1: reg_mat1<R, RB, 1, false, 0>(v, op, jm); break;
    case ARM_MAT1 + 2: reg_mat1<R, RB, 2, false, 0>(v, op, jm); break;
    case ARM_MAT1 + 3: if constexpr (RB > 3) reg_mat1<R, RB, 3, false, 0>(v, op, jm); break;
    case ARM_MAT1_REAL + 0: reg_mat1<R, RB, 0, false, 1>(v, op, jm); break;
    case ARM_MAT1_REAL + 1: reg_mat1<R, RB, 1, false, 1>(v, op, jm); break;
    case ARM_MAT1_REAL + 2: reg_mat1<R, RB, 2, false, 1>(v, op, jm); break;
    case ARM_MAT1_REAL + 3: if constexpr (RB > 3) reg_mat1<R, RB, 3, false, 1>(v, op, jm); break;
    case ARM_MAT1_RX + 0: reg_mat1<R, RB, 0, false, 2>(v, op, jm); break;
    case ARM_MAT1_RX + 1: reg_mat1<R, RB, 1, false, 2>(v, op, jm); break;
    case ARM_MAT1_RX + 2: reg_mat1<R, RB, 2, false, 2>(v, op, jm); break;
    case ARM_MAT1_RX + 3: if constexpr (RB > 3) reg_mat1<R, RB, 3, false, 2>(v, op, jm); break;
    case ARM_XPERM + 0: reg_xperm<R, RB, 0>(v, jm); break;
    case ARM_XPERM + 1: reg_xperm<R, RB, 1>(v, jm); break;
    case ARM_XPERM + 2: reg_xperm<R, RB, 2>(v, jm); break;
    case ARM_XPERM + 3: if constexpr (RB > 3) reg_xperm<R, RB, 3>(v, jm); break;
    case ARM_CX + 4 * 0 + 1: reg_cx<R, RB, 0, 1>(v); break;
    case ARM_CX + 4 * 0 + 2: reg_cx<R, RB, 0, 2>(v); break;
    case ARM_CX + 4 * 0 + 3: if constexpr (RB > 3) reg_cx<R, RB, 0, 3>(v); break;
    case ARM_CX + 4 * 1 + 0: reg_cx<R, RB, 1, 0>(v); break;
    case ARM_CX + 4 * 1 + 2: reg_cx<R, RB, 1, 2>(v); break;
    case ARM_CX + 4 * 1 + 3: if constexpr (RB > 3) reg_cx<R, RB, 1, 3>(v); break;
    case ARM_CX + 4 * 2 + 0: reg_cx<R, RB, 2, 0>(v); break;
    case ARM_CX + 4 * 2 + 1: reg_cx<R, RB, 2, 1>(v); break;
    case ARM_CX + 4 * 2 + 3: if constexpr (RB > 3) reg_cx<R, RB, 2, 3>(v); break;
    case ARM_CX + 4 * 3 + 0: if constexpr (RB > 3) reg_cx<R, RB, 3, 0>(v); break;
    case ARM_CX + 4 * 3 + 1: if constexpr (RB > 3) reg_cx<R, RB, 3, 1>(v); break;
    case ARM_CX + 4 * 3 + 2: if constexpr (RB > 3) reg_cx<R, RB, 3, 2>(v); break;
    case ARM_DIAG1 + 0: reg_diag1<R, RB, 0, false>(v, op); break;
    case ARM_DIAG1 + 1: reg_diag1<R, RB, 1, false>(v, op); break;
    case ARM_DIAG1 + 2: reg_diag1<R, RB, 2, false>(v, op); break;
    case ARM_DIAG1 + 3: if constexpr (RB > 3) reg_diag1<R, RB, 3, false>(v, op); break;
    case ARM_DIAG1_ONLY1 + 0: reg_diag1<R, RB, 0, true>(v, op); break;
    case ARM_DIAG1_ONLY1 + 1: reg_diag1<R, RB, 1, true>(v, op); break;
    case ARM_DIAG1_ONLY1 + 2: reg_diag1<R, RB, 2, true>(v, op); break;
    case ARM_DIAG1_ONLY1 + 3: if constexpr (RB > 3) reg_diag1<R, RB, 3, true>(v, op); break;
    case ARM_DIAG11 + 4 * 0 + 1: reg_diag11<R, RB, 0, 1>(v, op); break;
    case ARM_DIAG11 + 4 * 0 + 2: reg_diag11<R, RB, 0, 2>(v, op); break;
    case ARM_DIAG11 + 4 * 0 + 3: if constexpr (RB > 3) reg_diag11<R, RB, 0, 3>(v, op); break;
    case ARM_DIAG11 + 4 * 1 + 2: reg_diag11<R, RB, 1, 2>(v, op); break;
    case ARM_DIAG11 + 4 * 1 + 3: if constexpr (RB > 3) reg_diag11<R, RB, 1, 3>(v, op); break;
    case ARM_DIAG11 + 4 * 2 + 3: if constexpr (RB > 3) reg_diag11<R, RB, 2, 3>(v, op); break;
    default: {   // ARM_DIAG
        const int kg = global_bit(op.lq, base) | (global_bit(op.lq2, base) << 1);
        reg_diag<R, RB>(v, op, jm, op.cl, __builtin_amdgcn_readfirstlane(kg));
        break;
    }
    }
}

// One pass.  A tile enters and leaves in the IO layout, whose register bits
// are the top RB tile bits: the lanes of a wave then cover the low tile
// bits, which are runs of contiguous amplitudes.
template <typename R, bool NT, int TB, int RB>
__global__ __launch_bounds__(1 << (TB - RB)) void k_reg_pass(cplx_t<R> *__restrict__ a, BPass P,
                                                            const BOp *__restrict__ ops, uint64_t ntiles)
{
    constexpr int NV = 1 << RB, THREADS = 1 << (TB - RB), HSTEP = THREADS / 64, NHI = 1 << (TB - 6);
    static_assert(THREADS >= 64 && TB <= 13, "tile shape");
    __shared__ cplx_t<R> t[1 << TB];
    __shared__ uint64_t dlo[64], dhi[NHI];
    const int tid = threadIdx.x;
    // global offsets of tile-index bits 0..5 (dlo) and 6..TB-1 (dhi); the
    // tile set is indexed with constants only (P.nbits == TB), so it stays in
    // scalar registers
    for (int e = tid; e < 64 + NHI; e += THREADS) {
        uint64_t o = 0;
        if (e < 64) {
#pragma unroll
            for (int b = 0; b < 6; b++)
                if ((e >> b) & 1) o |= 1ull << P.qubits[b];
            dlo[e] = o;
        } else {
#pragma unroll
            for (int b = 6; b < TB; b++)
                if (((e - 64) >> (b - 6)) & 1) o |= 1ull << P.qubits[b];
            dhi[e - 64] = o;
        }
    }
    __syncthreads();
    constexpr int B = (int)sizeof(cplx_t<R>);
    RegLayout io{};
#pragma unroll
    for (int i = 0; i < RB; i++) io.r[i] = TB - RB + i;
    const uint32_t io_mask = ((1u << RB) - 1) << (TB - RB);
    uint32_t io_off[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) io_off[j] = (uint32_t)lds_slot<B>(j << (TB - RB)) * B;
    // IO layout: value j of thread tid is tile index tid + THREADS j
    const uint64_t lo = dlo[tid & 63];
    const int hi0 = tid >> 6;
    const BOp *pops = ops + P.first_op;
    cplx_t<R> v[NV], pre[NV];
    uint64_t ti = blockIdx.x;
    if (ti < ntiles) {
        const uint64_t b0 = tile_base_c<TB>(ti, P) | lo;
#pragma unroll
        for (int j = 0; j < NV; j++) pre[j] = ld<NT>(&a[b0 | dhi[hi0 + HSTEP * j]]);
    }
    for (; ti < ntiles; ti += gridDim.x) {
#pragma unroll
        for (int j = 0; j < NV; j++) v[j] = pre[j];
        const uint64_t nti = ti + gridDim.x;
        if (nti < ntiles) {
            const uint64_t b1 = tile_base_c<TB>(nti, P) | lo;
#pragma unroll
            for (int j = 0; j < NV; j++) pre[j] = ld<NT>(&a[b1 | dhi[hi0 + HSTEP * j]]);
        }
        const uint64_t base = tile_base_c<TB>(ti, P);
        int tloc = tid;
        uint32_t cur_off[NV];
#pragma unroll
        for (int j = 0; j < NV; j++) cur_off[j] = io_off[j];
        uint32_t cur_mask = io_mask;
        for (int k = 0; k < P.num_ops;) {
            const BOp &h = pops[k];   // BOP_PHASE header
            const uint32_t mask = (uint32_t)h.lq2;
            if (mask != cur_mask) {
                uint32_t nof[NV];
                const uint32_t *hs = (const uint32_t *)h.m;   // the new layout's slot byte offsets
#pragma unroll
                for (int j = 0; j < NV; j++) nof[j] = hs[j];
                relayout<R, RB>(t, v, tloc, cur_off, nof, unpack_layout(h.cl, RB));
#pragma unroll
                for (int j = 0; j < NV; j++) cur_off[j] = nof[j];
                cur_mask = mask;
            }
            const int cnt = h.lq;
            for (int q = 1; q <= cnt; q++) reg_op<R, RB>(v, pops[k + q], base);
            k += 1 + cnt;
        }
        if (cur_mask != io_mask) relayout<R, RB>(t, v, tloc, cur_off, io_off, io);
#pragma unroll
        for (int j = 0; j < NV; j++) st<NT>(&a[(base | lo) | dhi[hi0 + HSTEP * j]], v[j]);
    }
}

template <typename R, bool NT, int TB, int RB>
hipError_t run_reg_pass(void *amps, const BPass &P, const BOp *d_ops, uint64_t ntiles, int cus, hipStream_t st)
{
    constexpr int THREADS = 1 << (TB - RB);
    static std::atomic<int> per_cu_cached{0};
    int per_cu = per_cu_cached.load(std::memory_order_relaxed);
    if (!per_cu) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_reg_pass<R, NT, TB, RB>, THREADS, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        per_cu_cached.store(per_cu, std::memory_order_relaxed);
    }
    const uint64_t want = (uint64_t)cus * (uint64_t)per_cu;
    const unsigned grid = (unsigned)(ntiles < want ? ntiles : want);
    k_reg_pass<R, NT, TB, RB><<<grid, THREADS, 0, st>>>((cplx_t<R> *)amps, P, d_ops, ntiles);
    return hipGetLastError();
}

template <typename R, bool NT>
hipError_t reg_pass_shape(void *amps, const BPass &P, const BOp *d_ops, uint64_t ntiles, int cus, int reg_bits,
                          hipStream_t st)
{
    if (P.nbits == 13)
        return reg_bits == 3 ? run_reg_pass<R, NT, 13, 3>(amps, P, d_ops, ntiles, cus, st)
                             : run_reg_pass<R, NT, 13, 4>(amps, P, d_ops, ntiles, cus, st);
    return reg_bits == 3 ? run_reg_pass<R, NT, 12, 3>(amps, P, d_ops, ntiles, cus, st)
                         : run_reg_pass<R, NT, 12, 4>(amps, P, d_ops, ntiles, cus, st);
}

}  // namespace

// One register-phase pass (k_reg_pass) of the single-state blocked path:
// the pass's tile set has 12 or 13 qubits, n is at least that, and the
// program was planned for reg_bits (3 or 4) register bits.
hipError_t launch_reg_pass(void *amps, int prec, int n, const BPass &P, const BOp *d_ops, int cus, int reg_bits,
                           hipStream_t st)
{
    if ((P.nbits != 12 && P.nbits != 13) || n < P.nbits || (reg_bits != 3 && reg_bits != 4))
        return hipErrorInvalidValue;
    const uint64_t ntiles = 1ull << (n - P.nbits);
    const bool nt = ((1ull << n) * (prec == 1 ? 8ull : 16ull)) >= gate_tuning().nt_min_bytes;
    if (prec == 1)
        return nt ? reg_pass_shape<float, true>(amps, P, d_ops, ntiles, cus, reg_bits, st)
                  : reg_pass_shape<float, false>(amps, P, d_ops, ntiles, cus, reg_bits, st);
    return nt ? reg_pass_shape<double, true>(amps, P, d_ops, ntiles, cus, reg_bits, st)
              : reg_pass_shape<double, false>(amps, P, d_ops, ntiles, cus, reg_bits, st);
}

}  // namespace mlx
