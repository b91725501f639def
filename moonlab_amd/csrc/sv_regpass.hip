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

// bit of a global qubit code (-1 - q) of the tile's base index; 0 for none
__device__ __forceinline__ int global_bit(int32_t code, uint64_t base)
{
    return (code < 0 && code != INT32_MIN) ? (int)((base >> (-1 - code)) & 1ull) : 0;
}

// The ops of a register phase.  Each op is written from one register set into
// another (every value of `out` is written: computed, or copied from `in`),
// and the op loop alternates the two sets: each arm's results land where the
// next op reads them, so the merge after the arm dispatch needs no register
// copies; only the values an op leaves alone are copied.  (Updating one set
// in place made the compiler keep a second copy of the set and copy all of
// it back after every op.)
//
// 2x2 on register bit P for the register values j whose register-bit
// controls hold (bit j of jm; MASKED false: all).  FORM: 0 general (apply_op's
// BOP_MAT1 arithmetic, m[0..7]), 1 real entries (RY, H: m[0..3] = u00, u01,
// u10, u11), 2 real diagonal and imaginary off-diagonal (RX: m[0..3] = Re u00,
// Im u01, Im u10, Re u11); forms 1 and 2 skip the zero products.
template <typename R, int RB, int P, bool MASKED, int FORM>
__device__ __forceinline__ void pp_mat1(const cplx_t<R> (&in)[1 << RB], cplx_t<R> (&out)[1 << RB], const double *m,
                                        uint32_t jm)
{
    constexpr int NV = 1 << RB;
    // forms 1 and 2 keep the general form's names for their four entries:
    // m0, m2, m4, m6 (real) and m0, m3, m5, m6 (RX)
    const R m0 = (R)m[0];
    const R m1 = FORM == 0 ? (R)m[1] : R(0);
    const R m2 = FORM == 1 ? (R)m[1] : FORM == 0 ? (R)m[2] : R(0);
    const R m3 = FORM == 2 ? (R)m[1] : FORM == 0 ? (R)m[3] : R(0);
    const R m4 = FORM == 1 ? (R)m[2] : FORM == 0 ? (R)m[4] : R(0);
    const R m5 = FORM == 2 ? (R)m[2] : FORM == 0 ? (R)m[5] : R(0);
    const R m6 = FORM == 0 ? (R)m[6] : (R)m[3];
    const R m7 = FORM == 0 ? (R)m[7] : R(0);
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (j & (1 << P)) continue;
        const cplx_t<R> a = in[j], b = in[j | (1 << P)];
        cplx_t<R> x, y;
        if (FORM == 1) {
            x.x = fmar(m0, a.x, m2 * b.x);
            x.y = fmar(m0, a.y, m2 * b.y);
            y.x = fmar(m4, a.x, m6 * b.x);
            y.y = fmar(m4, a.y, m6 * b.y);
        } else if (FORM == 2) {
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
        if (MASKED && !((jm >> j) & 1u)) {
            out[j] = a;
            out[j | (1 << P)] = b;
        } else {
            out[j] = x;
            out[j | (1 << P)] = y;
        }
    }
}

// one 2x2 of form F (0 real: m[0..3] = u00, u01, u10, u11; 1 RX: m[0..3] =
// Re u00, Im u01, Im u10, Re u11) on the pair (a, b), the arithmetic of
// pp_mat1's forms 1 and 2
template <typename R, int F>
__device__ __forceinline__ void rot_pair(const R (&m)[4], cplx_t<R> &a, cplx_t<R> &b)
{
    cplx_t<R> x, y;
    if (F == 0) {
        x.x = fmar(m[0], a.x, m[1] * b.x);
        x.y = fmar(m[0], a.y, m[1] * b.y);
        y.x = fmar(m[2], a.x, m[3] * b.x);
        y.y = fmar(m[2], a.y, m[3] * b.y);
    } else {
        x.x = fmar(m[0], a.x, -(m[1] * b.y));
        x.y = fmar(m[0], a.y, m[1] * b.x);
        y.x = fmar(m[3], b.x, -(m[2] * a.y));
        y.y = fmar(m[3], b.y, m[2] * a.x);
    }
    a = x;
    b = y;
}

// two rotations on register bits P < Q in one pass over the values: form F
// on P (m[0..3]), then form G on Q (m[4..7]); they commute, so the order is
// the planner's choice
template <typename R, int RB, int P, int Q, int F, int G>
__device__ __forceinline__ void pp_rot2(const cplx_t<R> (&in)[1 << RB], cplx_t<R> (&out)[1 << RB], const double *m)
{
    constexpr int NV = 1 << RB;
    const R u[4] = {(R)m[0], (R)m[1], (R)m[2], (R)m[3]};
    const R v[4] = {(R)m[4], (R)m[5], (R)m[6], (R)m[7]};
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (j & ((1 << P) | (1 << Q))) continue;
        cplx_t<R> a = in[j], b = in[j | (1 << P)], c = in[j | (1 << Q)], d = in[j | (1 << P) | (1 << Q)];
        rot_pair<R, F>(u, a, b);
        rot_pair<R, F>(u, c, d);
        rot_pair<R, G>(v, a, c);
        rot_pair<R, G>(v, b, d);
        out[j] = a;
        out[j | (1 << P)] = b;
        out[j | (1 << Q)] = c;
        out[j | (1 << P) | (1 << Q)] = d;
    }
}

// controlled X on register bit P for the values selected by jm (any controls)
template <typename R, int RB, int P>
__device__ __forceinline__ void pp_xperm(const cplx_t<R> (&in)[1 << RB], cplx_t<R> (&out)[1 << RB], uint32_t jm)
{
    constexpr int NV = 1 << RB;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (j & (1 << P)) continue;
        const bool sw = (jm >> j) & 1u;
        out[j] = csel(sw, in[j | (1 << P)], in[j]);
        out[j | (1 << P)] = csel(sw, in[j], in[j | (1 << P)]);
    }
}

// CNOT with target register bit P and control register bit C: a fixed
// exchange of register values
template <typename R, int RB, int P, int C>
__device__ __forceinline__ void pp_cx(const cplx_t<R> (&in)[1 << RB], cplx_t<R> (&out)[1 << RB])
{
    constexpr int NV = 1 << RB;
#pragma unroll
    for (int j = 0; j < NV; j++) out[j] = in[(j & (1 << C)) ? (j ^ (1 << P)) : j];
}

// single-qubit diagonal on register bit P: value j *= d[bit P of j]
// (ONLY1: d0 is exactly 1 and those values are copied)
template <typename R, int RB, int P, bool ONLY1>
__device__ __forceinline__ void pp_diag1(const cplx_t<R> (&in)[1 << RB], cplx_t<R> (&out)[1 << RB], const double *m)
{
    constexpr int NV = 1 << RB;
    const R d0r = (R)m[0], d0i = (R)m[1], d1r = (R)m[2], d1i = (R)m[3];
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (j & (1 << P)) out[j] = cmul(in[j], d1r, d1i);
        else out[j] = ONLY1 ? in[j] : cmul(in[j], d0r, d0i);
    }
}

// two-qubit diagonal diag(1, 1, 1, d3) on register bits P < Q (CZ, CPHASE)
template <typename R, int RB, int P, int Q>
__device__ __forceinline__ void pp_diag11(const cplx_t<R> (&in)[1 << RB], cplx_t<R> (&out)[1 << RB], const double *m)
{
    constexpr int NV = 1 << RB;
    const R d3r = (R)m[0], d3i = (R)m[1];
#pragma unroll
    for (int j = 0; j < NV; j++) out[j] = ((j & (1 << P)) && (j & (1 << Q))) ? cmul(in[j], d3r, d3i) : in[j];
}

// any diagonal op: value j *= d[k], k = kg | bits 2j..2j+1 of kmap (the
// local qubits' bits of register value j); kg holds the global qubits' bits
template <typename R, int RB>
__device__ __forceinline__ void pp_diag(const cplx_t<R> (&in)[1 << RB], cplx_t<R> (&out)[1 << RB], const BOp &op,
                                        uint32_t jm, uint32_t kmap, int kg)
{
    constexpr int NV = 1 << RB;
    MLX_DIAG4(R, op);
#pragma unroll
    for (int j = 0; j < NV; j++) {
        const int k = kg | (int)((kmap >> (2 * j)) & 3u);
        const cplx_t<R> d = cmul(in[j], pick4(k, dr0, dr1, dr2, dr3), pick4(k, di0, di1, di2, di3));
        out[j] = csel(((jm >> j) & 1u) != 0, d, in[j]);
    }
}

template <typename R, int RB>
__device__ __forceinline__ void pp_copy(const cplx_t<R> (&in)[1 << RB], cplx_t<R> (&out)[1 << RB])
{
#pragma unroll
    for (int j = 0; j < (1 << RB); j++) out[j] = in[j];
}

// The fields of an op that every common arm reads (the arm, jm, cg and
// m[0..3]: one 64-byte scalar load), fetched one op ahead of their use.
struct OpHead {
    int32_t  arm;
    uint32_t jm;
    uint64_t cg;
    double   m[4];
};

// p itself, with an address that depends on `arm` (always >= 0): a fetch
// from it issues only once `arm` has arrived
__device__ __forceinline__ const BOp *after(const BOp *p, int32_t arm)
{
    return (const BOp *)((const char *)p + ((arm >> 31) & 64));
}

__device__ __forceinline__ OpHead op_head(const BOp *o)
{
    OpHead h;
    h.arm = o->pad;
    h.jm = o->pad2;
    h.cg = o->cg;
#pragma unroll
    for (int i = 0; i < 4; i++) h.m[i] = o->m[i];
    return h;
}

// One op, dispatched on its arm (BOp::pad, chosen by the host; the numbering
// is RegArm in batch_device.h).  The planner places every tile qubit an op
// touches in the phase's register bits, so the op is the same for every
// thread of the tile.  The general arms read m[4..7] (and ARM_DIAG its other
// fields) from the op itself.
template <typename R, int RB>
__device__ __forceinline__ void pp_op(const cplx_t<R> (&in)[1 << RB], cplx_t<R> (&out)[1 << RB], const OpHead &h,
                                      const BOp *op, uint64_t base)
{
    static_assert(RB == 3 || RB == 4, "register positions 0..RB-1");
    // arm -1: a global control is 0 for this whole tile
    const int arm = (base & h.cg) == h.cg ? h.arm : -1;
    const uint32_t jm = h.jm;
    const double *m = h.m;
    if (arm < 0) {
        pp_copy<R, RB>(in, out);
        return;
    }
    double m8[8];   // the general arms' eight entries
    auto full = [&]() -> const double * {
#pragma unroll
        for (int i = 0; i < 4; i++) m8[i] = h.m[i];
#pragma unroll
        for (int i = 4; i < 8; i++) m8[i] = op->m[i];
        return m8;
    };
#define MLX_PP4(ARM, CALL)                                                         \
    case ARM + 0: CALL(0); break;                                                  \
    case ARM + 1: CALL(1); break;                                                  \
    case ARM + 2: CALL(2); break;                                                  \
    case ARM + 3:                                                                  \
        if constexpr (RB > 3) CALL(3);                                             \
        else pp_copy<R, RB>(in, out);                                              \
        break;
#define MLX_PP_MASKED(P) pp_mat1<R, RB, P, true, 0>(in, out, full(), jm)
#define MLX_PP_MAT1(P) pp_mat1<R, RB, P, false, 0>(in, out, full(), jm)
#define MLX_PP_REAL(P) pp_mat1<R, RB, P, false, 1>(in, out, m, jm)
#define MLX_PP_RX(P) pp_mat1<R, RB, P, false, 2>(in, out, m, jm)
#define MLX_PP_XPERM(P) pp_xperm<R, RB, P>(in, out, jm)
#define MLX_PP_DIAG1(P) pp_diag1<R, RB, P, false>(in, out, m)
#define MLX_PP_ONLY1(P) pp_diag1<R, RB, P, true>(in, out, m)
#define MLX_PP_PAIR(ARM, P, Q, CALL)                                               \
    case ARM + 4 * P + Q:                                                          \
        if constexpr (RB > 3 || (P < 3 && Q < 3)) CALL<R, RB, P, Q>(in, out);      \
        else pp_copy<R, RB>(in, out);                                              \
        break;
#define MLX_PP_ROT2_FG(P, Q, F, G)                                                  \
    case ARM_ROT2 + 4 * rot2_pair(P, Q) + 2 * F + G:                               \
        if constexpr (RB > 3 || Q < 3) pp_rot2<R, RB, P, Q, F, G>(in, out, full());  \
        else pp_copy<R, RB>(in, out);                                              \
        break;
#define MLX_PP_ROT2(P, Q)                                                          \
    MLX_PP_ROT2_FG(P, Q, 0, 0)                                                     \
    MLX_PP_ROT2_FG(P, Q, 0, 1)                                                     \
    MLX_PP_ROT2_FG(P, Q, 1, 0)                                                     \
    MLX_PP_ROT2_FG(P, Q, 1, 1)
#define MLX_PP_PAIR_OP(ARM, P, Q, CALL)                                            \
    case ARM + 4 * P + Q:                                                          \
        if constexpr (RB > 3 || (P < 3 && Q < 3)) CALL<R, RB, P, Q>(in, out, m);   \
        else pp_copy<R, RB>(in, out);                                              \
        break;
    switch (arm) {
        MLX_PP4(ARM_MAT1_MASKED, MLX_PP_MASKED)
        MLX_PP4(ARM_MAT1, MLX_PP_MAT1)
        MLX_PP4(ARM_MAT1_REAL, MLX_PP_REAL)
        MLX_PP4(ARM_MAT1_RX, MLX_PP_RX)
        MLX_PP4(ARM_XPERM, MLX_PP_XPERM)
        MLX_PP_PAIR(ARM_CX, 0, 1, pp_cx)
        MLX_PP_PAIR(ARM_CX, 0, 2, pp_cx)
        MLX_PP_PAIR(ARM_CX, 0, 3, pp_cx)
        MLX_PP_PAIR(ARM_CX, 1, 0, pp_cx)
        MLX_PP_PAIR(ARM_CX, 1, 2, pp_cx)
        MLX_PP_PAIR(ARM_CX, 1, 3, pp_cx)
        MLX_PP_PAIR(ARM_CX, 2, 0, pp_cx)
        MLX_PP_PAIR(ARM_CX, 2, 1, pp_cx)
        MLX_PP_PAIR(ARM_CX, 2, 3, pp_cx)
        MLX_PP_PAIR(ARM_CX, 3, 0, pp_cx)
        MLX_PP_PAIR(ARM_CX, 3, 1, pp_cx)
        MLX_PP_PAIR(ARM_CX, 3, 2, pp_cx)
        MLX_PP4(ARM_DIAG1, MLX_PP_DIAG1)
        MLX_PP4(ARM_DIAG1_ONLY1, MLX_PP_ONLY1)
        MLX_PP_PAIR_OP(ARM_DIAG11, 0, 1, pp_diag11)
        MLX_PP_PAIR_OP(ARM_DIAG11, 0, 2, pp_diag11)
        MLX_PP_PAIR_OP(ARM_DIAG11, 0, 3, pp_diag11)
        MLX_PP_PAIR_OP(ARM_DIAG11, 1, 2, pp_diag11)
        MLX_PP_PAIR_OP(ARM_DIAG11, 1, 3, pp_diag11)
        MLX_PP_PAIR_OP(ARM_DIAG11, 2, 3, pp_diag11)
        MLX_PP_ROT2(0, 1)
        MLX_PP_ROT2(0, 2)
        MLX_PP_ROT2(0, 3)
        MLX_PP_ROT2(1, 2)
        MLX_PP_ROT2(1, 3)
        MLX_PP_ROT2(2, 3)
    default: {   // ARM_DIAG
        const int kg = global_bit(op->lq, base) | (global_bit(op->lq2, base) << 1);
        pp_diag<R, RB>(in, out, *op, jm, op->cl, __builtin_amdgcn_readfirstlane(kg));
        break;
    }
    }
#undef MLX_PP4
#undef MLX_PP_MASKED
#undef MLX_PP_MAT1
#undef MLX_PP_REAL
#undef MLX_PP_RX
#undef MLX_PP_XPERM
#undef MLX_PP_DIAG1
#undef MLX_PP_ONLY1
#undef MLX_PP_PAIR
#undef MLX_PP_PAIR_OP
#undef MLX_PP_ROT2_FG
#undef MLX_PP_ROT2
}

// One pass.  A tile enters and leaves in the IO layout, whose register bits
// are the top RB tile bits: the lanes of a wave then cover the low tile
// bits, which are runs of contiguous amplitudes.
template <typename R, bool NT, int TB, int RB, int WPE>
__global__ __launch_bounds__(1 << (TB - RB), WPE) void k_reg_pass(cplx_t<R> *__restrict__ a, BPass P,
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
    const TileWalk walk = tile_walk(ntiles);
    uint64_t ti = walk.t;
    if (ti < walk.end) {
        const uint64_t b0 = tile_base_c<TB>(ti, P) | lo;
#pragma unroll
        for (int j = 0; j < NV; j++) pre[j] = ld<NT>(&a[b0 | dhi[hi0 + HSTEP * j]]);
    }
    for (; ti < walk.end; ti += walk.step) {
#pragma unroll
        for (int j = 0; j < NV; j++) v[j] = pre[j];
        const uint64_t nti = ti + walk.step;
        if (nti < walk.end) {
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
            // ops alternate between the two register sets (v -> w -> v);
            // each op's head is loaded while the op before it runs: its
            // address depends on the previous head's arm (after()), so the
            // load issues only once that head has arrived, and the wait
            // before an op never includes the load of the op after it.  The
            // op after a phase's last op is the next phase's header (or the
            // spare op the host puts after the last op): fetched, never run.
            // (For complex64 this prefetch was once 10 % slower than loading
            // each head where it is used; on short-lived workgroups it is
            // 0.5 % faster, so both precisions share it.)
            cplx_t<R> w[NV];
            const BOp *p = pops + k + 1;
            OpHead h0 = op_head(p);
            for (int q = 1; q <= cnt; q += 2, p += 2) {
                const OpHead h1 = op_head(after(p + 1, h0.arm));
                pp_op<R, RB>(v, w, h0, p, base);
                if (q < cnt) {
                    h0 = op_head(after(p + 2, h1.arm));
                    pp_op<R, RB>(w, v, h1, p + 1, base);
                } else {
                    pp_copy<R, RB>(w, v);
                }
            }
            k += 1 + cnt;
        }
        if (cur_mask != io_mask) relayout<R, RB>(t, v, tloc, cur_off, io_off, io);
#pragma unroll
        for (int j = 0; j < NV; j++) st<NT>(&a[(base | lo) | dhi[hi0 + HSTEP * j]], v[j]);
    }
}

// Eight tiles per workgroup, XCD-contiguous (tile_walk): short-lived
// workgroups that the dispatcher keeps refilling stream HBM faster than one
// resident wave of persistent workgroups.  30-qubit random layers, depth 40:
// complex128 0.724 -> 0.679 s, complex64 0.370 -> 0.341 s (2, 4 and 16 tiles
// per workgroup: 0.703, 0.684 and 0.680 s complex128).
constexpr uint64_t kTilesPerGroup = 8;

template <typename R, bool NT, int TB, int RB, int WPE>
hipError_t run_reg_pass_w(void *amps, const BPass &P, const BOp *d_ops, uint64_t ntiles, hipStream_t st)
{
    constexpr int THREADS = 1 << (TB - RB);
    const unsigned grid = (unsigned)(ntiles >= kTilesPerGroup ? ntiles / kTilesPerGroup : 1);
    k_reg_pass<R, NT, TB, RB, WPE><<<grid, THREADS, 0, st>>>((cplx_t<R> *)amps, P, d_ops, ntiles);
    return hipGetLastError();
}

template <typename R, bool NT, int TB, int RB>
hipError_t run_reg_pass(void *amps, const BPass &P, const BOp *d_ops, uint64_t ntiles, hipStream_t st)
{
    // complex64 tiles of 2^(TB - RB) >= 512 threads: two workgroups fit a
    // CU's LDS, so 4 waves per SIMD, which needs <= 128 VGPRs
    constexpr int WPE = (sizeof(R) == 4 && TB - RB >= 9) ? 4 : 1;
    return run_reg_pass_w<R, NT, TB, RB, WPE>(amps, P, d_ops, ntiles, st);
}

template <typename R, bool NT>
hipError_t reg_pass_shape(void *amps, const BPass &P, const BOp *d_ops, uint64_t ntiles, int reg_bits, hipStream_t st)
{
    if (P.nbits == 13)
        return reg_bits == 3 ? run_reg_pass<R, NT, 13, 3>(amps, P, d_ops, ntiles, st)
                             : run_reg_pass<R, NT, 13, 4>(amps, P, d_ops, ntiles, st);
    return reg_bits == 3 ? run_reg_pass<R, NT, 12, 3>(amps, P, d_ops, ntiles, st)
                         : run_reg_pass<R, NT, 12, 4>(amps, P, d_ops, ntiles, st);
}

}  // namespace

// One register-phase pass (k_reg_pass) of the single-state blocked path:
// the pass's tile set has 12 or 13 qubits, n is at least that, and the
// program was planned for reg_bits (3 or 4) register bits.
hipError_t launch_reg_pass(void *amps, int prec, int n, const BPass &P, const BOp *d_ops, int reg_bits, hipStream_t st)
{
    if ((P.nbits != 12 && P.nbits != 13) || n < P.nbits || (reg_bits != 3 && reg_bits != 4))
        return hipErrorInvalidValue;
    const uint64_t ntiles = 1ull << (n - P.nbits);
    const bool nt = ((1ull << n) * (prec == 1 ? 8ull : 16ull)) >= gate_tuning().nt_min_bytes;
    if (prec == 1)
        return nt ? reg_pass_shape<float, true>(amps, P, d_ops, ntiles, reg_bits, st)
                  : reg_pass_shape<float, false>(amps, P, d_ops, ntiles, reg_bits, st);
    return nt ? reg_pass_shape<double, true>(amps, P, d_ops, ntiles, reg_bits, st)
              : reg_pass_shape<double, false>(amps, P, d_ops, ntiles, reg_bits, st);
}

}  // namespace mlx
