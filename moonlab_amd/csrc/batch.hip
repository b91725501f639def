// batch.hip -- one persistent launch that simulates thousands of small,
// uneven circuits.
//
// Each 1024-thread workgroup takes the next circuit from an atomic work queue
// (the host orders it longest-first) and runs it start to finish:
//   * n <= tile bits (12): the whole state lives in LDS (64 KiB at
//     complex128); gates are LDS sweeps; observables and shots are computed
//     from LDS.  HBM sees only the program and the results.
//   * n > 12: the state lives in the workgroup's HBM slot and every pass
//     streams it through LDS in 2^12-amplitude tiles (cache blocking): the
//     host planner packs each run of gates whose non-diagonal qubits fit in a
//     12-qubit tile set into one pass, so a circuit costs a few HBM sweeps
//     per layer instead of one per gate.
// Workgroups never wait on each other; every workgroup exits when the queue
// is empty, so the grid always drains.
//
// Reference counterparts: the CPU VQE loop (src/algorithms/vqe.c:764-781,
// 1287-1350, 2156-2257) and the one-block-per-search batch kernel
// (src/optimization/gpu/backends/gpu_cuda.cu:350).
#include <hip/hip_runtime.h>

#include <limits.h>

#include <atomic>

#include "batch_device.h"
#include "sv_device.h"
#include "sv_internal.h"

namespace mlx {
namespace {

constexpr int NT = kBatchThreads;
constexpr int NW = NT / 64;
constexpr int kMaxObs = 1 + kBatchMaxQubits + (kBatchMaxQubits - 1) + kBatchMaxQubits;

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

struct Shared {
    double2  tile[1 << kBatchMaxTileBits];   // 64 KiB: state tile, later the CDF
    uint64_t dlo[64], dhi[64];               // deposit tables of the current tile set
    double   red[NW][kMaxObs];               // per-wave partials
    double   acc[kMaxObs];                   // per-circuit accumulators
    double   res[8];                         // block_reduce results
    BOp      opbuf[kOpBuf];                  // op descriptors of the current chunk
    int      next;
    int      flag;
    int      last_nz;
};

// Block reduction of v[0..nv): wave butterflies, then waves summed in order.
// Result j lands in out[j] (for j < nv) on every thread after the call.
__device__ void block_reduce(Shared &sh, double *v, int nv, double *out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 0; j < nv; j++) {
        const double s = wsum(v[j]);
        if (lane == 0) sh.red[wave][j] = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        double s = 0.0;
        for (int w = 0; w < NW; w++) s += sh.red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

template <typename R>
__device__ __forceinline__ double p2(const cplx_t<R> &x)
{
    const double r = (double)x.x, i = (double)x.y;
    return fmar(r, r, i * i);
}

__device__ __forceinline__ int bit_of(int32_t code, int l, uint64_t base)
{
    // code >= 0: local tile bit; code < 0 (and not INT32_MIN): global qubit -1 - code
    if (code == INT32_MIN) return 0;
    return code >= 0 ? (l >> code) & 1 : (int)((base >> (-1 - code)) & 1ull);
}

template <typename R, int STRIDE = NT>
__device__ void apply_op(cplx_t<R> *t, int T, const BOp &op, uint64_t base)
{
    if ((base & op.cg) != op.cg) return;          // a global control is 0 for this whole tile
    const int tid = threadIdx.x;
    if (op.kind == BOP_MAT1) {
        const R m0 = (R)op.m[0], m1 = (R)op.m[1], m2 = (R)op.m[2], m3 = (R)op.m[3];
        const R m4 = (R)op.m[4], m5 = (R)op.m[5], m6 = (R)op.m[6], m7 = (R)op.m[7];
        const int half = T >> 1;
        for (int p = tid; p < half; p += STRIDE) {
            const int i0 = (int)ins0((uint64_t)p, op.lq), i1 = i0 | (1 << op.lq);
            if ((i0 & (int)op.cl) != (int)op.cl) continue;
            const cplx_t<R> a = t[i0], b = t[i1];
            cplx_t<R> x, y;
            x.x = fmar(m0, a.x, -(m1 * a.y)) + fmar(m2, b.x, -(m3 * b.y));
            x.y = fmar(m0, a.y, m1 * a.x) + fmar(m2, b.y, m3 * b.x);
            y.x = fmar(m4, a.x, -(m5 * a.y)) + fmar(m6, b.x, -(m7 * b.y));
            y.y = fmar(m4, a.y, m5 * a.x) + fmar(m6, b.y, m7 * b.x);
            t[i0] = x;
            t[i1] = y;
        }
    } else if (op.kind == BOP_XPERM) {
        const int half = T >> 1;
        for (int p = tid; p < half; p += STRIDE) {
            const int i0 = (int)ins0((uint64_t)p, op.lq), i1 = i0 | (1 << op.lq);
            if ((i0 & (int)op.cl) != (int)op.cl) continue;
            const cplx_t<R> a = t[i0];
            t[i0] = t[i1];
            t[i1] = a;
        }
    } else if (op.kind == BOP_DIAG) {
        MLX_DIAG4(R, op);
        for (int l = tid; l < T; l += STRIDE) {
            if ((l & (int)op.cl) != (int)op.cl) continue;
            const int k = (bit_of(op.lq2, l, base) << 1) | bit_of(op.lq, l, base);
            t[l] = cmul(t[l], pick4(k, dr0, dr1, dr2, dr3), pick4(k, di0, di1, di2, di3));
        }
    } else {   // BOP_SWAP
        const int b1 = 1 << op.lq, b2 = 1 << op.lq2;
        for (int l = tid; l < T; l += STRIDE) {
            if ((l & b1) && !(l & b2) && (l & (int)op.cl) == (int)op.cl) {
                const int j = (l ^ b1) | b2;
                const cplx_t<R> x = t[l];
                t[l] = t[j];
                t[j] = x;
            }
        }
    }
}

// 2x2 on the register pair (j, j | 1 << P) of a segment, the arithmetic of
// apply_op's BOP_MAT1 term by term.
template <typename R, int NR, int P>
__device__ __forceinline__ void seg_mat1(cplx_t<R> (&v)[1 << NR], const BOp &op, const int (&idx)[1 << NR])
{
    const R m0 = (R)op.m[0], m1 = (R)op.m[1], m2 = (R)op.m[2], m3 = (R)op.m[3];
    const R m4 = (R)op.m[4], m5 = (R)op.m[5], m6 = (R)op.m[6], m7 = (R)op.m[7];
#pragma unroll
    for (int j = 0; j < (1 << NR); j++) {
        if (j & (1 << P)) continue;
        if ((idx[j] & (int)op.cl) != (int)op.cl) continue;
        const cplx_t<R> a = v[j], b = v[j | (1 << P)];
        cplx_t<R> x, y;
        x.x = fmar(m0, a.x, -(m1 * a.y)) + fmar(m2, b.x, -(m3 * b.y));
        x.y = fmar(m0, a.y, m1 * a.x) + fmar(m2, b.y, m3 * b.x);
        y.x = fmar(m4, a.x, -(m5 * a.y)) + fmar(m6, b.x, -(m7 * b.y));
        y.y = fmar(m4, a.y, m5 * a.x) + fmar(m6, b.y, m7 * b.x);
        v[j] = x;
        v[j | (1 << P)] = y;
    }
}

// controlled X on register bit P (apply_op's BOP_XPERM)
template <typename R, int NR, int P>
__device__ __forceinline__ void seg_xperm(cplx_t<R> (&v)[1 << NR], const BOp &op, const int (&idx)[1 << NR])
{
#pragma unroll
    for (int j = 0; j < (1 << NR); j++) {
        if (j & (1 << P)) continue;
        if ((idx[j] & (int)op.cl) != (int)op.cl) continue;
        const cplx_t<R> a = v[j];
        v[j] = v[j | (1 << P)];
        v[j | (1 << P)] = a;
    }
}

// exchange of register bits P1 / P2 (apply_op's BOP_SWAP)
template <typename R, int NR, int P1, int P2>
__device__ __forceinline__ void seg_swap(cplx_t<R> (&v)[1 << NR], const BOp &op, const int (&idx)[1 << NR])
{
#pragma unroll
    for (int j = 0; j < (1 << NR); j++) {
        if (!(j & (1 << P1)) || (j & (1 << P2))) continue;
        if ((idx[j] & (int)op.cl) != (int)op.cl) continue;
        const int k = (j ^ (1 << P1)) | (1 << P2);
        const cplx_t<R> x = v[j];
        v[j] = v[k];
        v[k] = x;
    }
}

// A segment: nops ops whose non-diagonal targets lie in the NR local bits
// rq.  Each thread gathers the 2^NR amplitudes of one group from LDS,
// applies all ops in registers and writes them back: one LDS round trip and
// one barrier for the whole segment instead of one per op.
template <typename R, int STRIDE, int NR>
__device__ __forceinline__ void apply_segment(cplx_t<R> *t, int T, const BOp *ops, int nops, uint32_t packed,
                                              uint64_t base)
{
    int rq[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) rq[i] = (int)((packed >> (5 * i)) & 31u);
    const int groups = T >> NR;
    for (int g = threadIdx.x; g < groups; g += STRIDE) {
        int gb = g;
#pragma unroll
        for (int i = 0; i < NR; i++) gb = (int)ins0((uint64_t)gb, rq[i]);
        int idx[1 << NR];
        cplx_t<R> v[1 << NR];
#pragma unroll
        for (int j = 0; j < (1 << NR); j++) {
            idx[j] = gb;
#pragma unroll
            for (int i = 0; i < NR; i++)
                if (j & (1 << i)) idx[j] |= 1 << rq[i];
            v[j] = t[idx[j]];
        }
        for (int k = 0; k < nops; k++) {
            const BOp &op = ops[k];
            if ((base & op.cg) != op.cg) continue;   // a global control is 0 for this whole tile
            if (op.kind == BOP_DIAG) {
                MLX_DIAG4(R, op);
#pragma unroll
                for (int j = 0; j < (1 << NR); j++) {
                    if ((idx[j] & (int)op.cl) != (int)op.cl) continue;
                    const int d = (bit_of(op.lq2, idx[j], base) << 1) | bit_of(op.lq, idx[j], base);
                    v[j] = cmul(v[j], pick4(d, dr0, dr1, dr2, dr3), pick4(d, di0, di1, di2, di3));
                }
            } else if (op.kind == BOP_MAT1) {
                const int p = op.lq == rq[0] ? 0 : (NR > 2 && op.lq != rq[1]) ? 2 : 1;
                if (p == 0) seg_mat1<R, NR, 0>(v, op, idx);
                else if (p == 1) seg_mat1<R, NR, 1>(v, op, idx);
                else seg_mat1<R, NR, (NR > 2 ? 2 : 1)>(v, op, idx);
            } else if (op.kind == BOP_XPERM) {
                const int p = op.lq == rq[0] ? 0 : (NR > 2 && op.lq != rq[1]) ? 2 : 1;
                if (p == 0) seg_xperm<R, NR, 0>(v, op, idx);
                else if (p == 1) seg_xperm<R, NR, 1>(v, op, idx);
                else seg_xperm<R, NR, (NR > 2 ? 2 : 1)>(v, op, idx);
            } else {   // BOP_SWAP of the register bits of lq and lq2
                const int p1 = op.lq == rq[0] ? 0 : (NR > 2 && op.lq != rq[1]) ? 2 : 1;
                const int p2 = op.lq2 == rq[0] ? 0 : (NR > 2 && op.lq2 != rq[1]) ? 2 : 1;
                if (p1 == 0 && p2 == 1) seg_swap<R, NR, 0, 1>(v, op, idx);
                else if (p1 == 1 && p2 == 0) seg_swap<R, NR, 1, 0>(v, op, idx);
                else if (NR > 2) {
                    constexpr int L = NR > 2 ? 2 : 1;
                    if (p1 == 0 && p2 == 2) seg_swap<R, NR, 0, L>(v, op, idx);
                    else if (p1 == 2 && p2 == 0) seg_swap<R, NR, L, 0>(v, op, idx);
                    else if (p1 == 1 && p2 == 2) seg_swap<R, NR, 1, L>(v, op, idx);
                    else if (p1 == 2 && p2 == 1) seg_swap<R, NR, L, 1>(v, op, idx);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < (1 << NR); j++) t[idx[j]] = v[j];
    }
}

// Run a pass's op list on the tile in LDS: segment headers send their ops
// through registers, other ops run one by one.  Every step ends with a
// workgroup barrier (LDS only when the caller overlaps global traffic).
template <typename R, int STRIDE, int NR, bool LDS_ONLY>
__device__ __forceinline__ void run_pass_ops(cplx_t<R> *t, int T, const BOp *ops, int nops, uint64_t base)
{
    for (int k = 0; k < nops;) {
        const BOp &op = ops[k];
        if (op.kind == BOP_SEG) {
            apply_segment<R, STRIDE, NR>(t, T, ops + k + 1, op.lq, op.cl, base);
            k += 1 + op.lq;
        } else {
            apply_op<R, STRIDE>(t, T, op, base);
            k++;
        }
        if (LDS_ONLY) lds_barrier();
        else __syncthreads();
    }
}

// run_pass_ops for the batch kernel: a circuit's op descriptors are read
// once per tile, and a read straight from device memory would expose a full
// memory latency per op, so they are staged in LDS kOpBuf at a time (one
// cooperative 16-byte-per-thread copy, one barrier) and applied from there.
template <typename R, int NR>
__device__ void run_pass_ops_staged(Shared &sh, cplx_t<R> *t, int T, const BOp *ops, int nops, uint64_t base)
{
    constexpr int kWords = (int)(sizeof(BOp) / sizeof(uint4));
    static_assert(sizeof(BOp) % sizeof(uint4) == 0, "BOp is copied as 16-byte words");
    for (int k = 0; k < nops;) {
        const int cnt = nops - k < kOpBuf ? nops - k : kOpBuf;
        const uint4 *src = (const uint4 *)(ops + k);
        uint4 *dst = (uint4 *)sh.opbuf;
        for (int w = threadIdx.x; w < cnt * kWords; w += NT) dst[w] = src[w];
        __syncthreads();
        int j = 0;
        while (j < cnt) {
            const BOp &op = sh.opbuf[j];
            if (op.kind == BOP_SEG) {
                if (j + 1 + op.lq > cnt) break;   // the segment runs past this chunk: restage from it
                apply_segment<R, NT, NR>(t, T, &sh.opbuf[j + 1], op.lq, op.cl, base);
                j += 1 + op.lq;
            } else {
                apply_op<R, NT>(t, T, op, base);
                j++;
            }
            __syncthreads();
        }
        // every thread has passed the last barrier: opbuf may be overwritten.
        // The host refuses plans whose segments do not fit one chunk
        // (segments_fit_staging), so j > 0 here; should one ever reach the
        // device, its segment is skipped rather than restaged forever.
        k += j ? j : 1 + sh.opbuf[0].lq;
    }
}

__device__ void build_deposit(Shared &sh, const BPass &p)
{
    const int tid = threadIdx.x;
    if (tid < 128) {
        const int j = tid & 63, hi = tid >> 6;
        uint64_t o = 0;
        for (int b = 0; b < 6; b++) {
            const int qi = hi * 6 + b;
            if (qi < p.nbits && ((j >> b) & 1)) o |= 1ull << p.qubits[qi];
        }
        if (hi) sh.dhi[j] = o;
        else sh.dlo[j] = o;
    }
}

__device__ __forceinline__ uint64_t tile_base(uint64_t t, const BPass &p)
{
    for (int b = 0; b < p.nbits; b++) t = ins0(t, p.qubits[b]);
    return t;
}

// Observables of one tile, added into sh.acc (value layout of obs_size()).
// Measurement pass 0 uses the low `tb` qubits as the tile (tiles in index
// order: norm, P(1), ZZ and X on the low qubits); pass 1 uses the top `tb`
// qubits (X on the qubits pass 0 could not pair).  Values are reduced eight
// at a time to bound register use; the order of every sum is fixed.
template <typename R>
__device__ void measure_tile(Shared &sh, const cplx_t<R> *t, int T, int tb, int n, uint64_t base,
                             int flags, int mpass, double *tile_total)
{
    const int tid = threadIdx.x;
    const int nzz = n > 1 ? n - 1 : 0;
    const int xb = 1 + n + nzz;               // first X slot
    const int nv = xb + n;
    const int qoff = mpass == 0 ? 0 : n - tb;  // global qubit of local bit 0
    const int half = T >> 1;
    for (int c0 = 0; c0 < nv; c0 += 8) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = 0.0;
        bool any = false;
        for (int j = 0; j < 8 && c0 + j < nv; j++) {
            const int vi = c0 + j;
            if (vi == 0) any = any || mpass == 0;
            else if (vi < 1 + n) any = any || (mpass == 0 && (flags & BF_Z));
            else if (vi < xb) any = any || (mpass == 0 && (flags & BF_ZZ));
            else any = any || (flags & BF_X);
        }
        if (!any) continue;   // uniform across the block
        if (mpass == 0 && c0 < xb) {
            for (int l = tid; l < T; l += NT) {
                const double p = p2<R>(t[l]);
                const uint64_t full = base | (uint64_t)l;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int vi = c0 + j;
                    if (vi == 0) {
                        v[j] += p;
                    } else if (vi < 1 + n) {
                        if ((flags & BF_Z) && ((full >> (vi - 1)) & 1ull)) v[j] += p;
                    } else if (vi < xb) {
                        const int q = vi - 1 - n;
                        if (flags & BF_ZZ) v[j] += (((full >> q) ^ (full >> (q + 1))) & 1ull) ? -p : p;
                    }
                }
            }
        }
        if (flags & BF_X) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int vi = c0 + j;
                if (vi < xb || vi >= nv) continue;
                const int q = vi - xb, lb = q - qoff;
                if (lb < 0 || lb >= tb || (mpass == 1 && q < tb)) continue;
                double s2 = 0.0;
                for (int pp = tid; pp < half; pp += NT) {
                    const int i0 = (int)ins0((uint64_t)pp, lb), i1 = i0 | (1 << lb);
                    s2 += fmar((double)t[i0].x, (double)t[i1].x, (double)t[i0].y * (double)t[i1].y);
                }
                v[j] += s2 + s2;
            }
        }
        const int cnt = nv - c0 < 8 ? nv - c0 : 8;
        block_reduce(sh, v, cnt, sh.res);
        if (tid < cnt) sh.acc[c0 + tid] += sh.res[tid];
        if (c0 == 0 && tile_total && tid == 0) *tile_total = sh.res[0];
        __syncthreads();
    }
}

// Pauli strings over one tile of the low tb qubits (tiles in index order):
// term k adds sum_l Re[conj(a_(i^x)) a_i i^ny (-1)^popcount(i & z)], i =
// base | l, into out[k].  The partner i ^ x is in the tile when x has no bit
// at or above tb; otherwise it is read from the circuit's HBM slot (the
// final state, which measurement does not modify).  Terms are reduced eight
// at a time; every sum runs in a fixed order, and out[k] is updated by one
// thread, tile after tile.
template <typename R>
__device__ void measure_pauli_tile(Shared &sh, const cplx_t<R> *t, int T, uint64_t base, const cplx_t<R> *slot,
                                   const PauliT *terms, int nterms, double *out)
{
    const int tid = threadIdx.x;
    const uint64_t lowmask = (uint64_t)T - 1ull;
    for (int c0 = 0; c0 < nterms; c0 += 8) {
        const int cnt = nterms - c0 < 8 ? nterms - c0 : 8;
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            v[j] = 0.0;
            if (j >= cnt) continue;   // uniform across the block
            const uint64_t x = terms[c0 + j].x, z = terms[c0 + j].z;
            const int ny = terms[c0 + j].ny & 3;
            const uint64_t xl = x & lowmask, xh = x & ~lowmask;
            double s = 0.0;
            for (int l = tid; l < T; l += NT) {
                const cplx_t<R> a = t[l];
                const cplx_t<R> b = xh == 0 ? t[(uint64_t)l ^ xl] : slot[(base ^ xh) | ((uint64_t)l ^ xl)];
                const double br = (double)b.x, bi = (double)b.y, ar = (double)a.x, ai = (double)a.y;
                const double re = fmar(br, ar, bi * ai);     // Re conj(b) a
                const double im = fmar(br, ai, -(bi * ar));  // Im conj(b) a
                const double c = ny == 0 ? re : ny == 1 ? -im : ny == 2 ? -re : im;
                s += (__popcll((base | (uint64_t)l) & z) & 1) ? -c : c;
            }
            v[j] = s;
        }
        block_reduce(sh, v, cnt, sh.res);
        if (tid < cnt) out[c0 + tid] += sh.res[tid];
        __syncthreads();
    }
}

// Inclusive scan of p over T entries in LDS (as doubles), index order.
__device__ void lds_scan(Shared &sh, double *cdf, int T, double offset)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = T >= NT ? T / NT : 1;
    const int lo = tid * per;
    double s = 0.0;
    if (lo < T)
        for (int e = 0; e < per; e++) s += cdf[lo + e];
    double x = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh.red[wave][0] = x;
    __syncthreads();
    double run = offset;
    for (int w = 0; w < wave; w++) run += sh.red[w][0];
    run += x - s;
    __syncthreads();
    if (lo < T)
        for (int e = 0; e < per; e++) {
            run += cdf[lo + e];
            cdf[lo + e] = run;
        }
    __syncthreads();
}

template <typename R>
__device__ void run_circuit(Shared &sh, const BInst &I, const BPass *passes, const BOp *ops,
                            double2 *slot_raw, double *tile_sums, const double *draws, double *obs,
                            uint64_t *outcomes, const PauliT *terms, void *state_out)
{
    const int tid = threadIdx.x;
    const int n = I.n, tb = I.tile_bits, T = 1 << tb;
    cplx_t<R> *t = (cplx_t<R> *)sh.tile;
    cplx_t<R> *slot = (cplx_t<R> *)slot_raw;
    const uint64_t ntiles = 1ull << (n - tb);
    const bool resident = n == tb;
    for (int j = tid; j < kMaxObs; j += NT) sh.acc[j] = 0.0;

    // ---- gate passes ----
    for (int pi = 0; pi < I.num_passes; pi++) {
        const BPass P = passes[I.first_pass + pi];
        if (!resident) build_deposit(sh, P);
        __syncthreads();
        for (uint64_t ti = 0; ti < ntiles; ti++) {
            const uint64_t base = resident ? 0 : tile_base(ti, P);
            if (pi == 0) {
                for (int l = tid; l < T; l += NT) t[l] = cplx_t<R>{(base | l) == 0 ? R(1) : R(0), R(0)};
            } else if (!resident) {
                for (int l = tid; l < T; l += NT) t[l] = slot[base | sh.dlo[l & 63] | sh.dhi[l >> 6]];
            }
            __syncthreads();
            run_pass_ops_staged<R, kSegRegsBatch>(sh, t, T, ops + P.first_op, P.num_ops, base);
            if (!resident) {
                for (int l = tid; l < T; l += NT) slot[base | sh.dlo[l & 63] | sh.dhi[l >> 6]] = t[l];
                __syncthreads();
            }
        }
    }
    if (I.num_passes == 0) {   // empty circuit: |0..0>
        for (uint64_t ti = 0; ti < (resident ? 1 : ntiles); ti++) {
            for (int l = tid; l < T; l += NT)
                t[l] = cplx_t<R>{(ti == 0 && l == 0) ? R(1) : R(0), R(0)};
            __syncthreads();
            if (!resident)
                for (int l = tid; l < T; l += NT) slot[(ti << tb) | l] = t[l];
            __syncthreads();
        }
    }

    // ---- the final state, when asked for (geometric-tensor overlaps) ----
    if (I.state_offset >= 0) {
        cplx_t<R> *so = (cplx_t<R> *)state_out + I.state_offset;
        if (resident) {
            for (int l = tid; l < T; l += NT) so[l] = t[l];
        } else {
            const uint64_t dim = 1ull << n;
            for (uint64_t i = tid; i < dim; i += NT) so[i] = slot[i];
        }
        __syncthreads();
    }

    // ---- measurement pass 0: tiles of the low tb qubits, index order ----
    const bool want_shots = I.num_shots > 0;
    double *pauli_out = obs + I.obs_offset + obs_size(n);
    for (int k = tid; k < I.num_pauli; k += NT) pauli_out[k] = 0.0;
    __syncthreads();
    for (uint64_t ti = 0; ti < ntiles; ti++) {
        const uint64_t base = ti << tb;
        if (!resident) {
            for (int l = tid; l < T; l += NT) t[l] = slot[base | l];
            __syncthreads();
        }
        double tt = 0.0;
        measure_tile<R>(sh, t, T, tb, n, base, I.flags, 0, &tt);
        if (!resident && tid == 0 && want_shots) tile_sums[ti] = tt;
        __syncthreads();
        if (I.num_pauli > 0)
            measure_pauli_tile<R>(sh, t, T, base, slot, terms + I.pauli_offset, I.num_pauli, pauli_out);
    }
    // ---- measurement pass 1: X on the qubits above the low tile ----
    if ((I.flags & BF_X) && !resident) {
        BPass P;
        P.nbits = tb;
        for (int b = 0; b < tb; b++) P.qubits[b] = (int8_t)(n - tb + b);
        build_deposit(sh, P);
        __syncthreads();
        for (uint64_t ti = 0; ti < ntiles; ti++) {
            const uint64_t base = tile_base(ti, P);
            for (int l = tid; l < T; l += NT) t[l] = slot[base | sh.dlo[l & 63] | sh.dhi[l >> 6]];
            __syncthreads();
            measure_tile<R>(sh, t, T, tb, n, base, BF_X, 1, nullptr);
        }
    }
    const int nv = 1 + n + (n > 1 ? n - 1 : 0) + n;
    for (int j = tid; j < nv; j += NT) obs[I.obs_offset + j] = sh.acc[j];

    // ---- shots: first index with u < cdf_i, else the last nonzero ----
    if (!want_shots) return;
    double *cdf = (double *)sh.tile;
    if (resident) {
        double pv[(1 << kBatchMaxTileBits) / NT > 0 ? (1 << kBatchMaxTileBits) / NT : 1];
        int k = 0;
        for (int l = tid; l < T; l += NT) pv[k++] = p2<R>(t[l]);
        if (tid == 0) sh.last_nz = -1;
        __syncthreads();
        k = 0;
        int mylast = -1;
        for (int l = tid; l < T; l += NT) {
            cdf[l] = pv[k];
            if (pv[k] > 0.0) mylast = l;
            k++;
        }
        if (mylast >= 0) atomicMax(&sh.last_nz, mylast);
        __syncthreads();
        lds_scan(sh, cdf, T, 0.0);
        const int fallback = sh.last_nz < 0 ? 0 : sh.last_nz;
        for (int s = tid; s < I.num_shots; s += NT) {
            const double u = draws[I.draw_offset + s];
            int lo = 0, hi = T;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (u < cdf[mid]) hi = mid;
                else lo = mid + 1;
            }
            outcomes[I.out_offset + s] = (uint64_t)(lo < T ? lo : fallback);
        }
        __syncthreads();
        return;
    }
    // global state: tile prefix in place over tile_sums[ntiles .. 2 ntiles)
    double *incl = tile_sums + ntiles;
    if (tid == 0) {
        double run = 0.0;
        for (uint64_t i = 0; i < ntiles; i++) {
            run += tile_sums[i];
            incl[i] = run;
        }
    }
    __syncthreads();
    // each shot's starting tile
    const int per_thread = (I.num_shots + NT - 1) / NT;
    for (int rep = 0; rep < per_thread; rep++) {
        const int s = tid + rep * NT;
        uint64_t tile = ntiles;
        double u = 2.0;
        if (s < I.num_shots) {
            u = draws[I.draw_offset + s];
            uint64_t lo = 0, hi = ntiles;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (u < incl[mid]) hi = mid;
                else lo = mid + 1;
            }
            tile = lo;
        }
        uint64_t result = UINT64_MAX;
        // walk tiles in order; every thread joins every step so barriers match
        for (uint64_t ti = 0; ti < ntiles; ti++) {
            if (tid == 0) sh.flag = 0;
            __syncthreads();
            if (tile == ti && result == UINT64_MAX) sh.flag = 1;
            __syncthreads();
            const int need = sh.flag;
            __syncthreads();   // every thread has read the flag before thread 0 clears it again
            if (!need) continue;
            for (int l = tid; l < T; l += NT) cdf[l] = p2<R>(slot[(ti << tb) | l]);
            __syncthreads();
            lds_scan(sh, cdf, T, ti ? incl[ti - 1] : 0.0);
            if (tile == ti && result == UINT64_MAX) {
                int lo = 0, hi = T;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (u < cdf[mid]) hi = mid;
                    else lo = mid + 1;
                }
                if (lo < T) result = (ti << tb) | (uint64_t)lo;
                else tile = ti + 1;   // rounding left u past this tile: continue in the next
            }
            __syncthreads();
        }
        if (s < I.num_shots && result == UINT64_MAX) {
            // u >= total mass: the last index with nonzero probability
            uint64_t last = 0;
            for (uint64_t ti = ntiles; ti-- > 0;) {
                if (tile_sums[ti] > 0.0) {
                    for (int l = T - 1; l >= 0; l--)
                        if (p2<R>(slot[(ti << tb) | l]) > 0.0) {
                            last = (ti << tb) | (uint64_t)l;
                            break;
                        }
                    break;
                }
            }
            result = last;
        }
        if (s < I.num_shots) outcomes[I.out_offset + s] = result;
        __syncthreads();
    }
}

__global__ __launch_bounds__(NT) void k_batch(const BInst *__restrict__ insts,
                                              const int32_t *__restrict__ order, int ninst,
                                              int *__restrict__ counter, const BPass *__restrict__ passes,
                                              const BOp *__restrict__ ops, double2 *__restrict__ scratch,
                                              uint64_t slot_amps, double *__restrict__ tile_scratch,
                                              uint64_t tile_slot, const double *__restrict__ draws,
                                              double *__restrict__ obs, uint64_t *__restrict__ outcomes,
                                              const PauliT *__restrict__ terms, void *__restrict__ state_out)
{
    __shared__ Shared sh;
    double2 *slot = scratch + (uint64_t)blockIdx.x * slot_amps;
    double *ts = tile_scratch + (uint64_t)blockIdx.x * tile_slot;
    for (;;) {
        if (threadIdx.x == 0) sh.next = atomicAdd(counter, 1);
        __syncthreads();
        // every lane read the same queue slot: make that explicit, so the
        // instance, its passes and every op descriptor below are fetched with
        // scalar (wave-uniform) loads instead of one vector load per lane
        const int k = __builtin_amdgcn_readfirstlane(sh.next);
        __syncthreads();
        if (k >= ninst) break;
        const BInst I = insts[order[k]];
        if (I.prec == 1)
            run_circuit<float>(sh, I, passes, ops, slot, ts, draws, obs, outcomes, terms, state_out);
        else
            run_circuit<double>(sh, I, passes, ops, slot, ts, draws, obs, outcomes, terms, state_out);
        __syncthreads();
    }
}

__global__ void k_reset_counter(int *c) { *c = 0; }

// One cache-blocked pass over a large state: each workgroup streams tiles of
// 2^P.nbits amplitudes through LDS and applies the pass's ops there.  The
// next tile's loads are issued before the current tile's ops run, and the
// barriers order LDS only, so loads, LDS work and the previous tile's stores
// overlap.  NT: streaming accesses (runs are >= 16 amplitudes, whole lines).
template <typename R, bool NT, int THREADS, int TB>
__global__ __launch_bounds__(THREADS) void k_tile_pass(cplx_t<R> *__restrict__ a, BPass P,
                                                       const BOp *__restrict__ ops, uint64_t ntiles)
{
    constexpr int PER = (1 << TB) / THREADS;
    static_assert(PER >= 1 && (1 << TB) / 64 <= 64, "tile shape");
    __shared__ cplx_t<R> t[1 << TB];
    __shared__ uint64_t dlo[64], dhi[64];
    const int tid = threadIdx.x;
    const int T = 1 << P.nbits;   // P.nbits <= TB (checked by the launcher)
    if (tid < 128) {
        const int j = tid & 63, hi = tid >> 6;
        uint64_t o = 0;
        for (int b = 0; b < 6; b++) {
            const int qi = hi * 6 + b;
            if (qi < P.nbits && ((j >> b) & 1)) o |= 1ull << P.qubits[qi];
        }
        if (hi) dhi[j] = o;
        else dlo[j] = o;
    }
    __syncthreads();
    // thread tid moves tile amplitudes l = tid + THREADS i (l % 64 = tid % 64)
    const uint64_t lo = dlo[tid & 63];
    const int hi0 = tid >> 6;
    constexpr int HSTEP = THREADS / 64;
    cplx_t<R> pre[PER];
    uint64_t ti = blockIdx.x;
    if (ti < ntiles) {
        const uint64_t base = tile_base(ti, P) | lo;
#pragma unroll
        for (int i = 0; i < PER; i++)
            if (tid + THREADS * i < T) pre[i] = ld<NT>(&a[base | dhi[hi0 + HSTEP * i]]);
    }
    for (; ti < ntiles; ti += gridDim.x) {
#pragma unroll
        for (int i = 0; i < PER; i++)
            if (tid + THREADS * i < T) t[tid + THREADS * i] = pre[i];
        lds_barrier();
        const uint64_t nti = ti + gridDim.x;
        if (nti < ntiles) {
            const uint64_t nbase = tile_base(nti, P) | lo;
#pragma unroll
            for (int i = 0; i < PER; i++)
                if (tid + THREADS * i < T) pre[i] = ld<NT>(&a[nbase | dhi[hi0 + HSTEP * i]]);
        }
        const uint64_t base = tile_base(ti, P);
        run_pass_ops<R, THREADS, kSegRegsPass, true>(t, T, ops + P.first_op, P.num_ops, base);
#pragma unroll
        for (int i = 0; i < PER; i++)
            if (tid + THREADS * i < T) st<NT>(&a[(base | lo) | dhi[hi0 + HSTEP * i]], t[tid + THREADS * i]);
        lds_barrier();   // every thread has read its write-back values before the next tile lands
    }
}

// Persistent grid of one resident wave of workgroups for this instantiation.
template <typename R, bool NT, int THREADS, int TB>
hipError_t run_tile_pass(void *amps, const BPass &P, const BOp *d_ops, uint64_t ntiles, int cus, hipStream_t st)
{
    static std::atomic<int> per_cu_cached{0};
    int per_cu = per_cu_cached.load(std::memory_order_relaxed);
    if (!per_cu) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tile_pass<R, NT, THREADS, TB>, THREADS, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        per_cu_cached.store(per_cu, std::memory_order_relaxed);
    }
    const uint64_t want = (uint64_t)cus * (uint64_t)per_cu;
    const unsigned grid = (unsigned)(ntiles < want ? ntiles : want);
    k_tile_pass<R, NT, THREADS, TB><<<grid, THREADS, 0, st>>>((cplx_t<R> *)amps, P, d_ops, ntiles);
    return hipGetLastError();
}

template <typename R, bool NT>
hipError_t tile_pass_shape(void *amps, const BPass &P, const BOp *d_ops, uint64_t ntiles, int cus, int threads,
                           int tb, hipStream_t st)
{
    if (tb <= 11) {
        if (threads <= 256) return run_tile_pass<R, NT, 256, 11>(amps, P, d_ops, ntiles, cus, st);
        return run_tile_pass<R, NT, 512, 11>(amps, P, d_ops, ntiles, cus, st);
    }
    if (threads <= 256) return run_tile_pass<R, NT, 256, 12>(amps, P, d_ops, ntiles, cus, st);
    return run_tile_pass<R, NT, 512, 12>(amps, P, d_ops, ntiles, cus, st);
}

}  // namespace

// One pass of the single-state blocked path; tile_bits (11 or 12) is the
// LDS tile the kernel is built for (>= P.nbits), threads its workgroup size.
hipError_t launch_tile_pass(void *amps, int prec, int n, const BPass &P, const BOp *d_ops, int cus, int threads,
                            int tile_bits, hipStream_t st)
{
    if (P.nbits > tile_bits || tile_bits > kBatchMaxTileBits) return hipErrorInvalidValue;
    const uint64_t ntiles = 1ull << (n - P.nbits);
    const bool nt = ((1ull << n) * (prec == 1 ? 8ull : 16ull)) >= gate_tuning().nt_min_bytes;
    if (prec == 1)
        return nt ? tile_pass_shape<float, true>(amps, P, d_ops, ntiles, cus, threads, tile_bits, st)
                  : tile_pass_shape<float, false>(amps, P, d_ops, ntiles, cus, threads, tile_bits, st);
    return nt ? tile_pass_shape<double, true>(amps, P, d_ops, ntiles, cus, threads, tile_bits, st)
              : tile_pass_shape<double, false>(amps, P, d_ops, ntiles, cus, threads, tile_bits, st);
}

hipError_t launch_batch(const BInst *insts, const int32_t *order, int ninst, int *counter, const BPass *passes,
                        const BOp *ops, double2 *scratch, uint64_t slot_amps, double *tile_scratch,
                        uint64_t tile_slot, const double *draws, double *obs, uint64_t *outcomes,
                        const PauliT *terms, void *state_out, int grid, hipStream_t st)
{
    k_reset_counter<<<1, 1, 0, st>>>(counter);
    k_batch<<<grid, NT, 0, st>>>(insts, order, ninst, counter, passes, ops, scratch, slot_amps, tile_scratch,
                                 tile_slot, draws, obs, outcomes, terms, state_out);
    return hipGetLastError();
}

}  // namespace mlx
