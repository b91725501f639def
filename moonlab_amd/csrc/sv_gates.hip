// sv_gates.hip -- gfx950 gate kernels for the dense state vector.
//
// Every kernel is a streaming pass over HBM: each thread owns E pairs (or
// elements, or quadruples) spaced one block apart, issues all of its loads
// before any arithmetic, then stores.  Consecutive lanes take consecutive
// pair indices, so for any target qubit each wave reads whole contiguous
// runs: for q >= 2 two runs of >= 64 B per 64-lane wave (the |0> and |1>
// halves), for q < 2 one run that interleaves both halves.  Only the
// amplitudes a gate changes are touched: diagonal gates on one half (S, T,
// Z) read and write D/2 amplitudes, controlled gates enumerate only the
// control = 1 subspace (D/2^(k+1) pairs for k controls) by depositing the
// sub-index around the fixed bits, so no lane idles on a failed control test.
//
// Reference counterparts: src/backends/cuda_statevec.cu:81-390 (one thread
// per pair, half the CNOT threads idle, 2q matrix in a __constant__ symbol)
// and the CPU stride-block loops of src/quantum/gates.c.
#include "sv_device.h"
#include "sv_internal.h"

#include <math.h>

namespace mlx {

GateTuning &gate_tuning()
{
    // One pair per thread and 256-thread blocks measured fastest on MI355X for
    // every gate family at 30 qubits (2, 4, 8 pairs per thread: 2-10 % slower).
    // Blocked path: tile bits 0 = by precision (12 for complex128, 13 for
    // complex64), 512-thread LDS-engine workgroups, register-phase engine,
    // lookahead from 8 seed ops per pass and phase.
    static GateTuning t = {1, 256, 512ull << 20, 0, 512, 0, 8};
    return t;
}

namespace {

constexpr int kMaxBlock = 1024;

// Workgroup index of a 1-D or 2-D grid (grid_for), remapped so that each of
// the 8 XCDs streams one contiguous eighth of the launch: the dispatcher hands
// consecutive workgroups to the XCDs in turn, so workgroup b runs on XCD
// b mod 8 and takes chunk (b mod 8) * (nb / 8) + b / 8.  Measured on 30-qubit
// complex128 gates, this beats the plain order (all XCDs sharing one window)
// for most target qubits: 5.85 -> 6.0 TB/s averaged over the 1-qubit targets.
// A grid whose block count is not a multiple of 8 keeps the plain order.
__device__ __forceinline__ uint64_t block_id()
{
    const uint64_t nb = (uint64_t)gridDim.x * gridDim.y;
    const uint64_t b = (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
    if (nb & 7ull) return b;
    return (b & 7ull) * (nb >> 3) + (b >> 3);
}

// --- pair kernels -----------------------------------------------------------

template <typename R, typename Op, int E, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_pair(cplx_t<R> *__restrict__ a, uint64_t npairs,
                                                    int q, Op op)
{
    const uint64_t bs = blockDim.x;
    const uint64_t base = block_id() * bs * E + threadIdx.x;
    const uint64_t stride = 1ull << q;
    if (base + (uint64_t)(E - 1) * bs < npairs) {
        uint64_t i0[E];
        cplx_t<R> x0[E], x1[E];
#pragma unroll
        for (int k = 0; k < E; k++) {
            i0[k] = ins0(base + (uint64_t)k * bs, q);
            x0[k] = ld<NT>(&a[i0[k]]);
            x1[k] = ld<NT>(&a[i0[k] + stride]);
        }
#pragma unroll
        for (int k = 0; k < E; k++) {
            op(x0[k], x1[k]);
            st<NT>(&a[i0[k]], x0[k]);
            st<NT>(&a[i0[k] + stride], x1[k]);
        }
    } else {
        for (int k = 0; k < E; k++) {
            const uint64_t p = base + (uint64_t)k * bs;
            if (p >= npairs) break;
            const uint64_t j = ins0(p, q);
            cplx_t<R> x0 = ld<NT>(&a[j]), x1 = ld<NT>(&a[j + stride]);
            op(x0, x1);
            st<NT>(&a[j], x0);
            st<NT>(&a[j + stride], x1);
        }
    }
}

// Pair kernel for a target inside the wave (q < 6): every lane loads ONE
// amplitude of a contiguous run -- whole 128-byte lines per load, the access
// pattern of k_diag -- and takes its partner (lane ^ 2^q) through a cross-lane
// shuffle.  Both lanes of a pair run the same operator on the same (a0, a1)
// and each keeps its own half, so the results equal k_pair's bit for bit.
// Launched with dim a multiple of the block size (n >= 8), so every wave is
// full.

template <typename R, typename Op, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_pair_lane(cplx_t<R> *__restrict__ a, uint64_t dim, int q,
                                                         Op op)
{
    const uint64_t i = block_id() * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    const cplx_t<R> x = ld<NT>(&a[i]);
    cplx_t<R> y;
    y.x = __shfl_xor(x.x, 1 << q, 64);
    y.y = __shfl_xor(x.y, 1 << q, 64);
    const bool one = (i >> q) & 1ull;
    cplx_t<R> x0 = csel(one, y, x), x1 = csel(one, x, y);
    op(x0, x1);
    st<NT>(&a[i], csel(one, x1, x0));
}

// Pair kernel with one amplitude per thread for pairs that leave the wave
// (target >= 6): a workgroup of 2H threads takes H consecutive pairs of the
// enumeration p -> (i0, i0 | tbit), i0 = deposit(p, fixed) | cmask; threads
// [0, H) load the |0> amplitudes (contiguous runs), threads [H, 2H) the |1>
// amplitudes, the partner comes through LDS, each thread stores its own.  It
// moves the same bytes as a thread holding both amplitudes but with twice
// the loads in flight: 6.3-6.8 TB/s against 5.6-6.6 at 30 qubits
// complex128, every target.  Both lanes of a pair run the same operator on
// the same (a0, a1), so results equal k_pair's bit for bit.  Launched with
// npairs a multiple of H.
constexpr int kSplitPairs = 128;

template <typename R, typename Op, bool NT>
__global__ __launch_bounds__(2 * kSplitPairs) void k_pair_split(cplx_t<R> *__restrict__ a, uint64_t npairs,
                                                                BitSet fixed, uint64_t cmask, uint64_t tbit, Op op)
{
    __shared__ cplx_t<R> lds[2 * kSplitPairs];
    const int tid = threadIdx.x;
    const bool hi = tid >= kSplitPairs;
    const uint64_t p = block_id() * kSplitPairs + (uint64_t)(tid - (hi ? kSplitPairs : 0));
    if (p >= npairs) return;   // whole workgroups only: npairs is a multiple of kSplitPairs
    const uint64_t i = (deposit(p, fixed) | cmask) | (hi ? tbit : 0ull);
    const cplx_t<R> x = ld<NT>(&a[i]);
    lds[tid] = x;
    __syncthreads();
    const cplx_t<R> y = lds[tid ^ kSplitPairs];
    cplx_t<R> x0 = csel(hi, y, x), x1 = csel(hi, x, y);
    op(x0, x1);
    st<NT>(&a[i], csel(hi, x1, x0));
}

// Controlled pair whose touched runs are shorter than a 128-byte line (lowest
// control bit < 3 for complex128, < 4 for complex64): a kernel that reads only
// the control = 1 amplitudes still moves whole lines, in twice the requests.
// These read every amplitude with whole-line loads, apply the operator where
// every control bit is 1, and store either every lane (STORE_ALL: whole
// lines, the unchanged half written back as read) or only the changed lanes
// (runs of >= 64 bytes).  Target inside the wave: partner by shuffle.
template <typename R, typename Op, bool STORE_ALL>
__global__ __launch_bounds__(kMaxBlock) void k_ctrl_lane(cplx_t<R> *__restrict__ a, uint64_t dim,
                                                         uint64_t cmask, int t, Op op)
{
    const uint64_t i = block_id() * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    const cplx_t<R> x = ld<true>(&a[i]);
    cplx_t<R> y;
    y.x = __shfl_xor(x.x, 1 << t, 64);
    y.y = __shfl_xor(x.y, 1 << t, 64);
    const bool on = (i & cmask) == cmask;
    const bool one = (i >> t) & 1ull;
    cplx_t<R> x0 = csel(one, y, x), x1 = csel(one, x, y);
    op(x0, x1);
    const cplx_t<R> r = csel(on, csel(one, x1, x0), x);
    if (STORE_ALL || on) st<true>(&a[i], r);
}

// The same with the target outside the wave (t >= 6): pairs over the target
// bit only, both lines whole.
template <typename R, typename Op, bool STORE_ALL>
__global__ __launch_bounds__(kMaxBlock) void k_ctrl_lines(cplx_t<R> *__restrict__ a, uint64_t npairs,
                                                          uint64_t cmask, int t, Op op)
{
    const uint64_t p = block_id() * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const uint64_t i0 = ins0(p, t), i1 = i0 | (1ull << t);
    cplx_t<R> x0 = ld<true>(&a[i0]), x1 = ld<true>(&a[i1]);
    const bool on = (i0 & cmask) == cmask;
    if (on) op(x0, x1);
    if (STORE_ALL || on) {
        st<true>(&a[i0], x0);
        st<true>(&a[i1], x1);
    }
}

// Element op where every bit of `mask` is 1, over whole lines (phase-on-one
// and CZ-type gates whose touched runs are shorter than a line).
template <typename R, typename Op, bool STORE_ALL>
__global__ __launch_bounds__(kMaxBlock) void k_elem_lines(cplx_t<R> *__restrict__ a, uint64_t dim,
                                                          uint64_t mask, Op op)
{
    const uint64_t i = block_id() * blockDim.x + threadIdx.x;
    if (i >= dim) return;
    cplx_t<R> x = ld<true>(&a[i]);
    const bool on = (i & mask) == mask;
    if (on) op(x);
    if (STORE_ALL || on) st<true>(&a[i], x);
}

// Only the |1> half of each pair (phase-on-one gates).
template <typename R, typename Op, int E, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_half(cplx_t<R> *__restrict__ a, uint64_t nhalf,
                                                    int q, Op op)
{
    const uint64_t bs = blockDim.x;
    const uint64_t base = block_id() * bs * E + threadIdx.x;
    const uint64_t bit = 1ull << q;
    if (base + (uint64_t)(E - 1) * bs < nhalf) {
        uint64_t i[E];
        cplx_t<R> x[E];
#pragma unroll
        for (int k = 0; k < E; k++) {
            i[k] = ins0(base + (uint64_t)k * bs, q) | bit;
            x[k] = ld<NT>(&a[i[k]]);
        }
#pragma unroll
        for (int k = 0; k < E; k++) {
            op(x[k]);
            st<NT>(&a[i[k]], x[k]);
        }
    } else {
        for (int k = 0; k < E; k++) {
            const uint64_t p = base + (uint64_t)k * bs;
            if (p >= nhalf) break;
            const uint64_t j = ins0(p, q) | bit;
            cplx_t<R> x = ld<NT>(&a[j]);
            op(x);
            st<NT>(&a[j], x);
        }
    }
}

// diag(d0, d1) on qubit q over every amplitude: fully contiguous stream.
template <typename R, int E, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_diag(cplx_t<R> *__restrict__ a, uint64_t dim,
                                                    int q, R d0r, R d0i, R d1r, R d1i)
{
    const uint64_t bs = blockDim.x;
    const uint64_t base = block_id() * bs * E + threadIdx.x;
    const bool full = base + (uint64_t)(E - 1) * bs < dim;
    cplx_t<R> x[E];
#pragma unroll
    for (int k = 0; k < E; k++) {
        const uint64_t i = base + (uint64_t)k * bs;
        if (full || i < dim) x[k] = ld<NT>(&a[i]);
    }
#pragma unroll
    for (int k = 0; k < E; k++) {
        const uint64_t i = base + (uint64_t)k * bs;
        if (!(full || i < dim)) continue;
        const bool one = (i >> q) & 1ull;
        st<NT>(&a[i], cmul(x[k], one ? d1r : d0r, one ? d1i : d0i));
    }
}

// Controlled pair: indices with every bit of `fixed` zero are enumerated,
// then the control bits are set; the pair is (b, b | tbit).
template <typename R, typename Op, int E, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_ctrl_pair(cplx_t<R> *__restrict__ a, uint64_t nsub,
                                                         BitSet fixed, uint64_t cmask,
                                                         uint64_t tbit, Op op)
{
    const uint64_t bs = blockDim.x;
    const uint64_t base = block_id() * bs * E + threadIdx.x;
    if (base + (uint64_t)(E - 1) * bs < nsub) {
        uint64_t i0[E];
        cplx_t<R> x0[E], x1[E];
#pragma unroll
        for (int k = 0; k < E; k++) {
            i0[k] = deposit(base + (uint64_t)k * bs, fixed) | cmask;
            x0[k] = ld<NT>(&a[i0[k]]);
            x1[k] = ld<NT>(&a[i0[k] | tbit]);
        }
#pragma unroll
        for (int k = 0; k < E; k++) {
            op(x0[k], x1[k]);
            st<NT>(&a[i0[k]], x0[k]);
            st<NT>(&a[i0[k] | tbit], x1[k]);
        }
    } else {
        for (int k = 0; k < E; k++) {
            const uint64_t p = base + (uint64_t)k * bs;
            if (p >= nsub) break;
            const uint64_t j = deposit(p, fixed) | cmask;
            cplx_t<R> x0 = ld<NT>(&a[j]), x1 = ld<NT>(&a[j | tbit]);
            op(x0, x1);
            st<NT>(&a[j], x0);
            st<NT>(&a[j | tbit], x1);
        }
    }
}

// Element op where every bit of `mask` is 1 (CZ, CPHASE, MCZ).
template <typename R, typename Op, int E, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_mask_elem(cplx_t<R> *__restrict__ a, uint64_t nsub,
                                                         BitSet fixed, uint64_t mask, Op op)
{
    const uint64_t bs = blockDim.x;
    const uint64_t base = block_id() * bs * E + threadIdx.x;
    if (base + (uint64_t)(E - 1) * bs < nsub) {
        uint64_t i[E];
        cplx_t<R> x[E];
#pragma unroll
        for (int k = 0; k < E; k++) {
            i[k] = deposit(base + (uint64_t)k * bs, fixed) | mask;
            x[k] = ld<NT>(&a[i[k]]);
        }
#pragma unroll
        for (int k = 0; k < E; k++) {
            op(x[k]);
            st<NT>(&a[i[k]], x[k]);
        }
    } else {
        for (int k = 0; k < E; k++) {
            const uint64_t p = base + (uint64_t)k * bs;
            if (p >= nsub) break;
            const uint64_t j = deposit(p, fixed) | mask;
            cplx_t<R> x = ld<NT>(&a[j]);
            op(x);
            st<NT>(&a[j], x);
        }
    }
}

// Swap (base|cmask|b1) <-> (base|cmask|b2): SWAP and Fredkin.
template <typename R, int E, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_swap_bits(cplx_t<R> *__restrict__ a, uint64_t nsub,
                                                         BitSet fixed, uint64_t cmask,
                                                         uint64_t b1, uint64_t b2)
{
    const uint64_t bs = blockDim.x;
    const uint64_t base = block_id() * bs * E + threadIdx.x;
#pragma unroll
    for (int k = 0; k < E; k++) {
        const uint64_t p = base + (uint64_t)k * bs;
        if (p >= nsub) break;
        const uint64_t j = deposit(p, fixed) | cmask;
        const cplx_t<R> x0 = ld<NT>(&a[j | b1]), x1 = ld<NT>(&a[j | b2]);
        st<NT>(&a[j | b1], x1);
        st<NT>(&a[j | b2], x0);
    }
}

template <typename R> struct Mat4 {
    R m[32];
};

// 4x4 complex matvec on the quadruple (b, b|b0, b|b1, b|b0|b1), basis |q1 q0>.
template <typename R, int E, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_mat4(cplx_t<R> *__restrict__ a, uint64_t nquad,
                                                    BitSet fixed, uint64_t b0, uint64_t b1,
                                                    Mat4<R> M)
{
    const uint64_t bs = blockDim.x;
    const uint64_t base = block_id() * bs * E + threadIdx.x;
#pragma unroll
    for (int k = 0; k < E; k++) {
        const uint64_t p = base + (uint64_t)k * bs;
        if (p >= nquad) break;
        const uint64_t j = deposit(p, fixed);
        const uint64_t idx[4] = {j, j | b0, j | b1, j | b0 | b1};
        cplx_t<R> v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) v[c] = ld<NT>(&a[idx[c]]);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            R sr = 0, si = 0;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const R mr = M.m[r * 8 + c * 2], mi = M.m[r * 8 + c * 2 + 1];
                sr = fmar(mr, v[c].x, fmar(-mi, v[c].y, sr));
                si = fmar(mr, v[c].y, fmar(mi, v[c].x, si));
            }
            st<NT>(&a[idx[r]], cplx_t<R>{sr, si});
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_init_zero(cplx_t<R> *__restrict__ a, uint64_t dim)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride)
        a[i] = cplx_t<R>{i == 0 ? R(1) : R(0), R(0)};
}

template <typename R>
__global__ __launch_bounds__(256) void k_scale(cplx_t<R> *__restrict__ a, uint64_t dim, R f)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        cplx_t<R> x = a[i];
        x.x *= f;
        x.y *= f;
        a[i] = x;
    }
}

template <typename RS, typename RD>
__global__ __launch_bounds__(256) void k_convert(const cplx_t<RS> *__restrict__ s,
                                                 cplx_t<RD> *__restrict__ d, uint64_t count)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        d[i] = cplx_t<RD>{(RD)s[i].x, (RD)s[i].y};
}

// --- launch helpers -----------------------------------------------------------

// One-shot launch shape for `work` items, e per thread.  A launch holds at
// most 2^32 - 1 work-items per grid dimension, which 1-qubit gates on 33+
// qubits (complex128) exceed; such grids go 2-D, 65536 blocks per row, and
// block_id() linearizes them (blocks past the end fail the kernels' bounds
// checks).
inline dim3 grid_for(uint64_t work, int block, int e)
{
    const uint64_t per = (uint64_t)block * (uint64_t)e;
    const uint64_t blocks = (work + per - 1) / per;
    if (blocks * (uint64_t)block < (1ull << 31)) return dim3((unsigned)blocks, 1, 1);
    constexpr uint64_t kRow = 65536;
    return dim3((unsigned)kRow, (unsigned)((blocks + kRow - 1) / kRow), 1);
}

inline unsigned grid_stride_blocks(uint64_t work)
{
    const uint64_t b = (work + 255) / 256;
    return (unsigned)(b < 65536 ? (b ? b : 1) : 65536);
}

// Instantiate a launch for the tuned pairs-per-thread E and for streaming
// (NT) or cached access; NT is chosen when the state exceeds the
// last-level cache.
#define MLX_DISPATCH_E1(EV, NTV, ...)                                   \
    switch (EV) {                                                       \
    case 1: { constexpr int E = 1; constexpr bool NT = NTV; __VA_ARGS__; } break; \
    case 2: { constexpr int E = 2; constexpr bool NT = NTV; __VA_ARGS__; } break; \
    case 8: { constexpr int E = 8; constexpr bool NT = NTV; __VA_ARGS__; } break; \
    default: { constexpr int E = 4; constexpr bool NT = NTV; __VA_ARGS__; } break; \
    }
#define MLX_DISPATCH_E(EV, NTF, ...)                \
    do {                                            \
        if (NTF) { MLX_DISPATCH_E1(EV, true, __VA_ARGS__) } \
        else { MLX_DISPATCH_E1(EV, false, __VA_ARGS__) }    \
    } while (0)

// Streaming access pays only when the state is larger than the last-level
// cache AND every 128-byte line the kernel touches is written whole:
// non-temporal stores of partial lines (S on qubit 0, CNOT with control 0)
// measured 30-60 % slower than cached ones.  `restrict_bits` are the bits
// that select which amplitudes a launch touches (controls, the |1> bit of a
// phase-on-one gate); the touched runs are 2^ctz(restrict_bits) long.
inline bool use_nt(const SvDev &s, uint64_t restrict_bits = 0)
{
    const uint64_t ab = s.prec == MOONLAB_HIP_C64 ? 8ull : 16ull;
    if (((1ull << s.n) * ab) < gate_tuning().nt_min_bytes) return false;
    if (!restrict_bits) return true;
    return (1ull << __builtin_ctzll(restrict_bits)) * ab >= 128;
}

// Block size of the whole-line kernels (a multiple of the 64-lane wave).
constexpr int kLaneBlock = 256;

// A large state whose touched runs (2^ctz(restrict_bits) amplitudes) are
// shorter than a 128-byte line: the whole-line kernels take it.
inline bool partial_lines(const SvDev &s, uint64_t restrict_bits)
{
    const uint64_t ab = s.prec == MOONLAB_HIP_C64 ? 8ull : 16ull;
    if (s.n < 8 || !restrict_bits || ((1ull << s.n) * ab) < gate_tuning().nt_min_bytes) return false;
    return (1ull << __builtin_ctzll(restrict_bits)) * ab < 128;
}

// Only changed lanes are stored when they form runs of >= 64 bytes; shorter
// runs store whole lines (partial streaming stores of 16-32 byte runs
// measured 1.5x slower than writing the line back).
inline bool store_changed_only(const SvDev &s, uint64_t restrict_bits)
{
    const uint64_t ab = s.prec == MOONLAB_HIP_C64 ? 8ull : 16ull;
    return (1ull << __builtin_ctzll(restrict_bits)) * ab >= 64;
}

inline BitSet make_bitset(uint64_t bits)
{
    BitSet b;
    b.nb = 0;
    for (int i = 0; i < 64 && b.nb < 7; i++)
        if (bits & (1ull << i)) b.pos[b.nb++] = i;
    return b;
}

// The split pair kernel over npairs pairs (a multiple of kSplitPairs).
template <typename R, typename Op>
hipError_t run_split(const SvDev &s, const BitSet &fixed, uint64_t cmask, uint64_t tbit, uint64_t npairs, Op op)
{
    auto *a = (cplx_t<R> *)s.amps;
    const dim3 gp = grid_for(npairs * 2, 2 * kSplitPairs, 1);   // two threads per pair
    if (use_nt(s))
        k_pair_split<R, Op, true><<<gp, 2 * kSplitPairs, 0, s.stream>>>(a, npairs, fixed, cmask, tbit, op);
    else
        k_pair_split<R, Op, false><<<gp, 2 * kSplitPairs, 0, s.stream>>>(a, npairs, fixed, cmask, tbit, op);
    return hipGetLastError();
}

template <typename R, typename Op>
hipError_t run_pair(const SvDev &s, int q, Op op)
{
    const uint64_t np = (1ull << s.n) >> 1;
    const GateTuning t = gate_tuning();
    auto *a = (cplx_t<R> *)s.amps;
    if (q < 6 && s.n >= 8) {
        const uint64_t dim = 1ull << s.n;
        if (use_nt(s))
            k_pair_lane<R, Op, true><<<grid_for(dim, kLaneBlock, 1), kLaneBlock, 0, s.stream>>>(a, dim, q, op);
        else
            k_pair_lane<R, Op, false><<<grid_for(dim, kLaneBlock, 1), kLaneBlock, 0, s.stream>>>(a, dim, q, op);
        return hipGetLastError();
    }
    if (s.n >= 8) return run_split<R>(s, make_bitset(1ull << q), 0ull, 1ull << q, np, op);
    MLX_DISPATCH_E(t.ept, use_nt(s), k_pair<R, Op, E, NT><<<grid_for(np, t.block, E), t.block, 0, s.stream>>>(a, np, q, op));
    return hipGetLastError();
}

template <typename R, typename Op>
hipError_t run_elem_lines(const SvDev &s, uint64_t mask, Op op)
{
    const uint64_t dim = 1ull << s.n;
    auto *a = (cplx_t<R> *)s.amps;
    if (store_changed_only(s, mask))
        k_elem_lines<R, Op, false><<<grid_for(dim, kLaneBlock, 1), kLaneBlock, 0, s.stream>>>(a, dim, mask, op);
    else
        k_elem_lines<R, Op, true><<<grid_for(dim, kLaneBlock, 1), kLaneBlock, 0, s.stream>>>(a, dim, mask, op);
    return hipGetLastError();
}

template <typename R, typename Op>
hipError_t run_half(const SvDev &s, int q, Op op)
{
    if (partial_lines(s, 1ull << q)) return run_elem_lines<R>(s, 1ull << q, op);
    const uint64_t nh = (1ull << s.n) >> 1;
    const GateTuning t = gate_tuning();
    auto *a = (cplx_t<R> *)s.amps;
    MLX_DISPATCH_E(t.ept, use_nt(s, 1ull << q), k_half<R, Op, E, NT><<<grid_for(nh, t.block, E), t.block, 0, s.stream>>>(a, nh, q, op));
    return hipGetLastError();
}

template <typename R>
hipError_t run_diag(const SvDev &s, int q, const double d[4])
{
    const uint64_t dim = 1ull << s.n;
    const GateTuning t = gate_tuning();
    auto *a = (cplx_t<R> *)s.amps;
    MLX_DISPATCH_E(t.ept, use_nt(s), k_diag<R, E, NT><<<grid_for(dim, t.block, E), t.block, 0, s.stream>>>(
                              a, dim, q, (R)d[0], (R)d[1], (R)d[2], (R)d[3]));
    return hipGetLastError();
}


template <typename R, typename Op>
hipError_t run_ctrl_pair(const SvDev &s, uint64_t cmask, int target, Op op)
{
    if (partial_lines(s, cmask)) {
        auto *a = (cplx_t<R> *)s.amps;
        const bool all = !store_changed_only(s, cmask);
        if (target < 6) {
            const uint64_t dim = 1ull << s.n;
            const dim3 g = grid_for(dim, kLaneBlock, 1);
            if (all) k_ctrl_lane<R, Op, true><<<g, kLaneBlock, 0, s.stream>>>(a, dim, cmask, target, op);
            else k_ctrl_lane<R, Op, false><<<g, kLaneBlock, 0, s.stream>>>(a, dim, cmask, target, op);
        } else {
            const uint64_t np = (1ull << s.n) >> 1;
            const dim3 g = grid_for(np, kLaneBlock, 1);
            if (all) k_ctrl_lines<R, Op, true><<<g, kLaneBlock, 0, s.stream>>>(a, np, cmask, target, op);
            else k_ctrl_lines<R, Op, false><<<g, kLaneBlock, 0, s.stream>>>(a, np, cmask, target, op);
        }
        return hipGetLastError();
    }
    const uint64_t tbit = 1ull << target;
    const BitSet fixed = make_bitset(cmask | tbit);
    const uint64_t nsub = (1ull << s.n) >> fixed.nb;
    // a whole number of workgroups, and on a streamed state whole lines
    // (every fixed bit above the line: the control=1 runs are >= 128 bytes)
    if (nsub % kSplitPairs == 0 && (!use_nt(s) || use_nt(s, cmask | tbit)))
        return run_split<R>(s, fixed, cmask, tbit, nsub, op);
    const GateTuning t = gate_tuning();
    auto *a = (cplx_t<R> *)s.amps;
    MLX_DISPATCH_E(t.ept, use_nt(s, cmask), k_ctrl_pair<R, Op, E, NT><<<grid_for(nsub, t.block, E), t.block, 0, s.stream>>>(
                              a, nsub, fixed, cmask, tbit, op));
    return hipGetLastError();
}

template <typename R, typename Op>
hipError_t run_mask_elem(const SvDev &s, uint64_t mask, Op op)
{
    if (partial_lines(s, mask)) return run_elem_lines<R>(s, mask, op);
    const BitSet fixed = make_bitset(mask);
    const uint64_t nsub = (1ull << s.n) >> fixed.nb;
    const GateTuning t = gate_tuning();
    auto *a = (cplx_t<R> *)s.amps;
    MLX_DISPATCH_E(t.ept, use_nt(s, mask), k_mask_elem<R, Op, E, NT><<<grid_for(nsub, t.block, E), t.block, 0, s.stream>>>(
                              a, nsub, fixed, mask, op));
    return hipGetLastError();
}

template <typename R>
hipError_t run_swap(const SvDev &s, uint64_t cmask, int q1, int q2)
{
    const uint64_t b1 = 1ull << q1, b2 = 1ull << q2;
    const BitSet fixed = make_bitset(cmask | b1 | b2);
    const uint64_t nsub = (1ull << s.n) >> fixed.nb;
    const GateTuning t = gate_tuning();
    auto *a = (cplx_t<R> *)s.amps;
    MLX_DISPATCH_E(t.ept, use_nt(s, cmask | b1 | b2), k_swap_bits<R, E, NT><<<grid_for(nsub, t.block, E), t.block, 0, s.stream>>>(
                              a, nsub, fixed, cmask, b1, b2));
    return hipGetLastError();
}

template <typename R>
hipError_t run_mat4(const SvDev &s, int q0, int q1, const double m[32])
{
    const uint64_t b0 = 1ull << q0, b1 = 1ull << q1;
    const BitSet fixed = make_bitset(b0 | b1);
    const uint64_t nq = (1ull << s.n) >> 2;
    Mat4<R> M;
    for (int i = 0; i < 32; i++) M.m[i] = (R)m[i];
    const GateTuning t = gate_tuning();
    auto *a = (cplx_t<R> *)s.amps;
    MLX_DISPATCH_E(t.ept, use_nt(s), k_mat4<R, E, NT><<<grid_for(nq, t.block, E), t.block, 0, s.stream>>>(
                              a, nq, fixed, b0, b1, M));
    return hipGetLastError();
}

template <typename R> OpMat2<R> mat2(const double m[8])
{
    OpMat2<R> o;
    for (int i = 0; i < 8; i++) o.m[i] = (R)m[i];
    return o;
}

// One gate of the code table (include/moonlab_hip_statevec.h) at precision R.
template <typename R>
hipError_t gate_r(const SvDev &s, const moonlab_hip_gate_t &g)
{
    const int t = g.target, c = g.control, e = g.extra;
    switch (g.type) {
    case MOONLAB_HIP_GATE_I: return hipSuccess;
    case MOONLAB_HIP_GATE_X: return run_pair<R>(s, t, OpSwap<R>{});
    case MOONLAB_HIP_GATE_Y: return run_pair<R>(s, t, OpY<R>{});
    case MOONLAB_HIP_GATE_Z: return run_half<R>(s, t, ElNeg<R>{});
    case MOONLAB_HIP_GATE_H: return run_pair<R>(s, t, OpH<R>{(R)0.70710678118654752440});
    case MOONLAB_HIP_GATE_S: return run_half<R>(s, t, ElMulI<R>{});
    case MOONLAB_HIP_GATE_SDG: return run_half<R>(s, t, ElMulNegI<R>{});
    case MOONLAB_HIP_GATE_T:
        return run_half<R>(s, t, ElPhase<R>{(R)cos(M_PI / 4.0), (R)sin(M_PI / 4.0)});
    case MOONLAB_HIP_GATE_TDG:
        return run_half<R>(s, t, ElPhase<R>{(R)cos(-M_PI / 4.0), (R)sin(-M_PI / 4.0)});
    case MOONLAB_HIP_GATE_PHASE:
        return run_half<R>(s, t, ElPhase<R>{(R)cos(g.theta), (R)sin(g.theta)});
    case MOONLAB_HIP_GATE_RX:
        return run_pair<R>(s, t, OpRX<R>{(R)cos(g.theta / 2.0), (R)sin(g.theta / 2.0)});
    case MOONLAB_HIP_GATE_RY:
        return run_pair<R>(s, t, OpRY<R>{(R)cos(g.theta / 2.0), (R)sin(g.theta / 2.0)});
    case MOONLAB_HIP_GATE_RZ: {
        const double co = cos(g.theta / 2.0), si = sin(g.theta / 2.0);
        const double d[4] = {co, -si, co, si};
        return run_diag<R>(s, t, d);
    }
    case MOONLAB_HIP_GATE_U3: {
        const double co = cos(g.theta / 2.0), si = sin(g.theta / 2.0);
        const double m[8] = {co, 0.0, -cos(g.lam) * si, -sin(g.lam) * si,
                             cos(g.phi) * si, sin(g.phi) * si,
                             cos(g.phi + g.lam) * co, sin(g.phi + g.lam) * co};
        return run_pair<R>(s, t, mat2<R>(m));
    }
    case MOONLAB_HIP_GATE_CNOT: return run_ctrl_pair<R>(s, 1ull << c, t, OpSwap<R>{});
    case MOONLAB_HIP_GATE_CY: return run_ctrl_pair<R>(s, 1ull << c, t, OpNegY<R>{});
    case MOONLAB_HIP_GATE_CY_TRUE: return run_ctrl_pair<R>(s, 1ull << c, t, OpY<R>{});
    case MOONLAB_HIP_GATE_CZ: return run_mask_elem<R>(s, (1ull << c) | (1ull << t), ElNeg<R>{});
    case MOONLAB_HIP_GATE_SWAP: return run_swap<R>(s, 0ull, c, t);
    case MOONLAB_HIP_GATE_CPHASE:
        return run_mask_elem<R>(s, (1ull << c) | (1ull << t),
                                ElPhase<R>{(R)cos(g.theta), (R)sin(g.theta)});
    case MOONLAB_HIP_GATE_CRX:
        return run_ctrl_pair<R>(s, 1ull << c, t,
                                OpRX<R>{(R)cos(g.theta / 2.0), (R)sin(g.theta / 2.0)});
    case MOONLAB_HIP_GATE_CRY:
        return run_ctrl_pair<R>(s, 1ull << c, t,
                                OpRY<R>{(R)cos(g.theta / 2.0), (R)sin(g.theta / 2.0)});
    case MOONLAB_HIP_GATE_CRZ: {
        const R co = (R)cos(g.theta / 2.0), si = (R)sin(g.theta / 2.0);
        return run_ctrl_pair<R>(s, 1ull << c, t, OpDiag2<R>{co, -si, co, si});
    }
    case MOONLAB_HIP_GATE_TOFFOLI:
        return run_ctrl_pair<R>(s, (1ull << c) | (1ull << e), t, OpSwap<R>{});
    case MOONLAB_HIP_GATE_FREDKIN: return run_swap<R>(s, 1ull << c, t, e);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_gate(const SvDev &s, const moonlab_hip_gate_t &g)
{
    return s.prec == MOONLAB_HIP_C64 ? gate_r<float>(s, g) : gate_r<double>(s, g);
}

hipError_t launch_matrix_1q(const SvDev &s, int q, const double m[8])
{
    return s.prec == MOONLAB_HIP_C64 ? run_pair<float>(s, q, mat2<float>(m))
                                     : run_pair<double>(s, q, mat2<double>(m));
}

hipError_t launch_matrix_2q(const SvDev &s, int q0, int q1, const double m[32])
{
    return s.prec == MOONLAB_HIP_C64 ? run_mat4<float>(s, q0, q1, m)
                                     : run_mat4<double>(s, q0, q1, m);
}

hipError_t launch_controlled_1q(const SvDev &s, uint64_t cmask, int target, const double m[8])
{
    return s.prec == MOONLAB_HIP_C64 ? run_ctrl_pair<float>(s, cmask, target, mat2<float>(m))
                                     : run_ctrl_pair<double>(s, cmask, target, mat2<double>(m));
}

hipError_t launch_mcx(const SvDev &s, uint64_t cmask, int target)
{
    return s.prec == MOONLAB_HIP_C64 ? run_ctrl_pair<float>(s, cmask, target, OpSwap<float>{})
                                     : run_ctrl_pair<double>(s, cmask, target, OpSwap<double>{});
}

hipError_t launch_mcz(const SvDev &s, uint64_t mask)
{
    return s.prec == MOONLAB_HIP_C64 ? run_mask_elem<float>(s, mask, ElNeg<float>{})
                                     : run_mask_elem<double>(s, mask, ElNeg<double>{});
}

hipError_t launch_phase_mask(const SvDev &s, uint64_t mask, double pr, double pi)
{
    return s.prec == MOONLAB_HIP_C64
               ? run_mask_elem<float>(s, mask, ElPhase<float>{(float)pr, (float)pi})
               : run_mask_elem<double>(s, mask, ElPhase<double>{pr, pi});
}

hipError_t launch_cswap(const SvDev &s, uint64_t cmask, int t1, int t2)
{
    return s.prec == MOONLAB_HIP_C64 ? run_swap<float>(s, cmask, t1, t2)
                                     : run_swap<double>(s, cmask, t1, t2);
}

hipError_t launch_diag_1q(const SvDev &s, int q, const double d[4])
{
    return s.prec == MOONLAB_HIP_C64 ? run_diag<float>(s, q, d) : run_diag<double>(s, q, d);
}

hipError_t launch_init_zero(const SvDev &s)
{
    const uint64_t dim = 1ull << s.n;
    if (s.prec == MOONLAB_HIP_C64)
        k_init_zero<float><<<grid_stride_blocks(dim), 256, 0, s.stream>>>((float2 *)s.amps, dim);
    else
        k_init_zero<double><<<grid_stride_blocks(dim), 256, 0, s.stream>>>((double2 *)s.amps, dim);
    return hipGetLastError();
}

hipError_t launch_scale(const SvDev &s, double f)
{
    const uint64_t dim = 1ull << s.n;
    if (s.prec == MOONLAB_HIP_C64)
        k_scale<float><<<grid_stride_blocks(dim), 256, 0, s.stream>>>((float2 *)s.amps, dim, (float)f);
    else
        k_scale<double><<<grid_stride_blocks(dim), 256, 0, s.stream>>>((double2 *)s.amps, dim, f);
    return hipGetLastError();
}

hipError_t launch_convert(const void *src, int sp, void *dst, int dp, uint64_t count,
                          hipStream_t st)
{
    const unsigned g = grid_stride_blocks(count);
    if (sp == MOONLAB_HIP_C128 && dp == MOONLAB_HIP_C64)
        k_convert<double, float><<<g, 256, 0, st>>>((const double2 *)src, (float2 *)dst, count);
    else if (sp == MOONLAB_HIP_C64 && dp == MOONLAB_HIP_C128)
        k_convert<float, double><<<g, 256, 0, st>>>((const float2 *)src, (double2 *)dst, count);
    else if (sp == MOONLAB_HIP_C128)
        k_convert<double, double><<<g, 256, 0, st>>>((const double2 *)src, (double2 *)dst, count);
    else
        k_convert<float, float><<<g, 256, 0, st>>>((const float2 *)src, (float2 *)dst, count);
    return hipGetLastError();
}

}  // namespace mlx
