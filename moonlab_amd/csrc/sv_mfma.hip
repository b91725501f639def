// sv_mfma.hip -- fused k-qubit blocks (k = 3..5) contracted on the matrix cores.
//
// A fused block is a 2^k x 2^k complex unitary M applied to every group of
// 2^k amplitudes that share all bits outside the block's qubits Q.  Written
// as a real GEMM with interleaved (re, im) rows,
//     Y (2^(k+1) x G) = A (2^(k+1) x 2^(k+1)) * X (2^(k+1) x G),
//     A[2r+a][2c+b] = {Mr, -Mi; Mi, Mr}[a][b](r, c),  X[2c+b][g] = (re, im)_b of x_c(g),
// it maps onto v_mfma_{f32,f64}_16x16x4 with FP32 (complex64) or FP64
// (complex128) accumulation:
//   * A (the block's matrix) stays in VGPRs for the whole launch;
//   * a 256-thread workgroup stages a tile of 64 groups (2^(k+6) amplitudes)
//     in LDS: the tile set is Q plus the six lowest other qubits, so global
//     loads and stores are contiguous runs of >= 64 amplitudes; LDS indices
//     are padded by one slot per 32 to keep the strided B-operand reads
//     conflict-free when Q holds the low qubits;
//   * each wave contracts 16 groups: (2^(k+1)/16) row tiles x (2^(k+1)/4)
//     k-steps of 16x16x4 MFMAs; it reads and writes only its own groups, so
//     no barrier separates the contraction from the write-back.
// Operand layout of 16x16x4 (lane l): A[l % 16][l / 16], B[l / 16][l % 16];
// accumulator v = 0..3 holds D[4 (l / 16) + v][l % 16] for FP32 but
// D[4 v + l / 16][l % 16] for FP64.  The FP64 path loads A with its rows
// permuted (A row i = real row 4 (i % 4) + i / 4), which puts both results in
// the FP32 order, so one write-back serves both.
#include "sv_device.h"
#include "sv_internal.h"

#include <vector>

namespace mlx {

typedef float mfma_f4 __attribute__((ext_vector_type(4)));
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

namespace {

template <typename R> struct Acc;
template <> struct Acc<float> {
    using type = mfma_f4;
    static __device__ __forceinline__ type mma(float a, float b, type c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // real-matrix row that A row i (= lane % 16) carries
    static __device__ __forceinline__ int a_row(int i) { return i; }
};
template <> struct Acc<double> {
    using type = mfma_d4;
    static __device__ __forceinline__ type mma(double a, double b, type c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int a_row(int i) { return 4 * (i & 3) + (i >> 2); }
};

// Tile geometry, computed on the host.
struct DenseGeom {
    int32_t  tb;           // tile bits = k + 6
    int8_t   tq[16];       // tile-set qubits, ascending
    uint16_t offq[32];     // local (tile) index of amplitude r of a group, r < 2^k
    uint16_t offg[64];     // local index of group g within the tile
};

constexpr int kGroups = 64;
constexpr int kThreads = 256;

__device__ __forceinline__ int padded(int l) { return l + (l >> 5); }

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// waits for every outstanding global load and store of the wave, which would
// drain the prefetch of the next tile and the write-back of the last one.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Tile t's amplitude l lives at tile_base(t) | dlo[l % 64] | dhi[l / 64].
__device__ __forceinline__ uint64_t tile_base(uint64_t t, const DenseGeom &G)
{
    for (int b = 0; b < G.tb; b++) t = ins0(t, G.tq[b]);
    return t;
}

template <typename R, int K, bool NT>
__global__ __launch_bounds__(kThreads) void k_dense_mfma(cplx_t<R> *__restrict__ a, uint64_t ntiles,
                                                         DenseGeom G, const cplx_t<R> *__restrict__ mat)
{
    constexpr int D = 1 << K;
    constexpr int NRT = (2 * D) / 16;   // row tiles
    constexpr int NS = (2 * D) / 4;     // k-steps
    constexpr int T = kGroups * D;      // amplitudes per tile
    constexpr int PER = T / kThreads;   // amplitudes each thread moves per tile
    using AccT = Acc<R>;
    __shared__ cplx_t<R> lds[T + T / 32];
    __shared__ uint64_t dlo[64], dhi[64];
    __shared__ uint16_t offq[32], offg[64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    if (tid < 128) {
        const int j = tid & 63, hi = tid >> 6;
        uint64_t o = 0;
        for (int b = 0; b < 6; b++) {
            const int qi = hi * 6 + b;
            if (qi < G.tb && ((j >> b) & 1)) o |= 1ull << G.tq[qi];
        }
        if (hi) dhi[j] = o;
        else dlo[j] = o;
    } else if (tid < 128 + 32) {
        offq[tid - 128] = G.offq[tid - 128];
    } else if (tid < 128 + 32 + 64) {
        offg[tid - 160] = G.offg[tid - 160];
    }

    // A operands: rows 16 rt + a_row(l % 16), columns 4 s + l / 16 of the real matrix
    R A[NRT][NS];
#pragma unroll
    for (int rt = 0; rt < NRT; rt++)
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int row = 16 * rt + AccT::a_row(lane & 15), col = 4 * s + (lane >> 4);
            const cplx_t<R> m = mat[(row >> 1) * D + (col >> 1)];
            const int ra = row & 1, cb = col & 1;
            A[rt][s] = ra == cb ? m.x : (ra == 0 ? -m.y : m.y);
        }
    __syncthreads();

    const int g = wave * 16 + (lane & 15);
    const int go = offg[g];
    // thread tid moves tile amplitudes l = tid + kThreads i: l % 64 = tid % 64
    const uint64_t lo = dlo[tid & 63];
    const R *ldsr = (const R *)lds;

    // software pipeline: tile t + gridDim.x is in flight from global memory
    // while tile t is contracted and written back
    cplx_t<R> pre[PER];
    uint64_t t = blockIdx.x;
    if (t < ntiles) {
        const uint64_t base = tile_base(t, G) | lo;
#pragma unroll
        for (int i = 0; i < PER; i++) pre[i] = ld<NT>(&a[base | dhi[(tid >> 6) + 4 * i]]);
    }
    for (; t < ntiles; t += gridDim.x) {
#pragma unroll
        for (int i = 0; i < PER; i++) lds[padded(tid + kThreads * i)] = pre[i];
        lds_barrier();
        const uint64_t nt = t + gridDim.x;
        if (nt < ntiles) {
            const uint64_t nbase = tile_base(nt, G) | lo;
#pragma unroll
            for (int i = 0; i < PER; i++) pre[i] = ld<NT>(&a[nbase | dhi[(tid >> 6) + 4 * i]]);
        }

        R B[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int kk = 4 * s + (lane >> 4);
            B[s] = ldsr[2 * padded(offq[kk >> 1] | go) + (kk & 1)];
        }
        typename AccT::type acc[NRT];
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) {
            acc[rt] = typename AccT::type{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < NS; s++) acc[rt] = AccT::mma(A[rt][s], B[s], acc[rt]);
        }
        // a wave reads and writes only its own 16 groups: no barrier in between
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) {
            const int r0 = 8 * rt + 2 * (lane >> 4);
            lds[padded(offq[r0] | go)] = cplx_t<R>{acc[rt][0], acc[rt][1]};
            lds[padded(offq[r0 + 1] | go)] = cplx_t<R>{acc[rt][2], acc[rt][3]};
        }
        lds_barrier();
        const uint64_t base = tile_base(t, G) | lo;
#pragma unroll
        for (int i = 0; i < PER; i++) st<NT>(&a[base | dhi[(tid >> 6) + 4 * i]], lds[padded(tid + kThreads * i)]);
        lds_barrier();   // every wave has read its write-back values before the next tile lands
    }
}

// Tile set = Q plus the six lowest other qubits; local indices follow the
// ascending tile-set order.
bool make_geom(int n, int k, const int *qubits, DenseGeom &G)
{
    if (n < k + 6) return false;
    uint64_t qm = 0;
    for (int j = 0; j < k; j++) qm |= 1ull << qubits[j];
    uint64_t lm = 0;
    for (int q = 0, got = 0; q < n && got < 6; q++)
        if (!(qm & (1ull << q))) {
            lm |= 1ull << q;
            got++;
        }
    const uint64_t tm = qm | lm;
    G.tb = 0;
    int local[64];
    for (int q = 0; q < n; q++)
        if (tm & (1ull << q)) {
            local[q] = G.tb;
            G.tq[G.tb++] = (int8_t)q;
        }
    for (int r = 0; r < 32; r++) {
        uint32_t o = 0;
        for (int j = 0; j < k; j++)
            if ((r >> j) & 1) o |= 1u << local[qubits[j]];
        G.offq[r] = (uint16_t)(r < (1 << k) ? o : 0);
    }
    int lq[6], nl = 0;
    for (int q = 0; q < n; q++)
        if (lm & (1ull << q)) lq[nl++] = q;
    for (int gi = 0; gi < 64; gi++) {
        uint32_t o = 0;
        for (int j = 0; j < 6; j++)
            if ((gi >> j) & 1) o |= 1u << local[lq[j]];
        G.offg[gi] = (uint16_t)o;
    }
    return true;
}

int device_cus()
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (cus <= 0) cus = 256;
    }
    return cus;
}

// Persistent grid: as many workgroups as the device holds at once, so no
// workgroup waits for a slot while others still have tiles to take.
template <typename R, int K, bool NT>
int resident_blocks()
{
    static int cached = 0;   // one per kernel instantiation
    if (!cached) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dense_mfma<R, K, NT>, kThreads, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        cached = per_cu * device_cus();
    }
    return cached;
}

template <typename R, int K>
hipError_t run_mfma(const SvDev &s, const DenseGeom &G, const void *d_mat, bool nt)
{
    const uint64_t ntiles = (1ull << s.n) >> (K + 6);
    const int mb = nt ? resident_blocks<R, K, true>() : resident_blocks<R, K, false>();
    const unsigned grid = (unsigned)(ntiles < (uint64_t)mb ? ntiles : (uint64_t)mb);
    if (nt)
        k_dense_mfma<R, K, true><<<grid, kThreads, 0, s.stream>>>((cplx_t<R> *)s.amps, ntiles, G,
                                                                  (const cplx_t<R> *)d_mat);
    else
        k_dense_mfma<R, K, false><<<grid, kThreads, 0, s.stream>>>((cplx_t<R> *)s.amps, ntiles, G,
                                                                   (const cplx_t<R> *)d_mat);
    return hipGetLastError();
}

}  // namespace

// Dense k-qubit block on the matrix cores; false when the state is too small
// for a 2^(k+6)-amplitude tile (the caller then uses the VALU kernel).
bool mfma_dense_supported(const SvDev &s, int k) { return k >= 3 && k <= 5 && s.n >= k + 6; }

hipError_t launch_dense_mfma(const SvDev &s, int k, const int *qubits, const void *d_mat)
{
    DenseGeom G;
    if (!make_geom(s.n, k, qubits, G)) return hipErrorInvalidValue;
    const bool f = s.prec == MOONLAB_HIP_C64;
    const uint64_t bytes = (1ull << s.n) * (f ? 8ull : 16ull);
    const bool nt = bytes >= gate_tuning().nt_min_bytes;
    switch (k) {
    case 3: return f ? run_mfma<float, 3>(s, G, d_mat, nt) : run_mfma<double, 3>(s, G, d_mat, nt);
    case 4: return f ? run_mfma<float, 4>(s, G, d_mat, nt) : run_mfma<double, 4>(s, G, d_mat, nt);
    case 5: return f ? run_mfma<float, 5>(s, G, d_mat, nt) : run_mfma<double, 5>(s, G, d_mat, nt);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mlx
