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
//   * a 256-thread workgroup stages a tile of 64 * 2^gb groups in LDS (about
//     32 KiB): the tile set is Q plus the 6 + gb lowest other qubits, so
//     global accesses are contiguous runs of >= 64 amplitudes, moved as
//     16-byte slots (one complex128 or two complex64 amplitudes); LDS indices
//     are padded by one slot per 32 to keep the strided B-operand reads
//     conflict-free when Q holds the low qubits;
//   * each wave contracts 16 groups at a time: (2^(k+1)/16) row tiles x
//     (2^(k+1)/4) k-steps of 16x16x4 MFMAs; it reads and writes only its own
//     groups, so no barrier separates the contraction from the write-back;
//   * the next tile's global loads are issued before the current tile is
//     contracted, and the barriers order LDS only, so loads, MFMAs and the
//     previous tile's stores overlap.
// Operand layout of 16x16x4 (lane l): A[l % 16][l / 16], B[l / 16][l % 16];
// accumulator v = 0..3 holds D[4 (l / 16) + v][l % 16] for FP32 but
// D[4 v + l / 16][l % 16] for FP64.  The FP64 path loads A with its rows
// permuted (A row i = real row 4 (i % 4) + i / 4), which puts both results in
// the FP32 order, so one write-back serves both.
#include "sv_device.h"
#include "sv_internal.h"

#include <atomic>
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

// 16-byte global slot: one complex128 amplitude or two adjacent complex64
// amplitudes (tile bit 0 is always qubit 0, so tile amplitudes 2s and 2s + 1
// are neighbours in memory).
typedef float mfma_v4f __attribute__((ext_vector_type(4)));
template <typename R> struct Slot;
template <> struct Slot<double> {
    using type = double2;
    static constexpr int A = 1;
};
template <> struct Slot<float> {
    using type = float4;
    static constexpr int A = 2;
};

template <bool NT>
__device__ __forceinline__ double2 ld_slot(const double2 *p)
{
    return ld<NT>(p);
}
template <bool NT>
__device__ __forceinline__ float4 ld_slot(const float4 *p)
{
    if constexpr (NT) {
        const mfma_v4f v = __builtin_nontemporal_load((const mfma_v4f *)p);
        return float4{v.x, v.y, v.z, v.w};
    } else {
        return *p;
    }
}
template <bool NT>
__device__ __forceinline__ void st_slot(double2 *p, const double2 &x)
{
    st<NT>(p, x);
}
template <bool NT>
__device__ __forceinline__ void st_slot(float4 *p, const float4 &x)
{
    if constexpr (NT) {
        const mfma_v4f v = {x.x, x.y, x.z, x.w};
        __builtin_nontemporal_store(v, (mfma_v4f *)p);
    } else {
        *p = x;
    }
}
// slot <-> LDS at padded index i (a complex64 slot starts at an even tile
// index, so i + 1 never crosses a pad)
__device__ __forceinline__ void put_slot(double2 *l, int i, const double2 &x) { l[i] = x; }
__device__ __forceinline__ void put_slot(float2 *l, int i, const float4 &x)
{
    l[i] = float2{x.x, x.y};
    l[i + 1] = float2{x.z, x.w};
}
__device__ __forceinline__ double2 get_slot(const double2 *l, int i) { return l[i]; }
__device__ __forceinline__ float4 get_slot(const float2 *l, int i)
{
    const float2 u = l[i], v = l[i + 1];
    return float4{u.x, u.y, v.x, v.y};
}

// Tile geometry, computed on the host.  A tile holds 64 * 2^gb groups of 2^k
// amplitudes: the block's qubits plus the 6 + gb lowest other qubits.
struct DenseGeom {
    int32_t  tb;           // tile bits = k + 6 + gb (<= 12)
    int32_t  gb;
    int8_t   tq[16];       // tile-set qubits, ascending
    uint16_t offq[32];     // local index of amplitude r of a group, r < 2^k
    uint16_t offg_lo[64];  // local index contributed by group bits 0..5
    uint16_t offg_hi[8];   // local index contributed by group bits 6..6+gb-1
};

constexpr int kThreads = 256;

__device__ __forceinline__ int padded(int l) { return l + (l >> 5); }

// Tile t's amplitude l lives at tile_base(t) | dlo[l % 64] | dhi[l / 64].
__device__ __forceinline__ uint64_t tile_base(uint64_t t, const DenseGeom &G)
{
    for (int b = 0; b < G.tb; b++) t = ins0(t, G.tq[b]);
    return t;
}

template <typename R, int K, int GB, bool NT>
__global__ __launch_bounds__(kThreads) void k_dense_mfma(cplx_t<R> *__restrict__ a, uint64_t ntiles,
                                                         DenseGeom G, const cplx_t<R> *__restrict__ mat)
{
    constexpr int D = 1 << K;
    constexpr int GM = 1 << GB;         // 64-group sets per tile
    constexpr int NRT = (2 * D) / 16;   // row tiles
    constexpr int NS = (2 * D) / 4;     // k-steps
    constexpr int T = 64 * GM * D;      // amplitudes per tile
    constexpr int A = Slot<R>::A;       // amplitudes per 16-byte slot
    constexpr int PER = T / A / kThreads;   // slots each thread moves per tile
    constexpr int HSTEP = kThreads * A / 64;
    static_assert(T <= 4096 && PER >= 1 && T / 64 <= 64, "tile shape");
    using AccT = Acc<R>;
    using SlotT = typename Slot<R>::type;
    __shared__ cplx_t<R> lds[T + T / 32];
    __shared__ uint64_t dlo[64], dhi[64];
    __shared__ uint16_t offq[32], offg_lo[64], offg_hi[8];
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
    } else if (tid < 160) {
        offq[tid - 128] = G.offq[tid - 128];
    } else if (tid < 224) {
        offg_lo[tid - 160] = G.offg_lo[tid - 160];
    } else if (tid < 232) {
        offg_hi[tid - 224] = G.offg_hi[tid - 224];
    }

    // A operands: rows 16 rt + a_row(l % 16), columns 4 s + l / 16 of the real matrix
    R Am[NRT][NS];
#pragma unroll
    for (int rt = 0; rt < NRT; rt++)
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int row = 16 * rt + AccT::a_row(lane & 15), col = 4 * s + (lane >> 4);
            const cplx_t<R> m = mat[(row >> 1) * D + (col >> 1)];
            const int ra = row & 1, cb = col & 1;
            Am[rt][s] = ra == cb ? m.x : (ra == 0 ? -m.y : m.y);
        }
    __syncthreads();

    // slot s = tid + kThreads i holds tile amplitudes A s .. A s + A - 1:
    // (A s) % 64 = A (tid % (64 / A)) and (A s) / 64 = tid / (64 / A) + HSTEP i
    const uint64_t lo = dlo[A * (tid & (64 / A - 1))];
    const int hi0 = tid / (64 / A);
    const R *ldsr = (const R *)lds;

    // software pipeline: tile t + gridDim.x is in flight from global memory
    // while tile t is contracted and written back
    SlotT pre[PER];
    uint64_t t = blockIdx.x;
    if (t < ntiles) {
        const uint64_t base = tile_base(t, G) | lo;
#pragma unroll
        for (int i = 0; i < PER; i++) pre[i] = ld_slot<NT>((const SlotT *)&a[base | dhi[hi0 + HSTEP * i]]);
    }
    for (; t < ntiles; t += gridDim.x) {
#pragma unroll
        for (int i = 0; i < PER; i++) put_slot(lds, padded(A * (tid + kThreads * i)), pre[i]);
        lds_barrier();
        const uint64_t nt = t + gridDim.x;
        if (nt < ntiles) {
            const uint64_t nbase = tile_base(nt, G) | lo;
#pragma unroll
            for (int i = 0; i < PER; i++) pre[i] = ld_slot<NT>((const SlotT *)&a[nbase | dhi[hi0 + HSTEP * i]]);
        }

        // a wave reads and writes only its own groups: no barrier in between
#pragma unroll 1
        for (int j = 0; j < GM; j++) {
            const int g = 64 * j + 16 * wave + (lane & 15);
            const int go = offg_lo[g & 63] | offg_hi[g >> 6];
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
                for (int s = 0; s < NS; s++) acc[rt] = AccT::mma(Am[rt][s], B[s], acc[rt]);
            }
#pragma unroll
            for (int rt = 0; rt < NRT; rt++) {
                const int r0 = 8 * rt + 2 * (lane >> 4);
                lds[padded(offq[r0] | go)] = cplx_t<R>{acc[rt][0], acc[rt][1]};
                lds[padded(offq[r0 + 1] | go)] = cplx_t<R>{acc[rt][2], acc[rt][3]};
            }
        }
        lds_barrier();
        const uint64_t base = tile_base(t, G) | lo;
#pragma unroll
        for (int i = 0; i < PER; i++)
            st_slot<NT>((SlotT *)&a[base | dhi[hi0 + HSTEP * i]], get_slot(lds, padded(A * (tid + kThreads * i))));
        lds_barrier();   // every wave has read its write-back values before the next tile lands
    }
}

// Tile set = Q plus the 6 + gb lowest other qubits; local indices follow the
// ascending tile-set order.
bool make_geom(int n, int k, int gb, const int *qubits, DenseGeom &G)
{
    const int other = 6 + gb;
    if (n < k + other) return false;
    uint64_t qm = 0;
    for (int j = 0; j < k; j++) qm |= 1ull << qubits[j];
    uint64_t lm = 0;
    for (int q = 0, got = 0; q < n && got < other; q++)
        if (!(qm & (1ull << q))) {
            lm |= 1ull << q;
            got++;
        }
    const uint64_t tm = qm | lm;
    G.tb = 0;
    G.gb = gb;
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
    int lq[9], nl = 0;
    for (int q = 0; q < n; q++)
        if (lm & (1ull << q)) lq[nl++] = q;
    for (int gi = 0; gi < 64; gi++) {
        uint32_t o = 0;
        for (int j = 0; j < 6; j++)
            if ((gi >> j) & 1) o |= 1u << local[lq[j]];
        G.offg_lo[gi] = (uint16_t)o;
    }
    for (int h = 0; h < 8; h++) {
        uint32_t o = 0;
        for (int j = 0; j < gb; j++)
            if ((h >> j) & 1) o |= 1u << local[lq[6 + j]];
        G.offg_hi[h] = (uint16_t)(h < (1 << gb) ? o : 0);
    }
    return true;
}

// Cached launch facts are shared by every thread that launches; they are the
// same whichever thread computes them first.
int device_cus()
{
    static std::atomic<int> cached{0};
    int cus = cached.load(std::memory_order_relaxed);
    if (!cus) {
        int dev = 0;
        hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached.store(cus, std::memory_order_relaxed);
    }
    return cus;
}

// Persistent grid: as many workgroups as the device holds at once, so no
// workgroup waits for a slot while others still have tiles to take.
template <typename R, int K, int GB, bool NT>
int resident_blocks()
{
    static std::atomic<int> cached{0};   // one per kernel instantiation
    int blocks = cached.load(std::memory_order_relaxed);
    if (!blocks) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dense_mfma<R, K, GB, NT>, kThreads, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        blocks = per_cu * device_cus();
        cached.store(blocks, std::memory_order_relaxed);
    }
    return blocks;
}

template <typename R, int K, int GB>
hipError_t run_mfma(const SvDev &s, const DenseGeom &G, const void *d_mat, bool nt)
{
    const uint64_t ntiles = (1ull << s.n) >> G.tb;
    const int mb = nt ? resident_blocks<R, K, GB, true>() : resident_blocks<R, K, GB, false>();
    const unsigned grid = (unsigned)(ntiles < (uint64_t)mb ? ntiles : (uint64_t)mb);
    if (nt)
        k_dense_mfma<R, K, GB, true><<<grid, kThreads, 0, s.stream>>>((cplx_t<R> *)s.amps, ntiles, G,
                                                                      (const cplx_t<R> *)d_mat);
    else
        k_dense_mfma<R, K, GB, false><<<grid, kThreads, 0, s.stream>>>((cplx_t<R> *)s.amps, ntiles, G,
                                                                       (const cplx_t<R> *)d_mat);
    return hipGetLastError();
}

// Group-set bits for a 4096-amplitude complex64 / 2048-amplitude complex128
// tile (32 KiB): 2^(k + 6 + gb) amplitudes.
constexpr int wide_gb(bool f32, int k) { return (f32 ? 12 : 11) - 6 - k; }

template <typename R, int K>
hipError_t dispatch_k(const SvDev &s, const int *qubits, const void *d_mat, bool nt)
{
    constexpr int W = wide_gb(sizeof(R) == 4, K);
    DenseGeom G;
    if (make_geom(s.n, K, W, qubits, G)) return run_mfma<R, K, W>(s, G, d_mat, nt);
    if (make_geom(s.n, K, 0, qubits, G)) return run_mfma<R, K, 0>(s, G, d_mat, nt);
    return hipErrorInvalidValue;
}

}  // namespace

// Dense k-qubit block on the matrix cores; false when the state is too small
// for a 2^(k+6)-amplitude tile (the caller then uses the VALU kernel).
bool mfma_dense_supported(const SvDev &s, int k) { return k >= 3 && k <= 5 && s.n >= k + 6; }

hipError_t launch_dense_mfma(const SvDev &s, int k, const int *qubits, const void *d_mat)
{
    const bool f = s.prec == MOONLAB_HIP_C64;
    const uint64_t bytes = (1ull << s.n) * (f ? 8ull : 16ull);
    const bool nt = bytes >= gate_tuning().nt_min_bytes;
    switch (k) {
    case 3: return f ? dispatch_k<float, 3>(s, qubits, d_mat, nt) : dispatch_k<double, 3>(s, qubits, d_mat, nt);
    case 4: return f ? dispatch_k<float, 4>(s, qubits, d_mat, nt) : dispatch_k<double, 4>(s, qubits, d_mat, nt);
    case 5: return f ? dispatch_k<float, 5>(s, qubits, d_mat, nt) : dispatch_k<double, 5>(s, qubits, d_mat, nt);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mlx
