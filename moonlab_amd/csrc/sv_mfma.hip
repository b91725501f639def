// sv_mfma.hip -- fused k-qubit blocks (k = 3..5) contracted on the matrix cores.
//
// A fused block is a 2^k x 2^k complex unitary M applied to every group of
// 2^k amplitudes that share all bits outside the block's qubits Q.  Written
// as a real GEMM with interleaved (re, im) rows,
//     Y (2^(k+1) x G) = A (2^(k+1) x 2^(k+1)) * X (2^(k+1) x G),
//     A[2r+a][2c+b] = {Mr, -Mi; Mi, Mr}[a][b](r, c),  X[2c+b][g] = (re, im)_b of x_c(g),
// it maps onto v_mfma_{f32,f64}_16x16x4 with FP32 (complex64) or FP64
// (complex128) accumulation.  For k >= 4 the same product runs as three
// real 2^k x 2^k GEMMs (the 3M method, struct Block below), 3/4 of the
// real-embedding MFMAs:
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
//   * a workgroup streams a few tiles (4 complex64, 32 complex128) and exits;
//     the next tile's global loads are issued before the current tile is
//     contracted, and the barriers order LDS only, so loads, MFMAs and the
//     previous tile's stores overlap.
// Operand layout of 16x16x4 (lane l): A[l % 16][l / 16], B[l / 16][l % 16];
// accumulator v = 0..3 holds D[4 (l / 16) + v][l % 16] for FP32 but
// D[4 v + l / 16][l % 16] for FP64.  The FP64 path loads A with its rows
// permuted (A row i = real row 4 (i % 4) + i / 4), which puts both results in
// the FP32 order, so one write-back serves both.
#include "sv_device.h"
#include "sv_internal.h"

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
// TB = G.tb is a compile-time constant of every kernel, so the tile set is
// indexed with constants and read into scalar registers once; indexing it
// with a loop variable made the compiler fetch each byte with a vector
// load and wait for every outstanding load and store before using it.
template <int TB>
__device__ __forceinline__ uint64_t tile_base(uint64_t t, const DenseGeom &G)
{
#pragma unroll
    for (int b = 0; b < TB; b++) t = ins0(t, G.tq[b]);
    return t;
}

// The contraction of one 16-group column of a fused block, shared by both
// kernels.  k >= 4 blocks run as three real products (the 3M method):
//     P1 = Mr xr,  P2 = Mi xi,  P3 = (Mr + Mi)(xr + xi),
//     Re y = P1 - P2,  Im y = P3 - P1 - P2,
// each a 2^k x 2^k real GEMM: 3 * (2^k / 16) * (2^k / 4) MFMAs per 16
// groups, three quarters of the real-embedding form's (2^(k+1))^2 / 64, and
// the B operand is one complex LDS read per k-step instead of two halves.
// k = 3 (8 rows, less than one 16-row tile) keeps the real embedding
// A[2r+a][2c+b] = {Mr, -Mi; Mi, Mr}[a][b](r, c).
// Output o of a lane is block row out_row(o) of its group.
template <typename R, int K>
struct Block {
    static constexpr int D = 1 << K;
    static constexpr bool M3 = K >= 4;
    static constexpr int NP = M3 ? 3 : 1;                      // real products
    static constexpr int NRT = (M3 ? D : 2 * D) / 16;          // 16-row tiles
    static constexpr int NS = (M3 ? D : 2 * D) / 4;            // k-steps
    static constexpr int NOUT = M3 ? 4 * NRT : 2 * NRT;        // complex outputs per lane
    using AccT = Acc<R>;
    R Am[NP][NRT][NS];

    // Block row of accumulator v of lane group q = l / 16 in a 16-row tile
    // of the 3M products: (v & 1) + 2 q + 8 (v / 2), so one store (fixed v)
    // covers rows that differ in bits 1..2, as in the real-embedding form.
    static __device__ __forceinline__ int m3_row(int v, int q) { return (v & 1) + 2 * q + 8 * (v >> 1); }
    // block row that A row i carries: accumulator v holds D[4 q + v] for
    // FP32 and D[4 v + q] for FP64
    static __device__ __forceinline__ int m3_arow(int i)
    {
        if constexpr (sizeof(R) == 4) return m3_row(i & 3, i >> 2);
        else return m3_row(i >> 2, i & 3);
    }

    // A operands: rows 16 rt + (a_row or m3_arow)(l % 16), columns 4 s + l / 16
    __device__ __forceinline__ void load(const cplx_t<R> *__restrict__ mat, int lane)
    {
#pragma unroll
        for (int rt = 0; rt < NRT; rt++)
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int col = 4 * s + (lane >> 4);
                if constexpr (M3) {
                    const int row = 16 * rt + m3_arow(lane & 15);
                    const cplx_t<R> m = mat[row * D + col];
                    Am[0][rt][s] = m.x;
                    Am[1][rt][s] = m.y;
                    Am[2][rt][s] = m.x + m.y;
                } else {
                    const int row = 16 * rt + AccT::a_row(lane & 15);
                    const cplx_t<R> m = mat[(row >> 1) * D + (col >> 1)];
                    const int ra = row & 1, cb = col & 1;
                    Am[0][rt][s] = ra == cb ? m.x : (ra == 0 ? -m.y : m.y);
                }
            }
    }
    // real embedding: accumulator v of row tile rt holds real row
    // 16 rt + 4 (l / 16) + v (the FP64 A rows are permuted by a_row to give
    // the FP32 order), i.e. block row 8 rt + 2 (l / 16) + v / 2
    static __device__ __forceinline__ int out_row(int o, int lane)
    {
        if constexpr (M3) return 16 * (o >> 2) + m3_row(o & 3, lane >> 4);
        else return 8 * (o >> 1) + 2 * (lane >> 4) + (o & 1);
    }
    // get(c): amplitude c of this lane's group; put(o, y): output o
    template <class Get, class Put>
    __device__ __forceinline__ void apply(int lane, Get get, Put put) const
    {
        if constexpr (M3) {
            R br[NS], bi[NS], bs[NS];
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const cplx_t<R> x = get(4 * s + (lane >> 4));
                br[s] = x.x;
                bi[s] = x.y;
                bs[s] = x.x + x.y;
            }
            typename AccT::type p1[NRT], p2[NRT], p3[NRT];
#pragma unroll
            for (int rt = 0; rt < NRT; rt++) {
                p1[rt] = p2[rt] = p3[rt] = typename AccT::type{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    p1[rt] = AccT::mma(Am[0][rt][s], br[s], p1[rt]);
                    p2[rt] = AccT::mma(Am[1][rt][s], bi[s], p2[rt]);
                    p3[rt] = AccT::mma(Am[2][rt][s], bs[s], p3[rt]);
                }
            }
#pragma unroll
            for (int rt = 0; rt < NRT; rt++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    put(4 * rt + v, cplx_t<R>{p1[rt][v] - p2[rt][v], p3[rt][v] - p1[rt][v] - p2[rt][v]});
        } else {
            R B[NS];
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int kk = 4 * s + (lane >> 4);
                const cplx_t<R> x = get(kk >> 1);
                B[s] = (kk & 1) ? x.y : x.x;
            }
            typename AccT::type acc[NRT];
#pragma unroll
            for (int rt = 0; rt < NRT; rt++) {
                acc[rt] = typename AccT::type{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < NS; s++) acc[rt] = AccT::mma(Am[0][rt][s], B[s], acc[rt]);
            }
#pragma unroll
            for (int rt = 0; rt < NRT; rt++) {
                put(2 * rt, cplx_t<R>{acc[rt][0], acc[rt][1]});
                put(2 * rt + 1, cplx_t<R>{acc[rt][2], acc[rt][3]});
            }
        }
    }
};


template <typename R, int K, int GB, bool NT>
__global__ __launch_bounds__(kThreads) void k_dense_mfma(cplx_t<R> *__restrict__ a, uint64_t ntiles,
                                                         DenseGeom G, const cplx_t<R> *__restrict__ mat)
{
    constexpr int D = 1 << K;
    constexpr int GM = 1 << GB;         // 64-group sets per tile
    constexpr int T = 64 * GM * D;      // amplitudes per tile
    constexpr int A = Slot<R>::A;       // amplitudes per 16-byte slot
    constexpr int PER = T / A / kThreads;   // slots each thread moves per tile
    constexpr int HSTEP = kThreads * A / 64;
    static_assert(T <= 4096 && PER >= 1 && T / 64 <= 64, "tile shape");
    using SlotT = typename Slot<R>::type;
    __shared__ cplx_t<R> lds[T + T / 32];
    __shared__ uint64_t dlo[64], dhi[64];
    __shared__ uint16_t offq[32], offg_lo[64], offg_hi[8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // Index tables.  The tile set is read with constant indices (scalar
    // loads) and every per-thread load is issued before any is used: a
    // workgroup lives for a few tiles only, and the former loop over a
    // thread-dependent index fetched each qubit byte with its own wait.
    {
        const int j = tid & 63;
        const bool upper = tid >= 64;   // threads 64..127: tile bits 6..11
        uint64_t o = 0;
#pragma unroll
        for (int b = 0; b < 6; b++) {
            const int qlo = G.tq[b], qhi = G.tq[6 + b];
            const bool on = (j >> b) & 1;
            if (on && !upper && b < G.tb) o |= 1ull << qlo;
            if (on && upper && 6 + b < G.tb) o |= 1ull << qhi;
        }
        const uint16_t q = G.offq[tid & 31], gl = G.offg_lo[tid & 63], gh = G.offg_hi[tid & 7];
        if (tid < 64) dlo[j] = o;
        else if (tid < 128) dhi[j] = o;
        if (tid < 32) offq[tid] = q;
        else if (tid >= 128 && tid < 192) offg_lo[j] = gl;
        else if (tid >= 192 && tid < 200) offg_hi[tid - 192] = gh;
    }

    Block<R, K> blk;
    blk.load(mat, lane);
    __syncthreads();

    // slot s = tid + kThreads i holds tile amplitudes A s .. A s + A - 1:
    // (A s) % 64 = A (tid % (64 / A)) and (A s) / 64 = tid / (64 / A) + HSTEP i
    const uint64_t lo = dlo[A * (tid & (64 / A - 1))];
    const int hi0 = tid / (64 / A);

    // software pipeline: the walk's next tile is in flight from global
    // memory while tile t is contracted and written back
    SlotT pre[PER];
    const TileWalk walk{blockIdx.x, gridDim.x, ntiles};   // tiles b, b + grid, ...
    uint64_t t = walk.t;
    if (t < walk.end) {
        const uint64_t base = tile_base<K + 6 + GB>(t, G) | lo;
#pragma unroll
        for (int i = 0; i < PER; i++) pre[i] = ld_slot<NT>((const SlotT *)&a[base | dhi[hi0 + HSTEP * i]]);
    }
    for (; t < walk.end; t += walk.step) {
#pragma unroll
        for (int i = 0; i < PER; i++) put_slot(lds, padded(A * (tid + kThreads * i)), pre[i]);
        lds_barrier();
        const uint64_t nt = t + walk.step;
        if (nt < walk.end) {
            const uint64_t nbase = tile_base<K + 6 + GB>(nt, G) | lo;
#pragma unroll
            for (int i = 0; i < PER; i++) pre[i] = ld_slot<NT>((const SlotT *)&a[nbase | dhi[hi0 + HSTEP * i]]);
        }

        // a wave reads and writes only its own groups: no barrier in between
#pragma unroll 1
        for (int j = 0; j < GM; j++) {
            const int g = 64 * j + 16 * wave + (lane & 15);
            const int go = offg_lo[g & 63] | offg_hi[g >> 6];
            blk.apply(
                lane, [&](int c) { return lds[padded(offq[c] | go)]; },
                [&](int o, const cplx_t<R> &y) { lds[padded(offq[Block<R, K>::out_row(o, lane)] | go)] = y; });
        }
        lds_barrier();
        const uint64_t base = tile_base<K + 6 + GB>(t, G) | lo;
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

// Short-lived workgroups in plain order (workgroup b takes tiles b, b + grid,
// ...): kTilesPerGroup tiles each, so the dispatcher keeps refilling the CUs.
// A resident wave of persistent workgroups streamed HBM slower (fusion
// workload: complex64 4.39 -> 3.97 s at four tiles per workgroup; complex128
// 15.4-15.6 -> 14.8-15.2 s at 32, against the persistent ring kernel this
// replaced; fewer tiles per workgroup pay the A-operand and index-table
// set-up of every workgroup: complex128 at 2 tiles 18 % slower).
template <typename R>
constexpr uint64_t kTilesPerGroup = sizeof(R) == 4 ? 4 : 32;

template <typename R, int K, int GB>
hipError_t run_mfma(const SvDev &s, const DenseGeom &G, const void *d_mat, bool nt)
{
    const uint64_t ntiles = (1ull << s.n) >> G.tb;
    const unsigned grid = (unsigned)(ntiles >= kTilesPerGroup<R> ? ntiles / kTilesPerGroup<R> : 1);
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
