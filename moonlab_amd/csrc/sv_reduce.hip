// sv_reduce.hip -- FP64 reductions and the shot sampler.
//
// One streaming pass produces sum |a|^2, P(q = 1) for EVERY qubit and the
// per-tile probability mass the sampler needs (the reference makes n serial
// passes for <Z_q> over all q, measurement.c:73-107/126-132, and moves the
// whole state to the host first, measurement.c:46-52).
//
// Tiling: a tile is 2^12 consecutive amplitudes; a 256-thread block reads a
// tile with 16 coalesced loads per thread (element e*256 + tid).  Inside a
// tile, qubits 0..7 are thread-index bits, qubits 8..11 are the e bits, and
// qubits >= 12 are bits of the tile number, so per-qubit sums come from 11
// wave reductions per tile plus a per-tile total.  Each block walks a fixed
// set of tiles and leaves n+1 FP64 partials; a final one-block kernel sums
// the partials in block order.  Grid size and tile assignment depend only on
// n, so every result is bit-reproducible run to run and independent of what
// else shares the GPU.
//
// States below 2^12 amplitudes are reduced by one block that sums each
// qubit's probabilities sequentially in index order -- the reference's own
// summation order (measurement.c:99-103).
#include "sv_device.h"
#include "sv_internal.h"

#include <limits.h>

#include <mutex>

namespace mlx {
namespace {

constexpr int kTileBits = 12;
constexpr int kRedThreads = 256;
constexpr int kEpt = (1 << kTileBits) / kRedThreads;   // 16
constexpr int kMaxRedBlocks = 2048;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename R>
__device__ __forceinline__ double prob(const cplx_t<R> &x)
{
    const double r = (double)x.x, i = (double)x.y;
    return fmar(r, r, i * i);
}

// n >= 12: tiles of 4096 amplitudes.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_reduce_tiles(const cplx_t<R> *__restrict__ a, int n,
                                                              uint64_t ntiles,
                                                              double *__restrict__ partial,
                                                              double *__restrict__ tile_sums)
{
    __shared__ double red[4][11];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nvals = n + 1;
    double acc = 0.0;   // thread v <= n owns value v (0 = total, 1 + q = P(q))

    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const cplx_t<R> *t = a + (tile << kTileBits);
        double p[kEpt];
#pragma unroll
        for (int e = 0; e < kEpt; e++) p[e] = prob<R>(t[e * kRedThreads + tid]);
        double tot = 0.0, hb0 = 0.0, hb1 = 0.0, hb2 = 0.0, hb3 = 0.0;
#pragma unroll
        for (int e = 0; e < kEpt; e++) {
            tot += p[e];
            if (e & 1) hb0 += p[e];
            if (e & 2) hb1 += p[e];
            if (e & 4) hb2 += p[e];
            if (e & 8) hb3 += p[e];
        }
        double v[11];
        v[0] = tot;
#pragma unroll
        for (int b = 0; b < 6; b++) v[1 + b] = (lane >> b) & 1 ? tot : 0.0;
        v[7] = hb0; v[8] = hb1; v[9] = hb2; v[10] = hb3;
#pragma unroll
        for (int k = 0; k < 11; k++) v[k] = wave_sum(v[k]);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 11; k++) red[wave][k] = v[k];
        }
        __syncthreads();
        if (tid < nvals) {
            const double total = ((red[0][0] + red[1][0]) + red[2][0]) + red[3][0];
            double add;
            if (tid == 0) {
                add = total;
            } else {
                const int q = tid - 1;
                if (q < 6) add = ((red[0][1 + q] + red[1][1 + q]) + red[2][1 + q]) + red[3][1 + q];
                else if (q == 6) add = red[1][0] + red[3][0];
                else if (q == 7) add = red[2][0] + red[3][0];
                else if (q < 12) add = ((red[0][q - 1] + red[1][q - 1]) + red[2][q - 1]) + red[3][q - 1];
                else add = (tile >> (q - kTileBits)) & 1ull ? total : 0.0;
            }
            acc += add;
            if (tid == 0 && tile_sums) tile_sums[tile] = total;
        }
        __syncthreads();
    }
    if (tid < nvals) partial[(size_t)blockIdx.x * nvals + tid] = acc;
}

// n < 12: one block, sequential sums in index order.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_reduce_small(const cplx_t<R> *__restrict__ a, int n,
                                                              double *__restrict__ partial,
                                                              double *__restrict__ tile_sums)
{
    __shared__ double p[1 << kTileBits];
    const int dim = 1 << n;
    for (int i = threadIdx.x; i < dim; i += blockDim.x) p[i] = prob<R>(a[i]);
    __syncthreads();
    const int v = threadIdx.x;
    if (v <= n) {
        double s = 0.0;
        if (v == 0) {
            for (int i = 0; i < dim; i++) s += p[i];
            if (tile_sums) tile_sums[0] = s;
        } else {
            const int mk = 1 << (v - 1);
            for (int i = 0; i < dim; i++)
                if (i & mk) s += p[i];
        }
        partial[v] = s;
    }
}

__global__ void k_reduce_final(const double *__restrict__ partial, int blocks, int nvals,
                               double *__restrict__ out)
{
    const int v = threadIdx.x;
    if (v >= nvals) return;
    double s = 0.0;
    for (int b = 0; b < blocks; b++) s += partial[(size_t)b * nvals + v];
    out[v] = s;
}

// Inclusive prefix of the tile sums (one block, fixed chunking).
__global__ __launch_bounds__(1024) void k_tile_scan(const double *__restrict__ sums, uint64_t ntiles,
                                                    double *__restrict__ incl)
{
    __shared__ double wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t chunk = (ntiles + 1023) / 1024;
    const uint64_t lo = (uint64_t)tid * chunk;
    const uint64_t hi = lo + chunk < ntiles ? lo + chunk : ntiles;
    double s = 0.0;
    for (uint64_t i = lo; i < hi; i++) s += sums[i];
    // inclusive wave scan
    double x = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    double woff = 0.0;
    for (int w = 0; w < wave; w++) woff += wsum[w];
    double run = woff + (x - s);
    for (uint64_t i = lo; i < hi; i++) {
        run += sums[i];
        incl[i] = run;
    }
}

// One block per shot: upper_bound over the tile prefix, then a block scan of
// the chosen tile; the first index with u < cdf_i wins.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_sample(const cplx_t<R> *__restrict__ a, int tile_bits,
                                                        uint64_t ntiles,
                                                        const double *__restrict__ tile_incl,
                                                        const double *__restrict__ tile_sums,
                                                        const double *__restrict__ us, int nshots,
                                                        uint64_t *__restrict__ out)
{
    __shared__ uint64_t s_tile;
    __shared__ int s_first;
    __shared__ double wtot[4];
    const int shot = blockIdx.x;
    if (shot >= nshots) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double u = us[shot];
    const int tsize = 1 << tile_bits;
    const int ept = tsize >= kRedThreads ? tsize / kRedThreads : 1;

    if (tid == 0) {
        uint64_t lo = 0, hi = ntiles;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (u < tile_incl[mid]) hi = mid;
            else lo = mid + 1;
        }
        s_tile = lo;
    }
    __syncthreads();
    for (uint64_t t = s_tile; t < ntiles; t++) {
        const cplx_t<R> *ta = a + (t << tile_bits);
        const double base = t ? tile_incl[t - 1] : 0.0;
        const int first_idx = tid * ept;
        double p[kEpt];
        double s = 0.0;
        for (int e = 0; e < ept; e++) {
            const int i = first_idx + e;
            p[e] = i < tsize ? prob<R>(ta[i]) : 0.0;
            s += p[e];
        }
        double x = s;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wtot[wave] = x;
        if (tid == 0) s_first = INT_MAX;
        __syncthreads();
        double c = base;
        for (int w = 0; w < wave; w++) c += wtot[w];
        c += x - s;
        int mine = INT_MAX;
        for (int e = 0; e < ept; e++) {
            c += p[e];
            if (u < c && first_idx + e < tsize) { mine = first_idx + e; break; }
        }
        if (mine != INT_MAX) atomicMin(&s_first, mine);
        __syncthreads();
        const int f = s_first;
        __syncthreads();
        if (f != INT_MAX) {
            if (tid == 0) out[shot] = (t << tile_bits) + (uint64_t)f;
            return;
        }
    }
    // u >= total mass: the last index with nonzero probability, else 0
    // (measurement.c:292-297).
    if (tid == 0) {
        uint64_t t = ntiles;
        while (t > 0 && !(tile_sums[t - 1] > 0.0)) t--;
        s_tile = t;   // one past the last tile with mass, 0 when none
        s_first = -1;
    }
    __syncthreads();
    const uint64_t t = s_tile;
    if (t == 0) {
        if (tid == 0) out[shot] = 0;
        return;
    }
    const cplx_t<R> *ta = a + ((t - 1) << tile_bits);
    int mine = -1;
    for (int i = tid; i < tsize; i += blockDim.x)
        if (prob<R>(ta[i]) > 0.0) mine = i;
    if (mine >= 0) atomicMax(&s_first, mine);
    __syncthreads();
    if (tid == 0) out[shot] = ((t - 1) << tile_bits) + (uint64_t)(s_first < 0 ? 0 : s_first);
}

template <typename R>
__global__ __launch_bounds__(256) void k_probs(const cplx_t<R> *__restrict__ a, uint64_t dim,
                                               double *__restrict__ p)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride)
        p[i] = prob<R>(a[i]);
}

// Block-partial sums of a per-index quantity over a fixed grid-stride walk.
template <int NV>
__device__ __forceinline__ void block_store(double (&v)[NV], double *partial)
{
    __shared__ double red[16][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) red[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w][threadIdx.x];
        partial[(size_t)blockIdx.x * NV + threadIdx.x] = s;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_xy(const cplx_t<R> *__restrict__ a, uint64_t npairs, int q,
                                            double *__restrict__ partial)
{
    double v[2] = {0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += stride) {
        const uint64_t i = ins0(p, q), j = i | (1ull << q);
        const double ar = a[i].x, ai = a[i].y, br = a[j].x, bi = a[j].y;
        v[0] += fmar(ar, br, ai * bi);
        v[1] += fmar(ar, bi, -(ai * br));
    }
    block_store<2>(v, partial);
}

template <typename R>
__global__ __launch_bounds__(256) void k_zmask(const cplx_t<R> *__restrict__ a, uint64_t dim,
                                               uint64_t mask, double *__restrict__ partial)
{
    double v[1] = {0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        const double pr = prob<R>(a[i]);
        v[0] += (__popcll(i & mask) & 1) ? -pr : pr;
    }
    block_store<1>(v, partial);
}

// <P> for P = i^ny X^x Z^z (Y = iXZ on the qubits in x & z; x != 0).  Pairs
// (i, i ^ x) with bit p = lowest set bit of x clear read each amplitude once:
// with w = a_i conj(a_j) and s = (-1)^popcount(i & z), the pair contributes
// 2 s Re w (ny = 0 mod 4), -2 s Im w (1), -2 s Re w (2), 2 s Im w (3),
// i.e. s (cr Re w + ci Im w).
template <typename R>
__global__ __launch_bounds__(256) void k_pauli(const cplx_t<R> *__restrict__ a, uint64_t npairs, int p,
                                               uint64_t x, uint64_t z, double cr, double ci,
                                               double *__restrict__ partial)
{
    double v[1] = {0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < npairs; k += stride) {
        const uint64_t i = ins0(k, p), j = i ^ x;
        const double ar = a[i].x, ai = a[i].y, br = a[j].x, bi = a[j].y;
        const double wr = fmar(ar, br, ai * bi);
        const double wi = fmar(ai, br, -(ar * bi));
        const double c = fmar(cr, wr, ci * wi);
        v[0] += (__popcll(i & z) & 1) ? -c : c;
    }
    block_store<1>(v, partial);
}

template <typename R, typename R2>
__global__ __launch_bounds__(256) void k_inner(const cplx_t<R> *__restrict__ a,
                                               const cplx_t<R2> *__restrict__ b, uint64_t dim,
                                               double *__restrict__ partial)
{
    double v[2] = {0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
        v[0] += fmar(ar, br, ai * bi);
        v[1] += fmar(ar, bi, -(ai * br));
    }
    block_store<2>(v, partial);
}

inline int small_grid(uint64_t work, int blocks)
{
    const uint64_t need = (work + 255) / 256;
    return (int)(need < (uint64_t)blocks ? (need ? need : 1) : blocks);
}

}  // namespace

void reduce_geometry(int n, int *tile_bits, uint64_t *ntiles, int *blocks)
{
    if (n < kTileBits) {
        *tile_bits = n;
        *ntiles = 1;
        *blocks = 1;
        return;
    }
    *tile_bits = kTileBits;
    *ntiles = 1ull << (n - kTileBits);
    *blocks = (int)(*ntiles < (uint64_t)kMaxRedBlocks ? *ntiles : kMaxRedBlocks);
}

hipError_t launch_reduce_z(const SvDev &s, const ReduceWs &w, bool want_tiles)
{
    double *ts = want_tiles ? w.tile_sums : nullptr;
    if (s.n < kTileBits) {
        if (s.prec == MOONLAB_HIP_C64)
            k_reduce_small<float><<<1, kRedThreads, 0, s.stream>>>((const float2 *)s.amps, s.n, w.partial, ts);
        else
            k_reduce_small<double><<<1, kRedThreads, 0, s.stream>>>((const double2 *)s.amps, s.n, w.partial, ts);
    } else {
        if (s.prec == MOONLAB_HIP_C64)
            k_reduce_tiles<float><<<w.blocks, kRedThreads, 0, s.stream>>>(
                (const float2 *)s.amps, s.n, w.ntiles, w.partial, ts);
        else
            k_reduce_tiles<double><<<w.blocks, kRedThreads, 0, s.stream>>>(
                (const double2 *)s.amps, s.n, w.ntiles, w.partial, ts);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_final<<<1, 64, 0, s.stream>>>(w.partial, s.n < kTileBits ? 1 : w.blocks, s.n + 1, w.out);
    return hipGetLastError();
}

hipError_t launch_tile_scan(const ReduceWs &w, hipStream_t st)
{
    k_tile_scan<<<1, 1024, 0, st>>>(w.tile_sums, w.ntiles, w.tile_incl);
    return hipGetLastError();
}

hipError_t launch_sample(const SvDev &s, const ReduceWs &w, const double *d_u, int nshots,
                         uint64_t *d_out)
{
    if (nshots <= 0) return hipSuccess;
    if (s.prec == MOONLAB_HIP_C64)
        k_sample<float><<<nshots, kRedThreads, 0, s.stream>>>((const float2 *)s.amps, w.tile_bits, w.ntiles,
                                                               w.tile_incl, w.tile_sums, d_u, nshots, d_out);
    else
        k_sample<double><<<nshots, kRedThreads, 0, s.stream>>>((const double2 *)s.amps, w.tile_bits, w.ntiles,
                                                                w.tile_incl, w.tile_sums, d_u, nshots, d_out);
    return hipGetLastError();
}

hipError_t launch_probabilities(const SvDev &s, double *d_p)
{
    const uint64_t dim = 1ull << s.n;
    const uint64_t nb = (dim + 255) / 256;
    const unsigned g = (unsigned)(nb < 65536 ? nb : 65536);
    if (s.prec == MOONLAB_HIP_C64)
        k_probs<float><<<g, 256, 0, s.stream>>>((const float2 *)s.amps, dim, d_p);
    else
        k_probs<double><<<g, 256, 0, s.stream>>>((const double2 *)s.amps, dim, d_p);
    return hipGetLastError();
}

hipError_t launch_xy(const SvDev &s, int q, double *d_partial, int blocks, double *d_out)
{
    const uint64_t np = (1ull << s.n) >> 1;
    const int g = small_grid(np, blocks);
    if (s.prec == MOONLAB_HIP_C64)
        k_xy<float><<<g, 256, 0, s.stream>>>((const float2 *)s.amps, np, q, d_partial);
    else
        k_xy<double><<<g, 256, 0, s.stream>>>((const double2 *)s.amps, np, q, d_partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_final<<<1, 64, 0, s.stream>>>(d_partial, g, 2, d_out);
    return hipGetLastError();
}

hipError_t launch_pauli(const SvDev &s, uint64_t x, uint64_t z, double *d_partial, int blocks, double *d_out)
{
    const int ny = __builtin_popcountll(x & z) & 3;
    const double cr = ny == 0 ? 2.0 : ny == 2 ? -2.0 : 0.0;
    const double ci = ny == 1 ? -2.0 : ny == 3 ? 2.0 : 0.0;
    const int p = __builtin_ctzll(x);
    const uint64_t np = (1ull << s.n) >> 1;
    const int g = small_grid(np, blocks);
    if (s.prec == MOONLAB_HIP_C64)
        k_pauli<float><<<g, 256, 0, s.stream>>>((const float2 *)s.amps, np, p, x, z, cr, ci, d_partial);
    else
        k_pauli<double><<<g, 256, 0, s.stream>>>((const double2 *)s.amps, np, p, x, z, cr, ci, d_partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_final<<<1, 64, 0, s.stream>>>(d_partial, g, 1, d_out);
    return hipGetLastError();
}

hipError_t launch_zmask(const SvDev &s, uint64_t mask, double *d_partial, int blocks, double *d_out)
{
    const uint64_t dim = 1ull << s.n;
    const int g = small_grid(dim, blocks);
    if (s.prec == MOONLAB_HIP_C64)
        k_zmask<float><<<g, 256, 0, s.stream>>>((const float2 *)s.amps, dim, mask, d_partial);
    else
        k_zmask<double><<<g, 256, 0, s.stream>>>((const double2 *)s.amps, dim, mask, d_partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_final<<<1, 64, 0, s.stream>>>(d_partial, g, 1, d_out);
    return hipGetLastError();
}

hipError_t launch_inner(const SvDev &a, const SvDev &b, double *d_partial, int blocks, double *d_out)
{
    const uint64_t dim = 1ull << a.n;
    const int g = small_grid(dim, blocks);
    // products and sums in FP64 whatever the storage precision of either state
    const bool fa = a.prec == MOONLAB_HIP_C64, fb = b.prec == MOONLAB_HIP_C64;
    if (fa && fb)
        k_inner<float, float><<<g, 256, 0, a.stream>>>((const float2 *)a.amps, (const float2 *)b.amps, dim, d_partial);
    else if (fa)
        k_inner<float, double><<<g, 256, 0, a.stream>>>((const float2 *)a.amps, (const double2 *)b.amps, dim,
                                                        d_partial);
    else if (fb)
        k_inner<double, float><<<g, 256, 0, a.stream>>>((const double2 *)a.amps, (const float2 *)b.amps, dim,
                                                        d_partial);
    else
        k_inner<double, double><<<g, 256, 0, a.stream>>>((const double2 *)a.amps, (const double2 *)b.amps, dim,
                                                         d_partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_final<<<1, 64, 0, a.stream>>>(d_partial, g, 2, d_out);
    return hipGetLastError();
}

}  // namespace mlx

// ---------------------------------------------------------------------------
// Norm and <Z_q> without allocating: the block partials of k_reduce_tiles /
// k_reduce_small live in a fixed device array instead of a state's workspace.
//
// kNoAllocMaxQubits bounds the array: kMaxRedBlocks blocks x (n + 1) FP64
// partials for n <= 35 (complex64's limit) is 2048 x 36 x 8 B = 576 KiB per
// GPU, held from the first call on a device until the process ends.  Every
// partial a call reads is written earlier in the same call, so nothing needs
// clearing between calls.  All calls on one GPU share the array: calls that
// overlap in time on one GPU (two streams, two threads) race on it.
namespace {

constexpr int kNoAllocMaxQubits = MOONLAB_HIP_MAX_QUBITS_C64;
constexpr size_t kScratchDoubles = (size_t)mlx::kMaxRedBlocks * (kNoAllocMaxQubits + 1);

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices];   // per device: the array's device address

// The array's address on the current device, looked up once per device with
// that device current.  In host code `scratch` names a host-side stand-in,
// never device memory: only this address may reach a kernel or a copy.
hipError_t scratch_on_current_device(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

__global__ void k_z_from_sums(const double *__restrict__ partial, int blocks, int nvals, double *__restrict__ out)
{
    const int v = threadIdx.x;
    if (v >= nvals) return;
    double s = 0.0;
    for (int b = 0; b < blocks; b++) s += partial[(size_t)b * nvals + v];
    // 1 - 2 P(1): 2 s is exact, so this rounds once, as 1.0 - 2.0 * s does
    out[v] = v == 0 ? s : __builtin_fma(-2.0, s, 1.0);
}

}  // namespace

extern "C" moonlab_hip_status_t moonlab_hip_expect_z_all_noalloc(const void *amplitudes, uint32_t n_qubits,
                                                                 int precision, void *stream, double *out)
{
    using namespace mlx;
    if (!amplitudes || !out) return set_last_error(MOONLAB_HIP_ERR_BAD_ARG, "expect_z_all_noalloc: NULL pointer");
    if (precision != MOONLAB_HIP_C128 && precision != MOONLAB_HIP_C64)
        return set_last_error(MOONLAB_HIP_ERR_BAD_ARG, "expect_z_all_noalloc: precision is not C128 or C64");
    const int cap = precision == MOONLAB_HIP_C64 ? MOONLAB_HIP_MAX_QUBITS_C64 : MOONLAB_HIP_MAX_QUBITS_C128;
    if (n_qubits < 1 || (int)n_qubits > cap || (int)n_qubits > kNoAllocMaxQubits)
        return set_last_error(MOONLAB_HIP_ERR_BAD_ARG,
                              "expect_z_all_noalloc: n_qubits out of range (1..34 complex128, 1..35 complex64; "
                              "the fixed partial-sum array holds 2048 x 36 doubles)");
    double *partial = nullptr;
    hipError_t e = scratch_on_current_device(&partial);
    if (e != hipSuccess) return set_last_error(MOONLAB_HIP_ERR_DRIVER, "expect_z_all_noalloc: scratch lookup", e);
    const int n = (int)n_qubits;
    int tile_bits = 0, blocks = 0;
    uint64_t ntiles = 0;
    reduce_geometry(n, &tile_bits, &ntiles, &blocks);
    const hipStream_t st = (hipStream_t)stream;
    // the same kernels, grid and partial layout as moonlab_hip_expect_z_all,
    // so the sums are bit-identical
    if (n < kTileBits) {
        if (precision == MOONLAB_HIP_C64)
            k_reduce_small<float><<<1, kRedThreads, 0, st>>>((const float2 *)amplitudes, n, partial, nullptr);
        else
            k_reduce_small<double><<<1, kRedThreads, 0, st>>>((const double2 *)amplitudes, n, partial, nullptr);
    } else {
        if (precision == MOONLAB_HIP_C64)
            k_reduce_tiles<float><<<blocks, kRedThreads, 0, st>>>((const float2 *)amplitudes, n, ntiles, partial,
                                                                  nullptr);
        else
            k_reduce_tiles<double><<<blocks, kRedThreads, 0, st>>>((const double2 *)amplitudes, n, ntiles,
                                                                   partial, nullptr);
    }
    e = hipGetLastError();
    if (e == hipSuccess) {
        k_z_from_sums<<<1, 64, 0, st>>>(partial, blocks, n + 1, out);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return set_last_error(MOONLAB_HIP_ERR_DRIVER, "expect_z_all_noalloc: launch", e);
    return MOONLAB_HIP_OK;
}
