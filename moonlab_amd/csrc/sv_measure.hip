// sv_measure.hip -- collapsing measurements on the device: the outcome
// probabilities of k measured qubits (FP64, fixed summation order) and the
// collapse / renormalization pass.
//
// Reference counterparts: measurement_single_qubit, measurement_all_qubits
// and measurement_partial (src/quantum/measurement.c:181-231, 263-307,
// 320-410), which pull the state to the host and scan it sequentially.  Here
// the 2^k outcome sums are one pass over the state (each workgroup sums one
// outcome's amplitudes over a contiguous range of the other qubits' index),
// the host picks the outcome from 2^k doubles with the reference's rule, and
// a second pass zeroes the other outcomes and rescales the kept amplitudes by
// 1 / sqrt(norm) computed in FP64.
#include "sv_device.h"
#include "sv_internal.h"

#include <algorithm>

namespace mlx {
namespace {

constexpr int kMeasThreads = 256;
constexpr int kMeasMaxQubits = 20;   // outcome table of 2^k doubles

struct MeasGeom {
    int32_t k;
    int8_t  q[kMeasMaxQubits];    // outcome bit j = qubit q[j]
    int8_t  qs[kMeasMaxQubits];   // the same qubits, ascending
};

__device__ __forceinline__ double wave_total(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block b sums |a_i|^2 (FP64, re*re + im*im) for outcome o = b >> cb over
// the index range c = b & (2^cb - 1) of the unmeasured qubits; partial is
// laid out [c][o].
template <typename R>
__global__ __launch_bounds__(kMeasThreads) void k_outcome_probs(const cplx_t<R> *__restrict__ a, int n,
                                                                 MeasGeom g, int cb,
                                                                 double *__restrict__ partial)
{
    __shared__ double wsum[kMeasThreads / 64];
    const uint64_t b = blockIdx.x;
    const uint64_t o = b >> cb, c = b & ((1ull << cb) - 1ull);
    const uint64_t per = 1ull << (n - g.k - cb);
    uint64_t dep = 0;
    for (int j = 0; j < g.k; j++)
        if ((o >> j) & 1ull) dep |= 1ull << g.q[j];
    double acc = 0.0;
    for (uint64_t r = c * per + threadIdx.x; r < (c + 1) * per; r += kMeasThreads) {
        uint64_t i = r;
        for (int j = 0; j < g.k; j++) i = ins0(i, g.qs[j]);
        const cplx_t<R> x = a[i | dep];
        const double re = (double)x.x, im = (double)x.y;
        acc += fmar(re, re, im * im);
    }
    acc = wave_total(acc);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kMeasThreads / 64; w++) s += wsum[w];
        partial[(size_t)c * (1ull << g.k) + o] = s;
    }
}

// out[o] = sum over c of partial[c][o], c ascending
__global__ void k_sum_columns(const double *__restrict__ partial, uint64_t rows, uint64_t cols,
                              double *__restrict__ out)
{
    const uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= cols) return;
    double s = 0.0;
    for (uint64_t c = 0; c < rows; c++) s += partial[c * cols + o];
    out[o] = s;
}

// mode 0: zero the amplitudes whose masked bits differ from want; mode 1:
// also multiply the kept ones by inv (FP64); mode 2: the basis state want
// (mask = every bit) with amplitude exactly 1.
template <typename R>
__global__ __launch_bounds__(256) void k_collapse(cplx_t<R> *__restrict__ a, uint64_t dim, uint64_t mask,
                                                  uint64_t want, double inv, int mode)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        cplx_t<R> x;
        if ((i & mask) != want) {
            x.x = (R)0;
            x.y = (R)0;
        } else if (mode == 2) {
            x.x = (R)1;
            x.y = (R)0;
        } else {
            x = a[i];
            if (mode == 1) {
                x.x = (R)((double)x.x * inv);
                x.y = (R)((double)x.y * inv);
            }
        }
        a[i] = x;
    }
}

}  // namespace

hipError_t launch_outcome_probs(const SvDev &s, const int *qubits, int k, double *host_out)
{
    if (k < 1 || k > kMeasMaxQubits || k > s.n) return hipErrorInvalidValue;
    MeasGeom g{};
    g.k = k;
    for (int j = 0; j < k; j++) g.q[j] = (int8_t)qubits[j];
    std::copy(g.q, g.q + k, g.qs);
    std::sort(g.qs, g.qs + k);
    const int rest = s.n - k;
    const int cb = rest > 12 ? rest - 12 : 0;   // at most 4096 amplitudes per workgroup
    const uint64_t nout = 1ull << k, rows = 1ull << cb;
    double *d = nullptr;
    hipError_t e = hipMallocAsync((void **)&d, (size_t)(nout * rows + nout) * sizeof(double), s.stream);
    if (e != hipSuccess) return e;
    const uint64_t blocks = nout * rows;
    if (s.prec == MOONLAB_HIP_C64)
        k_outcome_probs<float><<<(unsigned)blocks, kMeasThreads, 0, s.stream>>>((const cplx_t<float> *)s.amps, s.n,
                                                                               g, cb, d);
    else
        k_outcome_probs<double><<<(unsigned)blocks, kMeasThreads, 0, s.stream>>>(
            (const cplx_t<double> *)s.amps, s.n, g, cb, d);
    e = hipGetLastError();
    if (e == hipSuccess) {
        k_sum_columns<<<(unsigned)((nout + 255) / 256), 256, 0, s.stream>>>(d, rows, nout, d + nout * rows);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(host_out, d + nout * rows, (size_t)nout * sizeof(double), hipMemcpyDeviceToHost,
                           s.stream);
    hipFreeAsync(d, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    return e;
}

hipError_t launch_collapse(const SvDev &s, uint64_t mask, uint64_t want, double inv, int mode)
{
    const uint64_t dim = 1ull << s.n;
    const uint64_t nb = (dim + 255) / 256;
    const unsigned grid = (unsigned)(nb < 65536 ? nb : 65536);
    if (s.prec == MOONLAB_HIP_C64)
        k_collapse<float><<<grid, 256, 0, s.stream>>>((cplx_t<float> *)s.amps, dim, mask, want, inv, mode);
    else
        k_collapse<double><<<grid, 256, 0, s.stream>>>((cplx_t<double> *)s.amps, dim, mask, want, inv, mode);
    return hipGetLastError();
}

}  // namespace mlx
