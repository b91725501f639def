// sv_dense.hip -- dense k-qubit (k = 3..5) block application.
//
// One thread per group of 2^k amplitudes that share every bit outside the
// block's qubits: gather, 2^k x 2^k complex matvec against the matrix held
// in LDS, scatter.  Consecutive threads take consecutive values of the free
// bits, so each of the 2^k loads of a wave covers one contiguous run.
#include "sv_device.h"
#include "sv_internal.h"

#include <vector>

namespace mlx {
namespace {

template <typename R, int K>
__global__ __launch_bounds__(256) void k_dense(cplx_t<R> *__restrict__ a, uint64_t ngroups, BitSet fixed,
                                               const cplx_t<R> *__restrict__ mat)
{
    constexpr int D = 1 << K;
    __shared__ cplx_t<R> m[D * D];
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) m[i] = mat[i];
    __syncthreads();
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const uint64_t base = deposit(g, fixed);
    uint64_t off[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
        uint64_t o = 0;
#pragma unroll
        for (int b = 0; b < K; b++)
            if (j & (1 << b)) o |= 1ull << fixed.pos[b];
        off[j] = base | o;
    }
    cplx_t<R> v[D];
#pragma unroll
    for (int j = 0; j < D; j++) v[j] = a[off[j]];
#pragma unroll
    for (int r = 0; r < D; r++) {
        R sr = 0, si = 0;
#pragma unroll
        for (int c = 0; c < D; c++) {
            const cplx_t<R> mm = m[r * D + c];
            sr = fmar(mm.x, v[c].x, fmar(-mm.y, v[c].y, sr));
            si = fmar(mm.x, v[c].y, fmar(mm.y, v[c].x, si));
        }
        a[off[r]] = cplx_t<R>{sr, si};
    }
}

template <typename R, int K>
hipError_t run_dense(const SvDev &s, const int *qubits, const void *d_mat)
{
    BitSet fixed;
    fixed.nb = K;
    for (int j = 0; j < K; j++) fixed.pos[j] = qubits[j];
    const uint64_t ng = (1ull << s.n) >> K;
    const unsigned grid = (unsigned)((ng + 255) / 256);
    k_dense<R, K><<<grid, 256, 0, s.stream>>>((cplx_t<R> *)s.amps, ng, fixed, (const cplx_t<R> *)d_mat);
    return hipGetLastError();
}

}  // namespace

bool mfma_dense_supported(const SvDev &s, int k);
hipError_t launch_dense_mfma(const SvDev &s, int k, const int *qubits, const void *d_mat);

// d_mat: device copy of the 2^k x 2^k matrix in the state's precision.  The
// matrix cores take every block whose tile (2^(k+6) amplitudes) fits the
// state; tiny states use the VALU kernel.
hipError_t launch_dense_dev(const SvDev &s, int k, const int *qubits, const void *d_mat)
{
    if (mfma_dense_supported(s, k)) return launch_dense_mfma(s, k, qubits, d_mat);
    const bool f = s.prec == MOONLAB_HIP_C64;
    switch (k) {
    case 3: return f ? run_dense<float, 3>(s, qubits, d_mat) : run_dense<double, 3>(s, qubits, d_mat);
    case 4: return f ? run_dense<float, 4>(s, qubits, d_mat) : run_dense<double, 4>(s, qubits, d_mat);
    case 5: return f ? run_dense<float, 5>(s, qubits, d_mat) : run_dense<double, 5>(s, qubits, d_mat);
    default: return hipErrorInvalidValue;
    }
}

// Host matrix (FP64 (re, im) pairs) -> device, in the state's precision, then apply.
// The copy is stream-ordered, so the buffer is released with hipFreeAsync after
// the kernel on the same stream.
hipError_t launch_dense(const SvDev &s, int k, const int *qubits, const double *mat)
{
    const size_t cnt = (size_t)1 << (2 * k);
    const size_t bytes = cnt * (s.prec == MOONLAB_HIP_C64 ? 8 : 16);
    void *d = nullptr;
    hipError_t e = hipMallocAsync(&d, bytes, s.stream);
    if (e != hipSuccess) return e;
    if (s.prec == MOONLAB_HIP_C64) {
        std::vector<float> h(2 * cnt);
        for (size_t i = 0; i < 2 * cnt; i++) h[i] = (float)mat[i];
        e = hipMemcpyAsync(d, h.data(), bytes, hipMemcpyHostToDevice, s.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s.stream);   // h is about to go out of scope
    } else {
        e = hipMemcpyAsync(d, mat, bytes, hipMemcpyHostToDevice, s.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    }
    if (e == hipSuccess) e = launch_dense_dev(s, k, qubits, d);
    hipFreeAsync(d, s.stream);
    return e;
}

}  // namespace mlx
