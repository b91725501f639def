// sv_gram.hip -- Gram matrix of a set of device states: G[a][b] = <S_a|S_b>
// in FP64, the overlaps the quantum geometric tensor is assembled from
// (the x-loops of vqe_compute_qgt / vqe_compute_berry_curvature,
// src/algorithms/vqe_qng.c:172-245, which the reference runs on the host
// once per (i, j) pair).
//
// S holds K states of `dim` complex128 amplitudes each, state k at
// S + k * dim.  A 256-thread workgroup computes one 16 x 16 block (a in rows
// 16 bi.., b in columns 16 bj..) of the upper triangle: the amplitudes of its
// 16 + 16 states stream through LDS in chunks of 64, and thread (r, c)
// accumulates conj(S_a[x]) S_b[x] over x in index order, so every entry is
// reproducible.  Blocks below the diagonal return at once; entries with
// a > b inside a diagonal block are written too (the host reads a <= b).
#include "sv_device.h"
#include "sv_internal.h"

namespace mlx {
namespace {

constexpr int kGramTile = 16;
constexpr int kGramChunk = 64;

__global__ __launch_bounds__(256) void k_gram(const double2 *__restrict__ S, int K, uint64_t dim,
                                              double2 *__restrict__ G)
{
    const int bi = blockIdx.y, bj = blockIdx.x;
    if (bi > bj) return;
    __shared__ double2 As[kGramTile][kGramChunk];
    __shared__ double2 Bs[kGramTile][kGramChunk + 1];   // +1: column reads of 16 rows hit 16 banks
    const int tid = threadIdx.x, r = tid / kGramTile, c = tid % kGramTile;
    const int a = bi * kGramTile + r, b = bj * kGramTile + c;
    double sr = 0.0, si = 0.0;
    for (uint64_t x0 = 0; x0 < dim; x0 += kGramChunk) {
        // 16 x 64 amplitudes of each side, 4 per thread, rows of consecutive x
        for (int e = tid; e < kGramTile * kGramChunk; e += 256) {
            const int row = e / kGramChunk, col = e % kGramChunk;
            const uint64_t x = x0 + (uint64_t)col;
            const int ka = bi * kGramTile + row, kb = bj * kGramTile + row;
            As[row][col] = (ka < K && x < dim) ? S[(uint64_t)ka * dim + x] : double2{0.0, 0.0};
            Bs[row][col] = (kb < K && x < dim) ? S[(uint64_t)kb * dim + x] : double2{0.0, 0.0};
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < kGramChunk; k++) {
            const double2 u = As[r][k], v = Bs[c][k];
            sr = fmar(u.x, v.x, fmar(u.y, v.y, sr));       // Re conj(u) v
            si = fmar(u.x, v.y, fmar(-u.y, v.x, si));      // Im conj(u) v
        }
        __syncthreads();
    }
    if (a < K && b < K) G[(uint64_t)a * K + b] = double2{sr, si};
}

}  // namespace

hipError_t launch_gram(const double2 *S, int K, uint64_t dim, double2 *G, hipStream_t st)
{
    if (K < 1) return hipSuccess;
    const int nb = (K + kGramTile - 1) / kGramTile;
    k_gram<<<dim3((unsigned)nb, (unsigned)nb, 1), 256, 0, st>>>(S, K, dim, G);
    return hipGetLastError();
}

}  // namespace mlx
