// sv_device.h -- device-side building blocks shared by the state-vector kernels.
//
// Amplitudes are interleaved (re, im): double2 for complex128, float2 for
// complex64.  Qubit q is bit q of the basis index (qubit 0 = LSB), the
// reference's convention (src/quantum/state.h:102-131).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlx {

template <typename R> struct Cplx;
template <> struct Cplx<double> { using type = double2; };
template <> struct Cplx<float> { using type = float2; };
template <typename R> using cplx_t = typename Cplx<R>::type;

// Fused multiply-add with the operation order fixed in the source, so the
// compiler's contraction choices cannot differ between kernel instantiations
// (launch shapes, tile sizes): every kernel that applies the same operator
// produces bit-identical amplitudes.
__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// c ? x : y per component: a select of whole structs can become a
// dynamically indexed stack slot (scratch memory) instead of two v_cndmasks.
template <typename T>
__device__ __forceinline__ T csel(bool c, const T &x, const T &y)
{
    T r;
    r.x = c ? x.x : y.x;
    r.y = c ? x.y : y.y;
    return r;
}

// Complex multiply (x.re, x.im) * (pr, pi) with a fixed FMA pattern.
template <typename T, typename R>
__device__ __forceinline__ T cmul(const T &x, R pr, R pi)
{
    T y;
    y.x = fmar(x.x, pr, -(x.y * pi));
    y.y = fmar(x.x, pi, x.y * pr);
    return y;
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// waits for every outstanding global load and store of the wave, which
// drains a prefetch of the next tile and the write-back of the last one.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Streaming (non-temporal) loads and stores: a gate pass over a state larger
// than the last-level cache has no reuse, and marking the traffic streaming
// lets L2 and the MALL skip allocating it.
typedef double mlx_v2d __attribute__((ext_vector_type(2)));
typedef float mlx_v2f __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ double2 ld(const double2 *p)
{
    if constexpr (NT) {
        const mlx_v2d v = __builtin_nontemporal_load((const mlx_v2d *)p);
        return double2{v.x, v.y};
    } else {
        return *p;
    }
}
template <bool NT>
__device__ __forceinline__ float2 ld(const float2 *p)
{
    if constexpr (NT) {
        const mlx_v2f v = __builtin_nontemporal_load((const mlx_v2f *)p);
        return float2{v.x, v.y};
    } else {
        return *p;
    }
}
template <bool NT>
__device__ __forceinline__ void st(double2 *p, const double2 &x)
{
    if constexpr (NT) {
        const mlx_v2d v = {x.x, x.y};
        __builtin_nontemporal_store(v, (mlx_v2d *)p);
    } else {
        *p = x;
    }
}
template <bool NT>
__device__ __forceinline__ void st(float2 *p, const float2 &x)
{
    if constexpr (NT) {
        const mlx_v2f v = {x.x, x.y};
        __builtin_nontemporal_store(v, (mlx_v2f *)p);
    } else {
        *p = x;
    }
}

// Insert a zero bit at position q of p (pair / subspace enumeration).
// The entries of a diagonal op are read with wave-uniform loads and the one
// an amplitude needs is selected in registers (an index that differs per lane
// would turn every read into a vector memory load).
// readfirstlane moves a value every lane holds into a scalar register; a
// select between such values stays a select instead of becoming a load from
// a selected address (which would keep the entries in scratch memory).
__device__ __forceinline__ float wave_uniform(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double wave_uniform(double x)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// entry k of four values held in registers
template <typename R>
__device__ __forceinline__ R pick4(int k, R a, R b, R c, R d)
{
    const R lo = (k & 1) ? b : a, hi = (k & 1) ? d : c;
    return (k & 2) ? hi : lo;
}

// The four entries of a diagonal op (amplitude *= d[k]), wave-uniform.
#define MLX_DIAG4(R, op)                                                                                  \
    const R dr0 = wave_uniform((R)(op).m[0]), di0 = wave_uniform((R)(op).m[1]);                           \
    const R dr1 = wave_uniform((R)(op).m[2]), di1 = wave_uniform((R)(op).m[3]);                           \
    const R dr2 = wave_uniform((R)(op).m[4]), di2 = wave_uniform((R)(op).m[5]);                           \
    const R dr3 = wave_uniform((R)(op).m[6]), di3 = wave_uniform((R)(op).m[7])

// The tiles a workgroup takes: t, t + step, ... below end.  The
// dispatcher hands consecutive workgroups to the 8 XCDs in turn, so when the
// grid and the tile count split evenly, workgroup b (on XCD b % 8) walks the
// b % 8-th eighth of the tiles and each XCD streams one contiguous range;
// otherwise workgroup b takes b, b + grid, ...
struct TileWalk {
    uint64_t t, step, end;
};
__device__ __forceinline__ TileWalk tile_walk(uint64_t ntiles)
{
    const uint64_t g = gridDim.x, b = blockIdx.x;
    if ((g & 7) == 0 && (ntiles & 7) == 0 && ntiles >= g) {
        const uint64_t per = ntiles >> 3;
        return TileWalk{(b & 7) * per + (b >> 3), g >> 3, ((b & 7) + 1) * per};
    }
    return TileWalk{b, g, ntiles};
}

__device__ __forceinline__ uint64_t ins0(uint64_t p, int q)
{
    const uint64_t lo = p & ((1ull << q) - 1ull);
    return ((p - lo) << 1) | lo;
}

// Up to 7 fixed bit positions, ascending.  Depositing a sub-index into the
// complement of these positions enumerates every basis index whose fixed
// bits are all zero.
struct BitSet {
    int nb;
    int pos[7];
};

__device__ __forceinline__ uint64_t deposit(uint64_t p, const BitSet &bs)
{
    for (int k = 0; k < bs.nb; k++) p = ins0(p, bs.pos[k]);
    return p;
}

// ---------------------------------------------------------------------------
// Pair operators: (a0, a1) -> (a0', a1') for the two amplitudes that differ
// only in the target bit.  Each mirrors the per-pair arithmetic of the
// reference gate it implements (src/quantum/gates.c lines cited).
// ---------------------------------------------------------------------------

template <typename R> struct OpSwap {            // X, CNOT, Toffoli (gates.c:123-154, 571-639)
    __device__ __forceinline__ void operator()(cplx_t<R> &a, cplx_t<R> &b) const
    {
        const cplx_t<R> t = a; a = b; b = t;
    }
};

template <typename R> struct OpY {               // Y = [[0,-i],[i,0]] (gates.c:156-192)
    __device__ __forceinline__ void operator()(cplx_t<R> &a, cplx_t<R> &b) const
    {
        const cplx_t<R> a0 = a;
        a.x = b.y;  a.y = -b.x;
        b.x = -a0.y; b.y = a0.x;
    }
};

template <typename R> struct OpNegY {            // -Y = S^dag X S: the job path's CY body
    __device__ __forceinline__ void operator()(cplx_t<R> &a, cplx_t<R> &b) const
    {                                             // (moonlab_qgtl_backend.c:138-144)
        const cplx_t<R> a0 = a;
        a.x = -b.y; a.y = b.x;
        b.x = a0.y; b.y = -a0.x;
    }
};

template <typename R> struct OpH {               // (a0 +- a1) * INV_SQRT2 (simd_ops.c:1026-1103)
    R r;
    __device__ __forceinline__ void operator()(cplx_t<R> &a, cplx_t<R> &b) const
    {
        const cplx_t<R> a0 = a;
        a.x = (a0.x + b.x) * r; a.y = (a0.y + b.y) * r;
        b.x = (a0.x - b.x) * r; b.y = (a0.y - b.y) * r;
    }
};

template <typename R> struct OpRX {              // [[c,-is],[-is,c]] (gates.c:385-427)
    R c, s;
    __device__ __forceinline__ void operator()(cplx_t<R> &a, cplx_t<R> &b) const
    {
        const cplx_t<R> a0 = a;
        const cplx_t<R> b0 = b;
        a.x = fmar(c, a0.x, s * b0.y);  a.y = fmar(c, a0.y, -(s * b0.x));
        b.x = fmar(c, b0.x, s * a0.y);  b.y = fmar(c, b0.y, -(s * a0.x));
    }
};

template <typename R> struct OpRY {              // [[c,-s],[s,c]] (gates.c:429-479)
    R c, s;
    __device__ __forceinline__ void operator()(cplx_t<R> &a, cplx_t<R> &b) const
    {
        const cplx_t<R> a0 = a;
        const cplx_t<R> b0 = b;
        a.x = fmar(c, a0.x, -(s * b0.x));  a.y = fmar(c, a0.y, -(s * b0.y));
        b.x = fmar(s, a0.x, c * b0.x);     b.y = fmar(s, a0.y, c * b0.y);
    }
};

template <typename R> struct OpDiag2 {           // diag(d0, d1) on a pair (controlled RZ)
    R d0r, d0i, d1r, d1i;
    __device__ __forceinline__ void operator()(cplx_t<R> &a, cplx_t<R> &b) const
    {
        a = cmul(a, d0r, d0i);
        b = cmul(b, d1r, d1i);
    }
};

template <typename R> struct OpMat2 {            // general 2x2, row-major (re, im) (gates.c:1381-1390)
    R m[8];
    __device__ __forceinline__ void operator()(cplx_t<R> &a, cplx_t<R> &b) const
    {
        const cplx_t<R> a0 = a, b0 = b;
        a.x = fmar(m[0], a0.x, -(m[1] * a0.y)) + fmar(m[2], b0.x, -(m[3] * b0.y));
        a.y = fmar(m[0], a0.y, m[1] * a0.x) + fmar(m[2], b0.y, m[3] * b0.x);
        b.x = fmar(m[4], a0.x, -(m[5] * a0.y)) + fmar(m[6], b0.x, -(m[7] * b0.y));
        b.y = fmar(m[4], a0.y, m[5] * a0.x) + fmar(m[6], b0.y, m[7] * b0.x);
    }
};

// ---------------------------------------------------------------------------
// Element operators for diagonal gates.
// ---------------------------------------------------------------------------

template <typename R> struct ElNeg {             // Z, CZ, MCZ (gates.c:194-224, 641-668)
    __device__ __forceinline__ void operator()(cplx_t<R> &a) const { a.x = -a.x; a.y = -a.y; }
};

template <typename R> struct ElMulI {            // S: times +i (gates.c:309-339)
    __device__ __forceinline__ void operator()(cplx_t<R> &a) const
    {
        const R r = a.x; a.x = -a.y; a.y = r;
    }
};

template <typename R> struct ElMulNegI {         // S-dagger: times -i (gates.c:341-371)
    __device__ __forceinline__ void operator()(cplx_t<R> &a) const
    {
        const R r = a.x; a.x = a.y; a.y = -r;
    }
};

template <typename R> struct ElPhase {           // T, P, CPHASE (gates.c:276-307, 737-765)
    R pr, pi;
    __device__ __forceinline__ void operator()(cplx_t<R> &a) const
    {
        a = cmul(a, pr, pi);
    }
};

}  // namespace mlx
