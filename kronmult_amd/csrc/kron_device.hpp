// kron_device.hpp -- device-side building blocks shared by the kernels.
//
// Index convention (the reference's, kroncommon.hpp:97-199): a vector of
// n1*...*nd elements is a column-major tensor X(j_d, ..., j_1) with j_d the
// fastest index; kron(A1,...,Ad) applies A1 to j_1 (slowest) and Ad to j_d.
// Contracting mode k (1-based) means
//     out(.., i_k, ..) = sum_j A_k(i_k, j) * in(.., j, ..),
// and the modes commute, so a kernel may contract them in any order.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

#include "kron_internal.hpp"

namespace kron {

// Fused multiply-add in the operands' own type.  __builtin_fma is the double
// builtin: float operands would be widened, summed in double and narrowed
// again, three instructions per term instead of one.
__device__ __forceinline__ float kfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double kfma(double a, double b, double c)
{
    return __builtin_fma(a, b, c);
}

constexpr int cpow(int n, int p)
{
    int r = 1;
    for (int i = 0; i < p; ++i) r *= n;
    return r;
}

constexpr int pow2ceil(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Pointers in the global address space: loads become global_load with an
// SGPR base and a 32-bit VGPR offset, atomics global_atomic_add_{f64,f32}.
// (Pointers read from the pointer arrays are otherwise generic -> flat_*.)
template <typename T>
struct GlobalAS {
    typedef __attribute__((address_space(1))) T type;
};
template <typename T>
using gptr = typename GlobalAS<T>::type *;

template <typename T>
__device__ __forceinline__ gptr<T> to_global(T *p)
{
    return (gptr<T>)p;
}

// Read-only inputs (factors, pointer arrays) through the constant address
// space: a workgroup-uniform address then becomes a scalar load (s_load),
// counted by lgkmcnt rather than vmcnt.  That matters: vmcnt retires in
// order, so a vector load of a factor issued after a prefetch of the next X
// would make the wait for the factor also wait for the whole prefetch.  The
// kernels never write these arrays, which is what the address space asserts.
template <typename T>
struct ConstAS {
    typedef __attribute__((address_space(4))) const T type;
};
template <typename T>
using cptr = typename ConstAS<T>::type *;

template <typename T>
__device__ __forceinline__ cptr<T> to_const(const T *p)
{
    return (cptr<T>)p;
}

// A workgroup-uniform value moved to SGPRs (v_readfirstlane).  Only valid
// when every active lane holds the same value.
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v)
{
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
    return reinterpret_cast<T *>(uniform_u64(reinterpret_cast<unsigned long long>(p)));
}

// Hardware FP atomic add (no-return global_atomic_add_f64 / _f32 on gfx950).
// Y must be ordinary device memory (hipMalloc / torch), where these are exact
// read-modify-write adds at device scope.
__device__ __forceinline__ void atomic_add(gptr<double> p, double v)
{
    __builtin_amdgcn_global_atomic_fadd_f64(p, v);
}
__device__ __forceinline__ void atomic_add(gptr<float> p, float v)
{
    __builtin_amdgcn_global_atomic_fadd_f32(p, v);
}
__device__ __forceinline__ void atomic_add(double *p, double v) { atomic_add(to_global(p), v); }
__device__ __forceinline__ void atomic_add(float *p, float v) { atomic_add(to_global(p), v); }

// Per-entry addressing of the two argument layouts.
template <typename T, Layout L>
struct Entries;

// kronmult6_batched.hpp:51-71: A (n,n,d,batch), X/Y (size, batch).
// The *_g accessors return global-address-space pointers; UNI asserts that b
// is the same on every lane of the workgroup (one entry per workgroup).
template <typename T>
struct Entries<T, Layout::Batched> {
    const T *A;
    const T *X;
    T *Y;
    int d;
    long long fstride;  // elements per factor (n*n)
    long long sx, sy;   // elements per X / Y vector
    __device__ __forceinline__ const T *factor(int b, int k) const
    {
        return A + ((long long)b * d + k) * fstride;
    }
    __device__ __forceinline__ const T *x(int b) const { return X + (long long)b * sx; }
    __device__ __forceinline__ T *y(int b) const { return Y + (long long)b * sy; }
    template <bool UNI>
    __device__ __forceinline__ auto factor_g(int b, int k) const
    {
        if constexpr (UNI)
            return to_const(factor(b, k));
        else
            return to_global(factor(b, k));
    }
    template <bool UNI>
    __device__ __forceinline__ gptr<const T> x_g(int b) const { return to_global(x(b)); }
    template <bool UNI>
    __device__ __forceinline__ gptr<T> y_g(int b) const { return to_global(y(b)); }
};

// kronmult_xbatched.hpp:79-83,123-135: Aarray[k + b*d], pX[b], pY[b].
template <typename T>
struct Entries<T, Layout::Ptr> {
    const T *const *pA;
    const T *const *pX;
    T *const *pY;
    int d;
    __device__ __forceinline__ const T *factor(int b, int k) const
    {
        return pA[(long long)b * d + k];
    }
    __device__ __forceinline__ const T *x(int b) const { return pX[b]; }
    __device__ __forceinline__ T *y(int b) const { return pY[b]; }
    template <bool UNI>
    __device__ __forceinline__ auto factor_g(int b, int k) const
    {
        if constexpr (UNI)
            return to_const(to_const(pA)[(long long)b * d + k]);
        else
            return to_global(factor(b, k));
    }
    template <bool UNI>
    __device__ __forceinline__ gptr<const T> x_g(int b) const
    {
        if constexpr (UNI)
            return to_global(to_const(pX)[b]);
        else
            return to_global(x(b));
    }
    template <bool UNI>
    __device__ __forceinline__ gptr<T> y_g(int b) const
    {
        if constexpr (UNI)
            return to_global(to_const(pY)[b]);
        else
            return to_global(y(b));
    }
};

// Y(idx) += v: device atomic when entries may share Y, else load-add-store.
template <bool SHARED, typename T, typename Acc, typename P>
__device__ __forceinline__ void y_accumulate(P y, long long idx, Acc v)
{
    if constexpr (SHARED) {
        atomic_add(y + idx, T(v));
    } else {
        y[idx] = T(Acc(y[idx]) + v);
    }
}

// A caller's scratch (W) as the kernels address it: the vector of entry b,
// which runs in launch-group slot `slot` (see the Scratch kinds).
template <typename T>
struct ScratchDesc {
    Scratch kind;
    T *W;                 // BatchedW base
    long long wstride;    // BatchedW elements per entry
    T *const *pW;         // PtrSlots / PtrSlot0 / PtrEntry
    T *Wbase;             // Linear
    long long slot_stride;
    __device__ __forceinline__ T *get(int slot, int b) const
    {
        switch (kind) {
        case Scratch::BatchedW: return W + (long long)b * wstride;
        case Scratch::PtrSlots: return pW[2 * slot];
        case Scratch::PtrSlot0: return pW[0];
        case Scratch::Linear: return Wbase + (long long)slot * slot_stride;
        case Scratch::PtrEntry: return pW[b];
        default: return nullptr;
        }
    }
};

// Quotient of small non-negative integers through the FP32 reciprocal:
// int((x + 1/2) * rcp(d)) == x / d for 0 <= x < 2^21.  (x + 1/2) / d lies at
// least 1/(2d) from every integer, and the rounding of rcp (1 ulp) and of the
// product moves it by less than (x + 1/2) * 2^-22 / d < 1/(2d).
struct FastDiv {
    float r;
    int d;
};
__device__ __forceinline__ FastDiv fast_div(int d) { return {__builtin_amdgcn_rcpf(float(d)), d}; }
__device__ __forceinline__ int quot(int x, FastDiv f) { return int((float(x) + 0.5f) * f.r); }
// Full-rate 24-bit multiply (LDS and in-vector offsets are far below 2^24).
__device__ __forceinline__ unsigned mul24(int a, int b) { return __umul24((unsigned)a, (unsigned)b); }

// Fiber x = o * inner + ii of a tensor viewed as outer x n x inner: o, and ii
// hidden from the optimiser (see fiber_base).
__device__ __forceinline__ int split_fiber(int x, FastDiv inner, int &ii)
{
    const int o = quot(x, inner);
    ii = x - (int)mul24(o, inner.d);
    asm volatile("" : "+v"(ii));
    return o;
}
// Offset o * n * inner + ii of the fiber's first element; element j is then
// j * inner further (a uniform add).  Hidden from the optimiser, which would
// otherwise fold the sum back into (o * n + j) * inner + ii: one quarter-rate
// 32- or 64-bit multiply per element.
__device__ __forceinline__ int fiber_base(int o, int ii, int n_inner)
{
    int base = (int)mul24(o, n_inner) + ii;
    asm volatile("" : "+v"(base));
    return base;
}

// dst = op(A) x_q src for one mode of an LDS tensor viewed as
// outer x n x inner (input) and outer x m x inner (output), column-major with
// the inner index fastest; A is m x n (n x m stored when TRANS) in LDS with
// leading dimension ldA.  One thread per fiber: n loads into registers
// (n <= NMAX), m outputs; consecutive threads take consecutive inner indices.
// Larger n falls back to one thread per output element.
template <typename T, typename Acc, bool TRANS, int NMAX>
__device__ __forceinline__ void contract_fibers(const T *__restrict__ src, T *__restrict__ dst,
                                                int outer, int n, int inner, int m,
                                                const T *__restrict__ Al, int ldA)
{
    const FastDiv di = fast_div(inner);
    if (n <= NMAX) {
        const int fibers = outer * inner;
        for (int f = threadIdx.x; f < fibers; f += blockDim.x) {
            int ii;
            const int o = split_fiber(f, di, ii);
            const T *s = src + fiber_base(o, ii, n * inner);
            T *dd = dst + fiber_base(o, ii, m * inner);
            Acc in[NMAX];
#pragma unroll
            for (int j = 0; j < NMAX; ++j) in[j] = (j < n) ? Acc(s[j * inner]) : Acc(0);
            for (int i = 0; i < m; ++i) {
                Acc acc = 0;
#pragma unroll
                for (int j = 0; j < NMAX; ++j)
                    if (j < n)
                        acc = kfma(TRANS ? Acc(Al[j + i * ldA]) : Acc(Al[i + j * ldA]),
                                            in[j], acc);
                dd[i * inner] = T(acc);
            }
        }
    } else {
        const int total = outer * m * inner;
        const FastDiv dm = fast_div(m);
        for (int w = threadIdx.x; w < total; w += blockDim.x) {
            int ii;
            const int rest = split_fiber(w, di, ii);
            const int o = quot(rest, dm);
            const int i = rest - (int)mul24(o, m);
            const T *s = src + fiber_base(o, ii, n * inner);
            Acc acc = 0;
            for (int j = 0; j < n; ++j)
                acc = kfma(TRANS ? Acc(Al[j + i * ldA]) : Acc(Al[i + j * ldA]),
                                    Acc(s[j * inner]), acc);
            dst[w] = T(acc);
        }
    }
}

// Square variant with compile-time N: the factor is read once from LDS into
// registers (N*N values) and each thread contracts whole fibers, N loads and
// N*N FMAs each.  SQ == 0 selects contract_fibers (runtime extents).
// dst = A x_q src for one mode of an LDS tensor viewed as outer x N x inner,
// with the N x N factor already in registers (a[i][j] = A(i, j)): each
// thread contracts whole fibers, N loads, N*N FMAs and N stores each.
template <typename T, typename Acc, int N>
__device__ __forceinline__ void contract_fibers_reg(const T *__restrict__ src,
                                                    T *__restrict__ dst, int outer, int inner,
                                                    const Acc (&a)[N][N])
{
    const int fibers = outer * inner;
    const FastDiv di = fast_div(inner);
    for (int f = threadIdx.x; f < fibers; f += blockDim.x) {
        int ii;
        const int o = split_fiber(f, di, ii);
        const int base = fiber_base(o, ii, N * inner);
        const T *s = src + base;
        T *dd = dst + base;
        Acc in[N];
#pragma unroll
        for (int j = 0; j < N; ++j) in[j] = Acc(s[j * inner]);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            Acc acc = a[i][0] * in[0];
#pragma unroll
            for (int j = 1; j < N; ++j) acc = kfma(a[i][j], in[j], acc);
            dd[i * inner] = T(acc);
        }
    }
}

// Square N = 9..15: the factor (N*N values) no longer fits one thread's
// registers, so it is cut into C = ceil(N / R) blocks of R rows.  Thread t
// keeps rows [c R, c R + R) (c = t mod C) in registers for the whole mode and
// walks the fibers t / C, t / C + S, ... (S = floor(256 / C) fibers per
// sweep; threads from S C on sit the mode out): each input element read from
// LDS feeds R FMAs, where one work item per output element costs two LDS
// reads (factor and input) per FMA.  Rows past N (the last block when R does
// not divide N) are computed from a clamped duplicate row and never stored.
template <typename T, typename Acc, bool TRANS, int N, int R>
__device__ __forceinline__ void contract_rowblocks(const T *__restrict__ src,
                                                   T *__restrict__ dst, int outer, int inner,
                                                   const T *__restrict__ Al)
{
    constexpr int C = (N + R - 1) / R;  // row blocks: blockDim / C fibers per sweep
    const int step = (int)blockDim.x / C;
    if ((int)threadIdx.x >= step * C) return;
    const int c = (int)threadIdx.x % C;
    const int i0 = c * R;
    Acc a[R][N];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = min(i0 + r, N - 1);
#pragma unroll
        for (int j = 0; j < N; ++j) a[r][j] = TRANS ? Acc(Al[j + i * N]) : Acc(Al[i + j * N]);
    }
    const int fibers = outer * inner;
    const FastDiv di = fast_div(inner);
    for (int f = (int)threadIdx.x / C; f < fibers; f += step) {
        int ii;
        const int o = split_fiber(f, di, ii);
        const int base = fiber_base(o, ii, N * inner);
        const T *s = src + base;
        T *dd = dst + base;
        Acc in[N];
#pragma unroll
        for (int j = 0; j < N; ++j) in[j] = Acc(s[j * inner]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Acc acc = a[r][0] * in[0];
#pragma unroll
            for (int j = 1; j < N; ++j) acc = kfma(a[r][j], in[j], acc);
            if (i0 + r < N) dd[(i0 + r) * inner] = T(acc);
        }
    }
}

// Square variant with compile-time N: the factor is read once from LDS into
// registers (N*N values), then contract_fibers_reg; N = 9..15 goes through
// contract_rowblocks.  SQ == 0 selects contract_fibers (runtime extents).
template <typename T, typename Acc, bool TRANS, int SQ>
__device__ __forceinline__ void contract_mode_lds(const T *__restrict__ src, T *__restrict__ dst,
                                                  int outer, int n, int inner, int m,
                                                  const T *__restrict__ Al, int ldA)
{
    if constexpr (SQ == 0) {
        contract_fibers<T, Acc, TRANS, 8>(src, dst, outer, n, inner, m, Al, ldA);
    } else if constexpr (SQ > 8) {
        // three-row blocks measured faster at N = 11, 12 (d = 3..4, FP64), two
        // rows elsewhere (the registers of a third row cost occupancy)
        constexpr int R = (SQ == 11 || SQ == 12) ? 3 : 2;
        contract_rowblocks<T, Acc, TRANS, SQ, R>(src, dst, outer, inner, Al);
    } else {
        constexpr int N = SQ;
        if ((int)threadIdx.x >= outer * inner) return;
        Acc a[N][N];
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j)
                a[i][j] = TRANS ? Acc(Al[j + i * N]) : Acc(Al[i + j * N]);
        contract_fibers_reg<T, Acc, N>(src, dst, outer, inner, a);
    }
}

// Copy an m x n column-major factor (leading dimension ld) into LDS, packed
// with leading dimension m.
template <typename T, typename PtrT>
__device__ __forceinline__ void stage_factor(T *dst, PtrT src, int m, int n, int ld)
{
    for (int w = threadIdx.x; w < m * n; w += blockDim.x) dst[w] = src[(w % m) + (w / m) * ld];
}

// Compile-time loop K = HI, HI-1, ..., LO calling f(integral_constant<K>).
template <int HI, int LO, typename F>
__device__ __forceinline__ void for_modes_desc(F &&f)
{
    if constexpr (HI >= LO) {
        f(std::integral_constant<int, HI>{});
        for_modes_desc<HI - 1, LO>(f);
    }
}

}  // namespace kron
