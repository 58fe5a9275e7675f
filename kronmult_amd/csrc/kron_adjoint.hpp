// kron_adjoint.hpp -- per-entry addressing of the adjoint's operands in the
// two argument layouts (kron_adjoint.hip, kron_adjoint_rows.hip).
//
// Batched (the contiguous layout of kronmult6_batched.hpp:51-71): results
// overwrite private dX vectors; the dA sums go into dA cleared beforehand.
// Ptr (the pointer arrays of kronmult_xbatched.hpp:79-83): factor k of entry
// b at pA[b*d + k] with leading dimension lda, X and dY through pX[b] and
// pdY[b]; K_b^T dY is added into pdX[b] and the gradient of factor k into
// pdA[b*d + k] (leading dimension ldda) with device atomics, so entries may
// share X vectors, dY vectors (shared Y targets in the forward) and factors.
//
// Every accessor taking the entry b assumes b is the same on all lanes of
// the workgroup: pointers are read from the arrays with scalar loads.
#pragma once

#include "kron_device.hpp"

namespace kron {

template <typename T, Layout LAY>
struct AdjIO;

template <typename T>
struct AdjIO<T, Layout::Batched> {
    const T *A, *X, *dY;
    T *dX, *dA;
    int d, n;
    long long E;
    static constexpr bool kAccumulate = false;
    __device__ __forceinline__ const T *fac(int b, int k) const
    {
        return A + ((long long)b * d + k) * n * n;
    }
    __device__ __forceinline__ int lda() const { return n; }
    // element w = k n^2 + i + j n of the entry's d stacked factors (column-major)
    __device__ __forceinline__ T fac_elem(int b, int w) const
    {
        return A[(long long)b * d * n * n + w];
    }
    __device__ __forceinline__ gptr<const T> x(int b) const { return to_global(X + b * E); }
    __device__ __forceinline__ gptr<const T> dy(int b) const { return to_global(dY + b * E); }
    __device__ __forceinline__ gptr<T> dx(int b) const { return to_global(dX + b * E); }
    __device__ __forceinline__ static void put_dx(gptr<T> p, int i, T v) { p[i] = v; }
    // dA_k(i, j) += v for w = k n^2 + i + j n
    __device__ __forceinline__ void add_da(int b, int w, T v) const
    {
        atomic_add(to_global(dA + (long long)b * d * n * n + w), v);
    }
};

template <typename T>
struct AdjIO<T, Layout::Ptr> {
    const T *const *pA, *const *pX, *const *pdY;
    T *const *pdX, *const *pdA;
    int d, n, lda_, ldda;
    static constexpr bool kAccumulate = true;
    __device__ __forceinline__ const T *fac(int b, int k) const
    {
        return to_const(pA)[(long long)b * d + k];
    }
    __device__ __forceinline__ int lda() const { return lda_; }
    __device__ __forceinline__ T fac_elem(int b, int w) const
    {
        const int nn = n * n, k = w / nn, e = w - k * nn, j = e / n;
        return to_global(pA[(long long)b * d + k])[e - j * n + j * lda_];
    }
    __device__ __forceinline__ gptr<const T> x(int b) const { return to_global(to_const(pX)[b]); }
    __device__ __forceinline__ gptr<const T> dy(int b) const
    {
        return to_global(to_const(pdY)[b]);
    }
    __device__ __forceinline__ gptr<T> dx(int b) const { return to_global(to_const(pdX)[b]); }
    __device__ __forceinline__ static void put_dx(gptr<T> p, int i, T v) { atomic_add(p + i, v); }
    __device__ __forceinline__ void add_da(int b, int w, T v) const
    {
        const int nn = n * n, k = w / nn, e = w - k * nn, j = e / n;
        atomic_add(to_global(pdA[(long long)b * d + k]) + (e - j * n + j * ldda), v);
    }
};

template <typename T>
inline AdjIO<T, Layout::Batched> batched_io(const AdjointArgs<T> &a)
{
    long long E = 1;
    for (int k = 0; k < a.d; ++k) E *= a.n;
    return AdjIO<T, Layout::Batched>{a.A, a.X, a.dY, a.dX, a.dA, a.d, a.n, E};
}

template <typename T>
inline AdjIO<T, Layout::Ptr> ptr_io(const AdjointArgs<T> &a)
{
    return AdjIO<T, Layout::Ptr>{a.pA, a.pX, a.pdY, a.pdX, a.pdA, a.d, a.n, a.lda, a.ldda};
}

}  // namespace kron
