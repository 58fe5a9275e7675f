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

// Hardware FP atomic add (global_atomic_add_f64 / _f32 on gfx950).  Y must be
// ordinary device memory (hipMalloc / torch), where these are exact adds.
__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }

// Per-entry addressing of the two argument layouts.
template <typename T, Layout L>
struct Entries;

// kronmult6_batched.hpp:51-71: A (n,n,d,batch), X/Y (size, batch).
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
};

// Y(idx) += v: device atomic when entries may share Y, else load-add-store.
template <bool SHARED, typename T, typename Acc>
__device__ __forceinline__ void y_accumulate(T *y, long long idx, Acc v)
{
    if constexpr (SHARED) {
        atomic_add(y + idx, T(v));
    } else {
        y[idx] = T(Acc(y[idx]) + v);
    }
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
