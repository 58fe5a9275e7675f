// kron_fast.hpp -- fused on-chip kernel for square factors, compile-time (D, N).
//
// Replaces the reference's tail recursion (kronmultv.hpp:19-148), which
// issues one small GEMM per vector per level with every intermediate going
// through global memory, by one pass per mode over an LDS-resident copy of
// the entry:
//
//   mode 2      : X (HBM, coalesced)      -> registers -> LDS
//   modes D..3  : LDS -> registers -> LDS (in place, one fiber per thread)
//   mode 1      : LDS -> registers -> Y   (coalesced; atomic when Y is shared)
//
// A "fiber" of mode k is the N elements that differ only in index k
// (stride N^(D-k)).  Each thread owns whole fibers, so a mode pass is N
// loads, N*N FMAs with the factor held in registers, and N stores, with no
// barrier inside the pass.  X is read once and Y updated once per entry; all
// other traffic stays in LDS.  The contraction order differs from the
// reference's (1, 2, ..., D), which changes only the rounding.
//
// Small entries share a 256-thread workgroup: G = pow2ceil(N^(D-1)) threads
// (at most 256) per entry, 256/G entries per workgroup.
#pragma once

#include "kron_device.hpp"

namespace kron {

template <int D, int N>
struct FastGeom {
    static constexpr int E = cpow(N, D);        // elements per entry
    static constexpr int F = E / N;              // fibers per mode
    static constexpr int G = F >= 256 ? 256 : pow2ceil(F);  // threads per entry
    static constexpr int P = 256 / G;            // entries per workgroup
    static constexpr size_t lds_bytes(size_t elem) { return (size_t)P * E * elem; }
};

// a[i][j] = A(i,j) (or A(j,i) when TRANS), A column-major with leading dim lda.
template <typename T, typename Acc, int N, bool TRANS>
__device__ __forceinline__ void load_factor(Acc (&a)[N][N], const T *__restrict__ Ap, int lda)
{
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i)
            a[i][j] = TRANS ? Acc(Ap[j + i * lda]) : Acc(Ap[i + j * lda]);
}

// One pass over the fibers of mode K owned by thread g of its entry.
template <typename Acc, int D, int N, int K, int G, typename LoadF, typename StoreF>
__device__ __forceinline__ void fiber_pass(int g, const Acc (&a)[N][N], LoadF &&ld, StoreF &&st)
{
    constexpr int F = cpow(N, D - 1);
    constexpr int S = cpow(N, D - K);
    constexpr int IT = (F + G - 1) / G;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int f = g + it * G;
        if ((F % G) == 0 || f < F) {
            const int base = (f % S) + (f / S) * (S * N);
            Acc in[N];
#pragma unroll
            for (int j = 0; j < N; ++j) in[j] = Acc(ld(base + j * S));
#pragma unroll
            for (int i = 0; i < N; ++i) {
                Acc s = a[i][0] * in[0];
#pragma unroll
                for (int j = 1; j < N; ++j) s = __builtin_fma(a[i][j], in[j], s);
                st(base + i * S, s);
            }
        }
    }
}

template <typename T, typename Acc, int D, int N, Layout LAY, bool SHARED, bool TRANS>
__global__ void __launch_bounds__(256)
kron_fast_kernel(Entries<T, LAY> ent, int lda, int batch)
{
    using Geo = FastGeom<D, N>;
    constexpr int E = Geo::E;
    constexpr int G = Geo::G;
    constexpr int P = Geo::P;
    extern __shared__ __align__(16) unsigned char kron_smem[];
    T *smem = reinterpret_cast<T *>(kron_smem);

    const int tid = threadIdx.x;
    const int slot = (P == 1) ? 0 : tid / G;
    const int g = (P == 1) ? tid : tid % G;
    const int b = blockIdx.x * P + slot;
    const bool active = b < batch;
    Acc a[N][N];

    if constexpr (D == 1) {
        if (active) {
            load_factor<T, Acc, N, TRANS>(a, ent.factor(b, 0), lda);
            const T *__restrict__ x = ent.x(b);
            T *y = ent.y(b);
            fiber_pass<Acc, D, N, 1, G>(
                g, a, [&](int i) { return x[i]; },
                [&](int i, Acc v) { y_accumulate<SHARED, T, Acc>(y, i, v); });
        }
    } else {
        T *buf = smem + slot * E;
        if (active) {
            load_factor<T, Acc, N, TRANS>(a, ent.factor(b, 1), lda);
            const T *__restrict__ x = ent.x(b);
            fiber_pass<Acc, D, N, 2, G>(
                g, a, [&](int i) { return x[i]; }, [&](int i, Acc v) { buf[i] = T(v); });
        }
        __syncthreads();
        for_modes_desc<D, 3>([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            if (active) {
                load_factor<T, Acc, N, TRANS>(a, ent.factor(b, K - 1), lda);
                fiber_pass<Acc, D, N, K, G>(
                    g, a, [&](int i) { return buf[i]; }, [&](int i, Acc v) { buf[i] = T(v); });
            }
            __syncthreads();
        });
        if (active) {
            load_factor<T, Acc, N, TRANS>(a, ent.factor(b, 0), lda);
            T *y = ent.y(b);
            fiber_pass<Acc, D, N, 1, G>(
                g, a, [&](int i) { return buf[i]; },
                [&](int i, Acc v) { y_accumulate<SHARED, T, Acc>(y, i, v); });
        }
    }
}

// Largest N with a compiled fast kernel for dimension D and element size.
// The bound keeps N^D elements of T within 160 KiB of LDS.
template <typename T>
constexpr int fast_max_n(int D)
{
    return (sizeof(T) == 8) ? (D == 6 ? 5 : D == 5 ? 7 : 8)
                            : (D == 6 ? 5 : 8);
}

template <typename T, typename Acc, int D, int N, Layout LAY, bool SHARED, bool TRANS>
kron_status launch_fast_one(const Entries<T, LAY> &ent, int lda, int batch, hipStream_t stream)
{
    using Geo = FastGeom<D, N>;
    const size_t lds = (D == 1) ? 0 : Geo::lds_bytes(sizeof(T));
    if ((long long)lds > (long long)max_lds_bytes()) return KRON_ENOTSUP;
    auto kern = kron_fast_kernel<T, Acc, D, N, LAY, SHARED, TRANS>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            set_last_hip_error((int)e);
            return KRON_EHIP;
        }
    }
    const long long nblocks = ((long long)batch + Geo::P - 1) / Geo::P;
    if (nblocks > 0x7fffffffLL) return KRON_EINVAL;
    hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(256), lds, stream, ent, lda, batch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    return KRON_OK;
}

// Runtime (D, N) -> compiled kernel.  KRON_ENOTSUP when N > fast_max_n.
template <typename T, typename Acc, Layout LAY, bool SHARED, bool TRANS>
kron_status launch_fast(int D, int N, const Entries<T, LAY> &ent, int lda, int batch,
                        hipStream_t stream)
{
    if (N < 1 || N > fast_max_n<T>(D)) return KRON_ENOTSUP;
#define KRON_FAST_CASE(DD, NN)                                                           \
    if (D == DD && N == NN) {                                                            \
        if constexpr (NN <= fast_max_n<T>(DD))                                           \
            return launch_fast_one<T, Acc, DD, NN, LAY, SHARED, TRANS>(ent, lda, batch,  \
                                                                       stream);          \
        else                                                                             \
            return KRON_ENOTSUP;                                                         \
    }
#define KRON_FAST_D(DD)                                                                  \
    KRON_FAST_CASE(DD, 1) KRON_FAST_CASE(DD, 2) KRON_FAST_CASE(DD, 3)                    \
    KRON_FAST_CASE(DD, 4) KRON_FAST_CASE(DD, 5) KRON_FAST_CASE(DD, 6)                    \
    KRON_FAST_CASE(DD, 7) KRON_FAST_CASE(DD, 8)
    KRON_FAST_D(1)
    KRON_FAST_D(2)
    KRON_FAST_D(3)
    KRON_FAST_D(4)
    KRON_FAST_D(5)
    KRON_FAST_D(6)
#undef KRON_FAST_D
#undef KRON_FAST_CASE
    return KRON_ENOTSUP;
}

}  // namespace kron
