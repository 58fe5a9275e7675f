// kron_plane.hpp -- the d = 6, n = 4 entry (4,096 elements) with two modes
// per LDS pass.
//
// The fiber scheme of kron_fast.hpp makes one LDS pass per mode (six passes,
// ten LDS accesses per element).  Here each of the 256 threads owns a 4 x 4
// "plane" of the entry (two modes varying, the other four fixed) and
// contracts both modes in registers, so an entry takes three passes and four
// LDS accesses per element:
//
//   pass A: X plane (modes 2, 4), prefetched from HBM one entry ahead -> LDS
//   pass B: LDS plane (modes 6, 5) -> LDS, in place
//   pass C: LDS plane (modes 3, 1) -> Y (or the running sum of a Y run)
//
// LDS layout: element e at e + e / 16 (one element of padding after every
// 16).  The lane digits are m6, m5 (or m4, m3 in pass B): with the padding
// the 16 lanes a 64-bit LDS access serves together hit 16 distinct bank
// slots, the HBM accesses of passes A and C are runs of 16 (X) and 64 (Y)
// consecutive elements, and the 16 elements of a thread's plane sit at
// compile-time offsets from one base address (pass A: 272 i + 17 j, pass B:
// i + 4 j, pass C: 68 i + 1088 j).
// Element index e = 1024 m1 + 256 m2 + 64 m3 + 16 m4 + 4 m5 + m6 (A1 acts on
// m1, the slowest index, as in the reference's layout).
#pragma once

#include "kron_fast.hpp"  // load_factor, Entries

namespace kron {

constexpr int kPlaneElems = 4096;
constexpr int kPlaneLds = kPlaneElems + kPlaneElems / 16;

__device__ __forceinline__ int plane_pad(int e) { return e + (e >> 4); }

// v[i][j] <- sum_k a[i][k] v[k][j]: the mode along the first plane index.
template <typename Acc>
__device__ __forceinline__ void plane_first(Acc (&v)[4][4], const Acc (&a)[4][4])
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        Acc t[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            Acc s = a[i][0] * v[0][j];
#pragma unroll
            for (int k = 1; k < 4; ++k) s = __builtin_fma(a[i][k], v[k][j], s);
            t[i] = s;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i][j] = t[i];
    }
}

// v[i][j] <- sum_k a[j][k] v[i][k]: the mode along the second plane index.
template <typename Acc>
__device__ __forceinline__ void plane_second(Acc (&v)[4][4], const Acc (&a)[4][4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        Acc t[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            Acc s = a[j][0] * v[i][0];
#pragma unroll
            for (int k = 1; k < 4; ++k) s = __builtin_fma(a[j][k], v[i][k], s);
            t[j] = s;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) v[i][j] = t[j];
    }
}

// Persistent workgroups, one entry at a time (as kron_fast_kernel with
// P == 1): GROUPED walks the Y-target order and adds each run's summed
// result to Y once; otherwise Y is updated per entry (atomically if SHARED).
template <typename T, typename Acc, Layout LAY, bool SHARED, bool TRANS, bool GROUPED>
__global__ void __launch_bounds__(256)
kron_plane_kernel(Entries<T, LAY> ent, int lda, int batch, GroupOrder ord)
{
    extern __shared__ __align__(16) unsigned char kron_smem[];
    T *const lds = reinterpret_cast<T *>(kron_smem);
    const int tid = threadIdx.x;
    const int l = tid & 63, w = tid >> 6;
    const int d0 = l & 3, d1 = (l >> 2) & 3, d2 = (l >> 4) & 3;
    // plane origins (lane digit d0 varies fastest): A (m1 = w, m3 = d2,
    // m5 = d1, m6 = d0), B (m1 = w, m2 = d2, m3 = d1, m4 = d0),
    // C (m2 = w, m4 = d2, m5 = d1, m6 = d0)
    const int eA = w * 1024 + d2 * 64 + d1 * 4 + d0;
    const int eB = w * 1024 + d2 * 256 + d1 * 64 + d0 * 16;
    const int eC = w * 256 + d2 * 16 + d1 * 4 + d0;
    T *const LA = lds + plane_pad(eA);
    T *const LB = lds + plane_pad(eB);
    T *const LC = lds + plane_pad(eC);

    int rbeg, rend, rstep;
    if constexpr (GROUPED) {
        rbeg = blockIdx.x * ord.chunk;
        rend = min(batch, rbeg + ord.chunk);
        rstep = 1;
    } else {
        rbeg = blockIdx.x;
        rend = batch;
        rstep = gridDim.x;
    }
    auto entry = [&](int r) -> int {
        if constexpr (GROUPED)
            return to_const(ord.perm)[r];
        else
            return r;
    };
    Acc a[4][4];
    T nx[4][4];  // the next entry's X plane (modes 2, 4)
    auto load_x = [&](int b) {
        const auto x = ent.template x_g<true>(b) + eA;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) nx[i][j] = x[i * 256 + j * 16];
    };
    Acc acc[GROUPED ? 4 : 1][4];  // running Y sum of the current run
    if constexpr (GROUPED) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0;
    }
    if (rbeg < rend) load_x(entry(rbeg));

    for (int r = rbeg; r < rend; r += rstep) {
        const int b = entry(r);
        Acc v[4][4];
        // pass A: modes 2 and 4 of X, into LDS
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[i][j] = Acc(nx[i][j]);
        if (r + rstep < rend) load_x(entry(r + rstep));
        load_factor<T, Acc, 4, TRANS, true>(a, ent.template factor_g<true>(b, 1), lda);
        plane_first(v, a);
        load_factor<T, Acc, 4, TRANS, true>(a, ent.template factor_g<true>(b, 3), lda);
        plane_second(v, a);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) LA[272 * i + 17 * j] = T(v[i][j]);
        __syncthreads();
        // pass B: modes 6 and 5, in place
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[i][j] = Acc(LB[i + 4 * j]);
        load_factor<T, Acc, 4, TRANS, true>(a, ent.template factor_g<true>(b, 5), lda);
        plane_first(v, a);
        load_factor<T, Acc, 4, TRANS, true>(a, ent.template factor_g<true>(b, 4), lda);
        plane_second(v, a);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) LB[i + 4 * j] = T(v[i][j]);
        __syncthreads();
        // pass C: modes 3 and 1, into Y
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[i][j] = Acc(LC[68 * i + 1088 * j]);
        load_factor<T, Acc, 4, TRANS, true>(a, ent.template factor_g<true>(b, 2), lda);
        plane_first(v, a);
        load_factor<T, Acc, 4, TRANS, true>(a, ent.template factor_g<true>(b, 0), lda);
        plane_second(v, a);
        if constexpr (GROUPED) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += v[i][j];
            const auto keys = to_const(ord.keys);
            const unsigned long long key = keys[r];
            if (r + 1 >= rend || keys[r + 1] != key) {
                const auto y = to_global(reinterpret_cast<T *>(key)) + eC;
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        atomic_add(y + (64 * i + 1024 * j), T(acc[i][j]));
                        acc[i][j] = 0;
                    }
            }
        } else {
            const auto y = ent.template y_g<true>(b) + eC;
            if constexpr (SHARED) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) atomic_add(y + (64 * i + 1024 * j), T(v[i][j]));
            } else {
                T yv[4][4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) yv[i][j] = y[64 * i + 1024 * j];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        y[64 * i + 1024 * j] = T(Acc(yv[i][j]) + v[i][j]);
            }
        }
        __syncthreads();  // LDS is rewritten by the next entry
    }
}

}  // namespace kron
