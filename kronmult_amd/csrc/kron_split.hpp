// kron_split.hpp -- device helpers shared by the compile-time two-pass
// kernels of entries larger than LDS (kron_split.hip: forward;
// kron_adjoint_rows.hip: the adjoint's row passes).
#pragma once

#include "kron_fast.hpp"

namespace kron {

// Raw buffer accesses (MUBUF): the wave-uniform base in a buffer resource,
// the lane's byte offset in one VGPR and a per-row byte offset in an SGPR, so
// that a column's R loads need no per-row address arithmetic.  Resource word
// 3 is the gfx9 raw-buffer format (32-bit elements, no swizzle); num_records
// bounds every access to the vector.
typedef unsigned int kron_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vec_rsrc(const void *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double buf_load(double *, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ float buf_load(float *, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(kron_u32x2, v), r, voff, soff, 0);
}
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

constexpr int kSplitTB1 = 448;  // pass-1 threads per workgroup

template <int D, int N>
struct SplitGeom {
    static constexpr int H = 2;                  // slow modes of pass 1
    static constexpr int R = cpow(N, H);         // rows
    static constexpr int C = cpow(N, D - H);     // columns (contiguous)
    static constexpr int TB1 = kSplitTB1;
    static constexpr int NB1 = (C + TB1 - 1) / TB1;  // pass-1 workgroups per entry
    static constexpr int PER2 = (C + 255) / 256;     // row elements per pass-2 thread
};


// In-place contraction of the fibers of stride S (within a block of S*N
// elements) of a register array: v[base + i*S] = sum_j a[i][j] v[base + j*S].
template <typename Acc, int N, int S, int LEN>
__device__ __forceinline__ void split_contract_regs(Acc (&v)[LEN], const Acc (&a)[N][N])
{
#pragma unroll
    for (int f = 0; f < LEN / N; ++f) {
        const int base = (f / S) * S * N + f % S;
        Acc out[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            Acc s = a[i][0] * v[base];
#pragma unroll
            for (int j = 1; j < N; ++j) s = kfma(a[i][j], v[base + j * S], s);
            out[i] = s;
        }
#pragma unroll
        for (int i = 0; i < N; ++i) v[base + i * S] = out[i];
    }
}

// One fast mode of a row in LDS: fibers of stride S, thread t owns fibers
// t, t + 256, ...; in place.
template <typename T, typename Acc, int N, int S, int C>
__device__ __forceinline__ void split_contract_lds(T *row, const Acc (&a)[N][N])
{
    constexpr int F = C / N;
    for (int f = threadIdx.x; f < F; f += 256) {
        const int base = (f / S) * (S * N) + f % S;
        Acc in[N];
#pragma unroll
        for (int j = 0; j < N; ++j) in[j] = Acc(row[base + j * S]);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            Acc s = a[i][0] * in[0];
#pragma unroll
            for (int j = 1; j < N; ++j) s = kfma(a[i][j], in[j], s);
            row[base + i * S] = T(s);
        }
    }
}

}  // namespace kron
