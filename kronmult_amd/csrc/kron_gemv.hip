// kron_gemv.hip -- d = 1 entries: Y(:,k) += A(k) X(:,k), one m x n factor.
//
// At d = 1 the Kronecker product is the factor itself, so an entry is a
// small matrix-vector product (kronmult1.hpp:17-64 reaches it through one
// kgemm_nn call per entry).  The fused kernels cover n <= 8; larger or
// rectangular factors would otherwise take the generic kernel, which gives
// every entry a whole workgroup and its own LDS staging (0.2-1.4 TB/s at
// n = 9..32).  Here a lane computes one output row: a segment of
// S = pow2ceil(m) <= 64 lanes takes one entry (64 / S entries per wave,
// rows beyond 64 in steps of 64), reads the factor column by column (the
// segment's rows of one column are contiguous: coalesced) and X as a
// broadcast within the segment, and adds its row into Y (an atomic when
// entries may share Y).  The factor is read once, so the kernel streams.
#include <algorithm>

#include "kron_device.hpp"
#include "kron_internal.hpp"

namespace kron {

template <typename T, Layout LAY, bool SHARED>
__global__ void __launch_bounds__(256)
kron_gemv_kernel(Entries<T, LAY> ent, int m, int n, int ld, int seg, long long batch)
{
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long long nwaves = (long long)gridDim.x * 4;
    const int per = 64 / seg;  // entries per wave and step
    const int e_in = lane / seg, i0 = lane % seg;
    for (long long b0 = wave * per; b0 < batch; b0 += nwaves * per) {
        const long long b = b0 + e_in;
        if (b >= batch) continue;
        const auto A = to_global(ent.factor((int)b, 0));
        const auto x = to_global(ent.x((int)b));
        const auto y = to_global(ent.y((int)b));
        for (int i = i0; i < m; i += seg) {
            T acc = 0;
#pragma unroll 8
            for (int j = 0; j < n; ++j) acc = kfma(A[i + (long long)j * ld], x[j], acc);
            y_accumulate<SHARED, T, T>(y, i, acc);
        }
    }
}

template <typename T, Layout LAY, bool SHARED>
static kron_status gemv_launch(const Entries<T, LAY> &e, int m, int n, int ld, int batch,
                               hipStream_t s)
{
    const int seg = std::min(64, (int)pow2ceil(m));
    const long long per_block = 4LL * (64 / seg);
    const void *kern = reinterpret_cast<const void *>(kron_gemv_kernel<T, LAY, SHARED>);
    const long long resident = persistent_blocks(kern, 0);
    const long long need = (batch + per_block - 1) / per_block;
    const long long grid = std::max(1LL, std::min(need, resident > 0 ? resident : 2048LL));
    hipLaunchKernelGGL((kron_gemv_kernel<T, LAY, SHARED>), dim3((unsigned)grid), dim3(256), 0, s,
                       e, m, n, ld, seg, (long long)batch);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error((int)err);
        return KRON_EHIP;
    }
    return KRON_OK;
}

template <typename T>
kron_status launch_gemv(const KronArgs<T> &a)
{
    if (a.d != 1 || a.trans) return KRON_ENOTSUP;
    const int m = a.m[0], n = a.n[0];
    const int ld = a.lda > 0 ? a.lda : m;
    if (m < 1 || n < 1 || ld < m) return KRON_EINVAL;
    if (a.batch == 0) return KRON_OK;
    if (a.layout == Layout::Batched) {
        Entries<T, Layout::Batched> e{a.A, a.X, a.Y, 1, (long long)ld * n, n, m};
        return gemv_launch<T, Layout::Batched, false>(e, m, n, ld, a.batch, a.stream);
    }
    Entries<T, Layout::Ptr> e{a.pA, a.pX, a.pY, 1};
    if (a.y_shared) return gemv_launch<T, Layout::Ptr, true>(e, m, n, ld, a.batch, a.stream);
    return gemv_launch<T, Layout::Ptr, false>(e, m, n, ld, a.batch, a.stream);
}

template kron_status launch_gemv<double>(const KronArgs<double> &);
template kron_status launch_gemv<float>(const KronArgs<float> &);

}  // namespace kron
