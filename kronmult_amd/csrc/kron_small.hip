// kron_small.hip -- small entries with runtime extents (d >= 3), off the
// fused kernels: mixed per-mode sizes in vbatched (kronmult_vbatched.hpp:
// 63-68), and any other pointer-layout entry whose intermediates are small.
//
// The generic kernel gives each such entry a whole workgroup, and most of its
// 256 threads idle (vbatched m = n = (4, 5, 6): 120 outputs per mode,
// 0.5 TB/s).  Here a wave takes an entry, as in the d = 2 kernel: X, the two
// ping-pong intermediates and all d factors sit in the wave's LDS slice, and
// the four waves of a workgroup step through entries together (workgroup
// barriers between the modes).  Mode q (slowest first) maps
//     in (outer x n_q x inner) -> out (outer x m_q x inner),
//     out(o, i, ii) = sum_j A_q(i, j) in(o, j, ii),
// outer = m_0..m_{q-1}, inner = n_{q+1}..n_{d-1}; a lane computes the output
// elements lane, lane + 64, ... (consecutive lanes on consecutive ii).  The
// last mode's outputs go straight into Y (an atomic add when entries may
// share Y).
#include <algorithm>

#include "kron_device.hpp"
#include "kron_internal.hpp"

namespace kron {

struct SmallShape {
    int d;
    int m[kMaxDim], n[kMaxDim];  // factor q is m[q] x n[q], slowest mode first
    int ld[kMaxDim];             // leading dimensions as stored
    int fac_off[kMaxDim];        // offset of factor q in the slice (packed, ld = m[q])
    int nx;                      // elements of X
    int buf;                     // elements of each ping-pong buffer
    int slice;                   // elements per wave: X, 2 buffers, factors
};

template <typename T, bool SHARED>
__global__ void __launch_bounds__(256)
kron_small_kernel(Entries<T, Layout::Ptr> ent, SmallShape sh, long long batch)
{
    extern __shared__ __align__(16) unsigned char kron_smem[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T *const xs = reinterpret_cast<T *>(kron_smem) + wid * sh.slice;
    T *const b0 = xs + sh.nx;
    T *const b1 = b0 + sh.buf;
    T *const fs = b1 + sh.buf;
    const int d = sh.d;
    for (long long g = (long long)blockIdx.x * 4; g < batch; g += (long long)gridDim.x * 4) {
        const long long b = g + wid;  // this wave's entry; idle past the end
        const bool live = b < batch;
        if (live) {
            const auto x = to_global(ent.x((int)b));
            for (int w = lane; w < sh.nx; w += 64) xs[w] = x[w];
            for (int q = 0; q < d; ++q) {
                const auto A = to_global(ent.factor((int)b, q));
                const int mq = sh.m[q], cnt = sh.m[q] * sh.n[q];
                for (int w = lane; w < cnt; w += 64)
                    fs[sh.fac_off[q] + w] = A[(w % mq) + (w / mq) * sh.ld[q]];
            }
        }
        __syncthreads();
        int inner = 1;
        for (int q = 1; q < d; ++q) inner *= sh.n[q];
        int outer = 1;
        const T *src = xs;
        for (int q = 0; q < d; ++q) {
            const int mq = sh.m[q], nq = sh.n[q];
            const int total = outer * mq * inner;
            T *dst = (q & 1) ? b1 : b0;
            if (live) {
                const T *Aq = fs + sh.fac_off[q];
                const FastDiv di = fast_div(inner), dm = fast_div(mq);
                const auto y = to_global(ent.y((int)b));
                for (int w = lane; w < total; w += 64) {
                    int ii;
                    const int rest = split_fiber(w, di, ii);
                    const int o = quot(rest, dm);
                    const int i = rest - (int)mul24(o, mq);
                    const T *s = src + fiber_base(o, ii, nq * inner);
                    T acc = 0;
                    for (int j = 0; j < nq; ++j)
                        acc = kfma(Aq[i + j * mq], s[j * inner], acc);
                    if (q + 1 < d)
                        dst[w] = acc;
                    else
                        y_accumulate<SHARED, T, T>(y, w, acc);
                }
            }
            __syncthreads();
            src = dst;
            outer *= mq;
            if (q + 1 < d) inner /= sh.n[q + 1];
        }
    }
}

// Per-wave LDS budget: four slices within 64 KB (two workgroups per CU).
constexpr size_t kSmallSliceBytes = 16384;

template <typename T>
kron_status launch_small(const KronArgs<T> &a)
{
    if (a.layout != Layout::Ptr || a.trans || a.d < 3) return KRON_ENOTSUP;
    SmallShape sh{};
    sh.d = a.d;
    long long nx = 1, buf = 0, fac = 0;
    for (int q = 0; q < a.d; ++q) {
        sh.m[q] = a.m[q];
        sh.n[q] = a.n[q];
        sh.ld[q] = a.lda > 0 ? a.lda : a.m[q];
        if (sh.m[q] < 1 || sh.n[q] < 1 || sh.ld[q] < sh.m[q]) return KRON_EINVAL;
        nx *= a.n[q];
    }
    // intermediate after mode q: m_0..m_q x n_{q+1}..n_{d-1}
    for (int q = 0; q < a.d; ++q) {
        long long s = 1;
        for (int l = 0; l < a.d; ++l) s *= (l <= q) ? a.m[l] : a.n[l];
        buf = std::max(buf, s);
    }
    for (int q = 0; q < a.d; ++q) {
        sh.fac_off[q] = (int)fac;
        fac += (long long)a.m[q] * a.n[q];
    }
    const long long slice = nx + 2 * buf + fac;
    if ((size_t)slice * sizeof(T) > kSmallSliceBytes) return KRON_ENOTSUP;
    if (a.batch == 0) return KRON_OK;
    sh.nx = (int)nx;
    sh.buf = (int)buf;
    sh.slice = (int)((slice + 1) & ~1LL);  // 16-byte aligned slices for doubles
    const size_t lds = (size_t)4 * sh.slice * sizeof(T);
    Entries<T, Layout::Ptr> e{a.pA, a.pX, a.pY, a.d};
    const void *kern = a.y_shared ? reinterpret_cast<const void *>(kron_small_kernel<T, true>)
                                  : reinterpret_cast<const void *>(kron_small_kernel<T, false>);
    const long long resident = persistent_blocks(kern, lds);
    const long long need = ((long long)a.batch + 3) / 4;
    const long long grid = std::max(1LL, std::min(need, resident > 0 ? resident : 2048LL));
    if (a.y_shared)
        hipLaunchKernelGGL((kron_small_kernel<T, true>), dim3((unsigned)grid), dim3(256), lds,
                           a.stream, e, sh, (long long)a.batch);
    else
        hipLaunchKernelGGL((kron_small_kernel<T, false>), dim3((unsigned)grid), dim3(256), lds,
                           a.stream, e, sh, (long long)a.batch);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error((int)err);
        return KRON_EHIP;
    }
    return KRON_OK;
}

template kron_status launch_small<double>(const KronArgs<double> &);
template kron_status launch_small<float>(const KronArgs<float> &);

}  // namespace kron
