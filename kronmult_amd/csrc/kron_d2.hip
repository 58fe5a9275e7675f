// kron_d2.hip -- d = 2 entries beyond the fused kernels: Y += (A1 (x) A2) X
// with factor extents up to 32 (n > 8, or rectangular), off the matrix cores.
//
// With A1 on the slow index an entry's X is the n1 x n2 matrix X[j1][j2]
// (row j1 contiguous) and
//     T[j1][k2] = sum_j2 A2(k2, j2) X[j1][j2],   Y[k1][k2] += sum_j1 A1(k1, j1) T[j1][k2]
// (kronmult2.hpp reaches the same two products through kgemm calls).  The
// generic kernel gives such a small entry a whole workgroup and stalls on it;
// here a wave takes an entry: X is copied into the wave's LDS slice with
// coalesced loads, lanes (p, k2) form T's column k2 for rows p, p + P, ...
// (P = 64 / m2 row groups; the factor's column j2 is contiguous over k2:
// coalesced) and then Y's column k2 the same way from T in LDS.  The four
// waves of a workgroup step through entries together (workgroup barriers
// order each wave's LDS slice between entries).  The factors are copied into
// the slice too.  When four such slices would pass 32 KB (square extents
// above 16 in FP64, above 22 in FP32), the whole workgroup takes one entry
// instead (P = 256 / m2 row groups), its factors still in LDS: read from
// global memory inside the FMA loops, their latency dominated (d2n24 FP64:
// 0.7 TB/s with a wave per entry and the factors in global memory).
#include <algorithm>

#include "kron_device.hpp"
#include "kron_internal.hpp"

namespace kron {

struct D2Shape {
    int m1, n1, m2, n2;  // A1: m1 x n1 (slow index), A2: m2 x n2 (fast index)
    int ld1, ld2;        // leading dimensions of A1, A2
    int slice;           // LDS elements per slice: X (n1 n2), T (n1 m2), A1, A2
    int per;             // entries per workgroup: 4 (a wave and a slice each) or 1
};

template <typename T, Layout LAY, bool SHARED>
__global__ void __launch_bounds__(256)
kron_d2_kernel(Entries<T, LAY> ent, D2Shape sh, long long batch)
{
    extern __shared__ __align__(16) unsigned char kron_smem[];
    const int tpe = 256 / sh.per;  // threads per entry: 64 or 256
    const int lane = threadIdx.x & (tpe - 1), wid = threadIdx.x / tpe;
    T *const xs = reinterpret_cast<T *>(kron_smem) + wid * sh.slice;
    T *const ts = xs + sh.n1 * sh.n2;
    T *const a1s = ts + sh.n1 * sh.m2;  // A1 (m1 x n1, ld m1)
    T *const a2s = a1s + sh.m1 * sh.n1; // A2 (m2 x n2, ld m2)
    const int P = tpe / sh.m2;  // row groups
    const int p = lane / sh.m2, k2 = lane % sh.m2;
    const bool act = p < P;
    const int nx = sh.n1 * sh.n2;
    const int per = sh.per;
    // private Y: a thread's first YPF outputs are read with the entry's other
    // loads, so phase 2 does not wait on them
    constexpr int YPF = SHARED ? 1 : 4;
    for (long long g = (long long)blockIdx.x * per; g < batch; g += (long long)gridDim.x * per) {
        const long long b = g + wid;  // this slice's entry; idle past the end
        const bool live = b < batch;
        T yv[YPF];
        if constexpr (!SHARED) {
            if (live && act) {
                const auto y = to_global(ent.y((int)b));
#pragma unroll
                for (int i = 0; i < YPF; ++i) {
                    const int k1 = p + i * P;
                    yv[i] = k1 < sh.m1 ? y[(long long)k1 * sh.m2 + k2] : T(0);
                }
            }
        }
        if (live) {
            const auto x = to_global(ent.x((int)b));
            for (int w = lane; w < nx; w += tpe) xs[w] = x[w];
            // both factors, compact (ld = rows), with coalesced column reads
            const auto A1 = to_global(ent.factor((int)b, 0));
            const auto A2 = to_global(ent.factor((int)b, 1));
            for (int w = lane; w < sh.m1 * sh.n1; w += tpe)
                a1s[w] = A1[(w % sh.m1) + (w / sh.m1) * sh.ld1];
            for (int w = lane; w < sh.m2 * sh.n2; w += tpe)
                a2s[w] = A2[(w % sh.m2) + (w / sh.m2) * sh.ld2];
        }
        __syncthreads();
        if (live && act) {
            for (int j1 = p; j1 < sh.n1; j1 += P) {
                T acc = 0;
#pragma unroll 4
                for (int j2 = 0; j2 < sh.n2; ++j2)
                    acc = kfma(a2s[k2 + j2 * sh.m2], xs[j1 * sh.n2 + j2], acc);
                ts[j1 * sh.m2 + k2] = acc;
            }
        }
        __syncthreads();
        if (live && act) {
            const auto y = to_global(ent.y((int)b));
            auto row = [&](int k1) __attribute__((always_inline)) {
                T acc = 0;
#pragma unroll 4
                for (int j1 = 0; j1 < sh.n1; ++j1)
                    acc = kfma(a1s[k1 + j1 * sh.m1], ts[j1 * sh.m2 + k2], acc);
                return acc;
            };
            int k1 = p;
            if constexpr (!SHARED) {
#pragma unroll
                for (int i = 0; i < YPF; ++i, k1 += P)
                    if (k1 < sh.m1) y[(long long)k1 * sh.m2 + k2] = yv[i] + row(k1);
            }
            for (; k1 < sh.m1; k1 += P)
                y_accumulate<SHARED, T, T>(y, (long long)k1 * sh.m2 + k2, row(k1));
        }
        __syncthreads();  // the slices are rewritten by the next entries
    }
}

template <typename T, Layout LAY, bool SHARED>
static kron_status d2_launch(const Entries<T, LAY> &e, const D2Shape &sh, int batch,
                             hipStream_t s)
{
    const size_t lds = (size_t)sh.per * sh.slice * sizeof(T);
    const void *kern = reinterpret_cast<const void *>(kron_d2_kernel<T, LAY, SHARED>);
    if (lds > 65536) {
        hipError_t err = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds);
        if (err != hipSuccess) {
            set_last_hip_error((int)err);
            return KRON_EHIP;
        }
    }
    const long long resident = persistent_blocks(kern, lds);
    const long long need = ((long long)batch + sh.per - 1) / sh.per;
    const long long grid = std::max(1LL, std::min(need, resident > 0 ? resident : 2048LL));
    hipLaunchKernelGGL((kron_d2_kernel<T, LAY, SHARED>), dim3((unsigned)grid), dim3(256), lds, s,
                       e, sh, (long long)batch);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error((int)err);
        return KRON_EHIP;
    }
    return KRON_OK;
}

template <typename T>
kron_status launch_d2(const KronArgs<T> &a)
{
    if (a.d != 2 || a.trans) return KRON_ENOTSUP;
    D2Shape sh;
    sh.m1 = a.m[0];
    sh.n1 = a.n[0];
    sh.m2 = a.m[1];
    sh.n2 = a.n[1];
    if (std::max(std::max(sh.m1, sh.n1), std::max(sh.m2, sh.n2)) > 32) return KRON_ENOTSUP;
    if (sh.m1 < 1 || sh.n1 < 1 || sh.m2 < 1 || sh.n2 < 1) return KRON_EINVAL;
    sh.ld1 = a.lda > 0 ? a.lda : sh.m1;
    sh.ld2 = a.lda > 0 ? a.lda : sh.m2;
    if (sh.ld1 < sh.m1 || sh.ld2 < sh.m2) return KRON_EINVAL;
    // four slices (a wave per entry) with the factors while they stay within
    // 32 KB (five workgroups per CU); above that one slice per workgroup,
    // factors included (at most 4 x 32 x 32 elements: 32 KB in FP64)
    const int base = sh.n1 * sh.n2 + sh.n1 * sh.m2;
    const int with_fac = base + sh.m1 * sh.n1 + sh.m2 * sh.n2;
    sh.per = (size_t)4 * with_fac * sizeof(T) <= 32768 ? 4 : 1;
    sh.slice = (with_fac + 1) & ~1;  // 16-byte aligned slices for doubles
    if (a.batch == 0) return KRON_OK;
    if (a.layout == Layout::Batched) {
        // batched layout: square factors, n x n each, entries n^2 apart
        Entries<T, Layout::Batched> e{a.A, a.X, a.Y, 2, (long long)sh.m1 * sh.n1,
                                      (long long)sh.n1 * sh.n2, (long long)sh.m1 * sh.m2};
        return d2_launch<T, Layout::Batched, false>(e, sh, a.batch, a.stream);
    }
    Entries<T, Layout::Ptr> e{a.pA, a.pX, a.pY, 2};
    if (a.y_shared) return d2_launch<T, Layout::Ptr, true>(e, sh, a.batch, a.stream);
    return d2_launch<T, Layout::Ptr, false>(e, sh, a.batch, a.stream);
}

template kron_status launch_d2<double>(const KronArgs<double> &);
template kron_status launch_d2<float>(const KronArgs<float> &);

}  // namespace kron
