// kgemm_batched.hip -- batched small GEMM with per-entry pointers and leading
// dimensions, uniform (mm, nn, kk).
//
// Replaces kgemm_nn_batched / kgemm_nt_batched (kgemm_nn_batched.hpp:8-79,
// kgemm_nt_batched.hpp:8-75) and their device GEMMs (kgemm_nn.hpp:227-406,
// kgemm_nt.hpp:215-385).  Each output element is one work item; consecutive
// threads take consecutive rows i, so loads of A(i,k) and the store of C(i,j)
// are coalesced and B(k,j) (or B(j,k)) is a broadcast; outputs of at most
// 128 elements pack several entries into a workgroup.  Accumulation is in
// double for both element types (kgemm_nn.hpp:688-701, kgemm_nt.hpp:557-571),
// summing k = 1..kk in order as the reference does.  Large outputs are split
// over several workgroups per entry.
//
// beta rules (kgemm_nn.hpp:391-400, kgemm_nt.hpp:371-379):
//   beta == 1 -> atomic add of alpha*AB (C may be shared between entries)
//   beta == 0 -> C = alpha*AB, C is never read
//   otherwise -> C = beta*C + alpha*AB
#include "kron_device.hpp"

namespace kron {

template <typename T, bool TRANSB, int BETA_MODE, bool PACKED>
__global__ void __launch_bounds__(256)
kgemm_batched_kernel(int mm, int nn, int kk, T alpha, const T *const *Aarray,
                     const int *ldAarray, const T *const *Barray, const int *ldBarray,
                     T beta, T *const *Carray, const int *ldCarray, int tiles, int epw,
                     int batch)
{
    // PACKED: small outputs (mm*nn <= 128), epw entries per workgroup, each
    // on mm*nn consecutive threads; otherwise `tiles` workgroups per entry
    const long long total = (long long)mm * nn;
    int b, tile, t0, nthr;
    if constexpr (PACKED) {
        b = blockIdx.x * epw + (int)(threadIdx.x / total);
        tile = 0;
        t0 = (int)(threadIdx.x % total);
        nthr = (int)total;
        if (threadIdx.x >= epw * total) return;
    } else {
        b = blockIdx.x / tiles;
        tile = blockIdx.x % tiles;
        t0 = threadIdx.x;
        nthr = blockDim.x;
    }
    if (b >= batch) return;
    const T *__restrict__ A = Aarray[b];
    const T *__restrict__ B = Barray[b];
    T *C = Carray[b];
    const int lda = ldAarray[b], ldb = ldBarray[b], ldc = ldCarray[b];
    const double dalpha = alpha, dbeta = beta;
    const long long step = (long long)tiles * nthr;
    for (long long w = (long long)tile * nthr + t0; w < total; w += step) {
        const int i = (int)(w % mm);
        const int j = (int)(w / mm);
        double c = 0;
        const T *ap = A + i;
        const T *bp = TRANSB ? (B + j) : (B + (long long)j * ldb);
        const long long binc = TRANSB ? ldb : 1;
        for (int k = 0; k < kk; ++k) {
            c = __builtin_fma(double(ap[(long long)k * lda]), double(bp[k * binc]), c);
        }
        T *cp = C + i + (long long)j * ldc;
        if constexpr (BETA_MODE == 1) {
            atomic_add(cp, T(dalpha * c));
        } else if constexpr (BETA_MODE == 0) {
            *cp = T(dalpha * c);
        } else {
            // nn rounds alpha*cij to T before the beta update (kgemm_nn.hpp:391,399)
            const double ab = TRANSB ? dalpha * c : double(T(dalpha * c));
            *cp = T(dbeta * double(*cp) + ab);
        }
    }
}

template <typename T>
kron_status launch_kgemm(bool transB, int mm, int nn, int kk, T alpha, const T *const *A,
                         const int *ldA, const T *const *B, const int *ldB, T beta,
                         T *const *C, const int *ldC, int batch, hipStream_t stream)
{
    if (batch == 0 || mm == 0 || nn == 0) return KRON_OK;
    const long long total = (long long)mm * nn;
    // about 4 outputs per thread, at most 1024 workgroups per entry
    long long tiles = (total + 1023) / 1024;
    if (tiles > 1024) tiles = 1024;
    if (tiles < 1) tiles = 1;
    // outputs of at most 128 elements: several entries per workgroup
    const int epw = total <= 128 ? (int)(256 / total) : 1;
    const long long grid = epw > 1 ? ((long long)batch + epw - 1) / epw : tiles * (long long)batch;
    if (grid > 0x7fffffffLL) return KRON_EINVAL;
    const int mode = (beta == T(1)) ? 1 : (beta == T(0)) ? 0 : 2;
#define KGEMM_LAUNCH(TB, MODE)                                                              \
    do {                                                                                    \
        if (epw > 1)                                                                        \
            hipLaunchKernelGGL((kgemm_batched_kernel<T, TB, MODE, true>),                   \
                               dim3((unsigned)grid), dim3(256), 0, stream, mm, nn, kk,      \
                               alpha, A, ldA, B, ldB, beta, C, ldC, (int)tiles, epw,        \
                               batch);                                                      \
        else                                                                                \
            hipLaunchKernelGGL((kgemm_batched_kernel<T, TB, MODE, false>),                  \
                               dim3((unsigned)grid), dim3(256), 0, stream, mm, nn, kk,      \
                               alpha, A, ldA, B, ldB, beta, C, ldC, (int)tiles, epw,        \
                               batch);                                                      \
    } while (0)
    if (transB) {
        if (mode == 1) KGEMM_LAUNCH(true, 1);
        else if (mode == 0) KGEMM_LAUNCH(true, 0);
        else KGEMM_LAUNCH(true, 2);
    } else {
        if (mode == 1) KGEMM_LAUNCH(false, 1);
        else if (mode == 0) KGEMM_LAUNCH(false, 0);
        else KGEMM_LAUNCH(false, 2);
    }
#undef KGEMM_LAUNCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    return KRON_OK;
}

template kron_status launch_kgemm<double>(bool, int, int, int, double, const double *const *,
                                          const int *, const double *const *, const int *,
                                          double, double *const *, const int *, int,
                                          hipStream_t);
template kron_status launch_kgemm<float>(bool, int, int, int, float, const float *const *,
                                         const int *, const float *const *, const int *, float,
                                         float *const *, const int *, int, hipStream_t);

}  // namespace kron
