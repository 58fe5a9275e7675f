// kgemm_batched.hip -- batched small GEMM with per-entry pointers and leading
// dimensions, uniform (mm, nn, kk).
//
// Replaces kgemm_nn_batched / kgemm_nt_batched (kgemm_nn_batched.hpp:8-79,
// kgemm_nt_batched.hpp:8-75) and their device GEMMs (kgemm_nn.hpp:227-406,
// kgemm_nt.hpp:215-385).  Accumulation is in double for both element types
// (kgemm_nn.hpp:688-701, kgemm_nt.hpp:557-571), summing k = 1..kk in order
// as the reference does.
//
// Register-tiled kernel (kgemm_tiled_kernel): a unit is one entry's column
// block of NJ columns and up to 256 of its rows; a thread owns one row i of
// the unit and keeps its NJ sums in registers, so each A(i, k) is loaded
// once and reused NJ times, and loads of A and of C are coalesced across
// rows.  The unit's B block (kk x NJ) is staged in LDS once per workgroup
// and read as broadcasts.  Units with fewer than 256 rows share a workgroup
// (several column blocks or entries).  Shapes whose staged B blocks would
// not fit fall back to kgemm_batched_kernel: one output element per thread,
// B read from global memory.
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
            c = kfma(double(ap[(long long)k * lda]), double(bp[k * binc]), c);
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

// C += / = alpha * A op(B) for units (entry b, column block cb, row chunk rc),
// unit u = (b * ncb + cb) * nrc + rc, R rows per unit, P units per workgroup
// (thread t: slot t / R, row t % R of it).
template <typename T, bool TRANSB, int BETA_MODE, int NJ>
__global__ void __launch_bounds__(256)
kgemm_tiled_kernel(int mm, int nn, int kk, T alpha, const T *const *Aarray, const int *ldAarray,
                   const T *const *Barray, const int *ldBarray, T beta, T *const *Carray,
                   const int *ldCarray, int R, int P, int ncb, int nrc, long long units)
{
    extern __shared__ __align__(16) unsigned char kgemm_smem[];
    T *Bs = reinterpret_cast<T *>(kgemm_smem);  // P blocks of kk x NJ, k-major
    const int t = threadIdx.x;
    const int slot = t / R, i0 = t - slot * R;
    const long long u = (long long)blockIdx.x * P + slot;
    const bool live = slot < P && u < units;
    int b = 0, cb = 0, rc = 0;
    if (live) {
        rc = (int)(u % nrc);
        const long long bc = u / nrc;
        cb = (int)(bc % ncb);
        b = (int)(bc / ncb);
    }
    const int j0 = cb * NJ;
    const int nj = live ? min(NJ, nn - j0) : 0;
    T *bs = Bs + (size_t)(slot < P ? slot : 0) * kk * NJ;
    // stage the unit's B block: element (k, jj) at bs[k * NJ + jj]
    if (live) {
        const T *B = Barray[b];
        const int ldb = ldBarray[b];
        for (int e = i0; e < kk * NJ; e += R) {
            // consecutive threads read consecutive addresses: j fastest for
            // B^T (row-major in j), k fastest for B
            int k, jj;
            if constexpr (TRANSB) {
                k = e / NJ;
                jj = e - k * NJ;
            } else {
                jj = e / kk;
                k = e - jj * kk;
            }
            T v = T(0);
            if (jj < nj) v = TRANSB ? B[(j0 + jj) + (long long)k * ldb]
                                    : B[k + (long long)(j0 + jj) * ldb];
            bs[k * NJ + jj] = v;
        }
    }
    __syncthreads();
    const int i = rc * R + i0;
    if (!live || i >= mm) return;
    const T *A = Aarray[b] + i;
    const long long lda = ldAarray[b];
    double c[NJ];
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) c[jj] = 0;
    int k = 0;
    for (; k + 8 <= kk; k += 8) {
        T a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = A[(k + q) * lda];
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int jj = 0; jj < NJ; ++jj)
                c[jj] = kfma(double(a[q]), double(bs[(k + q) * NJ + jj]), c[jj]);
    }
    for (; k < kk; ++k) {
        const double a = double(A[k * lda]);
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) c[jj] = kfma(a, double(bs[k * NJ + jj]), c[jj]);
    }
    T *C = Carray[b] + i + (long long)j0 * ldCarray[b];
    const long long ldc = ldCarray[b];
    const double dalpha = alpha, dbeta = beta;
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        if (jj < nj) {
            T *cp = C + jj * ldc;
            if constexpr (BETA_MODE == 1) {
                atomic_add(cp, T(dalpha * c[jj]));
            } else if constexpr (BETA_MODE == 0) {
                *cp = T(dalpha * c[jj]);
            } else {
                // nn rounds alpha*cij to T before the beta update (kgemm_nn.hpp:391,399)
                const double ab = TRANSB ? dalpha * c[jj] : double(T(dalpha * c[jj]));
                *cp = T(dbeta * double(*cp) + ab);
            }
        }
    }
}

// LDS budget of the tiled kernel's staged B blocks (bytes).
constexpr size_t kKgemmLds = 32768;

template <typename T>
kron_status launch_kgemm(bool transB, int mm, int nn, int kk, T alpha, const T *const *A,
                         const int *ldA, const T *const *B, const int *ldB, T beta,
                         T *const *C, const int *ldC, int batch, hipStream_t stream)
{
    if (batch == 0 || mm == 0 || nn == 0) return KRON_OK;
    // tiled kernel: NJ = the smallest of 1, 2, 4, 8, 16 covering min(nn, 16).
    // Outputs of at most 128 elements with fewer than 8 rows leave it too few
    // threads per entry (mm of them): those run packed, one output per thread
    // (4^3 x 1M entries 0.192 -> 0.160 ms, 4x8x4 x 512K 0.190 -> 0.145 ms; 8^3
    // stays tiled, 0.117 against 0.220 ms packed)
    if (!((long long)mm * nn <= 128 && mm < 8)) {
        const int want = nn < 16 ? nn : 16;
        int NJ = 1;
        while (NJ < want) NJ *= 2;
        const int R = mm < 256 ? mm : 256;
        const int P = 256 / R;
        const int ncb = (nn + NJ - 1) / NJ, nrc = (mm + R - 1) / R;
        const size_t lds = (size_t)P * kk * NJ * sizeof(T);
        const long long units = (long long)batch * ncb * nrc;
        const long long grid = (units + P - 1) / P;
        if (kk > 0 && lds <= kKgemmLds && grid <= 0x7fffffffLL) {
            const int mode = (beta == T(1)) ? 1 : (beta == T(0)) ? 0 : 2;
#define KGEMM_TILED(TB, MODE, NJV)                                                           \
    hipLaunchKernelGGL((kgemm_tiled_kernel<T, TB, MODE, NJV>), dim3((unsigned)grid), dim3(256), \
                       lds, stream, mm, nn, kk, alpha, A, ldA, B, ldB, beta, C, ldC, R, P, ncb, \
                       nrc, units)
#define KGEMM_TILED_NJ(TB, MODE)                                                             \
    switch (NJ) {                                                                           \
    case 1: KGEMM_TILED(TB, MODE, 1); break;                                                \
    case 2: KGEMM_TILED(TB, MODE, 2); break;                                                \
    case 4: KGEMM_TILED(TB, MODE, 4); break;                                                \
    case 8: KGEMM_TILED(TB, MODE, 8); break;                                                \
    default: KGEMM_TILED(TB, MODE, 16); break;                                              \
    }
            if (transB) {
                if (mode == 1) { KGEMM_TILED_NJ(true, 1) }
                else if (mode == 0) { KGEMM_TILED_NJ(true, 0) }
                else { KGEMM_TILED_NJ(true, 2) }
            } else {
                if (mode == 1) { KGEMM_TILED_NJ(false, 1) }
                else if (mode == 0) { KGEMM_TILED_NJ(false, 0) }
                else { KGEMM_TILED_NJ(false, 2) }
            }
#undef KGEMM_TILED_NJ
#undef KGEMM_TILED
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) {
                set_last_hip_error((int)e);
                return KRON_EHIP;
            }
            return KRON_OK;
        }
    }
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
