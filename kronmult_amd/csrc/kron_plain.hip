// kron_plain.hip -- the reference's per-entry vbatched structure, on the GPU.
//
// kronmult_vbatched.hpp:164-247 gives each batch entry one workgroup: X is
// copied into the entry's scratch slot (:187-192), then the modes are
// contracted one at a time, slowest first (kronmultv.hpp:19-148), every
// intermediate going through the slot's two global buffers, and the last
// mode is added into Y with atomics (kgemm_nn.hpp:392-394); sub-batches of
// Wcapacity / (2 sizeXW) entries run one after another (:157, :164).  This
// file restates that structure without the reference's code, as the plain
// per-entry kernel that the mixed-shape persistent launch (kron_mixed.hip)
// is measured against.  It is not on any fast path.
#include <algorithm>

#include "kron_device.hpp"

namespace kron {

struct PlainDims {
    int m[kMaxDim];
    int n[kMaxDim];
};

template <typename T>
__global__ void __launch_bounds__(256)
kron_plain_kernel(int d, PlainDims dims, const T *const *pA, const T *const *pX, T *const *pY,
                  T *W, long long slot, long long sizeX, int b0)
{
    const int iz = blockIdx.x;
    const int b = b0 + iz;
    T *src = W + (long long)iz * 2 * slot;
    T *dst = src + slot;
    const T *xin = pX[b];
    for (long long i = threadIdx.x; i < sizeX; i += blockDim.x) src[i] = xin[i];
    __syncthreads();
    for (int k = 0; k < d; ++k) {
        // modes before k are contracted (extents m), modes after k are not (n)
        int outer = 1, inner = 1;
        for (int r = 0; r < k; ++r) outer *= dims.m[r];
        for (int r = k + 1; r < d; ++r) inner *= dims.n[r];
        const int mk = dims.m[k], nk = dims.n[k];
        const T *A = pA[(long long)b * d + k];  // mk x nk, leading dimension mk
        const long long total = (long long)outer * mk * inner;
        const bool last = k == d - 1;
        T *y = pY[b];
        for (long long w = threadIdx.x; w < total; w += blockDim.x) {
            const long long ii = w % inner, rest = w / inner;
            const int i = (int)(rest % mk);
            const long long o = rest / mk;
            T acc = 0;
            for (int j = 0; j < nk; ++j) acc += A[i + (long long)j * mk] * src[(o * nk + j) * inner + ii];
            if (last)
                atomicAdd(y + w, acc);
            else
                dst[w] = acc;
        }
        __syncthreads();
        T *t = src;
        src = dst;
        dst = t;
    }
}

template <typename T>
kron_status launch_plain(int d, const int *m, const int *n, const T *const *pA,
                         const T *const *pX, T *const *pY, T *W, long long slot, int nslots,
                         int batch, hipStream_t stream)
{
    PlainDims dims{};
    long long sizeX = 1;
    for (int k = 0; k < kMaxDim; ++k) {
        dims.m[k] = k < d ? m[k] : 1;
        dims.n[k] = k < d ? n[k] : 1;
        if (k < d) sizeX *= n[k];
    }
    for (int b0 = 0; b0 < batch; b0 += nslots) {
        const int count = std::min(nslots, batch - b0);
        hipLaunchKernelGGL(kron_plain_kernel<T>, dim3((unsigned)count), dim3(256), 0, stream, d,
                           dims, pA, pX, pY, W, slot, sizeX, b0);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    return KRON_OK;
}

template kron_status launch_plain<double>(int, const int *, const int *, const double *const *,
                                          const double *const *, double *const *, double *,
                                          long long, int, int, hipStream_t);
template kron_status launch_plain<float>(int, const int *, const int *, const float *const *,
                                         const float *const *, float *const *, float *, long long,
                                         int, int, hipStream_t);

}  // namespace kron
