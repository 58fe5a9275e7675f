// kron_group.hip -- order the entries of a pointer-layout call by Y target.
//
// The reference merges every entry into its Y vector with one atomic add per
// element (kgemm_nn.hpp:392-394).  On MI355X FP64 atomics run at about
// 1.9e11 element-updates/s, so at d=6, n=4 the atomics alone bound a batch of
// 65,536 entries to ~1.4 ms.  Putting the entries with the same Y pointer
// next to each other lets a workgroup accumulate consecutive entries with the
// same target in registers and flush once per run.
//
// The order only has to make equal targets adjacent, so it is a counting sort
// by a hash of the pointer rather than a full sort: clear the bucket counts,
// histogram, exclusive scan (rocPRIM), scatter -- instead of a radix sort of
// 64-bit keys (about 45 us for 65,536 keys, launch-bound).  Entries in one
// bucket land in an arbitrary order; a bucket holding two targets interleaves
// them, which costs extra flushes but never correctness: the kernels flush
// whenever the full 64-bit target changes.  (The merged sums are not
// bit-reproducible in any case: a run split between two workgroups is added
// to Y by two atomics in either order.)
#include <algorithm>

#include <rocprim/device/device_scan.hpp>

#include "kron_internal.hpp"

namespace kron {

// Buckets: the power of two at least the batch, within [1024, 65536].  With
// T distinct targets about T^2 / (2 * buckets) pairs share a bucket.
static int group_bits(int batch)
{
    int bits = 10;
    while (bits < 16 && (1 << bits) < batch) ++bits;
    return bits;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

static size_t scan_temp_bytes(int nb)
{
    size_t bytes = 0;
    (void)rocprim::exclusive_scan(nullptr, bytes, (const int *)nullptr, (int *)nullptr, 0,
                                  (size_t)nb, rocprim::plus<int>(), hipStream_t(0));
    return bytes;
}

size_t group_workspace_bytes(int batch)
{
    const int nb = 1 << group_bits(batch);
    return align256((size_t)batch * 8) + align256((size_t)batch * 4) + 2 * align256((size_t)nb * 4) +
           align256(scan_temp_bytes(nb));
}

// Fibonacci hashing of the pointer (targets are at least 4-byte aligned).
__device__ __forceinline__ unsigned group_bucket(unsigned long long p, int bits)
{
    return (unsigned)(((p >> 2) * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

__global__ void __launch_bounds__(256) kron_group_clear(int *counts, int nb)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nb; i += gridDim.x * 256) counts[i] = 0;
}

__global__ void __launch_bounds__(256)
kron_group_hist(const unsigned long long *pY, int n, int bits, int *counts)
{
    for (int b = blockIdx.x * 256 + threadIdx.x; b < n; b += gridDim.x * 256)
        atomicAdd(counts + group_bucket(pY[b], bits), 1);
}

__global__ void __launch_bounds__(256)
kron_group_scatter(const unsigned long long *pY, int n, int bits, int *cursor, int *perm,
                   unsigned long long *keys)
{
    for (int b = blockIdx.x * 256 + threadIdx.x; b < n; b += gridDim.x * 256) {
        const unsigned long long p = pY[b];
        const int pos = atomicAdd(cursor + group_bucket(p, bits), 1);
        perm[pos] = b;
        keys[pos] = p;
    }
}

kron_status group_by_target(const void *const *pY, int batch, void *ws, size_t ws_bytes,
                            const unsigned long long **keys_sorted, const int **perm,
                            hipStream_t stream)
{
    if (batch < 1) return KRON_EINVAL;
    if (ws_bytes < group_workspace_bytes(batch)) return KRON_EWORKSPACE;
    const int bits = group_bits(batch);
    const int nb = 1 << bits;
    unsigned char *p = static_cast<unsigned char *>(ws);
    unsigned long long *keys_out = reinterpret_cast<unsigned long long *>(p);
    p += align256((size_t)batch * 8);
    int *perm_out = reinterpret_cast<int *>(p);
    p += align256((size_t)batch * 4);
    int *counts = reinterpret_cast<int *>(p);
    p += align256((size_t)nb * 4);
    int *cursor = reinterpret_cast<int *>(p);
    p += align256((size_t)nb * 4);
    size_t temp_bytes = scan_temp_bytes(nb);
    const auto keys_in = reinterpret_cast<const unsigned long long *>(pY);
    const unsigned grid = (unsigned)std::min(1024, (batch + 255) / 256);
    hipLaunchKernelGGL(kron_group_clear, dim3((unsigned)std::min(256, nb / 256)), dim3(256), 0,
                       stream, counts, nb);
    hipLaunchKernelGGL(kron_group_hist, dim3(grid), dim3(256), 0, stream, keys_in, batch, bits,
                       counts);
    hipError_t se = rocprim::exclusive_scan(p, temp_bytes, counts, cursor, 0, (size_t)nb,
                                            rocprim::plus<int>(), stream);
    if (se != hipSuccess) {
        set_last_hip_error((int)se);
        return KRON_EHIP;
    }
    hipLaunchKernelGGL(kron_group_scatter, dim3(grid), dim3(256), 0, stream, keys_in, batch, bits,
                       cursor, perm_out, keys_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    *keys_sorted = keys_out;
    *perm = perm_out;
    return KRON_OK;
}

}  // namespace kron
