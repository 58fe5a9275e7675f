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
// by a hash of the pointer rather than a full sort (a radix sort of 64-bit
// keys takes about 45 us for 65,536 keys, launch-bound).  Entries in one
// bucket land in an arbitrary order; a bucket holding two targets interleaves
// them, which costs extra flushes but never correctness: the kernels flush
// whenever the full 64-bit target changes.  (The merged sums are not
// bit-reproducible in any case: a run split between two workgroups is added
// to Y by two atomics in either order.)
#include <algorithm>

#include "kron_internal.hpp"

namespace kron {

// Two launches per call:
//   count:   rank[b] = the number of earlier arrivals in b's bucket (the
//            return value of the bucket counter's atomic add); the first
//            arrival in a bucket also appends the bucket to a list of used
//            buckets.  The last workgroup to finish gives every used bucket a
//            contiguous range of positions (wave prefix sums of the counts,
//            one shared-memory atomic per wave for the wave's base) and sets
//            their counts back to zero;
//   scatter: entry b goes to position base[bucket] + rank[b].
// The bucket counts, the used-bucket count and the finishing ticket live at
// the start of the workspace and are zero between calls: cleared once when
// the workspace is new, restored by every call's last workgroup.  The ranges
// land in an arbitrary order; equal targets are still contiguous.

// Buckets: the power of two at least the batch, within [1024, 65536]
// (group_bits).  With T distinct targets about T^2 / (2 * buckets) pairs
// share a bucket.  The workspace layout is group_workspace_bytes's.

// Fibonacci hashing of the pointer (targets are at least 4-byte aligned).
__device__ __forceinline__ unsigned group_bucket(unsigned long long p, int bits)
{
    return (unsigned)(((p >> 2) * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

__global__ void __launch_bounds__(256) kron_group_clear(int *zero, int n)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) zero[i] = 0;
}

__device__ __forceinline__ int wave_inclusive_sum(int x, int lane)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

constexpr int kCountThreads = 512;  // 64 -> 128 workgroups at 65,536 entries: more CUs on the atomics
constexpr int kCountUnroll = 8;  // used buckets per thread and step of the last workgroup

__global__ void __launch_bounds__(kCountThreads)
kron_group_count(const unsigned long long *pY, int n, int bits, int *counts, unsigned *ticket,
                 int *nused, int *used, int *rank, int *base)
{
    __shared__ int lused[kCountThreads];
    __shared__ int lcount, lbase, last, cursor;
    const int t = threadIdx.x, lane = t & 63;
    // 64-bit positions: b0 + t and the stride may pass INT_MAX near the end
    for (long long b0 = (long long)blockIdx.x * kCountThreads; b0 < n;
         b0 += (long long)gridDim.x * kCountThreads) {
        if (t == 0) lcount = 0;
        __syncthreads();
        const long long b = b0 + t;
        if (b < n) {
            const int k = (int)group_bucket(pY[b], bits);
            const int r = atomicAdd(counts + k, 1);
            rank[b] = r;
            if (r == 0) lused[atomicAdd(&lcount, 1)] = k;
        }
        __syncthreads();
        if (t == 0) lbase = atomicAdd(nused, lcount);
        __syncthreads();
        // wave 0 stores the list, so its own release fence below covers it
        if (t < 64)
            for (int i = t; i < lcount; i += 64) used[lbase + i] = lused[i];
        __syncthreads();  // lcount and lused are reused
    }

    // Data the last workgroup reads: the counts (atomics, complete: their
    // values were returned) and the used-bucket list (stored by wave 0).  One
    // agent-scope release fence by wave 0 publishes the list; on this device
    // an agent fence writes back the L2 of the workgroup's XCD, so it is
    // issued once per workgroup, not per wave.  The last workgroup's acquire
    // fence likewise invalidates the caches all its waves read through.
    __syncthreads();
    if (t == 0) {
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
        if (last) __threadfence();
        cursor = 0;
    }
    __syncthreads();
    if (!last) return;

    // the last workgroup: ranges of the m used buckets
    const int m = __hip_atomic_load(nused, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int i0 = 0; i0 < m; i0 += kCountThreads * kCountUnroll) {
        int k[kCountUnroll], c[kCountUnroll];
#pragma unroll
        for (int u = 0; u < kCountUnroll; ++u) {
            const int i = i0 + u * kCountThreads + t;
            k[u] = i < m ? used[i] : -1;
        }
#pragma unroll
        for (int u = 0; u < kCountUnroll; ++u) c[u] = k[u] >= 0 ? counts[k[u]] : 0;
#pragma unroll
        for (int u = 0; u < kCountUnroll; ++u) {
            if (i0 + u * kCountThreads >= m) break;
            const int incl = wave_inclusive_sum(c[u], lane);
            int wbase = 0;
            if (lane == 63) wbase = atomicAdd(&cursor, incl);
            wbase = __shfl(wbase, 63, 64);
            if (k[u] >= 0) {
                base[k[u]] = wbase + incl - c[u];
                counts[k[u]] = 0;
            }
        }
    }
    if (t == 0) {
        *ticket = 0u;
        *nused = 0;
    }
}

__global__ void __launch_bounds__(256)
kron_group_scatter(const unsigned long long *pY, int n, int bits, const int *rank,
                   const int *base, int *perm, unsigned long long *keys)
{
    for (long long b = (long long)blockIdx.x * 256 + threadIdx.x; b < n;
         b += (long long)gridDim.x * 256) {
        const unsigned long long p = pY[b];
        const int pos = base[group_bucket(p, bits)] + rank[b];
        perm[pos] = (int)b;
        keys[pos] = p;
    }
}

kron_status group_by_target(const void *const *pY, int batch, void *ws, size_t ws_bytes,
                            bool fresh, const unsigned long long **keys_sorted, const int **perm,
                            hipStream_t stream)
{
    if (batch < 1) return KRON_EINVAL;
    if (ws_bytes < group_workspace_bytes(batch)) return KRON_EWORKSPACE;
    const int bits = group_bits(batch);
    const int nb = 1 << bits;
    unsigned char *p = static_cast<unsigned char *>(ws);
    int *counts = reinterpret_cast<int *>(p);
    unsigned *ticket = reinterpret_cast<unsigned *>(p + kGroupCountBytes);
    int *nused = reinterpret_cast<int *>(p + kGroupCountBytes + 4);
    p += kGroupZeroBytes;
    int *base = reinterpret_cast<int *>(p);
    p += group_align256((size_t)nb * 4);
    int *used = reinterpret_cast<int *>(p);
    p += group_align256((size_t)nb * 4);
    int *rank = reinterpret_cast<int *>(p);
    p += group_align256((size_t)batch * 4);
    int *perm_out = reinterpret_cast<int *>(p);
    p += group_align256((size_t)batch * 4);
    unsigned long long *keys_out = reinterpret_cast<unsigned long long *>(p);
    const auto keys_in = reinterpret_cast<const unsigned long long *>(pY);
    if (fresh)
        hipLaunchKernelGGL(kron_group_clear, dim3(256), dim3(256), 0, stream, counts,
                           (int)(kGroupZeroBytes / 4));
    const unsigned gc = (unsigned)std::min(256, (batch + kCountThreads - 1) / kCountThreads);
    hipLaunchKernelGGL(kron_group_count, dim3(gc), dim3(kCountThreads), 0, stream, keys_in, batch,
                       bits, counts, ticket, nused, used, rank, base);
    const unsigned gs = (unsigned)std::min(1024, (batch + 255) / 256);
    hipLaunchKernelGGL(kron_group_scatter, dim3(gs), dim3(256), 0, stream, keys_in, batch, bits,
                       rank, base, perm_out, keys_out);
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
